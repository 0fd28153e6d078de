"""Serving simulate-and-image jobs with continuous batching.

Jobs (a configuration of the served model plus a step count) arrive over
time.  ``Server.serve`` admits every arrived job, first come first served,
as soon as its device memory fits the HBM budget, its particles fit the
pool and a replica slot is free; admitted jobs become replicas of one
resident ``EnsembleEngine`` and advance together.  The loop runs the
resident replicas up to the next event (a job reaching its step count or a
frame boundary) and streams back every frame that is due: the frames are
ended and compacted to their nonzero TCSPC bins on the device in one batch,
and their copy to host memory runs while the next batch of steps does (the
reference keeps every frame of a run, pipeline.py:645).  A job is complete
once its last frame has arrived; finished jobs free their memory, which
lets waiting jobs in.  A job's result depends only on its own configuration
(counter-based draws keyed per job), so it is the same alone and in any
mix.

``serve_plain`` runs the same jobs one after another on the single-
membrane path (one job per launch sequence), copying each frame out whole,
the baseline of the batched figure.
"""

from __future__ import annotations

import ctypes
import time
from collections import deque
from dataclasses import dataclass, field
from typing import Callable, Dict, List, Optional

import numpy as np

from .config import SimulationConfig
from .ensemble import EnsembleEngine, replica_bytes
from .microscope import FLIMFrame, SparseFrame


@dataclass
class Job:
    job_id: int
    config: SimulationConfig
    steps: int
    arrival: float = 0.0  # seconds after the start of the stream


@dataclass
class JobResult:
    """A served job: every frame in ``frame_list`` (``SparseFrame`` from the
    server, ``FLIMFrame`` from the plain path), ``frame`` the last one as a
    dense ``FLIMFrame``."""
    job_id: int
    arrival: float
    admitted: float
    finished: float
    steps: int
    particles: int
    frames: int
    frame_list: List = field(default_factory=list)
    counters: Dict[str, int] = field(default_factory=dict)
    state: Optional[Dict[str, np.ndarray]] = None

    @property
    def latency(self) -> float:
        return self.finished - self.arrival

    @property
    def frame(self) -> Optional[FLIMFrame]:
        if not self.frame_list:
            return None
        f = self.frame_list[-1]
        return f if isinstance(f, FLIMFrame) else f.to_frame()

    def sparse_frames(self) -> List[SparseFrame]:
        return [f if isinstance(f, SparseFrame) else SparseFrame.from_stack(
            f.flim_stack, f.frame_index, f.n_signal_photons, f.n_dark_counts, f.n_afterpulses)
            for f in self.frame_list]


def n_particles(config: SimulationConfig) -> int:
    return int(sum(config.initial_populations.values()) + sum(
        ex.cytosol_pool for ex in config.compartment_exchanges if ex.cytosol_pool > 0))


def frame_steps(config: SimulationConfig) -> int:
    return max(1, int(round(config.scan.frame_interval / config.bd_timestep)))


class Server:
    """Continuous batching over one resident ensemble.

    model              a configuration of the served model
    hbm_budget_bytes   device memory the resident jobs may hold together
    capacity           particles of the ensemble's pool
    max_jobs           resident jobs at most
    chunk              steps per launch batch at most
    """

    def __init__(self, model: SimulationConfig, hbm_budget_bytes: int, capacity: int,
                 max_jobs: int = 256, chunk: int = 50, precision: str = "fp32",
                 keep_state: bool = False, engine: Optional[EnsembleEngine] = None):
        # twice the slots: a finished job's slot is reusable only after a step
        self.engine = engine if engine is not None else EnsembleEngine(
            model, capacity, max_replicas=2 * max_jobs, precision=precision)
        self.budget = int(hbm_budget_bytes)
        self.capacity = int(capacity)
        self.max_jobs = int(max_jobs)
        self.chunk = int(chunk)
        self.precision = precision
        self.keep_state = keep_state
        self.used_bytes = 0
        self.peak_resident = 0
        self.launch_batches = 0

    def _fits(self, job: Job, resident: Dict) -> bool:
        b = replica_bytes(job.config, self.precision)
        if b > self.budget:
            raise ValueError(f"job {job.job_id} needs {b} bytes, more than the whole budget")
        if n_particles(job.config) > self.capacity:
            raise ValueError(f"job {job.job_id} has more particles than the pool holds")
        # a finished job's slot and pool space return at the next step
        return (self.used_bytes + b <= self.budget and len(resident) < self.max_jobs
                and self.engine.free_slots() > 0
                and n_particles(job.config) <= self.engine.free_particles())

    def serve(self, jobs: List[Job], realtime: bool = True,
              clock: Callable[[], float] = time.perf_counter,
              progress: Optional[Callable[[str], None]] = None) -> List[JobResult]:
        """Serve the jobs; with realtime, a job is not admitted before its
        arrival time (measured from the call).  Returns results in
        completion order; a job completes when its last frame is on the
        host."""
        eng = self.engine
        pending = deque(sorted(jobs, key=lambda j: (j.arrival, j.job_id)))
        resident: Dict[int, Dict] = {}
        results: List[JobResult] = []
        frames_of: Dict[int, List] = {}   # job id -> frames received so far
        collect = None  # in flight: (ticket, {slot: job id}, finished jobs' records)
        t0 = clock()

        def fetch(c):
            ticket, job_of, done_jobs = c
            got = eng.fetch_frames(ticket)
            for slot, fr in got.items():
                frames_of[job_of[slot]].append(fr)
            for slot, r in done_jobs:
                job = r["job"]
                fl = frames_of.pop(job.job_id)
                results.append(JobResult(job.job_id, job.arrival, r["admitted"], clock() - t0,
                                         job.steps, r["n"], len(fl), fl, r["ctr"], r["state"]))
                if progress:
                    progress(f"finished job {job.job_id} at {clock() - t0:.2f} s")

        while pending or resident or collect is not None:
            now = clock() - t0
            while pending and (not realtime or pending[0].arrival <= now):
                job = pending[0]
                if not self._fits(job, resident):
                    break  # first come, first served: later jobs wait behind it
                pending.popleft()
                slot = eng.add(job.config)
                b = replica_bytes(job.config, self.precision)
                self.used_bytes += b
                n = n_particles(job.config)
                resident[slot] = dict(job=job, done=0, bytes=b, n=n,
                                      fe=frame_steps(job.config), admitted=clock() - t0)
                frames_of[job.job_id] = []
                if progress:
                    progress(f"admitted job {job.job_id} ({n} particles, {job.steps} steps) "
                             f"at {clock() - t0:.2f} s, {len(resident)} resident")
            self.peak_resident = max(self.peak_resident, len(resident))
            if not resident:
                if collect is not None:
                    fetch(collect)
                    collect = None
                elif pending and (not realtime or pending[0].arrival <= clock() - t0):
                    # the head job has arrived but waits for the space of
                    # removed jobs: one empty step returns it
                    eng.run(1)
                elif pending:
                    wait = pending[0].arrival - (clock() - t0)
                    if wait > 0:
                        time.sleep(wait)
                continue
            k = self.chunk
            for r in resident.values():
                k = min(k, r["job"].steps - r["done"], r["fe"] - r["done"] % r["fe"])
            eng.run(k)
            self.launch_batches += 1
            # the previous batch's frames: their copy ran beside these steps
            if collect is not None:
                fetch(collect)
                collect = None
            finished, due = [], []
            for slot, r in resident.items():
                r["done"] += k
                if r["done"] >= r["job"].steps:
                    finished.append(slot)
                    due.append(slot)
                elif r["done"] % r["fe"] == 0:
                    due.append(slot)
            if due:
                ticket = eng.collect_frames(due)
                job_of = {slot: resident[slot]["job"].job_id for slot in due}
                done_jobs = []
                for slot in finished:
                    r = resident.pop(slot)
                    r["ctr"] = eng.counters(slot)
                    r["state"] = eng.state(slot) if self.keep_state else None
                    eng.remove(slot)
                    self.used_bytes -= r["bytes"]
                    done_jobs.append((slot, r))
                collect = (ticket, job_of, done_jobs)
        return results


def serve_plain(jobs: List[Job], realtime: bool = True,
                clock: Callable[[], float] = time.perf_counter,
                keep_state: bool = False,
                progress: Optional[Callable[[str], None]] = None) -> List[JobResult]:
    """The same jobs one at a time, first come first served, each on the
    single-membrane path (FP32, Philox, fused imaging), every frame ended on
    schedule and copied out (VirtualMicroscope.finalize_frame)."""
    from . import _hip
    from .dynamics import DynamicsEngine
    from .microscope import VirtualMicroscope
    from .particles import ParticleState
    from .photophysics import PhotonSimulator

    L = _hip.lib()
    results: List[JobResult] = []
    t0 = clock()
    tot, fms = ctypes.c_double(0), ctypes.c_double(0)
    for job in sorted(jobs, key=lambda j: (j.arrival, j.job_id)):
        if realtime:
            wait = job.arrival - (clock() - t0)
            if wait > 0:
                time.sleep(wait)
        admitted = clock() - t0
        cfg = job.config
        st = ParticleState(cfg)
        eng = DynamicsEngine(cfg, st, precision="fp32", rng="philox", pair_cap=False)
        PhotonSimulator(cfg, st, min(1.0, cfg.scan.dwell_time / cfg.scan.frame_interval))
        mic = VirtualMicroscope(cfg, st)
        fe = frame_steps(cfg)
        done = 0
        frame_list = []
        while done < job.steps:
            k = min(fe - done % fe, job.steps - done)
            _hip.check(L.edm_time_steps(eng.handle, k, 1, ctypes.byref(tot), ctypes.byref(fms),
                                        None))
            done += k
            if done % fe == 0 or done == job.steps:
                frame_list.append(mic.finalize_frame())
        # the steps ran through the C ABI: the host mirrors are stale
        st.mark_stale()
        eng._ctr_cache = None
        ctr = dict(n_catalytic=eng.n_catalytic, n_recruited=eng.n_recruited,
                   n_detached=eng.n_detached, n_converted=eng.n_converted)
        state = None
        if keep_state:
            state = dict(positions=st.positions.copy(), species_id=st.species_id.copy(),
                         is_alive=st.is_alive.copy())
        eng.close()
        results.append(JobResult(job.job_id, job.arrival, admitted, clock() - t0, job.steps,
                                 st.N, len(frame_list), frame_list, ctr, state))
        if progress:
            progress(f"plain job {job.job_id}: {st.N} particles, {job.steps} steps, "
                     f"done at {clock() - t0:.2f} s")
    return results


def latency_stats(results: List[JobResult]) -> Dict[str, float]:
    lat = np.array([r.latency for r in results]) if results else np.zeros(1)
    span = max(r.finished for r in results) - min(r.arrival for r in results) if results else 0
    return dict(jobs=len(results), makespan_s=float(span),
                jobs_per_s=float(len(results) / span) if span > 0 else 0.0,
                p50_s=float(np.percentile(lat, 50)), p99_s=float(np.percentile(lat, 99)),
                mean_s=float(lat.mean()))
