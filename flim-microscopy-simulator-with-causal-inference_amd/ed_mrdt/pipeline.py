"""SimulationPipeline with the reference's interface (pipeline.py), driving
the GPU engines, with its recording done on the device.

The reference records sub-frame data with (1 + n_species) numpy reductions
over the particle arrays after every BD step (pipeline.py:567-642) and
builds each frame's ground truth with a Python loop over all particles
(pipeline.py:357-455).  Here one census kernel per step appends a row of
counts to a device ring (``edm_record_step``; with Philox draws a whole
frame's steps run with no host synchronisation, ``edm_run_frame``), and the
frame-end FRET values come from one kernel over the dye carriers
(``edm_ground_truth``).  The host only turns rows into the reference's
``SubFrameData`` / ``GroundTruth`` fields: deltas of cumulative counters,
complexes = bonded / 2, the mean FRET efficiency.
"""

from __future__ import annotations

import ctypes
import math
import time
from dataclasses import dataclass
from typing import Callable, Dict, List, Optional

import numpy as np

from . import _hip
from ._hip import check, lib, ptr
from .config import ConfigError, SimulationConfig
from .dynamics import DynamicsEngine
from .microscope import FLIMFrame, VirtualMicroscope
from .particles import ParticleState
from .photophysics import PhotonSimulator

__all__ = ["GroundTruth", "SubFrameData", "FrameResult", "SimulationResult", "SubFrameRecorder",
           "census_to_subframe", "extract_ground_truth", "SimulationPipeline"]


@dataclass
class GroundTruth:
    """Exact physical state of one frame (reference pipeline.py:78-128)."""
    positions: np.ndarray
    species_id: np.ndarray
    bond_partner: np.ndarray
    is_alive: np.ndarray
    has_dye: np.ndarray
    dye_type_id: np.ndarray
    is_bleached: np.ndarray
    fret_efficiency_exact: np.ndarray
    donor_lifetime_exact: np.ndarray
    n_complexes: int = 0
    n_bleached: int = 0
    n_free_donors: int = 0
    species_counts: Optional[Dict[str, int]] = None


@dataclass
class SubFrameData:
    """Per-BD-step series (reference pipeline.py:131-156)."""
    n_complexes: np.ndarray
    n_bindings: np.ndarray
    n_dissociations: np.ndarray
    n_bleached: np.ndarray
    total_photons: np.ndarray
    mean_fret_E: np.ndarray
    dt_s: float
    steps_per_frame: int
    species_counts: Optional[Dict[str, np.ndarray]] = None
    n_catalytic: Optional[np.ndarray] = None
    n_recruited: Optional[np.ndarray] = None
    n_detached: Optional[np.ndarray] = None
    n_converted: Optional[np.ndarray] = None


@dataclass
class FrameResult:
    """One FLIM frame and its ground truth (reference pipeline.py:159-175)."""
    frame_index: int
    bio_time: float
    flim_frame: FLIMFrame
    ground_truth: GroundTruth
    n_photons_emitted: int = 0
    n_photons_detected: int = 0
    wall_time_s: float = 0.0

    @property
    def total_counts(self) -> int:
        return self.flim_frame.total_counts


@dataclass
class SimulationResult:
    """Complete simulation output (reference pipeline.py:178-231)."""
    config: SimulationConfig
    frames: List[FrameResult]
    total_wall_time_s: float = 0.0
    peak_ram_mb: float = 0.0
    total_photons_emitted: int = 0
    total_photons_detected: int = 0
    total_bindings: int = 0
    total_dissociations: int = 0
    total_bleached: int = 0
    subframe_data: Optional[SubFrameData] = None

    @property
    def n_frames(self) -> int:
        return len(self.frames)

    def get_flim_stack(self, frame_idx: int) -> np.ndarray:
        return self.frames[frame_idx].flim_frame.flim_stack

    def get_intensity_series(self) -> np.ndarray:
        return np.array([f.total_counts for f in self.frames])

    def get_complex_count_series(self) -> np.ndarray:
        return np.array([f.ground_truth.n_complexes for f in self.frames])

    def get_bleaching_series(self) -> np.ndarray:
        return np.array([f.ground_truth.n_bleached for f in self.frames])

    def summary(self) -> str:
        return "\n".join([
            "ED-MRDT Simulation Result",
            f"  Config: {self.config.name}",
            f"  Frames: {self.n_frames}",
            f"  Wall time: {self.total_wall_time_s:.1f} s ({self.total_wall_time_s / 60:.1f} min)",
            f"  Peak RAM: {self.peak_ram_mb:.1f} MB",
            f"  Photons emitted: {self.total_photons_emitted:,}",
            f"  Photons detected: {self.total_photons_detected:,}",
            f"  Bindings: {self.total_bindings:,}",
            f"  Dissociations: {self.total_dissociations:,}",
            f"  Final bleached: {self.total_bleached}",
        ])


def _donor_index(config: SimulationConfig, name: str) -> Optional[int]:
    """First fluorophore of that name (pipeline.py:386-390)."""
    for fi, f in enumerate(config.fluorophores):
        if f.name == name:
            return fi
    return None


def _photon_counters(h) -> np.ndarray:
    out = np.zeros(2, dtype=np.int64)
    check(lib().edm_photon_counters(h, ptr(out)), "edm_photon_counters")
    return out


def _event_counters(h) -> np.ndarray:
    out = np.zeros(6, dtype=np.int64)
    check(lib().edm_counters(h, ptr(out)), "edm_counters")
    return out


class SubFrameRecorder:
    """The per-BD-step recording of SimulationPipeline.run
    (pipeline.py:536-566, 567-642) on the device.

    ``record()`` after a step's dynamics + photon step appends one census row
    to the device ring (no host synchronisation); with ``auto=True`` the
    library appends a row after every step of ``edm_run_frame`` /
    ``edm_run`` itself.  Rows are converted when the ring is read out (when
    full, or by ``data()``)."""

    def __init__(self, config: SimulationConfig, dynamics: DynamicsEngine,
                 photon_sim: Optional[PhotonSimulator] = None, capacity: int = 4096,
                 auto: bool = False):
        self.config = config
        self.dynamics = dynamics
        self._h = dynamics.handle
        fp = config.fret_pairs[0] if config.fret_pairs else None
        self._mode = getattr(fp, "fret_mode", "bond") if fp is not None else None
        flags = _hip.REC_AUTO if auto else 0
        active = photon_sim is not None and getattr(photon_sim, "_active", False)
        if self._mode == "proximity" and active:
            flags |= _hip.REC_FRET_SUM
        self._fret_sum = bool(flags & _hip.REC_FRET_SUM)
        self.capacity = int(capacity)
        self.auto = auto
        check(lib().edm_record_open(self._h, self.capacity, flags), "edm_record_open")
        # cumulative counters when recording starts (pipeline.py:547-548, 562-565)
        self._prev_events = _event_counters(self._h)
        self._prev_photons = int(_photon_counters(self._h)[0])
        self._chunks: List[np.ndarray] = []
        self._pending = 0

    @property
    def room(self) -> int:
        """Rows the ring can still take before it must be read out."""
        return self.capacity - self._pending

    def record(self) -> None:
        if self._pending >= self.capacity:
            self.drain()
        check(lib().edm_record_step(self._h), "edm_record_step")
        self._pending += 1

    def note_auto_rows(self, n: int) -> None:
        """The library appended ``n`` rows by itself (EDM_REC_AUTO)."""
        self._pending += int(n)

    def drain(self) -> None:
        n = ctypes.c_int32(0)
        check(lib().edm_record_rows(self._h, ctypes.byref(n)), "edm_record_rows")
        rows = np.zeros(int(n.value), dtype=_hip.CENSUS_DTYPE)
        got = ctypes.c_int32(0)
        check(lib().edm_record_read(self._h, ptr(rows), rows.size, ctypes.byref(got)),
              "edm_record_read")
        if got.value:
            self._chunks.append(rows[:got.value])
        self._pending = 0

    def close(self) -> None:
        if self._h is not None:
            check(lib().edm_record_close(self._h), "edm_record_close")
            self._h = None

    def data(self, steps_per_frame: int) -> SubFrameData:
        """All rows so far as the reference's SubFrameData (pipeline.py:700-717)."""
        if self._h is not None:
            self.drain()
        rows = (np.concatenate(self._chunks) if self._chunks
                else np.zeros(0, dtype=_hip.CENSUS_DTYPE))
        return census_to_subframe(self.config, rows, self._prev_events, self._prev_photons,
                                  self.dynamics.dt, steps_per_frame, self._fret_sum)


def census_to_subframe(config: SimulationConfig, rows: np.ndarray, prev_events,
                       prev_photons: int, dt: float, steps_per_frame: int,
                       fret_sum: bool) -> SubFrameData:
    """SubFrameData (pipeline.py:700-717) from consecutive census rows (one
    per step): event and photon counts are differences of the cumulative
    counters, starting from ``prev_events`` / ``prev_photons`` (their values
    when recording began, pipeline.py:547-548, 562-565); complexes are bonded
    alive particles / 2 (:570); the mean FRET efficiency follows :588-617."""
    prev_events = np.asarray(prev_events, dtype=np.int64)
    ev = rows["events"]
    prev = np.vstack([prev_events[None, :], ev[:-1]]) if len(rows) else ev
    d_ev = (ev - prev).astype(np.int32)
    ph = rows["photons"]
    prev_ph = np.concatenate([[int(prev_photons)], ph[:-1]]) if len(rows) else ph
    n_bound = rows["n_bond_alive"] // 2
    n_tot = rows["n_donor0"]
    mean_E = np.zeros(len(rows), dtype=np.float64)
    if config.fret_pairs:
        fp = config.fret_pairs[0]
        mode = getattr(fp, "fret_mode", "bond")
        if mode != "proximity":
            E_const = 1.0 / (1.0 + (fp.intra_complex_distance / fp.R0) ** 6)
        for k in np.flatnonzero(n_tot > 0):
            if mode == "proximity":
                # np.mean of the donors' E (:600-606)
                mean_E[k] = rows["e_sum"][k] / int(n_tot[k]) if fret_sum else 0.0
            else:
                nb_, nt = int(n_bound[k]), int(n_tot[k])
                mean_E[k] = E_const * nb_ / (nb_ + (nt - nb_))
    sp = rows["species_alive"]
    return SubFrameData(
        n_complexes=n_bound.astype(np.int32),
        n_bindings=d_ev[:, 0].copy(),
        n_dissociations=d_ev[:, 1].copy(),
        n_bleached=rows["n_bleached"].astype(np.int32),
        total_photons=(ph - prev_ph).astype(np.int32),
        mean_fret_E=mean_E,
        dt_s=dt,
        steps_per_frame=int(steps_per_frame),
        species_counts={s.name: sp[:, i].astype(np.int32) for i, s in enumerate(config.species)},
        n_catalytic=d_ev[:, 2].copy(),
        n_recruited=d_ev[:, 3].copy(),
        n_detached=d_ev[:, 4].copy(),
        n_converted=d_ev[:, 5].copy(),
    )


def extract_ground_truth(config: SimulationConfig, state: ParticleState,
                         photon_sim: Optional[PhotonSimulator]) -> GroundTruth:
    """_extract_ground_truth (pipeline.py:357-455): FRET rates updated, exact
    E and tau_DA per donor and the frame's counts from one device pass."""
    dev = state._device
    if dev is None:
        raise RuntimeError("extract_ground_truth needs the state on the GPU")
    donors, taus = [], []
    rms = getattr(photon_sim, "rate_matrices", []) if photon_sim is not None else []
    for fp in config.fret_pairs:
        d = _donor_index(config, fp.donor_type)
        if d is None or d >= len(rms):
            continue
        rm = rms[d]
        tD = rm.get_lifetime(rm.excited_state)
        if tD <= 0 or math.isinf(tD):
            continue
        donors.append(d)
        taus.append(tD)
    if len(donors) > 8:
        raise ConfigError("the GPU path supports at most 8 FRET pairs")
    N = state.N
    E = np.zeros(N)
    tau = np.zeros(N)
    dn = np.asarray(donors + [0], dtype=np.int32)
    tn = np.asarray(taus + [0.0], dtype=np.float64)
    cen = _hip.Census()
    check(lib().edm_ground_truth(dev.handle, len(donors), ptr(dn), ptr(tn), ptr(E), ptr(tau),
                                 ctypes.byref(cen)), "edm_ground_truth")
    n_free = 0
    for fp in config.fret_pairs:
        for fi, f in enumerate(config.fluorophores):
            if f.name == fp.donor_type and fi < 8:
                n_free += int(cen.n_free_dye[fi])
    return GroundTruth(
        positions=state.positions.copy(),
        species_id=state.species_id.copy(),
        bond_partner=state.bond_partner.copy(),
        is_alive=state.is_alive.copy(),
        has_dye=state.has_dye.copy(),
        dye_type_id=state.dye_type_id.copy(),
        is_bleached=state.is_bleached.copy(),
        fret_efficiency_exact=E,
        donor_lifetime_exact=tau,
        n_complexes=int(cen.n_bond) // 2,
        n_bleached=int(cen.n_bleached),
        n_free_donors=n_free,
        species_counts={s.name: int(cen.species_alive[i]) for i, s in enumerate(config.species)},
    )


class SimulationPipeline:
    """Orchestrator BD -> photophysics -> microscope -> FLIM frames
    (reference pipeline.py:234-781) on the GPU engines.

    Extra keyword arguments select the engines' precision / draws
    (``DynamicsEngine``); with ``rng="philox"`` each frame's BD steps run as
    one device loop (``edm_run_frame``), recording included (``device_loop``
    False: step by step from the host, with the same results)."""

    def __init__(self, config: SimulationConfig, progress_callback: Optional[Callable] = None,
                 record_subframe: bool = False, precision: Optional[str] = None,
                 rng: Optional[str] = None, pair_cap: bool = True,
                 exact_force_order: bool = True, record_capacity: int = 65536,
                 device_loop: Optional[bool] = None):
        self.config = config
        self.record_subframe = record_subframe
        dt = config.bd_timestep
        frame_interval = config.scan.frame_interval
        n_frames = getattr(config, "n_frames", config.scan.n_frames)
        if frame_interval <= 0:
            raise ConfigError("frame_interval must be positive")
        if dt <= 0:
            raise ConfigError("bd_timestep must be positive")
        self.n_bd_per_frame = int(round(frame_interval / dt))
        if self.n_bd_per_frame < 1:
            raise ConfigError(f"frame_interval ({frame_interval}s) < bd_timestep ({dt}s)")
        actual = self.n_bd_per_frame * dt
        drift = abs(actual - frame_interval) / frame_interval
        if drift > 0.01:
            raise ConfigError(f"frame_interval ({frame_interval}s) is not an integer multiple of "
                              f"bd_timestep ({dt}s). Drift = {drift:.2%}")
        self.n_frames = n_frames
        self.dt = dt
        self.frame_interval = actual
        self._progress = progress_callback or self._default_progress
        self._engine_kw = dict(precision=precision, rng=rng, pair_cap=pair_cap,
                               exact_force_order=exact_force_order)
        self._record_capacity = int(record_capacity)
        self._device_loop = device_loop
        self._state: Optional[ParticleState] = None
        self._dynamics: Optional[DynamicsEngine] = None
        self._photon_sim: Optional[PhotonSimulator] = None
        self._microscope: Optional[VirtualMicroscope] = None
        self._recorder: Optional[SubFrameRecorder] = None

    def _init_components(self) -> None:
        """pipeline.py:317-353."""
        config = self.config
        self._state = ParticleState(config)
        self._dynamics = DynamicsEngine(config, self._state, **self._engine_kw)
        duty = min(1.0, config.scan.dwell_time / config.scan.frame_interval)
        self._photon_sim = PhotonSimulator(config, self._state, scan_duty_fraction=duty)
        self._microscope = VirtualMicroscope(config, self._state)

    def _extract_ground_truth(self) -> GroundTruth:
        return extract_ground_truth(self.config, self._state, self._photon_sim)

    @staticmethod
    def _get_ram_mb() -> float:
        try:
            import psutil
            return psutil.Process().memory_info().rss / 1e6
        except ImportError:
            try:
                with open("/proc/self/status") as f:
                    for line in f:
                        if line.startswith("VmRSS:"):
                            return float(line.split()[1]) / 1024.0
            except (FileNotFoundError, ValueError):
                return 0.0
        return 0.0

    @staticmethod
    def _default_progress(frame_idx: int, n_frames: int, msg: str) -> None:
        print(f"  [{frame_idx + 1}/{n_frames}] {msg}")

    # ---- the BD step loop of one frame (pipeline.py:556-642) ---------------
    def _frame_steps_host(self):
        """Step by step, as the reference: the parity path (numpy draws)."""
        emitted = detected = 0
        for _ in range(self.n_bd_per_frame):
            self._dynamics.step()
            batch = self._photon_sim.evolve_one_bd_step()
            emitted += batch.n_photons
            detected += self._microscope.process_photons(batch)
            if self._recorder is not None:
                self._recorder.record()
        return emitted, detected

    def _frame_steps_device(self):
        """Philox draws: the frame's steps as device loops with no host
        synchronisation, each at most as long as the recording ring has room."""
        h = self._dynamics.handle
        ne0, nd0 = _photon_counters(h)
        left = self.n_bd_per_frame
        rec = self._recorder
        while left > 0:
            k = left
            if rec is not None:
                if rec.room == 0:
                    rec.drain()
                k = min(k, rec.room)
            check(lib().edm_run_frame(h, int(k)), "edm_run_frame")
            if rec is not None:
                rec.note_auto_rows(k)
            left -= k
        dyn = self._dynamics
        dyn.n_steps_done += self.n_bd_per_frame
        dyn._ctr_cache = None
        self._state.mark_stale()
        ne1, nd1 = _photon_counters(h)
        emitted, detected = int(ne1 - ne0), int(nd1 - nd0)
        ps, mic = self._photon_sim, self._microscope
        if getattr(ps, "_active", False):
            ps._fused_prev = (int(ne1), int(nd1))
        mic.total_photons_received += emitted
        mic.total_photons_detected += detected
        mic._n_signal += detected
        mic._n_bd_steps_in_frame += self.n_bd_per_frame
        return emitted, detected

    def run(self) -> SimulationResult:
        """pipeline.py:489-738."""
        t0 = time.perf_counter()
        peak_ram = self._get_ram_mb()
        self._init_components()
        init_ram = self._get_ram_mb()
        peak_ram = max(peak_ram, init_ram)
        self._progress(-1, self.n_frames,
                       f"Initialized: {self._state.N} particles, {self.n_bd_per_frame} BD steps/frame, "
                       f"{self.n_frames} frames. RAM: {init_ram:.0f} MB")
        device_loop = self._dynamics.rng_mode == "philox"
        if self._device_loop is not None:
            if self._device_loop and not device_loop:
                raise ValueError("device_loop needs rng='philox' (numpy draws come from the host)")
            device_loop = self._device_loop
        if self.record_subframe:
            cap = min(self._record_capacity, max(1, self.n_frames * self.n_bd_per_frame))
            self._recorder = SubFrameRecorder(self.config, self._dynamics, self._photon_sim,
                                              capacity=cap, auto=device_loop)
        frames: List[FrameResult] = []
        tot_e = tot_d = 0
        for frame_i in range(self.n_frames):
            tf = time.perf_counter()
            if device_loop:
                fe, fd = self._frame_steps_device()
            else:
                fe, fd = self._frame_steps_host()
            flim_frame = self._microscope.finalize_frame()
            gt = self._extract_ground_truth()
            bio_time = (frame_i + 1) * self.frame_interval
            wall = time.perf_counter() - tf
            frames.append(FrameResult(frame_index=frame_i, bio_time=bio_time, flim_frame=flim_frame,
                                      ground_truth=gt, n_photons_emitted=fe,
                                      n_photons_detected=fd, wall_time_s=wall))
            tot_e += fe
            tot_d += fd
            cur_ram = self._get_ram_mb()
            peak_ram = max(peak_ram, cur_ram)
            elapsed = time.perf_counter() - t0
            eta = elapsed / (frame_i + 1) * (self.n_frames - frame_i - 1)
            self._progress(frame_i, self.n_frames,
                           f"t={bio_time:.1f}s | photons={fd:,} | counts={flim_frame.total_counts:,} | "
                           f"complexes={gt.n_complexes} | bleached={gt.n_bleached} | "
                           f"wall={wall:.1f}s | RAM={cur_ram:.0f}MB | ETA={eta:.0f}s")
            if cur_ram > self.config.ram_limit_gb * 1000:
                self._progress(frame_i, self.n_frames,
                               f"RAM limit exceeded ({cur_ram:.0f} MB > "
                               f"{self.config.ram_limit_gb * 1000:.0f} MB). Stopping.")
                break
        subframe = None
        if self._recorder is not None:
            subframe = self._recorder.data(self.n_bd_per_frame)
            self._recorder.close()
            self._recorder = None
        result = SimulationResult(
            config=self.config, frames=frames, total_wall_time_s=time.perf_counter() - t0,
            peak_ram_mb=peak_ram, total_photons_emitted=tot_e, total_photons_detected=tot_d,
            total_bindings=self._dynamics.n_bindings,
            total_dissociations=self._dynamics.n_dissociations,
            total_bleached=int(np.sum(self._state.is_bleached)), subframe_data=subframe)
        self._progress(self.n_frames - 1, self.n_frames, f"DONE. {result.summary()}")
        return result

    @classmethod
    def run_mini_demo(cls, config: SimulationConfig, n_frames: int = 2,
                      n_bd_steps_per_frame: Optional[int] = None, **kw) -> SimulationResult:
        """pipeline.py:742-767."""
        config.scan.n_frames = n_frames
        if n_bd_steps_per_frame is not None:
            config.scan.frame_interval = n_bd_steps_per_frame * config.bd_timestep
        return cls(config, **kw).run()

    def __repr__(self) -> str:
        total = self.n_bd_per_frame * self.n_frames
        return (f"SimulationPipeline(\n  config: {self.config.name}\n  frames: {self.n_frames}\n"
                f"  BD steps/frame: {self.n_bd_per_frame:,}\n  total BD steps: {total:,}\n"
                f"  bio time: {self.n_frames * self.frame_interval:.1f} s\n  dt: {self.dt:.1e} s\n)")
