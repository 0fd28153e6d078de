"""Benchmark of the per-timestep particle path on one MI355X.

Workload: the 1,000,000-particle membrane (100 x 100 um periodic box, FP32
positions, harmonic repulsion k = 1e7 kT/um^2 with r_c = 10 nm, the plain
PIP3/PTEN reaction set with Hill feedback on a 200 x 200 density grid,
PSS photon emission from PIP3-borne mTFP1, 1024 x 1024 px x 256-bin TCSPC
imaging with frames finalised on the device every 100 steps).

One step = cell binning + forces + Ermak-McCammon move + all reactions +
photon emission, detection, PSF pixel draw and TCSPC binning.  The state
is resident in HBM before the timed region starts.  ``value`` counts alive
particles per step (particle-steps/s).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]

Prints one JSON line.  The CPU baseline is the oracle (a C restatement of
the reference hot path, bd_step threaded like the reference's prange) timed
on a bounded sample of the same workload on this machine's host cores.

``recording``: the same steps with the per-step census of the pipeline's
sub-frame recording appended on the device, against the reference's numpy
recording of the 1M state on the host.
"""

from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "flim-microscopy-simulator-with-causal-inference_amd")
sys.path.insert(0, PKG)

METRIC = "particle-steps/s @1M; ensemble jobs/s + p99 latency, batched vs plain"
HBM_PEAK_GBPS = 8000.0
# algorithmic bytes of one force/BD launch per particle (FP32): read x, y
# (8), species + flags (2), particle index (4); write x', y' (8)
FORCE_BYTES_PER_PARTICLE = 22


def copy_probe(mib: int) -> int:
    """One device-to-device copy of ``mib`` MiB before the run: the kernel
    of known traffic (read mib, write mib) that scales the FETCH_SIZE and
    WRITE_SIZE counters of a rocprofv3 --pmc pass."""
    hip = ctypes.CDLL("libamdhip64.so")
    n = mib << 20
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    for p in (src, dst):
        if hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(n)) != 0:
            raise RuntimeError("hipMalloc failed in the copy probe")
    hip.hipMemset(src, 1, ctypes.c_size_t(n))
    hip.hipDeviceSynchronize()
    if hip.hipMemcpy(dst, src, ctypes.c_size_t(n), 3) != 0:  # device to device
        raise RuntimeError("hipMemcpy failed in the copy probe")
    hip.hipDeviceSynchronize()
    hip.hipFree(src)
    hip.hipFree(dst)
    return n


def traffic_from_profile():
    """HBM bytes per force launch measured by rocprofv3 --pmc FETCH_SIZE and
    WRITE_SIZE (scaled by the same counters on a copy of known size), as
    recorded in profiles/k_force_tile_1M_counters.json; None if absent."""
    path = os.path.join(ROOT, "profiles", "k_force_tile_1M_counters.json")
    if not os.path.exists(path):
        return None
    with open(path, "r", encoding="utf-8") as f:
        c = json.load(f)
    return c["fetch_kib"] * 1024.0 * c["fetch_scale"] + c["write_kib"] * 1024.0 * c["write_scale"]


def gpu_run(steps: int, warmup: int, n_alive: int, rec_steps: int = 0):
    import numpy as np
    from ed_mrdt import _hip, workloads
    from ed_mrdt.dynamics import DynamicsEngine
    from ed_mrdt.microscope import VirtualMicroscope
    from ed_mrdt.particles import ParticleState
    from ed_mrdt.photophysics import PhotonSimulator

    side = 100.0 * (n_alive / 1e6) ** 0.5
    cfg = workloads.large_membrane(side=side, n_alive=n_alive, pixels=1024)
    st = ParticleState(cfg)
    eng = DynamicsEngine(cfg, st, precision="fp32", rng="philox", pair_cap=False)
    duty = min(1.0, cfg.scan.dwell_time / cfg.scan.frame_interval)
    PhotonSimulator(cfg, st, duty)
    mic = VirtualMicroscope(cfg, st)
    L = _hip.lib()
    h = eng.handle
    frame_every = int(round(cfg.scan.frame_interval / cfg.bd_timestep))
    tot, fms = ctypes.c_double(0), ctypes.c_double(0)

    def run(k):
        """k steps in chunks ending on frame boundaries; returns (ms, force_ms)."""
        done = 0
        ms = fm = 0.0
        while done < k:
            chunk = min(frame_every, k - done)
            _hip.check(L.edm_time_steps(h, chunk, 1, ctypes.byref(tot), ctypes.byref(fms), None))
            ms += tot.value
            fm += fms.value
            done += chunk
            if done % frame_every == 0:
                t0 = time.perf_counter()
                _hip.check(L.edm_frame_noise(h, None, None))
                _hip.check(L.edm_frame_reset(h))
                _hip.check(L.edm_synchronize(h))
                ms += (time.perf_counter() - t0) * 1e3
        return ms, fm

    if warmup:
        run(warmup)
    ph0 = np.zeros(2, np.int64)
    _hip.check(L.edm_photon_counters(h, _hip.ptr(ph0)))
    st.mark_stale()
    alive0 = int(np.sum(st.is_alive))
    _hip.check(L.edm_synchronize(h))
    t0 = time.perf_counter()
    ms, force_ms = run(steps)
    _hip.check(L.edm_synchronize(h))
    wall = time.perf_counter() - t0
    st.mark_stale()
    alive1 = int(np.sum(st.is_alive))
    ph1 = np.zeros(2, np.int64)
    _hip.check(L.edm_photon_counters(h, _hip.ptr(ph1)))
    alive = 0.5 * (alive0 + alive1)
    out = dict(cfg=cfg, N=st.N, alive=alive, wall=wall, event_ms=ms, force_ms=force_ms,
               steps=steps, emitted=int(ph1[0] - ph0[0]), detected=int(ph1[1] - ph0[1]))
    if rec_steps:
        out["recording"] = recording_leg(L, h, run, rec_steps)
    return out


def recording_leg(L, h, run, k: int):
    """The same steps with the per-step census appended on the device
    (SimulationPipeline's sub-frame recording, EDM_REC_AUTO), against the
    same number of steps without it, both timed with HIP events."""
    from ed_mrdt import _hip
    import numpy as np

    ms_plain, _ = run(k)
    _hip.check(L.edm_record_open(h, k, _hip.REC_AUTO))
    ms_rec, _ = run(k)
    rows = np.zeros(k, dtype=_hip.CENSUS_DTYPE)
    n = ctypes.c_int32(0)
    _hip.check(L.edm_record_read(h, _hip.ptr(rows), k, ctypes.byref(n)))
    _hip.check(L.edm_record_close(h))
    if n.value != k:
        raise RuntimeError(f"recording leg: {n.value} census rows for {k} steps")
    return {"steps": k, "ms_per_step_plain": ms_plain / k, "ms_per_step_recording": ms_rec / k,
            "overhead_frac": ms_rec / ms_plain - 1.0,
            "last_row_alive": int(rows["species_alive"][-1].sum())}


def _progress(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def ensemble_bench(n_rep: int, steps: int, plain_every: int):
    """Batched ensemble of the 512-replica workload (box sides U[5,30] um,
    rates jittered, seeds 0..511, 64 x 64 px FLIM every 50 steps) against the
    plain path (one replica per launch sequence) on every plain_every-th
    replica.  Both run every step with imaging, end every frame on the
    device and copy each replica's final frame to the host."""
    import numpy as np
    from ed_mrdt import workloads
    from ed_mrdt.ensemble import EnsembleEngine
    from ed_mrdt.particles import ParticleState
    from ed_mrdt.serving import Job, serve_plain

    cfgs = [workloads.ensemble_replica(r) for r in range(n_rep)]
    states = [ParticleState(c) for c in cfgs]
    alive = sum(int(np.sum(s.is_alive)) for s in states)
    cap = sum(s.N for s in states) + 1024
    eng = EnsembleEngine(cfgs[0], cap, max_replicas=n_rep)
    slots = [eng.add(c, s) for c, s in zip(cfgs, states)]
    _progress(f"ensemble: {n_rep} replicas, {cap - 1024} particles admitted")
    fe = max(1, int(round(cfgs[0].scan.frame_interval / cfgs[0].bd_timestep)))
    eng.run(1)  # first step: layout, buffers
    eng.synchronize()
    t0 = time.perf_counter()
    done = 1
    while done < steps:
        k = min(fe - done % fe, steps - done)
        eng.run(k)
        done += k
        if done < steps and done % fe == 0:
            eng.end_frames(slots)
    counts = 0
    for sl in slots:
        counts += eng.frame(sl, download=True).total_counts
    eng.synchronize()
    t_b = time.perf_counter() - t0
    eng.close()
    _progress(f"ensemble: batched {steps} steps in {t_b:.2f} s")
    b_jobs = n_rep / t_b * (steps - 1) / steps  # the timed region ran steps - 1 steps
    if plain_every <= 0:  # no plain sample (profiling the batched path)
        return {"replicas": n_rep, "steps": steps, "particles": cap - 1024, "alive": alive,
                "batched_s": t_b, "batched_jobs_per_s": b_jobs,
                "batched_particle_steps_per_s": alive * (steps - 1) / t_b,
                "final_frame_counts": int(counts)}
    sample = [Job(i, c, steps) for i, c in enumerate(cfgs[::plain_every])]
    t1 = time.perf_counter()
    serve_plain(sample, realtime=False, progress=_progress)
    t_p = time.perf_counter() - t1
    _progress(f"ensemble: plain {len(sample)} replicas in {t_p:.2f} s")
    p_jobs = len(sample) / t_p
    return {"replicas": n_rep, "steps": steps, "particles": cap - 1024, "alive": alive,
            "batched_s": t_b, "batched_jobs_per_s": b_jobs,
            "batched_particle_steps_per_s": alive * (steps - 1) / t_b,
            "plain_sample_replicas": len(sample), "plain_s": t_p, "plain_jobs_per_s": p_jobs,
            "batched_over_plain": b_jobs / p_jobs, "final_frame_counts": int(counts)}


def served_bench(n_jobs: int, rate: float, budget_gib: float):
    """Served stream: Poisson arrivals at `rate` jobs/s of the first n_jobs
    jobs of the served workload (box side logU[5,50] um, 1k-250k particles
    clamped to 100 / um^2, 500-5000 steps, 64-256 px), continuous batching
    under an HBM budget vs the same arrivals served one job at a time."""
    import numpy as np
    from ed_mrdt import workloads
    from ed_mrdt.serving import Job, Server, latency_stats, n_particles, serve_plain

    g = np.random.default_rng(2026)
    arr = np.concatenate([[0.0], np.cumsum(g.exponential(1.0 / rate, n_jobs - 1))])
    jobs = []
    for j in range(n_jobs):
        cfg, steps = workloads.served_job(j)
        jobs.append(Job(j, cfg, steps, arrival=float(arr[j])))
    cap = sum(n_particles(jb.config) for jb in jobs) + 1024
    srv = Server(jobs[0].config, int(budget_gib * (1 << 30)), cap, max_jobs=n_jobs, chunk=100)
    res = srv.serve(jobs, realtime=True, progress=_progress)
    srv.engine.close()
    b = latency_stats(res)
    _progress(f"served: batched {n_jobs} jobs, {b['jobs_per_s']:.2f} jobs/s")
    p = latency_stats(serve_plain(jobs, realtime=True, progress=_progress))
    _progress(f"served: plain {n_jobs} jobs, {p['jobs_per_s']:.2f} jobs/s")
    work = sum(n_particles(jb.config) * jb.steps for jb in jobs)
    return {"jobs": n_jobs, "arrival_rate_per_s": rate, "hbm_budget_gib": budget_gib,
            "particle_steps": work, "batched": b, "plain": p,
            "peak_resident_jobs": srv.peak_resident,
            "batched_over_plain": b["jobs_per_s"] / p["jobs_per_s"] if p["jobs_per_s"] else None}


def cpu_baseline(n_alive: int, budget_s: float = 20.0):
    """Oracle steps (dynamics + photophysics) on the 1M workload until the
    budget is spent; returns alive particle-steps/s and the sample size."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import reference_model as rm
    from ed_mrdt import workloads

    side = 100.0 * (n_alive / 1e6) ** 0.5
    cfg = workloads.large_membrane(side=side, n_alive=n_alive, pixels=1024)
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    threads = max(1, min(threads, 16))
    st = rm.OracleState(cfg)
    dyn = rm.OracleDynamics(cfg, st, pair_cap=False, n_threads=threads)
    duty = min(1.0, cfg.scan.dwell_time / cfg.scan.frame_interval)
    ph = rm.OraclePhotons(cfg, st, duty)
    dyn.step()
    ph.evolve_one_bd_step()
    steps = 0
    ps = 0.0
    t0 = time.perf_counter()
    while True:
        a = int(np.sum(st.is_alive))
        dyn.step()
        ph.evolve_one_bd_step()
        steps += 1
        ps += a
        el = time.perf_counter() - t0
        if el >= budget_s or steps >= 200:
            break
    # the reference's per-step recording (numpy over the 1M arrays,
    # pipeline.py:567-642) on the same state
    import recording as orec
    sub = orec.OracleSubframe(cfg, dyn, ph)
    t1 = time.perf_counter()
    n_rec = 0
    while n_rec < 20 and time.perf_counter() - t1 < 5.0:
        sub.record(st, 0)
        n_rec += 1
    rec_ms = (time.perf_counter() - t1) * 1e3 / n_rec
    return dict(value=ps / el, steps=steps, seconds=el, threads=threads, record_ms=rec_ms,
                record_n=n_rec)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--ensemble-replicas", type=int, default=512)
    ap.add_argument("--ensemble-steps", type=int, default=5000)
    ap.add_argument("--plain-every", type=int, default=32)
    ap.add_argument("--served-jobs", type=int, default=48)
    ap.add_argument("--no-served", action="store_true")
    ap.add_argument("--ensemble-only", action="store_true",
                    help="only the ensemble (and served) legs, for profiling them")
    ap.add_argument("--record-steps", type=int, default=200,
                    help="steps of the recording leg (census every step); 0: none")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="copy this many MiB device-to-device first (counter calibration)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU only; multi-GPU runs are not "
                                   "part of it (run with --gpus 1)"}))
        sys.exit(2)
    if a.copy_probe_mib:
        copy_probe(a.copy_probe_mib)
    if a.ensemble_only:
        out = {"metric": METRIC, "ensemble": ensemble_bench(a.ensemble_replicas, a.ensemble_steps,
                                                            a.plain_every)}
        if not a.no_served:
            out["served"] = served_bench(a.served_jobs, 50.0, 96.0)
        print(json.dumps(out))
        return
    r = gpu_run(a.steps, a.warmup, a.n, a.record_steps)
    _progress(f"1M membrane: {r['steps']} steps in {r['wall']:.3f} s")
    traffic = traffic_from_profile()
    value = r["alive"] * r["steps"] / r["wall"]
    force_avg_s = r["force_ms"] / r["steps"] * 1e-3
    achieved = FORCE_BYTES_PER_PARTICLE * r["N"] / force_avg_s / 1e9
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": r["wall"] * 1e3 / r["steps"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "1M-particle PIP3/PTEN membrane, 100x100 um, harmonic k=1e7 kT/um^2, "
                               "200x200 density grid, 1024x1024x256 TCSPC frame every 100 steps",
                   "particles": r["N"], "alive_particles": int(r["alive"]),
                   "step": "bin+force+BD+reactions+PSS photons+PSF/TCSPC imaging"},
        "roofline": {"bound": "hbm", "kernel": "k_force_tile<float, true>",
                     "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBPS, "traffic": traffic,
                     "avg_launch_us": force_avg_s * 1e6,
                     "bytes_per_launch": FORCE_BYTES_PER_PARTICLE * r["N"]},
        "photons_per_s": r["emitted"] / max(r["wall"], 1e-9),
        "detected_per_s": r["detected"] / max(r["wall"], 1e-9),
    }
    if "recording" in r:
        out["recording"] = r["recording"]
    if not a.no_ensemble:
        out["ensemble"] = ensemble_bench(a.ensemble_replicas, a.ensemble_steps, a.plain_every)
        if not a.no_served:
            out["served"] = served_bench(a.served_jobs, 50.0, 96.0)
    if not a.no_cpu_baseline:
        c = cpu_baseline(a.n, a.cpu_budget)
        out["cpu_baseline"] = {"value": c["value"], "unit": "particle-steps/s", "cores": c["threads"],
                               "kind": "port",
                               "sample": f"{c['steps']} oracle steps (dynamics + photophysics, "
                                         f"no pair cap) of the same 1M workload in {c['seconds']:.1f} s"}
        out["speedup_vs_cpu"] = value / c["value"]
        if "recording" in out:
            out["recording"]["cpu_numpy_ms_per_step"] = c["record_ms"]
            out["recording"]["cpu_numpy_sample"] = (
                f"{c['record_n']} per-step recordings (numpy reductions of the reference's "
                f"pipeline) of the 1M oracle state")
    print(json.dumps(out))


if __name__ == "__main__":
    main()
