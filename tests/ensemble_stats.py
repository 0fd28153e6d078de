"""Long-run ensemble statistics shared by the oracle fixture script
(tests/golden/make_ensemble_stats.py) and the GPU test
(tests/test_gpu_statistics.py).

Each case is a system run for STEPS steps by REPLICAS independent replicas
(seeds 0..REPLICAS-1; the initial state of a seed is the same on both sides).
Observables of one replica, in this order:

* MSD (um^2) over the alive particles at each checkpoint, periodic jumps
  unwrapped step by step (validation.py:276-290)
* alive particles per species at each count checkpoint
* unordered pairs of alive particles per distance bin (uniform bins up to
  r_max) at the end: the RDF's counts (analysis.py:733-790)
* with imaging, from the donor channel's signal photons summed over every
  frame (the stacks before the frame-end noise): phasor g, s (harmonic 1)
  and the mean arrival time in ns (analysis.py:87-127, 497-566)
"""

from __future__ import annotations

import math

import numpy as np

from ed_mrdt import workloads

REPLICAS = 32
STEPS = 1000
MSD_AT = (100, 250, 500, 1000)
COUNT_AT = (250, 500, 1000)

CASES = {
    # name: (config of a seed, imaging, pair bins, r_max um)
    "binding": (lambda s: workloads.binding_fret(seed=s), True, 8, 0.4),
    "wave": (lambda s: workloads.wave_system(seed=s, n_steps=STEPS), True, 8, 0.2),
    "dense_harmonic": (lambda s: workloads.large_membrane(seed=s, side=14.0, n_alive=19_600,
                                                          pixels=64), False, 10, 0.05),
}


def observable_names(case):
    make, imaging, nb, r_max = CASES[case]
    cfg = make(0)
    names = [f"msd@{t}" for t in MSD_AT]
    names += [f"n[{s.name}]@{t}" for t in COUNT_AT for s in cfg.species]
    dr = r_max / nb
    names += [f"pairs[{k * dr:.3f},{(k + 1) * dr:.3f})" for k in range(nb)]
    if imaging:
        names += ["phasor g", "phasor s", "mean arrival ns"]
    return names


def frame_every(cfg) -> int:
    return max(1, int(round(cfg.scan.frame_interval / cfg.bd_timestep)))


def decay_moments(stack_ch: np.ndarray, bin_ns: float) -> np.ndarray:
    """[sum I, sum I cos phi, sum I sin phi, sum I t] of one channel's stack
    [nx, ny, n_bins], summed over pixels."""
    h = stack_ch.reshape(-1, stack_ch.shape[-1]).sum(axis=0).astype(np.float64)
    k = np.arange(h.size)
    phi = 2.0 * math.pi * (k + 0.5) / h.size
    return np.array([h.sum(), np.sum(h * np.cos(phi)), np.sum(h * np.sin(phi)),
                     np.sum(h * (k + 0.5) * bin_ns)])


def phasor_from(m: np.ndarray):
    return [m[1] / m[0], m[2] / m[0], m[3] / m[0]] if m[0] > 0 else [0.0, 0.0, 0.0]


def oracle_replica(case, seed):
    """The oracle's observables of one replica (FP64, the reference's draws)."""
    import reference_model as rm
    from scipy.spatial import cKDTree

    make, imaging, nb, r_max = CASES[case]
    cfg = make(seed)
    st = rm.OracleState(cfg)
    dyn = rm.OracleDynamics(cfg, st)
    if imaging:
        ph = rm.OraclePhotons(cfg, st, min(1.0, cfg.scan.dwell_time / cfg.scan.frame_interval))
        mic = rm.OracleMicroscope(cfg, st)
        fe = frame_every(cfg)
        mom = np.zeros(4)
    L = np.asarray(cfg.membrane.size, dtype=np.float64)
    prev = st.positions.copy()
    acc = np.zeros_like(prev)
    msd, counts = [], []
    n_sp = len(cfg.species)
    for k in range(1, STEPS + 1):
        dyn.step()
        if imaging:
            mic.process_photons(ph.evolve_one_bd_step())
        d = st.positions - prev
        d -= L * np.round(d / L)
        acc += d
        prev = st.positions.copy()
        if imaging and (k % fe == 0 or k == STEPS):
            mom += decay_moments(mic.flim[0], mic.bin_w * 1e9)
            mic.finalize_frame()
        alive = st.is_alive.astype(bool)
        if k in MSD_AT:
            msd.append(float(np.mean(np.sum(acc[alive] ** 2, axis=1))))
        if k in COUNT_AT:
            counts += np.bincount(st.species_id[alive], minlength=n_sp).tolist()
    alive = st.is_alive.astype(bool)
    tree = cKDTree(np.mod(st.positions[alive], L), boxsize=L)
    edges = np.linspace(0.0, r_max, nb + 1)
    pairs = np.diff(tree.count_neighbors(tree, edges).astype(np.float64)) / 2.0
    out = msd + counts + pairs.tolist()
    if imaging:
        out += phasor_from(mom)
    return np.array(out, dtype=np.float64)
