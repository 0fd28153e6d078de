"""Long-run statistics of the fast path (FP32, Philox draws) against the
oracle (FP64, the reference's numpy draws): independent replicas compared
through ensemble means of MSD, pair-distance counts (the RDF's histogram),
species-count trajectories and, with imaging, the phasor of the TCSPC
decay.  Two systems: the binding / FRET system with imaging (24 replicas x
100 steps) and the dense harmonic membrane of the 1M workload at 14 x 14 um
(12 replicas x 60 steps), whose stiff, displacement-capped dynamics make
FP32 and FP64 trajectories decorrelate within a few tens of steps, so that
only distributions can agree.

Bar: every observable's two ensemble means agree within 4.5 combined
standard errors (Welch), i.e. within the ensembles' confidence intervals.
The initial states are the same per seed on both sides (bit-exact
initialisation).
"""

from __future__ import annotations

import numpy as np
import pytest
from scipy.spatial import cKDTree

import reference_model as rm
from ed_mrdt import workloads
from ed_mrdt.dynamics import DynamicsEngine
from ed_mrdt.microscope import VirtualMicroscope
from ed_mrdt.particles import ParticleState
from ed_mrdt.photophysics import PhotonSimulator

pytestmark = pytest.mark.gpu

CASES = {
    # name: (config of a seed, seeds, steps, imaging, pair-distance bin edges in um)
    "binding": (lambda s: workloads.binding_fret(seed=s), 24, 100, True,
                np.array([0.0, 0.02, 0.05, 0.1, 0.2, 0.4])),
    "dense_harmonic": (lambda s: workloads.large_membrane(seed=s, side=14.0, n_alive=19_600,
                                                          pixels=64), 12, 60, False,
                       np.array([0.0, 0.005, 0.01, 0.02, 0.05, 0.1])),
}


def _observables(cfg, edges, pos0, alive0, pos, alive, species, flim, counts_mid):
    L = np.asarray(cfg.membrane.size, dtype=np.float64)
    keep = alive0 & alive
    d = pos[keep] - pos0[keep]
    d -= L * np.round(d / L)  # periodic box, displacements well below L / 2
    msd = float(np.mean(np.sum(d * d, axis=1)))
    tree = cKDTree(np.mod(pos[alive], L), boxsize=L)
    cum = tree.count_neighbors(tree, edges).astype(np.float64)
    rdf = np.diff(cum) / 2.0  # unordered pairs per distance bin
    n_sp = len(cfg.species)
    counts_end = np.bincount(species[alive], minlength=n_sp).astype(np.float64)
    out = [[msd], rdf, counts_mid, counts_end]
    if flim is not None:
        h = flim.reshape(-1, flim.shape[-1]).sum(axis=0).astype(np.float64)
        ph = 2.0 * np.pi * (np.arange(h.size) + 0.5) / h.size
        out.append([float(np.sum(h * np.cos(ph)) / h.sum()),
                    float(np.sum(h * np.sin(ph)) / h.sum())])
    return np.concatenate(out)


def _gpu_replica(case, seed):
    make, _, steps, imaging, edges = CASES[case]
    cfg = make(seed)
    st = ParticleState(cfg)
    pos0, alive0 = st.positions.copy(), st.is_alive.astype(bool).copy()
    eng = DynamicsEngine(cfg, st, precision="fp32", rng="philox")
    if imaging:
        ph = PhotonSimulator(cfg, st, min(1.0, cfg.scan.dwell_time / cfg.scan.frame_interval))
        mic = VirtualMicroscope(cfg, st)
    mid = None
    for k in range(steps):
        eng.step()
        if imaging:
            mic.process_photons(ph.evolve_one_bd_step())
        if k + 1 == steps // 2:
            a = st.is_alive.astype(bool)
            mid = np.bincount(st.species_id[a], minlength=len(cfg.species)).astype(np.float64)
    flim = mic.finalize_frame().flim_stack if imaging else None
    return _observables(cfg, edges, pos0, alive0, st.positions.astype(np.float64),
                        st.is_alive.astype(bool), st.species_id, flim, mid)


def _oracle_replica(case, seed):
    make, _, steps, imaging, edges = CASES[case]
    cfg = make(seed)
    ost = rm.OracleState(cfg)
    pos0, alive0 = ost.positions.copy(), ost.is_alive.astype(bool).copy()
    odyn = rm.OracleDynamics(cfg, ost)
    if imaging:
        oph = rm.OraclePhotons(cfg, ost, min(1.0, cfg.scan.dwell_time / cfg.scan.frame_interval))
        omic = rm.OracleMicroscope(cfg, ost)
    mid = None
    for k in range(steps):
        odyn.step()
        if imaging:
            omic.process_photons(oph.evolve_one_bd_step())
        if k + 1 == steps // 2:
            a = ost.is_alive.astype(bool)
            mid = np.bincount(ost.species_id[a], minlength=len(cfg.species)).astype(np.float64)
    flim = omic.finalize_frame()["flim_stack"] if imaging else None
    return _observables(cfg, edges, pos0, alive0, ost.positions, ost.is_alive.astype(bool),
                        ost.species_id, flim, mid)


@pytest.mark.parametrize("case", list(CASES))
def test_fp32_philox_statistics_match_oracle(oracle_lib, case):
    make, n, _, imaging, edges = CASES[case]
    a = np.array([_gpu_replica(case, s) for s in range(n)])
    b = np.array([_oracle_replica(case, s) for s in range(n)])
    ma, mb = a.mean(axis=0), b.mean(axis=0)
    se = np.sqrt(a.var(axis=0, ddof=1) / n + b.var(axis=0, ddof=1) / n)
    cfg = make(0)
    n_sp = len(cfg.species)
    nb = edges.size - 1
    names = (["msd"] + [f"pairs[{edges[i]:.3f},{edges[i + 1]:.3f})" for i in range(nb)]
             + [f"mid:{s.name}" for s in cfg.species] + [f"end:{s.name}" for s in cfg.species]
             + (["phasor g", "phasor s"] if imaging else []))
    assert len(names) == ma.size
    for i, nm in enumerate(names):
        print(f"{case} {nm:26s} fp32 {ma[i]:12.6g}  oracle {mb[i]:12.6g}  se {se[i]:10.3g}")
    bad = [nm for i, nm in enumerate(names)
           if abs(ma[i] - mb[i]) > 4.5 * se[i] + 1e-12 * abs(mb[i])]
    assert not bad, f"outside 4.5 standard errors: {bad}"
    # the observables move: diffusion and reactions happened
    assert ma[0] > 1e-4
    assert np.any(a[:, 1 + nb + n_sp:1 + nb + 2 * n_sp] != a[:, 1 + nb:1 + nb + n_sp])
