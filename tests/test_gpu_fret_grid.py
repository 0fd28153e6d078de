"""Proximity-mode FRET through the grid of target particles
(photophysics.py:592-683).  The reference scans every particle for every
donor; the device bins the map's alive targets on a grid of cells no
narrower than the proximity radius and searches a donor's 3 x 3 cells.

Checks, per case: the grid search and the scan (EDM_TUNE_FRET_SEARCH 1)
give identical per-particle rates on the same state, both classes (target
near / no target) occur, and in FP64 both equal the oracle's proximity
kernel (tests/test_oracle_kernels.py pins it) on the downloaded state.
Cases cover periodic and walled boxes, radii from 5 nm (grid capped by the
particle count) to more than a third of the box (one or two cells per
axis), and a 60 x 60 um system with 126k particles."""

from __future__ import annotations

import numpy as np
import pytest

from ed_mrdt import workloads
from ed_mrdt.dynamics import DynamicsEngine
from ed_mrdt.microscope import VirtualMicroscope
from ed_mrdt.particles import ParticleState
from ed_mrdt.photophysics import PhotonSimulator

pytestmark = pytest.mark.gpu

CASES = [
    # (label, side um, radius nm, reflective, precision, steps)
    ("periodic-50nm", 4.0, 50.0, False, "fp64", 20),
    ("walled-50nm", 4.0, 50.0, True, "fp64", 20),
    ("periodic-5nm", 4.0, 5.0, False, "fp64", 10),
    ("periodic-1.5um", 4.0, 1500.0, False, "fp64", 5),
    ("walled-1.5um", 4.0, 1500.0, True, "fp64", 5),
    ("periodic-100nm-fp32", 12.0, 100.0, False, "fp32", 30),
    ("large-60um-fp32", 60.0, 50.0, False, "fp32", 20),
]


def _rates(eng, sim, search):
    eng.set_tuning(fret_search=search)
    return sim._k_fret.copy()


@pytest.mark.parametrize("label,side,radius,walled,precision,steps", CASES,
                         ids=[c[0] for c in CASES])
def test_grid_search_equals_scan(oracle_lib, label, side, radius, walled, precision, steps):
    cfg = workloads.binding_fret(seed=17, proximity=True, side=side, reflective=walled)
    cfg.fret_pairs[0].proximity_radius = radius
    st = ParticleState(cfg)
    eng = DynamicsEngine(cfg, st, precision=precision, rng="philox")
    duty = min(1.0, cfg.scan.dwell_time / cfg.scan.frame_interval)
    sim = PhotonSimulator(cfg, st, duty)
    mic = VirtualMicroscope(cfg, st)
    print(f"[fret-grid] {label}: N={st.N} precision={precision}", flush=True)
    for _ in range(steps):
        eng.step()
        mic.process_photons(sim.evolve_one_bd_step())
    k_grid = _rates(eng, sim, 0)
    k_scan = _rates(eng, sim, 1)
    eng.set_tuning(fret_search=0)
    np.testing.assert_array_equal(k_grid, k_scan)
    vals = np.unique(k_grid[k_grid > 0])
    assert 1 <= vals.size <= 2
    if 50.0 <= radius < 1000.0:
        assert vals.size == 2, "both near and far donors expected"
    if precision == "fp64":
        from kernels import _compute_fret_rates_proximity_numba
        st.mark_stale()
        N = st.N
        k_or = np.zeros(N)
        acc = np.full(N, -1, np.int32)
        fp = cfg.fret_pairs[0]
        names = [s.name for s in cfg.species]
        rmat = sim.rate_matrices[sim._fret_maps[0]["donor_fluoro_id"]]
        tD = rmat.get_lifetime(rmat.excited_state)
        L = cfg.membrane.size
        per = cfg.membrane.periodic
        _compute_fret_rates_proximity_numba(
            k_or, acc, np.ascontiguousarray(st.positions), st.species_id.astype(np.int32),
            st.has_dye.astype(bool), st.is_alive.astype(bool), st.is_bleached.astype(bool),
            st.dye_type_id.astype(np.int32), N, tD, sim._fret_maps[0]["donor_fluoro_id"],
            names.index(fp.proximity_target_species), radius * 1e-3,
            sim._pp.fm_k_hi[0], sim._pp.fm_k_lo[0], L[0], L[1], bool(per[0]), bool(per[1]))
        np.testing.assert_array_equal(k_grid, k_or)
    eng.close()
