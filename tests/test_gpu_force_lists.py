"""Force sub-row lists against the strip kernel (FP32 / FP64, Philox).

Dense membranes, in steps that do not need the reference's in-cell order,
sort each cell by (force sub-row, particle index) and read a particle's
force candidates from global memory through the sub-row starts
(`k_sort_cells_fine`, `k_force_fine`); the strip kernel stages the same
candidates in LDS.  Both visit the candidates in the same order, so every
state, counter and frame is the same bit for bit — over hundreds of steps of
chaotic dense dynamics, where any difference in a force sum would grow —
periodic and walled, with imaging, up to the 1M-particle membrane."""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

from ed_mrdt import _hip, workloads
from ed_mrdt.dynamics import DynamicsEngine
from ed_mrdt.microscope import VirtualMicroscope
from ed_mrdt.particles import ParticleState
from ed_mrdt.photophysics import PhotonSimulator

pytestmark = pytest.mark.gpu


def _walled():
    cfg = workloads.large_membrane(side=14.0, n_alive=19_600, pixels=64)
    cfg.membrane.periodic = [False, False]
    return cfg


CASES = {
    "dense-14um": (lambda: workloads.large_membrane(side=14.0, n_alive=19_600, pixels=64),
                   "fp32", 600),
    "dense-walled": (_walled, "fp32", 400),
    "dense-14um-fp64": (lambda: workloads.large_membrane(side=14.0, n_alive=19_600, pixels=64),
                        "fp64", 200),
    "wave-13k": (lambda: workloads.wave_system(seed=1, n_steps=2000), "fp32", 300),
    "1M": (lambda: workloads.large_membrane(pixels=256), "fp32", 200),
}


def _run(case, lists):
    make, precision, steps = CASES[case]
    cfg = make()
    st = ParticleState(cfg)
    eng = DynamicsEngine(cfg, st, precision=precision, rng="philox", pair_cap=False)
    eng.set_tuning(force_lists=lists)
    PhotonSimulator(cfg, st, min(1.0, cfg.scan.dwell_time / cfg.scan.frame_interval))
    mic = VirtualMicroscope(cfg, st)
    L = _hip.lib()
    done, frames = 0, []
    while done < steps:
        k = min(100, steps - done)
        _hip.check(L.edm_run_frame(eng.handle, k), "edm_run_frame")
        done += k
        frames.append(mic.finalize_frame().flim_stack.copy())
    st.mark_stale()
    out = (st.positions.copy(), st.species_id.copy(), st.is_alive.copy(), st.ctmc_state.copy(),
           (eng.n_catalytic, eng.n_recruited, eng.n_detached, eng.n_converted), frames)
    eng.close()
    return out


@pytest.mark.parametrize("case", list(CASES))
def test_force_lists_equal_strips(case):
    print(f"[force-lists] {case}", flush=True)
    a = _run(case, 1)
    b = _run(case, 0)
    for x, y in zip(a[:4], b[:4]):
        np.testing.assert_array_equal(x, y)
    assert a[4] == b[4]
    for fa, fb in zip(a[5], b[5]):
        np.testing.assert_array_equal(fa, fb)
    assert sum(int(f.sum()) for f in a[5]) > 0
