"""Stored golden vectors (tests/golden/*.npz, written by
tests/golden/make_golden.py from the oracle).

CPU: the oracle reproduces its stored output exactly.
GPU: the HIP path, with the reference's numpy draws injected, reproduces
the stored output: flags, species, bonds, CTMC states, event counters and
the TCSPC stack bit-exact, positions within 1e-9 relative (FP64).
"""

from __future__ import annotations

import os

import numpy as np
import pytest

from conftest import ROOT

GOLDEN = os.path.join(ROOT, "tests", "golden")
import sys  # noqa: E402

sys.path.insert(0, GOLDEN)
import make_golden  # noqa: E402

INT_KEYS = ("species_id", "is_alive", "bond_partner", "ctmc_state", "is_bleached")


def _load(name):
    with np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.mark.parametrize("name", sorted(make_golden.CASES))
def test_oracle_reproduces_golden(oracle_lib, name):
    ref = _load(name)
    out = make_golden.run_case(name)
    for k in INT_KEYS + ("counters", "photons"):
        np.testing.assert_array_equal(out[k], ref[k], err_msg=k)
    np.testing.assert_array_equal(out["positions"], ref["positions"])
    if "flim_nz_index" in ref:
        np.testing.assert_array_equal(out["flim_nz_index"], ref["flim_nz_index"])
        np.testing.assert_array_equal(out["flim_nz_count"], ref["flim_nz_count"])
        np.testing.assert_array_equal(out["frame_noise"], ref["frame_noise"])


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(make_golden.CASES))
def test_hip_path_reproduces_golden(name):
    from ed_mrdt.dynamics import DynamicsEngine
    from ed_mrdt.microscope import VirtualMicroscope
    from ed_mrdt.particles import ParticleState
    from ed_mrdt.photophysics import PhotonSimulator

    ref = _load(name)
    fac, kw, steps, image = make_golden.CASES[name]
    cfg = fac(**kw)
    st = ParticleState(cfg)
    eng = DynamicsEngine(cfg, st, precision="fp64", rng="numpy", pair_cap=True)
    ph = PhotonSimulator(cfg, st, min(1.0, cfg.scan.dwell_time / cfg.scan.frame_interval))
    mic = VirtualMicroscope(cfg, st)
    emitted = detected = 0
    for _ in range(steps):
        eng.step()
        b = ph.evolve_one_bd_step()
        emitted += b.n_photons
        if image:
            detected += mic.process_photons(b)
    got = dict(species_id=st.species_id, is_alive=st.is_alive.astype(np.uint8),
               bond_partner=st.bond_partner, ctmc_state=st.ctmc_state,
               is_bleached=st.is_bleached.astype(np.uint8))
    for k in INT_KEYS:
        np.testing.assert_array_equal(got[k], ref[k], err_msg=k)
    scale = np.maximum(np.abs(ref["positions"]), 1.0)
    assert float(np.max(np.abs(st.positions - ref["positions"]) / scale)) <= 1e-9
    names = [str(s) for s in ref["counter_names"]]
    counters = dict(zip(names, ref["counters"].tolist()))
    assert eng.n_catalytic == counters["n_catalytic"]
    assert eng.n_recruited == counters["n_recruited"]
    assert eng.n_detached == counters["n_detached"]
    assert eng.n_converted == counters["n_converted"]
    assert eng.n_bindings == counters["n_bindings"]
    assert [emitted, detected if image else 0] == [int(ref["photons"][0]),
                                                   int(ref["photons"][1]) if image else 0]
    if image:
        fr = mic.finalize_frame()
        flat = fr.flim_stack.ravel()
        nz = np.flatnonzero(flat)
        np.testing.assert_array_equal(nz, ref["flim_nz_index"])
        np.testing.assert_array_equal(flat[nz], ref["flim_nz_count"])
        assert [fr.n_dark_counts, fr.n_afterpulses] == ref["frame_noise"].tolist()
