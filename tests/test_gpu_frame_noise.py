"""Frame-end detector noise of the Philox path (edm_frame_noise) against its
host restatement (tests/philox_ref.py) draw for draw: the same dark counts
and afterpulses land in the same bins, for the single-membrane kernel and
for the ensemble's, and the hook path's pixel totals equal numpy's sums."""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

import philox_ref
from ed_mrdt import _hip, workloads
from ed_mrdt._hip import check, lib, ptr
from ed_mrdt.dynamics import DynamicsEngine
from ed_mrdt.microscope import VirtualMicroscope
from ed_mrdt.particles import ParticleState
from ed_mrdt.photophysics import PhotonSimulator

pytestmark = pytest.mark.gpu


def _imaged(seed: int, steps: int, rng: str):
    cfg = workloads.binding_fret(seed=seed)
    st = ParticleState(cfg)
    eng = DynamicsEngine(cfg, st, precision="fp64", rng=rng)
    duty = min(1.0, cfg.scan.dwell_time / cfg.scan.frame_interval)
    ph = PhotonSimulator(cfg, st, duty)
    mic = VirtualMicroscope(cfg, st)
    for _ in range(steps):
        eng.step()
        mic.process_photons(ph.evolve_one_bd_step())
    return cfg, eng, mic


def _download(h, shape):
    a = np.empty(shape, dtype=np.uint32)
    check(lib().edm_frame_download(h, ptr(a), a.size), "edm_frame_download")
    return a


@pytest.mark.parametrize("seed", [4, 9])
def test_frame_noise_matches_restatement(seed):
    cfg, eng, mic = _imaged(seed, 30, "philox")
    h = eng.handle
    before = _download(h, mic.frame_shape)
    nb = mic.frame_shape[-1]
    flat = before.reshape(-1, nb)
    assert flat.sum() > 0
    ip = mic._ip
    nd, na = ctypes.c_int64(0), ctypes.c_int64(0)
    check(lib().edm_frame_noise(h, ctypes.byref(nd), ctypes.byref(na)), "edm_frame_noise")
    after = _download(h, mic.frame_shape).reshape(-1, nb)
    want, wnd, wna = philox_ref.frame_noise(flat, ip.dark_lambda, ip.p_ap, int(eng._params.seed),
                                            eng.n_steps_done)
    assert (nd.value, na.value) == (wnd, wna)
    assert wnd > 0 and wna > 0
    np.testing.assert_array_equal(after, want)


def test_hook_pixel_totals_are_row_sums():
    cfg, eng, mic = _imaged(6, 20, "numpy")
    h = eng.handle
    frame = _download(h, mic.frame_shape)
    n_ch, nx, ny, nb = mic.frame_shape
    tot = np.empty(n_ch * nx * ny, dtype=np.uint32)
    check(lib().edm_frame_pixel_totals(h, ptr(tot)), "edm_frame_pixel_totals")
    np.testing.assert_array_equal(tot, frame.reshape(-1, nb).sum(axis=1).astype(np.uint32))
    assert tot.sum() > 0
