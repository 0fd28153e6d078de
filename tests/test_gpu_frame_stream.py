"""Streamed frames of one membrane (edm_frame_collect / edm_frame_fetch)
against frames ended in place (finalize_frame: noise, dense copy, reset):
the same stacks bin for bin, the same dark counts, afterpulses and signal
photons per frame, with the next frame's steps queued before each fetch so
that the compaction on the second stream overlaps them.  The streamed frame
end reads only the pixel rows whose flag a count has set; the odd geometry
puts dark counts on rows no photon reached and rows that are not whole
wavefronts or 16-byte vectors."""

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

def _odd_geometry():
    """37 x 29 px (rows not a multiple of 64), 250 bins (not a multiple of 4),
    dark counts on a fifth of the pixels and 5 % afterpulses."""
    cfg = workloads.binding_fret(seed=5)
    cfg.scan.pixels_x, cfg.scan.pixels_y, cfg.scan.pixel_size = 37, 29, 108.0
    cfg.tcspc.n_bins = 250
    cfg.detector.dark_count_rate = 2e4
    cfg.detector.afterpulsing_prob = 0.05
    return cfg


CASES = {
    "binding-fret": (lambda: workloads.binding_fret(seed=4), 40, 5),
    "odd-geometry": (_odd_geometry, 30, 4),
    "wave-128px": (lambda: workloads.wave_system(seed=3, n_steps=1000), 100, 4),
    "membrane-512px": (lambda: workloads.large_membrane(side=31.6, n_alive=100_000,
                                                        pixels=512), 100, 3),
}


def _engine(cfg):
    st = ParticleState(cfg)
    eng = DynamicsEngine(cfg, st, precision="fp32", rng="philox", pair_cap=False)
    PhotonSimulator(cfg, st, min(1.0, cfg.scan.dwell_time / cfg.scan.frame_interval))
    return st, eng, VirtualMicroscope(cfg, st)


def _steps(eng, k):
    tot, fms = ctypes.c_double(0), ctypes.c_double(0)
    _hip.check(_hip.lib().edm_time_steps(eng.handle, k, 1, ctypes.byref(tot), ctypes.byref(fms),
                                         None), "edm_time_steps")


@pytest.mark.parametrize("case", list(CASES))
def test_streamed_frames_equal_frames_ended_in_place(case):
    make, fe, n_frames = CASES[case]
    print(f"[frame-stream] {case}", flush=True)
    _, ea, ma = _engine(make())
    dense, detected, prev = [], [], ma._detected_counter()
    for _ in range(n_frames):
        _steps(ea, fe)
        now = ma._detected_counter()
        detected.append(now - prev)  # what the pipeline adds to the frame's signal count
        prev = now
        dense.append(ma.finalize_frame())
    ea.close()
    _, eb, mb = _engine(make())
    sparse = []
    _steps(eb, fe)
    for k in range(n_frames):
        t = mb.collect_frame()
        if k + 1 < n_frames:
            _steps(eb, fe)  # queued before the fetch: overlaps the compaction
        sparse.append(mb.fetch_frame(t))
    eb.close()
    for k, (d, s) in enumerate(zip(dense, sparse)):
        np.testing.assert_array_equal(s.to_frame().flim_stack, d.flim_stack,
                                      err_msg=f"frame {k}")
        assert (s.frame_index, s.n_signal_photons, s.n_dark_counts, s.n_afterpulses) == (
            d.frame_index, detected[k], d.n_dark_counts, d.n_afterpulses), f"frame {k}"
        assert np.all(np.diff(s.flat_index.astype(np.int64)) > 0)
    assert sum(s.total_counts for s in sparse) > 0 and sum(detected) > 0
