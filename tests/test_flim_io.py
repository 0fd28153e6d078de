"""Streaming FLIM output (the reference's microscopy datasets, io.py:288-430):
frames written one at a time, dense or sparse, read back as the dense
[F, C, Px, Py, Tb] stack and the per-frame counts; no GPU needed."""

import numpy as np

from ed_mrdt import workloads
from ed_mrdt.flim_io import FLIMWriter, read_flim_stack, read_frames, read_metadata
from ed_mrdt.microscope import FLIMFrame, SparseFrame


def test_frames_round_trip(tmp_path):
    cfg = workloads.plain_membrane()
    g = np.random.default_rng(0)
    shape = (2, 16, 12, 256)
    dense = []
    for k in range(5):
        s = np.zeros(shape, np.uint32)
        idx = g.integers(0, s.size, 300)
        np.add.at(s.reshape(-1), idx, 1)
        if k == 3:
            s[:] = 0  # an empty frame
        dense.append(s)
    path = tmp_path / "run.npz"
    with FLIMWriter(path, cfg) as w:
        for k, s in enumerate(dense):
            if k % 2:
                w.write_frame(SparseFrame.from_stack(s, k, n_signal=k, n_dark=2 * k, n_ap=k % 3))
            else:
                w.write_frame(FLIMFrame.from_stack(s, k, n_signal=k, n_dark=2 * k, n_ap=k % 3))
    back = read_flim_stack(path)
    assert back.shape == (5,) + shape and back.dtype == np.uint32
    np.testing.assert_array_equal(back, np.stack(dense))
    frames = list(read_frames(path))
    assert [f.n_dark_counts for f in frames] == [0, 2, 4, 6, 8]
    assert [f.n_afterpulses for f in frames] == [0, 1, 2, 0, 1]
    np.testing.assert_array_equal(frames[2].to_frame().intensity, dense[2].sum(axis=-1))
    meta = read_metadata(path)
    assert meta["flim_stack_shape"] == (5,) + shape
    assert meta["species_names"][:2] == ["PIP2", "PIP3"]
    np.testing.assert_allclose(meta["bio_time"], 0.1 * np.arange(1, 6))
