"""Tuning knobs on the GPU: every setting gives bit-identical results (FP32,
Philox), and the autotuner returns one of its candidates."""

from __future__ import annotations

import numpy as np
import pytest

from ed_mrdt import workloads
from ed_mrdt.dynamics import DynamicsEngine
from ed_mrdt.particles import ParticleState
from ed_mrdt.tuning import CAT_LANES, CAT_SIDES, CAT_WALKS, FORCE_LISTS, TILE_ROWS, autotune

pytestmark = pytest.mark.gpu


def _run(cfg, steps, **tune):
    st = ParticleState(cfg)
    eng = DynamicsEngine(cfg, st, precision="fp32", rng="philox", pair_cap=False)
    eng.set_tuning(**tune)
    eng.run(steps)
    out = dict(positions=st.positions.copy(), species_id=st.species_id.copy(),
               is_alive=st.is_alive.copy(),
               counters=(eng.n_catalytic, eng.n_recruited, eng.n_detached, eng.n_converted))
    eng.close()
    return out


def test_tuning_leaves_results_unchanged():
    cfg = workloads.large_membrane(side=30.0, n_alive=90_000, pixels=64)
    ref = _run(cfg, 60, cat_side=0, tile_rows=64, cat_walk=0)
    assert ref["counters"][0] > 0, "the run must exercise catalysis"
    for tune in (dict(cat_side=1, tile_rows=64, cat_walk=0), dict(cat_side=1, tile_rows=16),
                 dict(cat_side=0, tile_rows=5, cat_walk=0), dict(cat_side=-1, tile_rows=32),
                 dict(cat_side=0, cat_walk=1), dict(cat_side=1, cat_walk=1),
                 dict(cat_side=0, cat_walk=1, cat_lanes=1),
                 dict(cat_side=1, cat_walk=1, cat_lanes=64),
                 dict(cat_side=-1, cat_walk=1, cat_lanes=4),
                 dict(cat_side=0, cat_lanes=1), dict(cat_side=1, cat_lanes=8, tile_rows=16),
                 dict(cat_side=0, force_lists=0), dict(cat_side=1, force_lists=0, tile_rows=16)):
        got = _run(cfg, 60, **tune)
        for k in ("positions", "species_id", "is_alive"):
            np.testing.assert_array_equal(got[k], ref[k], err_msg=f"{tune}: {k}")
        assert got["counters"] == ref["counters"], tune


def test_autotune_picks_a_candidate():
    cfg = workloads.large_membrane(side=30.0, n_alive=90_000, pixels=64)
    res = autotune(cfg, steps=5, warmup=1)
    assert res["tile_rows"] in TILE_ROWS and res["cat_side"] in CAT_SIDES
    assert res["cat_walk"] in CAT_WALKS and res["cat_lanes"] in CAT_LANES
    assert res["force_lists"] in FORCE_LISTS
    assert len(res["times_ms"]) == (len(TILE_ROWS) + len(CAT_SIDES) +
                                    len(CAT_WALKS) * len(CAT_LANES) + len(FORCE_LISTS))
    assert all(t > 0 for t in res["times_ms"].values())
