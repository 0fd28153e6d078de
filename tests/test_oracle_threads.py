"""The oracle's timing variant with every loop threaded (the CPU baseline's
all-loops figure): the threaded pair list is the serial one, pair for pair
and in order, with and without the cap truncating it; threaded forces equal
the serial sums to rounding; a few steps of the all-loops dynamics follow
the serial ones."""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

import reference_model as rm
from ed_mrdt import workloads

I, D, P, LL = ctypes.c_int, ctypes.c_double, rm.P, ctypes.c_longlong


def _cells(pos, nx, ny, csx, csy):
    N = pos.shape[0]
    head = np.empty(nx * ny, np.int32)
    nxt = np.empty(N, np.int32)
    rm.lib().orc_build_cell_list(P(pos), I(N), P(head), P(nxt), I(nx), I(ny), D(csx), D(csy))
    return head, nxt


@pytest.mark.parametrize("periodic", [True, False])
@pytest.mark.parametrize("cap", [None, 3000])
def test_threaded_pair_list_is_the_serial_one(periodic, cap):
    rng = np.random.default_rng(5)
    L, N, cut = 6.0, 4000, 0.15
    pos = np.ascontiguousarray(rng.uniform(0, L, size=(N, 2)))
    nx = ny = int(L / cut)
    cs = L / nx
    head, nxt = _cells(pos, nx, ny, cs, cs)
    mp = cap if cap else N * 64
    out = []
    for par in (False, True):
        pairs = np.zeros((mp, 2), np.int32)
        found = ctypes.c_longlong(0)
        args = (P(pos), I(N), P(head), P(nxt), I(nx), I(ny), D(cs), D(cs), D(L), D(L),
                I(int(periodic)), I(int(periodic)), D(cut), LL(mp), P(pairs),
                ctypes.byref(found))
        n = (rm.lib().orc_cell_list_pairs_par(*args, I(4)) if par
             else rm.lib().orc_cell_list_pairs(*args))
        out.append((n, found.value, pairs[:n].copy()))
    (n0, f0, p0), (n1, f1, p1) = out
    assert (n0, f0) == (n1, f1)
    np.testing.assert_array_equal(p0, p1)
    if cap:
        assert f0 > cap and n0 == cap
    assert n0 > 1000


def test_all_loops_steps_follow_the_serial_ones():
    cfg = workloads.plain_membrane(seed=2)
    runs = []
    for all_loops in (False, True):
        st = rm.OracleState(cfg)
        dyn = rm.OracleDynamics(cfg, st, pair_cap=True, n_threads=4, all_loops=all_loops)
        for _ in range(3):
            dyn.step()
        runs.append((st.positions.copy(), st.species_id.copy(), dyn.counters()))
    (pa, sa, ca), (pb, sb, cb) = runs
    np.testing.assert_allclose(pb, pa, rtol=1e-9, atol=1e-12)
    np.testing.assert_array_equal(sa, sb)
    assert ca == cb
