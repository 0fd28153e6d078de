"""Step loops replayed from recorded HIP graphs (edm_run, edm_run_frame,
edm_time_steps with EDM_TUNE_GRAPHS on) against the same steps launched one
by one: with Philox draws both run the same kernels with the same draw keys,
so every result is bit-identical — positions, species, flags, bonds, dye
states, event and photon counters, FLIM frames and the device observables.

Runs are split so that stretches start from every phase of the buffer
rotation, cross the longest recording (256 steps), include runs too short to
record (< 4 steps), end frames between runs, and change a tuning knob and
open the MSD tracker mid-run (both drop the recordings).  The
autocatalytic membrane also orders its conversions by the reference's pair
index, which step loops then compute on the device every step."""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

from ed_mrdt import _hip, workloads
from ed_mrdt.dynamics import DynamicsEngine
from ed_mrdt.microscope import VirtualMicroscope
from ed_mrdt.observables import MSDTracker
from ed_mrdt.particles import ParticleState
from ed_mrdt.photophysics import PhotonSimulator
from parity_util import autocatalytic_config

pytestmark = pytest.mark.gpu

CASES = {
    "plain-fp32": lambda: workloads.plain_membrane(seed=4),
    "binding-regions-fp32": lambda: workloads.binding_fret(seed=6),
    "proximity-walled-fp64": lambda: workloads.binding_fret(seed=8, proximity=True,
                                                            reflective=True),
    "wave-13k-fp32": lambda: workloads.wave_system(seed=2, n_steps=2000),
    "autocatalysis-fp32": lambda: autocatalytic_config(feedback=True, k_cat=2.0),
}
PRECISION = {"proximity-walled-fp64": "fp64"}

# (steps of the run, what happens after it)
SCHEDULE = [(7, "frame"), (3, None), (100, "frame"), (13, None), (300, "frame"),
            (5, "tune"), (64, None), (9, "msd"), (120, "frame"), (2, None), (50, "frame")]


def _run(case, graphs):
    cfg = CASES[case]()
    st = ParticleState(cfg)
    eng = DynamicsEngine(cfg, st, precision=PRECISION.get(case, "fp32"), rng="philox")
    eng.set_tuning(graphs=1 if graphs else 0)
    ph = PhotonSimulator(cfg, st, min(1.0, cfg.scan.dwell_time / cfg.scan.frame_interval))
    mic = VirtualMicroscope(cfg, st)
    L = _hip.lib()
    frames, msd = [], []
    tracker = None
    for k, (n, after) in enumerate(SCHEDULE):
        if k % 2 == 0:
            _hip.check(L.edm_run_frame(eng.handle, n), "edm_run_frame")
        else:
            tot, fms = ctypes.c_double(0), ctypes.c_double(0)
            _hip.check(L.edm_time_steps(eng.handle, n, 1, ctypes.byref(tot), ctypes.byref(fms),
                                        None), "edm_time_steps")
        if after == "frame":
            frames.append(mic.finalize_frame().flim_stack.copy())
        elif after == "tune":
            eng.set_tuning(tile_rows=16)
        elif after == "msd":
            tracker = MSDTracker(eng)
        if tracker is not None:
            msd.append(tracker.sums())
    st.mark_stale()
    out = dict(positions=st.positions.copy(), species=st.species_id.copy(),
               alive=st.is_alive.copy(), bond=st.bond_partner.copy(),
               state=st.ctmc_state.copy(), bleached=st.is_bleached.copy(),
               counters=(eng.n_catalytic, eng.n_recruited, eng.n_detached, eng.n_converted),
               frames=frames, msd=msd, replays=eng.graph_replays, status=eng.graph_status)
    eng.close()
    return out


@pytest.mark.parametrize("case", list(CASES))
def test_graph_replay_equals_launch_by_launch(case):
    print(f"[graphs] {case}", flush=True)
    a = _run(case, graphs=True)
    b = _run(case, graphs=False)
    assert b["replays"] == 0
    assert a["status"] == 0, f"graph recording failed with status {a['status']}"
    assert a["replays"] > 0, "no stretch ran from a recorded graph"
    for key in ("positions", "species", "alive", "bond", "state", "bleached"):
        np.testing.assert_array_equal(a[key], b[key], err_msg=key)
    assert a["counters"] == b["counters"]
    assert len(a["frames"]) == len(b["frames"])
    for fa, fb in zip(a["frames"], b["frames"]):
        np.testing.assert_array_equal(fa, fb)
    assert sum(int(f.sum()) for f in a["frames"]) > 0
    assert a["msd"] == b["msd"]
    print(f"[graphs] {case}: {a['replays']} replays, "
          f"{sum(n for n, _ in SCHEDULE)} steps", flush=True)


def test_device_decided_pair_cap_equals_host_decided():
    """The faithful pair cap (min(30 N, 500000) pairs, dynamics.py:220-278)
    on a 229k-slot membrane whose pair list (about 700k pairs within the
    150 nm contact radius) exceeds it: step loops decide on
    the device whether the cap truncates (no read-back, graph replays), the
    parity entry points read the count back and decide on the host; both
    give the same state bit for bit."""
    cfg = workloads.large_membrane(side=45.0, n_alive=200_000, pixels=64)
    L = _hip.lib()
    out = []
    for host_decided in (True, False):
        st = ParticleState(cfg)
        eng = DynamicsEngine(cfg, st, precision="fp32", rng="philox")
        print(f"[graphs] pair cap, host decided = {host_decided}, N={st.N}", flush=True)
        if host_decided:
            for _ in range(40):
                npairs = ctypes.c_int64(0)
                _hip.check(L.edm_step_begin(eng.handle, ctypes.byref(npairs)), "edm_step_begin")
                assert npairs.value == min(30 * st.N, 500_000), \
                    f"the cap must truncate ({npairs.value} pairs)"
                _hip.check(L.edm_step_finish(eng.handle, None, 0), "edm_step_finish")
        else:
            eng.set_tuning(graphs=1)
            _hip.check(L.edm_run(eng.handle, 40), "edm_run")
            assert eng.graph_replays > 0 and eng.graph_status == 0
        st.mark_stale()
        out.append((st.positions.copy(), st.species_id.copy(), st.is_alive.copy(),
                    eng.n_catalytic, eng.n_converted))
        eng.close()
    a, b = out
    for x, y in zip(a[:3], b[:3]):
        np.testing.assert_array_equal(x, y)
    assert a[3:] == b[3:]
