"""GPU parity of the wave system (13,000 particles, sinusoidal D(x), 128 x 128
px FLIM at 78 nm), of the adaptation field and of dense binding, against the
oracle with the reference's numpy draws injected on both sides (FP64).

Bar as in test_gpu_parity.py: species / alive / bonds / CTMC / bleach arrays,
event counters, photon batches and TCSPC stacks bit-exact; positions within
1e-9 relative; arrival times within 1e-12 relative.
"""

import numpy as np
import pytest

import reference_model as rm
from ed_mrdt import _hip, workloads
from ed_mrdt.dynamics import DynamicsEngine
from ed_mrdt.microscope import VirtualMicroscope
from ed_mrdt.particles import ParticleState
from ed_mrdt.photophysics import PhotonSimulator
from parity_util import compare_states, host_arrays, oracle_arrays

pytestmark = pytest.mark.gpu


def _pair(cfg, pair_cap=True, exact=False):
    st = ParticleState(cfg)
    eng = DynamicsEngine(cfg, st, precision="fp64", rng="numpy", pair_cap=pair_cap,
                         exact_force_order=exact)
    ost = rm.OracleState(cfg)
    odyn = rm.OracleDynamics(cfg, ost, pair_cap=pair_cap)
    return st, eng, ost, odyn


def _counters(eng):
    return (eng.n_bindings, eng.n_dissociations, eng.n_catalytic, eng.n_recruited,
            eng.n_detached, eng.n_converted)


def _ocounters(odyn):
    oc = odyn.counters()
    return tuple(oc[k] for k in ("n_bindings", "n_dissociations", "n_catalytic",
                                 "n_recruited", "n_detached", "n_converted"))


@pytest.mark.parametrize("n_steps,exact", [(100, True), (25, False)])
def test_wave_system_full_path_fp64(oracle_lib, n_steps, exact):
    """The wave system (sinusoidal D(x) = D_s (1 + 0.3 sin(kx x) sin(ky y))
    with its gradient drift, 13,000 alive particles + pools, the plain
    membrane's reaction network) with photons and the 128 x 128 px, 78 nm
    FLIM image, then a frame end with dark counts and afterpulses.  With the
    reference's pair-order force sums the run is bit-identical to the
    oracle, 100 steps.  The LDS-tiled force kernel sums a particle's pair
    forces in another order; at 130 particles / um^2 WCA contacts (which
    the 1 ms step turns into displacement-capped kicks) amplify that
    rounding, so it is held to the 1e-9 bar over 25 steps."""
    cfg = workloads.wave_system(n_steps=n_steps)
    assert cfg.diffusion_field.type == "sinusoidal"
    assert (cfg.scan.pixels_x, cfg.scan.pixels_y) == (128, 128)
    st, eng, ost, odyn = _pair(cfg, exact=exact)
    assert int(st.is_alive.sum()) == 13000
    duty = min(1.0, cfg.scan.dwell_time / cfg.scan.frame_interval)
    ph = PhotonSimulator(cfg, st, duty)
    mic = VirtualMicroscope(cfg, st)
    oph = rm.OraclePhotons(cfg, ost, duty)
    omic = rm.OracleMicroscope(cfg, ost)
    tot_ph = tot_det = 0
    for k in range(n_steps):
        eng.step()
        odyn.step()
        b = ph.evolve_one_bd_step()
        ob = oph.evolve_one_bd_step()
        assert b.n_photons == ob["n_photons"], f"step {k}: photon count"
        if b.n_photons:
            np.testing.assert_array_equal(b.particle_idx, ob["particle_idx"])
            np.testing.assert_array_equal(b.channel, ob["channel"])
            np.testing.assert_allclose(b.arrival_time, ob["arrival_time"], rtol=1e-12, atol=0)
        nd = mic.process_photons(b)
        assert nd == omic.process_photons(ob), f"step {k}: detected photons"
        tot_ph += b.n_photons
        tot_det += nd
        if (k + 1) % 25 == 0:
            compare_states(host_arrays(st), oracle_arrays(ost), label=f"step {k + 1}: ")
            assert _counters(eng) == _ocounters(odyn), f"step {k + 1}"
    fr = mic.finalize_frame()
    ofr = omic.finalize_frame()
    assert fr.n_dark_counts == ofr["n_dark"] and fr.n_afterpulses == ofr["n_ap"]
    np.testing.assert_array_equal(fr.flim_stack, ofr["flim_stack"])
    assert fr.flim_stack.shape[1:] == (128, 128, 256)
    assert tot_ph > 1000 and tot_det > 0 and eng.n_detached > 0
    print(f"wave system exact={exact}: photons {tot_ph}, detected {tot_det}, "
          f"events {_counters(eng)}, frame counts {fr.total_counts}")


def _adapt_config(fb_res):
    """The plain membrane with the adaptation field on (use_adaptation, the
    ad-hoc attributes dynamics.py:2078-2081 reads): a fast, strong
    adaptation (tau_w = 20 ms, K_w = 1 /um^2) of a PI3K catalysis made 2000x
    faster, so that the factor moves events within 60 steps.  fb_res is the
    feedback grids' resolution: 1.0 um matches the 1 um adaptation grid
    (the reference's production script sets 1.0); 0.5 um makes the
    reference read the feedback grid's indices on the adaptation array."""
    cfg = workloads.plain_membrane()
    cfg.use_adaptation = True
    cfg.adaptation_tau_w = 0.02
    cfg.adaptation_K_w = 1.0
    cfg.adaptation_n_w = 2.0
    for c in cfg.catalytic_reactions:
        if "PI3K" in c.enzyme:
            c.k_cat *= 2000.0
    for fb in cfg.feedback_rules:
        fb.grid_resolution = fb_res
    return cfg


@pytest.mark.parametrize("fb_res", [1.0, 0.5])
def test_adaptation_field_fp64(oracle_lib, fb_res):
    """update_adaptation_field (dynamics.py:1094-1112, 2228-2242) and the
    adapted PI3K catalysis (apply_catalytic_with_feedback_and_adaptation,
    dynamics.py:987-1090) against the oracle: the same catalytic events
    and states every 10 steps.  The same run without adaptation gives
    different events, so the factor is exercised."""
    cfg = _adapt_config(fb_res)
    st, eng, ost, odyn = _pair(cfg)
    assert "adapt" in eng.plan.kinds
    for k in range(60):
        eng.step()
        odyn.step()
        if (k + 1) % 10 == 0:
            compare_states(host_arrays(st), oracle_arrays(ost), label=f"step {k + 1}: ")
            assert _counters(eng) == _ocounters(odyn), f"step {k + 1}"
    cfg0 = _adapt_config(fb_res)
    cfg0.use_adaptation = False
    ost0 = rm.OracleState(cfg0)
    odyn0 = rm.OracleDynamics(cfg0, ost0)
    for _ in range(60):
        odyn0.step()
    assert odyn0.counters()["n_catalytic"] > eng.n_catalytic > 0
    print(f"adaptation fb_res={fb_res}: catalytic {eng.n_catalytic} "
          f"(without adaptation {odyn0.counters()['n_catalytic']})")


def _dense_binding_config(seed=5):
    """A 2 x 2 um box at ~1000 particles / um^2 with PHAkt + PIP3 binding at
    P_bind = 1 and a 150 nm contact: each free PHAkt has ~28 PIP3 within
    contact, so a step accepts far more candidate edges than there are
    particles, and the pair list (~40 pairs per particle) crosses the
    reference's min(30 N, 5e5) cap."""
    cfg = workloads.binding_fret(seed=seed, side=2.0, regions=False)
    pops = cfg.initial_populations
    pops["PIP2"] = 800
    pops["PIP3"] = 1600
    pops["PHAkt"] = 1400
    for r in cfg.reactions:
        r.kon = 1.0e4
        r.contact_radius = 150.0
    return cfg


@pytest.mark.parametrize("pair_cap", [True, False])
def test_dense_binding_edges_beyond_n_fp64(oracle_lib, pair_cap):
    """Greedy binding (dynamics.py:529-606) with more accepted candidate
    edges than particles, with and without the reference pair cap: the
    same bonds as the oracle's sequential pass."""
    cfg = _dense_binding_config()
    st, eng, ost, odyn = _pair(cfg, pair_cap=pair_cap)
    n = st.N
    n_pairs = eng.step_begin()
    eng.step_finish(eng._numpy_draws(n_pairs))
    odyn.step()
    found = odyn.last_n_found
    if pair_cap:
        assert found > 30 * n and n_pairs == 30 * n, (found, n_pairs, n)
    else:
        assert n_pairs == found > 30 * n
    compare_states(host_arrays(st), oracle_arrays(ost), label="step 1: ")
    assert _counters(eng) == _ocounters(odyn)
    # accepted edges of the first step: every listed PHAkt-PIP3 pair within
    # contact (P_bind = 1), far more than N
    assert eng.n_bindings > 0
    for k in range(1, 8):
        eng.step()
        odyn.step()
        compare_states(host_arrays(st), oracle_arrays(ost), label=f"step {k + 1}: ")
        assert _counters(eng) == _ocounters(odyn), f"step {k + 1}"
    print(f"dense binding pair_cap={pair_cap}: N {n}, pairs {n_pairs} of {found}, "
          f"bindings {eng.n_bindings}, dissociations {eng.n_dissociations}")


def test_dense_binding_edge_overflow_is_reported():
    """Philox draws (no pair list to bound the candidates): a step that
    accepts more candidate edges than the edge buffer (4 per particle)
    holds is reported as EDM_ECAPACITY at the next synchronising call,
    not truncated silently.  The box is denser still (6400 PIP3 in
    2 x 2 um, ~110 within contact of each PHAkt)."""
    cfg = _dense_binding_config(seed=6)
    cfg.initial_populations["PIP3"] = 6400
    st = ParticleState(cfg)
    eng = DynamicsEngine(cfg, st, precision="fp32", rng="philox", pair_cap=False)
    with pytest.raises(_hip.HipError, match="edge buffer"):
        eng.run(1)
        eng.synchronize()
