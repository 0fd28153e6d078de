"""GPU parity: the HIP path against the oracle on the same seeded inputs,
with the reference's numpy draws injected on both sides (rng="numpy").

Bar: particle species / alive / bond / CTMC / bleach arrays, reaction event
counters, photon batches and TCSPC stacks bit-exact; positions within 1e-9
relative (FP64).
"""

import numpy as np
import pytest

import reference_model as rm
from ed_mrdt import workloads
from ed_mrdt.dynamics import DynamicsEngine
from ed_mrdt.microscope import VirtualMicroscope
from ed_mrdt.particles import ParticleState
from ed_mrdt.photophysics import PhotonSimulator
from parity_util import autocatalytic_config, compare_states, host_arrays, oracle_arrays

pytestmark = pytest.mark.gpu


def _pair(cfg, pair_cap=True, exact=True):
    st = ParticleState(cfg)
    eng = DynamicsEngine(cfg, st, precision="fp64", rng="numpy", pair_cap=pair_cap,
                         exact_force_order=exact)
    ost = rm.OracleState(cfg)
    odyn = rm.OracleDynamics(cfg, ost, pair_cap=pair_cap)
    return st, eng, ost, odyn


def test_initial_state_matches_oracle(oracle_lib):
    cfg = workloads.binding_fret()
    st = ParticleState(cfg)
    ost = rm.OracleState(cfg)
    compare_states(host_arrays(st), oracle_arrays(ost), pos_rtol=0.0)


def test_pair_list_reference_order(oracle_lib):
    cfg = workloads.plain_membrane()
    st, eng, ost, odyn = _pair(cfg)
    for _ in range(3):
        eng.step()
        odyn.step()
    n = eng.step_begin()
    gp = eng.pairs()
    op = odyn.build_pairs()
    assert n == op.shape[0]
    np.testing.assert_array_equal(gp, op)
    eng.step_finish(eng._numpy_draws(n))


@pytest.mark.parametrize("exact", [True, False])
def test_dense_harmonic_tiles_fp64(oracle_lib, exact):
    """1M-workload density (100 / um^2, harmonic) on a 14 x 14 um box: 84
    cell rows, so the LDS-tiled kernel spans two strips per column."""
    cfg = workloads.large_membrane(side=14.0, n_alive=19_600, pixels=64)
    st, eng, ost, odyn = _pair(cfg, exact=exact)
    for k in range(20):
        eng.step()
        odyn.step()
    worst = compare_states(host_arrays(st), oracle_arrays(ost), label=f"exact={exact}: ")
    oc = odyn.counters()
    assert (eng.n_catalytic, eng.n_detached, eng.n_converted) == (
        oc["n_catalytic"], oc["n_detached"], oc["n_converted"])
    print(f"dense harmonic exact={exact}: worst relative position error {worst:.3e}")


@pytest.mark.parametrize("n_steps,exact", [(100, True), (40, False)])
def test_plain_membrane_dynamics_fp64(oracle_lib, n_steps, exact):
    cfg = workloads.plain_membrane()
    st, eng, ost, odyn = _pair(cfg, exact=exact)
    worst = 0.0
    for k in range(n_steps):
        eng.step()
        odyn.step()
        if (k + 1) % 10 == 0:
            worst = max(worst, compare_states(host_arrays(st), oracle_arrays(ost),
                                              label=f"step {k + 1}: "))
            oc = odyn.counters()
            assert eng.n_catalytic == oc["n_catalytic"]
            assert eng.n_recruited == oc["n_recruited"]
            assert eng.n_detached == oc["n_detached"]
            assert eng.n_converted == oc["n_converted"]
    print(f"plain membrane {n_steps} steps: worst relative position error {worst:.3e}")


@pytest.mark.parametrize("cat_side,tile_rows,cat_walk,cat_lanes",
                         [(0, 64, 0, 64), (1, 8, 0, 4), (0, 64, 1, 64), (1, 64, 1, 64),
                          (0, 64, 1, 4), (1, 64, 1, 1), (0, 16, 0, 1)])
def test_tuned_membrane_matches_oracle(oracle_lib, cat_side, tile_rows, cat_walk, cat_lanes):
    """Catalysis walked from either side, over LDS strips (short ones too)
    or work lists: still the oracle's events and positions (FP64, injected
    draws)."""
    cfg = workloads.plain_membrane()
    st, eng, ost, odyn = _pair(cfg, exact=False)
    tune = dict(tile_rows=tile_rows, cat_side=cat_side, cat_walk=cat_walk, cat_lanes=cat_lanes)
    eng.set_tuning(**tune)
    got = eng.tuning()
    assert {k: got[k] for k in tune} == tune
    for k in range(100):
        eng.step()
        odyn.step()
    compare_states(host_arrays(st), oracle_arrays(ost), label="step 100: ")
    oc = odyn.counters()
    assert eng.n_catalytic == oc["n_catalytic"] and oc["n_catalytic"] > 0
    assert (eng.n_recruited, eng.n_detached, eng.n_converted) == (
        oc["n_recruited"], oc["n_detached"], oc["n_converted"])


@pytest.mark.parametrize("feedback", [False, True])
def test_autocatalysis_membrane_fp64(oracle_lib, feedback):
    """PIP3 + PIP2 -> 2 PIP3 (enzyme == product) in the plain membrane: a
    PIP2 converted at pair p is an enzyme for the pairs after p, so
    conversions cascade in the reference's pair order; the oracle runs the
    reference's sequential loop (FP64, injected draws)."""
    cfg = autocatalytic_config(feedback)
    st, eng, ost, odyn = _pair(cfg, exact=False)
    pip3 = [s.name for s in cfg.species].index("PIP3")
    n0 = int(np.sum((st.species_id == pip3) & st.is_alive.astype(bool)))
    for k in range(60):
        eng.step()
        odyn.step()
        if (k + 1) % 20 == 0:
            compare_states(host_arrays(st), oracle_arrays(ost), label=f"step {k + 1}: ")
            oc = odyn.counters()
            assert (eng.n_catalytic, eng.n_recruited, eng.n_detached, eng.n_converted) == (
                oc["n_catalytic"], oc["n_recruited"], oc["n_detached"], oc["n_converted"])
    st.mark_stale()
    n1 = int(np.sum((st.species_id == pip3) & st.is_alive.astype(bool)))
    print(f"autocatalysis feedback={feedback}: PIP3 {n0} -> {n1}, "
          f"{eng.n_catalytic} catalytic events")
    assert n1 > 2 * n0, "the autocatalytic reaction must grow PIP3"


@pytest.mark.parametrize("variant", [dict(), dict(intramolecular=False),
                                     dict(reflective=True, harmonic=True),
                                     dict(intramolecular=False, exact=False)])
def test_binding_fret_full_path_fp64(oracle_lib, variant):
    variant = dict(variant)
    exact = variant.pop("exact", True)
    cfg = workloads.binding_fret(**variant)
    st, eng, ost, odyn = _pair(cfg, exact=exact)
    duty = min(1.0, cfg.scan.dwell_time / cfg.scan.frame_interval)
    ph = PhotonSimulator(cfg, st, duty)
    mic = VirtualMicroscope(cfg, st)
    oph = rm.OraclePhotons(cfg, ost, duty)
    omic = rm.OracleMicroscope(cfg, ost)
    tot_ph = tot_det = 0
    for k in range(25):
        eng.step()
        odyn.step()
        b = ph.evolve_one_bd_step()
        ob = oph.evolve_one_bd_step()
        assert b.n_photons == ob["n_photons"], f"step {k}: photon count"
        if b.n_photons:
            np.testing.assert_array_equal(b.particle_idx, ob["particle_idx"])
            np.testing.assert_array_equal(b.channel, ob["channel"])
            np.testing.assert_array_equal(b.is_fret, ob["is_fret"].astype(bool))
            np.testing.assert_allclose(b.arrival_time, ob["arrival_time"], rtol=1e-12, atol=0)
        nd = mic.process_photons(b)
        ond = omic.process_photons(ob)
        assert nd == ond, f"step {k}: detected photons"
        tot_ph += b.n_photons
        tot_det += nd
    compare_states(host_arrays(st), oracle_arrays(ost), label="final: ")
    oc = odyn.counters()
    assert (eng.n_bindings, eng.n_dissociations) == (oc["n_bindings"], oc["n_dissociations"])
    fr = mic.finalize_frame()
    ofr = omic.finalize_frame()
    assert fr.n_dark_counts == ofr["n_dark"] and fr.n_afterpulses == ofr["n_ap"]
    np.testing.assert_array_equal(fr.flim_stack, ofr["flim_stack"])
    assert tot_ph > 0 and tot_det > 0
    print(f"binding/FRET {variant}: photons {tot_ph}, detected {tot_det}, "
          f"bindings {eng.n_bindings}, counts {fr.total_counts}")


# um after 100 reaction-free steps: (median, 99th percentile, worst) of the
# distance between the FP32 and FP64 positions of a particle
# (measured: plain 1.6e-6 / 1.1e-4 / 8.2e-3, binding_regions 5.5e-7 / 1.1e-5 / 6.0e-5,
# wave 2.0e-6 / 5.9e-3 / 4.5e-2: the 13,000-particle wave system at 130 / um^2
# has more WCA contacts, whose displacement-capped kicks amplify rounding)
FP32_POS_BOUND = {"plain": (1e-5, 1e-3, 5e-2), "binding_regions": (1e-5, 1e-4, 1e-3),
                  "wave": (1e-5, 2e-2, 1e-1)}


@pytest.mark.parametrize("workload", list(FP32_POS_BOUND))
def test_fp32_positions_within_bound(oracle_lib, workload):
    """The FP32 dynamics (forces, Ermak-McCammon step, D(x), boundaries)
    with the reference's numpy noise injected, against the FP64 oracle, over
    100 steps with the reactions removed: the median, 99th percentile and
    worst distance between a particle's FP32 and FP64 positions within
    FP32_POS_BOUND (um).  With reactions, an event whose probability or
    contact test sits within FP32 rounding of its threshold can come out the
    other way and the recruitment cap carries that on; and the dense
    harmonic membrane (k = 1e7 kT/um^2, displacement-capped steps) is
    chaotic: its FP32 and FP64 trajectories part by a median 84 nm within
    100 steps.  Those are compared with the oracle through distributions
    (test_gpu_statistics.py)."""
    if workload == "plain":
        cfg = workloads.plain_membrane()
    elif workload == "wave":
        cfg = workloads.wave_system(n_steps=100)
    else:
        cfg = workloads.binding_fret(seed=4, reflective=True)
    cfg.reactions = []
    cfg.catalytic_reactions = []
    cfg.compartment_exchanges = []
    cfg.unimolecular_conversions = []
    st = ParticleState(cfg)
    eng = DynamicsEngine(cfg, st, precision="fp32", rng="numpy")
    ost = rm.OracleState(cfg)
    odyn = rm.OracleDynamics(cfg, ost)
    for _ in range(100):
        eng.step()
        odyn.step()
    g, o = host_arrays(st), oracle_arrays(ost)
    for k in ("species_id", "is_alive"):
        np.testing.assert_array_equal(np.asarray(g[k]), np.asarray(o[k]), err_msg=k)
    L = np.asarray(cfg.membrane.size, dtype=np.float64)
    d = np.asarray(g["positions"], dtype=np.float64) - o["positions"]
    d -= L * np.round(d / L)
    e = np.sqrt(np.sum(d * d, axis=1))
    p50, p99, worst = (float(v) for v in np.percentile(e, [50, 99, 100]))
    print(f"fp32 {workload}: position error after 100 steps: median {p50:.3e}, 99th percentile "
          f"{p99:.3e}, worst {worst:.3e} um")
    b50, b99, bmax = FP32_POS_BOUND[workload]
    assert p50 <= b50 and p99 <= b99 and worst <= bmax


def _local_config(sigma, k_on=150.0):
    """The plain membrane with density-biased PI3K recruitment
    (use_local_pi3k_recruitment; the PI3K exchange has the single rule
    pip3_activation) and a recruitment rate high enough that every step
    recruits."""
    cfg = workloads.plain_membrane()
    cfg.use_local_pi3k_recruitment = True
    cfg.recruit_sigma = sigma
    for ex in cfg.compartment_exchanges:
        if ex.species == "PI3K":
            ex.k_on = k_on
            ex.k_off = 20.0
    return cfg


@pytest.mark.parametrize("sigma", [0.5, 0.02])
def test_local_recruitment_fp64(oracle_lib, sigma):
    """Density-biased recruitment (dynamics.py:1116-1217, CDF :2355-2370)
    against the oracle with injected draws: the same particles come alive
    at the same sites (positions within 1e-9 relative), the same counts."""
    cfg = _local_config(sigma)
    st, eng, ost, odyn = _pair(cfg, exact=False)
    assert "recruit_local" in eng.plan.kinds
    for k in range(40):
        eng.step()
        odyn.step()
        if (k + 1) % 10 == 0:
            compare_states(host_arrays(st), oracle_arrays(ost), label=f"step {k + 1}: ")
            oc = odyn.counters()
            assert (eng.n_recruited, eng.n_detached, eng.n_catalytic, eng.n_converted) == (
                oc["n_recruited"], oc["n_detached"], oc["n_catalytic"], oc["n_converted"])
    assert eng.n_recruited >= 40
    print(f"local recruitment sigma={sigma}: recruited {eng.n_recruited}")


def test_local_recruitment_rejects_a_bad_reaction():
    """edm_create refuses a density-biased recruitment without its feedback
    rule or with a negative sigma."""
    import ctypes
    from ed_mrdt import _hip
    from ed_mrdt.dynamics import _Plan, build_params, derive_constants
    cfg = _local_config(0.3)
    c = derive_constants(cfg)
    plan = _Plan(cfg, c)
    k = plan.kinds.index("recruit_local")
    for bad in ("no_feedback", "negative_sigma"):
        p2 = _Plan(cfg, c)
        r = p2.reactions[k]
        if bad == "no_feedback":
            r.fb1 = _hip.no_feedback()
        else:
            r.contact = -1.0
        built = build_params(cfg, c, p2, ParticleState(cfg).N)
        h = ctypes.c_void_p()
        rc = _hip.lib().edm_create(ctypes.byref(built.params), ctypes.byref(h))
        assert rc == -1 and not h.value, bad
        assert b"local recruitment" in _hip.lib().edm_last_error()


def _recruit_sites(local, steps=30):
    """Philox run with every PIP3 moved into x < L/4 first: the x of every
    PI3K that comes alive, at the step it does (recruitment is the step's
    last move of a particle)."""
    cfg = _local_config(0.1)
    cfg.use_local_pi3k_recruitment = local
    st = ParticleState(cfg)
    names = [s.name for s in cfg.species]
    pip3 = st.species_id == names.index("PIP3")
    Lx = cfg.membrane.size[0]
    st.positions[pip3, 0] *= 0.25
    eng = DynamicsEngine(cfg, st, precision="fp64", rng="philox")
    pi3k = st.species_id == names.index("PI3K")
    alive = st.is_alive.copy()
    xs = []
    for _ in range(steps):
        eng.step()
        now = st.is_alive.copy()
        new = now & ~alive & (st.species_id == names.index("PI3K"))
        xs.extend(st.positions[new, 0].tolist())
        alive = now
    assert pi3k.sum() > 0
    return np.array(xs), Lx


def test_local_recruitment_philox_follows_the_sensor():
    """With the sensor (PIP3) confined to x < L/4, density-biased recruits
    land there (within a few sigma); uniform recruits spread over the box."""
    xs, L = _recruit_sites(True)
    band = (xs < 0.25 * L + 0.5) | (xs > L - 0.5)
    assert xs.size >= 30 and band.mean() > 0.9, (xs.size, band.mean())
    xu, _ = _recruit_sites(False)
    bu = (xu < 0.25 * L + 0.5) | (xu > L - 0.5)
    assert xu.size >= 30 and bu.mean() < 0.6, (xu.size, bu.mean())
    print(f"philox local: {band.mean():.3f} of {xs.size} in the band; uniform {bu.mean():.3f}")


def test_fp32_philox_large_runs():
    cfg = workloads.large_membrane(side=20.0, n_alive=40_000, pixels=128)
    st = ParticleState(cfg)
    eng = DynamicsEngine(cfg, st, precision="fp32", rng="philox", pair_cap=False)
    ph = PhotonSimulator(cfg, st, min(1.0, cfg.scan.dwell_time / cfg.scan.frame_interval))
    mic = VirtualMicroscope(cfg, st)
    n0 = int(st.is_alive.sum())
    for _ in range(50):
        eng.step()
        b = ph.evolve_one_bd_step()
        mic.process_photons(b)
    pos = st.positions
    assert np.all(np.isfinite(pos))
    assert np.all((pos >= 0) & (pos <= 20.0))
    n1 = int(st.is_alive.sum())
    assert abs(n1 - n0) < 0.2 * n0
    fr = mic.finalize_frame()
    assert fr.total_counts > 0
    print("fp32 philox: catalytic", eng.n_catalytic, "detached", eng.n_detached,
          "recruited", eng.n_recruited, "counts", fr.total_counts)
