"""Pins of the oracle's compositions and of its PSS photon kernel.

The oracle restates the reference's step, photon step and microscope from
the source text.  These tests pin what no single-kernel known answer
reaches:

* the PSS photon kernel (photophysics.py:339-483) on a hand-built set of
  emitters, with the expected photon list written out by hand: a FRET
  donor with its acceptor, an emitter in T1, one sitting in the bleached
  CTMC state, a dead one, one without a dye type, one already bleached,
  the max_photons cut and the exhaustion of the shared arrival array;
* the random-draw schedule: after ParticleState's init, one
  DynamicsEngine.step, one PhotonSimulator.evolve_one_bd_step and a
  microscope's process_photons / finalize_frame, each generator's state
  must equal a fresh numpy generator advanced by the calls the reference
  makes (particles.py:94-243, dynamics.py:2194-2438, photophysics.py:
  1136-1145, microscope.py:848, 371, 743, 859-920), with the pair count
  taken from an independent brute-force count;
* the reference's statistical photophysics and detector tests, restated
  with expectations computed independently from the CTMC tables
  (photon counts and bleaching from powers of the substochastic S0/T1
  block, exponential arrival times, FRET-quenched lifetime, dark counts
  and afterpulses).

Everything runs on the CPU.
"""

from __future__ import annotations

import ctypes
import math

import numpy as np
import pytest
from scipy import stats

import reference_model as rm
from ed_mrdt import workloads
from kernels import _poisson_draw

P, I, D = rm.P, rm.I, rm.D


@pytest.fixture(autouse=True, scope="module")
def _built(oracle_lib):
    return oracle_lib


# ----------------------------------------------------------------------
# PSS photon kernel, hand-built known answer (photophysics.py:339-483)
# ----------------------------------------------------------------------

def _pss_case():
    nf = 2
    # rows: from S0 (pss 0), from T1 (pss 1); columns: to S0, to T1, (rest: bleach)
    T = np.zeros((nf, 3, 3))
    Tf = np.zeros((nf, 3, 3))
    T[0, 0], T[0, 1] = [0.90, 0.06, 0.04], [0.30, 0.65, 0.05]
    Tf[0, 0], Tf[0, 1] = [0.80, 0.15, 0.05], [0.20, 0.70, 0.10]
    T[1, 0], T[1, 1] = [0.97, 0.02, 0.01], [0.50, 0.45, 0.05]
    Tf[1] = T[1]
    mean_ph = np.array([[3.0, 0.5], [2.6, 0.1]])
    mean_ph_f = np.array([[1.2, 0.2], [2.6, 0.1]])
    mean_fr = np.array([[0.0, 0.0], [0.0, 0.0]])
    mean_fr_f = np.array([[4.0, 0.0], [0.0, 0.0]])
    tau = np.array([2.5e-9, 3.0e-9])
    k_bound = np.array([2.0e9, 0.0])
    gs, es, t1, bl = (np.array(v, np.int32) for v in ([0, 0], [1, 1], [2, 2], [3, 3]))
    return nf, T, Tf, mean_ph, mean_ph_f, mean_fr, mean_fr_f, tau, k_bound, gs, es, t1, bl


def _seed_for(mean, want, fret=False):
    """A uniform seed whose Poisson draw (photophysics.py:487-521) is want."""
    for k in range(1, 5000):
        u = k / 5000.0
        if _poisson_draw(mean, u * 0.5 + 0.25 if fret else u) == want:
            return u
    raise AssertionError("no seed")


def test_pss_kernel_hand_built_emitters():
    nf, T, Tf, mph, mphf, mfr, mfrf, tau, kb, gs, es, t1, bl = _pss_case()
    N = 7
    ctmc = np.array([0, 2, 3, 0, 1, 0, 0], np.int8)
    bleached = np.array([0, 0, 0, 0, 0, 0, 1], np.uint8)
    has_dye = np.array([1, 1, 1, 1, 1, 1, 1], np.uint8)
    alive = np.array([1, 1, 1, 0, 1, 1, 1], np.uint8)
    dtype = np.array([0, 0, 0, 0, 1, -1, 0], np.int32)
    k_fret = np.zeros(N)
    k_fret[0] = 1.0e9  # frac = k / k_bound = 0.5
    acc = np.full(N, -1, np.int32)
    acc[0] = 4
    # emitter 0: mean_d = 0.5*3.0 + 0.5*1.2 = 2.1, mean_f = 0.5*0 + 0.5*4.0 = 2.0
    u0 = _seed_for(2.1, 3)
    # the FRET draw reuses the seed as u*0.5 + 0.25: pick u0 so it gives 2 there as well
    for k in range(1, 5000):
        u = k / 5000.0
        if _poisson_draw(2.1, u) == 3 and _poisson_draw(2.0, u * 0.5 + 0.25) == 2:
            u0 = u
            break
    u1 = _seed_for(0.5, 1)   # emitter 1 from T1: mean_ph[0, 1] = 0.5
    u4 = _seed_for(2.6, 4)   # emitter 4 (type 1, S1 -> pss 0): mean 2.6
    u_p = np.array([u0, u1, 0.5, 0.5, u4, 0.5, 0.5])
    u_t = np.array([0.5, 0.5, 0.5, 0.5, 0.995, 0.5, 0.5])
    u_a = np.array([0.11, 0.52, 0.93, 0.0, 0.37, 0.64, 0.28, 0.75])  # 8: runs out
    max_photons = 9
    op = np.full(16, -9, np.int32)
    ot = np.zeros(16)
    oc = np.zeros(16, np.int8)
    opl = np.zeros(16, np.int32)
    of = np.zeros(16, np.uint8)
    n = rm.lib().orc_pss_kernel(
        P(ctmc), P(bleached), P(has_dye), P(alive), P(dtype), P(k_fret), P(acc), I(nf),
        P(T), P(Tf), P(mph), P(mphf), P(mfr), P(mfrf), P(tau), P(kb), P(gs), P(es), P(t1),
        P(bl), P(op), P(ot), P(oc), P(opl), P(of), I(max_photons), I(N), P(u_p), P(u_t),
        P(u_a), I(len(u_a)))
    tau0 = 2.5e-9 / (1.0 + 1.0e9 * 2.5e-9)  # tau_DA of emitter 0
    want = [  # (particle, time, channel, is_fret), in emission order
        (0, -tau0 * math.log(0.11), 0, 0),
        (0, -tau0 * math.log(0.52), 0, 0),
        (0, -tau0 * math.log(0.93), 0, 0),
        (4, -3.0e-9 * math.log(1e-30), 1, 1),  # u = 0 clamped to 1e-30, acceptor's tau
        (4, -3.0e-9 * math.log(0.37), 1, 1),
        (1, -2.5e-9 * math.log(0.64), 0, 0),
        (4, -3.0e-9 * math.log(0.28), 1, 0),
        (4, -3.0e-9 * math.log(0.75), 1, 0),
        (4, 3.0e-9, 1, 0),  # arrival array exhausted: t = tau
    ]  # the 10th photon (emitter 4's fourth) is cut by max_photons = 9
    assert n == len(want) == max_photons
    np.testing.assert_array_equal(op[:n], [w[0] for w in want])
    np.testing.assert_allclose(ot[:n], [w[1] for w in want], rtol=1e-15, atol=0)
    np.testing.assert_array_equal(oc[:n], [w[2] for w in want])
    np.testing.assert_array_equal(of[:n], [w[3] for w in want])
    # CTMC: emitter 0 stays S0 (0.5 < T00 = 0.85); 1 stays T1 (0.30 <= 0.5 <
    # 0.95); 2 sat in the bleached state -> flagged, no transition; 4 bleaches
    # (0.995 >= 0.99); the dead, untyped and bleached ones are untouched
    assert ctmc.tolist() == [0, 2, 3, 0, 3, 0, 0]
    assert bleached.tolist() == [0, 0, 1, 0, 1, 0, 1]


def test_pss_kernel_transition_thresholds_and_interpolation():
    """T00 / T01 interpolated by frac = min(k / k_bound, 1) with the T01
    column of both tables (photophysics.py:419-427): thresholds probed just
    below and above each boundary."""
    nf, T, Tf, mph, mphf, mfr, mfrf, tau, kb, gs, es, t1, bl = _pss_case()
    for k, frac in ((0.0, 0.0), (1.0e9, 0.5), (5.0e9, 1.0)):
        T00 = T[0, 0, 0] * (1 - frac) + Tf[0, 0, 0] * frac
        T01 = T[0, 0, 1] * (1 - frac) + Tf[0, 0, 1] * frac
        for u, want in ((T00 - 1e-9, 0), (T00 + 1e-9, 2), (T00 + T01 - 1e-9, 2),
                        (T00 + T01 + 1e-9, 3)):
            ctmc = np.zeros(1, np.int8)
            bleached = np.zeros(1, np.uint8)
            buf = [np.zeros(4, np.int32), np.zeros(4), np.zeros(4, np.int8),
                   np.zeros(4, np.int32), np.zeros(4, np.uint8)]
            rm.lib().orc_pss_kernel(
                P(ctmc), P(bleached), P(np.ones(1, np.uint8)), P(np.ones(1, np.uint8)),
                P(np.zeros(1, np.int32)), P(np.array([k])), P(np.array([-1], np.int32)), I(nf),
                P(T), P(Tf), P(mph * 0), P(mphf * 0), P(mfr), P(mfrf), P(tau), P(kb), P(gs),
                P(es), P(t1), P(bl), *[P(b) for b in buf], I(4), I(1), P(np.array([0.5])),
                P(np.array([u])), P(np.zeros(1)), I(0))
            assert ctmc[0] == want and bleached[0] == (want == 3), (k, u)


# ----------------------------------------------------------------------
# Draw schedules
# ----------------------------------------------------------------------

def _same_state(a: np.random.Generator, b: np.random.Generator):
    assert a.bit_generator.state == b.bit_generator.state


def _brute_pairs(pos, L, cutoff, periodic):
    """Number of pairs i < j with d^2 < cutoff^2 (minimum image), by an
    O(N^2) numpy count."""
    n = len(pos)
    tot = 0
    for i0 in range(0, n, 512):
        d = pos[i0:i0 + 512, None, :] - pos[None, :, :]
        for ax in range(2):
            if periodic[ax]:
                s = d[..., ax]
                s[s > 0.5 * L[ax]] -= L[ax]
                s[s < -0.5 * L[ax]] += L[ax]
        d2 = np.sum(d * d, axis=-1)
        ii = np.arange(i0, min(n, i0 + 512))[:, None]
        jj = np.arange(n)[None, :]
        tot += int(np.count_nonzero((jj > ii) & (d2 < cutoff * cutoff)))
    return tot


def test_particle_state_init_draws():
    """particles.py:94-243: per species in config order x then y uniforms,
    then each cytosolic pool (exchange order), then per dye attachment one
    uniform per particle of that species (pools included)."""
    cfg = workloads.binding_fret(intramolecular=False)
    st = rm.OracleState(cfg, rng=np.random.default_rng(99))
    g = np.random.default_rng(99)
    Lx, Ly = cfg.membrane.size
    xs, sp = [], []
    names = [s.name for s in cfg.species]
    for k, s in enumerate(cfg.species):
        c = cfg.initial_populations.get(s.name, 0)
        xs.append(np.stack([g.uniform(0, Lx, c), g.uniform(0, Ly, c)], axis=1))
        sp += [k] * c
    for ex in cfg.compartment_exchanges:
        if ex.cytosol_pool > 0:
            xs.append(np.stack([g.uniform(0, Lx, ex.cytosol_pool),
                                g.uniform(0, Ly, ex.cytosol_pool)], axis=1))
            sp += [names.index(ex.species)] * ex.cytosol_pool
    pos = np.concatenate(xs)
    sp = np.array(sp)
    np.testing.assert_array_equal(st.positions, pos)
    np.testing.assert_array_equal(st.species_id, sp)
    dye = np.zeros(len(sp), bool)
    for at in cfg.dye_attachments:
        idx = np.where(sp == names.index(at.species))[0]
        dye[idx[g.random(len(idx)) < at.labeling_efficiency]] = True
    np.testing.assert_array_equal(st.has_dye.astype(bool), dye)
    assert 0 < dye.sum() < len(dye)
    # nothing else was drawn
    g2 = np.random.default_rng(99)
    rm.OracleState(cfg, rng=g2)
    _same_state(g, g2)


def _replay_step(g, cfg, n, n_pairs, local=False):
    """dynamics.py:2194-2438: the draw calls of one step."""
    g.standard_normal((n, 2))
    for _ in cfg.reactions:
        g.random(n_pairs)
        g.random(n)
    for _ in cfg.catalytic_reactions:
        g.random(n_pairs)
    for ex in cfg.compartment_exchanges:
        g.random(n)
        g.random((n, 2))
        if local and ex.feedback_rule and "pip3_activation" in ex.feedback_rule \
                and not ex.secondary_feedback_rule:
            g.random(n)
            g.standard_normal((n, 2))
        g.random(n)
    for _ in cfg.unimolecular_conversions:
        g.random(n)


def _dynamics_replay(cfg, n, n_pairs, local=False):
    g = np.random.default_rng(cfg.random_seed)
    _replay_step(g, cfg, n, n_pairs, local)
    return g


@pytest.mark.parametrize("case", ["plain", "binding", "local"])
def test_dynamics_step_draws(case):
    if case == "plain":
        cfg = workloads.plain_membrane()
    elif case == "binding":
        cfg = workloads.binding_fret()
    else:
        cfg = workloads.plain_membrane()
        cfg.use_local_pi3k_recruitment = True
        cfg.recruit_sigma = 0.3
    st = rm.OracleState(cfg)
    dyn = rm.OracleDynamics(cfg, st)
    L = np.asarray(cfg.membrane.size, dtype=float)
    n_pairs = min(_brute_pairs(st.positions, L, dyn.cutoff, cfg.membrane.periodic),
                  min(30 * st.N, 500000))
    dyn.step()
    assert dyn.last_pairs.shape[0] == n_pairs
    _same_state(dyn.rng, _dynamics_replay(cfg, st.N, n_pairs, local=case == "local"))
    # a second step, from the moved positions: both steps' calls in order
    n2 = min(_brute_pairs(st.positions, L, dyn.cutoff, cfg.membrane.periodic),
             min(30 * st.N, 500000))
    dyn.step()
    g2 = np.random.default_rng(cfg.random_seed)
    for npairs in (n_pairs, n2):
        _replay_step(g2, cfg, st.N, npairs, local=case == "local")
    _same_state(dyn.rng, g2)


def test_photon_step_draws():
    """photophysics.py:1136-1145: random(N), random(N), then random(est) with
    est = max(1000, int(N * max_t max(mean_ph[t, 0], mean_ph_fret[t, 0]) * 2.5))."""
    cfg = workloads.binding_fret(intramolecular=False)
    st = rm.OracleState(cfg)
    duty = min(1.0, cfg.scan.dwell_time / cfg.scan.frame_interval)
    ph = rm.OraclePhotons(cfg, st, duty)
    g = np.random.default_rng(cfg.random_seed + 1000)
    mm = max(max(float(ph.mean_ph[t, 0]), float(ph.mean_ph_f[t, 0])) for t in range(ph.nf))
    for _ in range(3):
        ph.evolve_one_bd_step()
        g.random(st.N)
        g.random(st.N)
        g.random(max(1000, int(st.N * mm * 2.5)))
    _same_state(ph.rng, g)


def test_microscope_draws_and_frame_noise():
    """microscope.py:1069-1203: per step random(n) for QE, random(n_survived)
    for the pixel draw, normal(0, sigma_irf, n_assigned) for the IRF (none
    when nothing survives); at the frame end the dark-count and afterpulse
    loops, replayed here on a copy of the stack."""
    cfg = workloads.binding_fret(intramolecular=False)
    st = rm.OracleState(cfg)
    dyn = rm.OracleDynamics(cfg, st)
    duty = min(1.0, cfg.scan.dwell_time / cfg.scan.frame_interval)
    ph = rm.OraclePhotons(cfg, st, duty)
    mic = rm.OracleMicroscope(cfg, st)
    g = np.random.default_rng(cfg.random_seed + 2000)
    for _ in range(6):
        dyn.step()
        b = ph.evolve_one_bd_step()
        na = mic.process_photons(b)
        n = b["n_photons"]
        if n == 0:
            continue
        ns = int(np.count_nonzero(g.random(n) < mic.qe))
        if ns == 0:
            continue
        g.random(ns)
        if na:
            g.normal(0.0, mic.sigma_irf, size=na)
    _same_state(mic.rng, g)
    stack = mic.flim.copy()
    out = mic.finalize_frame()
    nc, nx, ny, nb = stack.shape
    lam = mic.dark_rate * mic.dwell
    n_dark = n_ap = 0
    for c in range(nc):
        for i in range(nx):
            for j in range(ny):
                k = g.poisson(lam)
                if k > 0:
                    np.add.at(stack[c, i, j], g.integers(0, nb, size=k), 1)
                    n_dark += k
    for c in range(nc):
        for i in range(nx):
            for j in range(ny):
                tot = int(stack[c, i, j].sum())
                if tot == 0:
                    continue
                k = g.binomial(tot, mic.p_ap)
                if k > 0:
                    np.add.at(stack[c, i, j], g.integers(0, nb, size=k), 1)
                    n_ap += k
    np.testing.assert_array_equal(out["flim_stack"], stack)
    assert (out["n_dark"], out["n_ap"]) == (n_dark, n_ap)
    _same_state(mic.rng, g)


# ----------------------------------------------------------------------
# Statistical restatements (reference tests/test_photophysics.py:273-458,
# tests/test_microscope.py:284-303)
# ----------------------------------------------------------------------

def _donor_config(bleach=None, n=4000):
    cfg = workloads.plain_membrane(seed=5)
    if bleach is not None:
        for f in cfg.fluorophores:
            if f.name == "mTFP1":
                f.thermal_rates["S1"]["bleached"] = bleach
    # every PIP3 carries mTFP1; no binding, so no FRET
    cfg.initial_populations["PIP3"] = n
    cfg.reactions = []
    return cfg


def _ctmc_block(ph, t=0):
    T = ph.T_step[t]
    return np.array([[T[0, 0], T[0, 1]], [T[1, 0], T[1, 1]]])


def test_photon_counts_and_bleaching_follow_the_ctmc_tables():
    """Per step, the expected photons and surviving emitters of a population
    started in S0 follow the substochastic S0/T1 block M of T_step:
    alive_k = N0 e0 M^k 1, photons_k = N0 e0 M^k (mean_ph[t, 0], mean_ph[t, 1]).
    A fast-bleaching mTFP1 (S1 -> bleached 6e6 /s) makes the decay visible
    in 60 steps; every step is checked within 5 sigma (Poisson / binomial)."""
    cfg = _donor_config(bleach=6.0e6)
    st = rm.OracleState(cfg)
    duty = min(1.0, cfg.scan.dwell_time / cfg.scan.frame_interval)
    ph = rm.OraclePhotons(cfg, st, duty)
    donors = (st.has_dye == 1) & (st.is_alive == 1) & (st.dye_type_id == 0)
    n0 = int(donors.sum())
    M = _ctmc_block(ph)
    means = np.array([ph.mean_ph[0, 0], ph.mean_ph[0, 1]])
    v = np.array([1.0, 0.0])
    alive_hist, want_alive, ph_hist, want_ph = [], [], [], []
    for k in range(60):
        b = ph.evolve_one_bd_step()
        ph_hist.append(int(np.count_nonzero(b["channel"] == 0)))
        want_ph.append(n0 * float(v @ means))
        v = v @ M
        alive_hist.append(int(np.count_nonzero(donors & (st.is_bleached == 0))))
        want_alive.append(n0 * float(v.sum()))
    ph_hist, want_ph = np.array(ph_hist), np.array(want_ph)
    a, wa = np.array(alive_hist), np.array(want_alive)
    assert np.all(np.abs(ph_hist - want_ph) < 5 * np.sqrt(want_ph) + 1), (ph_hist, want_ph)
    p = wa / n0
    assert np.all(np.abs(a - wa) < 5 * np.sqrt(n0 * p * (1 - p)) + 1), (a, wa)
    assert a[-1] < 0.9 * n0  # bleaching is visible
    # log-linear decay, as the reference's bleaching test fits it
    slope = np.polyfit(np.arange(60), np.log(a / n0), 1)[0]
    lam = max(abs(np.linalg.eigvals(M)))
    assert slope == pytest.approx(math.log(lam), rel=0.15)


def test_donor_arrival_times_are_exponential_with_the_s1_lifetime():
    """photophysics.py:431-441 (B5): without FRET each donor photon's
    arrival time ~ Exp(tau_S1), tau_S1 = 1 / (k_S1->S0 + k_S1->T1 + k_S1->bl)."""
    cfg = _donor_config()
    st = rm.OracleState(cfg)
    ph = rm.OraclePhotons(cfg, st, 1.0)
    t = np.concatenate([ph.evolve_one_bd_step()["arrival_time"] for _ in range(8)])
    r = cfg.fluorophores[0].thermal_rates["S1"]
    tau_s1 = 1.0 / sum(r.values())
    assert len(t) > 5000
    assert np.mean(t) == pytest.approx(tau_s1, rel=4 / math.sqrt(len(t)))
    assert stats.kstest(t, "expon", args=(0, tau_s1)).pvalue > 0.001


def test_fret_quenches_the_donor_lifetime_and_feeds_the_acceptor():
    """photophysics.py:405-471 (B6): with k_FRET injected for every donor
    the donor photons arrive ~ Exp(tau_D / (1 + k tau_D)), their number falls
    to the frac-interpolated mean, and acceptor (FRET) photons appear with
    the acceptor's lifetime."""
    cfg = workloads.binding_fret(intramolecular=False)
    cfg.initial_populations["PHAkt"] = 3000
    st = rm.OracleState(cfg)
    ph = rm.OraclePhotons(cfg, st, 1.0)
    donors = (st.has_dye == 1) & (st.is_alive == 1) & (st.dye_type_id == 0)
    acc_i = int(np.where((st.dye_type_id == 1) & (st.has_dye == 1))[0][0])
    tau_d = ph.tau[0]
    k = 1.0 / tau_d  # r = R0
    frac = min(k / ph.k_bound[0], 1.0)

    def inject():
        ph.k_fret[:] = 0.0
        ph.acc[:] = -1
        ph.k_fret[donors] = k
        ph.acc[donors] = acc_i

    ph.update_fret_rates = inject
    bs = [ph.evolve_one_bd_step() for _ in range(3)]
    don = np.concatenate([b["arrival_time"][(b["channel"] == 0) & (b["is_fret"] == 0)]
                          for b in bs])
    fr = np.concatenate([b["arrival_time"][b["is_fret"] == 1] for b in bs])
    tau_da = tau_d / (1.0 + k * tau_d)
    assert tau_da == pytest.approx(tau_d / 2, rel=1e-12)
    assert np.mean(don) == pytest.approx(tau_da, rel=5 / math.sqrt(len(don)))
    assert stats.kstest(don, "expon", args=(0, tau_da)).pvalue > 0.001
    assert len(fr) > 100
    assert np.mean(fr) == pytest.approx(ph.tau[1], rel=5 / math.sqrt(len(fr)))
    # first step: all donors in S0, so E[donor photons] = n * interpolated mean
    md = ph.mean_ph[0, 0] * (1 - frac) + ph.mean_ph_f[0, 0] * frac
    n_first = int(np.count_nonzero((bs[0]["channel"] == 0) & (bs[0]["is_fret"] == 0)))
    want = int(donors.sum()) * md
    assert abs(n_first - want) < 5 * math.sqrt(want) + 1


def test_dark_counts_and_afterpulses_rates():
    """microscope.py:859-920 (reference test_microscope.py:284-303): dark
    counts ~ Poisson(rate x dwell) per pixel, spread over the bins;
    afterpulses ~ Binomial(pixel total, p_ap) in the pixel that had the
    counts."""
    cfg = workloads.plain_membrane()
    cfg.detector.dark_count_rate = 2.0e4  # 0.4 expected per pixel at 20 us
    st = rm.OracleState(cfg)
    mic = rm.OracleMicroscope(cfg, st)
    mic.flim[0, 10, 10, 50] = 10000
    out = mic.finalize_frame()
    n_pix = mic.flim.shape[0] * mic.nx * mic.ny
    lam = 2.0e4 * cfg.scan.dwell_time
    assert abs(out["n_dark"] - n_pix * lam) < 5 * math.sqrt(n_pix * lam)
    s = out["flim_stack"].astype(np.int64)
    s[0, 10, 10, 50] -= 10000
    # the afterpulses of the big pixel land in that pixel
    big = int(s[0, 10, 10].sum())
    exp_ap = 10000 * mic.p_ap
    assert out["n_ap"] >= big - 3  # a few more from dark-count pixels
    assert abs(big - exp_ap) < 5 * math.sqrt(exp_ap) + 3
    # dark counts are spread over the bins: a chi-square over 8 bin groups
    flat = s.reshape(-1, 256).sum(axis=0)
    grp = flat.reshape(8, 32).sum(axis=1)
    assert stats.chisquare(grp).pvalue > 1e-4
