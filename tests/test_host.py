"""Host-side logic of the product package, without a GPU.

The numbers that the reference's configuration fixes in closed form (cell
cutoff, displacement cap, Foerster efficiency and tau_DA, PSF width, PSF
window, IRF width, pulses per step) are checked against hand-derived
values; the product's constants are checked against the oracle's
independent restatement; the PSS tables against numpy's matrix power and an
explicit geometric sum; the particle initialisation bit for bit against the
oracle's (reference particles.py:94-243 draw order).
"""

from __future__ import annotations

import math

import numpy as np
import pytest

import reference_model as rm
from ed_mrdt import workloads
from ed_mrdt.config import ConfigError, config_from_dict
from ed_mrdt import _hip
from ed_mrdt.dynamics import _Plan, binding_probability, derive_constants
from ed_mrdt.microscope import PSF, ConfocalScanner, TCSPCAccumulator
from ed_mrdt.particles import ParticleState
from ed_mrdt.photophysics import RateMatrix, build_pss_matrices, fret_efficiency, fret_rate


@pytest.fixture(scope="module")
def c0():
    return workloads.plain_membrane()


# ----------------------------------------------------------------------
# Derived dynamics constants (reference dynamics.py:1724-2148)
# ----------------------------------------------------------------------

def test_plain_membrane_cell_grid_and_cap(c0):
    c = derive_constants(c0)
    # cutoff = 1.1 * max(contact 150 nm, sqrt(4 (D_E + D_S) dt), WCA cutoff)
    assert c["cutoff"] == pytest.approx(0.165, rel=1e-12)
    assert (c["nx"], c["ny"]) == (60, 60)
    # WCA present: cap = max(0.5 sigma_min, largest catalytic contact) = 0.15 um
    assert c["max_disp"] == pytest.approx(0.15, rel=1e-12)
    assert c["use_wca"] and not c["use_harmonic"]
    assert c["sigma"][0, 0] == pytest.approx(0.010) and c["epsilon"][0, 0] == 1.0


def test_large_membrane_constants():
    cfg = workloads.large_membrane(0)
    c = derive_constants(cfg)
    assert c["cutoff"] == pytest.approx(0.165, rel=1e-12)
    assert (c["nx"], c["ny"]) == (606, 606)
    assert c["use_harmonic"] and not c["use_wca"]
    assert c["k_harm"][0, 0] == pytest.approx(1e7)  # 10 kT/nm^2 in kT/um^2
    # harmonic: cap = max(5 sqrt(2 D_max dt), contact_max, sigma_min) = 0.15 um
    assert c["max_disp"] == pytest.approx(max(5 * math.sqrt(2 * 0.15 * 1e-3), 0.15, 0.01))


@pytest.mark.parametrize("name", ["plain_membrane", "wave_system", "binding_fret"])
def test_constants_match_oracle_restatement(name):
    cfg = getattr(workloads, name)()
    c = derive_constants(cfg)
    o = rm.OracleDynamics(cfg, rm.OracleState(cfg))
    assert c["cutoff"] == o.cutoff and (c["nx"], c["ny"]) == (o.nx, o.ny)
    assert c["max_disp"] == o.max_disp
    np.testing.assert_array_equal(c["sigma"], o.sig)
    np.testing.assert_array_equal(c["epsilon"], o.eps)
    np.testing.assert_array_equal(c["k_harm"], o.kh)
    for r, ob in zip([r for r in cfg.reactions if r.is_bimolecular], o.bind):
        s_eff, pb = binding_probability(cfg, r, cfg.bd_timestep)
        assert s_eff == ob["contact"] and pb == ob["prob"]


def test_binding_probability_collins_kimball():
    cfg = workloads.binding_fret()
    r = [r for r in cfg.reactions if r.is_bimolecular][0]
    sp = {s.name: s for s in cfg.species}
    Drel = sp[r.reactants[0]].diffusion_coefficient + sp[r.reactants[1]].diffusion_coefficient
    dt = cfg.bd_timestep
    contact = r.contact_radius / 1000.0
    s_eff = max(contact, math.sqrt(4 * Drel * dt))
    cb = cfg.initial_populations[r.reactants[1]] / (cfg.membrane.size[0] * cfg.membrane.size[1])
    R = min(1 / math.sqrt(math.pi * cb), min(cfg.membrane.size) / 2)
    k_int = 1.0 / (1.0 / r.kon - math.log(R / contact) / (2 * math.pi * Drel))
    got_s, got_p = binding_probability(cfg, r, dt)
    assert got_s == pytest.approx(s_eff, rel=1e-15)
    assert got_p == pytest.approx(min(1.0, k_int * dt / (math.pi * s_eff ** 2)), rel=1e-14)
    assert 0.0 < got_p <= 1.0


# ----------------------------------------------------------------------
# Particle initialisation (reference particles.py:94-243)
# ----------------------------------------------------------------------

@pytest.mark.parametrize("name", ["plain_membrane", "binding_fret"])
def test_particle_state_matches_oracle(name):
    cfg = getattr(workloads, name)()
    st = ParticleState(cfg)
    o = rm.OracleState(cfg)
    assert st.N == o.N
    np.testing.assert_array_equal(st.positions, o.positions)
    np.testing.assert_array_equal(st.species_id, o.species_id)
    np.testing.assert_array_equal(st.is_alive.astype(np.uint8), o.is_alive)
    np.testing.assert_array_equal(st.has_dye.astype(np.uint8), o.has_dye)
    np.testing.assert_array_equal(st.dye_type_id, o.dye_type_id)


def test_particle_state_counts_and_dtypes(c0):
    st = ParticleState(c0)
    n_init = sum(c0.initial_populations.values())
    n_pool = sum(ex.cytosol_pool for ex in c0.compartment_exchanges if ex.cytosol_pool > 0)
    assert st.N == n_init + n_pool == 4000
    assert int(st.is_alive.sum()) == n_init == 3500
    Lx, Ly = c0.membrane.size
    assert (st.positions >= 0).all() and (st.positions[:, 0] < Lx).all()
    assert (st.positions[:, 1] < Ly).all()
    assert st.positions.dtype == np.float64 and st.positions.shape == (st.N, 2)
    assert st.species_id.dtype == np.int32 and st.bond_partner.dtype == np.int32
    assert st.ctmc_state.dtype == np.int8 and st.is_alive.dtype == np.bool_
    assert ((st.species_id >= 0) & (st.species_id < len(c0.species))).all()
    a = ParticleState(c0)
    np.testing.assert_array_equal(a.positions, st.positions)  # same seed, same state


# ----------------------------------------------------------------------
# Photophysics host tables (reference photophysics.py:59-331)
# ----------------------------------------------------------------------

def test_fret_closed_forms():
    R0 = 5.4
    assert fret_efficiency(R0, R0) == pytest.approx(0.5, abs=1e-15)
    assert fret_efficiency(0.0, R0) == 1.0
    assert fret_efficiency(100.0, R0) < 1e-6
    E = fret_efficiency(5.0, R0)
    assert E == pytest.approx(0.6134, abs=5e-5)
    tau_D = 2.793e-9
    assert tau_D * (1 - E) == pytest.approx(1.080e-9, abs=5e-13)
    r = np.linspace(0.5, 30, 100)
    e = [fret_efficiency(v, R0) for v in r]
    assert all(b <= a for a, b in zip(e, e[1:]))
    for rr in (3.0, 5.0, 8.0):
        k = fret_rate(rr, R0, tau_D)
        assert fret_efficiency(rr, R0) == pytest.approx(k * tau_D / (1 + k * tau_D), rel=1e-12)
        assert tau_D * (1 - fret_efficiency(rr, R0)) == pytest.approx(
            1 / (1 / tau_D + k), rel=1e-12)


def _rm(cfg):
    return RateMatrix(cfg.fluorophores[0])


def test_pss_tables_stochastic_and_matrix_power(c0):
    rmx = _rm(c0)
    T_laser = 1.0 / c0.laser.repetition_rate
    n = max(1, int(round(c0.bd_timestep / T_laser)))
    assert n == 40000
    one = build_pss_matrices(rmx, 0.01, T_laser, 1)["T_step"]
    full = build_pss_matrices(rmx, 0.01, T_laser, n)["T_step"]
    np.testing.assert_allclose(one.sum(axis=1), 1.0, rtol=0, atol=1e-12)
    np.testing.assert_allclose(full.sum(axis=1), 1.0, rtol=0, atol=1e-9)
    np.testing.assert_allclose(full, np.linalg.matrix_power(one, n), rtol=1e-12, atol=1e-15)
    assert full[2].tolist() == [0.0, 0.0, 1.0]  # bleaching is absorbing


def test_pss_mean_photons_geometric_sum(c0):
    rmx = _rm(c0)
    T_laser = 1.0 / c0.laser.repetition_rate
    p_exc, n = 0.02, 7
    out = build_pss_matrices(rmx, p_exc, T_laser, n, QY_donor=0.4)
    T = build_pss_matrices(rmx, p_exc, T_laser, 1)["T_step"][:2, :2]
    S = sum(np.linalg.matrix_power(T, k) for k in range(n))  # explicit series
    kS1 = rmx.exit_rates[rmx.excited_state]
    p_rad = rmx.Q[rmx.excited_state, rmx.ground_state] / kS1
    # (I - T^n)(I - T)^-1 and the explicit sum agree to rounding of the inverse
    np.testing.assert_allclose(out["mean_ph"], S[:, 0] * p_exc * p_rad * 0.4, rtol=1e-7)
    assert out["tau_eff"] == pytest.approx(1.0 / kS1, rel=1e-15)


def test_pss_fret_quenches_donor(c0):
    rmx = _rm(c0)
    T_laser = 1.0 / c0.laser.repetition_rate
    n = 40000
    tau_D = rmx.get_lifetime(rmx.excited_state)
    assert tau_D == pytest.approx(2.793e-9, rel=1e-3)
    a = build_pss_matrices(rmx, 0.01, T_laser, n, k_fret=0.0, QY_donor=0.92)
    b = build_pss_matrices(rmx, 0.01, T_laser, n, k_fret=1 / tau_D, QY_donor=0.92)
    assert b["mean_ph"][0] < a["mean_ph"][0]
    assert b["tau_eff"] == pytest.approx(1 / (1 / tau_D + 1 / tau_D), rel=1e-3)
    assert (b["mean_fret"] >= 0).all() and b["mean_fret"][0] > 0
    assert (a["mean_fret"] == 0).all()


def test_pss_tables_match_oracle(c0):
    rmx = _rm(c0)
    ro = rm._Rates(c0.fluorophores[0])
    T_laser = 1.0 / c0.laser.repetition_rate
    for kf in (0.0, 3e8):
        a = build_pss_matrices(rmx, 0.013, T_laser, 40000, kf, 0.9, 0.7)
        T_step, mean_ph, mean_fr, tau = rm._pss(ro, 0.013, T_laser, 40000, kf, 0.9, 0.7)
        np.testing.assert_array_equal(a["T_step"], T_step)
        np.testing.assert_array_equal(a["mean_ph"], mean_ph)
        np.testing.assert_array_equal(a["mean_fret"], mean_fr)
        assert a["tau_eff"] == tau


# ----------------------------------------------------------------------
# Microscope constants (reference microscope.py:73-391, 650-712)
# ----------------------------------------------------------------------

def test_psf_sigma_and_window(c0):
    psf = PSF(c0)
    s_ill = 0.21 * 462.0 / 1.4
    s_det = 0.21 * 492.0 / 1.4
    expect = 1 / math.sqrt(1 / s_ill ** 2 + 1 / s_det ** 2) / 1000.0
    assert psf.get_sigma(0) == pytest.approx(expect, rel=1e-12)
    assert psf.get_sigma(0) * 1000 == pytest.approx(50.5, abs=0.05)
    assert psf.fwhm_um[0] == pytest.approx(expect * 2 * math.sqrt(2 * math.log(2)), rel=1e-12)
    assert psf.evaluate(0.0) == pytest.approx(1 / (2 * math.pi * expect ** 2), rel=1e-12)
    r = np.linspace(0, 10 * expect, 20001)
    integral = 2 * math.pi * np.trapezoid(psf.evaluate_array(r) * r, r)
    assert integral == pytest.approx(1.0, abs=1e-4)
    sc = ConfocalScanner(c0, psf)
    assert sc.psf_range_px == 2  # ceil(4 sigma / 156 nm)
    assert sc.pixel_cx[0] == pytest.approx(sc.offset_x + 0.078)


def test_large_frame_window():
    cfg = workloads.large_membrane(0)
    sc = ConfocalScanner(cfg, PSF(cfg))
    assert (sc.nx, sc.ny) == (1024, 1024)
    assert sc.psf_range_px == 3  # 97.7 nm pixels


def test_tcspc_settings(c0):
    t = TCSPCAccumulator(c0)
    assert t.n_bins == 256
    assert t.bin_width == pytest.approx(25e-9 / 256, rel=1e-15)
    f2s = 1 / (2 * math.sqrt(2 * math.log(2)))
    expect = math.sqrt((c0.laser.pulse_width * f2s) ** 2 + c0.detector.timing_jitter ** 2)
    assert t.sigma_irf == pytest.approx(expect, rel=1e-15)
    assert t.sigma_irf * 1e12 == pytest.approx(65.6, abs=0.1)
    assert t.n_channels == 2


# ----------------------------------------------------------------------
# Configuration (reference config.py:579-766)
# ----------------------------------------------------------------------

def _raw():
    return workloads.plain_raw()


@pytest.mark.parametrize("mutate, fragment", [
    (lambda d: d["species"][0].__setitem__("diffusion_coefficient", -1.0), "negative D"),
    (lambda d: d["catalytic_reactions"][0].__setitem__("substrate", "NOPE"), "not in species"),
    (lambda d: d["unimolecular_conversions"][0].__setitem__("feedback_rule", "missing"),
     "not found"),
    (lambda d: d.__setitem__("bd_timestep", 0.0), "bd_timestep must be positive"),
])
def test_config_validation_errors(mutate, fragment):
    d = _raw()
    mutate(d)
    with pytest.raises(ConfigError, match=fragment):
        config_from_dict(d).validate()


def test_plain_config_validates(c0):
    c0.validate()
    assert len(c0.species) == 5
    assert len(c0.catalytic_reactions) == 3 and len(c0.compartment_exchanges) == 3


# ---- Philox4x32-10 restatement used by the frame-noise GPU test -----------

def test_philox_known_answers():
    """Known-answer vectors of Philox4x32-10 published with Random123
    (kat_vectors: counter, key -> output)."""
    from philox_ref import philox4x32
    assert philox4x32((0, 0, 0, 0), 0, 0) == (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)
    assert philox4x32((0xffffffff,) * 4, 0xffffffff, 0xffffffff) == (
        0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)
    assert philox4x32((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), 0xa4093822,
                      0x299f31d0) == (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)


def test_frame_noise_restatement_limits():
    """Poisson and binomial inversion against scipy's distributions (CDF
    at the drawn value brackets the uniform)."""
    from scipy import stats
    from philox_ref import binomial_inv, poisson_inv
    for lam in (0.02, 0.7, 3.0):
        for u in (0.0, 0.3, 0.9, 0.999):
            k = poisson_inv(lam, u)
            assert stats.poisson.cdf(k, lam) >= u - 1e-12
            assert k == 0 or stats.poisson.cdf(k - 1, lam) < u + 1e-12
    for n, p in ((10, 0.005), (500, 0.005), (3000, 0.005)):
        for u in (0.0, 0.2, 0.95, 0.9999):
            k = binomial_inv(n, p, u)
            assert stats.binom.cdf(k, n, p) >= u - 1e-9
            assert k == 0 or stats.binom.cdf(k - 1, n, p) < u + 1e-9


def test_local_recruitment_plan():
    """dynamics.py:2034-2041, 2116-2119, 2354: with use_local_pi3k_recruitment
    only the exchange whose single rule is pip3_activation (PI3K) is planned
    as density-biased, carrying recruit_sigma; PTEN_T1 (two rules) and
    PTEN_T2 (none) keep their placement; without the flag nothing changes."""
    cfg = workloads.plain_membrane()
    base = _Plan(cfg, derive_constants(cfg))
    cfg.use_local_pi3k_recruitment = True
    cfg.recruit_sigma = 0.25
    plan = _Plan(cfg, derive_constants(cfg))
    rec = [(r, k) for r, k in zip(plan.reactions, plan.kinds) if k.startswith("recruit")]
    names = [s.name for s in cfg.species]
    assert [(names[r.a], k) for r, k in rec] == [
        ("PI3K", "recruit_local"), ("PTEN_T1", "recruit"), ("PTEN_T2", "recruit")]
    r = rec[0][0]
    assert r.kind == _hip.EDM_RX_RECRUIT and r.use_adapt == 1 and r.contact == 0.25
    assert r.fb1.sensor == names.index("PIP3") and r.fb2.sensor < 0
    assert all(r2.use_adapt == 0 for r2, _ in rec[1:])
    assert base.kinds.count("recruit_local") == 0
    assert [k for k in base.kinds] == [k.replace("recruit_local", "recruit") for k in plan.kinds]

