"""Caller-side recording of SimulationPipeline (pipeline.py:357-455 ground
truth, 567-642 per-step series), on the CPU:

* the oracle's restatement (oracle/recording.py) pinned by the reference's
  own ground-truth tests (tests/test_pipeline.py:307-380 of the reference),
  by closed forms (Förster's E(r) for a bound donor, the proximity
  biosensor's E_high / E_low, E + tau_DA / tau_D = 1) and by brute-force
  tallies of the state;
* the product's host conversion of census rows (``census_to_subframe``)
  against the oracle's per-step recording, on rows restated from the same
  oracle states.
"""

from __future__ import annotations

import numpy as np
import pytest

import recording as orec
import reference_model as rm
from ed_mrdt import _hip, workloads
from ed_mrdt.pipeline import census_to_subframe

N_STEPS = 40


def _run(cfg, n_steps=N_STEPS, fret_sum=False):
    ost = rm.OracleState(cfg)
    odyn = rm.OracleDynamics(cfg, ost)
    duty = min(1.0, cfg.scan.dwell_time / cfg.scan.frame_interval)
    oph = rm.OraclePhotons(cfg, ost, duty)
    c0 = dict(odyn.counters())
    sub = orec.OracleSubframe(cfg, odyn, oph)
    census, k_steps = [], []
    emitted = 0
    for k in range(n_steps):
        odyn.step()
        b = oph.evolve_one_bd_step()
        emitted += b["n_photons"]
        sub.record(ost, b["n_photons"])
        census.append(orec.census_row(cfg, ost, oph, odyn.counters(), emitted, k + 1, fret_sum))
        k_steps.append(oph.k_fret.copy())
    return dict(cfg=cfg, ost=ost, odyn=odyn, oph=oph, sub=sub, census=census, c0=c0,
                emitted=emitted, k_steps=k_steps)


@pytest.fixture(scope="module")
def bond_run(oracle_lib):
    return _run(workloads.binding_fret(seed=5))


@pytest.fixture(scope="module")
def prox_run(oracle_lib):
    return _run(workloads.binding_fret(seed=5, proximity=True), fret_sum=True)


def _mutual_pairs(bond, alive=None):
    n = 0
    for i in range(len(bond)):
        j = int(bond[i])
        if j > i and int(bond[j]) == i and (alive is None or (alive[i] and alive[j])):
            n += 1
    return n


def _gt(run):
    run["oph"].update_fret_rates()  # pipeline.py:649
    return orec.ground_truth(run["cfg"], run["ost"], run["oph"])


# ---- per-step series ------------------------------------------------------

@pytest.mark.parametrize("which", ["bond_run", "prox_run"])
def test_event_deltas_sum_to_the_cumulative_counters(request, which):
    r = request.getfixturevalue(which)
    a = r["sub"].arrays()
    c1 = r["odyn"].counters()
    for key in ("n_bindings", "n_dissociations", "n_catalytic", "n_recruited", "n_detached",
                "n_converted"):
        assert int(a[key].sum()) == c1[key] - r["c0"][key], key
    assert int(a["total_photons"].sum()) == r["emitted"] > 0
    assert a["n_complexes"].shape == (N_STEPS,)


def test_counts_match_brute_force_tallies(bond_run):
    r = bond_run
    st = r["ost"]
    row = r["sub"].rows[-1]
    alive = st.is_alive.astype(bool)
    # the reference halves the number of ALIVE bonded particles
    # (pipeline.py:570); a binding leaves only the complex particle alive and
    # both partners bonded, so this is half the complexes, as the reference has it
    n_alive_bonded = sum(1 for i in range(st.N) if alive[i] and st.bond_partner[i] >= 0)
    assert n_alive_bonded > 0
    assert row["n_complexes"] == n_alive_bonded // 2
    assert _mutual_pairs(st.bond_partner) == n_alive_bonded
    tally = {}
    for i in range(st.N):
        if alive[i]:
            name = r["cfg"].species[int(st.species_id[i])].name
            tally[name] = tally.get(name, 0) + 1
    assert {k: v for k, v in row["species_counts"].items() if v} == tally
    assert row["n_bleached"] == sum(1 for i in range(st.N) if st.is_bleached[i])


def test_bleached_count_never_decreases(bond_run):
    nb = bond_run["sub"].arrays()["n_bleached"]
    assert np.all(np.diff(nb) >= 0)


def test_bond_mode_mean_efficiency_bounds(bond_run):
    """Bond mode: mean E = E(r) x (bound fraction), so 0 when nothing is
    bound and never above Förster's E(r) of the complex distance."""
    fp = bond_run["cfg"].fret_pairs[0]
    E_r = fp.efficiency(fp.intra_complex_distance)
    a = bond_run["sub"].arrays()
    assert np.all(a["mean_fret_E"] <= E_r + 1e-15)
    assert np.all((a["mean_fret_E"] == 0.0) == (a["n_complexes"] == 0))
    assert np.any(a["n_complexes"] > 0)


def test_proximity_mean_efficiency_is_the_mix_of_E_high_and_E_low(prox_run):
    """Each donor's rate is the biosensor's k_high or k_low, so the mean E
    is (n_high E_high + n_low E_low) / n."""
    r = prox_run
    fp = r["cfg"].fret_pairs[0]
    Eh, El = fp.proximity_E_high, fp.proximity_E_low
    a = r["sub"].arrays()
    tau0 = r["oph"].tau[0]
    k_high = Eh / ((1.0 - Eh) * tau0)
    n_mixed = 0
    for s, row in enumerate(r["census"]):
        n = row["n_donor0"]
        if n == 0:
            continue
        st_k = r["k_steps"][s]
        # donors of this step: type-0 dye carriers that the census counted
        hi = np.isclose(st_k, k_high, rtol=1e-12, atol=0)
        lo = ~hi & (st_k > 0)
        n_hi = int(np.sum(hi))
        n_lo = int(np.sum(lo))
        if n_hi + n_lo != n:
            continue  # a donor bleached during this step's photon pass
        np.testing.assert_allclose(a["mean_fret_E"][s], (n_hi * Eh + n_lo * El) / n, rtol=1e-12)
        n_mixed += 0 < n_hi < n
    assert n_mixed > 0


# ---- ground truth ---------------------------------------------------------

@pytest.mark.parametrize("which", ["bond_run", "prox_run"])
def test_ground_truth_reference_properties(request, which):
    """tests/test_pipeline.py:307-380 of the reference."""
    r = request.getfixturevalue(which)
    g = _gt(r)
    N = r["ost"].N
    assert g["fret_efficiency_exact"].shape == (N,)
    assert g["donor_lifetime_exact"].shape == (N,)
    E, tau = g["fret_efficiency_exact"], g["donor_lifetime_exact"]
    assert np.all(E >= 0.0) and np.all(E <= 1.0)
    assert np.all(tau >= 0.0)
    # donors without FRET keep the donor-only lifetime: 1 / (sum of the
    # fluorescent state's exit rates) of the donor (the reference's test has
    # a 4.1 ns donor)
    fl = r["cfg"].fluorophores[0]
    s1 = next(st.name for st in fl.states if st.is_fluorescent)
    tau_D = 1.0 / sum(fl.thermal_rates[s1].values())
    free = (tau > 0) & (E < 0.01)
    assert np.any(free)
    assert np.all(np.abs(tau[free] - tau_D) / tau_D < 0.05)
    names = [s.name for s in r["cfg"].species]
    receptors = int(np.sum(np.isin(r["ost"].species_id,
                                   [names.index("PHAkt"), names.index("PHAkt_PIP3")])))
    assert 0 <= g["n_complexes"] <= receptors
    assert g["n_bleached"] >= 0


@pytest.mark.parametrize("which", ["bond_run", "prox_run"])
def test_ground_truth_identity_and_counts(request, which):
    r = request.getfixturevalue(which)
    g = _gt(r)
    st = r["ost"]
    E, tau = g["fret_efficiency_exact"], g["donor_lifetime_exact"]
    rmd = r["oph"].rms[0]
    tD = rmd.lifetime(rmd.es)
    donor = (st.has_dye == 1) & (st.is_bleached == 0) & (st.dye_type_id == 0)
    # 1 - tau_DA / tau_D = k tau_D / (1 + k tau_D) = E for every donor
    np.testing.assert_allclose(E[donor] + tau[donor] / tD, 1.0, rtol=0, atol=4e-16)
    assert np.all(E[~donor] == 0.0) and np.all(tau[~donor] == 0.0)
    assert g["n_complexes"] == _mutual_pairs(st.bond_partner)
    n_free = sum(1 for i in range(st.N)
                 if st.has_dye[i] and st.dye_type_id[i] == 0 and not st.is_bleached[i]
                 and st.bond_partner[i] < 0)
    assert g["n_free_donors"] == n_free


def test_ground_truth_bond_mode_is_foerster(bond_run):
    """A bound donor's rate is (R0/r)^6 / tau_D, so E = 1 / (1 + (r/R0)^6)."""
    g = _gt(bond_run)
    fp = bond_run["cfg"].fret_pairs[0]
    E = g["fret_efficiency_exact"]
    on = E > 0
    assert np.any(on)
    np.testing.assert_allclose(E[on], fp.efficiency(fp.intra_complex_distance), rtol=1e-12)
    st = bond_run["ost"]
    assert np.all(st.bond_partner[on] >= 0)


def test_ground_truth_proximity_mode_is_E_high_or_E_low(prox_run):
    g = _gt(prox_run)
    fp = prox_run["cfg"].fret_pairs[0]
    st = prox_run["ost"]
    donor = (st.has_dye == 1) & (st.is_bleached == 0) & (st.dye_type_id == 0) & (st.is_alive == 1)
    E = g["fret_efficiency_exact"][donor]
    hi = np.isclose(E, fp.proximity_E_high, rtol=1e-12, atol=0)
    lo = np.isclose(E, fp.proximity_E_low, rtol=1e-12, atol=0)
    assert np.all(hi | lo) and np.any(hi) and np.any(lo)


# ---- host conversion of census rows ---------------------------------------

def _rows(census):
    rows = np.zeros(len(census), dtype=_hip.CENSUS_DTYPE)
    for i, c in enumerate(census):
        for k, v in c.items():
            rows[i][k] = v
    return rows


@pytest.mark.parametrize("which,fret_sum", [("bond_run", False), ("prox_run", True)])
def test_census_rows_convert_to_the_reference_series(request, which, fret_sum):
    r = request.getfixturevalue(which)
    c0 = r["c0"]
    prev = [c0[k] for k in ("n_bindings", "n_dissociations", "n_catalytic", "n_recruited",
                            "n_detached", "n_converted")]
    sf = census_to_subframe(r["cfg"], _rows(r["census"]), prev, 0, r["cfg"].bd_timestep, 20,
                            fret_sum)
    want = r["sub"].arrays()
    for key in ("n_complexes", "n_bindings", "n_dissociations", "n_bleached", "total_photons",
                "n_catalytic", "n_recruited", "n_detached", "n_converted"):
        np.testing.assert_array_equal(getattr(sf, key), want[key], err_msg=key)
        assert getattr(sf, key).dtype == np.int32
    np.testing.assert_allclose(sf.mean_fret_E, want["mean_fret_E"], rtol=1e-12, atol=0)
    for name, arr in want["species_counts"].items():
        np.testing.assert_array_equal(sf.species_counts[name], arr, err_msg=name)
    assert sf.steps_per_frame == 20 and sf.dt_s == r["cfg"].bd_timestep
