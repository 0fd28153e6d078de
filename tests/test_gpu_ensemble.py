"""Batched ensemble on the GPU.

* A replica in the ensemble evolves bit for bit as the same configuration
  does on the single-membrane path (FP32, Philox): positions, species,
  flags, event counters and the TCSPC stack.
* A replica's result is the same alone and in a mix, including replicas
  admitted and removed while it runs (the serving requirement).
* Per-replica conservation laws hold across a 64-replica ensemble.
"""

from __future__ import annotations

import numpy as np
import pytest

from ed_mrdt import workloads
from ed_mrdt.dynamics import DynamicsEngine
from ed_mrdt.ensemble import EnsembleEngine
from ed_mrdt.microscope import VirtualMicroscope
from ed_mrdt.particles import ParticleState
from ed_mrdt.photophysics import PhotonSimulator

pytestmark = pytest.mark.gpu

KEYS = ("positions", "species_id", "is_alive", "is_bleached", "ctmc_state")


def _same(a, b, label):
    for k in KEYS:
        np.testing.assert_array_equal(np.asarray(a[k]), np.asarray(b[k]), err_msg=f"{label}: {k}")


def test_replica_matches_single_membrane_path():
    cfg = workloads.ensemble_replica(3)
    steps = 30
    st = ParticleState(cfg)
    eng = DynamicsEngine(cfg, st, precision="fp32", rng="philox", pair_cap=False)
    ph = PhotonSimulator(cfg, st, min(1.0, cfg.scan.dwell_time / cfg.scan.frame_interval))
    mic = VirtualMicroscope(cfg, st)
    for _ in range(steps):
        eng.step()
        mic.process_photons(ph.evolve_one_bd_step())
    fr = mic.finalize_frame()
    single = dict(positions=st.positions.copy(), species_id=st.species_id.copy(),
                  is_alive=st.is_alive.copy(), is_bleached=st.is_bleached.copy(),
                  ctmc_state=st.ctmc_state.copy())
    ens = EnsembleEngine(cfg, capacity=st.N + 16, max_replicas=4)
    k = ens.add(cfg)
    ens.run(steps)
    _same(ens.state(k), single, "ensemble vs single")
    c = ens.counters(k)
    assert (c["n_catalytic"], c["n_recruited"], c["n_detached"], c["n_converted"]) == (
        eng.n_catalytic, eng.n_recruited, eng.n_detached, eng.n_converted)
    efr = ens.frame(k)
    np.testing.assert_array_equal(efr.flim_stack, fr.flim_stack)
    assert efr.n_dark_counts == fr.n_dark_counts and efr.n_afterpulses == fr.n_afterpulses
    assert fr.total_counts > 0
    ens.close()
    eng.close()


def _alone(r, steps):
    cfg = workloads.ensemble_replica(r)
    e = EnsembleEngine(cfg, capacity=40_000, max_replicas=2)
    k = e.add(cfg)
    e.run(steps)
    out = e.state(k), e.counters(k), e.frame(k).flim_stack
    e.close()
    return out


def test_alone_equals_mix_with_admission_and_removal():
    model = workloads.ensemble_replica(0)
    e = EnsembleEngine(model, capacity=400_000, max_replicas=16)
    slots = {r: e.add(workloads.ensemble_replica(r)) for r in range(6)}
    e.run(20)
    late = e.add(workloads.ensemble_replica(9))  # admitted while the others run
    e.run(10)
    e.remove(slots[1])                           # leaves while the others run
    e.run(30)
    # replica 2: 60 steps in the mix; replica 9: 40 steps since admission
    for r, slot, steps in ((2, slots[2], 60), (9, late, 40), (5, slots[5], 60)):
        a_state, a_ctr, a_flim = _alone(r, steps)
        _same(e.state(slot), a_state, f"replica {r}")
        assert e.counters(slot) == a_ctr, f"replica {r} counters"
        np.testing.assert_array_equal(e.frame(slot).flim_stack, a_flim,
                                      err_msg=f"replica {r} TCSPC")
    info = e.info()
    assert info["replicas"] == 6
    e.close()


def test_ensemble_conservation_64_replicas():
    model = workloads.ensemble_replica(0)
    e = EnsembleEngine(model, capacity=2_500_000, max_replicas=64)
    cfgs = [workloads.ensemble_replica(r) for r in range(64)]
    init = {}
    for r, c in enumerate(cfgs):
        st = ParticleState(c)
        init[e.add(c, st)] = (st.species_id.copy(), st.is_alive.copy(), c)
    e.run(50)
    e.synchronize()
    total_cat = 0
    for slot, (sp0, al0, c) in init.items():
        s = e.state(slot)
        ctr = e.counters(slot)
        sid = {sp.name: i for i, sp in enumerate(c.species)}

        def cnt(a, *ns):
            return int(sum(np.sum(a == sid[n]) for n in ns))

        assert cnt(s["species_id"], "PIP2", "PIP3") == cnt(sp0, "PIP2", "PIP3")
        assert cnt(s["species_id"], "PI3K") == cnt(sp0, "PI3K")
        assert cnt(s["species_id"], "PTEN_T1", "PTEN_T2") == cnt(sp0, "PTEN_T1", "PTEN_T2")
        assert int(s["is_alive"].sum()) - int(al0.sum()) == ctr["n_recruited"] - ctr["n_detached"]
        L = c.membrane.size[0]
        assert (s["positions"] >= 0).all() and (s["positions"] < L).all()
        total_cat += ctr["n_catalytic"]
    assert total_cat > 0
    e.close()


def _local_replica(r):
    cfg = workloads.ensemble_replica(r)
    cfg.use_local_pi3k_recruitment = True
    cfg.recruit_sigma = 0.3
    for ex in cfg.compartment_exchanges:
        if ex.species == "PI3K":
            ex.k_on = 150.0
    return cfg


def test_local_recruitment_replicas_match_single_membrane_path():
    """Density-biased recruitment in the ensemble (a CDF per replica on its
    own grid): every replica of a two-replica ensemble evolves bit for bit as
    the same configuration does on the single-membrane path, also after the
    other replica (and its CDF) is removed mid-run."""
    steps = 30
    singles = {}
    for r in (3, 7):
        cfg = _local_replica(r)
        st = ParticleState(cfg)
        eng = DynamicsEngine(cfg, st, precision="fp32", rng="philox", pair_cap=False)
        assert "recruit_local" in eng.plan.kinds
        eng.run(steps if r == 7 else 2 * steps)
        singles[r] = (dict(positions=st.positions.copy(), species_id=st.species_id.copy(),
                           is_alive=st.is_alive.copy(), is_bleached=st.is_bleached.copy(),
                           ctmc_state=st.ctmc_state.copy()), eng.n_recruited)
        assert eng.n_recruited > 0
        eng.close()
    ens = EnsembleEngine(_local_replica(3), capacity=200_000, max_replicas=4, image=False)
    slots = {r: ens.add(_local_replica(r)) for r in (3, 7)}
    ens.run(steps)
    _same(ens.state(slots[7]), singles[7][0], "replica 7")
    assert ens.counters(slots[7])["n_recruited"] == singles[7][1]
    ens.remove(slots[7])
    ens.run(steps)
    _same(ens.state(slots[3]), singles[3][0], "replica 3 after the removal")
    assert ens.counters(slots[3])["n_recruited"] == singles[3][1]
    ens.close()


def _tuned_run(tune_schedule, n_rep=12, steps=(60, 40, 50)):
    cfgs = [workloads.ensemble_replica(r) for r in range(n_rep)]
    states = [ParticleState(c) for c in cfgs]
    ens = EnsembleEngine(cfgs[0], sum(s.N for s in states) + 1024, max_replicas=n_rep)
    slots = [ens.add(c, s) for c, s in zip(cfgs, states)]
    frames = []
    for k, n in enumerate(steps):
        if tune_schedule[k]:
            ens.set_tuning(**tune_schedule[k])
        ens.run(n)
        frames.append([ens.frame(sl).flim_stack.copy() for sl in slots])
    out = [ens.state(sl) for sl in slots], frames, [ens.counters(sl) for sl in slots]
    tuned = ens.tuning()
    ens.close()
    return out, tuned


def test_ensemble_tuning_leaves_results_unchanged():
    """Strip sizes, column pairing, catalysis lanes and the catalysis walk
    (strips or enzymes from the pass work list; set before the run and
    changed between frames) give the batched step's results bit for bit:
    states, counters and every frame of 12 replicas of 0.9k-31k particles."""
    (ref, _) = _tuned_run([None, None, None])
    sched = [dict(strip_particles=96, pair_below=0, cat_lanes=4, cat_walk=1),
             dict(strip_particles=384, pair_below=256, cat_lanes=64, cat_walk=0),
             dict(strip_particles=128, pair_below=1024, cat_lanes=1, cat_walk=1)]
    (got, tuned) = _tuned_run(sched)
    assert tuned == sched[-1]
    for i, (a, b) in enumerate(zip(ref[0], got[0])):
        _same(a, b, f"replica {i}")
    for fa, fb in zip(ref[1], got[1]):
        for x, y in zip(fa, fb):
            np.testing.assert_array_equal(x, y)
    assert ref[2] == got[2]
    assert sum(int(x.sum()) for x in ref[1][-1]) > 0


def test_ensemble_autotuner_picks_from_timed_candidates():
    from ed_mrdt import tuning
    cfgs = [workloads.ensemble_replica(r) for r in range(24)]
    states = [ParticleState(c) for c in cfgs]
    ens = EnsembleEngine(cfgs[0], sum(s.N for s in states) + 1024, max_replicas=24)
    for c, s in zip(cfgs, states):
        ens.add(c, s)
    res = tuning.autotune_ensemble(ens, steps=5, warmup=1)
    t = res["times_ms"]
    walk = min(((w, g) for w in tuning.ENS_CAT_WALK for g in tuning.ENS_CAT_LANES),
               key=lambda wg: t[f"cat_walk={wg[0]},cat_lanes={wg[1]}"])
    assert (res["cat_walk"], res["cat_lanes"]) == walk
    for name, cands in (("pair_below", tuning.ENS_PAIR_BELOW),
                        ("strip_particles", tuning.ENS_STRIP_PARTICLES)):
        assert res[name] == min(cands, key=lambda v: t[f"{name}={v}"])
    keys = ("strip_particles", "pair_below", "cat_lanes", "cat_walk")
    assert ens.tuning() == {k: res[k] for k in keys}
    print("[ens-tune]", {k: res[k] for k in keys},
          {k: round(v, 3) for k, v in t.items()}, flush=True)
    ens.close()


def _regions_adapt_replica(r):
    """An ensemble replica with two circular slow-diffusion regions of its
    own and the adaptation field on, strong enough to change the PI3K
    catalysis."""
    cfg = workloads.ensemble_replica(r)
    L = cfg.membrane.size[0]
    cfg.diffusion_field.type = "regions"
    cfg.diffusion_field.regions = [
        {"shape": "circle", "center": [L / 3, L / 2], "radius": L / 6, "D_factor": 0.2},
        {"shape": "circle", "center": [0.7 * L, 0.3 * L], "radius": L / 8, "D_factor": 2.0}]
    cfg.use_adaptation = True
    cfg.adaptation_tau_w = 0.02
    cfg.adaptation_K_w = 1.0
    cfg.adaptation_n_w = 2.0
    for c in cfg.catalytic_reactions:
        if "PI3K" in c.enzyme:
            c.k_cat *= 200.0
    return cfg


def test_regions_and_adaptation_replicas_match_single_membrane_path():
    """Circular D(x) regions (each replica its own) and the adaptation field
    (each replica its own on its grid) in the batch: every replica of a
    three-replica ensemble evolves bit for bit as on the single-membrane
    path (FP32, Philox), through the steps of a frame and its end; the same
    replicas without adaptation give other catalytic counts."""
    steps = 50
    rs = (2, 5, 9)
    singles = {}
    for r in rs:
        cfg = _regions_adapt_replica(r)
        st = ParticleState(cfg)
        eng = DynamicsEngine(cfg, st, precision="fp32", rng="philox", pair_cap=False)
        assert "adapt" in eng.plan.kinds
        ph = PhotonSimulator(cfg, st, min(1.0, cfg.scan.dwell_time / cfg.scan.frame_interval))
        mic = VirtualMicroscope(cfg, st)
        for _ in range(steps):
            eng.step()
            mic.process_photons(ph.evolve_one_bd_step())
        fr = mic.finalize_frame()
        singles[r] = (dict(positions=st.positions.copy(), species_id=st.species_id.copy(),
                           is_alive=st.is_alive.copy(), is_bleached=st.is_bleached.copy(),
                           ctmc_state=st.ctmc_state.copy()),
                      (eng.n_catalytic, eng.n_recruited, eng.n_detached, eng.n_converted),
                      fr.flim_stack.copy())
        eng.close()
    ens = EnsembleEngine(_regions_adapt_replica(rs[0]), capacity=400_000, max_replicas=4)
    slots = {r: ens.add(_regions_adapt_replica(r)) for r in rs}
    ens.run(steps)
    for r in rs:
        _same(ens.state(slots[r]), singles[r][0], f"replica {r}")
        c = ens.counters(slots[r])
        assert (c["n_catalytic"], c["n_recruited"], c["n_detached"],
                c["n_converted"]) == singles[r][1], r
        np.testing.assert_array_equal(ens.frame(slots[r]).flim_stack, singles[r][2])
    ens.close()
    cfg0 = _regions_adapt_replica(rs[0])
    cfg0.use_adaptation = False
    plain = EnsembleEngine(cfg0, capacity=400_000, max_replicas=4)
    k = plain.add(cfg0)
    plain.run(steps)
    assert plain.counters(k)["n_catalytic"] != singles[rs[0]][1][0]
    plain.close()


def test_ensemble_refuses_a_replica_with_more_than_8_regions():
    from ed_mrdt.config import ConfigError
    cfg = _regions_adapt_replica(1)
    ens = EnsembleEngine(cfg, capacity=100_000, max_replicas=2)
    bad = _regions_adapt_replica(1)
    bad.diffusion_field.regions = bad.diffusion_field.regions * 5
    with pytest.raises(ConfigError):
        ens.add(bad)
    ens.close()


def _single_run(cfg, steps):
    st = ParticleState(cfg)
    eng = DynamicsEngine(cfg, st, precision="fp32", rng="philox", pair_cap=False)
    ph = PhotonSimulator(cfg, st, min(1.0, cfg.scan.dwell_time / cfg.scan.frame_interval))
    mic = VirtualMicroscope(cfg, st)
    for _ in range(steps):
        eng.step()
        mic.process_photons(ph.evolve_one_bd_step())
    fr = mic.finalize_frame()
    out = dict(positions=st.positions.copy(), species_id=st.species_id.copy(),
               is_alive=st.is_alive.copy(), is_bleached=st.is_bleached.copy(),
               ctmc_state=st.ctmc_state.copy(), bond_partner=st.bond_partner.copy())
    ctr = (eng.n_bindings, eng.n_dissociations, eng.n_catalytic, eng.n_recruited,
           eng.n_detached, eng.n_converted)
    eng.close()
    return out, ctr, fr.flim_stack.copy()


@pytest.mark.parametrize("variant", [dict(), dict(intramolecular=False),
                                     dict(proximity=True)])
def test_binding_replicas_match_single_membrane_path(variant):
    """Binding and dissociation (PHAkt + PIP3 <-> PHAkt_PIP3, dynamics.py:
    529-654) with FRET -- bond mode (intramolecular, or a labelled PIP3
    acceptor found through the replica's slots) or the proximity biosensor
    (photophysics.py:592-683, each replica its own target grid) -- and
    circular D(x) regions in the batch: each replica of a three-replica
    ensemble (different boxes and seeds) evolves bit for bit as on the
    single-membrane path (FP32, Philox): positions, species, flags, bonds,
    event counters and the frame."""
    steps = 60
    cfgs = {k: workloads.binding_fret(seed=s, side=side, **variant)
            for k, (s, side) in enumerate(((3, 4.0), (4, 5.0), (7, 6.0)))}
    for c in cfgs.values():
        for r in c.reactions:
            r.koff = 40.0  # complexes dissociate within the run too
    singles = {k: _single_run(c, steps) for k, c in cfgs.items()}
    ens = EnsembleEngine(cfgs[0], capacity=200_000, max_replicas=4)
    slots = {k: ens.add(c) for k, c in cfgs.items()}
    ens.run(steps)
    for k in cfgs:
        s = ens.state(slots[k])
        _same(s, singles[k][0], f"replica {k}")
        np.testing.assert_array_equal(s["bond_partner"], singles[k][0]["bond_partner"])
        c = ens.counters(slots[k])
        got = (c["n_bindings"], c["n_dissociations"], c["n_catalytic"], c["n_recruited"],
               c["n_detached"], c["n_converted"])
        assert got == singles[k][1], (k, got, singles[k][1])
        np.testing.assert_array_equal(ens.frame(slots[k]).flim_stack, singles[k][2])
    assert all(v[1][0] > 0 and v[1][1] > 0 for v in singles.values()), \
        [v[1] for v in singles.values()]
    ens.close()


def test_ensemble_binding_edge_overflow_is_reported():
    """A replica whose step accepts more candidate binding edges than its
    edge region holds (2 n) makes the run fail with EDM_ECAPACITY instead of
    matching an incomplete set."""
    from ed_mrdt._hip import HipError
    cfg = workloads.binding_fret(seed=5, side=2.0, regions=False)
    pops = cfg.initial_populations
    pops["PIP2"], pops["PIP3"], pops["PHAkt"] = 800, 1600, 1400
    for r in cfg.reactions:
        r.kon = 1.0e4
        r.contact_radius = 150.0
    ens = EnsembleEngine(cfg, capacity=50_000, max_replicas=2)
    ens.add(cfg)
    with pytest.raises(HipError, match="edge region"):
        ens.run(3)
    ens.close()


def _autocat_replica(seed, side, feedback):
    """The plain membrane on a side x side box at its density with the
    autocatalytic PIP3 + PIP2 -> 2 PIP3 added after PI3K's reaction
    (optionally under the pip3_activation feedback)."""
    raw = workloads.rescale(workloads.plain_raw(), side, int(round(35.0 * side * side)))
    rx = {"name": "PIP3_autocatalysis", "enzyme": "PIP3", "substrate": "PIP2",
          "product": "PIP3", "k_cat": 8.0, "contact_radius": 150.0}
    if feedback:
        rx["feedback_rule"] = "pip3_activation"
    raw["catalytic_reactions"].insert(1, rx)
    raw["scan"].update(pixels_x=32, pixels_y=32, pixel_size=side * 1000.0 / 32,
                       frame_interval=0.05)
    raw["random_seed"] = seed
    return workloads.config_from_dict(raw)


@pytest.mark.parametrize("feedback", [False, True])
def test_autocatalysis_replicas_match_single_membrane_path(feedback):
    """Autocatalysis (enzyme == product) in the batch: each replica's
    conversions cascade in its own reference pair order (pair-list offsets
    over the replica's slot range), and every replica of a three-replica
    ensemble evolves bit for bit as on the single-membrane path (FP32,
    Philox)."""
    steps = 40
    cfgs = {k: _autocat_replica(s, side, feedback)
            for k, (s, side) in enumerate(((5, 6.0), (8, 9.0), (11, 7.5)))}
    singles = {k: _single_run(c, steps) for k, c in cfgs.items()}
    ens = EnsembleEngine(cfgs[0], capacity=200_000, max_replicas=4)
    slots = {k: ens.add(c) for k, c in cfgs.items()}
    ens.run(steps)
    for k, c in cfgs.items():
        s = ens.state(slots[k])
        _same(s, singles[k][0], f"replica {k}")
        cc = ens.counters(slots[k])
        got = (cc["n_bindings"], cc["n_dissociations"], cc["n_catalytic"], cc["n_recruited"],
               cc["n_detached"], cc["n_converted"])
        assert got == singles[k][1], (k, got, singles[k][1])
        np.testing.assert_array_equal(ens.frame(slots[k]).flim_stack, singles[k][2])
        pip3 = [sp.name for sp in c.species].index("PIP3")
        n0 = int(np.sum((ParticleState(c).species_id == pip3)))
        assert int(np.sum(s["species_id"] == pip3)) > 2 * n0, "PIP3 must grow"
    ens.close()
