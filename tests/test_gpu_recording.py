"""GPU parity of the caller-side recording (pipeline.py:357-455 ground truth,
567-642 per-step series): the device census rows and their conversion
against the oracle's recording of the same run (numpy draws injected on
both sides).

Bar: every count and event / photon delta bit-exact; FRET sums, mean E and
the ground-truth E / tau_DA arrays within 1e-12 relative (FP64; the mean's
summation order differs from numpy's).
"""

from __future__ import annotations

import numpy as np
import pytest

import recording as orec
import reference_model as rm
from ed_mrdt import _hip, workloads
from ed_mrdt.dynamics import DynamicsEngine
from ed_mrdt.microscope import VirtualMicroscope
from ed_mrdt.particles import ParticleState
from ed_mrdt.photophysics import PhotonSimulator
from ed_mrdt.pipeline import SimulationPipeline, SubFrameRecorder, extract_ground_truth
from parity_util import compare_states, host_arrays, oracle_arrays

pytestmark = pytest.mark.gpu

_INT_SERIES = ("n_complexes", "n_bindings", "n_dissociations", "n_bleached", "total_photons",
               "n_catalytic", "n_recruited", "n_detached", "n_converted")
_EVENTS = ("n_bindings", "n_dissociations", "n_catalytic", "n_recruited", "n_detached",
           "n_converted")


def _compare_series(sf, want, label=""):
    for key in _INT_SERIES:
        np.testing.assert_array_equal(getattr(sf, key), want[key], err_msg=label + key)
    np.testing.assert_allclose(sf.mean_fret_E, want["mean_fret_E"], rtol=1e-12, atol=0,
                               err_msg=label + "mean_fret_E")
    for name, arr in want["species_counts"].items():
        np.testing.assert_array_equal(sf.species_counts[name], arr, err_msg=label + name)


def _compare_gt(g, og, label=""):
    np.testing.assert_allclose(g.fret_efficiency_exact, og["fret_efficiency_exact"], rtol=1e-12,
                               atol=0, err_msg=label + "E")
    np.testing.assert_allclose(g.donor_lifetime_exact, og["donor_lifetime_exact"], rtol=1e-12,
                               atol=0, err_msg=label + "tau_DA")
    assert (g.n_complexes, g.n_bleached, g.n_free_donors) == (
        og["n_complexes"], og["n_bleached"], og["n_free_donors"]), label
    assert g.species_counts == og["species_counts"], label


@pytest.mark.parametrize("proximity", [False, True])
def test_step_recording_and_ground_truth_match_oracle(oracle_lib, proximity):
    cfg = workloads.binding_fret(seed=5, proximity=proximity)
    duty = min(1.0, cfg.scan.dwell_time / cfg.scan.frame_interval)
    st = ParticleState(cfg)
    eng = DynamicsEngine(cfg, st, precision="fp64", rng="numpy")
    ph = PhotonSimulator(cfg, st, duty)
    mic = VirtualMicroscope(cfg, st)
    ost = rm.OracleState(cfg)
    odyn = rm.OracleDynamics(cfg, ost)
    oph = rm.OraclePhotons(cfg, ost, duty)
    omic = rm.OracleMicroscope(cfg, ost)
    # a ring of 16 rows: 40 steps read it out twice on the way
    rec = SubFrameRecorder(cfg, eng, ph, capacity=16)
    osub = orec.OracleSubframe(cfg, odyn, oph)
    census = []
    emitted = 0
    for k in range(40):
        eng.step()
        odyn.step()
        b = ph.evolve_one_bd_step()
        ob = oph.evolve_one_bd_step()
        assert b.n_photons == ob["n_photons"], f"step {k}"
        mic.process_photons(b)
        omic.process_photons(ob)
        rec.record()
        osub.record(ost, ob["n_photons"])
        emitted += ob["n_photons"]
        census.append(orec.census_row(cfg, ost, oph, odyn.counters(), emitted, k + 1, proximity))
    sf = rec.data(steps_per_frame=40)
    rows = np.concatenate(rec._chunks)
    rec.close()
    assert rows.shape == (40,)
    for i, c in enumerate(census):
        for key in ("step", "n_bond_alive", "n_bond", "n_bleached", "n_donor0", "photons"):
            assert int(rows[key][i]) == c[key], (i, key)
        np.testing.assert_array_equal(rows["n_free_dye"][i], c["n_free_dye"])
        np.testing.assert_array_equal(rows["species_alive"][i], c["species_alive"])
        np.testing.assert_array_equal(rows["events"][i], c["events"])
        np.testing.assert_allclose(rows["e_sum"][i], c["e_sum"], rtol=1e-12, atol=0)
    _compare_series(sf, osub.arrays())
    assert np.any(sf.n_complexes > 0) and np.any(sf.total_photons > 0)
    if proximity:
        assert np.any(sf.mean_fret_E > 0.02)
    compare_states(host_arrays(st), oracle_arrays(ost))
    g = extract_ground_truth(cfg, st, ph)
    oph.update_fret_rates()
    _compare_gt(g, orec.ground_truth(cfg, ost, oph))
    assert np.any(g.fret_efficiency_exact > 0)


def _oracle_pipeline(cfg, n_frames):
    """SimulationPipeline.run (pipeline.py:489-738) on the oracle."""
    ost = rm.OracleState(cfg)
    odyn = rm.OracleDynamics(cfg, ost)
    duty = min(1.0, cfg.scan.dwell_time / cfg.scan.frame_interval)
    oph = rm.OraclePhotons(cfg, ost, duty)
    omic = rm.OracleMicroscope(cfg, ost)
    osub = orec.OracleSubframe(cfg, odyn, oph)
    n_bd = int(round(cfg.scan.frame_interval / cfg.bd_timestep))
    frames = []
    for _ in range(n_frames):
        fe = fd = 0
        for _ in range(n_bd):
            odyn.step()
            ob = oph.evolve_one_bd_step()
            fe += ob["n_photons"]
            fd += omic.process_photons(ob)
            osub.record(ost, ob["n_photons"])
        fr = omic.finalize_frame()
        oph.update_fret_rates()
        frames.append(dict(flim=fr["flim_stack"], gt=orec.ground_truth(cfg, ost, oph),
                           emitted=fe, detected=fd, state=oracle_arrays(ost)))
    return frames, osub.arrays()


def _short_frames(cfg, n_bd):
    cfg.scan.frame_interval = n_bd * cfg.bd_timestep
    return cfg


@pytest.mark.parametrize("proximity", [False, True])
def test_pipeline_frames_match_oracle(oracle_lib, proximity):
    cfg = _short_frames(workloads.binding_fret(seed=7, proximity=proximity), 12)
    cfg.scan.n_frames = 3
    res = SimulationPipeline(cfg, progress_callback=lambda *a: None, record_subframe=True,
                             precision="fp64", rng="numpy", record_capacity=10).run()
    ocfg = _short_frames(workloads.binding_fret(seed=7, proximity=proximity), 12)
    oframes, osf = _oracle_pipeline(ocfg, 3)
    assert res.n_frames == 3
    for f, of in zip(res.frames, oframes):
        np.testing.assert_array_equal(f.flim_frame.flim_stack, of["flim"])
        assert (f.n_photons_emitted, f.n_photons_detected) == (of["emitted"], of["detected"])
        _compare_gt(f.ground_truth, of["gt"], label=f"frame {f.frame_index}: ")
        compare_states(dict(positions=f.ground_truth.positions,
                            species_id=f.ground_truth.species_id,
                            is_alive=f.ground_truth.is_alive.astype(np.uint8),
                            bond_partner=f.ground_truth.bond_partner,
                            ctmc_state=of["state"]["ctmc_state"],
                            is_bleached=f.ground_truth.is_bleached.astype(np.uint8)),
                       of["state"], label=f"frame {f.frame_index}: ")
    sf = res.subframe_data
    assert sf.steps_per_frame == 12 and sf.n_complexes.shape == (36,)
    _compare_series(sf, osf)


def test_device_frame_loop_equals_host_steps():
    """Philox draws: a frame's steps as one device loop (edm_run_frame with
    the census appended on the device) give the same frames, series and
    ground truth as stepping from the host."""
    out = []
    for device_loop in (True, False):
        cfg = _short_frames(workloads.binding_fret(seed=11, proximity=True), 25)
        cfg.scan.n_frames = 2
        res = SimulationPipeline(cfg, progress_callback=lambda *a: None, record_subframe=True,
                                 precision="fp64", rng="philox", device_loop=device_loop,
                                 record_capacity=16).run()
        out.append(res)
    a, b = out
    for fa, fb in zip(a.frames, b.frames):
        np.testing.assert_array_equal(fa.flim_frame.flim_stack, fb.flim_frame.flim_stack)
        assert (fa.n_photons_emitted, fa.n_photons_detected) == (fb.n_photons_emitted,
                                                                 fb.n_photons_detected)
        np.testing.assert_array_equal(fa.ground_truth.positions, fb.ground_truth.positions)
        np.testing.assert_array_equal(fa.ground_truth.fret_efficiency_exact,
                                      fb.ground_truth.fret_efficiency_exact)
        assert fa.ground_truth.species_counts == fb.ground_truth.species_counts
    for key in _INT_SERIES:
        np.testing.assert_array_equal(getattr(a.subframe_data, key), getattr(b.subframe_data, key))
    np.testing.assert_array_equal(a.subframe_data.mean_fret_E, b.subframe_data.mean_fret_E)
    assert a.total_photons_emitted > 0


def test_reference_pipeline_properties():
    """tests/test_pipeline.py:307-380 of the reference on the GPU pipeline."""
    cfg = _short_frames(workloads.binding_fret(seed=2), 10)
    cfg.scan.n_frames = 2
    res = SimulationPipeline(cfg, progress_callback=lambda *a: None, rng="philox").run()
    N = res.frames[0].ground_truth.positions.shape[0]
    names = [s.name for s in cfg.species]
    for fr in res.frames:
        gt = fr.ground_truth
        assert gt.positions.shape == (N, 2)
        for arr in (gt.species_id, gt.bond_partner, gt.fret_efficiency_exact,
                    gt.donor_lifetime_exact):
            assert arr.shape == (N,)
        assert np.all(gt.fret_efficiency_exact >= 0) and np.all(gt.fret_efficiency_exact <= 1)
        assert np.all(gt.donor_lifetime_exact >= 0)
        receptors = int(np.sum(np.isin(gt.species_id, [names.index("PHAkt"),
                                                       names.index("PHAkt_PIP3")])))
        assert 0 <= gt.n_complexes <= receptors
        assert gt.n_bleached >= 0
    assert res.get_intensity_series().shape == (2,)
    assert res.subframe_data is None


def test_large_recording_matches_final_state():
    """40,000 particles, FP32 Philox, 60 steps in device loops with the
    census every step: the last row holds the counts of the downloaded
    state, and the deltas add up to the engine's counters."""
    cfg = workloads.large_membrane(side=20.0, n_alive=40_000, pixels=128)
    cfg.scan.frame_interval = 30 * cfg.bd_timestep
    cfg.scan.n_frames = 2
    pipe = SimulationPipeline(cfg, progress_callback=lambda *a: None, record_subframe=True,
                              precision="fp32", rng="philox", pair_cap=False,
                              record_capacity=25)
    res = pipe.run()
    sf = res.subframe_data
    assert sf.n_complexes.shape == (60,)
    st = pipe._state
    alive = st.is_alive.astype(bool)
    for i, s in enumerate(cfg.species):
        assert int(sf.species_counts[s.name][-1]) == int(np.sum(alive & (st.species_id == i)))
    assert int(sf.n_bleached[-1]) == int(np.sum(st.is_bleached))
    eng = pipe._dynamics
    assert int(sf.n_catalytic.sum()) == eng.n_catalytic > 0
    assert int(sf.n_detached.sum()) == eng.n_detached
    assert int(sf.n_recruited.sum()) == eng.n_recruited
    assert int(sf.total_photons.sum()) == res.total_photons_emitted > 0
    gt = res.frames[-1].ground_truth
    assert gt.species_counts == {s.name: int(sf.species_counts[s.name][-1]) for s in cfg.species}
