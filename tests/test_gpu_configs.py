"""The small baseline workloads at their full size on the GPU.

* The plain PIP3/PTEN membrane (10 x 10 um, 3,500 alive particles, WCA,
  64 x 64 px at 156 nm, a frame every 100 steps) for its whole run of
  1,000 steps and 10 frames, FP64 with the reference's numpy draws injected
  on both sides: per-step photon and detection counts, every frame's TCSPC
  stack with its dark counts and afterpulses, and the reaction counters
  bit-exact against the oracle, positions within 1e-9 relative.
* The 512-replica batch (box sides U[5, 30] um, 35 / um^2): every replica
  conserves its particle classes and keeps its particles in its box, and
  sampled replicas evolve bit for bit as they do alone on the
  single-membrane path, through the steps of a frame and its end.
"""

from __future__ import annotations

import numpy as np
import pytest

import reference_model as rm
from ed_mrdt import workloads
from ed_mrdt.dynamics import DynamicsEngine
from ed_mrdt.ensemble import EnsembleEngine
from ed_mrdt.microscope import VirtualMicroscope
from ed_mrdt.particles import ParticleState
from ed_mrdt.photophysics import PhotonSimulator
from parity_util import compare_states, host_arrays, oracle_arrays

pytestmark = pytest.mark.gpu


def test_plain_membrane_full_run_with_imaging_fp64(oracle_lib):
    cfg = workloads.plain_membrane()
    fe = int(round(cfg.scan.frame_interval / cfg.bd_timestep))
    n_steps = cfg.scan.n_frames * fe
    assert (cfg.scan.pixels_x, cfg.scan.pixels_y, cfg.scan.pixel_size) == (64, 64, 156.0)
    assert n_steps == 1000
    duty = min(1.0, cfg.scan.dwell_time / cfg.scan.frame_interval)
    st = ParticleState(cfg)
    eng = DynamicsEngine(cfg, st, precision="fp64", rng="numpy", pair_cap=True,
                         exact_force_order=True)
    ph = PhotonSimulator(cfg, st, duty)
    mic = VirtualMicroscope(cfg, st)
    ost = rm.OracleState(cfg)
    odyn = rm.OracleDynamics(cfg, ost, pair_cap=True)
    oph = rm.OraclePhotons(cfg, ost, duty)
    omic = rm.OracleMicroscope(cfg, ost)
    tot_ph = tot_det = tot_counts = 0
    worst = 0.0
    for k in range(n_steps):
        eng.step()
        odyn.step()
        b = ph.evolve_one_bd_step()
        ob = oph.evolve_one_bd_step()
        assert b.n_photons == ob["n_photons"], f"step {k}: photons"
        nd = mic.process_photons(b)
        assert nd == omic.process_photons(ob), f"step {k}: detected photons"
        tot_ph += b.n_photons
        tot_det += nd
        if (k + 1) % fe == 0:
            fr = mic.finalize_frame()
            ofr = omic.finalize_frame()
            f = (k + 1) // fe
            assert (fr.n_dark_counts, fr.n_afterpulses) == (ofr["n_dark"], ofr["n_ap"]), f
            np.testing.assert_array_equal(fr.flim_stack, ofr["flim_stack"], err_msg=f"frame {f}")
            tot_counts += fr.total_counts
            worst = max(worst, compare_states(host_arrays(st), oracle_arrays(ost),
                                              label=f"frame {f}: "))
            oc = odyn.counters()
            assert (eng.n_catalytic, eng.n_recruited, eng.n_detached, eng.n_converted) == (
                oc["n_catalytic"], oc["n_recruited"], oc["n_detached"], oc["n_converted"]), f
    assert tot_ph > 0 and tot_det > 0 and tot_counts > 0
    print(f"plain membrane 1000 steps, 10 frames: photons {tot_ph}, detected {tot_det}, "
          f"frame counts {tot_counts}, worst relative position error {worst:.3e}")
    eng.close()


def _single(cfg, steps, fe):
    st = ParticleState(cfg)
    eng = DynamicsEngine(cfg, st, precision="fp32", rng="philox", pair_cap=False)
    ph = PhotonSimulator(cfg, st, min(1.0, cfg.scan.dwell_time / cfg.scan.frame_interval))
    mic = VirtualMicroscope(cfg, st)
    frames = []
    for k in range(steps):
        eng.step()
        mic.process_photons(ph.evolve_one_bd_step())
        if (k + 1) % fe == 0:
            frames.append(mic.finalize_frame().flim_stack.copy())
    out = dict(positions=st.positions.copy(), species_id=st.species_id.copy(),
               is_alive=st.is_alive.copy(), is_bleached=st.is_bleached.copy(),
               ctmc_state=st.ctmc_state.copy())
    ctr = (eng.n_catalytic, eng.n_recruited, eng.n_detached, eng.n_converted)
    eng.close()
    return out, ctr, frames


def test_512_replica_batch():
    n_rep = 512
    cfgs = [workloads.ensemble_replica(r) for r in range(n_rep)]
    fe = int(round(cfgs[0].scan.frame_interval / cfgs[0].bd_timestep))
    steps = fe + 10
    states = [ParticleState(c) for c in cfgs]
    init = [(s.species_id.copy(), s.is_alive.copy()) for s in states]
    ens = EnsembleEngine(cfgs[0], sum(s.N for s in states) + 1024, max_replicas=n_rep)
    slots = [ens.add(c, s) for c, s in zip(cfgs, states)]
    assert ens.info()["replicas"] == n_rep
    ens.run(fe)
    frames = {r: ens.frame(slots[r]).flim_stack.copy() for r in (0, 137, 300, 511)}
    ens.run(steps - fe)
    total_cat = 0
    for r, (c, sl) in enumerate(zip(cfgs, slots)):
        s = ens.state(sl)
        ctr = ens.counters(sl)
        sid = {sp.name: i for i, sp in enumerate(c.species)}
        sp0, al0 = init[r]

        def cnt(a, *ns):
            return int(sum(np.sum(a == sid[n]) for n in ns))

        assert cnt(s["species_id"], "PIP2", "PIP3") == cnt(sp0, "PIP2", "PIP3"), r
        assert cnt(s["species_id"], "PI3K") == cnt(sp0, "PI3K"), r
        assert cnt(s["species_id"], "PTEN_T1", "PTEN_T2") == cnt(sp0, "PTEN_T1", "PTEN_T2"), r
        assert int(s["is_alive"].sum()) - int(al0.sum()) == (
            ctr["n_recruited"] - ctr["n_detached"]), r
        L = c.membrane.size[0]
        assert (s["positions"] >= 0).all() and (s["positions"] < L).all(), r
        total_cat += ctr["n_catalytic"]
    assert total_cat > 0
    for r in (0, 137, 300, 511):
        single, sctr, sframes = _single(cfgs[r], steps, fe)
        s = ens.state(slots[r])
        for k in ("positions", "species_id", "is_alive", "is_bleached", "ctmc_state"):
            np.testing.assert_array_equal(np.asarray(s[k]), single[k], err_msg=f"replica {r}: {k}")
        c = ens.counters(slots[r])
        assert (c["n_catalytic"], c["n_recruited"], c["n_detached"], c["n_converted"]) == sctr, r
        np.testing.assert_array_equal(frames[r], sframes[0], err_msg=f"replica {r}: frame")
    ens.close()
