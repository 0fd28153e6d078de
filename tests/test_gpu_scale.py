"""The throughput path (FP32, Philox draws, no pair cap) at the full 1M
workload, checked through properties that do not depend on the size:

* particle conservation: the total, PIP2 + PIP3 (catalysis only
  interconverts them), PI3K and PTEN_T1 + PTEN_T2 (alive + dormant);
* alive(t) - alive(0) = recruited - detached, exactly;
* positions finite and inside the box;
* detected photons, dark counts and afterpulses land in the TCSPC stack
  except arrival times beyond the 25 ns range (about e^-9 of the photons);
* the run is deterministic: the same seed gives bit-identical state.
"""

from __future__ import annotations

import numpy as np
import pytest

from ed_mrdt import workloads
from ed_mrdt.dynamics import DynamicsEngine
from ed_mrdt.microscope import VirtualMicroscope
from ed_mrdt.particles import ParticleState
from ed_mrdt.photophysics import PhotonSimulator

pytestmark = pytest.mark.gpu

STEPS = 20


def _run(cfg):
    st = ParticleState(cfg)
    sp0 = st.species_id.copy()
    alive0 = st.is_alive.copy()
    eng = DynamicsEngine(cfg, st, precision="fp32", rng="philox", pair_cap=False)
    ph = PhotonSimulator(cfg, st, min(1.0, cfg.scan.dwell_time / cfg.scan.frame_interval))
    mic = VirtualMicroscope(cfg, st)
    detected = 0
    for _ in range(STEPS):
        eng.step()
        b = ph.evolve_one_bd_step()
        detected += mic.process_photons(b)
    fr = mic.finalize_frame()
    return st, eng, sp0, alive0, detected, fr


def test_full_size_invariants_and_determinism():
    cfg = workloads.large_membrane(0)
    st, eng, sp0, alive0, detected, fr = _run(cfg)
    names = [s.name for s in cfg.species]
    sid = {n: i for i, n in enumerate(names)}
    sp, alive = st.species_id, st.is_alive
    assert st.N == len(sp0) and int(alive0.sum()) >= 1_000_000

    def count(a, *ns):
        return int(sum(np.sum(a == sid[n]) for n in ns))

    assert count(sp, "PIP2", "PIP3") == count(sp0, "PIP2", "PIP3")
    assert count(sp, "PI3K") == count(sp0, "PI3K")
    assert count(sp, "PTEN_T1", "PTEN_T2") == count(sp0, "PTEN_T1", "PTEN_T2")
    assert int(alive.sum()) - int(alive0.sum()) == eng.n_recruited - eng.n_detached
    assert eng.n_catalytic > 0 and eng.n_recruited > 0 and eng.n_detached > 0
    pos = st.positions
    L = cfg.membrane.size[0]
    assert np.isfinite(pos).all() and (pos >= 0).all() and (pos < L).all()
    assert detected > 0
    expected = detected + fr.n_dark_counts + fr.n_afterpulses
    assert 0 <= expected - fr.total_counts <= 1e-3 * detected + 10
    pos1, sp1, fl1 = pos.copy(), sp.copy(), alive.copy()
    eng.close()
    st2, eng2, *_ = _run(cfg)
    np.testing.assert_array_equal(st2.positions, pos1)
    np.testing.assert_array_equal(st2.species_id, sp1)
    np.testing.assert_array_equal(st2.is_alive, fl1)
    eng2.close()
    print(f"1M x {STEPS} steps: catalytic {eng.n_catalytic}, recruited {eng.n_recruited}, "
          f"detached {eng.n_detached}, converted {eng.n_converted}, detected {detected}")
