"""The device observables (ed_mrdt.observables) against the same statistics
computed with numpy on the oracle's state, on one FP64 run with the
reference's draws (both sides hold the same positions to 1e-9): MSD with
periodic unwrapping (validation.py:276-290), pair-distance counts and g(r)
(analysis.py:733-790), and the per-pixel TCSPC moments behind
mean_arrival_time and phasor_transform (analysis.py:87-127, 497-566)."""

from __future__ import annotations

import math

import numpy as np
import pytest
from scipy.spatial import cKDTree

import reference_model as rm
from ed_mrdt import workloads
from ed_mrdt.dynamics import DynamicsEngine
from ed_mrdt.microscope import VirtualMicroscope
from ed_mrdt.observables import (MSDTracker, flim_moments, mean_arrival_time,
                                 pair_distance_histogram, phasor_transform, rdf)
from ed_mrdt.particles import ParticleState
from ed_mrdt.photophysics import PhotonSimulator

pytestmark = pytest.mark.gpu


def test_observables_match_numpy_on_the_oracle_state(oracle_lib):
    cfg = workloads.binding_fret(side=4.0)
    st = ParticleState(cfg)
    eng = DynamicsEngine(cfg, st, precision="fp64", rng="numpy", exact_force_order=True)
    duty = min(1.0, cfg.scan.dwell_time / cfg.scan.frame_interval)
    ph = PhotonSimulator(cfg, st, duty)
    mic = VirtualMicroscope(cfg, st)
    ost = rm.OracleState(cfg)
    odyn = rm.OracleDynamics(cfg, ost)
    oph = rm.OraclePhotons(cfg, ost, duty)
    omic = rm.OracleMicroscope(cfg, ost)
    tracker = MSDTracker(eng)
    L = np.asarray(cfg.membrane.size)
    prev = ost.positions.copy()
    acc = np.zeros_like(prev)
    names = [s.name for s in cfg.species]
    for k in range(30):
        eng.step()
        odyn.step()
        mic.process_photons(ph.evolve_one_bd_step())
        omic.process_photons(oph.evolve_one_bd_step())
        d = ost.positions - prev
        d -= L * np.round(d / L)
        acc += d
        prev = ost.positions.copy()
        if (k + 1) % 10 == 0:
            alive = ost.is_alive.astype(bool)
            for sp in (None, ["PIP2", "PIP3"], ["PHAkt"]):
                sel = alive.copy()
                if sp is not None:
                    sel &= np.isin(ost.species_id, [names.index(s) for s in sp])
                want = float(np.mean(np.sum(acc[sel] ** 2, axis=1)))
                s, c = tracker.sums(sp)
                assert c == int(sel.sum())
                assert s / c == pytest.approx(want, rel=1e-9)
    # pair-distance counts of the alive particles against a k-d tree
    alive = ost.is_alive.astype(bool)
    for sp, nb, r_max in ((None, 20, 0.5), (["PIP2"], 7, 0.35)):
        sel = alive.copy()
        if sp is not None:
            sel &= np.isin(ost.species_id, [names.index(s) for s in sp])
        tree = cKDTree(np.mod(ost.positions[sel], L), boxsize=L)
        want = np.diff(tree.count_neighbors(tree, np.linspace(0, r_max, nb + 1))) // 2
        got = pair_distance_histogram(eng, sp, nb, r_max)
        np.testing.assert_array_equal(got, want)
        assert want.sum() > 0
    r, g = rdf(eng, None, 20, 0.5)
    Nsel = int(alive.sum())
    rho = Nsel / (L[0] * L[1])
    np.testing.assert_allclose(g, 2 * want_all(ost, L, 20, 0.5) / (Nsel * rho * 2 * np.pi * r *
                                                                   0.025), rtol=1e-12)
    # TCSPC moments, mean arrival time and phasor of the current frame
    stack = omic.flim
    assert stack.sum() > 50
    m = flim_moments(mic, 0)
    h = stack[0].astype(np.float64)
    nbins = h.shape[-1]
    phi = 2 * math.pi * (np.arange(nbins) + 0.5) / nbins
    bin_ns = cfg.tcspc.time_range / nbins * 1e9
    np.testing.assert_array_equal(m[..., 0], h.sum(-1))
    np.testing.assert_allclose(m[..., 1], h @ np.cos(phi), rtol=1e-12, atol=1e-9)
    np.testing.assert_allclose(m[..., 2], h @ np.sin(phi), rtol=1e-12, atol=1e-9)
    np.testing.assert_allclose(m[..., 3], h @ ((np.arange(nbins) + 0.5) * bin_ns), rtol=1e-12)
    tau, I = mean_arrival_time(mic, 0, min_counts=3)
    ok = h.sum(-1) >= 3
    assert ok.any()
    np.testing.assert_allclose(tau[ok], (h @ ((np.arange(nbins) + 0.5) * bin_ns))[ok] /
                               h.sum(-1)[ok], rtol=1e-12)
    assert np.all(np.isnan(tau[~ok]))
    p = phasor_transform(mic, 0, min_counts=3)
    np.testing.assert_allclose(p.g[ok], (h @ np.cos(phi))[ok] / h.sum(-1)[ok], rtol=1e-12,
                               atol=1e-12)
    assert np.all(p.g[~ok] == 0)


def want_all(ost, L, nb, r_max):
    alive = ost.is_alive.astype(bool)
    tree = cKDTree(np.mod(ost.positions[alive], L), boxsize=L)
    return np.diff(tree.count_neighbors(tree, np.linspace(0, r_max, nb + 1))) / 2.0
