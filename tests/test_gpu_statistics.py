"""Long-run statistics of the fast path (FP32, Philox draws) against the
oracle's ensembles (FP64, the reference's numpy draws), with the
statistics computed on the device (ed_mrdt.observables: MSD with
periodic unwrapping, pair-distance histogram, TCSPC moments) and the
species counts read from the state.

Systems (tests/ensemble_stats.py): the binding / FRET system with imaging,
the 13,000-particle wave system with its sinusoidal D(x) and 128 x 128 px
imaging, and the dense harmonic membrane of the 1M workload at 14 x 14 um,
whose displacement-capped dynamics decorrelate FP32 from FP64 within tens
of steps.  32 replicas per side, 1000 steps each; MSD at 100, 250, 500 and
1000 steps, species counts at 250, 500 and 1000, pair distances at the end,
phasor and mean arrival time over every frame.

Bar: for every observable o of a case, the two ensemble means agree within
the 95 % confidence interval of their difference, Bonferroni-corrected for
the K observables of the case (the family-wise error stays 5 %):
|mean_gpu - mean_oracle| <= z(1 - 0.025 / K) * sqrt(se_gpu^2 + se_oracle^2).
The oracle's ensembles are fixtures written by
tests/golden/make_ensemble_stats.py.

Power: with 32 replicas a side the bar's half-width on the MSD at 1000
steps is about 1.4 % of the MSD per unit z (z = 3.2 for the 36 observables
of the binding case), so a diffusion coefficient off by 5 % (a dropped
sqrt(2) is 41 %) or a binding probability off by a quarter fails it; the
pair bins catch a contact radius or a WCA sigma off by a few per cent.
"""

from __future__ import annotations

import ctypes
import os

import numpy as np
import pytest
from scipy.stats import norm

import ensemble_stats as es
from ed_mrdt import _hip
from ed_mrdt.dynamics import DynamicsEngine
from ed_mrdt.microscope import VirtualMicroscope
from ed_mrdt.observables import MSDTracker, flim_moments, pair_distance_histogram
from ed_mrdt.particles import ParticleState
from ed_mrdt.photophysics import PhotonSimulator

pytestmark = pytest.mark.gpu

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                       "oracle_ensemble_stats.npz")


def _run_steps(eng, k, imaging):
    tot, fms = ctypes.c_double(0), ctypes.c_double(0)
    _hip.check(_hip.lib().edm_time_steps(eng.handle, int(k), 1 if imaging else 0,
                                         ctypes.byref(tot), ctypes.byref(fms), None))


def gpu_replica(case, seed):
    make, imaging, nb, r_max = es.CASES[case]
    cfg = make(seed)
    st = ParticleState(cfg)
    eng = DynamicsEngine(cfg, st, precision="fp32", rng="philox")
    if imaging:
        PhotonSimulator(cfg, st, min(1.0, cfg.scan.dwell_time / cfg.scan.frame_interval))
        mic = VirtualMicroscope(cfg, st)
        fe = es.frame_every(cfg)
        mom = np.zeros(4)
    tracker = MSDTracker(eng)
    marks = sorted(set(es.MSD_AT) | set(es.COUNT_AT) |
                   (set(range(fe, es.STEPS + 1, fe)) if imaging else set()) | {es.STEPS})
    msd, counts = [], []
    done = 0
    for t in marks:
        _run_steps(eng, t - done, imaging)
        done = t
        if imaging and (t % fe == 0 or t == es.STEPS):
            mom += flim_moments(mic, 0).reshape(-1, 4).sum(axis=0)
            mic.finalize_frame()
        if t in es.MSD_AT:
            msd.append(tracker.msd())
        if t in es.COUNT_AT:
            st.mark_stale()
            alive = st.is_alive.astype(bool)
            counts += np.bincount(st.species_id[alive], minlength=len(cfg.species)).tolist()
    pairs = pair_distance_histogram(eng, None, nb, r_max).astype(np.float64)
    out = msd + counts + pairs.tolist()
    if imaging:
        out += es.phasor_from(mom)
    eng.close()
    return np.array(out, dtype=np.float64)


@pytest.mark.parametrize("case", list(es.CASES))
def test_fp32_philox_statistics_match_oracle_ensembles(case):
    b = np.load(FIXTURE)[case]
    n = es.REPLICAS
    assert b.shape[0] == n
    a = np.array([gpu_replica(case, s) for s in range(n)])
    names = es.observable_names(case)
    assert a.shape == b.shape == (n, len(names))
    K = len(names)
    z = norm.ppf(1.0 - 0.025 / K)
    ma, mb = a.mean(axis=0), b.mean(axis=0)
    se = np.sqrt(a.var(axis=0, ddof=1) / n + b.var(axis=0, ddof=1) / n)
    bad = []
    for i, nm in enumerate(names):
        tol = z * se[i] if se[i] > 0 else 0.0
        ok = abs(ma[i] - mb[i]) <= tol + 1e-12 * abs(mb[i])
        print(f"{case} {nm:26s} fp32 {ma[i]:12.6g}  oracle {mb[i]:12.6g}  "
              f"|d|/se {abs(ma[i] - mb[i]) / se[i] if se[i] > 0 else 0.0:5.2f}"
              f"{'' if ok else '  <-- outside'}")
        if not ok:
            bad.append(nm)
    assert not bad, f"outside the Bonferroni-corrected 95 % interval (z = {z:.2f}): {bad}"
    # the observables move: diffusion happened and the MSD grows
    assert np.all(np.diff(ma[:len(es.MSD_AT)]) > 0) and ma[0] > 1e-4
