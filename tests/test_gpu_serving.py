"""Serving on the GPU: every job's result is the same in the continuously
batched mix as alone on the single-membrane path — every FLIM frame (the
served frames stream back sparse, the plain path copies each stack whole),
the frame's dark counts and afterpulses, event counters and particle state,
bit for bit (FP32, Philox)."""

from __future__ import annotations

import numpy as np
import pytest

from ed_mrdt import workloads
from ed_mrdt.serving import Job, Server, n_particles, serve_plain

pytestmark = pytest.mark.gpu


def _jobs(n_jobs, max_particles, max_px, min_particles=0, px=None, start=0):
    jobs, j = [], start
    while len(jobs) < n_jobs:
        cfg, steps = workloads.served_job(j)
        n = n_particles(cfg)
        if min_particles <= n <= max_particles and cfg.scan.pixels_x <= max_px and (
                px is None or cfg.scan.pixels_x == px):
            jobs.append(Job(j, cfg, 120 + 37 * len(jobs), arrival=0.0))
        j += 1
    return jobs


def _same_results(mixed, alone):
    for jid, a in alone.items():
        m = mixed[jid]
        fa, fm = a.sparse_frames(), m.sparse_frames()
        assert m.frames == a.frames == len(fa) == len(fm) == (a.steps - 1) // 100 + 1
        for k, (x, y) in enumerate(zip(fa, fm)):
            assert y.same_counts(x), f"job {jid} frame {k}"
            assert (y.frame_index, y.n_dark_counts, y.n_afterpulses) == (
                x.frame_index, x.n_dark_counts, x.n_afterpulses), f"job {jid} frame {k}"
        assert sum(f.n_signal_photons for f in fm) > 0
        np.testing.assert_array_equal(m.frame.flim_stack, a.frame.flim_stack)
        for k in ("n_catalytic", "n_recruited", "n_detached", "n_converted"):
            assert m.counters[k] == a.counters[k], f"job {jid} {k}"
        for k in ("positions", "species_id", "is_alive"):
            np.testing.assert_array_equal(m.state[k], a.state[k], err_msg=f"job {jid} {k}")


def test_each_job_same_alone_and_in_the_mix():
    jobs = _jobs(6, 30_000, 128)
    srv = Server(jobs[0].config, hbm_budget_bytes=8 << 30, capacity=400_000, max_jobs=4,
                 chunk=40, keep_state=True)
    mixed = {r.job_id: r for r in srv.serve(jobs, realtime=False)}
    alone = {r.job_id: r for r in serve_plain(jobs, realtime=False, keep_state=True)}
    assert srv.peak_resident >= 2  # jobs really overlapped on the device
    _same_results(mixed, alone)
    srv.engine.close()


def test_large_and_256_px_jobs_same_alone_and_in_the_mix():
    """The served workload's extremes: a 256 x 256 px job and jobs of
    100k-250k particles, admitted beside small ones."""
    big = _jobs(1, 250_000, 256, min_particles=100_000)
    wide = _jobs(1, 60_000, 256, px=256)
    small = _jobs(3, 20_000, 128, start=40)
    jobs = big + wide + small
    assert max(n_particles(j.config) for j in jobs) >= 100_000
    assert any(j.config.scan.pixels_x == 256 for j in jobs)
    cap = sum(n_particles(j.config) for j in jobs) + 1024
    srv = Server(jobs[0].config, hbm_budget_bytes=16 << 30, capacity=cap, max_jobs=5,
                 chunk=50, keep_state=True)
    mixed = {r.job_id: r for r in srv.serve(jobs, realtime=False)}
    alone = {r.job_id: r for r in serve_plain(jobs, realtime=False, keep_state=True)}
    assert srv.peak_resident >= 3
    _same_results(mixed, alone)
    srv.engine.close()
