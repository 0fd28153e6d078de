"""Serving on the GPU: every job's result is the same in the continuously
batched mix as alone on the single-membrane path — final FLIM stack,
event counters and particle state bit for bit (FP32, Philox)."""

from __future__ import annotations

import numpy as np
import pytest

from ed_mrdt import workloads
from ed_mrdt.serving import Job, Server, n_particles, serve_plain

pytestmark = pytest.mark.gpu


def _small_jobs(n_jobs=6, max_particles=30_000):
    jobs, j = [], 0
    while len(jobs) < n_jobs:
        cfg, steps = workloads.served_job(j)
        if n_particles(cfg) <= max_particles and cfg.scan.pixels_x <= 128:
            jobs.append(Job(j, cfg, 120 + 37 * len(jobs), arrival=0.0))
        j += 1
    return jobs


def test_each_job_same_alone_and_in_the_mix():
    jobs = _small_jobs()
    model = jobs[0].config
    srv = Server(model, hbm_budget_bytes=8 << 30, capacity=400_000, max_jobs=4, chunk=40,
                 keep_state=True)
    mixed = {r.job_id: r for r in srv.serve(jobs, realtime=False)}
    alone = {r.job_id: r for r in serve_plain(jobs, realtime=False, keep_state=True)}
    assert srv.peak_resident >= 2  # jobs really overlapped on the device
    for jid, a in alone.items():
        m = mixed[jid]
        assert m.frames == a.frames
        np.testing.assert_array_equal(m.frame.flim_stack, a.frame.flim_stack,
                                      err_msg=f"job {jid} FLIM")
        assert m.frame.n_dark_counts == a.frame.n_dark_counts
        for k in ("n_catalytic", "n_recruited", "n_detached", "n_converted"):
            assert m.counters[k] == a.counters[k], f"job {jid} {k}"
        for k in ("positions", "species_id", "is_alive"):
            np.testing.assert_array_equal(m.state[k], a.state[k], err_msg=f"job {jid} {k}")
    srv.engine.close()
