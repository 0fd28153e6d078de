"""The serving loop's scheduling, without a GPU: a stand-in engine records
what the server asks of it, and the checks are on the schedule itself
(admission order and HBM budget, launch batches that stop at every job end
and frame boundary, batched frame ends, completion)."""

from __future__ import annotations

import pytest

from ed_mrdt import serving, workloads
from ed_mrdt.serving import Job, Server


class _Engine:
    def __init__(self):
        self.replicas = {}
        self.log = []
        self.step = 0
        self.next_slot = 0
        self.live = {}

    def add(self, config):
        k = self.next_slot
        self.next_slot += 1
        self.replicas[k] = dict(frame_index=0)
        self.live[k] = self.step
        self.log.append(("add", k, self.step))
        return k

    def run(self, k):
        assert k > 0
        self.step += k
        self.log.append(("run", k))

    def end_frames(self, slots):
        for s in slots:
            self.replicas[s]["frame_index"] += 1
        self.log.append(("end", tuple(sorted(slots)), self.step))

    def collect_frames(self, slots):
        self.end_frames(slots)
        self.tickets = getattr(self, "tickets", {})
        t = len(self.tickets)
        self.tickets[t] = {s: ("frame", s, self.replicas[s]["frame_index"] - 1) for s in slots}
        assert sum(1 for v in self.tickets.values() if v is not None) <= 2
        return t

    def fetch_frames(self, ticket):
        got, self.tickets[ticket] = self.tickets[ticket], None
        return got

    def frame(self, slot, download=True):
        self.replicas[slot]["frame_index"] += 1
        return None

    def counters(self, slot):
        return {"steps": self.step - self.live[slot]}

    def state(self, slot):
        return None

    def remove(self, slot):
        self.log.append(("remove", slot, self.step))
        del self.live[slot]

    def free_slots(self):
        return 64 - len(self.live)

    def free_particles(self):
        return 10**9


def _jobs():
    base = workloads.plain_membrane()
    steps = [250, 120, 400, 60, 300, 90]
    return [Job(j, base, s, arrival=0.0) for j, s in enumerate(steps)]


def test_schedule_respects_budget_fifo_and_frame_boundaries(monkeypatch):
    eng = _Engine()
    jobs = _jobs()
    per_job = serving.replica_bytes(jobs[0].config)
    srv = Server(jobs[0].config, hbm_budget_bytes=int(2.5 * per_job), capacity=10**7,
                 max_jobs=8, chunk=64, engine=eng)
    res = srv.serve(jobs, realtime=False)
    assert sorted(r.job_id for r in res) == list(range(6))
    # at most two jobs fit the budget at once, admitted first come first served
    adds = [e for e in eng.log if e[0] == "add"]
    assert [a[1] for a in adds] == list(range(6))
    live, peak = 0, 0
    for e in eng.log:
        live += {"add": 1, "remove": -1}.get(e[0], 0)
        peak = max(peak, live)
    assert peak == 2 and srv.peak_resident == 2
    # every job ran exactly its steps; launch batches never exceed the chunk
    for r in res:
        assert r.counters["steps"] == r.steps
    assert all(e[1] <= 64 for e in eng.log if e[0] == "run")
    # frames: one per 100 steps (frame_interval 0.1 s at dt = 1 ms) plus the final
    # one, every one of them returned, in order
    for r in res:
        assert r.frames == (r.steps - 1) // 100 + 1 == len(r.frame_list)
        assert [f[2] for f in r.frame_list] == list(range(r.frames))
    # a job completes only once its last frame was fetched, after the next launch
    assert all(r.finished >= r.admitted for r in res)


def test_oversized_job_is_rejected():
    eng = _Engine()
    job = _jobs()[0]
    srv = Server(job.config, hbm_budget_bytes=1000, capacity=10**7, engine=eng)
    with pytest.raises(ValueError):
        srv.serve([job], realtime=False)


def test_latency_stats():
    r = [serving.JobResult(i, 0.1 * i, 0.1 * i, 0.1 * i + 0.5 + 0.1 * (i % 3), 100, 10, 1)
         for i in range(10)]
    s = serving.latency_stats(r)
    assert s["jobs"] == 10
    assert s["p50_s"] == pytest.approx(0.6)
    assert s["p99_s"] <= 0.7 + 1e-12 and s["p99_s"] >= 0.69
    # first arrival 0.0, last completion job 8 at 0.8 + 0.5 + 0.2 = 1.5 s
    assert s["makespan_s"] == pytest.approx(1.5)
    assert s["jobs_per_s"] == pytest.approx(10 / 1.5)
