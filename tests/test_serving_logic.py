"""Admission and continuous-batching logic of the server, on the CPU.

The device arenas are replaced by a host stand-in with the same admit / run /
drain / retire interface and first-fit capacity accounting, so the queueing
policy (admission by footprint, waiting, backfill, retirement at the job's
duration, rejection of jobs larger than an arena) is tested without a GPU.
"""
import numpy as np
import pytest

from naer_amd import _abi, serving


class FakeArena:
    def __init__(self, cap, plastic):
        self.cap = cap
        self.plastic = plastic
        self.used_n = self.used_s = 0
        self.slots = {}
        self.steps = 0
        self.admitted = []

    def admit(self, net):
        job = net._job
        n = (job.n + 1 + 255) // 256 * 256
        if self.used_n + n > self.cap.neurons or self.used_s + job.n_synapses() > self.cap.synapses:
            return None
        slot = max(self.slots, default=-1) + 1
        self.slots[slot] = (n, job.n_synapses())
        self.used_n += n
        self.used_s += job.n_synapses()
        self.admitted.append(job.index)
        return slot

    def retire(self, slot):
        n, s = self.slots.pop(slot)
        self.used_n -= n
        self.used_s -= s

    def run(self, steps):
        self.steps += steps

    def drain(self):
        return {}

    def close(self):
        pass


class FakeClock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        self.t += 0.001
        return self.t

    def sleep(self, dt):
        self.t += dt


@pytest.fixture(autouse=True)
def no_device_networks(monkeypatch):
    # the fake arena only needs the job spec, not a device description
    def network(self, mode=_abi.MODE_DETERMINISTIC, record_capacity=0):
        class N:
            pass
        n = N()
        n._job = self
        return n
    monkeypatch.setattr(serving.JobSpec, "network", network)


def make_server(budget):
    return serving.Server(budget, arena_factory=lambda cap, plastic: FakeArena(cap, plastic), graphs=False)


def test_budget_split_respects_budget():
    for budget in (1e9, 50e9, 200e9):
        s, p = serving.split_budget(budget)
        total = (s.neurons * serving.neuron_bytes(_abi.MODE_DETERMINISTIC, False) + s.synapses * 8 +
                 p.neurons * serving.neuron_bytes(_abi.MODE_DETERMINISTIC, True) + p.synapses * 8)
        assert total <= budget * 1.001
        assert s.synapses > 0 and p.synapses > 0
        assert p.synapses < s.synapses  # 30 % of the jobs are plastic


def test_footprint_counts_synapses_and_stdp_state():
    a = serving.JobSpec(0, 0.0, 10_000, 1000, 1.0, False, 1)
    b = serving.JobSpec(1, 0.0, 10_000, 1000, 1.0, True, 1)
    assert serving.footprint(a) >= 10_000 * 1000 * 8
    # STDP keeps per-neuron spike records (176 bytes), no per-synapse index
    assert serving.footprint(b) - serving.footprint(a) == (10_001 + 255) // 256 * 256 * 176


def test_jobs_wait_for_capacity_and_all_complete():
    jobs = serving.served_jobs(40, rate_per_s=1000.0, n_range=(2000, 20000), k_range=(500, 2000))
    # room for only a few jobs at a time
    budget = 3 * max(serving.footprint(j) for j in jobs)
    srv = make_server(budget)
    clock = FakeClock()
    out = srv.serve(jobs, clock=clock, sleep=clock.sleep)
    static, plastic = srv.arenas
    assert out["completed"] + len(out["rejected"]) == len(jobs)
    assert out["completed"] > 0
    # never more resident than capacity (checked by the fake) and queueing happened
    assert any(r.admitted_s > r.arrival_s + 0.01 for r in out["results"].values() if r.admitted_s >= 0)
    assert static.used_n == 0 and plastic.used_n == 0  # everything retired
    for r in out["results"].values():
        if r.done_s >= 0:
            job = jobs[r.index]
            assert r.done_s >= r.admitted_s >= r.arrival_s
    assert out["p99_s"] >= out["p50_s"] > 0


def test_oversized_job_is_rejected_not_queued_forever():
    jobs = [serving.JobSpec(0, 0.0, 200_000, 2000, 0.2, False, 1),
            serving.JobSpec(1, 0.0, 2_000, 500, 0.2, False, 2)]
    srv = make_server(serving.footprint(jobs[1]) * 4)
    clock = FakeClock()
    out = srv.serve(jobs, clock=clock, sleep=clock.sleep)
    assert out["rejected"] == [0]
    assert out["completed"] == 1


def test_job_stream_statistics():
    jobs = serving.served_jobs(4000, rate_per_s=10.0)
    n = np.array([j.n for j in jobs])
    k = np.array([j.indegree for j in jobs])
    gaps = np.diff([0.0] + [j.arrival_s for j in jobs])
    assert 2000 <= n.min() and n.max() <= 200_000
    assert abs(np.median(np.log(n)) - 0.5 * (np.log(2000) + np.log(200_000))) < 0.1
    assert 500 <= k.min() and k.max() <= 2000
    assert abs(gaps.mean() - 0.1) < 0.01          # Poisson arrivals at 10 jobs/s
    assert abs(np.mean([j.plastic for j in jobs]) - 0.3) < 0.03
