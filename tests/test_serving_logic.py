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
    """Host stand-in of naer_amd.arena.Arena.  Each resident job 'fires'
    deterministically from its own seed and local steps, so its raster
    depends on nothing else; `alloc_fail` jobs fail to build (as a device
    allocation error would), `poison` jobs report a non-finite state after
    their first chunk."""

    def __init__(self, cap, plastic, alloc_fail=(), poison=()):
        self.cap = cap
        self.plastic = plastic
        self.used_n = self.used_s = 0
        self.slots = {}
        self.steps = 0
        self.admitted = []
        self.alloc_fail = set(alloc_fail)
        self.poison = set(poison)

    def admit(self, net):
        job = net._job
        if job.index in self.alloc_fail:
            raise _abi.NaerError(_abi.NAER_ERR_NOMEM, "hipMalloc failed (injected)")
        n = serving.arena_neurons(job)
        if n > self.cap.neurons or job.n_synapses() > self.cap.synapses:
            raise _abi.NaerError(_abi.NAER_ERR_INVALID, "larger than the arena")
        if self.used_n + n > self.cap.neurons or self.used_s + job.n_synapses() > self.cap.synapses:
            return None
        slot = max(self.slots, default=-1) + 1
        self.slots[slot] = dict(n=n, syn=job.n_synapses(), job=job, t=0, start=self.steps)
        self.used_n += n
        self.used_s += job.n_synapses()
        self.admitted.append(job.index)
        return slot

    def retire(self, slot):
        e = self.slots.pop(slot)
        self.used_n -= e["n"]
        self.used_s -= e["syn"]

    def run(self, steps):
        self.steps += steps

    def drain(self):
        out = {}
        for slot, e in self.slots.items():
            local = self.steps - e["start"]
            steps = np.arange(e["t"], local, dtype=np.uint32)
            rng = np.random.default_rng(e["job"].seed)
            fire = rng.integers(0, 2**32, size=local, dtype=np.uint64)[e["t"]:]
            sel = steps[(fire % 7) == 0]
            sp = np.zeros(len(sel), dtype=_abi.SPIKE_DTYPE)
            sp["step"] = sel
            sp["neuron"] = (fire[(fire % 7) == 0] >> np.uint64(8)) % np.uint64(e["job"].n)
            out[slot] = sp
            e["t"] = local
        return out

    def check(self):
        return [s for s, e in self.slots.items() if e["job"].index in self.poison and e["t"] > 0]

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


def make_server(budget, **fake):
    return serving.Server(budget, arena_factory=lambda cap, plastic: FakeArena(cap, plastic, **fake), graphs=False,
                          keep_rasters=True)


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
        if r.status == "done":
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


# --- failure isolation -------------------------------------------------------
ISO_JOBS = serving.served_jobs(24, rate_per_s=200.0, n_range=(2000, 20000), k_range=(500, 1500))


def serve_iso(**kw):
    fake = {k: kw.pop(k) for k in ("alloc_fail", "poison") if k in kw}
    budget = 4 * max(serving.footprint(j) for j in ISO_JOBS)
    srv = make_server(budget, **fake)
    clock = FakeClock()
    return srv.serve(ISO_JOBS, clock=clock, sleep=clock.sleep, **kw)


def assert_others_unchanged(out, base, affected):
    for i, r in out["results"].items():
        if i in affected:
            continue
        b = base["results"][i]
        assert r.status == b.status == "done", (i, r.status)
        assert np.array_equal(r.raster, b.raster), i


def test_injected_allocation_failure_rejects_that_job_only():
    base = serve_iso()
    assert base["completed"] == len(ISO_JOBS)
    out = serve_iso(alloc_fail={3, 11})
    for i in (3, 11):
        assert out["results"][i].status == "rejected" and "injected" in out["results"][i].error
    assert sorted(out["rejected"]) == [3, 11]
    assert_others_unchanged(out, base, {3, 11})


def test_cancel_and_deadline():
    base = serve_iso()
    # cancel one job before it can start and one while it runs; a deadline
    # shorter than the slowest jobs times those out
    first = min(ISO_JOBS, key=lambda j: j.arrival_s)
    out = serve_iso(cancel_at={first.index: first.arrival_s + 1e-4, 5: 0.0})
    assert out["results"][5].status == "cancelled" and out["results"][5].admitted_s < 0
    assert out["results"][first.index].status == "cancelled"
    assert_others_unchanged(out, base, {first.index, 5})
    lat = sorted(r.latency_s for r in base["results"].values())
    limit = lat[len(lat) // 2]
    out = serve_iso(deadline_s=limit)
    late = {i for i, r in out["results"].items() if r.status == "timeout"}
    assert late and len(late) < len(ISO_JOBS)
    assert all(base["results"][i].latency_s >= limit * 0.5 for i in late)
    for i, r in out["results"].items():
        if i not in late:
            assert np.array_equal(r.raster, base["results"][i].raster), i


def test_watchdog_retires_a_non_finite_job():
    base = serve_iso()
    out = serve_iso(poison={7})
    assert out["results"][7].status == "failed" and "finite" in out["results"][7].error
    assert_others_unchanged(out, base, {7})
    assert out["statuses"]["failed"] == 1


def test_job_of_exactly_the_arena_capacity_is_rejected():
    """The arena takes n + 1 neurons rounded up to 256: a job of n ==
    capacity neurons can never be admitted and must not wait forever."""
    srv = make_server(50e9)
    cap = srv.capacity[0]
    jobs = [serving.JobSpec(0, 0.0, cap.neurons, 500, 0.2, False, 1),
            serving.JobSpec(1, 0.0, 2000, 500, 0.2, False, 2)]
    clock = FakeClock()
    out = srv.serve(jobs, clock=clock, sleep=clock.sleep)
    assert out["results"][0].status == "rejected"
    assert out["results"][1].status == "done"


class BusyArena(FakeArena):
    """FakeArena whose enqueued chunk stays 'running' for `polls` calls of
    idle() (naer_net_idle), as a device chunk does until it completes."""

    def __init__(self, cap, plastic, polls, **kw):
        super().__init__(cap, plastic, **kw)
        self.polls = polls
        self.left = 0
        self.other = None
        self.overlapped = 0  # chunks enqueued while the other arena's chunk was running

    def run(self, steps):
        assert self.left == 0, "a chunk was enqueued before the last one was collected"
        super().run(steps)
        self.left = self.polls
        if self.other is not None and self.other.left:
            self.overlapped += 1

    def idle(self):
        if self.left:
            self.left -= 1
            return False
        return True

    def drain(self):
        assert self.left == 0, "drained while the chunk was still running"
        return super().drain()


def test_arenas_are_serviced_independently():
    """Each arena is collected and stepped again as soon as its own chunk is
    done: a slow plastic arena does not hold back the static one, and every
    job's raster is the one it gets with arenas that are always done."""
    jobs = serving.served_jobs(30, rate_per_s=500.0, n_range=(2000, 20000), k_range=(500, 1500),
                               stdp_fraction=0.5)
    budget = 6 * max(serving.footprint(j) for j in jobs)
    base = make_server(budget)
    clock = FakeClock()
    ref = base.serve(jobs, clock=clock, sleep=clock.sleep)
    srv = serving.Server(budget, graphs=False, keep_rasters=True,
                         arena_factory=lambda cap, plastic: BusyArena(cap, plastic, polls=12 if plastic else 2))
    static, plastic = srv.arenas
    static.other, plastic.other = plastic, static
    clock = FakeClock()
    out = srv.serve(jobs, clock=clock, sleep=clock.sleep)
    assert out["completed"] == ref["completed"] == len(jobs)
    for i, r in out["results"].items():
        assert r.status == "done" and np.array_equal(r.raster, ref["results"][i].raster), i
    # the static arena was stepped again while the plastic one's chunk still ran
    assert static.overlapped > 10
    assert static.used_n == 0 and plastic.used_n == 0
