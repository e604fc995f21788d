"""Serving with continuous batching on the device.

Jobs of different sizes, durations and plasticity arrive while others run;
they are admitted into two arenas as capacity frees up.  Each job's raster in
the mix must equal its raster when simulated alone (deterministic mode, bit
for bit), and the queueing must actually have happened (some jobs admitted
late, slots reused).
"""
import numpy as np
import pytest

from helpers import first_divergence
from naer_amd import _abi, serving

pytestmark = pytest.mark.gpu


def test_mix_equals_alone():
    jobs = serving.served_jobs(10, rate_per_s=400.0, seed=99, n_range=(300, 3000), k_range=(50, 300),
                               dur_range=(0.02, 0.05), stdp_fraction=0.4)
    budget = 6 * max(serving.footprint(j) for j in jobs)
    srv = serving.Server(budget, mode=_abi.MODE_DETERMINISTIC, chunk=50, keep_rasters=True,
                         record_per_chunk=1 << 20, synapses_per_neuron=40, stdp_fraction=0.4)
    out = srv.serve(jobs)
    srv.close()
    assert out["completed"] == len(jobs), out["rejected"]
    res = out["results"]
    assert any(r.admitted_s > r.arrival_s for r in res.values())
    assert sum(r.spikes for r in res.values()) > 500
    for job in jobs:
        alone = serving.run_alone(job, _abi.MODE_DETERMINISTIC)
        mixed = res[job.index].raster
        assert first_divergence(mixed, alone) is None, (job, len(mixed), len(alone))


def test_arena_slot_reuse_and_capacity_errors():
    from naer_amd.arena import Arena
    jobs = serving.served_jobs(3, rate_per_s=1.0, seed=5, n_range=(500, 800), k_range=(50, 60),
                               dur_range=(0.01, 0.01), stdp_fraction=0.0)
    cap_n = 1024
    a = Arena(cap_n, 10 * 800 * 60, dt_ms=0.1, mode=_abi.MODE_DETERMINISTIC, delay_steps=(15, 15))
    s0 = a.admit(jobs[0].network())
    assert s0 is not None
    # 1024 neurons hold one job of up to 800 neurons (ranges are 256-aligned, one entry more than the job)
    assert a.admit(jobs[1].network()) is None
    a.run(30)
    first = a.drain()[s0]
    assert first["step"].max() < 30
    a.retire(s0)
    s1 = a.admit(jobs[1].network())
    assert s1 == s0
    a.run(30)
    sp = a.drain()
    alone = serving.run_alone(serving.JobSpec(1, 0, jobs[1].n, jobs[1].indegree, 0.003, False, jobs[1].seed))
    assert first_divergence(np.sort(sp[s1][sp[s1]["step"] < 30], order=["step", "neuron", "time_us"]),
                            alone) is None
    a.close()


def test_plastic_arena_batches_and_rejects_short_refractory():
    from naer_amd import models
    from naer_amd.arena import Arena
    from naer_amd.network import NaerError
    a = Arena(4096, 4096 * 300, dt_ms=0.1, mode=_abi.MODE_DETERMINISTIC, delay_steps=(15, 15), plastic=True,
              stdp=models.PAIR_STDP, refractory_min=20)
    assert a.stats().batch == 15  # min(delay 15, refractory 20 + 1)
    ok = models.balanced(1000, 100, 0.1, 5.0, 20_000.0, 0.1, delay_ms=1.5, plastic=True, seed=3,
                         mode=_abi.MODE_DETERMINISTIC)
    short = models.balanced(1000, 100, 0.1, 5.0, 20_000.0, 0.1, delay_ms=1.5, plastic=True, seed=4,
                            mode=_abi.MODE_DETERMINISTIC)
    for p in short._pops:
        p.params["t_ref_ms"] = 1.0  # 10 steps: a neuron could fire twice in a 15-step batch
    assert a.admit(ok) is not None
    with pytest.raises(NaerError):
        a.admit(short)
    a.close()


@pytest.mark.parametrize("graphs", [False, True])
@pytest.mark.parametrize("case", ["at_start", "mid_batch", "reused_range", "neighbour_retires", "two_in_retired_range"])
def test_plastic_job_in_arena_equals_alone(case, graphs):
    """A plastic job in a batching plastic arena (chunks of 50 steps against
    15-step batches) gives its alone raster: admitted at step 0, in the
    middle of a batch, into the ranges of a retired job, or beside a job that
    retires while it runs."""
    from naer_amd import models
    from naer_amd.arena import Arena
    job = serving.JobSpec(0, 0.0, 700, 150, 0.06, True, 1234)
    other = serving.JobSpec(1, 0.0, 900, 200, 0.06, True, 999)
    a = Arena(4096, 4096 * 250, dt_ms=0.1, mode=_abi.MODE_DETERMINISTIC, delay_steps=(15, 15), plastic=True,
              stdp=models.PAIR_STDP, refractory_min=20, record_capacity=1 << 20)
    if graphs:
        a.set_graph(50)
    so = None
    if case in ("mid_batch", "neighbour_retires"):
        so = a.admit(other.network())
        a.run(7)
    elif case in ("reused_range", "two_in_retired_range"):
        big = serving.JobSpec(2, 0.0, 2500, 200, 0.06, True, 777) if case == "two_in_retired_range" else other
        so = a.admit(big.network())
        a.run(120)
        a.drain()
        a.retire(so)
        so = None
        if case == "two_in_retired_range":
            # a smaller job first: `job` lands behind it, with free neurons
            # between their ranges, inside the retired job's ranges
            a.admit(serving.JobSpec(3, 0.0, 300, 100, 0.06, True, 555).network())
    s = a.admit(job.network())
    assert s is not None
    got = []
    for k in range(job.steps() // 50 + 1):
        a.run(50)
        sp = a.drain()
        if s in sp:
            got.append(sp[s])
        if case == "neighbour_retires" and k == 3:
            a.retire(so)
    mixed = np.sort(np.concatenate(got), order=["step", "neuron", "time_us"])
    mixed = mixed[mixed["step"] < job.steps()]
    alone = serving.run_alone(job, _abi.MODE_DETERMINISTIC)
    assert len(alone) > 300
    assert first_divergence(mixed, alone) is None, (case, len(mixed), len(alone))
    a.close()


@pytest.mark.parametrize("plastic", [False, True])
def test_job_in_large_arena_equals_alone(plastic):
    """An arena with the neuron capacity of the served benchmark (2e7): its
    buckets only fit the histogram with 1024-neuron tiles, and a
    histogram/placement block spans ~80 spike-list segments (staged 64 at a
    time).  A job in it gives its alone raster."""
    from naer_amd import models
    from naer_amd.arena import Arena
    a = Arena(20_000_000, 50_000_000, dt_ms=0.1, mode=_abi.MODE_DETERMINISTIC, delay_steps=(15, 15),
              plastic=plastic, stdp=models.PAIR_STDP if plastic else None, refractory_min=20 if plastic else 0,
              record_capacity=1 << 20)
    st = a.stats()
    assert st.batch == 15
    assert a.schedule()["bucketed"] == 1 and a.schedule()["segments_per_block"] > 64
    job = serving.JobSpec(0, 0.0, 3000, 300, 0.03, plastic, 4321)
    first = serving.JobSpec(1, 0.0, 2000, 200, 0.03, plastic, 99)
    a.admit(first.network())  # the job lands behind another one
    a.run(11)
    s = a.admit(job.network())
    got = []
    for _ in range(job.steps() // 50 + 1):
        a.run(50)
        sp = a.drain()
        if s in sp:
            got.append(sp[s])
    mixed = np.sort(np.concatenate(got), order=["step", "neuron", "time_us"])
    mixed = mixed[mixed["step"] < job.steps()]
    alone = serving.run_alone(job, _abi.MODE_DETERMINISTIC)
    assert len(alone) > 300
    assert first_divergence(mixed, alone) is None, (len(mixed), len(alone))
    a.close()


def test_watchdog_and_cancel_leave_the_other_jobs_unchanged(monkeypatch):
    """One job's drive (-3e38 mV per event, finite) drives its potentials to
    -inf: the watchdog retires it as failed after a chunk.  Another job is
    cancelled while it runs.  Every other job finishes with the raster it
    has alone."""
    jobs = serving.served_jobs(8, rate_per_s=400.0, seed=123, n_range=(300, 2000), k_range=(50, 200),
                               dur_range=(0.02, 0.05), stdp_fraction=0.25)
    poisoned, cancelled = jobs[2].index, jobs[4].index
    jobs[4].duration_s = 2.0  # long enough to be running when it is cancelled
    build = serving.JobSpec.network

    def network(self, mode=_abi.MODE_DETERMINISTIC, record_capacity=0):
        net = build(self, mode, record_capacity)
        if self.index == poisoned:
            for p in net._pops:
                p.params["ext_weight"] = -3e38
        return net
    monkeypatch.setattr(serving.JobSpec, "network", network)
    budget = 6 * max(serving.footprint(j) for j in jobs)
    srv = serving.Server(budget, mode=_abi.MODE_DETERMINISTIC, chunk=50, keep_rasters=True,
                         record_per_chunk=1 << 20, synapses_per_neuron=40, stdp_fraction=0.25)
    out = srv.serve(jobs, cancel_at={cancelled: jobs[4].arrival_s + 0.05})
    srv.close()
    res = out["results"]
    assert res[poisoned].status == "failed", res[poisoned]
    assert res[cancelled].status == "cancelled" and res[cancelled].admitted_s >= 0
    monkeypatch.setattr(serving.JobSpec, "network", build)
    for job in jobs:
        if job.index in (poisoned, cancelled):
            continue
        assert res[job.index].status == "done", res[job.index]
        alone = serving.run_alone(job, _abi.MODE_DETERMINISTIC)
        assert first_divergence(res[job.index].raster, alone) is None, job


def test_admission_errors_tell_permanent_from_transient():
    from naer_amd.arena import Arena
    from naer_amd.network import NaerError
    jobs = serving.served_jobs(2, rate_per_s=1.0, seed=8, n_range=(600, 700), k_range=(50, 60),
                               dur_range=(0.01, 0.01), stdp_fraction=0.0)
    a = Arena(1024, 10 * 800 * 60, dt_ms=0.1, mode=_abi.MODE_DETERMINISTIC, delay_steps=(15, 15))
    big = serving.JobSpec(9, 0.0, 1024, 50, 0.01, False, 3)  # 1024 + 1 neurons round up to 1280
    with pytest.raises(NaerError, match="exceeds the arena"):
        a.admit(big.network())
    assert a.admit(jobs[0].network()) is not None
    assert a.admit(jobs[1].network()) is None  # fits once the first retires
    a.close()


@pytest.mark.parametrize("what", ["weight", "v_rest", "ext_weight", "tau"])
def test_non_finite_parameters_are_rejected(what):
    from naer_amd import models
    from naer_amd.network import NaerError
    net = models.balanced(1000, 100, 0.1, 5.0, 20_000.0, 0.1, mode=_abi.MODE_DETERMINISTIC)
    if what == "weight":
        net._conns[0].defaultWeights = float("inf")
    elif what == "tau":
        net._pops[0].params["tau_m_ms"] = float("nan")
    else:
        net._pops[1].params[what] = float("nan")
    with pytest.raises(NaerError, match="finite"):
        net.compile()


@pytest.mark.parametrize("direct", [False, True])
def test_active_tiles_shrink_and_grow_with_the_jobs(direct):
    """The update, the scatter and the bucketed passes cover the arena's tiles
    up to the last occupied one.  Retiring the highest job shrinks that
    prefix (its tiles' spike-list counts go stale), and a job admitted into
    the freed range grows it again: the jobs below it and the new job give
    their alone rasters.  direct: the placement writes the events straight
    into the tile buckets (counters for the occupied prefix only)."""
    from naer_amd.arena import Arena
    a = Arena(16_384, 4 * 3000 * 300, dt_ms=0.1, mode=_abi.MODE_DETERMINISTIC, delay_steps=(15, 15),
              record_capacity=1 << 20)
    if direct:
        sc = a.schedule()
        a.set_schedule(sc["task_synapses"], sc["segments_per_block"], 0)
        assert a.schedule()["refine_split"] == 0
    specs = [serving.JobSpec(i, 0.0, 3000, 300, 0.03, False, 700 + i) for i in range(4)]
    slots = [a.admit(s.network()) for s in specs[:3]]
    assert None not in slots
    begin, size, _ = a.layout()
    assert begin[slots[2]] > begin[slots[1]] > begin[slots[0]]  # first fit, low to high
    got = {s: [] for s in slots}

    def run(n):
        a.run(n)
        for s, sp in a.drain().items():
            got.setdefault(s, []).append(sp)

    run(90)
    a.retire(slots[2])  # the highest job: the covered prefix shrinks
    got.pop(slots[2])
    run(120)
    s3 = a.admit(specs[3].network())  # first fit: into the freed range, the prefix grows again
    assert a.layout()[0][s3] == begin[slots[2]]
    got[s3] = []
    for _ in range(specs[3].steps() // 60 + 1):
        run(60)
    for spec, s in ((specs[0], slots[0]), (specs[1], slots[1]), (specs[3], s3)):
        mixed = np.sort(np.concatenate(got[s]), order=["step", "neuron", "time_us"])
        mixed = mixed[mixed["step"] < spec.steps()]
        alone = serving.run_alone(spec, _abi.MODE_DETERMINISTIC)
        assert len(alone) > 100
        assert first_divergence(mixed, alone) is None, (spec, len(mixed), len(alone))
    a.close()
