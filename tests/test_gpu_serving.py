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
                         record_per_chunk=1 << 20, synapses_per_neuron=40)
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
