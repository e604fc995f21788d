"""The benchmark configurations at their stated sizes (deterministic mode).

  plasticity   N = 1e5, in-degree 1000 (1e8 synapses), pair STDP on E->E:
               raster, plastic weights and traces bit-exact with the oracle
               over 300 steps;
  sweep        all 256 networks of the sweep (N log-uniform in 1k-50k,
               K = 0.1 N, half with STDP) packed as the sweep packs them,
               every raster identical to the network alone, and the two
               smallest (one static, one plastic) identical to the oracle;
  served       jobs drawn with the served stream's ranges (2k-200k neurons,
               in-degree 500-2000, 30 % STDP; durations shortened) through
               the continuous-batching server with the 200 GB budget: every
               job's raster equals its raster alone.
"""
import numpy as np
import pytest

import oracle as O
from helpers import first_divergence, oracle_of
from naer_amd import _abi, models, serving, sweep

pytestmark = pytest.mark.gpu


def test_plasticity_network_full_size_bit_exact():
    net = models.plastic_100k(mode=_abi.MODE_DETERMINISTIC, record_capacity=4_000_000)
    net.compile()
    assert net.getNeurons() == 100_000 and net.getSynapses() == 100_000_000
    o = oracle_of(net, O.FIXED, threads=16)
    net.run(300)
    o.run(300)
    gs, os_ = net.spikes(), o.spikes()
    assert len(gs) > 1000
    assert first_divergence(gs, os_) is None
    st, ost = net.state(), o.state()
    for k in ("v", "ref", "pre", "post"):
        np.testing.assert_array_equal(st[k].view(np.uint32), ost[k].view(np.uint32), err_msg=k)
    _, col, w = net.csr()
    _, ocol, ow = o.csr()
    np.testing.assert_array_equal(col, ocol)
    np.testing.assert_array_equal(w.view(np.uint32), ow.view(np.uint32))
    assert net.stats().events == o.event_count
    net.close()
    o.close()


SWEEP_STEPS = 300


def test_sweep_all_256_networks_packed_equal_alone():
    specs = sweep.sweep_specs(256)
    packed = sweep.run_packed(specs, SWEEP_STEPS, _abi.MODE_DETERMINISTIC)
    seq = sweep.run_sequential(specs, SWEEP_STEPS, _abi.MODE_DETERMINISTIC)
    same = [s["index"] for s in specs
            if first_divergence(packed["rasters"][s["index"]], seq["rasters"][s["index"]]) is None]
    assert len(same) == 256
    assert sum(len(r) for r in seq["rasters"].values()) > 100_000
    for plastic in (False, True):
        s = min((x for x in specs if x["plastic"] == plastic), key=lambda x: x["n"])
        o = oracle_of(sweep.make_network(s, _abi.MODE_DETERMINISTIC, sweep.expected_record(s, SWEEP_STEPS)), O.FIXED,
                      threads=16)
        o.run(SWEEP_STEPS)
        assert first_divergence(seq["rasters"][s["index"]], o.spikes()) is None, s
        o.close()


def test_served_stream_jobs_of_the_stated_sizes_mix_equals_alone():
    jobs = serving.served_jobs(8, rate_per_s=50.0, seed=0xC0F4, dur_range=(0.02, 0.04))
    assert max(j.n for j in jobs) > 20_000 and max(j.indegree for j in jobs) > 1000
    assert any(j.plastic for j in jobs) and not all(j.plastic for j in jobs)
    srv = serving.Server(200e9, mode=_abi.MODE_DETERMINISTIC, chunk=100, keep_rasters=True)
    out = srv.serve(jobs)
    srv.close()
    assert out["completed"] == len(jobs), out["statuses"]
    res = out["results"]
    # several jobs were resident together
    spans = sorted((r.admitted_s, r.done_s) for r in res.values())
    assert any(b[0] < a[1] for a, b in zip(spans, spans[1:]))
    for job in jobs:
        alone = serving.run_alone(job, _abi.MODE_DETERMINISTIC)
        assert len(alone) > 0
        assert first_divergence(res[job.index].raster, alone) is None, job
