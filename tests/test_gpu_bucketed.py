"""Bucketed spike delivery (aligned batches of B steps whose events are sorted
by arrival batch and target tile, then summed in LDS by the update) against
the oracle, in the cases where it hands over to the delay ring: partial
batches from run lengths that are not multiples of B, batches that exceed the
event buffer, graph-captured chunks, and at the full 1M-neuron size the
conservation of events and run-to-run determinism.

Bar: NAER_MODE_DETERMINISTIC rasters, potentials and counters bit-exact with
the oracle's 32.32 fixed-point step.
"""
import ctypes
import os

import numpy as np
import pytest

import oracle as O
from helpers import first_divergence, oracle_of
from naer_amd import _abi, models
from naer_amd.network import Network, NeuronType

pytestmark = pytest.mark.gpu


def multi_delay_net(record=2_000_000, plastic=False, random_weights=False):
    """8,000 neurons, in-degree 800, delays 1.5-6 ms on a 0.5 ms grid: B = 15
    steps per batch, events spread over M = 4 later batches.  With pair STDP
    on E->E (plastic=True) the batches are as long (refractory 20 steps) and the
    placement applies lazy STDP.  Constant weights per source population are
    delivered as 4-byte class records; random_weights=True (E->E weights
    uniform in [0.05, 0.15) mV) keeps the 8-byte records with weights."""
    if random_weights:
        net = Network(mode=_abi.MODE_DETERMINISTIC, seed=0xB0C4E7)
        lif = dict(models.BRUNEL_LIF, ext_rate_hz=20_000.0, ext_weight=0.1)
        E = net.addPopulation(NeuronType.NOP, 6400, **lif)
        I = net.addPopulation(NeuronType.NOP, 1600, **lif)
        d = (1.5, 6.0, 0.5)
        net.addConnectionFixedIndegree(E, E, 640, 0.0, 0.05, 0.15, seed=11, delay_ms=d)
        net.addConnectionFixedIndegree(E, I, 640, 0.1, seed=12, delay_ms=d)
        net.addConnectionFixedIndegree(I, E, 160, -0.5, seed=13, delay_ms=d)
        net.addConnectionFixedIndegree(I, I, 160, -0.5, seed=14, delay_ms=d)
    else:
        net = models.balanced(8000, 800, 0.1, 5.0, 20_000.0, 0.1, delay_ms=(1.5, 6.0, 0.5), plastic=plastic,
                              mode=_abi.MODE_DETERMINISTIC, seed=0xB0C4E7)
    net.record_capacity = record
    return net


def assert_same(net, o):
    gs, os_ = net.spikes(), o.spikes()
    assert len(gs) > 500
    assert first_divergence(gs, os_) is None
    st, ost = net.state(), o.state()
    np.testing.assert_array_equal(st["v"].view(np.uint32), ost["v"].view(np.uint32))
    np.testing.assert_array_equal(st["ref"], ost["ref"])
    s = net.stats()
    assert s.spikes == o.spike_count and s.events == o.event_count


def test_batch_shape():
    net = multi_delay_net()
    net.compile()
    s = net.stats()
    assert s.batch == 15 and s.ring_depth == 61
    assert net.schedule()["record_bytes"] == 4
    net.close()
    for kw in (dict(random_weights=True), dict(plastic=True)):
        net = multi_delay_net(**kw)
        net.compile()
        assert net.schedule()["record_bytes"] == 8, kw
        net.close()


@pytest.mark.parametrize("random_weights", [False, True])
def test_irregular_run_lengths_bit_exact(random_weights):
    net = multi_delay_net(random_weights=random_weights)
    net.compile()
    o = oracle_of(net, O.FIXED)
    chunks = [1, 14, 15, 7, 30, 23, 45, 3, 12, 100, 50, 15, 15, 29, 1, 140]
    for c in chunks:
        net.run(c)
    o.run(sum(chunks))
    assert_same(net, o)
    # partial launches defer their spikes to the launch that completes the
    # aligned batch: one bucketed delivery per batch, none through the ring
    s = net.stats()
    assert s.kernel_calls[_abi.K_BUCKET_PLACE] == sum(chunks) // 15
    assert s.kernel_calls[_abi.K_SCATTER] == sum(chunks) // 15
    net.close()


@pytest.mark.parametrize("cap,random_weights,reserve", [(2_000, False, 0), (150_000, False, 0), (150_000, True, 0),
                                                        (2_000, False, 1), (400_000, False, 1)])
def test_event_buffer_overflow_falls_back_to_ring(cap, random_weights, reserve):
    # a small event buffer: every batch (2_000) or the busier ones (150_000;
    # 400_000 with reserved placement, whose layout needs about twice the
    # events) go through the ring from the placement kernel
    os.environ["NAER_BUCKET_EVENTS"] = str(cap)
    try:
        net = multi_delay_net(random_weights=random_weights)
        net.compile()
    finally:
        del os.environ["NAER_BUCKET_EVENTS"]
    if reserve:
        s = net.schedule()
        net.set_schedule(s["task_synapses"], s["segments_per_block"], 4, reserve=1)
    o = oracle_of(net, O.FIXED)
    net.run(450)
    o.run(450)
    if cap < 10_000:
        assert net.stats().events / 30 > cap  # no batch fits
    assert_same(net, o)
    net.close()


def test_graph_chunks_with_unaligned_runs():
    net = multi_delay_net()
    net.compile()
    net.set_graph(25)  # captured chunks round up to whole batches (30 steps)
    o = oracle_of(net, O.FIXED)
    chunks = [37, 61, 100, 8, 94]
    for c in chunks:
        net.run(c)
    o.run(sum(chunks))
    assert_same(net, o)
    net.close()


def _row_ptr(net):
    rp = np.zeros(net.getNeurons() + 1, np.uint64)
    _abi.check(_abi.lib().naer_net_read_csr(net._h(), rp.ctypes.data, None, None))
    return rp


def test_full_size_events_conserved_and_deterministic():
    """1M neurons, 2e9 synapses: every recorded spike delivers exactly its
    out-degree, and two runs, the second with the other placement (counted
    two-level or reserved), give the same raster bit for bit."""
    rasters = []
    for k in range(2):
        net = models.large_1m(mode=_abi.MODE_DETERMINISTIC, record_capacity=2_000_000)
        net.compile()
        s = net.schedule()
        assert s["parts_max"] >= 4
        if k == 0:
            assert s["reserve"] == 1  # the default of this net (558 group buckets)
        if k == 1:
            res = 1 - s["reserve"]
            net.set_schedule(s["task_synapses"], s["segments_per_block"], 4 if res else 3, reserve=res)
        net.run(400)
        net.finish()
        net.spikes()  # drain the warm-up
        e0 = net.stats().events
        net.run(300)
        net.finish()
        sp = net.spikes()
        de = net.stats().events - e0
        rp = _row_ptr(net)
        deg = (rp[1:] - rp[:-1]).astype(np.int64)
        assert len(sp) > 10_000
        assert int(deg[sp["neuron"]].sum()) == de
        assert deg.sum() == 2_000_000_000
        rasters.append(sp)
        net.close()
    assert first_divergence(rasters[0], rasters[1]) is None



@pytest.mark.parametrize("kw", [{}, {"plastic": True}, {"random_weights": True}])
def test_direct_placement_matches_oracle(kw):
    """Schedule refine_split 0: the placement writes every event straight into
    its tile bucket (no refine pass), for 4-byte class records, 8-byte
    records with weights and plastic rows (lazy STDP in the placement)."""
    net = multi_delay_net(**kw)
    net.compile()
    o = oracle_of(net, O.FIXED)
    s = net.schedule()
    net.set_schedule(s["task_synapses"], s["segments_per_block"], 0)
    assert net.schedule()["refine_split"] == 0
    net.run(463)  # a partial batch at the end
    o.run(463)
    assert_same(net, o)
    if kw.get("plastic"):
        np.testing.assert_array_equal(net.csr()[2].view(np.uint32), o.csr()[2].view(np.uint32))
    net.close()


@pytest.mark.parametrize("kw,margin", [({}, None), ({"plastic": True}, None), ({"random_weights": True}, None),
                                       ({}, "0.3"), ({"plastic": True}, "0.3"), ({"random_weights": True}, "0")])
def test_reserved_placement_matches_oracle(kw, margin):
    """Reserved placement (no counting passes: regions reserved from the
    batch's event count and the net's synapse shares, a counting sort per
    part of a group bucket): 4-byte class records, 8-byte records with
    weights, plastic rows.  A margin of 0.3 sends the events beyond the
    regions into the spill regions (and some through the ring), 0 sends all
    but 8 per region and 64 per group bucket through the ring."""
    if margin is not None:
        os.environ["NAER_RESERVE_MARGIN"] = margin
    try:
        net = multi_delay_net(**kw)
        net.compile()
    finally:
        os.environ.pop("NAER_RESERVE_MARGIN", None)
    o = oracle_of(net, O.FIXED)
    s = net.schedule()
    assert s["parts_max"] >= 4 and s["reserve"] in (0, 1)
    for parts in (4, 1, s["parts_max"]):
        net.set_schedule(s["task_synapses"], s["segments_per_block"], parts, reserve=1)
        assert net.schedule()["reserve"] == 1 and net.schedule()["refine_split"] == parts
        net.run(151)  # ends inside a batch (plastic nets deliver it through the ring)
    o.run(453)
    assert_same(net, o)
    if kw.get("plastic"):
        np.testing.assert_array_equal(net.csr()[2].view(np.uint32), o.csr()[2].view(np.uint32))
    net.close()


def test_reserved_placement_switches_and_graphs():
    """Batches in flight keep the layout they were placed with when the
    schedule switches between reserved and counted placement; deferred
    launches and graph-captured chunks with reserved placement."""
    net = multi_delay_net()
    net.compile()
    o = oracle_of(net, O.FIXED)
    s = net.schedule()
    total = 0
    for i, c in enumerate([37, 16, 61, 15, 8, 94, 45]):
        net.set_schedule(s["task_synapses"], s["segments_per_block"], 2 if i % 2 == 0 else s["refine_split"],
                         reserve=1 if i % 2 == 0 else 0)
        net.run(c)
        total += c
    net.set_schedule(s["task_synapses"], s["segments_per_block"], 3, reserve=1)
    net.set_graph(25)  # whole batches: 30 steps
    for c in [37, 61, 100]:
        net.run(c)
        total += c
    o.run(total)
    assert_same(net, o)
    net.close()


def test_reserved_placement_refused_where_unavailable():
    net = multi_delay_net()
    net.compile()
    s = net.schedule()
    with pytest.raises(_abi.NaerError):
        net.set_schedule(s["task_synapses"], s["segments_per_block"], s["parts_max"] + 1, reserve=1)
    with pytest.raises(_abi.NaerError):
        net.set_schedule(s["task_synapses"], s["segments_per_block"], 0, reserve=1)
    net.close()


def test_direct_placement_refused_beyond_its_counters():
    """More tile buckets than the placement's 8192 LDS counters (1M neurons in
    1024-neuron tiles x 9 arrival batches): direct placement is refused."""
    net = models.large_1m(indegree=10)
    net.compile()
    s = net.schedule()
    assert s["bucketed"] == 1
    with pytest.raises(_abi.NaerError):
        net.set_schedule(s["task_synapses"], s["segments_per_block"], 0)
    net.set_schedule(s["task_synapses"], s["segments_per_block"], s["refine_split"])
    net.close()


def test_autotune_and_schedules_leave_results_unchanged():
    net = multi_delay_net()
    net.compile()
    o = oracle_of(net, O.FIXED)
    net.run(100)
    chosen = net.autotune(1)  # real steps, aligned batches
    assert chosen["bucketed"] == 1 and chosen["task_synapses"] in (256, 512)
    steps = int(net.stats().step)
    assert steps > 100
    net.set_schedule(256, 1, 3, run_aggregate=1, reserve=0)
    net.run(60)
    net.set_schedule(512, 2, 2, run_aggregate=2, reserve=1)  # reserved placement, 2 parts
    net.run(40)
    net.set_schedule(512, 4, 1, run_aggregate=3, reserve=0)
    net.run(40)
    net.set_schedule(512, 4, 0, run_aggregate=0, reserve=0)  # direct placement
    net.run(40)
    net.set_schedule(512, 4, 1, run_aggregate=0, reserve=0)
    net.run(200)
    o.run(steps + 380)
    assert_same(net, o)
    with pytest.raises(Exception):
        net.set_schedule(128, 1, 1)
    with pytest.raises(Exception):
        net.set_schedule(256, 1, 1, run_aggregate=4)
    net.close()
