"""Lazy pair STDP on the device against the oracle's eager statement.

The device applies potentiation when the pre-synaptic neuron next fires (or
at a flush of all plastic synapses, before any neuron can have fired 32 times
since the last one, and before weights are read), from each neuron's last 32
spikes; plastic nets run batches of B <= the fewest refractory steps + 1
steps.  The oracle potentiates at every post-synaptic spike through a reverse
index.  These cases drive the device through what the lazy form depends on:
partial batches (run lengths off the batch grid, delivered through the
ring), graph-captured chunks (the flush condition is evaluated on the
device), weight reads in mid-run (a forced flush), many flush windows, and
short refractory periods (short batches, frequent flushes).

Bar: NAER_MODE_DETERMINISTIC rasters, potentials, STDP traces and weights
bit-exact with the oracle's fixed-point step.
"""
import numpy as np
import pytest

import oracle as O
from helpers import first_divergence, oracle_of
from naer_amd import _abi, models
from naer_amd.network import Network, NeuronType

pytestmark = pytest.mark.gpu


def plastic_net(n=3000, indegree=300, t_ref_ms=2.0, delay_ms=(1.0, 5.0, 1.0), drive_hz=15_000.0, seed=0x57D9):
    """Balanced E/I net with pair STDP on E->E (weights uniform in [0, wmax])."""
    net = Network(seed=seed, mode=_abi.MODE_DETERMINISTIC, stdp=models.PAIR_STDP, record_capacity=2_000_000)
    lif = dict(models.BRUNEL_LIF, ext_rate_hz=drive_hz, ext_weight=0.1, t_ref_ms=t_ref_ms)
    n_e = int(n * 0.8)
    k_e = int(indegree * 0.8)
    E = net.addPopulation(NeuronType.STDP, n_e, **lif)
    I = net.addPopulation(NeuronType.NOP, n - n_e, **lif)
    wmax = net.stdp["w_max"]
    net.addConnectionFixedIndegree(E, E, k_e, 0.0, 0.0, wmax, seed=seed + 1, delay_ms=delay_ms)
    net.addConnectionFixedIndegree(E, I, k_e, 0.1, seed=seed + 2, delay_ms=delay_ms)
    net.addConnectionFixedIndegree(I, E, indegree - k_e, -0.5, seed=seed + 3, delay_ms=delay_ms)
    net.addConnectionFixedIndegree(I, I, indegree - k_e, -0.5, seed=seed + 4, delay_ms=delay_ms)
    return net


def assert_same(net, o, min_spikes=1000):
    gs, os_ = net.spikes(), o.spikes()
    assert len(gs) > min_spikes
    assert first_divergence(gs, os_) is None
    st, ost = net.state(), o.state()
    for k in ("v", "pre", "post"):
        np.testing.assert_array_equal(st[k].view(np.uint32), ost[k].view(np.uint32), err_msg=k)
    np.testing.assert_array_equal(st["ref"], ost["ref"])
    _, _, w = net.csr()
    _, _, ow = o.csr()
    np.testing.assert_array_equal(w.view(np.uint32), ow.view(np.uint32))


def test_irregular_runs_graphs_and_weight_reads():
    net = plastic_net()
    net.compile()
    s = net.stats()
    assert s.batch == 10  # d_min 10 steps, 20 refractory steps
    o = oracle_of(net, O.FIXED)
    done = 0
    for c in [7, 23, 100, 1, 59, 333]:
        net.run(c)
        o.run(c)
        done += c
        if c == 100:  # a weight read in mid-run (forced flush), then on
            _, _, w = net.csr()
            _, _, ow = o.csr()
            np.testing.assert_array_equal(w.view(np.uint32), ow.view(np.uint32))
    net.set_graph(40)  # chunks of 40 steps; the flush condition is evaluated on the device
    for c in [45, 400, 81, 800]:
        net.run(c)
        o.run(c)
        done += c
    assert done > 2 * 31 * 21  # several flush windows of 651 steps
    assert_same(net, o)
    net.close()


@pytest.mark.parametrize("t_ref_ms, batch", [(0.3, 4), (0.0, 1)])
def test_short_refractory_short_batches(t_ref_ms, batch):
    # B = refractory steps + 1 (4 or 1) and a flush every 31 * B steps
    net = plastic_net(n=2000, indegree=200, t_ref_ms=t_ref_ms, drive_hz=18_000.0)
    net.compile()
    assert net.stats().batch == batch
    o = oracle_of(net, O.FIXED)
    for c in [13, 250, 3, 334]:
        net.run(c)
        o.run(c)
    assert_same(net, o)
    net.close()
