"""Parity of the HIP path (through the C ABI) with the CPU oracle.

Bar: bit-exact for RNG draws, connectivity and every integer quantity; in
NAER_MODE_DETERMINISTIC also bit-exact potentials, rasters and inputs (the
32.32 fixed-point ring makes the sums order independent and the oracle
reproduces the fp32 arithmetic of the update).  NAER_MODE_FAST sums with fp32
atomics in hardware order: its per-step inputs must agree with the exact
(fp64) sum of the same events to 1e-6 of the summed |w|, its rasters match
until the first floating-point divergence.
"""
import numpy as np
import pytest

import oracle as O
from helpers import first_divergence, oracle_of, raster
from naer_amd import _abi, models
from naer_amd.network import philox as gpu_philox, ref_poisson_epoch as gpu_poisson_epoch

pytestmark = pytest.mark.gpu

SEED32 = 0x15AC7EBD


def test_philox_device_vs_oracle(golden):
    rng = np.random.default_rng(7)
    ctr = rng.integers(0, 2**32, (200_000, 4), dtype=np.uint64).astype(np.uint32)
    key = rng.integers(0, 2**32, (200_000, 2), dtype=np.uint64).astype(np.uint32)
    np.testing.assert_array_equal(gpu_philox(ctr, key), O.philox(ctr, key))
    for v in golden["philox_kat"]:
        out = gpu_philox([int(x, 16) for x in v["ctr"]], [int(x, 16) for x in v["key"]])
        assert [int(x) for x in out[0]] == [int(x, 16) for x in v["out"]]


@pytest.mark.parametrize("t0", [0, 1500, 4_500_000, 9_998_500])
def test_reference_poisson_epoch_bit_exact(t0, golden):
    n, n_pois = 20_000, 22_360
    times, count = gpu_poisson_epoch(n, n_pois, SEED32, 20.0, t0, t0 + 1500)
    for idx in range(n):
        ref = O.ref_poisson_epoch(idx, n_pois, SEED32, 20.0, t0, t0 + 1500)
        assert count[idx] == len(ref), idx
        assert np.array_equal(times[idx, : count[idx]], ref), idx
    if t0 == 0:
        want = golden["poisson_first_isi"]["times"]
        big, cnt = gpu_poisson_epoch(8, n_pois, SEED32, 20.0, 0, 10**9)
        assert [int(x) for x in big[:, 0]] == want


def test_bernoulli_connectivity_brunel_1e6(golden):
    net = models.brunel(1e6)
    net.compile()
    g = golden["brunel_counts"][0]
    assert net.getNeurons() == g["N"]
    assert net.getSynapses() == g["S"]
    rp, col, w = net.csr()
    o = oracle_of(net)
    orp, ocol, ow = o.csr()
    np.testing.assert_array_equal(rp, orp)
    np.testing.assert_array_equal(col, ocol)
    np.testing.assert_array_equal(w.view(np.uint32), ow.view(np.uint32))
    # the P->E rows of the reference probe: POISSON neurons are addressed after
    # the NOP populations ([STDP | NOP | POISSON], Network.cpp:269-285)
    P, E = net._pops[0], net._pops[1]
    for src, key in ((0, "src0_below200"), (1, "src1_below200")):
        s = P.address + src
        row = col[rp[s]:rp[s + 1]] & 0xFFFFFF
        dst = sorted(int(x) - E.address for x in row if E.address <= x < E.address + 200)
        assert dst == golden["bernoulli_rows"][key]


def test_fixed_indegree_connectivity_random_weights_and_delays():
    net = models.plastic_100k(n=3000, indegree=300)
    net.compile()
    rp, col, w = net.csr()
    o = oracle_of(net)
    orp, ocol, ow = o.csr()
    assert net.getSynapses() == 3000 * 300
    np.testing.assert_array_equal(rp, orp)
    np.testing.assert_array_equal(col, ocol)
    np.testing.assert_array_equal(w.view(np.uint32), ow.view(np.uint32))
    # every target has exactly its in-degree
    tgt = col & 0xFFFFFF
    assert np.all(np.bincount(tgt, minlength=3000) == 300)
    delays = (col >> 24) & 0x7F
    assert set(np.unique(delays)) == {10, 20, 30, 40, 50}


def test_multi_group_network_csr_and_deterministic_step():
    """50,000 neurons span four 2^14-neuron groups: the canonical row order
    (group, word, weight) and the bucketed delivery over several groups and
    arrival batches, bit-exact with the oracle in deterministic mode."""
    net = models.large_1m(n=50_000, indegree=200, mode=_abi.MODE_DETERMINISTIC, record_capacity=4_000_000)
    net.compile()
    rp, col, w = net.csr()
    o = oracle_of(net, O.FIXED)
    orp, ocol, ow = o.csr()
    np.testing.assert_array_equal(rp, orp)
    np.testing.assert_array_equal(col, ocol)
    np.testing.assert_array_equal(w.view(np.uint32), ow.view(np.uint32))
    assert int(np.max(col & 0xFFFFFF)) >= 3 << 14
    net.run(600)
    o.run(600)
    gs, os_ = net.spikes(), o.spikes()
    assert len(gs) > 1000
    assert first_divergence(gs, os_) is None
    np.testing.assert_array_equal(net.state()["v"].view(np.uint32), o.state()["v"].view(np.uint32))
    assert net.stats().events == o.event_count
    net.close()
    o.close()


def _run_pair(net, steps, precision, probe=None):
    net.record_capacity = 4_000_000
    net.compile()
    o = oracle_of(net, precision)
    if probe:
        net.set_probe(*probe)
        o.set_probe(*probe)
    net.run(steps)
    o.run(steps)
    return o


def test_deterministic_mode_bit_exact_balanced():
    net = models.balanced_10k(mode=_abi.MODE_DETERMINISTIC)
    o = _run_pair(net, 1200, O.FIXED, probe=(400, 100))
    gs, os_ = net.spikes(), o.spikes()
    assert len(gs) > 1000
    assert first_divergence(gs, os_) is None
    st, ost = net.state(), o.state()
    np.testing.assert_array_equal(st["v"].view(np.uint32), ost["v"].view(np.uint32))
    np.testing.assert_array_equal(st["ref"], ost["ref"])
    used, exact, absw = o.probe()
    gp = net.probe()
    np.testing.assert_array_equal(gp.view(np.uint32), used.view(np.uint32))
    # per-step synaptic input sums vs the fp64 sum of the same events
    err = np.abs(gp.astype(np.float64) - exact)
    assert np.all(err <= 1e-6 * np.maximum(absw, 1e-30) + 1e-12)
    s = net.stats()
    assert s.spikes == o.spike_count and s.events == o.event_count


def test_fast_mode_inputs_and_raster():
    net = models.balanced_10k(mode=_abi.MODE_FAST)
    o = _run_pair(net, 600, O.FLOAT, probe=(0, 60))
    gs, os_ = net.spikes(), o.spikes()
    d = first_divergence(gs, os_)
    # rasters agree at least until the first step whose inputs differ
    gp = net.probe()
    used, exact, absw = o.probe()
    err = np.abs(gp.astype(np.float64) - exact)
    assert np.all(err <= 1e-6 * np.maximum(absw, 1e-30) + 1e-12)
    if d is not None:
        assert raster(gs)[d][0] > 20  # no divergence within the first 2 ms
    # population rate within 2% of the oracle's
    assert abs(len(gs) - len(os_)) <= 0.02 * len(os_)


def test_reference_brunel_model_with_poisson_sources():
    net = models.brunel(2e6, mode=_abi.MODE_DETERMINISTIC)
    o = _run_pair(net, 600, O.FIXED)
    gs, os_ = net.spikes(), o.spikes()
    pois = [p for p in net._pops if p.type == "POISSON"][0]
    is_p = (gs["neuron"] >= pois.address) & (gs["neuron"] < pois.address + pois.size)
    assert is_p.sum() > 100
    assert first_divergence(gs, os_) is None
    # POISSON spike times are the reference's train, epoch by epoch
    n_pois = pois.size
    for sp in gs[is_p][:300]:
        idx = int(sp["neuron"]) - pois.address
        t0 = (int(sp["time_us"]) // 1500) * 1500
        tr = O.ref_poisson_epoch(idx, n_pois, SEED32, 20.0, t0, t0 + 1500)
        assert int(sp["time_us"]) in set(int(x) for x in tr)


def test_stdp_deterministic_bit_exact():
    net = models.plastic_100k(n=2000, indegree=200, mode=_abi.MODE_DETERMINISTIC)
    o = _run_pair(net, 400, O.FIXED)
    gs, os_ = net.spikes(), o.spikes()
    assert len(gs) > 100
    assert first_divergence(gs, os_) is None
    _, _, w = net.csr()
    _, _, ow = o.csr()
    np.testing.assert_array_equal(w.view(np.uint32), ow.view(np.uint32))
    st, ost = net.state(), o.state()
    np.testing.assert_array_equal(st["pre"].view(np.uint32), ost["pre"].view(np.uint32))
    np.testing.assert_array_equal(st["post"].view(np.uint32), ost["post"].view(np.uint32))
