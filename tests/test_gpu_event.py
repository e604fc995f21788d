"""The event-driven mode (NAER_MODE_EVENT: the reference's epoch,
src/Network.cpp:122-256, and neuron step, lif.cl:15-34, with stdp.cl's
plasticity) on the GPU equals the oracle's restatement (orc_ev_run, pinned by
test_oracle_event.py and test_oracle_event_stdp.py) bit for bit: rasters
(window, neuron, microsecond), final potentials and, with plastic
connections, weights and post traces, on the reference's brunel and brunel+
models and fixed in-degree networks with equal-time inputs."""
import numpy as np
import pytest

from helpers import first_divergence, oracle_of
from naer_amd import _abi, models
from naer_amd.network import Network, NeuronType

pytestmark = pytest.mark.gpu


def compare(net, windows, chunks=None):
    """Raster, potentials and counts equal the oracle's; with plastic
    connections also every weight and the post traces (bit for bit)."""
    o = oracle_of(net, threads=1)
    o.run_events(windows)
    want, want_st = o.spikes(), o.event_state()
    net.compile()
    for c in chunks or (windows,):
        net.run(c)
    net.finish()
    got = net.spikes()
    st = net.stats()
    assert first_divergence(got, want) is None
    state = net.state()
    assert np.array_equal(state["v"], want_st["v"])
    assert st.step == windows and st.time_us == windows * 1500
    assert st.spikes == o.spike_count and st.events == o.event_count
    if st.n_plastic:
        np.testing.assert_array_equal(state["post"].view(np.uint32), want_st["trace"].view(np.uint32))
        _, col, w = net.csr()
        _, ocol, ow = o.csr()
        np.testing.assert_array_equal(col, ocol)
        np.testing.assert_array_equal(w.view(np.uint32), ow.view(np.uint32))
    o.close()
    return got


@pytest.mark.parametrize("synapses", [1_000_000, 10_000_000])
def test_brunel_event_mode_equals_oracle(synapses):
    net = models.brunel(synapses, mode=_abi.MODE_EVENT, record_capacity=1 << 21)
    sp = compare(net, 40, chunks=(13, 1, 26))
    n_lif = net.getNeurons() - sum(p.size for p in net._pops if p.type == NeuronType.POISSON)
    assert np.count_nonzero(sp["neuron"] < n_lif) > 100  # the LIF populations fire


@pytest.mark.parametrize("synapses", [1_000_000, 10_000_000])
def test_brunel_plus_event_mode_stdp_equals_oracle(synapses):
    """main.cpp:12-30 with stdp.cl's plasticity on E -> E: weights, traces,
    potentials and raster bit-exact."""
    net = models.brunel_plus(synapses, mode=_abi.MODE_EVENT, record_capacity=1 << 21)
    sp = compare(net, 40, chunks=(13, 1, 26))
    assert net.stats().n_plastic > 0.1 * net.getSynapses()
    assert len(sp) > 1000


def test_event_stdp_with_refractory_inputs_and_clamping():
    """STDP populations with 1.5 ms refractory sources (inputs inside the
    targets' 2 ms refractory period clear their traces), weights drawn above
    the 0.3 mV bound (clamped at their first update), and equal-time inputs."""
    net = Network(dt_ms=0.1, mode=_abi.MODE_EVENT, record_capacity=1 << 20)
    P = net.addPopulation(NeuronType.POISSON, 400, src_rate_hz=3000.0)
    E = net.addPopulation(NeuronType.STDP, 300, tau_m_ms=20.0, v_th=0.002, v_reset=0.0, v_rest=0.0, t_ref_ms=1.5,
                          v0_min=0.0, v0_max=0.002)
    F = net.addPopulation(NeuronType.STDP, 100, tau_m_ms=20.0, v_th=0.002, v_reset=0.0, v_rest=0.0, t_ref_ms=2.0,
                          v0_min=0.0, v0_max=0.002)
    net.addConnectionFixedIndegree(P, E, 40, 0.0, 0.0, 0.0004, seed=11)
    net.addConnectionFixedIndegree(P, F, 40, 0.0, 0.0, 0.0004, seed=12)
    net.addConnectionFixedIndegree(E, F, 30, 0.0, 0.0001, 0.0008, seed=13)  # plastic, partly above the bound
    net.addConnectionFixedIndegree(E, E, 10, 0.0, -0.0002, 0.0005, seed=14)  # plastic
    net.addConnectionFixedIndegree(F, E, 10, -0.0003, seed=15)  # plastic (F fires at least 2 ms apart)
    sp = compare(net, 30)
    assert np.count_nonzero(sp["neuron"] < 400) > 300


def test_fixed_indegree_event_mode_with_equal_time_inputs():
    """Many sources at a high rate onto few targets: inputs of one target at
    the same microsecond are common, and their order (source, row position)
    decides the float sums."""
    net = Network(dt_ms=0.1, mode=_abi.MODE_EVENT, record_capacity=1 << 20)
    P = net.addPopulation(NeuronType.POISSON, 3000, src_rate_hz=2000.0)
    L = net.addPopulation(NeuronType.NOP, 200, tau_m_ms=10.0, v_th=20.0, v_reset=10.0, v_rest=0.0, t_ref_ms=1.0,
                          v0_min=0.0, v0_max=20.0)
    net.addConnectionFixedIndegree(P, L, 600, 0.0, 0.05, 0.3, seed=5)  # weight 0: drawn in [0.05, 0.3)
    net.addConnectionFixedIndegree(P, L, 200, 0.0, -0.4, -0.1, seed=6)
    net.addConnectionFixedIndegree(L, L, 20, 0.5, seed=7)
    sp = compare(net, 25)
    assert np.count_nonzero(sp["neuron"] < 200) > 200


@pytest.mark.parametrize("t_ref_ms,v_rest", [(0.0, 0.0), (0.5, -5.0), (3.0, 2.0)])
def test_event_mode_refractory_and_rest_edges(t_ref_ms, v_rest):
    """t_ref = 0 lets a neuron fire several times per window (and at the same
    microsecond as its input); v_rest != 0 decays towards it."""
    net = Network(dt_ms=0.1, mode=_abi.MODE_EVENT, record_capacity=1 << 20)
    P = net.addPopulation(NeuronType.POISSON, 1000, src_rate_hz=1500.0)
    L = net.addPopulation(NeuronType.NOP, 300, tau_m_ms=15.0, v_th=12.0, v_reset=v_rest + 1.0, v_rest=v_rest,
                          t_ref_ms=t_ref_ms, v0_min=v_rest, v0_max=12.0)
    net.addConnectionFixedIndegree(P, L, 200, 0.0, 0.1, 0.6, seed=8)
    net.addConnectionFixedIndegree(L, L, 30, -0.7, seed=9)
    sp = compare(net, 30)
    lif = sp[sp["neuron"] < 300]
    assert len(lif) > 300
    if t_ref_ms == 0.0:
        per = np.unique(np.stack([lif["neuron"], lif["step"]], 1), axis=0, return_counts=True)[1]
        assert per.max() > 1  # several spikes of one neuron in one window


def test_event_mode_rejects_what_it_does_not_model(tmp_path):
    bad = Network(dt_ms=0.1, mode=_abi.MODE_EVENT)
    bad.addPopulation(NeuronType.NOP, 10, ext_rate_hz=1000.0, ext_weight=0.1)
    with pytest.raises(_abi.NaerError, match="without drive"):
        bad.compile()
    fast = Network(dt_ms=0.1, mode=_abi.MODE_EVENT)
    S = fast.addPopulation(NeuronType.STDP, 10, t_ref_ms=1.0)
    fast.addConnectionFixedIndegree(S, S, 2, 0.0001)
    with pytest.raises(_abi.NaerError, match="at least 1.5 ms"):
        fast.compile()
    net = models.brunel(100_000, mode=_abi.MODE_EVENT)
    net.compile()
    with pytest.raises(_abi.NaerError, match="graphs"):
        net.set_graph(100)
    with pytest.raises(_abi.NaerError, match="checkpointed"):
        net.save(tmp_path / "x.ckpt")
    net.step()  # one 1500 us window
    net.finish()
    assert net.getTime() == 1500
