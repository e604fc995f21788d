"""The oracle's event-driven mode (orc_ev_run: the reference's epoch,
src/Network.cpp:122-256, and neuron step, lif.cl:15-34) pinned on networks
small enough for closed forms:
  * sub-threshold: V after the inputs of one source equals sum_i w e^(-(t_last - t_i)/tau);
  * refractory: with every input above threshold the neuron fires at exactly
    the inputs separated by at least t_ref from the previous spike (inputs in
    between are dropped);
  * delay: a spike at t reaches its targets at t + 1500 us;
  * the decay tables equal exp(-dt/tau) to 1e-12 relative (both round the
    argument) and underflow where exp does.
POISSON trains come from the reference generator (pinned by golden vectors
in test_oracle_reference_vectors.py)."""
import math

import numpy as np
import pytest

import oracle as O
from helpers import oracle_of
from naer_amd import _abi
from naer_amd.network import Network, NeuronType

TAU = 20.0
WINDOW = 1500


def chain(w1, w2=None, v_th=1.0, rate=3000.0, t_ref=2.0, n_src=1):
    net = Network(dt_ms=0.1, mode=_abi.MODE_EVENT, record_capacity=1 << 16)
    P = net.addPopulation(NeuronType.POISSON, n_src, src_rate_hz=rate)
    A = net.addPopulation(NeuronType.NOP, 1, tau_m_ms=TAU, v_th=v_th, v_reset=0.0, v_rest=0.0, t_ref_ms=t_ref)
    net.addConnectionFixedIndegree(P, A, n_src, w1, seed=3)
    B = None
    if w2 is not None:
        B = net.addPopulation(NeuronType.NOP, 1, tau_m_ms=TAU, v_th=v_th, v_reset=0.0, v_rest=0.0, t_ref_ms=t_ref)
        net.addConnectionFixedIndegree(A, B, 1, w2, seed=4)
    return net, P, A, B


def run(net, epochs):
    o = oracle_of(net, threads=1)
    o.run_events(epochs)
    return o, o.spikes(), o.event_state()


def test_subthreshold_potential_is_the_decayed_sum():
    net, P, A, _ = chain(0.01, v_th=1e9)
    o, sp, st = run(net, 20)
    src = sp[sp["neuron"] == P.address]
    times = src["time_us"].astype(np.int64)
    assert len(times) > 50
    t_last = int(times[-1])
    want = sum(0.01 * math.exp(-(t_last - t) / (TAU * 1000.0)) for t in times)
    a = A.address
    assert st["t_last"][a] == t_last
    assert st["v"][a] == pytest.approx(want, rel=1e-5)
    # the float recursion of lif.cl with exact exp
    v = np.float32(0.0)
    prev = 0
    for t in times:
        v = np.float32(float(v) * math.exp(-(int(t) - prev) / (TAU * 1000.0)))
        v = np.float32(v + np.float32(0.01))
        prev = int(t)
    assert st["v"][a] == v


def test_refractory_inputs_are_dropped():
    net, P, A, _ = chain(1.0, v_th=0.5, rate=5000.0, t_ref=2.0)
    o, sp, st = run(net, 20)
    src = sp[sp["neuron"] == P.address]["time_us"].astype(np.int64)
    out = sp[sp["neuron"] == A.address]["time_us"].astype(np.int64)
    want, free_at = [], 0
    for t in src:
        if t >= free_at:
            want.append(int(t))
            free_at = int(t) + 2000
    assert len(want) > 10 and len(want) < len(src)
    assert out.tolist() == want


def test_spikes_arrive_after_the_axonal_delay():
    net, P, A, B = chain(1.0, w2=1.0, v_th=0.5, rate=400.0, t_ref=2.0)
    o, sp, st = run(net, 30)
    a = sp[sp["neuron"] == A.address]
    b = sp[sp["neuron"] == B.address]
    assert len(a) > 3
    # B fires at every A spike that arrives within the simulated windows
    arrive = [int(t) + WINDOW for t in a["time_us"] if int(t) + WINDOW < 30 * WINDOW]
    assert b["time_us"].astype(np.int64).tolist() == arrive
    # recorded step = the window of the spike
    assert (b["step"] == b["time_us"] // WINDOW).all()


def test_decay_tables():
    lo, hi = O.ev_tables(TAU * 1000.0)
    rng = np.random.default_rng(1)
    for dt in list(rng.integers(0, len(hi) * 4096, 2000)) + [0, 1, 4095, 4096, 4097]:
        d = hi[dt >> 12] * lo[dt & 4095]
        e = math.exp(-dt / (TAU * 1000.0))
        # both round their argument -dt/tau (relative error ~ dt/tau * 2^-53):
        # far below the float precision of V
        assert abs(d - e) <= 1e-12 * e or (e < 1e-300 and abs(d - e) < 1e-300)
    # H stops where exp underflows to zero
    assert math.exp(-(len(hi) * 4096) / (TAU * 1000.0)) == 0.0
    assert hi[-1] > 0.0
