"""The oracle's event-driven STDP (stdp.cl:8-140 with basic_stdp.cl's rule)
pinned by closed forms on networks of two STDP neurons A -> B, where A is
driven by a POISSON source:
  * without post spikes the trace stays 0 and every plastic input scales w by
    1 - alpha lambda = 1 - 2.02 * 0.01: w = w0 (1 - 0.0202)^m after m inputs;
  * when every input fires B, the trace obeys tau_k = e^(-(t_k - t_{k-1} - t_ref)/T)
    (tau_{k-1} + 1) (decay from the end of B's refractory period, where the
    history puts B's last spike), w_k = w_{k-1} (1 - 0.0202 e^(-tau_k/100)),
    the first input clamps w to +-0.0003, and B's history holds its
    inter-spike intervals after the refractory period in 4 us units; the LTP
    sum contributes nothing (the unsigned difference of a later post time
    wraps and the increment underflows), so w never grows;
  * an input inside B's refractory period clears the trace (the unsigned
    t - last wraps to a huge positive float);
and the single-precision exponential of the rule (orc_expf) within 2 ulp of
exp, with its overflow and underflow edges."""
import math

import numpy as np
import pytest

import oracle as O
from helpers import oracle_of
from naer_amd import _abi
from naer_amd.network import Network, NeuronType

WINDOW = 1500
T_STDP = 20000.0
ALPHA_LAMBDA = float(np.float32(2.02)) * 0.01


def pair(w_ab, b_th, rate=600.0, a_ref=2.0, b_ref=2.0):
    net = Network(dt_ms=0.1, mode=_abi.MODE_EVENT, record_capacity=1 << 16)
    P = net.addPopulation(NeuronType.POISSON, 1, src_rate_hz=rate)
    A = net.addPopulation(NeuronType.STDP, 1, tau_m_ms=20.0, v_th=0.5, v_reset=0.0, v_rest=0.0, t_ref_ms=a_ref)
    B = net.addPopulation(NeuronType.STDP, 1, tau_m_ms=20.0, v_th=b_th, v_reset=0.0, v_rest=0.0, t_ref_ms=b_ref)
    net.addConnectionFixedIndegree(P, A, 1, 1.0, seed=3)   # every P spike fires A outside its refractory period
    net.addConnectionFixedIndegree(A, B, 1, w_ab, seed=4)  # STDP -> STDP: plastic
    return net, P, A, B


def run(net, epochs):
    o = oracle_of(net, threads=1)
    o.run_events(epochs)
    rp, col, w = o.csr()
    return o, o.spikes(), o.event_state(), rp, col, w


def ab_weight(rp, col, w, A, B):
    row = slice(int(rp[A.address]), int(rp[A.address + 1]))
    hit = np.nonzero((col[row] & 0xFFFFFF) == B.address)[0]
    assert len(hit) == 1
    assert col[row][hit[0]] & 0x80000000  # marked plastic
    return float(w[row][hit[0]])


def arrivals(sp, A, epochs):
    a = sp[sp["neuron"] == A.address]["time_us"].astype(np.int64)
    return [int(t) + WINDOW for t in a if int(t) + WINDOW < epochs * WINDOW]


def test_expf_within_two_ulp_and_its_edges():
    rng = np.random.default_rng(7)
    xs = np.concatenate([rng.uniform(-87.0, 88.7, 20000), np.linspace(-1.0, 1.0, 2001)]).astype(np.float32)
    for x in xs:
        got = np.float32(O.expf(float(x)))
        want = math.exp(float(x))
        ulp = float(np.spacing(np.float32(want)))
        assert abs(float(got) - want) <= 2.0 * ulp, x
    assert O.expf(0.0) == 1.0
    assert O.expf(88.7229) == math.inf and O.expf(1e6) == math.inf
    assert O.expf(-104.0) == 0.0 and O.expf(-math.inf) == 0.0
    assert math.isnan(O.expf(math.nan))
    # the LTP factor of a later post time: (float)(2^32 - 500) / 20000 overflows
    assert O.expf(float(np.float32(2**32 - 500)) / 20000.0) == math.inf


def test_weight_decays_geometrically_without_post_spikes():
    epochs = 60
    w0 = 2e-4
    net, P, A, B = pair(w0, b_th=1e9)
    o, sp, st, rp, col, w = run(net, epochs)
    m = len(arrivals(sp, A, epochs))
    assert m > 20
    assert not (sp["neuron"] == B.address).any()
    assert st["trace"][B.address] == 0.0
    got = ab_weight(rp, col, w, A, B)
    assert got == pytest.approx(w0 * (1.0 - ALPHA_LAMBDA) ** m, rel=2e-7 * m)
    o.close()


def test_trace_recurrence_when_every_input_fires_the_target():
    epochs = 40
    t_ref = 2000
    net, P, A, B = pair(0.01, b_th=1e-4, rate=300.0)
    o, sp, st, rp, col, w = run(net, epochs)
    t = arrivals(sp, A, epochs)
    assert len(t) > 8
    # A fires at least t_ref apart, so B takes every input outside its refractory period
    assert all(b - a >= t_ref for a, b in zip(t, t[1:]))
    b_spikes = sp[sp["neuron"] == B.address]["time_us"].astype(np.int64).tolist()
    assert b_spikes == t
    tau, wk, hist, last = 0.0, 0.01, [], 0
    for k, tk in enumerate(t):
        tau = math.exp(-(tk - last) / T_STDP) * tau
        wk = wk * (1.0 - ALPHA_LAMBDA * math.exp(-tau / 100.0))
        wk = max(min(wk, 0.0003), -0.0003)
        hist.insert(0, (tk - last) // 4)
        tau += 1.0
        last = tk + t_ref
        if k == 0:
            assert wk == 0.0003  # clamped
    assert ab_weight(rp, col, w, A, B) == pytest.approx(wk, rel=1e-5)
    assert st["trace"][B.address] == pytest.approx(tau, rel=1e-5)
    # history: 0, then the intervals newest first
    want_hist = [0] + hist[:9]
    assert st["hist"][B.address].tolist() == want_hist
    o.close()


def test_input_inside_the_refractory_period_clears_the_trace():
    epochs = 60
    # A may fire 1.5 ms apart, B is refractory for 2 ms after each spike
    net, P, A, B = pair(0.01, b_th=1e-4, rate=2000.0, a_ref=1.5, b_ref=2.0)
    o, sp, st, rp, col, w = run(net, epochs)
    t = arrivals(sp, A, epochs)
    gaps = [b - a for a, b in zip(t, t[1:])]
    first_in_ref = next(i + 1 for i, g in enumerate(gaps) if g < 2000)
    # it is the last input of its window, so the state after that window is the one to check
    stop = t[first_in_ref] // WINDOW + 1
    assert first_in_ref + 1 == len(t) or t[first_in_ref + 1] >= stop * WINDOW
    t = t[: first_in_ref + 1]
    # replay up to and including the first input inside B's refractory period
    tau, wk, last = 0.0, 0.01, 0
    for k, tk in enumerate(t):
        inside = k == first_in_ref
        tau = 0.0 if inside else math.exp(-(tk - last) / T_STDP) * tau
        wk = max(min(wk * (1.0 - ALPHA_LAMBDA * math.exp(-tau / 100.0)), 0.0003), -0.0003)
        if not inside:
            tau += 1.0
            last = tk + 2000
    assert math.exp(-tau / 100.0) == 1.0  # the cleared trace: the full 2.02 % step
    # run the oracle only as far as that input's window
    net2, P2, A2, B2 = pair(0.01, b_th=1e-4, rate=2000.0, a_ref=1.5, b_ref=2.0)
    o2, sp2, st2, rp2, col2, w2 = run(net2, stop)
    assert ab_weight(rp2, col2, w2, A2, B2) == pytest.approx(wk, rel=1e-5)
    assert st2["trace"][B2.address] == 0.0
    o.close()
    o2.close()


def test_plastic_weights_never_grow_on_brunel_plus():
    from naer_amd import models
    net = models.brunel_plus(2_000_000, mode=_abi.MODE_EVENT, record_capacity=1 << 22)
    o = oracle_of(net, threads=1)
    rp0, col0, w0 = o.csr()
    w0 = w0.copy()
    o.run_events(12)
    rp, col, w = o.csr()
    pl = (col & 0x80000000) != 0
    assert pl.sum() > 10000
    assert (w[pl] <= w0[pl]).all()
    changed = w[pl] < w0[pl]
    assert changed.sum() > 1000
    # at most one 2.02 % step per window and synapse
    assert (w[pl] >= np.minimum(w0[pl], 0.0003) * (1.0 - ALPHA_LAMBDA) ** 12 * (1 - 1e-6)).all()
    np.testing.assert_array_equal(w[~pl], w0[~pl])
    o.close()
