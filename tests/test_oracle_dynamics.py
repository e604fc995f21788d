"""Pins the oracle's clock-driven step against what the mathematics fixes.

The reference has no tests and no golden vectors for its neuron dynamics, and
its event-driven LIF (lif.cl, nop.cl) cannot be built here.  These checks tie
the oracle's step (the object the HIP kernels are compared with bit for bit)
to closed forms and distributions that a dropped term, a wrong sign or an
off-by-one in the step would violate:

  free decay          V_t = v_rest + (V_0 - v_rest) P^t, P = exp(-dt / tau_m)
                      (dt is the fp32 value of dt_ms; P in fp64 for the fp64
                      variant, rounded to fp32 for the fp32 ones)
  refractoriness      a spike at step 0 holds V at v_reset for t_ref / dt steps
  delayed delivery    a spike of step 0 over a synapse of delay d is input at step d
  leak-then-input     V_d = v_rest + (V_{d-1} - v_rest) P + w
  Poisson drive       per-step drive counts are Poisson(rate dt) (chi-square)
  pair STDP           pre->post interval dt: dw = +A+ exp(-dt / tau+);
                      post->pre interval dt: dw = -A- exp(-dt / tau-)
All run on the CPU (no GPU).
"""
import math

import numpy as np
import pytest
from scipy import stats

import oracle as O
from naer_amd import _abi
from naer_amd.network import Network, NeuronType

DT_MS = 0.1
DT32 = float(np.float32(DT_MS))  # dt as the step holds it


def prop(tau_ms, precision=O.DOUBLE):
    """The step's propagator exp(-dt / tau): fp64, or rounded to fp32."""
    p = math.exp(-DT32 / tau_ms)
    return p if precision == O.DOUBLE else float(np.float32(p))


def lif(**kw):
    p = dict(tau_m_ms=20.0, v_th=20.0, v_reset=10.0, v_rest=0.0, t_ref_ms=2.0, v0_min=0.0, v0_max=0.0,
             ext_rate_hz=0.0, ext_weight=0.0)
    p.update(kw)
    return p


def oracle(net, precision=O.DOUBLE):
    return O.OracleNet(net.description(), precision, threads=1)


@pytest.mark.parametrize("precision", [O.DOUBLE, O.FIXED, O.FLOAT])
def test_free_decay_closed_form(precision):
    net = Network(dt_ms=DT_MS, mode=_abi.MODE_DETERMINISTIC)
    net.addPopulation(NeuronType.NOP, 256, **lif(v_rest=-3.0, v0_min=-2.0, v0_max=19.0))
    o = oracle(net, precision)
    v0 = o.v64() if precision == O.DOUBLE else o.state()["v"].astype(np.float64)
    assert v0.min() >= -2.0 and v0.max() < 19.0 and np.ptp(v0) > 10.0  # drawn over the range
    steps = 250
    o.run(steps)
    v = o.v64() if precision == O.DOUBLE else o.state()["v"].astype(np.float64)
    expect = -3.0 + (v0 + 3.0) * prop(20.0, precision) ** steps
    # and the propagator is the exponential's, not a linearisation
    assert prop(20.0) ** steps == pytest.approx(math.exp(-steps * DT_MS / 20.0), rel=1e-6)
    tol = 1e-12 if precision == O.DOUBLE else 2e-5
    np.testing.assert_allclose(v, expect, rtol=0, atol=tol)
    assert o.spike_count == 0
    o.close()


def test_refractory_hold_and_release():
    net = Network(dt_ms=DT_MS, mode=_abi.MODE_DETERMINISTIC, record_capacity=4096)
    net.addPopulation(NeuronType.NOP, 32, **lif(v0_min=25.0, v0_max=30.0))  # above threshold
    o = oracle(net)
    n_ref = int(round(2.0 / DT_MS))
    trace = []
    for _ in range(n_ref + 6):
        o.run(1)
        trace.append(o.v64().copy())
    sp = o.spikes()
    # every neuron fires once, at step 0, and never again
    assert len(sp) == 32 and (sp["step"] == 0).all()
    assert sorted(sp["neuron"].tolist()) == list(range(32))
    trace = np.array(trace)  # trace[t] = V after step t
    # steps 0 .. n_ref: V held at v_reset; from step n_ref + 1 it decays to v_rest = 0
    np.testing.assert_array_equal(trace[: n_ref + 1], 10.0)
    P = prop(20.0)
    for k in range(1, 6):
        np.testing.assert_allclose(trace[n_ref + k], 10.0 * P ** k, rtol=1e-12)
    o.close()


@pytest.mark.parametrize("delay_steps", [1, 7, 40])
def test_delay_delivery_and_leak_then_input(delay_steps):
    net = Network(dt_ms=DT_MS, mode=_abi.MODE_DETERMINISTIC)
    a = net.addPopulation(NeuronType.NOP, 1, **lif(v0_min=25.0, v0_max=25.0, t_ref_ms=1000.0))
    b = net.addPopulation(NeuronType.NOP, 1, **lif(v_rest=1.0, v0_min=4.0, v0_max=4.0))
    w = 0.75
    net.addConnectionFixedIndegree(a, b, 1, w, seed=11, delay_ms=delay_steps * DT_MS)
    o = oracle(net)
    steps = delay_steps + 5
    o.set_probe(0, steps)
    vb = []
    for _ in range(steps):
        o.run(1)
        vb.append(o.v64()[1])
    used, exact, _ = o.probe()
    inp = exact[:, 1]
    expect_in = np.zeros(steps)
    expect_in[delay_steps] = w
    np.testing.assert_allclose(inp, expect_in, rtol=0, atol=1e-9)
    assert exact[:, 0].max() == 0.0  # the source gets nothing
    P = prop(20.0)
    # V_b after step t: leak toward v_rest = 1 from V0 = 4, the input added after the leak
    v = 4.0
    for t in range(steps):
        v = 1.0 + (v - 1.0) * P + (w if t == delay_steps else 0.0)
        assert vb[t] == pytest.approx(v, rel=1e-12), t
    o.close()


def test_drive_counts_are_poisson():
    rate_hz, w_ext = 20_000.0, 0.25
    lam = rate_hz * DT_MS * 1e-3
    net = Network(dt_ms=DT_MS, mode=_abi.MODE_DETERMINISTIC, seed=0x1234_5678_9ABC)
    # threshold out of reach and no leak toward a moving target: V_t - V_{t-1} P = k w_ext
    net.addPopulation(NeuronType.NOP, 4000, **lif(v_th=1e12, ext_rate_hz=rate_hz, ext_weight=w_ext))
    o = oracle(net)
    P = prop(20.0)
    counts = []
    v_prev = o.v64().copy()
    for _ in range(12):  # 12 steps: three Philox blocks of four words per neuron
        o.run(1)
        v = o.v64().copy()
        k = (v - v_prev * P) / w_ext
        assert np.abs(k - np.round(k)).max() < 1e-6
        counts.append(np.round(k).astype(np.int64))
        v_prev = v
    counts = np.concatenate(counts)
    n = len(counts)
    assert abs(counts.mean() - lam) < 4 * math.sqrt(lam / n)
    assert abs(counts.var() - lam) < 0.05 * lam
    kmax = 9
    obs = np.bincount(np.minimum(counts, kmax), minlength=kmax + 1)
    pmf = stats.poisson.pmf(np.arange(kmax), lam)
    exp_ = n * np.append(pmf, 1.0 - pmf.sum())
    chi2 = ((obs - exp_) ** 2 / exp_).sum()
    assert stats.chi2.sf(chi2, kmax) > 1e-3, (obs, exp_)
    # consecutive counts of a neuron are uncorrelated
    c = counts.reshape(12, -1)
    r = np.corrcoef(c[:-1].ravel(), c[1:].ravel())[0, 1]
    assert abs(r) < 4 / math.sqrt(c[:-1].size)
    o.close()


STDP = dict(tau_plus_ms=16.8, tau_minus_ms=33.7, a_plus=0.01, a_minus=0.012, w_min=0.0, w_max=5.0)


def plastic_pair(pre_first, interval_steps):
    """Two STDP neurons pre -> post (plastic, weight 1, delay 1 step).  The
    neuron that fires first is forced (V0 above threshold); the other is made
    to fire interval_steps later by a strong static synapse from a forced NOP
    driver neuron.  Refractory periods are long, so each fires exactly once."""
    net = Network(dt_ms=DT_MS, mode=_abi.MODE_DETERMINISTIC, stdp=STDP, record_capacity=64)
    long_ref = dict(t_ref_ms=1000.0)
    pre_v0 = 25.0 if pre_first else 0.0
    post_v0 = 0.0 if pre_first else 25.0
    pre = net.addPopulation(NeuronType.STDP, 1, **lif(v0_min=pre_v0, v0_max=pre_v0, **long_ref))
    post = net.addPopulation(NeuronType.STDP, 1, **lif(v0_min=post_v0, v0_max=post_v0, **long_ref))
    drv = net.addPopulation(NeuronType.NOP, 1, **lif(v0_min=25.0, v0_max=25.0, **long_ref))
    late = post if pre_first else pre
    net.addConnectionFixedIndegree(drv, late, 1, 50.0, seed=3, delay_ms=interval_steps * DT_MS)
    net.addConnectionFixedIndegree(pre, post, 1, 1.0, seed=5, delay_ms=DT_MS)
    return net


@pytest.mark.parametrize("interval", [3, 20, 60])
def test_pair_stdp_potentiation_closed_form(interval):
    net = plastic_pair(True, interval)
    o = oracle(net, O.DOUBLE)
    o.run(interval + 30)
    sp = o.spikes()
    steps = {int(n): int(s) for s, n in zip(sp["step"], sp["neuron"])}
    assert steps[0] == 0 and steps[1] == interval  # pre at 0, post at interval
    rp, col, w = o.csr()
    plastic = w[rp[0]:rp[1]]  # the pre neuron's row: its one plastic synapse
    assert len(plastic) == 1
    expect = 1.0 + STDP["a_plus"] * math.exp(-interval * DT_MS / STDP["tau_plus_ms"])
    assert float(plastic[0]) == pytest.approx(expect, rel=1e-6)
    o.close()


@pytest.mark.parametrize("interval", [4, 25, 90])
def test_pair_stdp_depression_closed_form(interval):
    net = plastic_pair(False, interval)
    o = oracle(net, O.DOUBLE)
    o.run(interval + 30)
    sp = o.spikes()
    steps = {int(n): int(s) for s, n in zip(sp["step"], sp["neuron"])}
    assert steps[1] == 0 and steps[0] == interval  # post at 0, pre at interval
    rp, col, w = o.csr()
    plastic = w[rp[0]:rp[1]]
    expect = 1.0 - STDP["a_minus"] * math.exp(-interval * DT_MS / STDP["tau_minus_ms"])
    assert float(plastic[0]) == pytest.approx(expect, rel=1e-6)
    o.close()


def test_pair_stdp_weight_bounds():
    net = plastic_pair(True, 2)
    net.stdp = dict(STDP, a_plus=10.0)  # one pairing would overshoot w_max
    o = oracle(net, O.DOUBLE)
    o.run(40)
    rp, col, w = o.csr()
    assert float(w[rp[0]]) == STDP["w_max"]
    o.close()


def two_spike_pair(pre_twice, k1, k2):
    """The neuron that fires twice (pre if pre_twice, else post) fires at step 0
    (V0 above threshold) and at k1 (a strong synapse from a forced NOP driver);
    the other neuron is made to fire at k2 by a second driver.  Refractory
    5 steps for the twice-firing neuron, long for the others."""
    net = Network(dt_ms=DT_MS, mode=_abi.MODE_DETERMINISTIC, stdp=STDP, record_capacity=64)
    twice = dict(t_ref_ms=0.5, v0_min=25.0, v0_max=25.0)
    once = dict(t_ref_ms=1000.0, v0_min=0.0, v0_max=0.0)
    pre = net.addPopulation(NeuronType.STDP, 1, **lif(**(twice if pre_twice else once)))
    post = net.addPopulation(NeuronType.STDP, 1, **lif(**(once if pre_twice else twice)))
    d1 = net.addPopulation(NeuronType.NOP, 1, **lif(v0_min=25.0, v0_max=25.0, t_ref_ms=1000.0))
    d2 = net.addPopulation(NeuronType.NOP, 1, **lif(v0_min=25.0, v0_max=25.0, t_ref_ms=1000.0))
    net.addConnectionFixedIndegree(d1, pre if pre_twice else post, 1, 50.0, seed=3, delay_ms=k1 * DT_MS)
    net.addConnectionFixedIndegree(d2, post if pre_twice else pre, 1, 50.0, seed=4, delay_ms=k2 * DT_MS)
    net.addConnectionFixedIndegree(pre, post, 1, 1.0, seed=5, delay_ms=DT_MS)
    return net


@pytest.mark.parametrize("precision", [O.DOUBLE, O.FLOAT, O.FIXED])
@pytest.mark.parametrize("pre_twice", [True, False])
def test_pair_stdp_traces_sum_over_spikes(precision, pre_twice):
    # the trace at the second neuron's spike sums both earlier spikes:
    # x = dec^k2 + dec^(k2 - k1); a dropped accumulation or a trace reset at
    # each spike would give one term only
    k1, k2 = 30, 70
    net = two_spike_pair(pre_twice, k1, k2)
    o = oracle(net, precision)
    o.run(k2 + 20)
    sp = o.spikes()
    fired = {}
    for s, nrn in zip(sp["step"], sp["neuron"]):
        fired.setdefault(int(nrn), []).append(int(s))
    first, second = (0, 1) if pre_twice else (1, 0)
    assert fired[first] == [0, k1] and fired[second] == [k2]
    rp, col, w = o.csr()
    tau = STDP["tau_plus_ms"] if pre_twice else STDP["tau_minus_ms"]
    x = math.exp(-k2 * DT_MS / tau) + math.exp(-(k2 - k1) * DT_MS / tau)
    expect = 1.0 + STDP["a_plus"] * x if pre_twice else 1.0 - STDP["a_minus"] * x
    assert float(w[rp[0]]) == pytest.approx(expect, rel=2e-6)
    o.close()
