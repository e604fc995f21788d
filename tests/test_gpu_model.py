"""Neuron models given as source and compiled with hipRTC (naer_model_create),
the reference's NeuronType source-string path (NeuronType.h:14,
processing_neuron_group.cl:8-82).

Parity is pinned two ways:
  * a model restating the built-in LIF step must give the built-in LIF
    network's raster and potentials bit for bit (deterministic mode), with
    model neurons as sources and targets, over bucketed and ring batches,
    graphs, packs and both tile widths; the built-in LIF is itself pinned
    against the oracle by the other GPU tests;
  * a model with spike-frequency adaptation is replayed in float32 numpy from
    the synaptic input the probe records, and must give the same spikes and
    final potentials.
"""
import math

import numpy as np
import pytest

from naer_amd import _abi, neuron_models
from naer_amd.arena import Arena
from naer_amd.network import Network, NeuronModel, NeuronType
from naer_amd.pack import Pack

pytestmark = pytest.mark.gpu

DT = 0.1

# The built-in LIF step written as a model: s[0] = V, s[1] = refractory steps
# left; p = {P, v_th, v_reset, v_rest, n_ref}.
LIF_SRC = r"""
__device__ bool naer_model_step(float* s, float I, float ext, const float* p, uint32_t t)
{
    if (s[1] > 0.0f) {
        s[1] = s[1] - 1.0f;
        return false;
    }
    const float x = s[0] - p[3];
    const float xp = x * p[0];
    float v = xp + p[3];
    v = v + I;
    v = v + ext;
    const bool f = v > p[1];
    s[0] = f ? p[2] : v;
    if (f)
        s[1] = p[4];
    return f;
}
"""

# LIF with an adaptation current a (s[2]) that decays by p[5] per step and
# jumps by p[6] at each spike.
ADAPT_SRC = r"""
__device__ bool naer_model_step(float* s, float I, float ext, const float* p, uint32_t t)
{
    const float a = s[2];
    s[2] = a * p[5];
    if (s[1] > 0.0f) {
        s[1] = s[1] - 1.0f;
        return false;
    }
    const float x = s[0] - p[3];
    const float xp = x * p[0];
    float v = xp + p[3];
    v = v + I;
    v = v + ext;
    v = v - a;
    const bool f = v > p[1];
    s[0] = f ? p[2] : v;
    if (f) {
        s[1] = p[4];
        s[2] = s[2] + p[6];
    }
    return f;
}
"""

LIF = dict(tau_m_ms=20.0, v_th=20.0, v_reset=10.0, v_rest=0.0, t_ref_ms=2.0, v0_min=0.0, v0_max=20.0,
           ext_rate_hz=20000.0, ext_weight=0.1)


def lif_params(q):
    """The runtime's propagator and refractory steps, from the same float inputs."""
    dt = float(np.float32(DT))
    P = np.float32(math.exp(-dt / float(np.float32(q["tau_m_ms"]))))
    n_ref = float(round(float(np.float32(q["t_ref_ms"])) / dt))
    return [float(P), q["v_th"], q["v_reset"], q["v_rest"], n_ref]


_models = {}


def model(name):
    if name not in _models:
        src, ns = {"lif": (LIF_SRC, 2), "adapt": (ADAPT_SRC, 3)}[name]
        _models[name] = NeuronModel(src, ns, name)
    return _models[name]


def ei_net(n, as_model, mode=_abi.MODE_DETERMINISTIC, delay_ms=1.5, seed=7, record=1 << 21):
    """E (LIF) and I (LIF, or the LIF model) with all four projections."""
    net = Network(dt_ms=DT, mode=mode, seed=seed, record_capacity=record)
    n_e, n_i = int(n * 0.8), n - int(n * 0.8)
    E = net.addPopulation(NeuronType.NOP, n_e, **LIF)
    if as_model:
        I = net.addPopulation(model("lif"), n_i, params=lif_params(LIF),
                              **{k: LIF[k] for k in ("v0_min", "v0_max", "ext_rate_hz", "ext_weight")})
    else:
        I = net.addPopulation(NeuronType.NOP, n_i, **LIF)
    K = min(max(10, n // 10), 500)
    ke, ki = int(K * 0.8), K - int(K * 0.8)
    for dst in (E, I):
        net.addConnectionFixedIndegree(E, dst, ke, 0.2, seed=11 + (dst is I), delay_ms=delay_ms)
        net.addConnectionFixedIndegree(I, dst, ki, -1.0, seed=21 + (dst is I), delay_ms=delay_ms)
    return net


def run(net, chunks, graph=0):
    net.compile()
    if graph:
        net.set_graph(graph)
    for c in chunks:
        net.run(c)
    net.finish()
    return net.spikes(), net.state()["v"]


@pytest.mark.parametrize("graph", [0, 100])
def test_lif_model_equals_builtin_lif(graph):
    chunks = (500, 37, 713)  # full bucketed batches and partial ring batches
    ref_sp, ref_v = run(ei_net(1500, False), chunks)
    sp, v = run(ei_net(1500, True), chunks, graph)
    assert len(ref_sp) > 1000
    n_e = 1200
    assert np.count_nonzero(ref_sp["neuron"] >= n_e) > 100  # the I population fires
    assert np.array_equal(sp, ref_sp)
    assert np.array_equal(v, ref_v)


def test_lif_model_equals_builtin_wide_tiles():
    """3e5 neurons with 1 ms delays: 1024-neuron update tiles."""
    chunks = (200, 7)
    ref_sp, ref_v = run(ei_net(300_000, False, delay_ms=1.0), chunks)
    net = ei_net(300_000, True, delay_ms=1.0)
    sp, v = run(net, chunks)
    assert len(ref_sp) > 1000
    assert np.array_equal(sp, ref_sp)
    assert np.array_equal(v, ref_v)


def test_model_net_in_pack_equals_alone():
    steps = 600
    alone = [run(ei_net(1200, True, seed=3), (steps,))[0], run(ei_net(900, False, seed=4), (steps,))[0]]
    p = Pack([ei_net(1200, True, seed=3), ei_net(900, False, seed=4)])
    p.build()
    p.run(steps)
    packed = p.spikes()
    p.close()
    for i in range(2):
        assert np.array_equal(packed[i], alone[i]), i


def test_fast_mode_runs_model():
    sp, v = run(ei_net(1500, True, mode=_abi.MODE_FAST), (1000,))
    ref, _ = run(ei_net(1500, False, mode=_abi.MODE_FAST), (1000,))
    assert abs(len(sp) - len(ref)) < 0.05 * len(ref)
    assert np.all(np.isfinite(v))


def test_adaptation_model_matches_numpy_replay():
    """Model neurons driven by POISSON sources and LIF neurons; the probe
    gives the input each model neuron received at every step."""
    steps = 1500
    net = Network(dt_ms=DT, mode=_abi.MODE_DETERMINISTIC, seed=5, record_capacity=1 << 20)
    src = net.addPopulation(NeuronType.POISSON, 400, src_rate_hz=400.0)
    L = net.addPopulation(NeuronType.NOP, 300, **LIF)
    params = lif_params(LIF) + [0.99, 0.8]  # a decays by 0.99 per step, +0.8 mV per spike
    A = net.addPopulation(model("adapt"), 250, params=params, v0_min=0.0, v0_max=20.0)
    net.addConnectionFixedIndegree(src, L, 40, 1.0, seed=1)
    net.addConnectionFixedIndegree(src, A, 60, 1.2, seed=2)
    net.addConnectionFixedIndegree(L, A, 30, 0.5, seed=3)
    net.addConnectionFixedIndegree(A, L, 20, -0.5, seed=4)
    net.compile()
    base = A.address
    v0 = net.state()["v"][base: base + 250].copy()
    net.set_probe(0, steps)
    net.run(steps)
    net.finish()
    sp = net.spikes()
    I = net.probe()[:, base: base + 250]
    v_end = net.state()["v"][base: base + 250]

    f = np.float32
    P, v_th, v_reset, v_rest, n_ref, dec, jump = (f(x) for x in params)
    V = v0.astype(np.float32)
    R = np.zeros(250, np.float32)
    a = np.zeros(250, np.float32)
    want = []
    for t in range(steps):
        a_old = a
        a = (a_old * dec).astype(np.float32)
        refr = R > 0
        v = ((V - v_rest) * P + v_rest).astype(np.float32)
        v = (v + I[t]).astype(np.float32)
        v = (v + f(0.0)).astype(np.float32)
        v = (v - a_old).astype(np.float32)
        fired = (~refr) & (v > v_th)
        V = np.where(refr, V, np.where(fired, v_reset, v)).astype(np.float32)
        R = np.where(refr, R - f(1.0), np.where(fired, n_ref, R)).astype(np.float32)
        a = np.where(fired, (a + jump).astype(np.float32), a)
        want += [(t, base + int(j)) for j in np.nonzero(fired)[0]]
    got = [(int(s["step"]), int(s["neuron"])) for s in sp if base <= s["neuron"] < base + 250]
    assert len(want) > 200
    assert got == sorted(want)
    assert np.array_equal(v_end, V)


def test_shipped_lif_model_equals_builtin_lif():
    """neuron_models' lif source and parameter helper give the built-in LIF."""
    q = LIF
    params = neuron_models.lif_params(q["tau_m_ms"], q["v_th"], q["v_reset"], q["v_rest"], q["t_ref_ms"], DT)
    assert params == lif_params(q)
    _models["lif"] = neuron_models.model("lif")  # ei_net's model: the shipped source
    try:
        ref_sp, ref_v = run(ei_net(1000, False, seed=9), (900,))
        sp, v = run(ei_net(1000, True, seed=9), (900,))
    finally:
        del _models["lif"]
    assert len(ref_sp) > 500
    assert np.array_equal(sp, ref_sp)
    assert np.array_equal(v, ref_v)


def test_cuba_lif_matches_numpy_replay():
    """The shipped current-based LIF, replayed in float32 numpy from the
    probed input: the synaptic current decays by P11 and takes the step's
    arrivals; V integrates the current of the step's start through P21.  (The
    drive of a model population is pinned by the LIF tests above.)"""
    steps = 1500
    params = neuron_models.cuba_lif_params(tau_m_ms=20.0, tau_s_ms=5.0, v_th=20.0, v_reset=10.0, v_rest=0.0,
                                           t_ref_ms=2.0, dt_ms=DT)
    p11, p22 = math.exp(-DT / 5.0), math.exp(-DT / 20.0)
    assert params[6] == pytest.approx(20.0 * (p11 - p22) / (5.0 - 20.0), rel=1e-6)
    # for dt << tau_s the step charges V by ~ (dt / tau_s) * current
    assert params[6] == pytest.approx(DT / 5.0, rel=0.02)
    net = Network(dt_ms=DT, mode=_abi.MODE_DETERMINISTIC, seed=6, record_capacity=1 << 20)
    src = net.addPopulation(NeuronType.POISSON, 400, src_rate_hz=400.0)
    L = net.addPopulation(NeuronType.NOP, 300, **LIF)
    C = net.addPopulation(neuron_models.model("cuba_lif"), 250, params=params, v0_min=0.0, v0_max=20.0)
    net.addConnectionFixedIndegree(src, L, 40, 1.0, seed=1)
    net.addConnectionFixedIndegree(src, C, 60, 1.5, seed=2)
    net.addConnectionFixedIndegree(L, C, 30, 0.5, seed=3)
    net.addConnectionFixedIndegree(C, L, 20, -0.5, seed=4)
    net.compile()
    base = C.address
    v0 = net.state()["v"][base: base + 250].copy()
    net.set_probe(0, steps)
    net.run(steps)
    net.finish()
    sp = net.spikes()
    I = net.probe()[:, base: base + 250]
    v_end = net.state()["v"][base: base + 250]

    f = np.float32
    P22, v_th, v_reset, v_rest, n_ref, P11, P21 = (f(x) for x in params)
    V = v0.astype(np.float32)
    R = np.zeros(250, np.float32)
    c = np.zeros(250, np.float32)
    want = []
    for t in range(steps):
        c_old = c
        c = ((c_old * P11).astype(np.float32) + I[t]).astype(np.float32)
        refr = R > 0
        v = ((V - v_rest) * P22 + v_rest).astype(np.float32)
        v = (v + (c_old * P21).astype(np.float32)).astype(np.float32)
        v = (v + f(0.0)).astype(np.float32)
        fired = (~refr) & (v > v_th)
        V = np.where(refr, V, np.where(fired, v_reset, v)).astype(np.float32)
        R = np.where(refr, R - f(1.0), np.where(fired, n_ref, R)).astype(np.float32)
        want += [(t, base + int(j)) for j in np.nonzero(fired)[0]]
    got = [(int(s["step"]), int(s["neuron"])) for s in sp if base <= s["neuron"] < base + 250]
    assert len(want) > 200
    assert got == sorted(want)
    assert np.array_equal(v_end, V)


def test_model_populations_rejected_where_unsupported(tmp_path):
    net = ei_net(500, True)
    net.compile()
    with pytest.raises(_abi.NaerError, match="checkpointed"):
        net.save(tmp_path / "x.ckpt")
    arena = Arena(10_000, 1_000_000, delay_steps=(15, 15))
    with pytest.raises(_abi.NaerError, match="arenas"):
        arena.admit(ei_net(500, True))
    arena.close()
    bad = Network(dt_ms=DT, mode=_abi.MODE_DETERMINISTIC, stdp=dict(tau_plus_ms=20.0, tau_minus_ms=20.0,
                  a_plus=0.01, a_minus=0.01, w_min=0.0, w_max=1.0))
    S = bad.addPopulation(NeuronType.STDP, 100, **LIF)
    M = bad.addPopulation(model("lif"), 100, params=lif_params(LIF))
    bad.addConnectionFixedIndegree(S, M, 10, 0.1)
    bad.addConnectionFixedIndegree(M, S, 10, 0.1)
    # S -> M and M -> S are static: plasticity needs STDP at both ends
    bad.compile()
    assert bad.stats().n_plastic == 0
    bad.run(100)
    bad.finish()
