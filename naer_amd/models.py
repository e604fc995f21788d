"""Model builders.

brunel / brunel_plus / vogels follow the reference's driver models
(src/main.cpp:12-69).
The workload builders describe the benchmark networks in mV and ms:

    balanced_10k      N=10,000 (8,000 E / 2,000 I), in-degree 1,000, J_E=+0.1,
                      J_I=-0.5 mV, delay 1.5 ms, 20 kHz x 0.1 mV drive, 1 s
    plastic_100k      N=100,000, in-degree 1,000, delays 1-5 ms (whole ms),
                      15 kHz drive, additive pair STDP on E->E
    large_1m          N=1,000,000, in-degree 2,000 (2e9 synapses), delays
                      1-8 ms (whole ms), J_E=0.05 mV, g=5, drive for ~5 Hz
"""
import math

from . import _abi
from .network import REFERENCE_LIF, REFERENCE_SEED, Network, NeuronType

# Brunel LIF parameters used by every workload (tau_m 20 ms, theta 20 mV,
# reset 10 mV, t_ref 2 ms, V0 ~ U[0, 20) mV).
BRUNEL_LIF = dict(tau_m_ms=20.0, v_th=20.0, v_reset=10.0, v_rest=0.0, t_ref_ms=2.0,
                  v0_min=0.0, v0_max=20.0)

# Additive pair STDP of the plastic workload: tau+ = tau- = 20 ms,
# A+ = 0.01 wmax, A- = 0.0105 wmax, wmax = 0.3 mV, clip [0, wmax].
STDP_WMAX = 0.3
PAIR_STDP = dict(tau_plus_ms=20.0, tau_minus_ms=20.0, a_plus=0.01 * STDP_WMAX,
                 a_minus=0.0105 * STDP_WMAX, w_min=0.0, w_max=STDP_WMAX)


def brunel(num_synapses, **net_kw):
    """src/main.cpp:32-50: N = floor(sqrt(S/0.05)); P:E:I = N/2 : 0.4N : 0.1N;
    six pairwise-Bernoulli connections with p = 0.1, J_E = 2/N, J_I = -10/N (volts)."""
    N = math.floor(math.sqrt(num_synapses / 0.05))
    net = Network(**net_kw)
    P = net.addPopulation(NeuronType.POISSON, int(N / 2))
    E = net.addPopulation(NeuronType.NOP, int(N * 4 / 10))
    I = net.addPopulation(NeuronType.NOP, int(N / 10))
    net.addConnection(P, E, 2.0 / N, 0.1)
    net.addConnection(P, I, 2.0 / N, 0.1)
    net.addConnection(E, E, 2.0 / N, 0.1)
    net.addConnection(E, I, 2.0 / N, 0.1)
    net.addConnection(I, E, -10.0 / N, 0.1)
    net.addConnection(I, I, -10.0 / N, 0.1)
    return net


def vogels(num_synapses, tau_s_ms=5.0, **net_kw):
    """src/main.cpp:52-69 (commented out there: its CubaLIF type is not
    defined): brunel's populations and connections with E and I current-based
    LIF neurons, the library's cuba_lif model with the reference's LIF
    constants (lif.cl:3-6) and a synaptic time constant tau_s."""
    from . import neuron_models
    N = math.floor(math.sqrt(num_synapses / 0.05))
    net = Network(**net_kw)
    cuba = neuron_models.model("cuba_lif")
    q = REFERENCE_LIF
    params = neuron_models.cuba_lif_params(q["tau_m_ms"], tau_s_ms, q["v_th"], q["v_reset"], q["v_rest"],
                                           q["t_ref_ms"], net.dt_ms)
    P = net.addPopulation(NeuronType.POISSON, int(N / 2))
    E = net.addPopulation(cuba, int(N * 4 / 10), params=params)
    I = net.addPopulation(cuba, int(N / 10), params=params)
    net.addConnection(P, E, 2.0 / N, 0.1)
    net.addConnection(P, I, 2.0 / N, 0.1)
    net.addConnection(E, E, 2.0 / N, 0.1)
    net.addConnection(E, I, 2.0 / N, 0.1)
    net.addConnection(I, E, -10.0 / N, 0.1)
    net.addConnection(I, I, -10.0 / N, 0.1)
    return net


def brunel_plus(num_synapses, **net_kw):
    """src/main.cpp:12-30: as brunel with E an STDP population."""
    N = math.floor(math.sqrt(num_synapses / 0.05))
    net_kw.setdefault("stdp", PAIR_STDP)
    net = Network(**net_kw)
    P = net.addPopulation(NeuronType.POISSON, int(N / 2))
    E = net.addPopulation(NeuronType.STDP, int(N * 4 / 10))
    I = net.addPopulation(NeuronType.NOP, int(N / 10))
    net.addConnection(P, E, 2.0 / N, 0.1)
    net.addConnection(P, I, 2.0 / N, 0.1)
    net.addConnection(E, E, 2.0 / N, 0.1)
    net.addConnection(E, I, 2.0 / N, 0.1)
    net.addConnection(I, E, -10.0 / N, 0.1)
    net.addConnection(I, I, -10.0 / N, 0.1)
    return net


def balanced(n, indegree, j_e, g, drive_hz, drive_w, delay_ms=1.5, plastic=False, seed=REFERENCE_SEED,
             exc_frac=0.8, stdp=None, i_model=None, **net_kw):
    """Brunel-style E/I network with fixed in-degree: every neuron receives
    exc_frac*indegree synapses from E (weight j_e) and the rest from I (weight
    -g*j_e).  With plastic=True, E->E synapses start uniform in [0, wmax] and
    follow additive pair STDP.  i_model: the I population as a source-defined
    model with the LIF step (a NeuronModel compiled from the shipped "lif"
    source), the same neurons as the built-in LIF."""
    n_e = int(round(n * exc_frac))
    n_i = n - n_e
    k_e = int(round(indegree * exc_frac))
    k_i = indegree - k_e
    if plastic:
        net_kw.setdefault("stdp", stdp or PAIR_STDP)
    net = Network(seed=seed, **net_kw)
    lif = dict(BRUNEL_LIF, ext_rate_hz=drive_hz, ext_weight=drive_w)
    E = net.addPopulation(NeuronType.STDP if plastic else NeuronType.NOP, n_e, **lif)
    if i_model is not None:
        from . import neuron_models
        q = BRUNEL_LIF
        mp = neuron_models.lif_params(q["tau_m_ms"], q["v_th"], q["v_reset"], q["v_rest"], q["t_ref_ms"], net.dt_ms)
        I = net.addPopulation(i_model, n_i, params=mp, v0_min=q["v0_min"], v0_max=q["v0_max"],
                              ext_rate_hz=drive_hz, ext_weight=drive_w)
    else:
        I = net.addPopulation(NeuronType.NOP, n_i, **lif)
    s = seed
    if plastic:
        wmax = net.stdp["w_max"]
        net.addConnectionFixedIndegree(E, E, k_e, 0.0, 0.0, wmax, seed=s + 1, delay_ms=delay_ms)
    else:
        net.addConnectionFixedIndegree(E, E, k_e, j_e, seed=s + 1, delay_ms=delay_ms)
    net.addConnectionFixedIndegree(E, I, k_e, j_e, seed=s + 2, delay_ms=delay_ms)
    net.addConnectionFixedIndegree(I, E, k_i, -g * j_e, seed=s + 3, delay_ms=delay_ms)
    net.addConnectionFixedIndegree(I, I, k_i, -g * j_e, seed=s + 4, delay_ms=delay_ms)
    return net


def balanced_10k(n=10_000, **kw):
    """Balanced 10k network: K=1000, J_E=0.1, J_I=-0.5 mV, delay 1.5 ms,
    20 kHz x 0.1 mV Poisson drive per neuron."""
    return balanced(n, 1000, 0.1, 5.0, 20_000.0, 0.1, delay_ms=1.5, **kw)


def plastic_100k(n=100_000, indegree=1000, **kw):
    """Plastic 100k network: K=1000, delays uniform over whole ms in 1-5 ms,
    15 kHz x 0.1 mV drive, STDP on E->E."""
    return balanced(n, indegree, 0.1, 5.0, 15_000.0, 0.1, delay_ms=(1.0, 5.0, 1.0), plastic=True, **kw)


# Drive of the 1M network, chosen for a mean rate near 5 Hz (J_ext = J_E).
LARGE_DRIVE_HZ = 19_700.0
LARGE_DRIVE_W = 0.05


def large_1m(n=1_000_000, indegree=2000, drive_hz=LARGE_DRIVE_HZ, **kw):
    """1M network: K=2000 (2e9 synapses), delays uniform over whole ms in 1-8 ms,
    J_E=0.05 mV, g=5."""
    return balanced(n, indegree, 0.05, 5.0, drive_hz, LARGE_DRIVE_W, delay_ms=(1.0, 8.0, 1.0), **kw)
