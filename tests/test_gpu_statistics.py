"""The HIP step against the fp64 variant of the oracle (ORC_DOUBLE: fp64
state, fp64 input sums): rasters agree until the first floating-point
divergence, and afterwards the population statistics agree within stated
tolerances (rate, CV of the inter-spike intervals, the STDP weight
distribution).  These pin how close the fp32 device arithmetic stays to
fp64, beyond the bit-exact checks against the fp32/fixed-point oracle."""
import numpy as np
import pytest

import oracle as O
from helpers import first_divergence, oracle_of, raster
from naer_amd import _abi, models

pytestmark = pytest.mark.gpu

STEPS = 5000  # 0.5 s


def cv_isi(sp, n_neurons, min_spikes=4):
    """Mean coefficient of variation of the inter-spike intervals over the
    neurons with at least min_spikes spikes."""
    order = np.lexsort((sp["step"], sp["neuron"]))
    nrn, stp = sp["neuron"][order].astype(np.int64), sp["step"][order].astype(np.float64)
    cvs = []
    bounds = np.searchsorted(nrn, np.arange(n_neurons + 1))
    for j in range(n_neurons):
        s = stp[bounds[j]:bounds[j + 1]]
        if len(s) >= min_spikes:
            isi = np.diff(s)
            cvs.append(isi.std() / isi.mean())
    return float(np.mean(cvs)), len(cvs)


def test_balanced_1k_raster_and_statistics_vs_fp64():
    # the configs' small network (J_E 0.1 mV, J_I -0.5 mV, in-degree 1000,
    # delay 1.5 ms, 20 kHz x 0.1 mV drive) at N = 1000
    net = models.balanced_10k(n=1000, mode=_abi.MODE_DETERMINISTIC, record_capacity=2_000_000)
    net.compile()
    o = oracle_of(net, O.DOUBLE)
    net.run(STEPS)
    o.run(STEPS)
    gs, os_ = net.spikes(), o.spikes()
    assert len(gs) > 5000
    d = first_divergence(gs, os_)
    if d is not None:
        # identical rasters up to the first divergence, which takes a while
        assert raster(gs)[d][0] >= 50, f"rasters diverge at step {raster(gs)[d][0]}"
    rate_g = len(gs) / 1000 / (STEPS * 1e-4)
    rate_o = len(os_) / 1000 / (STEPS * 1e-4)
    assert abs(rate_g - rate_o) <= 0.05 * rate_o, (rate_g, rate_o)
    cv_g, n_g = cv_isi(gs, 1000)
    cv_o, n_o = cv_isi(os_, 1000)
    print(f"first divergence step {None if d is None else raster(gs)[d][0]}, rate {rate_g:.2f} vs {rate_o:.2f} Hz, "
          f"CV-ISI {cv_g:.3f} vs {cv_o:.3f}")
    assert n_g > 500 and n_o > 500
    assert abs(cv_g - cv_o) <= 0.05, (cv_g, cv_o)
    net.close()


def test_stdp_weight_distribution_vs_fp64():
    # the configs' plastic network (pair STDP on E->E, w0 ~ U[0, wmax]) at N = 2000
    net = models.plastic_100k(n=2000, indegree=200, mode=_abi.MODE_DETERMINISTIC, record_capacity=2_000_000)
    net.compile()
    w0 = net.csr()[2].copy()
    o = oracle_of(net, O.DOUBLE)
    steps = 3000
    net.run(steps)
    o.run(steps)
    rp, col, w = net.csr()
    _, ocol, ow = o.csr()
    np.testing.assert_array_equal(col, ocol)  # same synapses in the same order
    plastic = (col & 0x80000000) != 0
    assert plastic.sum() > 100_000
    wmax = models.PAIR_STDP["w_max"]
    hg, edges = np.histogram(w[plastic], bins=20, range=(0.0, wmax))
    ho, _ = np.histogram(ow[plastic], bins=edges)
    tv = 0.5 * np.abs(hg / hg.sum() - ho / ho.sum()).sum()
    print(f"weight histogram TV distance {tv:.4f}, mean {w[plastic].mean():.5f} vs {ow[plastic].mean():.5f}")
    assert tv < 0.02, tv
    assert abs(w[plastic].mean() - ow[plastic].mean()) <= 0.01 * ow[plastic].mean()
    # plasticity acted: the weights moved away from their initial values
    assert np.abs(w[plastic] - w0[plastic]).mean() > 1e-3 * wmax
    net.close()
