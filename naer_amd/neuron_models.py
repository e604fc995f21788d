"""Neuron models the library ships as source (naer_model_source in
include/naer/naer.h) and their parameter vectors.

    lif       the built-in LIF step as a model (s = {V, refractory steps})
    cuba_lif  current-based LIF: a spike of weight w adds w to a synaptic
              current that decays with tau_s and charges V by w in total
              (s = {V, refractory steps, current}); the reference's vogels
              model names such a CubaLIF type (src/main.cpp:52-69)

Parameters are computed as the runtime computes its LIF constants: double
arithmetic on the float32 inputs, rounded to float32.
"""
import ctypes
import math

import numpy as np

from . import _abi
from ._abi import check

N_STATE = {"lif": 2, "cuba_lif": 3}
_cache = {}


def source(name):
    out = ctypes.c_char_p()
    check(_abi.lib().naer_model_source(name.encode(), ctypes.byref(out)))
    return out.value.decode()


def model(name):
    """The compiled model (one per process and name)."""
    from .network import NeuronModel
    if name not in _cache:
        _cache[name] = NeuronModel(source(name), N_STATE[name], name)
    return _cache[name]


def _f(x):
    return float(np.float32(x))


def _decay(dt_ms, tau_ms):
    return math.exp(-_f(dt_ms) / _f(tau_ms))


def lif_params(tau_m_ms, v_th, v_reset, v_rest, t_ref_ms, dt_ms):
    n_ref = float(math.floor(_f(t_ref_ms) / _f(dt_ms) + 0.5))  # llround for positive values
    return [_f(_decay(dt_ms, tau_m_ms)), _f(v_th), _f(v_reset), _f(v_rest), n_ref]


def cuba_lif_params(tau_m_ms, tau_s_ms, v_th, v_reset, v_rest, t_ref_ms, dt_ms):
    if _f(tau_s_ms) == _f(tau_m_ms):
        raise ValueError("cuba_lif needs tau_s != tau_m")
    p22 = _decay(dt_ms, tau_m_ms)
    p11 = _decay(dt_ms, tau_s_ms)
    p21 = _f(tau_m_ms) * (p11 - p22) / (_f(tau_s_ms) - _f(tau_m_ms))
    return lif_params(tau_m_ms, v_th, v_reset, v_rest, t_ref_ms, dt_ms) + [_f(p11), _f(p21)]
