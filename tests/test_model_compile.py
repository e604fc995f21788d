"""Source-defined neuron models compile without a GPU (naer_model_create runs
hipRTC only), report compiler errors, and the descriptor layout the Python
binding uses matches the C header."""
import ctypes
import os
import subprocess

import pytest

from naer_amd import _abi
from naer_amd.network import NeuronModel

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

pytestmark = pytest.mark.skipif(not os.path.exists(_abi.LIB_PATH), reason="libnaer.so not built")

STEP = r"""
__device__ bool naer_model_step(float* s, float I, float ext, const float* p, uint32_t t)
{
    s[0] = s[0] * p[0] + I + ext;
    const bool f = s[0] > p[1];
    if (f)
        s[0] = p[2];
    return f;
}
"""


def test_model_compiles_without_device():
    m = NeuronModel(STEP, 1)
    assert m._handle
    m.close()


@pytest.mark.parametrize("name", ["lif", "cuba_lif"])
def test_shipped_models_compile(name):
    from naer_amd import neuron_models
    src = neuron_models.source(name)
    assert "naer_model_step" in src
    m = NeuronModel(src, neuron_models.N_STATE[name], name)
    m.close()


def test_unknown_shipped_model():
    from naer_amd import neuron_models
    with pytest.raises(_abi.NaerError, match="cuba_lif"):
        neuron_models.source("hodgkin_huxley")


def test_cuba_lif_parameters():
    """Closed forms of the exact step: P22 = e^(-dt/tau_m), P11 = e^(-dt/tau_s),
    P21 = tau_m (P11 - P22) / (tau_s - tau_m); as tau_s -> 0 the current model
    becomes a voltage jump one step later (P21 -> P22)."""
    import math
    from naer_amd import neuron_models
    p = neuron_models.cuba_lif_params(20.0, 5.0, 20.0, 10.0, 0.0, 2.0, 0.1)
    assert p[4] == 20.0
    assert p[0] == pytest.approx(math.exp(-0.1 / 20.0), rel=1e-7)
    assert p[5] == pytest.approx(math.exp(-0.1 / 5.0), rel=1e-7)
    p11, p22 = math.exp(-0.1 / 5.0), math.exp(-0.1 / 20.0)
    assert p[6] == pytest.approx(20.0 * (p11 - p22) / (5.0 - 20.0), rel=1e-5)
    q = neuron_models.cuba_lif_params(20.0, 1e-4, 20.0, 10.0, 0.0, 2.0, 0.1)
    assert q[6] == pytest.approx(q[0], rel=1e-3)
    with pytest.raises(ValueError):
        neuron_models.cuba_lif_params(20.0, 20.0, 20.0, 10.0, 0.0, 2.0, 0.1)


def test_compile_error_carries_the_compiler_log():
    bad = STEP.replace("p[2]", "p[2] + undeclared_thing")
    with pytest.raises(_abi.NaerError) as e:
        NeuronModel(bad, 1)
    assert e.value.code == _abi.NAER_ERR_INVALID
    assert "undeclared_thing" in str(e.value)


def test_missing_step_function_is_an_error():
    with pytest.raises(_abi.NaerError):
        NeuronModel("__device__ float unrelated(float x) { return x; }", 1)


@pytest.mark.parametrize("n_state", [0, _abi.MODEL_STATE + 1])
def test_state_size_is_bounded(n_state):
    with pytest.raises(_abi.NaerError, match="n_state"):
        NeuronModel(STEP, n_state)


def test_pop_desc_layout_matches_header(tmp_path):
    src = tmp_path / "layout.c"
    src.write_text(
        '#include <stddef.h>\n#include <stdio.h>\n#include "naer/naer.h"\n'
        "int main(void) {\n"
        '  printf("%zu %zu %zu %zu %zu\\n", sizeof(naer_pop_desc), offsetof(naer_pop_desc, src_rate_hz),\n'
        "         offsetof(naer_pop_desc, model), offsetof(naer_pop_desc, params), sizeof(naer_net_desc));\n"
        "  return 0;\n}\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]
    P = _abi.PopDesc
    assert got == [ctypes.sizeof(P), P.src_rate_hz.offset, P.model.offset, P.params.offset,
                   ctypes.sizeof(_abi.NetDesc)]


def test_model_populations_sit_between_nop_and_poisson():
    """Address order STDP | NOP | models (grouped per model, in order of
    first use) | POISSON, and the descriptor carries the model and params."""
    from naer_amd.network import Network, NeuronType
    a, b = NeuronModel(STEP, 1, "a"), NeuronModel(STEP, 1, "b")
    net = Network(mode=_abi.MODE_DETERMINISTIC)
    P = net.addPopulation(NeuronType.POISSON, 5)
    A1 = net.addPopulation(a, 3, params=[0.9, 1.0, 0.0])
    N1 = net.addPopulation(NeuronType.NOP, 4)
    B1 = net.addPopulation(b, 2, params=[0.5])
    S1 = net.addPopulation(NeuronType.STDP, 6)
    A2 = net.addPopulation(a, 7, params=[0.8])
    d = net.description()
    assert [S1.address, N1.address, A1.address, A2.address, B1.address, P.address] == [0, 6, 10, 13, 20, 22]
    kinds = [d.pops[i].kind for i in range(d.n_pops)]
    assert kinds == [_abi.POP_LIF, _abi.POP_LIF, _abi.POP_MODEL, _abi.POP_MODEL, _abi.POP_MODEL, _abi.POP_POISSON]
    assert d.pops[2].model == a._handle.value and d.pops[4].model == b._handle.value
    assert list(d.pops[2].params)[:3] == pytest.approx([0.9, 1.0, 0.0])
    with pytest.raises(ValueError):
        net.addPopulation(a, 1, params=[0.0] * (_abi.MODEL_PARAMS + 1))
