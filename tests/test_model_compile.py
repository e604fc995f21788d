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
