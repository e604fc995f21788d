"""The C++ drop-in (include/naer/Network.h over the C ABI) through the CLI
driver with the reference's interface (src/main.cpp: --synapses, --model;
vogels runs the library's current-based LIF model, compiled with hipRTC):
the driver's network has the size and gives the spike and event counts of the
same model built through the Python binding, which is bit-exact with the
oracle (test_gpu_parity.py).  Deterministic mode, so the counts are exact."""
import os
import re
import subprocess

import pytest

from naer_amd import _abi, models

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "naer_amd", "bin", "naer")
# the driver's brunel+ STDP (volts, as its weights): src/main.cpp:12-30 scaled to w_max = 0.3 mV
DRIVER_STDP = dict(tau_plus_ms=20.0, tau_minus_ms=20.0, a_plus=0.01 * 0.0003, a_minus=0.0105 * 0.0003,
                   w_min=0.0, w_max=0.0003)


def run_driver(*args):
    out = subprocess.run([DRIVER, *args], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    n, s = map(int, re.search(r"\(N, S\): \((\d+), (\d+)\)", out.stdout).groups())
    t_us, spikes, events = map(int, re.search(r"Simulated: (\d+) us, spikes (\d+), synaptic events (\d+)",
                                              out.stdout).groups())
    return n, s, t_us, spikes, events


@pytest.mark.parametrize("model", ["brunel", "brunel+", "vogels"])
def test_driver_equals_python_binding(model):
    n, s, t_us, spikes, events = run_driver("--synapses", "1000000", "--model", model, "--mode", "det",
                                            "--seconds", "0.03")
    assert t_us == 30_000  # 20 epochs of 1500 us
    if model == "brunel":
        net = models.brunel(1_000_000, mode=_abi.MODE_DETERMINISTIC)
    elif model == "vogels":
        net = models.vogels(1_000_000, mode=_abi.MODE_DETERMINISTIC)
    else:
        net = models.brunel_plus(1_000_000, mode=_abi.MODE_DETERMINISTIC, stdp=DRIVER_STDP)
    net.compile()
    net.run(t_us // 100)
    st = net.stats()
    assert (n, s) == (net.getNeurons(), net.getSynapses())
    assert spikes > 100
    assert (spikes, events) == (int(st.spikes), int(st.events))
    net.close()


def test_driver_event_mode_equals_python_binding():
    """--mode event: the reference's event-driven step, 20 windows of 1500 us."""
    n, s, t_us, spikes, events = run_driver("--synapses", "1000000", "--model", "brunel", "--mode", "event",
                                            "--seconds", "0.03")
    assert t_us == 30_000
    net = models.brunel(1_000_000, mode=_abi.MODE_EVENT)
    net.compile()
    net.run(20)
    st = net.stats()
    assert (n, s) == (net.getNeurons(), net.getSynapses())
    assert spikes > 100
    assert (spikes, events) == (int(st.spikes), int(st.events))
    net.close()
