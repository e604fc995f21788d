"""Spike raster files written by the library (naer_net_raster_*) against the
oracle's raster, and the file layout naer_amd/raster.py reads."""
import numpy as np
import pytest

import oracle as O
from helpers import oracle_of
from test_gpu_bucketed import multi_delay_net

pytestmark = pytest.mark.gpu


def test_raster_file_equals_oracle(tmp_path):
    """Spikes written through naer_net_raster_* (drains between runs) equal
    the oracle's raster, in (step, neuron) order."""
    from naer_amd.raster import read_raster

    net = multi_delay_net(record=400_000)
    net.compile()
    o = oracle_of(net, O.FIXED)
    path = tmp_path / "spikes.naer"
    net.raster_open(path)
    for c in (100, 37, 163):
        net.run(c)
        net.raster_drain()
    total = net.raster_close()
    o.run(300)
    h, rec = read_raster(path)
    assert h["records"] == total and h["neurons"] == net.getNeurons() and h["dt_ms"] == pytest.approx(0.1)
    ref = o.spikes()
    assert total == len(ref) > 500
    np.testing.assert_array_equal(rec["step"], ref["step"])
    np.testing.assert_array_equal(rec["neuron"], ref["neuron"])
    net.close()
