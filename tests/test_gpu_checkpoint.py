"""Checkpoint and resume (naer_net_save / naer_net_load): a net built from
the same description, loaded from a checkpoint taken mid-run (also mid-batch),
continues bit for bit as the saved net does, in deterministic mode."""
import numpy as np
import pytest

from helpers import first_divergence
from naer_amd import _abi
from naer_amd.network import NaerError
from test_gpu_bucketed import multi_delay_net

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("plastic", [False, True])
def test_resume_continues_bit_for_bit(tmp_path, plastic):
    path = tmp_path / "net.ckpt"
    a = multi_delay_net(plastic=plastic)
    a.compile()
    a.run(157)  # not a multiple of the 15-step batch
    a.save(path)
    a.spikes()  # the saved spikes stay recorded in the checkpoint as well
    a.run(143)
    sa = a.spikes()
    st_a = a.state()
    w_a = a.csr()[2] if plastic else None
    b = multi_delay_net(plastic=plastic)
    b.compile()
    b.load(path)
    first = b.spikes()  # the undrained spikes of the first 157 steps
    assert len(first) > 0 and first["step"].max() < 157
    b.run(143)
    sb = b.spikes()
    assert len(sa) > 200
    assert first_divergence(sa, sb) is None
    st_b = b.state()
    for k in ("v", "ref", "pre", "post"):
        np.testing.assert_array_equal(st_a[k].view(np.uint32), st_b[k].view(np.uint32))
    if plastic:
        np.testing.assert_array_equal(w_a.view(np.uint32), b.csr()[2].view(np.uint32))
    assert a.stats().events == b.stats().events and a.stats().step == b.stats().step
    a.close()
    b.close()


def test_load_rejects_another_shape(tmp_path):
    path = tmp_path / "net.ckpt"
    a = multi_delay_net()
    a.compile()
    a.run(30)
    a.save(path)
    a.close()
    from naer_amd import models
    b = models.balanced(4000, 400, 0.1, 5.0, 20_000.0, 0.1, mode=_abi.MODE_DETERMINISTIC)
    b.compile()
    with pytest.raises(NaerError):
        b.load(path)
    b.close()
