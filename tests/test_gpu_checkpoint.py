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


@pytest.mark.parametrize("plastic,random_weights,reserve",
                         [(False, False, False), (True, False, False), (False, True, False), (False, False, True),
                          (True, False, True)])
def test_resume_continues_bit_for_bit(tmp_path, plastic, random_weights, reserve):
    """reserve: the saved net places its batches in reserved regions; the
    loading net keeps its default schedule and reads the batches in flight
    by the layout they were placed with."""
    path = tmp_path / "net.ckpt"
    a = multi_delay_net(plastic=plastic, random_weights=random_weights)
    a.compile()
    if reserve:
        s = a.schedule()
        a.set_schedule(s["task_synapses"], s["segments_per_block"], 4, reserve=1)
    a.run(157)  # not a multiple of the 15-step batch
    a.save(path)
    a.spikes()  # the saved spikes stay recorded in the checkpoint as well
    a.run(143)
    sa = a.spikes()
    st_a = a.state()
    w_a = a.csr()[2] if plastic else None
    b = multi_delay_net(plastic=plastic, random_weights=random_weights)
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


def test_load_rejects_another_description(tmp_path):
    """Same shape, another seed: the description hash in the header differs."""
    path = tmp_path / "net.ckpt"
    a = multi_delay_net()
    a.compile()
    a.run(30)
    a.save(path)
    a.close()
    b = multi_delay_net()
    b.seed += 1
    b.compile()
    with pytest.raises(NaerError, match="description"):
        b.load(path)
    b.close()


# byte offset of the saved device clock: magic (8), shape (11 x 8 + 8 x 4),
# step count (8), dropped spikes (8)
CLOCK_OFFSET = 8 + 120 + 16


@pytest.mark.parametrize("field,delta", [(0, 1), (8, 10**6), (28, 1)])
def test_load_rejects_corrupt_state_before_touching_the_device(tmp_path, field, delta):
    """Corrupt clock, batch start or list parity: the load fails on the dry
    pass, and the net keeps its own state."""
    path = tmp_path / "net.ckpt"
    a = multi_delay_net()
    a.compile()
    a.run(157)
    a.save(path)
    a.close()
    raw = bytearray(path.read_bytes())
    at = CLOCK_OFFSET + field
    width = 8 if field < 24 else 4
    v = int.from_bytes(raw[at:at + width], "little") + delta
    raw[at:at + width] = v.to_bytes(width, "little")
    bad = tmp_path / "bad.ckpt"
    bad.write_bytes(bytes(raw))
    b = multi_delay_net()
    b.compile()
    b.run(20)
    before = b.state()
    with pytest.raises(NaerError, match="corrupt"):
        b.load(bad)
    after = b.state()
    np.testing.assert_array_equal(before["v"].view(np.uint32), after["v"].view(np.uint32))
    assert b.stats().step == 20
    b.load(path)  # the intact file still loads
    assert b.stats().step == 157
    b.close()
