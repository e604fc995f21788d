"""Raster file format (CPU): the reader checks the header and record count;
a file written by the Python writer round-trips ordered by (step, neuron)."""
import numpy as np
import pytest

from naer_amd.raster import HEADER_BYTES, read_raster, write_raster


def test_round_trip_sorted(tmp_path):
    rng = np.random.default_rng(5)
    steps = rng.integers(0, 1000, 5000).astype(np.uint32)
    neurons = rng.integers(0, 70000, 5000).astype(np.uint32)
    p = tmp_path / "r.naer"
    write_raster(p, steps, neurons, 0.1, 70000)
    h, rec = read_raster(p)
    assert h == dict(dt_ms=0.1, neurons=70000, records=5000)
    key = rec["step"].astype(np.uint64) << 32 | rec["neuron"]
    assert np.all(np.diff(key.astype(np.int64)) >= 0)
    exp = np.sort(steps.astype(np.uint64) << 32 | neurons)
    np.testing.assert_array_equal(key, exp)
    assert p.stat().st_size == HEADER_BYTES + 8 * 5000


def test_reader_rejects_bad_files(tmp_path):
    p = tmp_path / "bad.naer"
    p.write_bytes(b"NOTARASTER" + bytes(40))
    with pytest.raises(ValueError):
        read_raster(p)
    good = tmp_path / "g.naer"
    write_raster(good, np.arange(3, dtype=np.uint32), np.arange(3, dtype=np.uint32), 0.1, 3)
    raw = good.read_bytes()
    (tmp_path / "trunc.naer").write_bytes(raw[:-3])
    with pytest.raises(ValueError):
        read_raster(tmp_path / "trunc.naer")
    (tmp_path / "short.naer").write_bytes(raw[:-8])
    with pytest.raises(ValueError):
        read_raster(tmp_path / "short.naer")


def test_sort_spikes_equals_the_field_sort():
    """The library's raster sort (two stable integer argsorts) orders records
    as numpy's sort by the fields (step, neuron, time_us), including several
    spikes of one neuron in one step (POISSON sources)."""
    from naer_amd import _abi
    rng = np.random.default_rng(7)
    n = 20_000
    sp = np.zeros(n, dtype=_abi.SPIKE_DTYPE)
    sp["step"] = rng.integers(0, 300, n)
    sp["neuron"] = rng.integers(0, 50, n)
    sp["time_us"] = rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)
    want = np.sort(sp, order=["step", "neuron", "time_us"])
    assert np.array_equal(_abi.sort_spikes(sp), want)
    assert len(_abi.sort_spikes(sp[:1])) == 1 and len(_abi.sort_spikes(sp[:0])) == 0


def test_step_counts_outside_uint32_are_refused():
    from naer_amd import _abi
    assert _abi.steps_arg(0) == 0 and _abi.steps_arg(2**32 - 1) == 2**32 - 1
    for bad in (-1, -480, 2**32):
        with pytest.raises(ValueError):
            _abi.steps_arg(bad)
