"""Spike raster files (naer_net_raster_* in include/naer/naer.h).

Layout: a 40-byte little-endian header ("NAERRST1", uint32 version 1, uint32
bytes per record 8, float64 dt_ms, uint64 neurons, uint64 records) followed by
{uint32 step, uint32 neuron} records ordered by (step, neuron).
"""
import numpy as np

MAGIC = b"NAERRST1"
HEADER_BYTES = 40
HEADER = np.dtype([("magic", "S8"), ("version", "<u4"), ("record_bytes", "<u4"), ("dt_ms", "<f8"),
                   ("neurons", "<u8"), ("records", "<u8")])
RECORD = np.dtype([("step", "<u4"), ("neuron", "<u4")])


def read_raster(path):
    """(header dict, structured array of (step, neuron)); checks the header
    and that the record count matches the file."""
    with open(path, "rb") as f:
        raw = f.read()
    if len(raw) < HEADER_BYTES:
        raise ValueError(f"{path}: shorter than a raster header")
    h = np.frombuffer(raw[:HEADER_BYTES], HEADER, count=1)[0]
    if h["magic"] != MAGIC or h["version"] != 1 or h["record_bytes"] != RECORD.itemsize:
        raise ValueError(f"{path}: not a version-1 raster file")
    body = raw[HEADER_BYTES:]
    if len(body) % RECORD.itemsize:
        raise ValueError(f"{path}: truncated record")
    rec = np.frombuffer(body, RECORD)
    if int(h["records"]) != len(rec):
        raise ValueError(f"{path}: header says {int(h['records'])} records, file holds {len(rec)}")
    header = dict(dt_ms=float(h["dt_ms"]), neurons=int(h["neurons"]), records=int(h["records"]))
    return header, rec


def write_raster(path, steps, neurons, dt_ms, n_neurons):
    """Writes a raster file from arrays (sorted here); the same layout as the
    library's writer, for tools and tests."""
    rec = np.zeros(len(steps), RECORD)
    rec["step"] = steps
    rec["neuron"] = neurons
    rec = np.sort(rec, order=["step", "neuron"])
    h = np.zeros(1, HEADER)
    h["magic"] = MAGIC
    h["version"] = 1
    h["record_bytes"] = RECORD.itemsize
    h["dt_ms"] = dt_ms
    h["neurons"] = n_neurons
    h["records"] = len(rec)
    with open(path, "wb") as f:
        f.write(h.tobytes())
        f.write(rec.tobytes())
