"""Shared helpers of the parity tests: build the same description on the
device (naer_amd) and in the CPU oracle."""
import numpy as np

import oracle as O


def oracle_of(net, precision=O.FIXED, threads=8):
    return O.OracleNet(net.description(), precision, threads)


def raster(sp):
    return np.stack([sp["step"], sp["neuron"], sp["time_us"]], axis=1).astype(np.int64)


def first_divergence(a, b):
    """Index of the first differing spike of two sorted rasters (None if equal)."""
    ra, rb = raster(a), raster(b)
    n = min(len(ra), len(rb))
    bad = np.nonzero((ra[:n] != rb[:n]).any(axis=1))[0]
    if len(bad):
        return int(bad[0])
    return None if len(ra) == len(rb) else n


def rates_hz(sp, n_neurons, steps, dt_ms):
    return len(sp) / n_neurons / (steps * dt_ms * 1e-3)
