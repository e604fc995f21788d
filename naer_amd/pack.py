"""Packs: several independent networks simulated by the same launches.

Wraps naer_pack_create (include/naer/naer.h).  Each network keeps its own
seeds, local neuron ids and local clock, so in deterministic mode its raster is
bit-identical to the raster it produces alone.
"""
import ctypes

import numpy as np

from . import _abi
from ._abi import check
from .network import Context


class Pack:
    """networks: list of naer_amd.network.Network (descriptions, not compiled)."""

    def __init__(self, networks, device=0, profiling=False):
        self.networks = list(networks)
        self._ctx = Context.get(device)
        descs = (_abi.NetDesc * len(self.networks))()
        self._keep = []
        for i, net in enumerate(self.networks):
            d = net.description()
            self._keep.append(net._keep)
            descs[i] = d
        self._descs = descs
        h = ctypes.c_void_p()
        check(_abi.lib().naer_pack_create(self._ctx.handle, descs, len(self.networks), ctypes.byref(h)))
        self._handle = h
        self.record_capacity = sum(n.record_capacity for n in self.networks)
        n = len(self.networks)
        self.begin = np.zeros(n, np.uint32)
        self.size = np.zeros(n, np.uint32)
        self.t_start = np.zeros(n, np.uint64)
        check(_abi.lib().naer_pack_layout(h, n, self.begin.ctypes.data, self.size.ctypes.data,
                                          self.t_start.ctypes.data, None))
        if profiling:
            check(_abi.lib().naer_net_set_profiling(h, 1))

    def build(self):
        check(_abi.lib().naer_net_build(self._handle))

    def autotune(self, batches=2):
        """Try delivery schedules on the running pack (real steps) and keep
        the fastest; returns the chosen schedule."""
        return _abi.autotune(self._handle, batches)

    def schedule(self):
        return _abi.get_schedule(self._handle)

    def set_schedule(self, task_synapses, segments_per_block, refine_split, run_aggregate=None):
        """refine_split 0: direct placement (events straight into their tile buckets)."""
        _abi.set_schedule(self._handle, task_synapses, segments_per_block, refine_split, run_aggregate)

    def set_graph(self, chunk_steps):
        check(_abi.lib().naer_net_set_graph(self._handle, int(chunk_steps)))

    def run(self, n_steps):
        check(_abi.lib().naer_net_step(self._handle, _abi.steps_arg(n_steps)))

    def sync(self):
        check(_abi.lib().naer_net_sync(self._handle))

    def stats(self):
        s = _abi.Stats()
        check(_abi.lib().naer_net_query(self._handle, ctypes.byref(s)))
        return s

    def spikes(self):
        """Per-network rasters (local neuron ids, local steps), each sorted."""
        cap = self.record_capacity
        out = np.empty(max(1, cap), dtype=_abi.SPIKE_DTYPE)
        got = ctypes.c_uint64()
        check(_abi.lib().naer_net_read_spikes(self._handle, out.ctypes.data, cap, ctypes.byref(got)))
        out = out[: got.value]
        net_of = np.searchsorted(self.begin.astype(np.int64), out["neuron"].astype(np.int64),
                                 side="right") - 1
        res = []
        for i in range(len(self.networks)):
            sp = out[net_of == i].copy()
            sp["neuron"] -= self.begin[i]
            sp["step"] -= np.uint32(self.t_start[i])
            res.append(_abi.sort_spikes(sp))
        return res

    def close(self):
        if self._handle is not None:
            _abi.lib().naer_net_destroy(self._handle)
            self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
