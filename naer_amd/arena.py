"""Arena: preallocated device capacity in which networks (jobs) are admitted
and retired between steps (naer_arena_* in include/naer/naer.h)."""
import ctypes

import numpy as np

from . import _abi
from ._abi import NaerError, check
from .network import Context


class Arena:
    """Continuous-batching container.

    max_neurons / max_synapses                 capacities (device memory is
                                               allocated up front)
    delay_steps                                (min, max) delay every job must use
    plastic                                    jobs may carry STDP
    refractory_min                             plastic: fewest refractory steps a
                                               job's LIF populations may have
    """

    def __init__(self, max_neurons, max_synapses, dt_ms=0.1, mode=_abi.MODE_DETERMINISTIC,
                 delay_steps=(15, 15), max_row=8192, plastic=False, stdp=None, refractory_min=0,
                 record_capacity=1 << 20, poisson=False, device=0):
        self._ctx = Context.get(device)
        d = _abi.ArenaDesc()
        d.max_neurons = int(max_neurons)
        d.max_synapses = int(max_synapses)
        d.refractory_min = int(refractory_min)
        d.dt_ms = dt_ms
        d.mode = mode
        d.delay_min, d.delay_max = delay_steps
        d.max_row = max_row
        d.plastic = 1 if plastic else 0
        if stdp:
            d.stdp = _abi.StdpDesc(**stdp)
        d.record_capacity = int(record_capacity)
        d.poisson = 1 if poisson else 0
        self.desc = d
        self.dt_ms = dt_ms
        self.record_capacity = int(record_capacity)
        h = ctypes.c_void_p()
        check(_abi.lib().naer_arena_create(self._ctx.handle, ctypes.byref(d), ctypes.byref(h)))
        self._handle = h
        self.slots = {}  # slot -> network description keep-alive

    def admit(self, network):
        """Build `network` (an uncompiled naer_amd.network.Network) into the
        arena; returns its slot, or None when it does not fit now.  A job
        that can never be admitted (larger than the arena, settings the arena
        was not created with) raises NaerError (NAER_ERR_INVALID), as does a
        failed build (e.g. NAER_ERR_NOMEM)."""
        d = network.description()
        s = ctypes.c_uint32()
        rc = _abi.lib().naer_arena_admit(self._handle, ctypes.byref(d), ctypes.byref(s))
        if rc == _abi.NAER_ERR_CAPACITY:
            return None
        check(rc)
        self.slots[s.value] = network
        return s.value

    def check(self):
        """Slots of the active jobs holding a potential that is not finite
        (naer_net_check_finite)."""
        cap = max(1, len(self.slots))
        out = (ctypes.c_uint32 * cap)()
        n = ctypes.c_uint32()
        check(_abi.lib().naer_net_check_finite(self._handle, out, cap, ctypes.byref(n)))
        return [int(out[i]) for i in range(min(cap, n.value))]

    def retire(self, slot):
        """Free the job's ranges.  Drain first: spikes still in the record
        buffer are attributed by neuron range."""
        check(_abi.lib().naer_arena_retire(self._handle, int(slot)))
        self.slots.pop(slot, None)

    def layout(self):
        """(neuron_begin, n_neurons, t_start) per slot, retired slots included."""
        n = ctypes.c_uint32()
        check(_abi.lib().naer_pack_layout(self._handle, 0, None, None, None, ctypes.byref(n)))
        n = n.value
        begin = np.zeros(max(1, n), np.uint32)
        size = np.zeros(max(1, n), np.uint32)
        t0 = np.zeros(max(1, n), np.uint64)
        check(_abi.lib().naer_pack_layout(self._handle, n, begin.ctypes.data, size.ctypes.data, t0.ctypes.data,
                                          None))
        return begin[:n], size[:n], t0[:n]

    def set_graph(self, chunk_steps):
        check(_abi.lib().naer_net_set_graph(self._handle, int(chunk_steps)))

    def run(self, n_steps):
        check(_abi.lib().naer_net_step(self._handle, _abi.steps_arg(n_steps)))

    def idle(self):
        """True when the arena's enqueued steps are done (naer_net_idle; does not wait)."""
        v = ctypes.c_int()
        check(_abi.lib().naer_net_idle(self._handle, ctypes.byref(v)))
        return v.value != 0

    def sync(self):
        check(_abi.lib().naer_net_sync(self._handle))

    def schedule(self):
        """Delivery schedule of the arena (naer_net_get_schedule)."""
        return _abi.get_schedule(self._handle)

    def set_schedule(self, task_synapses, segments_per_block, refine_split, run_aggregate=None):
        """refine_split 0: direct placement for the batches whose placement
        counters fit the arena's occupied prefix of tiles."""
        _abi.set_schedule(self._handle, task_synapses, segments_per_block, refine_split, run_aggregate)

    def stats(self):
        s = _abi.Stats()
        check(_abi.lib().naer_net_query(self._handle, ctypes.byref(s)))
        return s

    def drain(self):
        """Recorded spikes since the last drain, split per active slot:
        {slot: array(step, neuron, time_us)} with local neuron ids and local steps."""
        cap = self.record_capacity
        out = np.empty(max(1, cap), dtype=_abi.SPIKE_DTYPE)
        got = ctypes.c_uint64()
        check(_abi.lib().naer_net_read_spikes(self._handle, out.ctypes.data, cap, ctypes.byref(got)))
        out = out[: got.value]
        begin, size, t0 = self.layout()
        res = {}
        for slot in self.slots:
            m = (out["neuron"] >= begin[slot]) & (out["neuron"] < begin[slot] + size[slot])
            sp = out[m].copy()
            sp["neuron"] -= begin[slot]
            sp["step"] -= np.uint32(t0[slot])
            res[slot] = sp
        return res

    def dropped(self):
        return int(self.stats().dropped)

    def close(self):
        if self._handle is not None:
            _abi.lib().naer_net_destroy(self._handle)
            self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
