"""Python mirror of the reference's model-description API (Naer::Network).

Reference: include/naer/Network.h:15-56, Population.h:9-17, Connection.h:8-30
and src/Network.cpp.  Method names, argument meaning and the address layout
(populations grouped by type as [STDP | NOP | POISSON], src/Network.cpp:
269-285) follow the reference; failures raise RuntimeError subclasses as the
reference's ``throw std::runtime_error`` does.  Everything runs through the C
ABI of libnaer.so (include/naer/naer.h).

Additions the reference lacks, needed for the clock-driven step: LIF
parameters per population, a Poisson drive per neuron, fixed in-degree
connections with delays, and the STDP parameters.
"""
import ctypes

import numpy as np

from . import _abi
from ._abi import NaerError, check

REFERENCE_SEED = 0xBF785E3215AC7EBD          # src/Network.cpp:91
REFERENCE_POISSON_SEED = REFERENCE_SEED & 0xFFFFFFFF  # uint32 truncation, src/Network.cpp:124
REFERENCE_EPOCH_US = 1500                      # NAER_AXONAL_DELAY, naer_types.h:4


class NeuronType:
    """Population types of the reference (Network.h:82-84).

    NOP and STDP are both leaky integrate-and-fire populations; synapses from an
    STDP population onto an STDP population are plastic (stdp.cl:100 applies
    plasticity when the input spike comes from an STDP neuron).  POISSON is a
    spike source with the reference's exponential-ISI trains.
    """
    STDP = "STDP"
    NOP = "NOP"
    POISSON = "POISSON"
    ORDER = (STDP, NOP, POISSON)


class NeuronModel:
    """A neuron model given as HIP source, compiled with hipRTC
    (naer_model_create): the reference's NeuronType source strings
    (NeuronType.h:14), whose dynamics function its processing kernel calls
    (processing_neuron_group.cl:8-82).

    `source` defines

        __device__ bool naer_model_step(float* s, float I, float ext,
                                        const float* p, uint32_t t);

    called per neuron and step with the neuron's `n_state` floats (s[0] = the
    membrane potential, drawn in [v0_min, v0_max) at build), the step's
    synaptic input I, its Poisson drive ext, the population's params p and the
    network-local step t; it returns true when the neuron fires.  Pass the
    model as the type of addPopulation, with params=[...] (up to 16 floats).
    """

    def __init__(self, source, n_state, name="model"):
        h = ctypes.c_void_p()
        check(_abi.lib().naer_model_create(source.encode(), int(n_state), ctypes.byref(h)))
        self._handle = h
        self.source = source
        self.n_state = int(n_state)
        self.name = name

    def __repr__(self):
        return f"NeuronModel({self.name!r}, n_state={self.n_state})"

    def close(self):
        if self._handle is not None:
            _abi.lib().naer_model_destroy(self._handle)
            self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# Reference LIF constants (lif.cl:3-6), in volts and microseconds there.
REFERENCE_LIF = dict(tau_m_ms=20.0, v_th=0.020, v_reset=0.0, v_rest=0.0, t_ref_ms=2.0,
                     v0_min=0.0, v0_max=0.0, ext_rate_hz=0.0, ext_weight=0.0)


class Population:
    """Population.h:9-17: type, baseAddress, size (+ the LIF parameters)."""

    def __init__(self, type_, base_address, size, params):
        self.type = type_
        self.baseAddress = base_address
        self.size = size
        self.params = params

    def __repr__(self):
        return f"Population({self.type}, base={self.baseAddress}, size={self.size})"


class Connection:
    """Connection.h:8-30 plus the rule, in-degree and delay range."""

    def __init__(self, source, destination, defaultWeights, minRandomWeights, maxRandomWeights,
                 probability, seed, rule=_abi.RULE_BERNOULLI, indegree=0, delay=None):
        self.source = source
        self.destination = destination
        self.defaultWeights = defaultWeights
        self.minRandomWeights = minRandomWeights
        self.maxRandomWeights = maxRandomWeights
        self.probability = probability
        self.seed = seed
        self.rule = rule
        self.indegree = indegree
        self.delay = delay  # (min_steps, max_steps, quantum) or None = network default


class Context:
    """One HIP device (replaces the OpenCL set-up of Network::Network,
    src/Network.cpp:14-47)."""

    _shared = {}

    def __init__(self, device=0):
        h = ctypes.c_void_p()
        check(_abi.lib().naer_ctx_create(device, ctypes.byref(h)))
        self.handle = h
        self.device = device

    @classmethod
    def get(cls, device=0):
        if device not in cls._shared:
            cls._shared[device] = Context(device)
        return cls._shared[device]


class Network:
    """Naer::Network over the C ABI.

    dt_ms      step of the clock-driven update (0.1 ms)
    mode       _abi.MODE_FAST (fp32 atomics) or _abi.MODE_DETERMINISTIC
    step_us    simulated time advanced by step(); the reference advances one
               1500 us epoch (src/Network.cpp:125)
    """

    def __init__(self, device=0, dt_ms=0.1, mode=_abi.MODE_FAST, seed=REFERENCE_SEED,
                 poisson_seed=REFERENCE_POISSON_SEED, record_capacity=0, stdp=None,
                 default_delay_ms=REFERENCE_EPOCH_US / 1000.0, step_us=REFERENCE_EPOCH_US):
        self.device = device
        self.dt_ms = float(dt_ms)
        self.mode = mode
        self.seed = seed
        self.poisson_seed = poisson_seed
        self.record_capacity = record_capacity
        self.stdp = dict(stdp) if stdp else dict(tau_plus_ms=20.0, tau_minus_ms=20.0, a_plus=0.0,
                                                 a_minus=0.0, w_min=0.0, w_max=0.0)
        d = int(round(default_delay_ms / self.dt_ms))
        self.default_delay = (d, d, 1)
        self.step_us = step_us
        self._pops = []
        self._conns = []
        self._handle = None
        self._ctx = None
        self._type_counts = {t: 0 for t in NeuronType.ORDER}
        self.nop_time = self.stdp_time = self.poisson_time = 0.0
        self.scan_time = self.scatter_time = self.offset_time = 0.0

    # --- description (Network.cpp:76-104) ---------------------------------
    def addPopulation(self, type_, size, **params):
        """type_: a NeuronType, or a NeuronModel (then params=[...] holds the
        model's parameters; v0_min/v0_max/ext_rate_hz/ext_weight apply)."""
        model = isinstance(type_, NeuronModel)
        if not model and type_ not in NeuronType.ORDER:
            raise ValueError(f"unknown neuron type {type_!r}")
        if self._handle is not None:
            raise NaerError(_abi.NAER_ERR_STATE, "network already compiled")
        p = dict(REFERENCE_LIF)
        p["src_rate_hz"] = 20.0  # NAER_F_POISSON, input_neuron_group.cl:4
        if model:
            p["params"] = ()
        unknown = set(params) - set(p)
        if unknown:
            raise TypeError(f"unknown population parameters {sorted(unknown)}")
        p.update(params)
        if model:
            if len(p["params"]) > _abi.MODEL_PARAMS:
                raise ValueError(f"a model takes at most {_abi.MODEL_PARAMS} parameters")
            self._type_counts.setdefault(type_, 0)
        pop = Population(type_, self._type_counts[type_], int(size), p)
        self._type_counts[type_] += int(size)
        self._pops.append(pop)
        return pop

    def _types(self):
        """Address order of the population types: STDP, NOP, the models in
        order of first use, POISSON (Network.cpp:269-285 for the reference's
        three)."""
        models = [t for t in self._type_counts if isinstance(t, NeuronModel)]
        return [NeuronType.STDP, NeuronType.NOP, *models, NeuronType.POISSON]

    def addConnection(self, source, destination, defaultWeights, *args, delay_ms=None):
        """addConnection(src, dst, w, p) or addConnection(src, dst, w, wmin, wmax, p, seed)
        (Network.h:23-36); defaultWeights == 0 draws weights in [wmin, wmax)."""
        if len(args) == 1:
            wmin, wmax, p, seed = 0.0, 1.0, args[0], REFERENCE_SEED  # Network.cpp:90
        elif len(args) == 4:
            wmin, wmax, p, seed = args
        else:
            raise TypeError("addConnection(src, dst, w, p) or (src, dst, w, wmin, wmax, p, seed)")
        self._conns.append(Connection(source, destination, float(defaultWeights), float(wmin),
                                      float(wmax), float(p), int(seed), delay=self._delay(delay_ms)))

    def addConnectionFixedIndegree(self, source, destination, indegree, weight, wmin=0.0, wmax=0.0,
                                   seed=REFERENCE_SEED, delay_ms=None):
        """Every neuron of `destination` receives `indegree` synapses from sources
        drawn uniformly (with replacement) from `source`.  delay_ms: a number, or
        (min, max, quantum) in ms for delays uniform on that grid."""
        self._conns.append(Connection(source, destination, float(weight), float(wmin), float(wmax),
                                      0.0, int(seed), rule=_abi.RULE_FIXED_INDEGREE,
                                      indegree=int(indegree), delay=self._delay(delay_ms)))

    def _delay(self, delay_ms):
        if delay_ms is None:
            return None
        if isinstance(delay_ms, (tuple, list)):
            lo, hi, q = delay_ms
        else:
            lo = hi = q = delay_ms
        to = lambda x: int(round(x / self.dt_ms))
        return (to(lo), to(hi), max(1, to(q)))

    # --- compile (Network.cpp:267-308) -------------------------------------
    def ordered_populations(self):
        """Populations in address order: by type [STDP | NOP | POISSON], then
        insertion order, as Network::compile lays them out (Network.cpp:269-285)."""
        return [p for t in self._types() for p in self._pops if p.type is t or p.type == t]

    def _desc(self):
        order = self.ordered_populations()
        type_base = {}
        acc = 0
        for t in self._types():
            type_base[t] = acc
            acc += self._type_counts[t]
        for p in self._pops:
            p.address = type_base[p.type] + p.baseAddress
        index = {id(p): i for i, p in enumerate(order)}
        pops = (_abi.PopDesc * max(1, len(order)))()
        for i, p in enumerate(order):
            q = p.params
            if isinstance(p.type, NeuronModel):
                kind = _abi.POP_MODEL
            else:
                kind = _abi.POP_POISSON if p.type == NeuronType.POISSON else _abi.POP_LIF
            pops[i] = _abi.PopDesc(
                kind, p.size,
                q["tau_m_ms"], q["v_th"], q["v_reset"], q["v_rest"], q["t_ref_ms"], q["v0_min"],
                q["v0_max"], q["ext_rate_hz"], q["ext_weight"], q["src_rate_hz"])
            if kind == _abi.POP_MODEL:
                pops[i].model = p.type._handle
                for k, x in enumerate(q["params"]):
                    pops[i].params[k] = float(x)
        conns = (_abi.ConnDesc * max(1, len(self._conns)))()
        for i, c in enumerate(self._conns):
            dl = c.delay or self.default_delay
            plastic = int(c.source.type == NeuronType.STDP and c.destination.type == NeuronType.STDP)
            conns[i] = _abi.ConnDesc(index[id(c.source)], index[id(c.destination)], c.rule,
                                     c.indegree, c.probability, c.defaultWeights,
                                     c.minRandomWeights, c.maxRandomWeights, dl[0], dl[1], dl[2],
                                     plastic, 0, c.seed & 0xFFFFFFFFFFFFFFFF)
        d = _abi.NetDesc()
        d.pops = pops
        d.n_pops = len(order)
        d.conns = conns
        d.n_conns = len(self._conns)
        d.stdp = _abi.StdpDesc(**self.stdp)
        d.dt_ms = self.dt_ms
        d.mode = self.mode
        d.seed = self.seed
        d.poisson_seed = self.poisson_seed
        d.record_capacity = self.record_capacity
        self._keep = (pops, conns)
        return d

    def description(self):
        """The naer_net_desc this network compiles to (ctypes struct)."""
        return self._desc()

    def compile(self):
        if self._handle is not None:
            raise NaerError(_abi.NAER_ERR_STATE, "network already compiled")
        self._ctx = Context.get(self.device)
        d = self._desc()
        h = ctypes.c_void_p()
        check(_abi.lib().naer_net_create(self._ctx.handle, ctypes.byref(d), ctypes.byref(h)))
        self._handle = h
        check(_abi.lib().naer_net_build(h))

    # --- simulation (Network.cpp:122-265) ----------------------------------
    def _h(self):
        if self._handle is None:
            raise NaerError(_abi.NAER_ERR_STATE, "network not compiled")
        return self._handle

    def step(self):
        """Advance one reference epoch (step_us of simulated time)."""
        if self.mode == _abi.MODE_EVENT:
            check(_abi.lib().naer_net_step(self._h(), 1))
            return
        n = int(round(self.step_us / (self.dt_ms * 1000.0)))
        check(_abi.lib().naer_net_step(self._h(), n))

    def run(self, n_steps):
        """Advance n_steps steps of dt (asynchronous); in the event-driven mode
        (mode=_abi.MODE_EVENT) n_steps windows of 1500 us."""
        check(_abi.lib().naer_net_step(self._h(), _abi.steps_arg(n_steps)))

    def finish(self):
        check(_abi.lib().naer_net_sync(self._h()))
        self._refresh_timers()

    def schedule(self):
        """The delivery schedule (naer_schedule) as a dict."""
        return _abi.get_schedule(self._h())

    def set_schedule(self, task_synapses, segments_per_block, refine_split, run_aggregate=None, reserve=None):
        """reserve 1: reserved placement (refine_split: its parts per group bucket)."""
        _abi.set_schedule(self._h(), task_synapses, segments_per_block, refine_split, run_aggregate, reserve)

    def autotune(self, batches=2):
        """Try delivery schedules on the running simulation (real steps) and
        keep the fastest; returns the chosen schedule."""
        return _abi.autotune(self._h(), batches)

    def set_graph(self, chunk_steps):
        """Replay steps as captured HIP graphs of chunk_steps steps (0 = off)."""
        check(_abi.lib().naer_net_set_graph(self._h(), int(chunk_steps)))

    def set_profiling(self, on=True):
        check(_abi.lib().naer_net_set_profiling(self._h(), 1 if on else 0))

    def stats(self):
        s = _abi.Stats()
        check(_abi.lib().naer_net_query(self._h(), ctypes.byref(s)))
        return s

    def _refresh_timers(self):
        s = self.stats()
        # the reference's six timers (Network.h:50-55) map onto the three step
        # kernels: neuron update (nop/stdp/poisson), scatter (scan/scatter/offsets)
        self.nop_time = s.kernel_ms[_abi.K_UPDATE] / 1e3
        self.scatter_time = s.kernel_ms[_abi.K_SCATTER] / 1e3
        self.stdp_time = s.kernel_ms[_abi.K_POTENTIATE] / 1e3

    def getTime(self):
        """Simulated time in microseconds (uint32 in the reference)."""
        return int(self.stats().time_us) & 0xFFFFFFFF

    def getSynapses(self):
        return int(self.stats().n_synapses)

    def getNeurons(self):
        return sum(p.size for p in self._pops)

    # --- read-back ------------------------------------------------------------
    def spikes(self, cap=None):
        """Drain recorded spikes; structured array (step, neuron, time_us) sorted
        by (step, neuron, time_us)."""
        cap = self.record_capacity if cap is None else cap
        out = np.empty(max(1, cap), dtype=_abi.SPIKE_DTYPE)
        n = ctypes.c_uint64()
        check(_abi.lib().naer_net_read_spikes(self._h(), out.ctypes.data, cap, ctypes.byref(n)))
        return _abi.sort_spikes(out[: n.value])

    # --- checkpoint / resume (naer_net_save / naer_net_load) ----------------
    def save(self, path):
        """Write the dynamic state (at the current step) to `path`."""
        check(_abi.lib().naer_net_save(self._h(), str(path).encode()))

    def load(self, path):
        """Restore a state saved by a net built from the same description."""
        check(_abi.lib().naer_net_load(self._h(), str(path).encode()))

    # --- spike raster files (naer_net_raster_*) ---------------------------
    def raster_open(self, path):
        """Start a raster file; needs record_capacity > 0 (naer_amd/raster.py
        reads it)."""
        check(_abi.lib().naer_net_raster_open(self._h(), str(path).encode()))

    def raster_drain(self):
        """Append the spikes recorded since the last drain; returns their number."""
        n = ctypes.c_uint64()
        check(_abi.lib().naer_net_raster_drain(self._h(), ctypes.byref(n)))
        return n.value

    def raster_close(self):
        """Drain and finish the file; returns the total number of records."""
        n = ctypes.c_uint64()
        check(_abi.lib().naer_net_raster_close(self._h(), ctypes.byref(n)))
        return n.value

    def state(self):
        N = self.getNeurons()
        v = np.zeros(N, np.float32)
        r = np.zeros(N, np.int32)
        pre = np.zeros(N, np.float32)
        post = np.zeros(N, np.float32)
        check(_abi.lib().naer_net_read_state(self._h(), v.ctypes.data, r.ctypes.data,
                                             pre.ctypes.data, post.ctypes.data))
        return dict(v=v, ref=r, pre=pre, post=post)

    def csr(self):
        N = self.getNeurons()
        S = self.getSynapses()
        rp = np.zeros(N + 1, np.uint64)
        col = np.zeros(max(1, S), np.uint32)
        w = np.zeros(max(1, S), np.float32)
        check(_abi.lib().naer_net_read_csr(self._h(), rp.ctypes.data, col.ctypes.data, w.ctypes.data))
        return rp, col[:S], w[:S]

    def set_probe(self, t_begin, n_steps):
        check(_abi.lib().naer_net_set_probe(self._h(), int(t_begin), int(n_steps)))
        self._probe_n = int(n_steps)

    def probe(self):
        out = np.zeros((self._probe_n, self.getNeurons()), np.float32)
        check(_abi.lib().naer_net_read_probe(self._h(), out.ctypes.data))
        return out

    def close(self):
        if self._handle is not None:
            _abi.lib().naer_net_destroy(self._handle)
            self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def philox(ctr, key, device=0):
    """Philox4x32-10 blocks computed on the device (parity entry point)."""
    ctr = np.ascontiguousarray(ctr, dtype=np.uint32).reshape(-1, 4)
    key = np.ascontiguousarray(key, dtype=np.uint32).reshape(-1, 2)
    out = np.zeros_like(ctr)
    check(_abi.lib().naer_philox(Context.get(device).handle, ctr.ctypes.data, key.ctypes.data,
                                 out.ctypes.data, ctr.shape[0]))
    return out


def ref_poisson_epoch(n, n_pois_total, seed, rate_hz, t0, t1, device=0):
    """The reference's POISSON train of one epoch on the device: (times[n,10], count[n])."""
    times = np.zeros((n, 10), np.uint32)
    count = np.zeros(n, np.uint32)
    check(_abi.lib().naer_ref_poisson_epoch(Context.get(device).handle, n, n_pois_total, seed,
                                            rate_hz, t0, t1, times.ctypes.data, count.ctypes.data))
    return times, count
