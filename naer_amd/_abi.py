"""ctypes mirror of the C ABI in include/naer/naer.h.

The shared library is loaded from naer_amd/lib/libnaer.so (built in-tree by
``python -m naer_amd.build``).  There is no fallback: when the library is
missing or the device is absent, calls raise NaerError.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# NAER_LIB: load another build of the library (e.g. a variant under test)
LIB_PATH = os.environ.get("NAER_LIB") or os.path.join(_HERE, "lib", "libnaer.so")

NAER_OK = 0
NAER_ERR_INVALID = -1
NAER_ERR_HIP = -2
NAER_ERR_NOMEM = -3
NAER_ERR_CAPACITY = -4
NAER_ERR_STATE = -5
NAER_ERR_NODEVICE = -6

POP_LIF = 0
POP_POISSON = 1
POP_MODEL = 2
MODEL_STATE = 8
MODEL_PARAMS = 16
RULE_BERNOULLI = 0
RULE_FIXED_INDEGREE = 1
MODE_FAST = 0
MODE_DETERMINISTIC = 1
MODE_EVENT = 2

STREAM_V0 = 1
STREAM_DRIVE = 2
STREAM_CONN = 3
DRIVE_TABLE = 64

K_UPDATE, K_SCATTER, K_POTENTIATE = 0, 1, 2
K_BUCKET_HIST, K_BUCKET_SCAN, K_BUCKET_PLACE, K_BUCKET_REFINE = 3, 4, 5, 6
K_MODEL = 7

u32, u64, f32, f64 = ctypes.c_uint32, ctypes.c_uint64, ctypes.c_float, ctypes.c_double


class PopDesc(ctypes.Structure):
    _fields_ = [
        ("kind", u32), ("size", u32),
        ("tau_m_ms", f32), ("v_th", f32), ("v_reset", f32), ("v_rest", f32), ("t_ref_ms", f32),
        ("v0_min", f32), ("v0_max", f32), ("ext_rate_hz", f32), ("ext_weight", f32),
        ("src_rate_hz", f32),
        ("model", ctypes.c_void_p), ("params", f32 * 16),
    ]


class ConnDesc(ctypes.Structure):
    _fields_ = [
        ("src_pop", u32), ("dst_pop", u32), ("rule", u32), ("indegree", u32),
        ("probability", f64),
        ("weight", f32), ("wmin", f32), ("wmax", f32),
        ("delay_min", u32), ("delay_max", u32), ("delay_quantum", u32),
        ("plastic", u32), ("_pad", u32),
        ("seed", u64),
    ]


class StdpDesc(ctypes.Structure):
    _fields_ = [("tau_plus_ms", f32), ("tau_minus_ms", f32), ("a_plus", f32), ("a_minus", f32),
                ("w_min", f32), ("w_max", f32)]


class NetDesc(ctypes.Structure):
    _fields_ = [
        ("pops", ctypes.POINTER(PopDesc)), ("n_pops", u32), ("n_conns", u32),
        ("conns", ctypes.POINTER(ConnDesc)),
        ("stdp", StdpDesc),
        ("dt_ms", f32), ("mode", u32), ("seed", u64),
        ("poisson_seed", u32), ("record_capacity", u32),
    ]


class ArenaDesc(ctypes.Structure):
    _fields_ = [
        ("max_neurons", u64), ("max_synapses", u64), ("refractory_min", u32),
        ("dt_ms", f32), ("mode", u32), ("delay_min", u32), ("delay_max", u32), ("max_row", u32),
        ("plastic", u32), ("stdp", StdpDesc), ("record_capacity", u32), ("poisson", u32),
    ]


class Spike(ctypes.Structure):
    _fields_ = [("step", u32), ("neuron", u32), ("time_us", u32), ("_pad", u32)]


class Schedule(ctypes.Structure):
    _fields_ = [("bucketed", ctypes.c_uint32), ("task_synapses", ctypes.c_uint32),
                ("segments_per_block", ctypes.c_uint32), ("refine_split", ctypes.c_uint32),
                ("run_aggregate", ctypes.c_uint32), ("reserve", ctypes.c_uint32),
                ("parts_max", ctypes.c_uint32), ("record_bytes", ctypes.c_uint32),
                ("delivery_us", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Stats(ctypes.Structure):
    _fields_ = [
        ("n_neurons", u64), ("n_synapses", u64), ("n_plastic", u64), ("step", u64),
        ("time_us", u64), ("spikes", u64), ("events", u64), ("dropped", u64),
        ("ring_depth", u32), ("batch", u32), ("max_row", u32), ("kernels", u32),
        ("kernel_ms", f64 * 8), ("kernel_calls", u64 * 8), ("device_bytes", u64),
    ]


SPIKE_DTYPE = [("step", "<u4"), ("neuron", "<u4"), ("time_us", "<u4"), ("_pad", "<u4")]


def sort_spikes(sp):
    """Spike records sorted by (step, neuron, time_us): two stable integer
    argsorts (time_us, then a 64-bit (step, neuron) key), ~10x faster than a
    structured-array sort by fields on millions of records."""
    import numpy as np
    if len(sp) < 2:
        return sp.copy()
    i = np.argsort(sp["time_us"], kind="stable")
    key = (sp["step"].astype(np.uint64) << np.uint64(32)) | sp["neuron"].astype(np.uint64)
    return sp[i[np.argsort(key[i], kind="stable")]]


class NaerError(RuntimeError):
    """A C ABI call returned an error (the message is naer_last_error())."""

    def __init__(self, code, msg):
        super().__init__(f"naer error {code}: {msg}")
        self.code = code


_lib = None

_SIGS = {
    "naer_ctx_create": [ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)],
    "naer_ctx_destroy": [ctypes.c_void_p],
    "naer_net_create": [ctypes.c_void_p, ctypes.POINTER(NetDesc), ctypes.POINTER(ctypes.c_void_p)],
    "naer_pack_create": [ctypes.c_void_p, ctypes.POINTER(NetDesc), u32, ctypes.POINTER(ctypes.c_void_p)],
    "naer_net_get_schedule": [ctypes.c_void_p, ctypes.c_void_p],
    "naer_net_set_schedule": [ctypes.c_void_p, ctypes.c_void_p],
    "naer_net_autotune": [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p],
    "naer_net_save": [ctypes.c_void_p, ctypes.c_char_p],
    "naer_net_load": [ctypes.c_void_p, ctypes.c_char_p],
    "naer_net_raster_open": [ctypes.c_void_p, ctypes.c_char_p],
    "naer_net_raster_drain": [ctypes.c_void_p, ctypes.c_void_p],
    "naer_net_raster_close": [ctypes.c_void_p, ctypes.c_void_p],
    "naer_pack_layout": [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                         ctypes.c_void_p],
    "naer_arena_create": [ctypes.c_void_p, ctypes.POINTER(ArenaDesc), ctypes.POINTER(ctypes.c_void_p)],
    "naer_arena_admit": [ctypes.c_void_p, ctypes.POINTER(NetDesc), ctypes.POINTER(u32)],
    "naer_arena_retire": [ctypes.c_void_p, u32],
    "naer_net_check_finite": [ctypes.c_void_p, ctypes.c_void_p, u32, ctypes.POINTER(u32)],
    "naer_net_build": [ctypes.c_void_p],
    "naer_net_step": [ctypes.c_void_p, u32],
    "naer_net_sync": [ctypes.c_void_p],
    "naer_net_idle": [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)],
    "naer_net_set_graph": [ctypes.c_void_p, u32],
    "naer_net_query": [ctypes.c_void_p, ctypes.POINTER(Stats)],
    "naer_net_set_profiling": [ctypes.c_void_p, ctypes.c_int],
    "naer_net_read_spikes": [ctypes.c_void_p, ctypes.c_void_p, u64, ctypes.POINTER(u64)],
    "naer_net_read_state": [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                            ctypes.c_void_p],
    "naer_net_read_csr": [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p],
    "naer_net_set_probe": [ctypes.c_void_p, u64, u32],
    "naer_net_read_probe": [ctypes.c_void_p, ctypes.c_void_p],
    "naer_net_destroy": [ctypes.c_void_p],
    "naer_philox": [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, u64],
    "naer_ref_poisson_epoch": [ctypes.c_void_p, u32, u32, u32, f32, u32, u32, ctypes.c_void_p,
                               ctypes.c_void_p],
    "naer_ref_poisson_epoch_noalloc": [ctypes.c_void_p, u32, u32, u32, f32, u32, u32, ctypes.c_void_p,
                               ctypes.c_void_p],
    "naer_model_create": [ctypes.c_char_p, u32, ctypes.POINTER(ctypes.c_void_p)],
    "naer_model_destroy": [ctypes.c_void_p],
    "naer_model_source": [ctypes.c_char_p, ctypes.POINTER(ctypes.c_char_p)],
}

# every entry point include/naer/naer.h declares
EXPORTS = sorted(list(_SIGS) + ["naer_last_error", "naer_version"])


def lib():
    """The loaded libnaer.so (raises NaerError if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise NaerError(NAER_ERR_STATE,
                            f"{LIB_PATH} is missing; build it with `python -m naer_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, args in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = ctypes.c_int
        L.naer_last_error.restype = ctypes.c_char_p
        L.naer_last_error.argtypes = []
        L.naer_version.restype = ctypes.c_char_p
        L.naer_version.argtypes = []
        _lib = L
    return _lib


def steps_arg(n_steps):
    """A step count for naer_net_step (uint32): negative or larger counts are refused."""
    n = int(n_steps)
    if n < 0 or n > 0xFFFFFFFF:
        raise ValueError(f"step count {n} outside [0, 2^32)")
    return n


def get_schedule(handle):
    s = Schedule()
    check(lib().naer_net_get_schedule(handle, ctypes.byref(s)))
    return s.as_dict()


def set_schedule(handle, task_synapses, segments_per_block, refine_split, run_aggregate=None, reserve=None):
    """run_aggregate / reserve None keep the current counter mode / placement."""
    cur = get_schedule(handle)
    if run_aggregate is None:
        run_aggregate = cur["run_aggregate"]
    if reserve is None:
        reserve = cur["reserve"]
    s = Schedule(0, int(task_synapses), int(segments_per_block), int(refine_split), int(run_aggregate),
                 int(reserve), 0, 0, 0.0)
    check(lib().naer_net_set_schedule(handle, ctypes.byref(s)))


def autotune(handle, batches=2):
    s = Schedule()
    check(lib().naer_net_autotune(handle, int(batches), ctypes.byref(s)))
    return s.as_dict()


def check(rc):
    if rc != NAER_OK:
        raise NaerError(rc, lib().naer_last_error().decode(errors="replace"))
    return rc
