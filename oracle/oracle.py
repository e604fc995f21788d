"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker; never by the product path.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")

FIXED, FLOAT, DOUBLE = 0, 1, 2

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run `make -C oracle`")
        L = ctypes.CDLL(LIB_PATH)
        v, p, u32, u64 = ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64
        L.orc_philox_n.argtypes = [p, p, p, u64]
        L.orc_ref_uniform.argtypes = [u32, u32, u32]
        L.orc_ref_uniform.restype = ctypes.c_double
        L.orc_ref_isi_us.argtypes = [ctypes.c_float, u32, u32, u32]
        L.orc_ref_isi_us.restype = u32
        L.orc_ref_poisson_gid.argtypes = [u32]
        L.orc_ref_poisson_gid.restype = u32
        L.orc_ref_poisson_epoch.argtypes = [u32, u32, u32, ctypes.c_float, u32, u32, p]
        L.orc_ref_poisson_epoch.restype = ctypes.c_int
        L.orc_ref_bernoulli_row.argtypes = [u32, u32, ctypes.c_double, u32, p, u32]
        L.orc_ref_bernoulli_row.restype = u32
        L.orc_net_create.argtypes = [p, ctypes.c_int, ctypes.c_int]
        L.orc_net_create.restype = v
        L.orc_error.restype = ctypes.c_char_p
        L.orc_net_build.argtypes = [v]
        L.orc_net_run.argtypes = [v, u32]
        L.orc_net_destroy.argtypes = [v]
        for name in ("orc_net_n_neurons", "orc_net_n_synapses", "orc_net_step_count",
                     "orc_net_spike_count", "orc_net_event_count"):
            getattr(L, name).argtypes = [v]
            getattr(L, name).restype = u64
        L.orc_net_ring_depth.argtypes = [v]
        L.orc_net_ring_depth.restype = u32
        L.orc_net_read_csr.argtypes = [v, p, p, p]
        L.orc_net_read_state.argtypes = [v, p, p, p, p]
        L.orc_net_read_state_f64.argtypes = [v, p]
        L.orc_net_read_spikes.argtypes = [v, p, u64]
        L.orc_net_read_spikes.restype = u64
        L.orc_net_set_probe.argtypes = [v, u64, u32]
        L.orc_net_read_probe.argtypes = [v, p, p, p]
        L.orc_ev_run.argtypes = [v, ctypes.POINTER(ctypes.c_void_p), u32]
        L.orc_ev_run.restype = ctypes.c_int
        L.orc_ev_read_state.argtypes = [v, u32, p, p]
        L.orc_ev_read_stdp.argtypes = [v, u32, p, p]
        L.orc_expf.argtypes = [ctypes.c_float]
        L.orc_expf.restype = ctypes.c_float
        L.orc_ev_free.argtypes = [v]
        L.orc_ev_tables.argtypes = [ctypes.c_double, p, p, u32, ctypes.POINTER(u32)]
        L.orc_set_conn_mutation.argtypes = [u32]
        _lib = L
    return _lib


def expf(x):
    """orc_expf: the single-precision e^x of the event-driven STDP."""
    return float(lib().orc_expf(float(x)))


# Planted mistakes of the fixed in-degree rule (naer_oracle.h), for tests
# that must fail on them.
MUT_NONE, MUT_SRC_RANGE, MUT_SRC_BASE, MUT_WORDS_SWAPPED, MUT_DELAY_REUSE, MUT_WEIGHT_REUSE, MUT_WEIGHT_SPAN = range(7)


def set_conn_mutation(m):
    """Plant mistake m in the oracle's fixed in-degree rule (MUT_NONE: the
    rule as specified) for the networks built afterwards."""
    lib().orc_set_conn_mutation(int(m))


def philox(ctr, key):
    ctr = np.ascontiguousarray(ctr, dtype=np.uint32).reshape(-1, 4)
    key = np.ascontiguousarray(key, dtype=np.uint32).reshape(-1, 2)
    out = np.zeros_like(ctr)
    lib().orc_philox_n(ctr.ctypes.data, key.ctypes.data, out.ctypes.data, ctr.shape[0])
    return out


def ref_uniform(gid, seed, it):
    return lib().orc_ref_uniform(gid, seed, it)


def ref_isi_us(rate, gid, seed, it):
    return lib().orc_ref_isi_us(rate, gid, seed, it)


def ref_poisson_epoch(idx, n_pois, seed, rate, t0, t1):
    buf = np.zeros(10, np.uint32)
    n = lib().orc_ref_poisson_epoch(idx, n_pois, seed, rate, t0, t1, buf.ctypes.data)
    return buf[:n].copy()


def ev_tables(tau_us, cap_hi=8192):
    """The event-driven mode's decay tables L[l] = exp(-l/T), H[h] =
    exp(-4096 h/T) and the number of nonzero H entries."""
    lo = np.zeros(4096, np.float64)
    hi = np.zeros(cap_hi, np.float64)
    n = ctypes.c_uint32()
    lib().orc_ev_tables(float(tau_us), lo.ctypes.data, hi.ctypes.data, cap_hi, ctypes.byref(n))
    return lo, hi[: n.value]


def ref_bernoulli_row(src, n_dst, p, seed32):
    buf = np.zeros(max(1, n_dst), np.uint32)
    n = lib().orc_ref_bernoulli_row(src, n_dst, p, seed32, buf.ctypes.data, n_dst)
    return buf[:n].copy()


SPIKE_DTYPE = [("step", "<u4"), ("neuron", "<u4"), ("time_us", "<u4"), ("_pad", "<u4")]


class OracleNet:
    """The CPU oracle of one network description (a naer_net_desc ctypes struct)."""

    def __init__(self, desc, precision=FIXED, threads=1):
        self._desc = desc  # keep the arrays alive
        h = lib().orc_net_create(ctypes.byref(desc), precision, threads)
        if not h:
            raise RuntimeError("oracle: " + lib().orc_error().decode())
        self.h = h
        self.precision = precision
        self.record_capacity = desc.record_capacity
        if lib().orc_net_build(h) != 0:
            raise RuntimeError("oracle: " + lib().orc_error().decode())
        self.N = lib().orc_net_n_neurons(h)
        self.S = lib().orc_net_n_synapses(h)

    def run(self, steps):
        if lib().orc_net_run(self.h, int(steps)) != 0:
            raise RuntimeError("oracle: " + lib().orc_error().decode())

    @property
    def steps(self):
        return lib().orc_net_step_count(self.h)

    @property
    def spike_count(self):
        return lib().orc_net_spike_count(self.h)

    @property
    def event_count(self):
        return lib().orc_net_event_count(self.h)

    @property
    def ring_depth(self):
        return lib().orc_net_ring_depth(self.h)

    def csr(self):
        rp = np.zeros(self.N + 1, np.uint64)
        col = np.zeros(max(1, self.S), np.uint32)
        w = np.zeros(max(1, self.S), np.float32)
        lib().orc_net_read_csr(self.h, rp.ctypes.data, col.ctypes.data, w.ctypes.data)
        return rp, col[: self.S], w[: self.S]

    def state(self):
        N = self.N
        v = np.zeros(N, np.float32)
        r = np.zeros(N, np.int32)
        pre = np.zeros(N, np.float32)
        post = np.zeros(N, np.float32)
        lib().orc_net_read_state(self.h, v.ctypes.data, r.ctypes.data, pre.ctypes.data, post.ctypes.data)
        return dict(v=v, ref=r, pre=pre, post=post)

    def v64(self):
        v = np.zeros(self.N, np.float64)
        lib().orc_net_read_state_f64(self.h, v.ctypes.data)
        return v

    def spikes(self):
        out = np.zeros(max(1, self.record_capacity), dtype=SPIKE_DTYPE)
        n = lib().orc_net_read_spikes(self.h, out.ctypes.data, self.record_capacity)
        return np.sort(out[:n], order=["step", "neuron", "time_us"])

    def set_probe(self, t_begin, n):
        lib().orc_net_set_probe(self.h, int(t_begin), int(n))
        self._probe_n = int(n)

    def probe(self):
        shape = (self._probe_n, self.N)
        used = np.zeros(shape, np.float32)
        exact = np.zeros(shape, np.float64)
        absw = np.zeros(shape, np.float64)
        lib().orc_net_read_probe(self.h, used.ctypes.data, exact.ctypes.data, absw.ctypes.data)
        return used, exact, absw

    # --- event-driven mode (orc_ev_run) -----------------------------------
    def run_events(self, epochs):
        """Windows of 1500 us of the reference's event-driven step."""
        if not hasattr(self, "_ev"):
            self._ev = ctypes.c_void_p()
        if lib().orc_ev_run(self.h, ctypes.byref(self._ev), int(epochs)) != 0:
            raise RuntimeError("oracle: " + lib().orc_error().decode())

    def event_state(self):
        """Potentials, last input times and, for neurons of populations with
        a plastic connection, the post trace and inter-spike history."""
        v = np.zeros(self.N, np.float32)
        t_last = np.zeros(self.N, np.uint32)
        trace = np.zeros(self.N, np.float32)
        hist = np.zeros((self.N, 10), np.uint16)
        lib().orc_ev_read_state(self._ev, self.N, v.ctypes.data, t_last.ctypes.data)
        lib().orc_ev_read_stdp(self._ev, self.N, trace.ctypes.data, hist.ctypes.data)
        return dict(v=v, t_last=t_last, trace=trace, hist=hist)

    def close(self):
        if getattr(self, "_ev", None):
            lib().orc_ev_free(self._ev)
            self._ev = None
        if self.h:
            lib().orc_net_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
