// kernels_event.hip — the event-driven mode: the reference's epoch
// (src/Network.cpp:122-256) and neuron step (lif.cl:15-34) at microsecond
// resolution.
//
// One window of 1500 us per epoch:
//   k_ev_poisson   POISSON spikes of the window (input_neuron_group.cl:26-39)
//                  appended to the window's input list, which already holds
//                  the LIF spikes of the previous window (arriving 1500 us
//                  after they were fired)
//   (radix sort of the inputs by (time, source))
//   k_ev_rowlen    row length of every input; an exclusive scan gives each
//                  input's range of events
//   k_ev_expand    events {target << 11 | time offset, weight} in input
//                  order, i.e. by (time, source, row position)
//   (stable radix sort of the events by target alone: each target's events
//    keep the order (time, source, row position))
//   k_ev_bounds    each target's run of sorted events
//   k_ev_process   a lane per LIF neuron applies its events in order and
//                  appends its spikes to the next window's input list
// These replace histogram_scan / scatter / distribute_offsets
// (spike_aggregation.cl:3-65), which order a window's spikes by time.  The
// sorts and the scan are this file's own: an LSD radix sort of 32-bit keys
// (and 32-bit values) in passes of up to 8 bits (11 for up to 2^20 items) —
// per tile of 4096 items a digit histogram, one flat scan of the histograms,
// and a stable scatter that ranks equal digits with wave ballots and writes
// each tile out in digit order — and a three-kernel exclusive scan of 64-bit
// row lengths.
#include <algorithm>

#include <hip/hip_runtime.h>

#include "naer_device.hpp"
#include "naer_expf.hpp"
#include "naer_kernels.hpp"

namespace naer {

namespace {

__device__ __forceinline__ uint32_t ev_pop(const EvArgs& a, uint32_t j)
{
    uint32_t p = a.blk_pop[j >> a.tw_shift] & ~kTileIdle;
    while (p + 1 < a.n_pops && j >= a.pops[p + 1].begin)
        ++p;
    return p;
}

__device__ __forceinline__ void ev_record(const EvArgs& a, uint32_t j, uint32_t time)
{
    if (a.rec_cap) {
        const unsigned long long r = atomicAdd(a.rec_n, 1ull);
        if (r < a.rec_cap) {
            naer_spike sp;
            sp.step = a.epoch;
            sp.neuron = j;
            sp.time_us = time;
            sp._pad = 0;
            a.rec[r] = sp;
        }
    }
}

// An input of the window starting at t_window: time offset << kEvSrcBits |
// source; plastic nets: its slot, and a plastic neuron's payload (stdp.cl:123-127).
__device__ __forceinline__ void ev_append(const EvArgs& a, uint32_t j, uint32_t time, uint32_t t_window,
                                          bool payload = false, uint32_t prev = 0, float trace = 0.0f)
{
    const uint32_t k = atomicAdd(a.n_out, 1u);
    if (k < a.in_cap) {
        a.out_keys[k] = (time - t_window) << kEvSrcBits | j;
        if (a.out_vals) {
            a.out_vals[k] = k;
            if (payload) {
                a.out_prev[k] = prev;
                a.out_trace[k] = trace;
            }
        }
    } else {
        atomicAdd(a.lost, 1u);
    }
}

// POISSON sources: the reference train of the window [t0, t0 + 1500).
__global__ __launch_bounds__(256) void k_ev_poisson(const EvArgs a)
{
    const uint32_t j = blockIdx.x * 256 + threadIdx.x;
    if (j >= a.N)
        return;
    const uint32_t p = ev_pop(a, j);
    const PopParams& pp = a.pops[p];
    if (pp.kind != NAER_POP_POISSON || j < pp.begin || j >= pp.end)
        return;
    const uint32_t idx = pp.pois_base + (j - pp.begin);
    const uint32_t gid = ref_poisson_gid(idx);
    const uint32_t t0 = a.t0, t1 = a.t0 + kEpochUs;
    uint32_t time = t0 + ref_isi_us(pp.neg_rate_per_us, gid, pp.pois_seed, t0 * pp.n_pois + idx);
    uint32_t spikes = 0;
    for (uint32_t m = 0; m < kMaxPoissonPerEpoch && time < t1; ++m) {
        ev_append(a, j, time, t0);
        ev_record(a, j, time);
        ++spikes;
        time += ref_isi_us(pp.neg_rate_per_us, gid, pp.pois_seed, time * pp.n_pois + idx);
    }
    if (spikes)
        atomicAdd(a.spikes, (unsigned long long)spikes);
}

__global__ __launch_bounds__(256) void k_ev_rowlen(const EvArgs a, uint32_t n_in)
{
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i > n_in)
        return;
    unsigned long long len = 0;
    if (i < n_in) {
        const uint32_t s = a.in_keys[i] & ((1u << kEvSrcBits) - 1u);
        len = a.row_ptr[s + 1] - a.row_ptr[s];
    }
    a.row_len[i] = len;
}

// One block per input spike walks its row.  Plastic nets also write every
// event's synapse (r < 2^31, | kPlasticBit for a plastic one) and its input's
// slot; a plastic synapse occurs at most once per window (its source fires at
// most once), so the weight read here is the one its event updates.
template <bool PL>
__global__ __launch_bounds__(256) void k_ev_expand(const EvArgs a)
{
    const uint32_t key = a.in_keys[blockIdx.x];
    const uint32_t s = key & ((1u << kEvSrcBits) - 1u);
    const uint32_t dt = key >> kEvSrcBits;  // < 1500
    const uint32_t slot = PL ? a.in_vals[blockIdx.x] : 0u;
    const unsigned long long r0 = a.row_ptr[s], r1 = a.row_ptr[s + 1];
    const unsigned long long e0 = a.row_off[blockIdx.x];
    for (unsigned long long r = r0 + threadIdx.x; r < r1; r += 256) {
        const uint32_t c = a.col[r];
        const uint32_t tgt = c & 0xFFFFFFu;
        const unsigned long long e = e0 + (r - r0);
        a.ev_keys[e] = tgt << 11 | dt;
        a.ev_w[e] = a.w[r];
        if (PL) {
            a.ev_syn[e] = (uint32_t)r | (c & kPlasticBit);
            a.ev_slot[e] = slot;
        }
    }
}

__device__ __forceinline__ double ev_decay(const double* H, const double* L, uint32_t n_hi, long long dt)
{
    const unsigned long long h = (unsigned long long)dt >> 12;
    return h < n_hi ? H[h] * L[dt & 4095] : 0.0;
}

// Boundaries of each target's run of sorted events: begin[t] and end[t] of
// the targets that have events (the step kernel clears them after use, so
// targets without events read an empty range).
// A thread per aligned quad of keys (the key arrays hold whole quads).
__global__ __launch_bounds__(256) void k_ev_bounds(const EvArgs a)
{
    const uint32_t i0 = (blockIdx.x * 256 + threadIdx.x) * 4;
    if (i0 >= a.n_ev)
        return;
    const uint4 k = *reinterpret_cast<const uint4*>(a.ev_keys + i0);
    const uint32_t before = i0 ? a.ev_keys[i0 - 1] >> 11 : 0xFFFFFFFFu;
    const uint32_t after = i0 + 4 < a.n_ev ? a.ev_keys[i0 + 4] >> 11 : 0xFFFFFFFFu;
    const uint32_t t[6] = { before, k.x >> 11, k.y >> 11, k.z >> 11, k.w >> 11, after };
#pragma unroll
    for (uint32_t m = 0; m < 4; ++m) {
        const uint32_t i = i0 + m;
        if (i >= a.n_ev)
            break;
        const uint32_t nxt = i + 1 < a.n_ev ? t[m + 2] : 0xFFFFFFFFu;
        if (t[m] != t[m + 1])
            a.ev_begin[t[m + 1]] = i;
        if (nxt != t[m + 1])
            a.ev_end[t[m + 1]] = i + 1;
    }
}

// lif.cl:15-34 per event, in time order; the decay e^(-dt/tau) is H[dt >> 12]
// * L[dt & 4095] from the population's tables.  One lane per target walks its
// run of sorted events in batches of aligned quads (16-byte loads) through a
// three-stage pipeline: batch b + 2 is loaded, batch b + 1 is decoded (event
// times, each event's predecessor, the decay from it and, plastic, its
// input's payload) and batch b is stepped.  Loads are unconditional (indices
// clamped into range) and their results used an iteration later, so none is
// waited for early.  A block whose LIF targets share one population keeps
// its L table and the head of its H table in LDS.  The step looks the decay
// up again only when t_last is not the previous event's time (after a spike
// or a dropped input).
constexpr uint32_t kEvHiLds = 64;    // H entries kept in LDS (dt < 2^18 us)
constexpr uint32_t kEvStepThreads = 64;  // one wave per block: the waves spread over every CU
constexpr uint32_t kEvSpkLane = 32;       // spikes a lane buffers in LDS (two batches' worth)

// stdp.cl:8-30 for one plastic event (the reference's update_synapse with
// basic_stdp.cl's rule), in the reference's integer and float types:
//   LTP over the inter-spike history while the post spike time (int) is
//   above the input's previous time (int); each step evaluates
//   exp((float)(uint)(prev - post) / T) (the unsigned difference of a later
//   post time wraps, so the factor overflows to +inf and the increment is 0);
//   the post trace decays from the last spike time (uint difference; a
//   negative one wraps and clears the trace); LTD with alpha * lambda in
//   double; hard bounds +-W_MAX.
// The LTP sum adds exactly 0 while every time is below kEvLtpZeroBelow: a
// step runs only when post > prev as ints, and then (uint)(prev - post) is
// at least 2^31 - 2^21 (both real times: the difference wraps; either one a
// wrapped "last" >= 2^32 - 2^18 * 10: it is at least that far from any real
// time), so e^((float)diff / 20000) overflows to +inf, pre = +inf (pre_trace
// >= 1: it is a trace just after its +1), e^(-pre / 100) = 0 and w + 0 = w.
constexpr uint32_t kEvLtpZeroBelow = (1u << 31) - (1u << 21);

__device__ __forceinline__ float ev_stdp_update(float w, const uint32_t (&hh)[kEvHist], uint32_t last,
                                                uint32_t t, uint32_t prev, float pre_trace, float& trace,
                                                bool ltp_zero)
{
    int32_t post = (int32_t)last;
    if (!ltp_zero) {
#pragma unroll
        for (uint32_t h = 1; h < kEvHist; ++h) {
            if (!(post > (int32_t)prev && hh[h] > 0u))
                break;
            const float pre = naer_expf((float)(prev - (uint32_t)post) / kEvStdpTau) * pre_trace + 1.0f;
            const double inc = kEvStdpLambda * (double)(1.0f - w) * (double)naer_expf(-pre / kEvStdpConstDt);
            w = (float)((double)w + inc);
            post -= (int32_t)(hh[h] * kEvHistUnit);
        }
    }
    const float dt = (float)(t - last);
    trace = dt > 0.0f ? naer_expf(-dt / kEvStdpTau) * trace : trace;
    const double dec = (double)kEvStdpAlpha * kEvStdpLambda * (double)w * (double)naer_expf(-trace / kEvStdpConstDt);
    w = (float)((double)w - dec);
    return fmaxf(fminf(w, kEvStdpWMax), -kEvStdpWMax);
}

// PL: the net has plastic connections.  A neuron of a population with a
// plastic connection then follows stdp.cl:94-128: last = t_last - 4 hist[0];
// a plastic event first updates its synapse (written back), then the neuron
// steps with the updated weight; hist[0] = (t - last) / 4 (16 bits); a spike
// shifts the history, clears hist[0], adds 1 to the trace and sends
// {prev = last, trace} with the spike.  Its other inputs (static synapses)
// step and keep the history the same way.
template <bool PL>
__global__ __launch_bounds__(kEvStepThreads) void k_ev_process(const EvArgs a)
{
    constexpr uint32_t QB = PL ? 2 : 4;  // quads per batch
    constexpr uint32_t EB = 4 * QB;      // events per batch
    __shared__ double s_lo[kEvLo];
    __shared__ double s_hi[kEvHiLds];
    __shared__ uint32_t s_pop[2];
    // each lane's spikes {time; plastic: last, trace}, written out together
    // after the stepping (a global atomic per spike inside the serial loop
    // would stall the wave and drain its prefetched loads).  Every step
    // stores its candidate entry at the lane's next slot and a spike keeps
    // it (no branch); a lane that might fill its slots in the next batch
    // sends them out first.
    __shared__ uint32_t s_spk_t[kEvSpkLane][kEvStepThreads];
    __shared__ uint32_t s_spk_last[PL ? kEvSpkLane : 1][kEvStepThreads];
    __shared__ float s_spk_tr[PL ? kEvSpkLane : 1][kEvStepThreads];
    __shared__ unsigned long long s_base[2];
    const uint32_t j = blockIdx.x * kEvStepThreads + threadIdx.x;
    uint32_t lo = 0, hi = 0, p = 0;
    if (j < a.N) {
        lo = a.ev_begin[j];
        hi = a.ev_end[j];
        if (lo != hi) {
            a.ev_begin[j] = 0;  // ready for the next window
            a.ev_end[j] = 0;
            p = ev_pop(a, j);
            const PopParams& pp = a.pops[p];
            if (pp.kind != NAER_POP_LIF || j < pp.begin || j >= pp.end)
                hi = lo;
        }
    }
    // one population for the block's stepping lanes: its tables go to LDS
    if (threadIdx.x == 0) {
        s_pop[0] = 0xFFFFFFFFu;
        s_pop[1] = 0;
    }
    __syncthreads();
    if (lo != hi) {
        atomicMin(&s_pop[0], p);
        atomicMax(&s_pop[1], p);
    }
    __syncthreads();
    if (s_pop[0] == 0xFFFFFFFFu)
        return;  // nothing to step in this block
    const bool lds = s_pop[0] == s_pop[1];
    if (lds) {
        const uint32_t q = s_pop[0];
        const uint32_t nh = min(a.n_hi[q], kEvHiLds);
        for (uint32_t x = threadIdx.x; x < kEvLo; x += kEvStepThreads)
            s_lo[x] = a.dec_lo[(size_t)q * kEvLo + x];
        for (uint32_t x = threadIdx.x; x < kEvHiLds; x += kEvStepThreads)
            s_hi[x] = x < nh ? a.dec_hi[(size_t)q * kEvHiCap + x] : 0.0;
    }
    __syncthreads();
    const double* L = a.dec_lo + (size_t)p * kEvLo;
    const double* H = a.dec_hi + (size_t)p * kEvHiCap;
    uint32_t n_hi = 1, t_ref = 0, t_last = 0;
    float v_rest = 0.0f, v_th = 0.0f, v_reset = 0.0f, V = 0.0f;
    bool plastic = false;
    float trace = 0.0f;
    uint32_t hh[kEvHist];
#pragma unroll
    for (uint32_t h = 0; h < kEvHist; ++h)
        hh[h] = 0;
    if (lo != hi) {
        n_hi = a.n_hi[p];
        t_ref = a.t_ref[p];
        v_rest = a.pops[p].v_rest;
        v_th = a.pops[p].v_th;
        v_reset = a.pops[p].v_reset;
        V = a.V[j];
        t_last = a.t_last[j];
        if (PL && a.pop_plastic[p]) {
            plastic = true;
            trace = a.trace[j];
#pragma unroll
            for (uint32_t h = 0; h < kEvHist / 2; ++h) {
                const uint32_t x = a.hist[(size_t)j * (kEvHist / 2) + h];
                hh[2 * h] = x & 0xFFFFu;
                hh[2 * h + 1] = x >> 16;
            }
        }
    }
    // e^(-dt/tau) for dt >= 0 (0 past the tables), from LDS where it can
    auto decay = [&](uint32_t dt) -> double {
        const uint32_t h = dt >> 12;
        if (h >= n_hi)
            return 0.0;
        if (lds && h < kEvHiLds)
            return s_hi[h] * s_lo[dt & 4095];
        return H[h] * L[dt & 4095];
    };
    // batches start at the aligned quad holding lo; event e of the lane is
    // valid when lo <= e < hi
    const uint32_t first = lo & ~3u;
    struct Raw {
        uint4 key[QB];
        float4 w[QB];
        uint4 syn[QB], slot[QB];
    };
    struct Dec {
        uint32_t t[EB], tp[EB], syn[EB], slot[EB];
        float w[EB];
        double d[EB];
    };
    auto load = [&](uint32_t c, Raw& r) {
#pragma unroll
        for (uint32_t q = 0; q < QB; ++q) {
            // past the lane's run: reread its first quad (in range, unused)
            const uint32_t e = c + 4 * q < hi ? c + 4 * q : first;
            r.key[q] = *reinterpret_cast<const uint4*>(a.ev_keys + e);
            r.w[q] = *reinterpret_cast<const float4*>(a.ev_w + e);
            if (PL) {
                r.syn[q] = *reinterpret_cast<const uint4*>(a.ev_syn + e);
                r.slot[q] = *reinterpret_cast<const uint4*>(a.ev_slot + e);
            }
        }
    };
    auto pick = [](const uint4& v, uint32_t m) { return m == 0 ? v.x : m == 1 ? v.y : m == 2 ? v.z : v.w; };
    auto pickf = [](const float4& v, uint32_t m) { return m == 0 ? v.x : m == 1 ? v.y : m == 2 ? v.z : v.w; };
    // decode batch c: times, predecessors (t_before: the one before the
    // batch), decays; returns the batch's last valid event time
    auto decode = [&](uint32_t c, const Raw& r, uint32_t t_before, Dec& k) -> uint32_t {
        uint32_t tp = t_before;
#pragma unroll
        for (uint32_t x = 0; x < EB; ++x) {
            const uint32_t e = c + x;
            const bool valid = e >= lo && e < hi;
            const uint32_t t = valid ? a.t0 + (pick(r.key[x / 4], x % 4) & 2047u) : 0u;
            k.t[x] = t;
            k.w[x] = pickf(r.w[x / 4], x % 4);
            k.tp[x] = tp;
            k.d[x] = valid && t >= tp ? decay(t - tp) : 0.0;
            k.syn[x] = PL && plastic && valid ? pick(r.syn[x / 4], x % 4) : 0u;
            k.slot[x] = PL ? pick(r.slot[x / 4], x % 4) : 0u;
            if (valid)
                tp = t;
        }
        return tp;
    };
    const bool ltp_zero = a.t0 + 2 * kEpochUs <= kEvLtpZeroBelow;
    Dec k0;
    Raw r1;
    uint32_t t_end;
    {
        Raw r0;
        load(first, r0);
        t_end = decode(first, r0, t_last, k0);
    }
    load(first + EB, r1);
    uint32_t ns = 0;  // the lane's buffered spikes
    unsigned long long n_spikes = 0;
    auto send = [&](uint32_t m) {  // the lane's first m buffered spikes, one by one
        for (uint32_t i = 0; i < m; ++i) {
            const uint32_t t = s_spk_t[i][threadIdx.x];
            ev_append(a, j, t + kEpochUs, a.t0 + kEpochUs, PL, PL ? s_spk_last[i][threadIdx.x] : 0u,
                      PL ? s_spk_tr[i][threadIdx.x] : 0.0f);
            ev_record(a, j, t);
        }
    };
    for (uint32_t c = first;; c += EB) {
        if (!__any(c < hi))
            break;
        if (ns > kEvSpkLane - EB) {  // rare: only short refractory periods fire this often
            send(ns);
            n_spikes += ns;
            ns = 0;
        }
        Dec k1;
        const uint32_t t_next = decode(c + EB, r1, t_end, k1);
        Raw r2;
        load(c + 2 * EB, r2);
        // plastic: the payloads of this batch's plastic events
        uint32_t prev[EB];
        float ptr[EB];
#pragma unroll
        for (uint32_t x = 0; x < EB; ++x) {
            prev[x] = 0;
            ptr[x] = 0.0f;
            if (PL && (k0.syn[x] & kPlasticBit)) {
                const uint32_t sl = k0.slot[x] < a.in_cap ? k0.slot[x] : 0u;
                prev[x] = a.in_prev[sl];
                // stdp.cl:84 holds the pre trace in a uint
                ptr[x] = (float)(uint32_t)a.in_trace[sl];
            }
        }
#pragma unroll
        for (uint32_t x = 0; x < EB; ++x) {
            const bool valid = c + x >= lo && c + x < hi;
            const uint32_t tk = k0.t[x];
            float wk = k0.w[x];
            uint32_t last = 0;
            if (PL && plastic) {
                last = t_last - hh[0] * kEvHistUnit;
                if (valid && (k0.syn[x] & kPlasticBit)) {
                    wk = ev_stdp_update(wk, hh, last, tk, prev[x], ptr[x], trace, ltp_zero);
                    a.w_mut[k0.syn[x] & ~kPlasticBit] = wk;
                }
            }
            // the step, computed on every lane and kept where it applies
            // (refractory inputs are dropped)
            const bool take = valid && tk >= t_last;
            double dk = k0.d[x];
            if (take && k0.tp[x] != t_last)  // after a spike or a dropped input
                dk = decay(tk - t_last);
            float xv = V - v_rest;
            xv = (float)((double)xv * dk);
            float v = xv + v_rest;
            v = v + wk;
            const bool spiked = take && v > v_th;
            V = take ? (spiked ? v_reset : v) : V;
            t_last = take ? (spiked ? tk + t_ref : tk) : t_last;
            if (PL && plastic && valid) {
                hh[0] = ((tk - last) / kEvHistUnit) & 0xFFFFu;
                if (spiked) {
#pragma unroll
                    for (uint32_t h = kEvHist - 1; h > 0; --h)
                        hh[h] = hh[h - 1];
                    hh[0] = 0;
                    trace += 1.0f;
                }
            }
            s_spk_t[ns][threadIdx.x] = tk;
            if (PL) {
                s_spk_last[ns][threadIdx.x] = last;
                s_spk_tr[ns][threadIdx.x] = trace;
            }
            ns += spiked ? 1u : 0u;
        }
        k0 = k1;
        r1 = r2;
        t_end = t_next;
    }
    if (lo != hi) {
        a.V[j] = V;
        a.t_last[j] = t_last;
        if (PL && plastic) {
            a.trace[j] = trace;
#pragma unroll
            for (uint32_t h = 0; h < kEvHist / 2; ++h)
                a.hist[(size_t)j * (kEvHist / 2) + h] = hh[2 * h] | hh[2 * h + 1] << 16;
        }
    }
    // the buffered spikes: one reservation per block (one wave) in the next
    // window's input list and in the record
    n_spikes += ns;
    uint32_t before = ns;  // inclusive prefix over the wave's lanes
    for (uint32_t off = 1; off < kWave; off <<= 1) {
        const uint32_t y = (uint32_t)__shfl_up((int)before, off);
        if (threadIdx.x >= off)
            before += y;
    }
    const uint32_t total = (uint32_t)__shfl((int)before, kWave - 1);
    before -= ns;
    unsigned long long all_spikes = n_spikes;
    for (uint32_t off = kWave / 2; off; off >>= 1)
        all_spikes += __shfl_xor(all_spikes, off);
    if (threadIdx.x == 0) {
        if (total) {
            s_base[0] = atomicAdd(a.n_out, total);
            s_base[1] = a.rec_cap ? atomicAdd(a.rec_n, (unsigned long long)total) : 0ull;
        }
        if (all_spikes)
            atomicAdd(a.spikes, all_spikes);
    }
    __syncthreads();
    for (uint32_t q = 0; q < ns; ++q) {
        const uint32_t t = s_spk_t[q][threadIdx.x];
        const uint32_t i = before + q;
        const unsigned long long k = s_base[0] + i;
        if (k < a.in_cap) {
            a.out_keys[k] = (t - a.t0) << kEvSrcBits | j;  // (t + 1500) - (t0 + 1500)
            if (a.out_vals) {
                a.out_vals[k] = (uint32_t)k;
                a.out_prev[k] = PL ? s_spk_last[q][threadIdx.x] : 0u;
                a.out_trace[k] = PL ? s_spk_tr[q][threadIdx.x] : 0.0f;
            }
        } else {
            atomicAdd(a.lost, 1u);
        }
        if (a.rec_cap) {
            const unsigned long long r = s_base[1] + i;
            if (r < a.rec_cap) {
                naer_spike sp;
                sp.step = a.epoch;
                sp.neuron = j;
                sp.time_us = t;
                sp._pad = 0;
                a.rec[r] = sp;
            }
        }
    }
}

}  // namespace

hipError_t ev_launch(const EvArgs& a, int which, uint32_t n, hipStream_t s)
{
    const uint32_t gn = (a.N + 255) / 256;
    switch (which) {
    case kEvPoisson:
        hipLaunchKernelGGL(k_ev_poisson, dim3(gn), dim3(256), 0, s, a);
        break;
    case kEvRowLen:
        hipLaunchKernelGGL(k_ev_rowlen, dim3(n / 256 + 1), dim3(256), 0, s, a, n);
        break;
    case kEvExpand:
        if (n && a.ev_syn)
            hipLaunchKernelGGL(k_ev_expand<true>, dim3(n), dim3(256), 0, s, a);
        else if (n)
            hipLaunchKernelGGL(k_ev_expand<false>, dim3(n), dim3(256), 0, s, a);
        break;
    case kEvProcess:
        if (a.n_ev)
            hipLaunchKernelGGL(k_ev_bounds, dim3((a.n_ev + 1023) / 1024), dim3(256), 0, s, a);
        if (a.ev_syn)
            hipLaunchKernelGGL(k_ev_process<true>, dim3((a.N + kEvStepThreads - 1) / kEvStepThreads),
                               dim3(kEvStepThreads), 0, s, a);
        else
            hipLaunchKernelGGL(k_ev_process<false>, dim3((a.N + kEvStepThreads - 1) / kEvStepThreads),
                               dim3(kEvStepThreads), 0, s, a);
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

namespace {

// ---------------------------------------------------------------------------
// LSD radix sort of 32-bit keys (optionally with 32-bit values), stable.  A
// pass sorts on `bits` bits (at most log2(DMAX)) at `shift`, over tiles of
// kRsTile items (4 waves x 16 rounds of 64 items; wave w holds the tile's
// w-th quarter, so the order (wave, round, lane) is the item order):
//   rs_count    per tile, the count of every digit (LDS), stored digit-major
//               cnt[d * tiles + t]
//   (exclusive scan of cnt as one flat array: cnt[d * tiles + t] becomes the
//    output position of tile t's first item of digit d)
//   rs_scatter  ranks every item within its wave (lanes of equal digit found
//               with one ballot per digit bit; per-wave digit counters in
//               LDS), orders the tile by digit in LDS, then writes it out in
//               that order, so consecutive lanes store consecutive positions
//               within each digit's run
// ---------------------------------------------------------------------------
constexpr uint32_t kRsThreads = 256, kRsWaves = kRsThreads / kWave, kRsRounds = 16;
constexpr uint32_t kRsTile = kRsThreads * kRsRounds;  // 4096 items
constexpr uint32_t kRsSmallDigits = 256, kRsLargeDigits = 2048;
// 11-bit digits (fewer passes) while the counts stay small: up to 2^20 items
constexpr uint32_t kRsLargeDigitsMaxItems = 1u << 20;
constexpr uint32_t kRsMaxVals = 3;

template <uint32_t DMAX>
__global__ __launch_bounds__(kRsThreads) void rs_count(const uint32_t* keys, uint32_t n, uint32_t shift,
                                                      uint32_t bits, uint32_t tiles, uint32_t* cnt)
{
    __shared__ uint32_t h[DMAX];
    const uint32_t digits = 1u << bits;
    for (uint32_t d = threadIdx.x; d < digits; d += kRsThreads)
        h[d] = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * kRsTile;
    // all loads in flight before the counting
    uint32_t key[kRsRounds];
#pragma unroll
    for (uint32_t k = 0; k < kRsRounds; ++k) {
        const uint32_t i = base + k * kRsThreads + threadIdx.x;
        key[k] = i < n ? keys[i] : 0u;
    }
#pragma unroll
    for (uint32_t k = 0; k < kRsRounds; ++k)
        if (base + k * kRsThreads + threadIdx.x < n)
            atomicAdd(&h[(key[k] >> shift) & (digits - 1u)], 1u);
    __syncthreads();
    for (uint32_t d = threadIdx.x; d < digits; d += kRsThreads)
        cnt[(size_t)d * tiles + blockIdx.x] = h[d];
}

// Exclusive prefix over the block's DMAX digit slots of v (each thread owns
// DMAX / kRsThreads consecutive slots); s is kRsThreads words of scratch.
template <uint32_t DMAX>
__device__ __forceinline__ void rs_block_digit_scan(uint32_t* v, uint32_t* s)
{
    constexpr uint32_t per = DMAX / kRsThreads;
    uint32_t sum = 0;
    for (uint32_t k = 0; k < per; ++k)
        sum += v[threadIdx.x * per + k];
    // wave-inclusive scan of the per-thread sums, then the waves' totals
    const uint32_t lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
    uint32_t x = sum;
    for (uint32_t off = 1; off < kWave; off <<= 1) {
        const uint32_t y = (uint32_t)__shfl_up((int)x, off);
        if (lane >= off)
            x += y;
    }
    if (lane == kWave - 1)
        s[w] = x;
    __syncthreads();
    uint32_t run = x - sum;
    for (uint32_t i = 0; i < w; ++i)
        run += s[i];
    for (uint32_t k = 0; k < per; ++k) {
        const uint32_t c = v[threadIdx.x * per + k];
        v[threadIdx.x * per + k] = run;
        run += c;
    }
    __syncthreads();
}

// the value arrays a sort moves along with its keys
struct RsVals {
    const uint32_t* in[kRsMaxVals];
    uint32_t* out[kRsMaxVals];
};

template <uint32_t DMAX, uint32_t NV>
__global__ __launch_bounds__(kRsThreads) void rs_scatter(const uint32_t* kin, uint32_t* kout, const RsVals vals,
                                                        uint32_t n, uint32_t shift, uint32_t bits, uint32_t tiles,
                                                        const uint32_t* cnt)
{
    __shared__ uint32_t wc[kRsWaves][DMAX];  // per-wave digit counts, then the wave's offset within the digit
    __shared__ uint32_t ls[DMAX];            // the tile's digit counts, then their exclusive prefix
    __shared__ uint32_t gs[DMAX];            // output position of the tile's first item of each digit
    __shared__ uint32_t sk[kRsTile];
    __shared__ uint32_t sv[NV ? NV : 1][NV ? kRsTile : 1];
    __shared__ uint32_t s_scan[kRsWaves];
    const uint32_t digits = 1u << bits;
    const uint32_t lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
    for (uint32_t d = threadIdx.x; d < DMAX; d += kRsThreads) {
        for (uint32_t q = 0; q < kRsWaves; ++q)
            wc[q][d] = 0;
        gs[d] = d < digits ? cnt[(size_t)d * tiles + blockIdx.x] : 0u;
    }
    const uint32_t base = blockIdx.x * kRsTile;
    const uint32_t wbase = base + w * (kRsTile / kRsWaves);
    uint32_t key[kRsRounds], rank[kRsRounds];
    uint32_t val[NV ? NV : 1][kRsRounds];
    for (uint32_t r = 0; r < kRsRounds; ++r) {
        const uint32_t i = wbase + r * kWave + lane;
        key[r] = i < n ? kin[i] : 0u;
        for (uint32_t v = 0; v < NV; ++v)
            val[v][r] = i < n ? vals.in[v][i] : 0u;
    }
    __syncthreads();
    const unsigned long long below = (1ull << lane) - 1ull;
    for (uint32_t r = 0; r < kRsRounds; ++r) {
        const bool valid = wbase + r * kWave + lane < n;
        const uint32_t d = (key[r] >> shift) & (digits - 1u);
        unsigned long long same = __ballot(valid);
        for (uint32_t b = 0; b < bits; ++b) {
            const unsigned long long set = __ballot((d >> b) & 1u);
            same &= ((d >> b) & 1u) ? set : ~set;
        }
        const uint32_t leader = valid ? (uint32_t)__builtin_ctzll(same) : 0u;
        uint32_t at = 0;
        if (valid && lane == leader) {
            at = wc[w][d];
            wc[w][d] = at + (uint32_t)__popcll(same);
        }
        at = (uint32_t)__shfl((int)at, (int)leader);
        rank[r] = at + (uint32_t)__popcll(same & below);
    }
    __syncthreads();
    for (uint32_t d = threadIdx.x; d < DMAX; d += kRsThreads) {
        uint32_t run = 0;
        for (uint32_t q = 0; q < kRsWaves; ++q) {
            const uint32_t c = wc[q][d];
            wc[q][d] = run;
            run += c;
        }
        ls[d] = run;
    }
    __syncthreads();
    rs_block_digit_scan<DMAX>(ls, s_scan);
    for (uint32_t r = 0; r < kRsRounds; ++r) {
        if (wbase + r * kWave + lane < n) {
            const uint32_t d = (key[r] >> shift) & (digits - 1u);
            const uint32_t lp = ls[d] + wc[w][d] + rank[r];
            sk[lp] = key[r];
            for (uint32_t v = 0; v < NV; ++v)
                sv[v][lp] = val[v][r];
        }
    }
    __syncthreads();
    const uint32_t items = min(kRsTile, n - base);
    for (uint32_t k = threadIdx.x; k < items; k += kRsThreads) {
        const uint32_t x = sk[k];
        const uint32_t d = (x >> shift) & (digits - 1u);
        const uint32_t pos = gs[d] + (k - ls[d]);
        kout[pos] = x;
        for (uint32_t v = 0; v < NV; ++v)
            vals.out[v][pos] = sv[v][k];
    }
}

// ---------------------------------------------------------------------------
// Exclusive scan of n values (32- or 64-bit) in chunks of kScanChunk: chunk
// sums, their exclusive prefix (one block), then every chunk's own prefix.
// A chunk is staged in LDS so that the loads and stores coalesce.
// ---------------------------------------------------------------------------
constexpr uint32_t kScanThreads = 256, kScanPer = 16, kScanChunk = kScanThreads * kScanPer;

// exclusive prefix of x over the block; *total: the block's sum
template <typename T>
__device__ __forceinline__ T block_excl_scan(T x, T* s_wave, T* total)
{
    const uint32_t lane = threadIdx.x % kWave, w = threadIdx.x / kWave, waves = blockDim.x / kWave;
    T v = x;
    for (uint32_t off = 1; off < kWave; off <<= 1) {
        const T y = __shfl_up(v, off);
        if (lane >= off)
            v += y;
    }
    if (lane == kWave - 1)
        s_wave[w] = v;
    __syncthreads();
    T before = 0, sum = 0;
    for (uint32_t i = 0; i < waves; ++i) {
        before += i < w ? s_wave[i] : T(0);
        sum += s_wave[i];
    }
    *total = sum;
    __syncthreads();
    return before + v - x;
}

__device__ __forceinline__ uint32_t scan_slot(uint32_t i) { return i + (i >> 5); }  // padded against bank conflicts

template <typename T>
__global__ __launch_bounds__(kScanThreads) void sc_partials(const T* in, uint32_t n, T* part)
{
    __shared__ T s_wave[kScanThreads / kWave];
    const uint32_t base = blockIdx.x * kScanChunk;
    T sum = 0;
    for (uint32_t k = 0; k < kScanPer; ++k) {
        const uint32_t i = base + k * kScanThreads + threadIdx.x;
        sum += i < n ? in[i] : T(0);
    }
    T total;
    (void)block_excl_scan<T>(sum, s_wave, &total);
    if (threadIdx.x == 0)
        part[blockIdx.x] = total;
}

template <typename T>
__global__ __launch_bounds__(1024) void sc_partial_prefix(T* part, uint32_t chunks)
{
    __shared__ T s_wave[1024 / kWave];
    T carry = 0;
    for (uint32_t c0 = 0; c0 < chunks; c0 += 1024) {
        const uint32_t c = c0 + threadIdx.x;
        const T x = c < chunks ? part[c] : T(0);
        T total;
        const T ex = block_excl_scan<T>(x, s_wave, &total);
        if (c < chunks)
            part[c] = carry + ex;
        carry += total;
    }
}

template <typename T>
__global__ __launch_bounds__(kScanThreads) void sc_final(const T* in, uint32_t n, const T* part, T* out)
{
    __shared__ T s[kScanChunk + kScanChunk / 32];
    __shared__ T s_wave[kScanThreads / kWave];
    const uint32_t base = blockIdx.x * kScanChunk;
    for (uint32_t k = 0; k < kScanPer; ++k) {
        const uint32_t j = k * kScanThreads + threadIdx.x;
        s[scan_slot(j)] = base + j < n ? in[base + j] : T(0);
    }
    __syncthreads();
    T v[kScanPer];
    T sum = 0;
    for (uint32_t k = 0; k < kScanPer; ++k) {
        v[k] = s[scan_slot(threadIdx.x * kScanPer + k)];
        sum += v[k];
    }
    T total;
    T run = part[blockIdx.x] + block_excl_scan<T>(sum, s_wave, &total);
    for (uint32_t k = 0; k < kScanPer; ++k) {
        s[scan_slot(threadIdx.x * kScanPer + k)] = run;
        run += v[k];
    }
    __syncthreads();
    for (uint32_t k = 0; k < kScanPer; ++k) {
        const uint32_t j = k * kScanThreads + threadIdx.x;
        if (base + j < n)
            out[base + j] = s[scan_slot(j)];
    }
}

template <typename T>
hipError_t scan_launch(const T* in, T* out, uint32_t n, T* part, hipStream_t s)
{
    if (n == 0)
        return hipSuccess;
    const uint32_t chunks = (n + kScanChunk - 1) / kScanChunk;
    hipLaunchKernelGGL(sc_partials<T>, dim3(chunks), dim3(kScanThreads), 0, s, in, n, part);
    hipLaunchKernelGGL(sc_partial_prefix<T>, dim3(1), dim3(1024), 0, s, part, chunks);
    hipLaunchKernelGGL(sc_final<T>, dim3(chunks), dim3(kScanThreads), 0, s, in, n, part, out);
    return hipGetLastError();
}

uint64_t rs_tiles(uint64_t n) { return (n + kRsTile - 1) / kRsTile; }

// counts of a pass over n items and the partial sums of their scan
uint64_t rs_count_words(uint64_t n)
{
    const uint64_t c = std::max(rs_tiles(n) * kRsSmallDigits,
                                rs_tiles(std::min<uint64_t>(n, kRsLargeDigitsMaxItems)) * kRsLargeDigits);
    return c + (c + kScanChunk - 1) / kScanChunk + 1;
}

}  // namespace

uint64_t ev_sort_space(uint32_t n_max) { return rs_count_words(n_max); }

uint64_t ev_scan_space(uint32_t n_max) { return ((uint64_t)n_max + kScanChunk - 1) / kScanChunk + 1; }

template <uint32_t DMAX>
void rs_scatter_launch(uint32_t nv, const uint32_t* src, uint32_t* dst, const RsVals& v, uint32_t n, uint32_t shift,
                       uint32_t bits, uint32_t tiles, const uint32_t* cnt, hipStream_t s)
{
    switch (nv) {
    case 0:
        hipLaunchKernelGGL((rs_scatter<DMAX, 0>), dim3(tiles), dim3(kRsThreads), 0, s, src, dst, v, n, shift, bits,
                           tiles, cnt);
        break;
    case 1:
        hipLaunchKernelGGL((rs_scatter<DMAX, 1>), dim3(tiles), dim3(kRsThreads), 0, s, src, dst, v, n, shift, bits,
                           tiles, cnt);
        break;
    default:
        hipLaunchKernelGGL((rs_scatter<DMAX, 3>), dim3(tiles), dim3(kRsThreads), 0, s, src, dst, v, n, shift, bits,
                           tiles, cnt);
        break;
    }
}

hipError_t ev_radix_sort(uint32_t* keys, uint32_t* keys_alt, uint32_t* const* vals, uint32_t* const* vals_alt,
                         uint32_t nv, uint32_t n, uint32_t lo_bit, uint32_t hi_bit, uint32_t* space, bool* in_alt,
                         hipStream_t s)
{
    *in_alt = false;
    if (nv == 2 || nv > kRsMaxVals)
        return hipErrorInvalidValue;
    if (n == 0 || hi_bit <= lo_bit)
        return hipSuccess;
    const uint32_t tiles = (uint32_t)rs_tiles(n);
    const bool large = n <= kRsLargeDigitsMaxItems;
    const uint32_t max_bits = large ? 11 : 8;
    // equal passes: e.g. 18 bits as 6 + 6 + 6 rather than 8 + 8 + 2
    const uint32_t total = hi_bit - lo_bit;
    const uint32_t passes = (total + max_bits - 1) / max_bits;
    const uint32_t per = (total + passes - 1) / passes;
    uint32_t* src = keys;
    uint32_t* dst = keys_alt;
    RsVals v{};
    for (uint32_t q = 0; q < nv; ++q) {
        v.in[q] = vals[q];
        v.out[q] = vals_alt[q];
    }
    for (uint32_t shift = lo_bit; shift < hi_bit; shift += per) {
        const uint32_t bits = std::min(per, hi_bit - shift);
        const uint32_t n_cnt = (1u << bits) * tiles;
        uint32_t* cnt = space;
        uint32_t* part = space + n_cnt;
        if (large)
            hipLaunchKernelGGL(rs_count<kRsLargeDigits>, dim3(tiles), dim3(kRsThreads), 0, s, src, n, shift, bits,
                               tiles, cnt);
        else
            hipLaunchKernelGGL(rs_count<kRsSmallDigits>, dim3(tiles), dim3(kRsThreads), 0, s, src, n, shift, bits,
                               tiles, cnt);
        hipError_t e = scan_launch<uint32_t>(cnt, cnt, n_cnt, part, s);
        if (e != hipSuccess)
            return e;
        if (large)
            rs_scatter_launch<kRsLargeDigits>(nv, src, dst, v, n, shift, bits, tiles, cnt, s);
        else
            rs_scatter_launch<kRsSmallDigits>(nv, src, dst, v, n, shift, bits, tiles, cnt, s);
        std::swap(src, dst);
        for (uint32_t q = 0; q < nv; ++q) {
            uint32_t* const was_in = const_cast<uint32_t*>(v.in[q]);
            v.in[q] = v.out[q];
            v.out[q] = was_in;
        }
        *in_alt = !*in_alt;
    }
    return hipGetLastError();
}

hipError_t ev_scan_u64(const unsigned long long* in, unsigned long long* out, uint32_t n, unsigned long long* space,
                       hipStream_t s)
{
    return scan_launch<unsigned long long>(in, out, n, space, s);
}

}  // namespace naer
