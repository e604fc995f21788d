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
//   k_ev_process   a group of 8 lanes per LIF neuron finds its range of sorted
//                  events (8-ary search), applies them in order and appends
//                  its spikes to the next window's input list
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

// lif.cl:15-34 per event, in time order; the decay e^(-dt/tau) is H[dt >> 12]
// * L[dt & 4095] from the population's tables.  A target per group of 8
// lanes (8 per wave): the group loads 8 events at a time (coalesced) and
// looks up each event's decay from the previous event's time in parallel; the
// step itself runs in order on values shuffled within the group, and looks
// the decay up again only when the neuron's t_last is not the previous
// event's time (after a spike or a dropped input).
constexpr uint32_t kEvGroup = 8;
static_assert(kEvGroup >= 2 && kEvGroup < 32 && (kEvGroup & (kEvGroup - 1)) == 0,
              "a step group is a power of two of lanes below 32 (its ballot mask is 32-bit)");

// First index of keys[0, n) whose target (key >> 11) is >= tgt (n if none),
// found by the kEvGroup lanes of a group together: each round probes 8 points
// of the remaining range.  Every lane of the wave takes part (ballots).
__device__ __forceinline__ uint32_t group_lower_bound(const uint32_t* keys, uint32_t n, uint32_t tgt, uint32_t gl)
{
    const uint32_t base = (threadIdx.x & 63u) & ~(kEvGroup - 1);
    uint32_t lo = 0, hi = n;  // keys[< lo] are below tgt; hi == n or keys[hi] is not
    for (;;) {
        const bool busy = hi - lo > kEvGroup;
        if (!__any(busy))
            break;
        const unsigned long long span = hi - lo;
        // probes q(g) = lo + (g + 1) span / 9, strictly increasing and < hi
        auto q = [&](uint32_t g) { return lo + (uint32_t)((g + 1) * span / (kEvGroup + 1)); };
        const bool ge = busy && (keys[q(gl)] >> 11) >= tgt;
        const uint32_t bits = (uint32_t)(__ballot(ge) >> base) & ((1u << kEvGroup) - 1);
        if (busy) {
            if (bits == 0) {
                lo = q(kEvGroup - 1) + 1;
            } else {
                const uint32_t f = (uint32_t)__ffs(bits) - 1;
                const uint32_t nlo = f ? q(f - 1) + 1 : lo;
                hi = q(f);
                lo = nlo;
            }
        }
    }
    const uint32_t i = lo + gl;
    const bool ge = i < hi && (keys[i] >> 11) >= tgt;
    const uint32_t bits = (uint32_t)(__ballot(ge) >> base) & ((1u << kEvGroup) - 1);
    return bits ? lo + (uint32_t)__ffs(bits) - 1 : hi;
}

// stdp.cl:8-30 for one plastic event (the reference's update_synapse with
// basic_stdp.cl's rule), in the reference's integer and float types:
//   LTP over the inter-spike history while the post spike time (int) is
//   above the input's previous time (int); each step evaluates
//   exp((float)(uint)(prev - post) / T) (the unsigned difference of a later
//   post time wraps, so the factor overflows to +inf and the increment is 0);
//   the post trace decays from the last spike time (uint difference; a
//   negative one wraps and clears the trace); LTD with alpha * lambda in
//   double; hard bounds +-W_MAX.
__device__ __forceinline__ float ev_stdp_update(float w, const uint32_t (&hh)[kEvHist], uint32_t last,
                                                uint32_t t, uint32_t prev, float pre_trace, float& trace)
{
    int32_t post = (int32_t)last;
#pragma unroll
    for (uint32_t h = 1; h < kEvHist; ++h) {
        if (!(post > (int32_t)prev && hh[h] > 0u))
            break;
        const float pre = naer_expf((float)(prev - (uint32_t)post) / kEvStdpTau) * pre_trace + 1.0f;
        const double inc = kEvStdpLambda * (double)(1.0f - w) * (double)naer_expf(-pre / kEvStdpConstDt);
        w = (float)((double)w + inc);
        post -= (int32_t)(hh[h] * kEvHistUnit);
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
__global__ __launch_bounds__(256) void k_ev_process(const EvArgs a)
{
    const uint32_t gl = threadIdx.x & (kEvGroup - 1);
    const uint32_t j = blockIdx.x * (256 / kEvGroup) + threadIdx.x / kEvGroup;
    uint32_t p = 0;
    // the target's range of sorted events (groups past N search nothing)
    const uint32_t n_ev = j < a.N ? a.n_ev : 0u;
    uint32_t lo = group_lower_bound(a.ev_keys, n_ev, j, gl);
    uint32_t hi = group_lower_bound(a.ev_keys, n_ev, j + 1, gl);
    if (j < a.N) {
        if (lo != hi) {
            p = ev_pop(a, j);
            const PopParams& pp = a.pops[p];
            if (pp.kind != NAER_POP_LIF || j < pp.begin || j >= pp.end)
                hi = lo;
        }
    }
    if (!__any(lo != hi))
        return;
    const double* L = a.dec_lo + (size_t)p * kEvLo;
    const double* H = a.dec_hi + (size_t)p * kEvHiCap;
    uint32_t n_hi = 0, t_ref = 0, t_last = 0;
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
    // Software pipeline over chunks of kEvGroup events: while chunk c is
    // stepped, chunk c + 1's decays (and, plastic, its inputs' payloads) are
    // looked up and chunk c + 2's events are loaded.
    struct Ev {
        uint32_t t;
        float w;
        uint32_t syn, slot;
    };
    auto load = [&](uint32_t c, Ev& e) {
        const uint32_t i = c + gl;
        e.t = 0;
        e.w = 0.0f;
        e.syn = 0;
        e.slot = 0;
        if (i < hi) {
            e.t = a.t0 + (a.ev_keys[i] & 2047u);
            e.w = a.ev_w[i];
            if (PL && plastic) {
                e.syn = a.ev_syn[i];
                e.slot = a.ev_slot[i];
            }
        }
    };
    // the decay of each event of chunk c from the previous event's time
    // (t_before: the last event of the chunk before)
    auto lookup = [&](uint32_t c, const Ev& e, uint32_t t_before, uint32_t& tp, double& d, uint32_t& prev,
                      float& ptr) {
        tp = (uint32_t)__shfl_up((int)e.t, 1, kEvGroup);
        if (gl == 0)
            tp = t_before;
        const long long dtp = (long long)e.t - (long long)tp;
        d = (c + gl < hi && dtp >= 0) ? ev_decay(H, L, n_hi, dtp) : 0.0;
        prev = 0;
        ptr = 0.0f;
        if (PL && c + gl < hi && (e.syn & kPlasticBit)) {
            prev = a.in_prev[e.slot];
            // stdp.cl:84 holds the pre trace in a uint
            ptr = (float)(uint32_t)a.in_trace[e.slot];
        }
    };
    Ev e0, e1;
    uint32_t tp0, prev0 = 0;
    float ptr0 = 0.0f;
    double d0;
    load(lo, e0);
    lookup(lo, e0, t_last, tp0, d0, prev0, ptr0);
    load(lo + kEvGroup, e1);
    uint32_t fired = 0;
    for (uint32_t c = lo;; c += kEvGroup) {
        const bool more = c < hi;
        if (!__any(more))
            break;
        const uint32_t n = more ? min(kEvGroup, hi - c) : 0u;
        // next chunk's decays start from this chunk's last event
        const uint32_t t_end = (uint32_t)__shfl((int)e0.t, (int)(n ? n - 1 : 0), kEvGroup);
        uint32_t tp1, prev1;
        float ptr1;
        double d1;
        lookup(c + kEvGroup, e1, t_end, tp1, d1, prev1, ptr1);
        Ev e2;
        load(c + 2 * kEvGroup, e2);
        for (uint32_t k = 0; k < kEvGroup; ++k) {
            // every lane takes part in the shuffles; the step is predicated
            const uint32_t tk = (uint32_t)__shfl((int)e0.t, (int)k, kEvGroup);
            const uint32_t tpk = (uint32_t)__shfl((int)tp0, (int)k, kEvGroup);
            float wk = __shfl(e0.w, (int)k, kEvGroup);
            const double dpk = __shfl(d0, (int)k, kEvGroup);
            uint32_t synk = 0, prevk = 0;
            float ptrk = 0.0f;
            if (PL) {
                synk = (uint32_t)__shfl((int)e0.syn, (int)k, kEvGroup);
                prevk = (uint32_t)__shfl((int)prev0, (int)k, kEvGroup);
                ptrk = __shfl(ptr0, (int)k, kEvGroup);
            }
            if (k < n) {
                uint32_t last = 0;
                if (PL && plastic) {
                    last = t_last - hh[0] * kEvHistUnit;
                    if (synk & kPlasticBit) {
                        wk = ev_stdp_update(wk, hh, last, tk, prevk, ptrk, trace);
                        if (gl == 0)
                            a.w_mut[synk & ~kPlasticBit] = wk;
                    }
                }
                bool spiked = false;
                const long long dt = (long long)tk - (long long)t_last;
                if (dt >= 0) {
                    const double dk = tpk == t_last ? dpk : ev_decay(H, L, n_hi, dt);
                    float x = V - v_rest;
                    x = (float)((double)x * dk);
                    float v = x + v_rest;
                    v = v + wk;
                    t_last = tk;
                    if (v > v_th) {
                        v = v_reset;
                        t_last = tk + t_ref;
                        spiked = true;
                    }
                    V = v;
                }
                if (PL && plastic) {
                    hh[0] = ((tk - last) / kEvHistUnit) & 0xFFFFu;
                    if (spiked) {
#pragma unroll
                        for (uint32_t h = kEvHist - 1; h > 0; --h)
                            hh[h] = hh[h - 1];
                        hh[0] = 0;
                        trace += 1.0f;
                    }
                }
                if (spiked) {
                    if (gl == 0) {
                        ev_append(a, j, tk + kEpochUs, a.t0 + kEpochUs, PL && plastic, last, trace);
                        ev_record(a, j, tk);
                    }
                    ++fired;
                }
            }
        }
        e0 = e1;
        tp0 = tp1;
        d0 = d1;
        prev0 = prev1;
        ptr0 = ptr1;
        e1 = e2;
    }
    if (gl == 0 && lo != hi) {
        a.V[j] = V;
        a.t_last[j] = t_last;
        if (PL && plastic) {
            a.trace[j] = trace;
#pragma unroll
            for (uint32_t h = 0; h < kEvHist / 2; ++h)
                a.hist[(size_t)j * (kEvHist / 2) + h] = hh[2 * h] | hh[2 * h + 1] << 16;
        }
        if (fired)
            atomicAdd(a.spikes, (unsigned long long)fired);
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
        if (a.ev_syn)
            hipLaunchKernelGGL(k_ev_process<true>, dim3((a.N + 255 / kEvGroup) / (256 / kEvGroup)), dim3(256), 0, s,
                               a);
        else
            hipLaunchKernelGGL(k_ev_process<false>, dim3((a.N + 255 / kEvGroup) / (256 / kEvGroup)), dim3(256), 0, s,
                               a);
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
    for (uint32_t k = 0; k < kRsRounds; ++k) {
        const uint32_t i = base + k * kRsThreads + threadIdx.x;
        if (i < n)
            atomicAdd(&h[(keys[i] >> shift) & (digits - 1u)], 1u);
    }
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
