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
// The sorts and the scan are rocPRIM's (through hipCUB).
#include <algorithm>

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include "naer_device.hpp"
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

__device__ __forceinline__ void ev_append(const EvArgs& a, uint32_t j, uint32_t time)
{
    const uint32_t k = atomicAdd(a.n_out, 1u);
    if (k < a.in_cap)
        a.out_keys[k] = (unsigned long long)time << 32 | j;
    else
        atomicAdd(a.lost, 1u);
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
        ev_append(a, j, time);
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
        const uint32_t s = (uint32_t)a.in_keys[i];
        len = a.row_ptr[s + 1] - a.row_ptr[s];
    }
    a.row_len[i] = len;
}

// One block per input spike walks its row.
__global__ __launch_bounds__(256) void k_ev_expand(const EvArgs a)
{
    const unsigned long long key = a.in_keys[blockIdx.x];
    const uint32_t s = (uint32_t)key;
    const uint32_t dt = (uint32_t)(key >> 32) - a.t0;  // < 1500
    const unsigned long long r0 = a.row_ptr[s], r1 = a.row_ptr[s + 1];
    const unsigned long long e0 = a.row_off[blockIdx.x];
    for (unsigned long long r = r0 + threadIdx.x; r < r1; r += 256) {
        const uint32_t tgt = a.col[r] & 0xFFFFFFu;
        const unsigned long long e = e0 + (r - r0);
        a.ev_keys[e] = tgt << 11 | dt;
        a.ev_w[e] = a.w[r];
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
    if (lo != hi) {
        n_hi = a.n_hi[p];
        t_ref = a.t_ref[p];
        v_rest = a.pops[p].v_rest;
        v_th = a.pops[p].v_th;
        v_reset = a.pops[p].v_reset;
        V = a.V[j];
        t_last = a.t_last[j];
    }
    // Software pipeline over chunks of kEvGroup events: while chunk c is
    // stepped, chunk c + 1's decays are looked up and chunk c + 2's events
    // are loaded.
    auto load = [&](uint32_t c, uint32_t& t, float& w) {
        const uint32_t i = c + gl;
        t = 0;
        w = 0.0f;
        if (i < hi) {
            t = a.t0 + (a.ev_keys[i] & 2047u);
            w = a.ev_w[i];
        }
    };
    // the decay of each event of chunk c from the previous event's time
    // (t_before: the last event of the chunk before)
    auto lookup = [&](uint32_t c, uint32_t t, uint32_t t_before, uint32_t& tp, double& d) {
        tp = (uint32_t)__shfl_up((int)t, 1, kEvGroup);
        if (gl == 0)
            tp = t_before;
        const long long dtp = (long long)t - (long long)tp;
        d = (c + gl < hi && dtp >= 0) ? ev_decay(H, L, n_hi, dtp) : 0.0;
    };
    uint32_t t0c, t1c, tp0;
    float w0c, w1c;
    double d0;
    load(lo, t0c, w0c);
    lookup(lo, t0c, t_last, tp0, d0);
    load(lo + kEvGroup, t1c, w1c);
    uint32_t fired = 0;
    for (uint32_t c = lo;; c += kEvGroup) {
        const bool more = c < hi;
        if (!__any(more))
            break;
        const uint32_t n = more ? min(kEvGroup, hi - c) : 0u;
        // next chunk's decays start from this chunk's last event
        const uint32_t t_end = (uint32_t)__shfl((int)t0c, (int)(n ? n - 1 : 0), kEvGroup);
        uint32_t tp1;
        double d1;
        lookup(c + kEvGroup, t1c, t_end, tp1, d1);
        uint32_t t2c;
        float w2c;
        load(c + 2 * kEvGroup, t2c, w2c);
        for (uint32_t k = 0; k < kEvGroup; ++k) {
            // every lane takes part in the shuffles; the step is predicated
            const uint32_t tk = (uint32_t)__shfl((int)t0c, (int)k, kEvGroup);
            const uint32_t tpk = (uint32_t)__shfl((int)tp0, (int)k, kEvGroup);
            const float wk = __shfl(w0c, (int)k, kEvGroup);
            const double dpk = __shfl(d0, (int)k, kEvGroup);
            if (k < n) {
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
                        if (gl == 0) {
                            ev_append(a, j, tk + kEpochUs);
                            ev_record(a, j, tk);
                        }
                        ++fired;
                    }
                    V = v;
                }
            }
        }
        t0c = t1c;
        w0c = w1c;
        tp0 = tp1;
        d0 = d1;
        t1c = t2c;
        w1c = w2c;
    }
    if (gl == 0 && lo != hi) {
        a.V[j] = V;
        a.t_last[j] = t_last;
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
        if (n)
            hipLaunchKernelGGL(k_ev_expand, dim3(n), dim3(256), 0, s, a);
        break;
    case kEvProcess:
        hipLaunchKernelGGL(k_ev_process, dim3((a.N + 255 / kEvGroup) / (256 / kEvGroup)), dim3(256), 0, s, a);
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// Temporary storage of the sorts and the scan for up to n_in inputs and
// n_ev events (null `temp`: the size only).
hipError_t ev_sort_inputs(void* temp, size_t* bytes, const unsigned long long* in, unsigned long long* out,
                          uint32_t n, hipStream_t s)
{
    // keys: time << 32 | source
    return hipcub::DeviceRadixSort::SortKeys(temp, *bytes, in, out, (int)n, 0, 64, s);
}

hipError_t ev_scan_rows(void* temp, size_t* bytes, const unsigned long long* in, unsigned long long* out, uint32_t n,
                        hipStream_t s)
{
    return hipcub::DeviceScan::ExclusiveSum(temp, *bytes, in, out, (int)n, s);
}

hipError_t ev_sort_events(void* temp, size_t* bytes, const uint32_t* kin, uint32_t* kout, const float* vin,
                          float* vout, uint32_t n, uint32_t n_neurons, hipStream_t s)
{
    // keys: target << 11 | time offset (< 2048), sorted on the target bits
    // only (radix sort is stable, so the expansion order survives)
    int bits = 11;
    for (uint32_t m = n_neurons ? n_neurons - 1 : 0; m; m >>= 1)
        ++bits;
    return hipcub::DeviceRadixSort::SortPairs(temp, *bytes, kin, kout, vin, vout, (int)n, 11, std::max(bits, 12), s);
}

}  // namespace naer
