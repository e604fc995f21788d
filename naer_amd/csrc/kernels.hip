// kernels.hip — gfx950 kernels of the simulation step and the connectivity
// build.  Host launch wrappers at the bottom (declared in naer_kernels.hpp).
//
// Reference kernels these replace (include/naer/kernels/ of the reference):
//   k_update     <- simulate_nop_neurons (nop.cl:7-90), simulate_stdp_neurons'
//                   neuron part (stdp.cl:32-140), neuron_dynamics (lif.cl:15-34)
//                   and simulate_poisson_neurons (input_neuron_group.cl:6-45);
//                   its ballot compaction replaces histogram_scan/scatter
//                   (spike_aggregation.cl:3-65): with a fixed step there is no
//                   time sort, only a per-step spike list.
//   k_scatter    <- the per-spike synapse segment walk of nop.cl:51-82 /
//                   stdp.cl:86-105 plus distribute_offsets
//                   (spike_aggregation.cl:20-32): the CSR row of a source is
//                   row_ptr[s]..row_ptr[s+1], no per-group offset table.
//   lazy_stdp / k_flush <- the lazy LTP of update_synapse (stdp.cl:19-26):
//                   potentiation is applied when the pre-synaptic neuron next
//                   fires, from each post neuron's last spikes (bounded by a
//                   flush instead of a truncated history).
//   k_bucket_*   <- the delivery of a batch of steps: events sorted by arrival
//                   batch and target tile, summed in the update's LDS.
//   k_conn_*     <- count_connections / integrate_connections /
//                   generate_connections (network.cl:14-107).
// k_update also stages the input of populations of source-defined models
// (NeuronType.h:14), whose own kernels naer_model.hip compiles; the
// reference's event-driven step itself is in kernels_event.hip.
#include <mutex>

#include "naer_device.hpp"
#include "naer_kernels.hpp"

namespace naer {

namespace {

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

__device__ __forceinline__ long long to_fixed(float w)
{
    return __double2ll_rn((double)w * 4294967296.0);
}

// value of the first active lane (wave-uniform values into scalar registers)
__device__ __forceinline__ uint32_t rfl(uint32_t x) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)x); }
__device__ __forceinline__ float rflf(float x) { return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(x))); }

constexpr uint32_t kCdfHead = 8;  // drive table entries compared in registers

// (double)a * 2^-32 as hi + lo * 2^-32: one rounding of the same exact value
// (scaling by 2^-32 commutes with it), so the same double and float, with one
// double-precision instruction fewer than converting the 64-bit integer
__device__ __forceinline__ float from_fixed(long long a)
{
    return (float)__fma_rn((double)(uint32_t)a, 0x1p-32, (double)(int32_t)(a >> 32));
}

__device__ __forceinline__ float clip(float w, float lo, float hi)
{
    return w < lo ? lo : (w > hi ? hi : w);
}

// Adds 1 to cnt[key] for every active lane and returns the lane's slot (the
// counter's old value plus the lane's rank among the lanes of that key).
// AGG: one LDS atomic per run of equal keys on consecutive lanes (the active
// lanes must be a prefix of the wave): sorted CSR rows of a small network put
// whole waves on one tile, and lane-by-lane atomics on one LDS word
// serialise.  Without AGG (rows spread over many tiles) a plain atomic per
// lane keeps the unrolled atomics of a task in flight together.
// More run heads per wave than HEADS: lane-by-lane atomics.  The histogram
// and refine counters (tile keys) aggregate only long runs; the placement's
// group keys come in runs of a few synapses (rows are ordered by group and
// delay), where one atomic per run still beats one per lane.
constexpr uint32_t kRunHeads = 16;
constexpr uint32_t kPlaceRunHeads = 40;

template <bool POS, bool AGG, uint32_t HEADS = kRunHeads>
__device__ __forceinline__ uint32_t run_add(uint32_t* cnt, uint32_t key)
{
    if (!AGG)
        return atomicAdd(&cnt[key], 1u);
    const uint32_t lane = lane_id();
    // previous lane's key (DPP wave_shr:1; lane 0 is a head regardless)
    const uint32_t prev = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)key, 0x138, 0xF, 0xF, false);
    const unsigned long long heads = __ballot(lane == 0 || prev != key);
    if (__popcll(heads) > HEADS)
        return atomicAdd(&cnt[key], 1u);
    const unsigned long long act = __ballot(1);
    const unsigned long long upto = (2ull << lane) - 1ull;  // lanes <= lane (all for lane 63)
    const unsigned long long above = heads & ~upto;
    const uint32_t next = above ? (uint32_t)__builtin_ctzll(above) : 64u - (uint32_t)__builtin_clzll(act);
    uint32_t old = 0;
    if ((heads >> lane) & 1ull)
        old = atomicAdd(&cnt[key], next - lane);
    if (!POS)
        return 0;
    const uint32_t hl = 63u - (uint32_t)__builtin_clzll(heads & upto);  // lane 0 is always a head
    return (uint32_t)__shfl((int)old, (int)hl) + (lane - hl);
}

// dec^k of a trace from its table (k >= the table: below 2^-40, taken as 0)
__device__ __forceinline__ float dec_pow(const float* tab, uint32_t n, uint32_t k)
{
    return k < n ? tab[k] : 0.0f;
}

// The traces' tables as a kernel reads them: the first n_l entries of each
// from an LDS copy (the placement kernel stages them), the rest from memory.
// The LDS pointers carry their address space, so the lookups are ds_read
// (a generic pointer makes them flat loads, which also wait on the
// outstanding global loads).
typedef const __attribute__((address_space(3))) float LdsFloat;
typedef __attribute__((address_space(3))) void LdsSink;

struct TraceTabs {
    LdsFloat *tp_l, *tm_l;
    uint32_t n_l;
};

__device__ __forceinline__ float dec_pow2(LdsFloat* lds, uint32_t n_l, const float* tab, uint32_t n, uint32_t k)
{
    return k < n_l ? lds[k] : (k < n ? tab[k] : 0.0f);
}

constexpr uint32_t kLdsTrace = 6144;  // table entries staged per trace (tau up to ~22 dt)

// Lazy pair STDP of one plastic synapse i -> j (pair_stdp in the oracle is
// the eager statement).  Applies, in time order, the potentiations of j's
// spikes with steps in [L, T) -- pre trace from i's spike at tp with trace
// a_s = xp (has_p), or 0 -- and with DEP the depression of i's spike at T,
// post trace of j at T.  j has at most one spike at or after T (one spike
// per batch), so its last spike before T is s1 or s0 of pj = pst[2 j]; s2, s3
// come from pb = pst[2 j + 1] (both loaded by the caller, one line), older
// spikes from the ring, and only when s3 itself is in [L, T).
template <bool DEP>
__device__ __forceinline__ float lazy_stdp(const StdpArgs& s, const TraceTabs& tt, uint32_t j, const uint4& pj,
                                           const uint4& pb, float w, int32_t T, int32_t L, bool has_p, int32_t tp,
                                           float xp)
{
    auto potentiate = [&](int32_t t2) {
        float x = 0.0f;
        if (has_p) {
            if (t2 == tp) {
                x = xp;
            } else {
                const float a = xp + 1.0f;
                x = a * dec_pow2(tt.tp_l, tt.n_l, s.tp, s.n_tp, (uint32_t)(t2 - tp));
            }
        }
        const float dw = s.a_plus * x;
        w = clip(w + dw, s.w_min, s.w_max);
    };
    const bool has1 = pj.x != kNoSpike, has0 = pj.z != kNoSpike;
    const int32_t s1 = (int32_t)pj.x, s0 = (int32_t)pj.z;
    const bool pend1 = has1 && s1 >= T;  // j's newest spike is not before T
    if (has0 && s0 >= L) {
        // spikes older than s0 may lie in the window too: s2, s3 from the
        // record's second half (same line), older ones from the ring; in order
        const uint4& b = pb;
        const uint32_t n = b.z;  // >= 2
        if (n >= 3 && (int32_t)b.x >= L) {
            if (n >= 4 && (int32_t)b.y >= L) {
                const uint32_t* hj = s.hist + (size_t)j * kSpikeHist;
                uint32_t lo = n - 4;  // s3
                while (lo > 0 && lo + kSpikeHist > n && (int32_t)hj[(lo - 1) & (kSpikeHist - 1)] >= L)
                    --lo;
                for (uint32_t k = lo; k + 4 < n; ++k)
                    potentiate((int32_t)hj[k & (kSpikeHist - 1)]);
                potentiate((int32_t)b.y);
            }
            potentiate((int32_t)b.x);
        }
        potentiate(s0);
    }
    if (has1 && !pend1 && s1 >= L)
        potentiate(s1);
    if (DEP) {
        // j's last spike before T
        const bool has = pend1 ? has0 : has1;
        float y = 0.0f;
        if (has) {
            const int32_t sy = pend1 ? s0 : s1;
            const float a = __uint_as_float(pend1 ? pj.w : pj.y) + 1.0f;
            y = a * dec_pow2(tt.tm_l, tt.n_l, s.tm, s.n_tm, (uint32_t)(T - sy));
        }
        const float dw = s.a_minus * y;
        w = clip(w - dw, s.w_min, s.w_max);
    }
    return w;
}

// The pre-synaptic side of a spike of i at step t1, right after its batch's
// update (lat[i] is that spike): potentiation is pending for post spikes from
// L = max(i's previous spike, the last flush of row i) on.
struct PreSide {
    int32_t t1, L, tp;
    float xp;
    bool has_p;
};

__device__ __forceinline__ PreSide pre_side(const StdpArgs& s, uint32_t i, uint32_t t1)
{
    PreSide p;
    const uint4 li = s.lat[i];
    const uint32_t s0 = s.pst[2 * (size_t)i].z;
    const int32_t f = (int32_t)s.flushed[i];
    p.t1 = (int32_t)t1;
    p.has_p = li.x >= 2;
    p.tp = (int32_t)s0;
    p.xp = __uint_as_float(li.w);
    p.L = p.has_p ? max(p.tp, f) : f;
    return p;
}

// Population of neuron j: populations are contiguous ranges; the host gives
// the first population of every update block.
__device__ __forceinline__ uint32_t find_pop(const PopParams* pops, uint32_t n_pops, uint32_t first,
                                             uint32_t j)
{
    uint32_t p = first;
    while (p + 1 < n_pops && j >= pops[p + 1].begin)
        ++p;
    return p;
}

// Block-wide spike staging: wave ballots reserve slots of the block's list
// segment through an LDS counter; the first kLdsSpikes entries are staged in
// LDS and copied out coalesced at the end, the rest go straight to the
// segment.  No global atomics.
struct Stage {
    naer_spike* list;  // LDS
    uint32_t* n;       // LDS counter
    naer_spike* seg;   // this block's global list segment
    uint32_t seg_cap;
};

// Ballot compaction of the lanes with `fire` set.
__device__ __forceinline__ void emit(Stage& sg, bool fire, uint32_t neuron, uint32_t step,
                                     uint32_t time_us)
{
    const unsigned long long mask = __ballot(fire);
    if (mask == 0)
        return;
    const uint32_t leader = __ffsll((long long)mask) - 1;
    uint32_t base = 0;
    if (lane_id() == leader)
        base = atomicAdd(sg.n, (uint32_t)__popcll(mask));
    base = __builtin_amdgcn_readlane(base, leader);
    if (fire) {
        const uint32_t lt = __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                                      __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
        const uint32_t slot = base + lt;
        naer_spike sp;
        sp.step = step;
        sp.neuron = neuron;
        sp.time_us = time_us;
        sp._pad = 0;
        if (slot < kLdsSpikes)
            sg.list[slot] = sp;
        else if (slot < sg.seg_cap)
            sg.seg[slot] = sp;
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// Update: one block per tile of TW neurons, one thread per neuron, B
// consecutive steps per launch.  Synaptic input of a step = the neuron's
// delay-ring slot (events delivered by atomics) + the events of the tile's
// buckets (events sorted by the bucketed delivery, summed here in LDS).  Ring
// slots and buckets of all B steps are read up front: nothing can arrive
// during the batch because B <= the minimum delay.
// ---------------------------------------------------------------------------
// EXTRA: the rarely used population kinds (POISSON sources, source-defined
// models) are compiled in; nets of LIF populations only run without them.
// REC4: the bucketed events are 4-byte class records (kRec4*).
// Workgroups are dealt to the 8 XCDs round-robin (block b to XCD b % 8);
// xcd_block gives XCD x the x-th contiguous share of the logical blocks, so
// that neighbouring tiles and segments are handled under one L2: with the
// row prefetch, the update of a tile and the delivery blocks that walk its
// spikes' rows (which the update loads into the caches) share an XCD, and
// the rows of one network,
// whose targets' STDP records share lines, are walked under one L2.  A
// bijection of [0, nb).
constexpr uint32_t kXcds = 8;

__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t nb)
{
    const uint32_t x = b % kXcds, i = b / kXcds;
    const uint32_t q = nb / kXcds, r = nb % kXcds;
    return x * q + min(x, r) + i;
}

template <bool DET, bool PLASTIC, bool BUCKET, bool EXTRA, uint32_t TW, bool REC4>
__global__ __launch_bounds__(TW, (BUCKET && !PLASTIC && !EXTRA && TW == 1024) ? 8 : 1) void k_update(const UpdateArgs a)
{
    static_assert(!REC4 || (BUCKET && !PLASTIC), "class records: bucketed static delivery only");
    using RingT = typename std::conditional<DET, long long, float>::type;
    // bucketed input accumulates in 32.32 fixed point in both modes: integer
    // LDS atomics run at several times the rate of fp32 ones on gfx950
    using AccT = unsigned long long;
    extern __shared__ unsigned long long s_dyn[];
    AccT* acc = (AccT*)s_dyn;  // [B][TW] when BUCKET
    __shared__ naer_spike s_list[kLdsSpikes];
    __shared__ float s_cdf[kStagePops * NAER_DRIVE_TABLE];
    __shared__ uint32_t s_n;
    __shared__ unsigned long long s_events;
    __shared__ unsigned long long s_base;
    // the block's tile: with the row prefetch, XCD-contiguous ranges of tiles
    // (an XCD's update blocks load the rows its placement blocks walk); else
    // block b (packs: XCD-contiguous ranges put a few networks' tiles on each
    // XCD, measured 13 % slower on the sweep)
    const uint32_t tile = a.prefetch ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x;
    const uint32_t bp = a.n_pops ? a.blk_pop[tile] : kTileIdle;
    if (bp & kTileIdle) {
        // no population on this tile (an arena's free range): no neuron to
        // update and no spike this batch; block 0 still does the batch's
        // bookkeeping (as below)
        if (threadIdx.x == 0) {
            DevState* st = a.st;
            const unsigned long long t0 = st->clock;
            const uint32_t pl = (uint32_t)((t0 / a.Bfull) & 1u);
            if (!a.defer || t0 % a.Bfull == 0) {
                a.seg_count[pl * a.n_seg + tile] = 0;
                a.seg_ev[pl * a.n_seg + tile] = 0;
            }
            if (blockIdx.x == 0) {
                st->cur = pl;
                st->t0_cur = t0;
                if (PLASTIC && t0 + a.B > st->t_flush + (st->window >> 1))
                    st->t_flush = t0;
            }
        }
        return;
    }
    if (threadIdx.x == 0) {
        s_n = 0;
        s_events = 0;
    }
    const uint32_t p_first = bp;
    {
        const uint32_t np = a.n_pops ? min(kStagePops, a.n_pops - p_first) : 0u;
        for (uint32_t i = threadIdx.x; i < np * NAER_DRIVE_TABLE; i += TW)
            s_cdf[i] = a.cdf[(size_t)p_first * NAER_DRIVE_TABLE + i];
    }
    DevState* st = a.st;
    const unsigned long long t0 = st->clock;
    // bucketed events: the M buckets (batches A-1 .. A-M) of this tile,
    // flattened: event i of the tile lies in bucket m with
    // s_pre[m] <= i < s_pre[m + 1], at ev + s_evbase[m] + i
    // (reserved placement: `parts` ranges per batch, those of the tile in
    // the parts of its group bucket)
    __shared__ uint32_t s_pre[kResMaxRanges + 1];
    __shared__ __attribute__((aligned(16))) long long s_evbase[kResMaxRanges];  // 1 KB: also the prefetch's sink
    static_assert(sizeof(s_evbase) >= kWave * 16, "the prefetch writes 16 bytes per lane");
    __shared__ uint32_t s_live[kResMaxRanges];
    __shared__ uint32_t s_ring_live;
    const unsigned long long A = t0 / a.Bfull;
    const uint32_t o0 = (uint32_t)(t0 - A * a.Bfull);
    // reserved placement: this batch's flags start clear (its placement sets them)
    if (BUCKET && a.reserve && blockIdx.x == 0 && threadIdx.x == 0 && (!a.defer || o0 == 0))
        a.flags[(uint32_t)(A % a.R)] = 0;
    // the thread's neuron state, loaded first so that its latency overlaps
    // the event phase
    const uint32_t j = tile * TW + threadIdx.x;
    float V_in = 0.0f;
    int32_t ref_in = 0, last_in = 0;
    if (j < a.N) {
        V_in = a.V[j];
        ref_in = a.ref[j];
        last_in = a.last[j];
    }
    __shared__ unsigned long long s_cls[REC4 ? kRec4StagedClasses : 1];
    if (REC4) {
        for (uint32_t i = threadIdx.x; i < min(a.n_cls, kRec4StagedClasses); i += TW)
            s_cls[i] = a.cls_fx[i];
    }
    if (BUCKET) {
        for (uint32_t i = threadIdx.x; i < a.B * TW; i += TW)
            acc[i] = (AccT)0;
        const uint32_t n_rng = a.M * a.parts;
        for (uint32_t x = threadIdx.x; x < n_rng; x += TW) {
            const uint32_t m = x / a.parts + 1, y = x - (m - 1) * a.parts;
            uint32_t lo = 0, hi = 0, live = 0;
            uint32_t buf = 0;
            if (A >= m) {
                buf = (uint32_t)((A - m) % a.R);
                // bounds are read together with the flag (in bounds either
                // way) and dropped when the batch went through the ring
                const uint32_t f = a.flags[buf];
                uint32_t b_lo = 0, b_hi = 0;
                if (f & 4u) {
                    const uint32_t gt = 1u << a.gt_shift;
                    const uint32_t kc = (m - 1) * (a.NT >> a.gt_shift) + (tile >> a.gt_shift);
                    const uint32_t* tp = a.tb +
                                         (((size_t)buf * (a.NBK >> a.gt_shift) + kc) * a.parts + y) * (gt + 1) +
                                         (tile & (gt - 1));
                    b_lo = tp[0];
                    b_hi = tp[1];
                } else if (y == 0) {
                    const uint32_t* bs = a.bs + (size_t)buf * (a.NBK + 1);
                    const uint32_t k = (m - 1) * a.NT + tile;
                    b_lo = bs[k];
                    b_hi = bs[k + 1];
                }
                if (f & 1u) {
                    lo = b_lo;
                    hi = b_hi;
                }
                // bit 1: some of the batch's events went through the ring
                live = (f & 3u) != 1u ? 1u : 0u;
            }
            s_pre[x + 1] = hi - lo;
            s_evbase[x] = (long long)buf * (long long)a.e_cap + lo;
            s_live[x] = live;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t live = 0;
            s_pre[0] = 0;
            for (uint32_t x = 0; x < n_rng; ++x) {
                const uint32_t n = s_pre[x + 1];
                s_evbase[x] -= s_pre[x];
                s_pre[x + 1] = s_pre[x] + n;
                live |= s_live[x];
            }
            s_ring_live = live;
        }
        __syncthreads();
        const uint32_t total = s_pre[n_rng];
        constexpr uint32_t U = 8;
        uint32_t m = 0;  // bucket of the thread's current event (its indexes only grow)
        if constexpr (REC4) {
            const uint32_t* ev4 = (const uint32_t*)a.ev;
            for (uint32_t i0 = threadIdx.x; i0 < total; i0 += TW * U) {
                uint32_t e[U];
#pragma unroll
                for (uint32_t u = 0; u < U; ++u) {
                    const uint32_t i = i0 + u * TW;
                    e[u] = 0u;
                    if (i < total) {
                        while (i >= s_pre[m + 1])
                            ++m;
                        e[u] = ev4[s_evbase[m] + i];
                    }
                }
#pragma unroll
                for (uint32_t u = 0; u < U; ++u) {
                    const uint32_t o = ((e[u] >> kRec4SlotShift) & 0xFu) - o0;
                    if (i0 + u * TW < total && o < a.B) {
                        const uint32_t c = e[u] >> kRec4ClassShift;
                        const unsigned long long x = c < kRec4StagedClasses ? s_cls[c] : a.cls_fx[c];
                        atomicAdd(&acc[o * TW + (e[u] & kRec4OffsetMask)], x);
                    }
                }
            }
        }
        for (uint32_t i0 = threadIdx.x; !REC4 && i0 < total; i0 += TW * U) {
            uint2 e[U];
#pragma unroll
            for (uint32_t u = 0; u < U; ++u) {
                const uint32_t i = i0 + u * TW;
                e[u] = make_uint2(0xFFFFFFFFu, 0u);
                if (i < total) {
                    while (i >= s_pre[m + 1])
                        ++m;
                    e[u] = a.ev[s_evbase[m] + i];
                }
            }
#pragma unroll
            for (uint32_t u = 0; u < U; ++u) {
                const uint32_t o = (e[u].x >> 16) - o0;
                if (o < a.B) {
                    const float w = __uint_as_float(e[u].y);
                    atomicAdd(&acc[o * TW + (e[u].x & 0xFFFFu)], (unsigned long long)to_fixed(w));
                }
            }
        }
        __syncthreads();
    } else {
        __syncthreads();
    }
    // ring slots of the batch hold events only if a batch they can come
    // from went through the ring
    const bool ring_live = !BUCKET || s_ring_live != 0;

    // the list segments of aligned batch A; a launch that continues the
    // batch (deferred delivery) appends to what the earlier launches wrote
    const uint32_t pl = (uint32_t)(A & 1u);
    naer_spike* seg = (pl ? a.list1 : a.list0) + (size_t)tile * a.seg_cap;
    const bool append = a.defer && o0 > 0;
    const uint32_t base = append ? min(a.seg_count[pl * a.n_seg + tile], a.seg_cap) : 0u;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st->cur = pl;
        st->t0_cur = t0;
        // the flush kernel before this update applied the pending potentiation
        // under the same condition
        if (PLASTIC && t0 + a.B > st->t_flush + (st->window >> 1))
            st->t_flush = t0;
    }
    Stage sg{ s_list, &s_n, seg + base, a.seg_cap - base };

    uint32_t my_events = 0;  // <= B * longest row
    // neurons outside every population (free ranges of an arena) stay idle
    const uint32_t p = (j < a.N && a.n_pops) ? find_pop(a.pops, a.n_pops, p_first, j) : 0u;
    bool todo = j < a.N && a.n_pops && j >= a.pops[p].begin && j < a.pops[p].end;
    // The lanes of a wave are processed population by population, so the
    // population's constants are wave-uniform (scalar registers); a wave
    // almost always lies in one population.
    for (;;) {
        const unsigned long long pending = __ballot(todo);
        if (pending == 0)
            break;
        const uint32_t pu = (uint32_t)__builtin_amdgcn_readlane((int)p, (int)(__ffsll((long long)pending) - 1));
        if (!(todo && p == pu))
            continue;
        todo = false;
        const PopParams& pp = a.pops[pu];
        const uint32_t jl = j - rfl(pp.net_base);                          // neuron id within its net
        const uint32_t t0l = (uint32_t)(t0 - ((unsigned long long)rfl(pp.t_start_hi) << 32 | rfl(pp.t_start_lo)));
        const uint32_t kind = rfl(pp.kind);
        // a model neuron gets its input exactly as a LIF neuron does and
        // leaves its dynamics to the model's kernel
        const bool model = EXTRA && kind == NAER_POP_MODEL;
        if (kind == NAER_POP_LIF || model) {
            RingT* ring = (RingT*)a.ring;
            uint32_t slot0 = (uint32_t)(t0 % a.D);
            // ring slots of the batch: into registers, or (bucketed) added to
            // the thread's own LDS accumulators
            RingT in[BUCKET ? 1 : kMaxBatch];
            if (!BUCKET) {
#pragma unroll
                for (uint32_t s = 0; s < kMaxBatch; ++s) {
                    if (s < a.B) {
                        uint32_t sl = slot0 + s;
                        sl = sl >= a.D ? sl - a.D : sl;
                        in[s] = ring[(size_t)sl * a.N + j];
                    }
                }
            } else if (ring_live) {
                for (uint32_t s = 0; s < a.B; ++s) {
                    uint32_t sl = slot0 + s;
                    sl = sl >= a.D ? sl - a.D : sl;
                    const RingT r = ring[(size_t)sl * a.N + j];
                    acc[s * TW + threadIdx.x] += DET ? (AccT)r : (AccT)to_fixed((float)r);
                }
            }
            if (ring_live) {
                for (uint32_t s = 0; s < a.B; ++s) {
                    uint32_t sl = slot0 + s;
                    sl = sl >= a.D ? sl - a.D : sl;
                    ring[(size_t)sl * a.N + j] = (RingT)0;
                }
            }
            float V = V_in;
            int32_t ref = ref_in;
            int32_t last = last_in;
            uint4 lt = make_uint4(0u, 0u, 0u, 0u);  // newest two spikes (plastic nets)
            uint4 ps = make_uint4(kNoSpike, 0u, kNoSpike, 0u);
            bool lt_new = false;
            if (PLASTIC) {
                lt = a.sd.lat[j];
                ps = a.sd.pst[2 * (size_t)j];
            }
            // population constants in scalar registers for the whole batch
            const float P = rflf(pp.P), v_th = rflf(pp.v_th), v_reset = rflf(pp.v_reset);
            const float v_rest = rflf(pp.v_rest), w_ext = rflf(pp.w_ext);
            const int32_t n_ref = (int32_t)rfl((uint32_t)pp.n_ref);
            const bool has_drive = rfl(pp.has_drive) != 0;
            const float* cdf = pu - p_first < kStagePops ? s_cdf + (pu - p_first) * NAER_DRIVE_TABLE
                                                         : a.cdf + rfl(pp.cdf_offset);
            // the table's first kCdfHead entries, also in scalar registers: the
            // count k = #{i < kCdfHead : u >= cdf[i]} is the search result
            // whenever it is below kCdfHead (the table is nondecreasing)
            float head[kCdfHead];
#pragma unroll
            for (uint32_t i = 0; i < kCdfHead; ++i)
                head[i] = rflf(cdf[i]);
            const uint32_t s_lo = rfl(pp.seed_lo), s_hi = rfl(pp.seed_hi);
            U4 rnd{ 0, 0, 0, 0 };
            uint32_t rnd_blk = 0xFFFFFFFFu;
            // the drive of step s: k * w_ext, k drawn by inverting the Poisson
            // CDF at the step's counter-based uniform
            auto drive = [&](uint32_t s) -> float {
                const uint32_t tl = t0l + s;
                const uint32_t blk = (uint32_t)(tl >> 2);
                if (blk != rnd_blk) {
                    rnd = philox(U4{ jl, blk, NAER_STREAM_DRIVE, 0u }, s_lo, s_hi);
                    rnd_blk = blk;
                }
                const float u = unit24(word_of(rnd, (uint32_t)tl & 3u));
                uint32_t k = 0;
#pragma unroll
                for (uint32_t i = 0; i < kCdfHead; ++i)
                    k += u >= head[i] ? 1u : 0u;
                if (k == kCdfHead) {
                    while (u >= cdf[k])
                        ++k;
                }
                return (float)k * w_ext;
            };
            // synaptic input of the next step (steps are taken in order)
            auto input = [&](uint32_t s) -> float {
                float I;
                if (BUCKET) {
                    I = from_fixed((long long)acc[s * TW + threadIdx.x]);
                } else {
                    const RingT cur = in[0];
#pragma unroll
                    for (uint32_t q = 0; q + 1 < (BUCKET ? 1 : kMaxBatch); ++q)
                        in[q] = in[q + 1];
                    I = DET ? from_fixed((long long)cur) : (float)cur;
                }
                const unsigned long long t = t0 + s;
                if (EXTRA && a.probe && t - a.probe_t0 < a.probe_n)
                    a.probe[(size_t)(t - a.probe_t0) * a.N + j] = I;
                return I;
            };
            unsigned long long row_beg = 0;  // the row of the neuron's latest spike
            uint32_t row_len = 0;
            if (model) {
                // the drive is drawn every step (counter-based: the same values
                // a LIF neuron draws when it is not refractory)
#pragma nounroll
                for (uint32_t s = 0; s < a.B; ++s) {
                    const float I = input(s);
                    a.cin[(size_t)s * a.N + j] = make_float2(I, has_drive ? drive(s) : 0.0f);
                }
            } else {
#pragma nounroll
                for (uint32_t s = 0; s < a.B; ++s) {
                    const unsigned long long t = t0 + s;
                    const float I = input(s);
                    bool fired = false;
                    if (ref > 0) {
                        --ref;
                    } else {
                        const float x = V - v_rest;
                        const float xp = x * P;
                        float v = xp + v_rest;
                        v = v + I;
                        if (has_drive) {
                            const float e = drive(s);
                            v = v + e;
                        }
                        fired = v > v_th;
                        V = fired ? v_reset : v;
                        if (fired) {
                            ref = n_ref;
                            last = (int32_t)t;
                            row_beg = a.row_ptr[j];
                            row_len = (uint32_t)(a.row_ptr[j + 1] - row_beg);
                            my_events += row_len;
                            if (PLASTIC) {
                                // traces at t (spikes before t), from the previous spike
                                const uint32_t ns = lt.x;
                                float xa = 0.0f, ya = 0.0f;
                                if (ns) {
                                    const uint32_t k = (uint32_t)t - lt.y;
                                    const float ax = __uint_as_float(lt.z) + 1.0f;
                                    xa = ax * dec_pow(a.sd.tp, a.sd.n_tp, k);
                                    const float ay = __uint_as_float(ps.y) + 1.0f;
                                    ya = ay * dec_pow(a.sd.tm, a.sd.n_tm, k);
                                }
                                a.sd.hist[(size_t)j * kSpikeHist + (ns & (kSpikeHist - 1))] = (uint32_t)t;
                                uint4* pb = a.sd.pst + 2 * (size_t)j + 1;
                                const uint4 b = *pb;
                                *pb = make_uint4(ps.z, b.x, ns + 1u, 0u);  // s0 -> s2, s2 -> s3
                                lt = make_uint4(ns + 1u, (uint32_t)t, __float_as_uint(xa), lt.z);
                                ps = make_uint4((uint32_t)t, __float_as_uint(ya), ps.x, ps.y);
                                lt_new = true;
                            }
                        }
                    }
                    emit(sg, fired, j, (uint32_t)t, (uint32_t)(t0l + s) * a.dt_us);
                    if (BUCKET && a.prefetch) {
                        // the delivery's first pass reads the rows of the
                        // batch's spikes: the wave loads them now, into a
                        // dead LDS array (the event phase's bases), so that
                        // they are in the caches by then; the wave waits for
                        // them only at its next LDS atomic (a later spike's
                        // staging) or the block's final barrier
                        unsigned long long fm = __ballot(fired);
                        while (fm) {
                            const int l = __builtin_ctzll(fm);
                            fm &= fm - 1ull;
                            const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)row_beg, l);
                            const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(row_beg >> 32), l);
                            const uint32_t len = (uint32_t)__builtin_amdgcn_readlane((int)row_len, l);
                            // 16-byte loads over the row's aligned interior
                            // (LDS: 1 KB per wave instruction), 4-byte
                            // loads for its ends
                            const unsigned long long b = ((unsigned long long)hi << 32) | lo, e = b + len;
                            const unsigned long long a0 = (b + 3) & ~3ull, a1 = e & ~3ull;
                            const uint32_t ln = lane_id();
                            if (a0 >= a1) {
                                for (uint32_t q = ln; q < len; q += kWave)
                                    __builtin_amdgcn_global_load_lds((const void*)(a.col + b + q), (LdsSink*)s_evbase,
                                                                     4, 0, 0);
                            } else {
                                if (ln < a0 - b)
                                    __builtin_amdgcn_global_load_lds((const void*)(a.col + b + ln), (LdsSink*)s_evbase,
                                                                     4, 0, 0);
                                if (ln < e - a1)
                                    __builtin_amdgcn_global_load_lds((const void*)(a.col + a1 + ln), (LdsSink*)s_evbase,
                                                                     4, 0, 0);
                                for (unsigned long long q = a0 + 4ull * ln; q < a1; q += 4ull * kWave)
                                    __builtin_amdgcn_global_load_lds((const void*)(a.col + q), (LdsSink*)s_evbase, 16,
                                                                     0, 0);
                            }
                        }
                    }
                }
                a.V[j] = V;
                a.ref[j] = ref;
                a.last[j] = last;
            }
            if (PLASTIC && lt_new) {
                a.sd.lat[j] = lt;
                a.sd.pst[2 * (size_t)j] = ps;
            }
        } else if (EXTRA) {
            // POISSON source: the reference train (input_neuron_group.cl:26-39)
            // evaluated window by window; state = next spike time and the
            // number already emitted in the current 1500 us epoch.
            uint32_t nxt = a.pnext[j], cnt = a.pcnt[j];
            const uint32_t idx = rfl(pp.pois_base) + (j - rfl(pp.begin));
            const uint32_t gid = ref_poisson_gid(idx);
            const double neg_rate = pp.neg_rate_per_us;
            const uint32_t pois_seed = rfl(pp.pois_seed), n_pois = rfl(pp.n_pois);
#pragma nounroll
            for (uint32_t s = 0; s < a.B; ++s) {
                const unsigned long long t = t0 + s;
                const uint32_t T = (uint32_t)(t0l + s) * a.dt_us;
                if (T % kEpochUs == 0) {
                    nxt = T + ref_isi_us(neg_rate, gid, pois_seed, T * n_pois + idx);
                    cnt = 0;
                }
                const uint32_t end = T + a.dt_us;
                bool pend = cnt < kMaxPoissonPerEpoch && nxt < end;
                while (__any(pend)) {
                    emit(sg, pend, j, (uint32_t)t, nxt);
                    if (pend) {
                        my_events += (uint32_t)(a.row_ptr[j + 1] - a.row_ptr[j]);
                        ++cnt;
                        nxt = nxt + ref_isi_us(neg_rate, gid, pois_seed, nxt * n_pois + idx);
                        pend = cnt < kMaxPoissonPerEpoch && nxt < end;
                    }
                }
            }
            a.pnext[j] = nxt;
            a.pcnt[j] = cnt;
        }
    }
    if (my_events)
        atomicAdd(&s_events, (unsigned long long)my_events);
    __syncthreads();
    const uint32_t total = s_n;
    const uint32_t room = a.seg_cap - base;
    const uint32_t n = total < room ? total : room;
    const uint32_t staged = n < kLdsSpikes ? n : kLdsSpikes;
    for (uint32_t i = threadIdx.x; i < staged; i += TW)
        seg[base + i] = s_list[i];
    if (threadIdx.x == 0) {
        a.seg_count[pl * a.n_seg + tile] = base + n;
        unsigned long long* sev = a.seg_ev + pl * a.n_seg + tile;
        *sev = (append ? *sev : 0ull) + s_events;
        a.spk_part[tile] += total;
        a.ev_part[tile] += s_events;
        if (total > n)
            atomicAdd(&st->overflow, total - n);
        s_base = (a.rec_cap && n) ? atomicAdd(&st->rec_n, (unsigned long long)n) : 0ull;
    }
    if (a.rec_cap && n) {
        // spikes beyond the LDS stage were written to the segment directly
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < n; i += TW) {
            const unsigned long long r = s_base + i;
            if (r < a.rec_cap)
                a.rec[r] = i < kLdsSpikes ? s_list[i] : seg[base + i];
        }
    }
}

// ---------------------------------------------------------------------------
// Scatter: grid (x: the update blocks' list segments, one block each, its
// waves taking the segment's spikes in turn; y: 256-synapse chunk of the row).  Each lane loads 4 synapses before
// issuing its 4 accumulations.  The last kernel of a batch advances the
// device clock.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void batch_epilogue(DevState* st, uint32_t B)
{
    st->clock += B;
    st->batch += 1;
}


// A launch of `b` steps whose spikes stay in the lists until their aligned
// batch is complete (deferred delivery): only the clock advances.
__global__ __launch_bounds__(kWave) void k_clock_advance(DevState* st, uint32_t b)
{
    if (threadIdx.x == 0)
        batch_epilogue(st, b);
}

template <bool DET, bool PLASTIC>
__global__ __launch_bounds__(kBlock) void k_scatter(const ScatterArgs a)
{
    using RingT = typename std::conditional<DET, unsigned long long, float>::type;
    DevState* st = a.st;
    const uint32_t cur = st->cur;
    const naer_spike* list = cur ? a.list1 : a.list0;
    const uint32_t* counts = a.seg_count + cur * a.n_seg;
    if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) {
        if (a.flags)  // this batch's events go through the ring, not the buckets
            a.flags[(st->t0_cur / a.Bfull) % a.R] = 0;
        if (a.is_last)
            batch_epilogue(st, a.B);
    }

    RingT* ring = (RingT*)a.ring;
    const uint32_t lane = lane_id();
    const unsigned long long chunk_off = (unsigned long long)blockIdx.y * kChunk;
    {
        const uint32_t g = blockIdx.x;  // one block per list segment, its waves over the spikes
        const uint32_t cnt = counts[g];
        const naer_spike* seg = list + (size_t)g * a.seg_cap;
        for (uint32_t i = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave); i < cnt; i += kBlock / kWave) {
            const naer_spike sp = seg[i];
            const uint32_t s = sp.neuron;
            const unsigned long long re = a.row_ptr[s + 1];
            const unsigned long long beg = a.row_ptr[s] + chunk_off;
            if (beg >= re)
                continue;
            const uint32_t tD = sp.step % a.D;
            PreSide ps{};
            if (PLASTIC)
                ps = pre_side(a.sd, s, sp.step);
            uint32_t c[kChunkUnroll];
            float wv[kChunkUnroll];
#pragma unroll
            for (uint32_t u = 0; u < kChunkUnroll; ++u) {
                const unsigned long long q = beg + u * kWave + lane;
                c[u] = q < re ? a.col[q] : 0u;
                wv[u] = q < re ? a.w[q] : 0.0f;
            }
#pragma unroll
            for (uint32_t u = 0; u < kChunkUnroll; ++u) {
                const unsigned long long q = beg + u * kWave + lane;
                if (q < re) {
                    const uint32_t tgt = c[u] & kTargetMask;
                    uint32_t sl = tD + ((c[u] >> 24) & 0x7Fu);
                    sl = sl >= a.D ? sl - a.D : sl;
                    float x = wv[u];
                    if (PLASTIC && (c[u] & kPlasticBit)) {
                        x = lazy_stdp<true>(a.sd, TraceTabs{ nullptr, nullptr, 0u }, tgt, a.sd.pst[2 * (size_t)tgt],
                                            a.sd.pst[2 * (size_t)tgt + 1], x, ps.t1, ps.L, ps.has_p, ps.tp, ps.xp);
                        a.w[q] = x;
                    }
                    RingT* dst = ring + (size_t)sl * a.N + tgt;
                    if (DET)
                        atomicAdd((unsigned long long*)dst, (unsigned long long)to_fixed(x));
                    else
                        unsafeAtomicAdd((float*)dst, x);
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Flush of pending potentiation, the first kernel of a plastic batch: when
// the batch would end more than `window` steps after the last flush (or
// `force`), every plastic synapse i -> j takes the potentiations of j's
// spikes since max(i's last spike, the last flush).  One wave per source row
// (grid-stride); the update that follows records the new flush step under
// the same condition.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_flush(const FlushArgs a)
{
    const DevState* st = a.st;
    const unsigned long long clock = st->clock, t_flush = st->t_flush;
    const uint32_t period = st->window >> 1;
    if (!a.force && clock + a.B <= t_flush + period)
        return;
    // a row is flushed now unless its pending range [L, ...) stays within
    // the window until after the next flush (at most period + B steps away):
    // rows of neurons that fired recently are left to their next spike
    const long long keep = (long long)clock + period + 2 * (long long)a.B - (long long)st->window;
    const uint32_t lane = lane_id();
    const uint32_t wpb = kBlock / kWave;
    const uint32_t waves = gridDim.x * wpb;
    for (uint32_t i = blockIdx.x * wpb + rfl(threadIdx.x / kWave); i < a.N; i += waves) {
        if (!a.n_pops)
            break;
        const uint32_t p = find_pop(a.pops, a.n_pops, a.blk_pop[i >> a.tw_shift] & ~kTileIdle, i);
        if (a.pops[p].kind != NAER_POP_LIF || i < a.pops[p].begin || i >= a.pops[p].end)
            continue;  // plastic synapses leave LIF neurons of active networks only
        const unsigned long long rb = a.row_ptr[i], re = a.row_ptr[i + 1];
        if (rb == re)
            continue;
        const uint4 li = a.sd.lat[i];
        const bool has_p = li.x != 0;
        if (!has_p)
            continue;  // no spike of i yet: nothing to potentiate
        const int32_t tp = (int32_t)li.y;
        const float xp = __uint_as_float(li.z);
        const int32_t L = max(tp, (int32_t)a.sd.flushed[i]);
        if (!a.force && (long long)L >= keep)
            continue;
        constexpr uint32_t U = 4;  // synapses per lane in flight
        for (unsigned long long q0 = rb; q0 < re; q0 += U * kWave) {
            uint32_t c[U];
            float w[U];
            uint4 pj[U], pb[U];
#pragma unroll
            for (uint32_t u = 0; u < U; ++u) {
                const unsigned long long q = q0 + u * kWave + lane;
                c[u] = q < re ? a.col[q] : 0u;
            }
#pragma unroll
            for (uint32_t u = 0; u < U; ++u) {
                const unsigned long long q = q0 + u * kWave + lane;
                const bool pl = q < re && (c[u] & kPlasticBit);
                const size_t r = 2 * (size_t)(c[u] & kTargetMask);
                w[u] = pl ? a.w[q] : 0.0f;
                pj[u] = pl ? a.sd.pst[r] : make_uint4(kNoSpike, 0u, kNoSpike, 0u);
                pb[u] = pl ? a.sd.pst[r + 1] : make_uint4(kNoSpike, kNoSpike, 0u, 0u);
            }
#pragma unroll
            for (uint32_t u = 0; u < U; ++u) {
                const unsigned long long q = q0 + u * kWave + lane;
                if (q < re && (c[u] & kPlasticBit))
                    a.w[q] = lazy_stdp<false>(a.sd, TraceTabs{ nullptr, nullptr, 0u }, c[u] & kTargetMask, pj[u],
                                              pb[u], w[u], 0x7FFFFFFF, L, has_p, tp, xp);
            }
        }
        if (lane == 0)
            a.sd.flushed[i] = (uint32_t)clock;
    }
}

// ---------------------------------------------------------------------------
// Bucketed delivery (see BucketArgs): histogram, column scan, bucket starts,
// placement.  Block b walks the spikes of list segments [b*G, (b+1)*G) as
// tasks t = i*C + c (spike i, (64 U)-synapse chunk c of its row; U = the
// schedule's synapses per lane, 4 or 8); wave w takes
// tasks w, w + waves, ...  Both passes use the same split, so each block
// places exactly the events it counted.
// ---------------------------------------------------------------------------
constexpr uint32_t kBucketThreads = 1024;
constexpr uint32_t kArrivalLut = 256;  // bound of the offsets r + d < kMaxBatch + 128

// Arrival batch and slot of an event at offset v = r + d (< kArrivalLut)
// from its aligned batch's first step: v / B and v % B by a multiply-shift,
// exact for v < 256 and B <= 16 (the rounding error of ceil(2^16 / B) adds
// less than 256 / 2^16 < 1 / B).
struct Arrival {
    uint32_t B, magic;
    __device__ __forceinline__ explicit Arrival(uint32_t b) : B(b), magic((65535u + b) / b) {}
    __device__ __forceinline__ uint32_t batch(uint32_t v) const { return (v * magic) >> 16; }
    __device__ __forceinline__ uint32_t slot(uint32_t v, uint32_t m) const { return v - m * B; }
};

// Segment ranges of the histogram/placement blocks, balanced by the batch's
// synaptic events: segment s goes to block
//   min(NB - 1, floor((prefix(s) + events(s) / 2) * NB / total)),
// a nondecreasing map, so every block gets a contiguous (possibly empty)
// range; blk_seg[b] = first segment of block b, blk_seg[NB] = n_act (the
// segments [n_act, n_seg) are idle and get no block).  One
// block of 1024 threads, each owning a contiguous run of segments.
// Walking a segment costs something even without events (its count is read
// in a window of kMaxSegPerBlock, under a barrier): each segment weighs its
// events plus this many, so blocks of a sparse arena do not get thousands of
// empty segments.
constexpr unsigned long long kSegWeight = 64;

__global__ __launch_bounds__(1024) void k_bucket_split(const BucketArgs a)
{
    __shared__ unsigned long long s_part[1024];
    const uint32_t cur = a.st->cur;
    const unsigned long long* seg_ev = a.seg_ev + (size_t)cur * a.n_seg;
    auto weight = [&](uint32_t k) { return seg_ev[k] + kSegWeight; };
    const uint32_t n_act = a.n_act;
    const uint32_t per = (n_act + 1023) / 1024;
    const uint32_t lo = min(threadIdx.x * per, n_act), hi = min(lo + per, n_act);
    unsigned long long sum = 0;
    for (uint32_t k = lo; k < hi; ++k)
        sum += weight(k);
    s_part[threadIdx.x] = sum;
    __syncthreads();
    for (uint32_t off = 1; off < 1024; off <<= 1) {
        const unsigned long long v = threadIdx.x >= off ? s_part[threadIdx.x - off] : 0ull;
        __syncthreads();
        s_part[threadIdx.x] += v;
        __syncthreads();
    }
    const unsigned long long total = s_part[1023];
    unsigned long long run = s_part[threadIdx.x] - sum;  // exclusive prefix of segment lo
    const unsigned long long nb = a.NB;
    auto block_of = [&](uint32_t k, unsigned long long pre, unsigned long long e) -> uint32_t {
        const unsigned long long b = total ? (2ull * pre + e) * nb / (2ull * total) : (unsigned long long)k * nb / n_act;
        return (uint32_t)min(b, nb - 1ull);
    };
    uint32_t prev = 0;
    if (lo > 0 && lo < hi) {
        const unsigned long long e = weight(lo - 1);
        prev = block_of(lo - 1, run - e, e);
    }
    for (uint32_t k = lo; k < hi; ++k) {
        const unsigned long long e = weight(k);
        const uint32_t b = block_of(k, run, e);
        // blocks (prev, b] start at k (blocks [0, b] for the first segment)
        for (uint32_t q = k == 0 ? 0u : prev + 1u; q <= b; ++q)
            a.blk_seg[q] = k;
        prev = b;
        run += e;
    }
    if (hi == n_act && lo < hi) {
        for (uint32_t q = prev + 1u; q <= a.NB; ++q)
            a.blk_seg[q] = n_act;
    }
}

// Rows of the block's spikes, staged in LDS kRowRound at a time: one round
// of parallel loads (spike, row bounds), then the chunk tasks read LDS only.
constexpr uint32_t kRowRound = 1024;

struct RowStage {
    unsigned long long beg[kRowRound];
    uint32_t len[kRowRound];
    uint32_t r[kRowRound];    // spike step - first step of the batch
    uint32_t nrn[kRowRound];  // the spiking neuron
};

// ... and the class of each spike's neuron (class-weight delivery)
struct RowStageCls : RowStage {
    uint32_t cls[kRowRound];
};

// Calls f(r, neuron, q_begin, q_end, i) for every non-empty (64 U)-synapse chunk of the
// rows of the block's spikes (i: the spike's entry in the row stage); f runs
// on a whole wave (q range wave-uniform).  CLS: the stage also holds the
// class of each spike's neuron.
// The block's segments are staged kMaxSegPerBlock at a time.

template <uint32_t U, bool CLS, typename RS, typename F>
__device__ __forceinline__ void for_block_chunks(const BucketArgs& a, uint32_t cur, uint32_t* s_off,
                                                 unsigned long long t0, RS& rs, F f)
{
    constexpr uint32_t kTask = kWave * U;  // synapses of a row per task
    const naer_spike* list = cur ? a.list1 : a.list0;
    const uint32_t lb = xcd_block(blockIdx.x, gridDim.x);
    const uint32_t seg_lo = a.blk_seg[lb], seg_hi = a.blk_seg[lb + 1];
    const uint32_t nwv = blockDim.x / kWave;
    for (uint32_t w0 = seg_lo; w0 < seg_hi; w0 += kMaxSegPerBlock) {
        const uint32_t G = min(kMaxSegPerBlock, seg_hi - w0);
        if (threadIdx.x < G)
            s_off[threadIdx.x + 1] = a.seg_count[cur * a.n_seg + w0 + threadIdx.x];
        __syncthreads();
        if (threadIdx.x == 0) {
            s_off[0] = 0;
            for (uint32_t g = 1; g <= G; ++g)
                s_off[g] += s_off[g - 1];
        }
        __syncthreads();
        const uint32_t n_spk = s_off[G];
        for (uint32_t base = 0; base < n_spk; base += kRowRound) {
            const uint32_t cnt = min(kRowRound, n_spk - base);
            for (uint32_t i = threadIdx.x; i < cnt; i += blockDim.x) {
                const uint32_t x = base + i;
                uint32_t lo = 0, hi = G;  // segment g with s_off[g] <= x < s_off[g + 1]
                while (hi - lo > 1) {
                    const uint32_t mid = (lo + hi) >> 1;
                    if (s_off[mid] <= x)
                        lo = mid;
                    else
                        hi = mid;
                }
                const naer_spike sp = list[(size_t)(w0 + lo) * a.seg_cap + (x - s_off[lo])];
                const unsigned long long rb = a.row_ptr[sp.neuron];
                rs.beg[i] = rb;
                rs.len[i] = (uint32_t)(a.row_ptr[sp.neuron + 1] - rb);
                rs.r[i] = (uint32_t)(sp.step - t0);
                rs.nrn[i] = sp.neuron;
                if constexpr (CLS)
                    rs.cls[i] = a.ncls[sp.neuron];
            }
            __syncthreads();
            // tasks per spike: those of the round's longest row, not of the
            // longest row the net may hold (an arena's max_row is far above
            // its jobs' rows, which would leave most tasks empty); every wave
            // takes the maximum over the staged rows itself
            uint32_t lmax = 0;
            for (uint32_t i = lane_id(); i < cnt; i += kWave)
                lmax = max(lmax, rs.len[i]);
            for (uint32_t off = kWave / 2; off > 0; off >>= 1)
                lmax = max(lmax, (uint32_t)__shfl_xor((int)lmax, (int)off));
            const uint32_t C = rfl((lmax + kTask - 1) / kTask);
            const uint32_t n_tasks = cnt * C;
            for (uint32_t t = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave); t < n_tasks; t += nwv) {
                const uint32_t i = t / C, c = t - i * C;
                const uint32_t off = c * kTask, len = rs.len[i];
                if (off < len)
                    f(rs.r[i], rs.nrn[i], rs.beg[i] + off, rs.beg[i] + min(len, off + kTask), i);
            }
            __syncthreads();
        }
        __syncthreads();  // every thread has read s_off[G] before the next window
    }
}

template <uint32_t U, bool AGG>
__global__ __launch_bounds__(kBucketThreads) void k_bucket_hist(const BucketArgs a)
{
    extern __shared__ uint32_t s_hist[];
    __shared__ uint32_t s_off[kMaxSegPerBlock + 1];
    __shared__ RowStage rs;
    const DevState* st = a.st;
    const uint32_t cur = st->cur;
    const unsigned long long t0 = st->t0_cur - st->t0_cur % a.Bfull;  // the aligned batch (deferred launches)
    for (uint32_t i = threadIdx.x; i < a.NBK; i += blockDim.x)
        s_hist[i] = 0;
    const Arrival arr(a.Bfull);
    __syncthreads();
    const uint32_t lane = lane_id();
    for_block_chunks<U, false>(a, cur, s_off, t0, rs, [&](uint32_t r, uint32_t, unsigned long long beg,
                                                          unsigned long long end, uint32_t) {
        uint32_t c[U];
#pragma unroll
        for (uint32_t u = 0; u < U; ++u) {
            const unsigned long long q = beg + u * kWave + lane;
            c[u] = q < end ? a.col[q] : 0u;
        }
#pragma unroll
        for (uint32_t u = 0; u < U; ++u) {
            const unsigned long long q = beg + u * kWave + lane;
            if (q < end) {
                const uint32_t v = r + ((c[u] >> 24) & 0x7Fu);
                const uint32_t T = (c[u] & kTargetMask) >> a.tw_shift;
                run_add<false, AGG>(s_hist, (arr.batch(v) - 1u) * a.NT + T);
            }
        }
    });
    // count rows over the active tiles' columns only (idle tiles get no
    // events; the column scan takes their totals as zero)
    uint32_t* h = a.hist + (size_t)blockIdx.x * a.NBK;
    const uint32_t nt = a.n_act, ft = a.M * nt;
    for (uint32_t x = threadIdx.x; x < ft; x += blockDim.x) {
        const uint32_t m = x / nt, i = m * a.NT + (x - m * nt);
        h[i] = s_hist[i];
    }
    if (a.direct)
        return;  // the column scan takes the placement's prefixes from the tile rows
    const uint32_t nbc = a.NBK >> a.gt_shift;
    const uint32_t ng = a.NT >> a.gt_shift, ng_act = (nt + (1u << a.gt_shift) - 1) >> a.gt_shift;
    uint32_t* hc = a.hist_c + (size_t)blockIdx.x * nbc;
    for (uint32_t x = threadIdx.x; x < a.M * ng_act; x += blockDim.x) {
        const uint32_t m = x / ng_act, i = m * ng + (x - m * ng_act);
        uint32_t sum = 0;
        for (uint32_t t = 0; t < (1u << a.gt_shift); ++t)
            sum += s_hist[(i << a.gt_shift) + t];
        hc[i] = sum;
    }
}

// Blocks [0, ceil((NBK >> gt_shift) / 64)): hist_c[b][kc] <- sum of
// hist_c[b'][kc] for b' < b.  The remaining blocks: totals[k] <- sum over b
// of hist_f[b][k].  Direct placement: every block does both on the tile
// rows, hist_f[b][k] <- its prefix, totals[k] <- the column's sum.  A block
// covers 64 columns with 16 row groups.
constexpr uint32_t kScanCols = 64, kScanGroups = 16;

__global__ __launch_bounds__(kScanCols * kScanGroups) void k_bucket_colscan(const BucketArgs a)
{
    __shared__ uint32_t s_part[kScanGroups][kScanCols];
    // direct placement: one set of blocks over the tile columns, prefixes in
    // place (the placement's) and totals
    const uint32_t nbc = a.NBK >> a.gt_shift;
    const uint32_t coarse_blocks = a.direct ? 0u : (nbc + kScanCols - 1) / kScanCols;
    const bool coarse = a.direct || blockIdx.x < coarse_blocks;
    uint32_t* hist = (coarse && !a.direct) ? a.hist_c : a.hist;
    const uint32_t width = (coarse && !a.direct) ? nbc : a.NBK;
    const uint32_t col = threadIdx.x % kScanCols, grp = threadIdx.x / kScanCols;
    const uint32_t k = (blockIdx.x - (coarse && !a.direct ? 0u : coarse_blocks)) * kScanCols + col;
    const uint32_t rows = (a.NB + kScanGroups - 1) / kScanGroups;
    const uint32_t b0 = min(grp * rows, a.NB), b1 = min(b0 + rows, a.NB);
    // columns of idle tiles (group columns of idle groups) were not written
    // by the histogram: their totals are zero and their prefixes unused
    const uint32_t per_m = (coarse && !a.direct) ? a.NT >> a.gt_shift : a.NT;
    const uint32_t act = (coarse && !a.direct) ? (a.n_act + (1u << a.gt_shift) - 1) >> a.gt_shift : a.n_act;
    const bool live = k < width && k % per_m < act;
    uint32_t sum = 0;
    if (live) {
        for (uint32_t b = b0; b < b1; b += 8) {
            uint32_t c[8];
#pragma unroll
            for (uint32_t u = 0; u < 8; ++u)
                c[u] = b + u < b1 ? hist[(size_t)(b + u) * width + k] : 0u;
#pragma unroll
            for (uint32_t u = 0; u < 8; ++u)
                sum += c[u];
        }
    }
    s_part[grp][col] = sum;
    __syncthreads();
    if (grp == 0) {
        uint32_t run = 0;
        for (uint32_t g = 0; g < kScanGroups; ++g) {
            const uint32_t x = s_part[g][col];
            s_part[g][col] = run;
            run += x;
        }
        if ((!coarse || a.direct) && k < width)
            a.totals[k] = run;
    }
    if (!coarse)
        return;
    __syncthreads();
    if (!live)
        return;
    uint32_t run = s_part[grp][col];
    for (uint32_t b = b0; b < b1; b += 8) {
        uint32_t c[8];
#pragma unroll
        for (uint32_t u = 0; u < 8; ++u)
            c[u] = b + u < b1 ? hist[(size_t)(b + u) * width + k] : 0u;
#pragma unroll
        for (uint32_t u = 0; u < 8; ++u) {
            if (b + u < b1)
                hist[(size_t)(b + u) * width + k] = run;
            run += c[u];
        }
    }
}

// Exclusive scan of the bucket totals (one block, staged in LDS) into the
// bucket starts of the batch's buffer, and the flag saying whether the batch
// fits the buffer.
__global__ __launch_bounds__(1024) void k_bucket_starts(const BucketArgs a)
{
    extern __shared__ uint32_t s_tot[];  // [NBK]
    __shared__ unsigned long long s_part[1024];
    for (uint32_t k = threadIdx.x; k < a.NBK; k += 1024)
        s_tot[k] = a.totals[k];
    __syncthreads();
    const uint32_t per = (a.NBK + 1023) / 1024;
    const uint32_t lo = min(threadIdx.x * per, a.NBK), hi = min(lo + per, a.NBK);
    unsigned long long sum = 0;
    for (uint32_t k = lo; k < hi; ++k)
        sum += s_tot[k];
    s_part[threadIdx.x] = sum;
    __syncthreads();
    for (uint32_t off = 1; off < 1024; off <<= 1) {
        const unsigned long long v = threadIdx.x >= off ? s_part[threadIdx.x - off] : 0ull;
        __syncthreads();
        s_part[threadIdx.x] += v;
        __syncthreads();
    }
    unsigned long long run = s_part[threadIdx.x] - sum;  // exclusive
    for (uint32_t k = lo; k < hi; ++k) {
        const uint32_t x = s_tot[k];
        s_tot[k] = (uint32_t)min(run, 0xFFFFFFFFull);
        run += x;
    }
    __syncthreads();
    const unsigned long long t0 = a.st->t0_cur;
    const uint32_t buf = (uint32_t)((t0 / a.Bfull) % a.R);
    uint32_t* bs = a.bs + (size_t)buf * (a.NBK + 1);
    for (uint32_t k = threadIdx.x; k < a.NBK; k += 1024) {
        bs[k] = s_tot[k];
        a.gcur[k] = s_tot[k];
    }
    if (threadIdx.x == 1023) {
        const unsigned long long total = s_part[1023];
        bs[a.NBK] = (uint32_t)min(total, 0xFFFFFFFFull);
        a.flags[buf] = total <= a.e_cap ? 1u : 0u;
    }
}

// REC4: class-weight delivery (4-byte records, no weight reads); static
// networks only.
template <bool DET, uint32_t U, bool AGG, bool PLASTIC, bool REC4>
__global__ __launch_bounds__(kBucketThreads) void k_bucket_place(const BucketArgs a)
{
    static_assert(!(PLASTIC && REC4), "plastic synapses carry their own weights");
    using RingT = typename std::conditional<DET, unsigned long long, float>::type;
    using Stage = typename std::conditional<REC4, RowStageCls, RowStage>::type;
    extern __shared__ uint32_t s_cur[];
    __shared__ uint32_t s_off[kMaxSegPerBlock + 1];
    __shared__ Stage rs;
    DevState* st = a.st;
    const uint32_t cur = st->cur;
    const unsigned long long t0 = st->t0_cur - st->t0_cur % a.Bfull;  // the aligned batch (deferred launches)
    const uint32_t buf = (uint32_t)((t0 / a.Bfull) % a.R);
    const uint32_t nbc = a.NBK >> a.gt_shift;
    const uint32_t ng = a.NT >> a.gt_shift;
    const uint32_t g_shift = a.tw_shift + a.gt_shift;
    // reserved placement: s_cur[kc] runs from the region's start to its end
    // s_end[kc] (s_beg[kc]: the start)
    uint32_t* s_end = s_cur + a.n_cnt;
    uint32_t* s_beg = s_end + a.n_cnt;
    uint32_t* s_spb = s_beg + a.n_cnt;  // the group bucket's spill region [s_spb, s_spe)
    uint32_t* s_spe = s_spb + a.n_cnt;
    const uint32_t lb = xcd_block(blockIdx.x, gridDim.x);  // the block's segments (as for_block_chunks)
    bool ok;
    if (a.reserve) {
        // P_b, the block's events and the batch's, from the segments' counts
        __shared__ unsigned long long s_red[3][kBucketThreads / kWave];
        const unsigned long long* sev = a.seg_ev + (size_t)cur * a.n_seg;
        const uint32_t s_lo = a.blk_seg[lb], s_hi = a.blk_seg[lb + 1];
        // the shares of this thread's counters, loaded with the counts
        unsigned long long cv[kResCntPerThread], ccv[kResCntPerThread];
#pragma unroll
        for (uint32_t k = 0; k < kResCntPerThread; ++k) {
            const uint32_t i = threadIdx.x + k * kBucketThreads;
            cv[k] = i < nbc ? a.rc[i] : 0ull;
            ccv[k] = i < nbc ? a.rcc[i] : 0ull;
        }
        const unsigned long long cc_all = a.rcc[nbc];
        unsigned long long pre = 0, mine = 0, tot = 0;
        for (uint32_t k = threadIdx.x; k < a.n_act; k += blockDim.x) {
            const unsigned long long e = sev[k];
            tot += e;
            pre += k < s_lo ? e : 0ull;
            mine += (k >= s_lo && k < s_hi) ? e : 0ull;
        }
        for (uint32_t o = kWave / 2; o > 0; o >>= 1) {
            tot += __shfl_xor(tot, (int)o);
            pre += __shfl_xor(pre, (int)o);
            mine += __shfl_xor(mine, (int)o);
        }
        if (lane_id() == 0) {
            s_red[0][threadIdx.x / kWave] = tot;
            s_red[1][threadIdx.x / kWave] = pre;
            s_red[2][threadIdx.x / kWave] = mine;
        }
        __syncthreads();
        tot = pre = mine = 0;
        for (uint32_t w = 0; w < blockDim.x / kWave; ++w) {
            tot += s_red[0][w];
            pre += s_red[1][w];
            mine += s_red[2][w];
        }
        ok = tot <= a.e_cap && res_base(cc_all, tot, nbc, a.NB) <= a.e_cap;
        if (threadIdx.x == 0) {
            a.pb[lb] = pre;
            if (lb == 0) {
                a.pb[a.NB] = tot;
                // bit 0: bucketed, bit 2: tile ranges in tb (the update of
                // the batch cleared the flags)
                if (ok)
                    atomicOr(&a.flags[buf], 5u);
            }
        }
        if (ok) {
#pragma unroll
            for (uint32_t k = 0; k < kResCntPerThread; ++k) {
                const uint32_t i = threadIdx.x + k * kBucketThreads;
                if (i < nbc) {
                    const unsigned long long c = cv[k];
                    const unsigned long long base = res_base(ccv[k], tot, i, a.NB);
                    const uint32_t b0 = (uint32_t)(base + ((c * pre) >> kResShift) + (unsigned long long)kResPad * lb);
                    s_beg[i] = b0;
                    s_cur[i] = b0;
                    s_end[i] = (uint32_t)(base + ((c * (pre + mine)) >> kResShift) +
                                          (unsigned long long)kResPad * (lb + 1));
                    const uint32_t sb = (uint32_t)(base + ((c * tot) >> kResShift) + (unsigned long long)kResPad * a.NB);
                    s_spb[i] = sb;
                    s_spe[i] = sb + (uint32_t)((c * tot) >> (kResShift + 2)) + kResSpillPad;
                }
            }
        }
    } else {
        ok = a.flags[buf] != 0;
        if (ok) {
            const uint32_t* bs = a.bs + (size_t)buf * (a.NBK + 1);
            const uint32_t* h = a.direct ? a.hist + (size_t)blockIdx.x * a.NBK : a.hist_c + (size_t)blockIdx.x * nbc;
            for (uint32_t i = threadIdx.x; i < a.n_cnt; i += blockDim.x)
                s_cur[i] = bs[i << a.gt_shift] + h[i];
        }
    }
    const Arrival arr(a.Bfull);
    TraceTabs tt{ nullptr, nullptr, 0u };
    if (PLASTIC) {
        __shared__ float s_tp[kLdsTrace], s_tm[kLdsTrace];
        const uint32_t n_l = min(kLdsTrace, min(a.sd.n_tp, a.sd.n_tm));
        for (uint32_t i = threadIdx.x; i < n_l; i += blockDim.x) {
            s_tp[i] = a.sd.tp[i];
            s_tm[i] = a.sd.tm[i];
        }
        tt = TraceTabs{ (LdsFloat*)s_tp, (LdsFloat*)s_tm, n_l };
    }
    __syncthreads();
    if (a.is_last && blockIdx.x == 0 && threadIdx.x == 0)
        batch_epilogue(st, a.B);
    // direct: the group buckets are the tile buckets of the batch's buffer
    uint2* stage = a.direct ? a.ev + (size_t)buf * a.e_cap : a.stage;
    uint32_t* stage4 = a.direct ? (uint32_t*)a.ev + (size_t)buf * a.e_cap : (uint32_t*)a.stage;
    RingT* ring = (RingT*)a.ring;
    const uint32_t lane = lane_id();
    for_block_chunks<U, REC4>(a, cur, s_off, t0, rs, [&](uint32_t r, uint32_t nrn, unsigned long long beg,
                                                         unsigned long long end, uint32_t si) {
        uint32_t c[U];
        float w[U];
        uint32_t cls = 0;
        if constexpr (REC4) {
            cls = rfl(rs.cls[si]);
            const float wc = a.cls_w[cls];  // the ring fallback's weight
#pragma unroll
            for (uint32_t u = 0; u < U; ++u) {
                const unsigned long long q = beg + u * kWave + lane;
                c[u] = q < end ? a.col[q] : 0u;
                w[u] = wc;
            }
        } else {
#pragma unroll
            for (uint32_t u = 0; u < U; ++u) {
                const unsigned long long q = beg + u * kWave + lane;
                c[u] = q < end ? a.col[q] : 0u;
                w[u] = q < end ? a.w[q] : 0.0f;
            }
        }
        if (PLASTIC) {
            // lazy pair STDP of the row's plastic synapses before they are placed
            const PreSide ps = pre_side(a.sd, nrn, (uint32_t)(t0 + r));
            // the targets' newest two spikes for all U synapses first, then the walks
            uint4 pj[U], pb[U];
#pragma unroll
            for (uint32_t u = 0; u < U; ++u) {
                const unsigned long long q = beg + u * kWave + lane;
                const bool pl = q < end && (c[u] & kPlasticBit);
                const size_t rj = 2 * (size_t)(c[u] & kTargetMask);
                pj[u] = pl ? a.sd.pst[rj] : make_uint4(kNoSpike, 0u, kNoSpike, 0u);
                pb[u] = pl ? a.sd.pst[rj + 1] : make_uint4(kNoSpike, kNoSpike, 0u, 0u);
            }
#pragma unroll
            for (uint32_t u = 0; u < U; ++u) {
                const unsigned long long q = beg + u * kWave + lane;
                if (q < end && (c[u] & kPlasticBit)) {
                    w[u] = lazy_stdp<true>(a.sd, tt, c[u] & kTargetMask, pj[u], pb[u], w[u], ps.t1, ps.L, ps.has_p,
                                           ps.tp, ps.xp);
                    a.w[q] = w[u];
                }
            }
        }
        if (ok) {
            // (storing each record right after its position was measured
            // faster than all positions first, then all stores)
#pragma unroll
            for (uint32_t u = 0; u < U; ++u) {
                const unsigned long long q = beg + u * kWave + lane;
                if (q < end) {
                    const uint32_t d = (c[u] >> 24) & 0x7Fu;
                    const uint32_t tgt = c[u] & kTargetMask;
                    const uint32_t v = r + d;
                    const uint32_t g = tgt >> g_shift;
                    const uint32_t m = arr.batch(v);
                    const uint32_t key = (m - 1u) * ng + g;
                    uint32_t pos = run_add<true, AGG, kPlaceRunHeads>(s_cur, key);
                    const uint32_t rec = (tgt - (g << g_shift)) | (arr.slot(v, m) << (REC4 ? kRec4SlotShift : 16u)) |
                                         (REC4 ? cls << kRec4ClassShift : 0u);
                    bool store = true;
                    if (a.reserve && pos >= s_end[key]) {
                        // the region is full: the bucket's spill region, else the ring
                        const uint32_t sp = atomicAdd(&a.spill[key], 1u);
                        const uint32_t sb = s_spb[key];
                        pos = sb + sp;
                        if (sp >= s_spe[key] - sb) {
                            store = false;
                            uint32_t sl = (uint32_t)((t0 + r) % a.D) + ((c[u] >> 24) & 0x7Fu);
                            sl = sl >= a.D ? sl - a.D : sl;
                            RingT* dst = ring + (size_t)sl * a.N + tgt;
                            if (DET)
                                atomicAdd((unsigned long long*)dst, (unsigned long long)to_fixed(w[u]));
                            else
                                unsafeAtomicAdd((float*)dst, w[u]);
                            atomicOr(&a.flags[buf], 2u);
                        }
                    }
                    if (store) {
                        if constexpr (REC4)
                            stage4[pos] = rec;
                        else
                            stage[pos] = make_uint2(rec, __float_as_uint(w[u]));
                    }
                }
            }
        } else {
            // the batch does not fit the event buffer: deliver through the ring
            const uint32_t slot0 = (uint32_t)((t0 + r) % a.D);
#pragma unroll
            for (uint32_t u = 0; u < U; ++u) {
                const unsigned long long q = beg + u * kWave + lane;
                if (q < end) {
                    const uint32_t d = (c[u] >> 24) & 0x7Fu;
                    const uint32_t tgt = c[u] & kTargetMask;
                    uint32_t sl = slot0 + d;
                    sl = sl >= a.D ? sl - a.D : sl;
                    RingT* dst = ring + (size_t)sl * a.N + tgt;
                    if (DET)
                        atomicAdd((unsigned long long*)dst, (unsigned long long)to_fixed(w[u]));
                    else
                        unsafeAtomicAdd((float*)dst, w[u]);
                }
            }
        }
    });
    if (a.reserve && ok) {
        // the events each region holds (the refine reads them)
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < nbc; i += blockDim.x)
            a.rcnt[(size_t)i * a.NB + lb] = min(s_cur[i], s_end[i]) - s_beg[i];
    }
}

// The staging buffer holds the batch's events in group-bucket order (group
// bucket kc = (m - 1) * groups + g covers buckets [kc << gt_shift,
// (kc + 1) << gt_shift)).  With `balance` (packs, arenas) block b takes the
// b-th of gridDim.x equal parts of the whole buffer, whatever the spread of
// the events over groups; otherwise block b takes part b % RS of group bucket
// b / RS.  It moves the events of each group bucket it overlaps to their tile's bucket of
// ev[E % R], kRefineChunk at a time: counted per tile in LDS, a base per tile
// taken from the buckets' global cursors, written.
constexpr uint32_t kRefineThreads = 512;
constexpr uint32_t kRefineChunk = 4096;

template <bool AGG, bool REC4>
__global__ __launch_bounds__(kRefineThreads) void k_bucket_refine(const BucketArgs a)
{
    using Rec = typename std::conditional<REC4, uint32_t, uint2>::type;
    __shared__ Rec s_ev[kRefineChunk];
    __shared__ uint16_t s_rank[kRefineChunk];
    __shared__ uint32_t s_cnt[kMaxGroupTiles];
    __shared__ uint32_t s_base[kMaxGroupTiles];
    const unsigned long long t0 = a.st->t0_cur;
    const uint32_t buf = (uint32_t)((t0 / a.Bfull) % a.R);
    if (!a.flags[buf])
        return;
    const uint32_t* bs = a.bs + (size_t)buf * (a.NBK + 1);
    const uint32_t gt = 1u << a.gt_shift;
    const uint32_t ngb = a.NBK >> a.gt_shift;
    uint32_t a0, a1, lo_k;
    if (a.balance) {
        const uint32_t total = bs[a.NBK];
        a0 = (uint32_t)(((unsigned long long)total * blockIdx.x) / gridDim.x);
        a1 = (uint32_t)(((unsigned long long)total * (blockIdx.x + 1)) / gridDim.x);
        if (a0 >= a1)
            return;
        // first group bucket that ends after a0
        uint32_t hi_k = ngb - 1;
        lo_k = 0;
        while (lo_k < hi_k) {
            const uint32_t mid = (lo_k + hi_k) >> 1;
            if (bs[(mid + 1) << a.gt_shift] > a0)
                hi_k = mid;
            else
                lo_k = mid + 1;
        }
    } else {
        // a single network: RS equal parts of each group bucket
        lo_k = blockIdx.x / a.RS;
        const uint32_t y = blockIdx.x - lo_k * a.RS;
        const uint32_t lo = bs[lo_k << a.gt_shift], len = bs[(lo_k + 1) << a.gt_shift] - lo;
        a0 = lo + (uint32_t)(((unsigned long long)len * y) / a.RS);
        a1 = lo + (uint32_t)(((unsigned long long)len * (y + 1)) / a.RS);
        if (a0 >= a1)
            return;
    }
    Rec* ev = (Rec*)a.ev + (size_t)buf * a.e_cap;
    const Rec* stage = (const Rec*)a.stage;
    const uint32_t tmask = (1u << a.tw_shift) - 1u;
    // the record's target offset within its group, and the record with the
    // offset within its tile
    auto offset = [](const Rec& e) -> uint32_t {
        if constexpr (REC4)
            return e & kRec4OffsetMask;
        else
            return e.x & 0xFFFFu;
    };
    auto retile = [&](const Rec& e) -> Rec {
        if constexpr (REC4)
            return (e & tmask) | (e & ~kRec4OffsetMask);
        else
            return make_uint2((e.x & tmask) | (e.x & 0xFFFF0000u), e.y);
    };
    for (uint32_t kc = lo_k; kc < ngb && bs[kc << a.gt_shift] < a1; ++kc) {
        const uint32_t k0 = kc << a.gt_shift;
        const uint32_t g0 = max(bs[k0], a0), g1 = min(bs[k0 + gt], a1);
        for (uint32_t r0 = g0; r0 < g1; r0 += kRefineChunk) {
            const uint32_t n = min(kRefineChunk, g1 - r0);
            if (threadIdx.x < gt)
                s_cnt[threadIdx.x] = 0;
            __syncthreads();
            {
                // the chunk's records into registers first (all loads in
                // flight together), then ranked per tile
                constexpr uint32_t kPer = kRefineChunk / kRefineThreads;
                Rec e[kPer];
#pragma unroll
                for (uint32_t k = 0; k < kPer; ++k) {
                    const uint32_t i = threadIdx.x + k * kRefineThreads;
                    if (i < n)
                        e[k] = stage[r0 + i];
                }
#pragma unroll
                for (uint32_t k = 0; k < kPer; ++k) {
                    const uint32_t i = threadIdx.x + k * kRefineThreads;
                    if (i < n) {
                        s_ev[i] = e[k];
                        s_rank[i] = (uint16_t)run_add<true, AGG>(s_cnt, offset(e[k]) >> a.tw_shift);
                    }
                }
            }
            __syncthreads();
            if (threadIdx.x < gt && s_cnt[threadIdx.x])
                s_base[threadIdx.x] = atomicAdd(&a.gcur[k0 + threadIdx.x], s_cnt[threadIdx.x]);
            __syncthreads();
            for (uint32_t i = threadIdx.x; i < n; i += kRefineThreads) {
                const Rec e = s_ev[i];
                const uint32_t t = offset(e) >> a.tw_shift;
                ev[s_base[t] + s_rank[i]] = retile(e);
            }
            __syncthreads();
        }
    }
}

// Reserved placement's refine: block (kc, y) takes part y of group bucket
// kc, the regions of placement blocks [y NB / RS, (y + 1) NB / RS) (the
// last part also the bucket's spill region), and writes their events sorted
// by tile into the part's span of ev[E % R], which starts where its first
// region starts: counted per tile in LDS, scanned, placed by LDS cursors.  A
// part of at most kResCache events is read once (kept in LDS), a larger one
// twice.  The tile ranges go to tb[E % R][kc][y].
constexpr uint32_t kResThreads = 512;
constexpr uint32_t kResCacheBytes = 24576;
constexpr uint32_t kResLoads = 8;  // events a thread loads at once
constexpr uint32_t kResChunks = 256;  // 64-event chunks whose first region is tabled

template <bool REC4>
__global__ __launch_bounds__(kResThreads) void k_bucket_refine_res(const BucketArgs a)
{
    using Rec = typename std::conditional<REC4, uint32_t, uint2>::type;
    constexpr uint32_t kCache = kResCacheBytes / sizeof(Rec);
    __shared__ Rec s_ev[kCache];
    __shared__ uint32_t s_q[kResMaxPartRegions + 1];  // exclusive prefix of the regions' counts
    __shared__ uint32_t s_src[kResMaxPartRegions];    // their starts in the staging buffer
    __shared__ uint32_t s_cnt[kMaxGroupTiles];
    __shared__ uint32_t s_pos[kMaxGroupTiles + 1];
    __shared__ uint32_t s_wsum[kResThreads / kWave];
    __shared__ uint16_t s_reg[kResChunks];
    const uint32_t ngb = a.NBK >> a.gt_shift;
    const uint32_t gt = 1u << a.gt_shift;
    const uint32_t kc = blockIdx.x / a.RS, y = blockIdx.x - kc * a.RS;
    const uint32_t b0 = (uint32_t)((unsigned long long)a.NB * y / a.RS);
    const uint32_t b1 = (uint32_t)((unsigned long long)a.NB * (y + 1) / a.RS);
    const bool last = y + 1 == a.RS;
    const uint32_t nr = b1 - b0 + (last ? 1u : 0u);
    // everything below depends only on these loads (issued together); the
    // placement used the buckets iff the batch fit, which E decides
    const unsigned long long t0 = a.st->t0_cur;
    const unsigned long long E = a.pb[a.NB];
    const unsigned long long c = a.rc[kc];
    const unsigned long long Cc = a.rcc[kc], cc_all = a.rcc[ngb];
    if (!(E <= a.e_cap && res_base(cc_all, E, ngb, a.NB) <= a.e_cap))
        return;  // the batch went through the ring
    const uint32_t buf = (uint32_t)((t0 / a.Bfull) % a.R);
    const unsigned long long base = res_base(Cc, E, kc, a.NB);
    const uint32_t out = (uint32_t)(base + ((c * a.pb[b0]) >> kResShift) + (unsigned long long)kResPad * b0);
    const uint32_t lane = lane_id(), wave = threadIdx.x / kWave;
    // region counts and starts; exclusive scan of the counts (two per thread)
    uint32_t n2[2];
#pragma unroll
    for (uint32_t h = 0; h < 2; ++h) {
        const uint32_t i = 2 * threadIdx.x + h;
        n2[h] = 0;
        if (i < nr) {
            const uint32_t b = b0 + i;
            if (b < b1) {
                s_src[i] = (uint32_t)(base + ((c * a.pb[b]) >> kResShift) + (unsigned long long)kResPad * b);
                n2[h] = a.rcnt[(size_t)kc * a.NB + b];
            } else {
                s_src[i] = (uint32_t)(base + ((c * E) >> kResShift) + (unsigned long long)kResPad * a.NB);
                n2[h] = min(a.spill[kc], (uint32_t)((c * E) >> (kResShift + 2)) + kResSpillPad);
            }
        }
    }
    uint32_t incl = n2[0] + n2[1];
    for (uint32_t o = 1; o < kWave; o <<= 1) {
        const uint32_t v = (uint32_t)__shfl_up((int)incl, o);
        incl += lane >= o ? v : 0u;
    }
    if (lane == kWave - 1)
        s_wsum[wave] = incl;
    if (threadIdx.x < gt)
        s_cnt[threadIdx.x] = 0;
    __syncthreads();
    uint32_t wpre = 0;
    for (uint32_t w = 0; w < wave; ++w)
        wpre += s_wsum[w];
    {
        const uint32_t ex = wpre + incl - n2[0] - n2[1];
        const uint32_t i = 2 * threadIdx.x;
        if (i < nr)
            s_q[i] = ex;
        if (i + 1 < nr)
            s_q[i + 1] = ex + n2[0];
        if (i + 1 == nr || i + 2 == nr)
            s_q[nr] = ex + n2[0] + n2[1];
    }
    if (nr == 0 && threadIdx.x == 0)
        s_q[0] = 0;
    __syncthreads();
    const uint32_t n = s_q[nr];
    const bool cached = n <= kCache;
    // region of the first event of each 64-event chunk
    for (uint32_t ch = threadIdx.x; ch < kResChunks && (ch << 6) < n; ch += blockDim.x) {
        const uint32_t i = ch << 6;
        uint32_t lo = 0, hi = nr;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (s_q[mid] <= i)
                lo = mid;
            else
                hi = mid;
        }
        s_reg[ch] = lo;
    }
    __syncthreads();
    if (last && threadIdx.x == 0)
        a.spill[kc] = 0;  // read above (before the barrier); the next batch's placement starts over
    const Rec* stage = (const Rec*)a.stage;
    Rec* ev = (Rec*)a.ev + (size_t)buf * a.e_cap;
    const uint32_t tmask = (1u << a.tw_shift) - 1u;
    auto offset = [](const Rec& e) -> uint32_t {
        if constexpr (REC4)
            return e & kRec4OffsetMask;
        else
            return e.x & 0xFFFFu;
    };
    auto retile = [&](const Rec& e) -> Rec {
        if constexpr (REC4)
            return (e & tmask) | (e & ~kRec4OffsetMask);
        else
            return make_uint2((e.x & tmask) | (e.x & 0xFFFF0000u), e.y);
    };
    // calls f(event, index in the part) for every event of the part's
    // regions: thread t takes events t, t + 512, ... (kResLoads loads in
    // flight), each finding its region by a binary search of the prefix
    auto walk = [&](auto f) {
        for (uint32_t i0 = threadIdx.x; i0 < n; i0 += blockDim.x * kResLoads) {
            Rec e[kResLoads];
#pragma unroll
            for (uint32_t u = 0; u < kResLoads; ++u) {
                const uint32_t i = i0 + u * blockDim.x;
                if (i < n) {
                    // the last region r with s_q[r] <= i: from the region of
                    // the event's 64-event chunk (regions hold ~80 events)
                    uint32_t lo;
                    if ((i >> 6) < kResChunks) {
                        lo = s_reg[i >> 6];
                        while (s_q[lo + 1] <= i)
                            ++lo;
                    } else {
                        lo = 0;
                        uint32_t hi = nr;
                        while (hi - lo > 1) {
                            const uint32_t mid = (lo + hi) >> 1;
                            if (s_q[mid] <= i)
                                lo = mid;
                            else
                                hi = mid;
                        }
                    }
                    e[u] = stage[s_src[lo] + (i - s_q[lo])];
                }
            }
#pragma unroll
            for (uint32_t u = 0; u < kResLoads; ++u) {
                const uint32_t i = i0 + u * blockDim.x;
                if (i < n)
                    f(e[u], i);
            }
        }
    };
    walk([&](const Rec& e, uint32_t i) {
        if (cached)
            s_ev[i] = e;
        atomicAdd(&s_cnt[offset(e) >> a.tw_shift], 1u);
    });
    __syncthreads();
    if (wave == 0) {
        // exclusive scan of the tile counts (gt <= 64: one wave)
        const uint32_t x = lane < gt ? s_cnt[lane] : 0u;
        uint32_t inc = x;
        for (uint32_t o = 1; o < kWave; o <<= 1) {
            const uint32_t v = (uint32_t)__shfl_up((int)inc, o);
            inc += lane >= o ? v : 0u;
        }
        uint32_t* tb = a.tb + (((size_t)buf * ngb + kc) * a.parts + y) * (gt + 1);
        if (lane < gt) {
            s_pos[lane] = inc - x;
            tb[lane] = out + inc - x;
        }
        if (lane == 0)
            tb[gt] = out + n;
        // the part slots this schedule does not use hold empty ranges
        if (y == 0) {
            for (uint32_t i = a.RS * (gt + 1) + lane; i < a.parts * (gt + 1); i += kWave)
                tb[i] = 0;
        }
    }
    __syncthreads();
    auto put = [&](const Rec& e) {
        const uint32_t p = atomicAdd(&s_pos[offset(e) >> a.tw_shift], 1u);
        ev[out + p] = retile(e);
    };
    if (cached) {
        for (uint32_t i = threadIdx.x; i < n; i += blockDim.x)
            put(s_ev[i]);
    } else {
        walk([&](const Rec& e, uint32_t) { put(e); });
    }
}

// Zeroes the weight of every bucketed event whose target lies in [j_begin,
// j_end): run when a network is (re)initialised on that range, so events of
// a network that held the range before cannot reach the new one (the ring
// cells of the range are cleared by k_init_state).  Grid: (tiles overlapping
// the range, R * M buckets of each tile).
__global__ __launch_bounds__(kBlock) void k_bucket_scrub(const UpdateArgs a, uint32_t j_begin, uint32_t j_end)
{
    const uint32_t T = (j_begin >> a.tw_shift) + blockIdx.x;
    const uint32_t buf = blockIdx.y / a.M;
    const uint32_t m = blockIdx.y % a.M;
    if (!a.flags[buf])
        return;
    const uint32_t* bs = a.bs + (size_t)buf * (a.NBK + 1);
    const uint32_t k = m * a.NT + T;
    const uint32_t lo = bs[k], hi = bs[k + 1];
    uint2* ev = (uint2*)a.ev + (size_t)buf * a.e_cap;
    const uint32_t base = T << a.tw_shift;
    for (uint32_t i = lo + threadIdx.x; i < hi; i += kBlock) {
        const uint32_t j = base + (ev[i].x & 0xFFFFu);
        if (j >= j_begin && j < j_end)
            ev[i].y = 0u;
    }
}

// ---------------------------------------------------------------------------
// Connectivity build
// ---------------------------------------------------------------------------
__device__ __forceinline__ float draw_weight(const ConnArgs& c, uint32_t word)
{
    if (c.weight != 0.0f)
        return c.weight;
    const float span = c.wmax - c.wmin;
    const float x = span * unit24(word);
    return c.wmin + x;
}

template <bool PLACE>
__global__ __launch_bounds__(kBlock) void k_conn_fixed(const ConnArgs c, uint32_t* deg,
                                                       unsigned long long* cursor, uint32_t* col,
                                                       float* w)
{
    const uint32_t k0 = (uint32_t)c.seed, k1 = (uint32_t)(c.seed >> 32);
    for (uint32_t j = blockIdx.x; j < c.n_dst; j += gridDim.x) {
        for (uint32_t k = threadIdx.x; k < c.K; k += kBlock) {
            const U4 r = philox(U4{ j, k, NAER_STREAM_CONN, c.conn_index }, k0, k1);
            const uint32_t src = c.src_begin + below(r.x, c.n_src);
            if (!PLACE) {
                atomicAdd(&deg[src], 1u);
            } else {
                const uint32_t del = c.dmin + c.dq * below(r.y, c.dchoices);
                const unsigned long long pos = atomicAdd(&cursor[src], 1ull);
                col[pos] = (c.dst_begin + j) | (del << 24) | c.plastic_bit;
                w[pos] = draw_weight(c, r.z);
            }
        }
    }
}

template <bool PLACE>
__global__ __launch_bounds__(kBlock) void k_conn_bernoulli(const ConnArgs c, uint32_t* deg,
                                                           unsigned long long* cursor,
                                                           uint32_t* col, float* w)
{
    const uint32_t seed32 = (uint32_t)c.seed;
    for (uint32_t s = blockIdx.x; s < c.n_src; s += gridDim.x) {
        uint32_t cnt = 0;
        for (uint32_t j = threadIdx.x; j < c.n_dst; j += kBlock) {
            const U4 r = philox(U4{ j, 0u, 0u, 0u }, s * 32u + j % 32u, seed32);
            if (!((double)r.z / 4294967295.0 < c.p))
                continue;
            if (!PLACE) {
                ++cnt;
            } else {
                const uint32_t del = c.dmin + c.dq * below(r.y, c.dchoices);
                const unsigned long long pos = atomicAdd(&cursor[c.src_begin + s], 1ull);
                col[pos] = (c.dst_begin + j) | (del << 24) | c.plastic_bit;
                w[pos] = draw_weight(c, r.w);
            }
        }
        if (!PLACE && cnt)
            atomicAdd(&deg[c.src_begin + s], cnt);
    }
}

// Canonical row order: ascending (row_key(word) << 32 | weight bits), bitonic
// sort of one row per block in LDS.  row_key orders a row by the target's
// 2^14-neuron group within its network first, then by the word (plastic bit,
// delay, target): the bucketed placement keys events by (arrival batch,
// group), so a wave's 64 consecutive synapses form runs of equal keys (one
// LDS atomic and one contiguous store run each) instead of 64 scattered ones.
// net_base: the network's first neuron, so the order is a function of the
// network alone (not of where a pack or an arena places it).
__device__ __forceinline__ uint32_t row_key(uint32_t word, uint32_t net_base)
{
    const uint32_t local = (word & kTargetMask) - net_base;
    return ((local >> kGroupNeuronShift) << 22) | ((word >> 31) << 21) | (((word >> 24) & 0x7Fu) << 14) |
           (local & ((1u << kGroupNeuronShift) - 1u));
}

__device__ __forceinline__ uint32_t row_word(uint32_t key, uint32_t net_base)
{
    const uint32_t local = ((key >> 22) << kGroupNeuronShift) | (key & ((1u << kGroupNeuronShift) - 1u));
    return (local + net_base) | (((key >> 14) & 0x7Fu) << 24) | (((key >> 21) & 1u) << 31);
}

__global__ __launch_bounds__(512) void k_sort_rows(const unsigned long long* row_ptr, uint32_t* col,
                                                   float* w, uint32_t row_begin, uint32_t row_end,
                                                   uint32_t net_base)
{
    extern __shared__ unsigned long long s_keys[];
    for (uint32_t r = row_begin + blockIdx.x; r < row_end; r += gridDim.x) {
        const unsigned long long beg = row_ptr[r];
        const uint32_t len = (uint32_t)(row_ptr[r + 1] - beg);
        if (len <= 1)
            continue;
        uint32_t L = 1;
        while (L < len)
            L <<= 1;
        for (uint32_t i = threadIdx.x; i < L; i += blockDim.x)
            s_keys[i] = i < len ? ((unsigned long long)row_key(col[beg + i], net_base) << 32) |
                                      (unsigned long long)__float_as_uint(w[beg + i])
                                : ~0ull;
        __syncthreads();
        for (uint32_t k = 2; k <= L; k <<= 1) {
            for (uint32_t jj = k >> 1; jj > 0; jj >>= 1) {
                for (uint32_t i = threadIdx.x; i < L; i += blockDim.x) {
                    const uint32_t ixj = i ^ jj;
                    if (ixj > i) {
                        const unsigned long long x = s_keys[i], y = s_keys[ixj];
                        const bool up = (i & k) == 0;
                        if ((x > y) == up) {
                            s_keys[i] = y;
                            s_keys[ixj] = x;
                        }
                    }
                }
                __syncthreads();
            }
        }
        for (uint32_t i = threadIdx.x; i < len; i += blockDim.x) {
            col[beg + i] = row_word((uint32_t)(s_keys[i] >> 32), net_base);
            w[beg + i] = __uint_as_float((uint32_t)s_keys[i]);
        }
        __syncthreads();
    }
}

// Plastic synapses of the synapse range [begin, end), added to *count.
__global__ __launch_bounds__(kBlock) void k_count_plastic(const uint32_t* col, unsigned long long begin,
                                                          unsigned long long end, unsigned long long* count)
{
    uint32_t k = 0;
    for (unsigned long long i = begin + (unsigned long long)blockIdx.x * kBlock + threadIdx.x; i < end;
         i += (unsigned long long)gridDim.x * kBlock)
        k += (col[i] & kPlasticBit) ? 1u : 0u;
    for (uint32_t off = kWave / 2; off > 0; off >>= 1)
        k += __shfl_down(k, off);
    if (lane_id() == 0 && k)
        atomicAdd(count, (unsigned long long)k);
}

// Synapses of [0, S) per (delay d - dmin, target group): out[(d - dmin) * ng + g]
// (nd * ng bins, at most kSynStatsBins, counted in LDS and added once per block).
__global__ __launch_bounds__(kBlock) void k_syn_stats(const uint32_t* col, unsigned long long S, uint32_t dmin,
                                                      uint32_t nd, uint32_t ng, unsigned long long* out)
{
    extern __shared__ uint32_t s_bin[];
    const uint32_t nb = nd * ng;
    for (uint32_t i = threadIdx.x; i < nb; i += kBlock)
        s_bin[i] = 0;
    __syncthreads();
    for (unsigned long long i = (unsigned long long)blockIdx.x * kBlock + threadIdx.x; i < S;
         i += (unsigned long long)gridDim.x * kBlock) {
        const uint32_t c = col[i];
        const uint32_t d = ((c >> 24) & 0x7Fu) - dmin;
        const uint32_t g = (c & kTargetMask) >> kGroupNeuronShift;
        if (d < nd && g < ng)
            atomicAdd(&s_bin[d * ng + g], 1u);
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < nb; i += kBlock)
        if (s_bin[i])
            atomicAdd(&out[i], (unsigned long long)s_bin[i]);
}

// Initial state of neurons [j_begin, j_end): potentials from the V0 stream,
// empty refractory/trace/POISSON state and cleared ring cells.
template <bool DET>
__global__ __launch_bounds__(kBlock) void k_init_state(const UpdateArgs a, uint32_t j_begin, uint32_t j_end)
{
    const uint32_t j = j_begin + blockIdx.x * kBlock + threadIdx.x;
    if (j >= j_end)
        return;
    float v = 0.0f;
    if (a.n_pops) {
        const uint32_t p = find_pop(a.pops, a.n_pops, a.blk_pop[j >> a.tw_shift] & ~kTileIdle, j);
        const PopParams pp = a.pops[p];
        if ((pp.kind == NAER_POP_LIF || pp.kind == NAER_POP_MODEL) && j >= pp.begin && j < pp.end) {
            const U4 r = philox(U4{ j - pp.net_base, 0u, NAER_STREAM_V0, 0u }, pp.seed_lo, pp.seed_hi);
            const float span = pp.v0_max - pp.v0_min;
            const float x = span * unit24(r.x);
            v = pp.v0_min + x;
        }
    }
    a.V[j] = v;
    a.ref[j] = 0;
    a.last[j] = -1;
    if (a.sd.lat) {
        a.sd.flushed[j] = 0u;
        a.sd.lat[j] = make_uint4(0u, 0u, 0u, 0u);
        a.sd.pst[2 * (size_t)j] = make_uint4(kNoSpike, 0u, kNoSpike, 0u);
        a.sd.pst[2 * (size_t)j + 1] = make_uint4(kNoSpike, kNoSpike, 0u, 0u);
    }
    a.pnext[j] = 0;
    a.pcnt[j] = 0;
    using RingT = typename std::conditional<DET, long long, float>::type;
    RingT* ring = (RingT*)a.ring;
    for (uint32_t sl = 0; sl < a.D; ++sl)
        ring[(size_t)sl * a.N + j] = (RingT)0;
}

// ---------------------------------------------------------------------------
// Generators exposed for parity checks
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_philox(const uint32_t* ctr, const uint32_t* key,
                                                   uint32_t* out, unsigned long long n)
{
    const unsigned long long i = (unsigned long long)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n)
        return;
    const U4 r = philox(U4{ ctr[4 * i], ctr[4 * i + 1], ctr[4 * i + 2], ctr[4 * i + 3] },
                        key[2 * i], key[2 * i + 1]);
    out[4 * i] = r.x;
    out[4 * i + 1] = r.y;
    out[4 * i + 2] = r.z;
    out[4 * i + 3] = r.w;
}

__global__ __launch_bounds__(kBlock) void k_ref_poisson_epoch(uint32_t n, uint32_t n_pois,
                                                              uint32_t seed, double neg_rate,
                                                              uint32_t t0, uint32_t t1,
                                                              uint32_t* times, uint32_t* count)
{
    const uint32_t idx = blockIdx.x * kBlock + threadIdx.x;
    if (idx >= n)
        return;
    const uint32_t gid = ref_poisson_gid(idx);
    uint32_t time = t0 + ref_isi_us(neg_rate, gid, seed, t0 * n_pois + idx);
    uint32_t c = 0;
    for (uint32_t m = 0; m < kMaxPoissonPerEpoch; ++m) {
        if (time < t1) {
            times[(size_t)kMaxPoissonPerEpoch * idx + c++] = time;
            time += ref_isi_us(neg_rate, gid, seed, time * n_pois + idx);
        }
    }
    count[idx] = c;
}

// Fills n words with v (clears memory without hipMemset, which a captured
// graph must not contain).
__global__ __launch_bounds__(kBlock) void k_fill_u32(uint32_t* p, unsigned long long n, uint32_t v)
{
    for (unsigned long long i = (unsigned long long)blockIdx.x * kBlock + threadIdx.x; i < n;
         i += (unsigned long long)gridDim.x * kBlock)
        p[i] = v;
}

// flags[c] = 1 when a neuron of [256 c, 256 c + 256) that belongs to a
// population of an active network holds a potential that is not finite
// (free ranges of an arena are not looked at).
__global__ __launch_bounds__(kBlock) void k_check_finite(const UpdateArgs a, uint32_t* flags)
{
    const uint32_t j = blockIdx.x * kBlock + threadIdx.x;
    bool bad = false;
    if (j < a.N && a.n_pops) {
        const uint32_t p = find_pop(a.pops, a.n_pops, a.blk_pop[j >> a.tw_shift] & ~kTileIdle, j);
        bad = j >= a.pops[p].begin && j < a.pops[p].end && !isfinite(a.V[j]);
    }
    const int any = __syncthreads_or(bad ? 1 : 0);
    if (threadIdx.x == 0)
        flags[blockIdx.x] = any ? 1u : 0u;
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
static inline uint32_t cdiv(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

// Raising a kernel's dynamic-LDS limit: only ever raised, under a lock, so
// nets stepped from several host threads never lower it under each other.
static std::mutex g_lds_mu;

template <bool DET, bool PLASTIC, bool BUCKET, bool EXTRA, uint32_t TW, bool REC4>
static hipError_t launch_update_t(const UpdateArgs& a, hipStream_t s)
{
    const size_t lds = BUCKET ? (size_t)a.B * TW * sizeof(unsigned long long) : 0;
    static size_t lds_set = 0;  // per instantiation: largest size allowed so far
    if (lds > 32768) {
        std::lock_guard<std::mutex> lock(g_lds_mu);
        if (lds > lds_set) {
            hipError_t e = hipFuncSetAttribute((const void*)k_update<DET, PLASTIC, BUCKET, EXTRA, TW, REC4>,
                                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            if (e != hipSuccess)
                return e;
            lds_set = lds;
        }
    }
    hipLaunchKernelGGL((k_update<DET, PLASTIC, BUCKET, EXTRA, TW, REC4>), dim3(a.n_act), dim3(TW), lds, s, a);
    return hipGetLastError();
}

template <bool DET, bool PLASTIC, bool BUCKET, bool REC4>
static hipError_t launch_update_tw(const UpdateArgs& a, bool extra, hipStream_t s)
{
    if (a.tw_shift == 10)
        return extra ? launch_update_t<DET, PLASTIC, BUCKET, true, 1024, REC4>(a, s)
                    : launch_update_t<DET, PLASTIC, BUCKET, false, 1024, REC4>(a, s);
    if (a.tw_shift == 9)
        return extra ? launch_update_t<DET, PLASTIC, BUCKET, true, 512, REC4>(a, s)
                    : launch_update_t<DET, PLASTIC, BUCKET, false, 512, REC4>(a, s);
    return extra ? launch_update_t<DET, PLASTIC, BUCKET, true, 256, REC4>(a, s)
                : launch_update_t<DET, PLASTIC, BUCKET, false, 256, REC4>(a, s);
}

hipError_t launch_update(const UpdateArgs& a, bool det, bool plastic, bool bucket, bool extra, bool rec4,
                         hipStream_t s)
{
    if (a.tw_shift < 8 || a.tw_shift > 10 || (rec4 && (plastic || !bucket)))
        return hipErrorInvalidValue;
    if (rec4)
        return det ? launch_update_tw<true, false, true, true>(a, extra, s)
                   : launch_update_tw<false, false, true, true>(a, extra, s);
    if (plastic && bucket)
        return det ? launch_update_tw<true, true, true, false>(a, extra, s)
                   : launch_update_tw<false, true, true, false>(a, extra, s);
    if (plastic)
        return det ? launch_update_tw<true, true, false, false>(a, extra, s)
                   : launch_update_tw<false, true, false, false>(a, extra, s);
    if (bucket)
        return det ? launch_update_tw<true, false, true, false>(a, extra, s)
                   : launch_update_tw<false, false, true, false>(a, extra, s);
    return det ? launch_update_tw<true, false, false, false>(a, extra, s)
               : launch_update_tw<false, false, false, false>(a, extra, s);
}

template <bool AGG>
static hipError_t hist_lds_attributes(size_t lds)
{
    hipError_t e = hipFuncSetAttribute((const void*)k_bucket_hist<4, AGG>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e == hipSuccess)
        e = hipFuncSetAttribute((const void*)k_bucket_hist<8, AGG>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)lds);
    return e;
}

static hipError_t bucket_lds_attributes(const BucketArgs& a)
{
    const size_t lds = (size_t)a.NBK * sizeof(uint32_t);
    static size_t lds_set = 0;
    std::lock_guard<std::mutex> lock(g_lds_mu);
    if (lds > 32768 && lds > lds_set) {
        hipError_t e = hist_lds_attributes<false>(lds);
        if (e == hipSuccess)
            e = hist_lds_attributes<true>(lds);
        if (e == hipSuccess)
            e = hipFuncSetAttribute((const void*)k_bucket_starts, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)lds);
        if (e != hipSuccess)
            return e;
        lds_set = lds;
    }
    return hipSuccess;
}

template <bool DET, uint32_t U, bool AGG>
static void launch_place(const BucketArgs& a, bool plastic, size_t lds_c, hipStream_t s)
{
    if (plastic)
        hipLaunchKernelGGL((k_bucket_place<DET, U, AGG, true, false>), dim3(a.NB), dim3(kBucketThreads), lds_c, s,
                           a);
    else if (a.ncls)
        hipLaunchKernelGGL((k_bucket_place<DET, U, AGG, false, true>), dim3(a.NB), dim3(kBucketThreads), lds_c, s,
                           a);
    else
        hipLaunchKernelGGL((k_bucket_place<DET, U, AGG, false, false>), dim3(a.NB), dim3(kBucketThreads), lds_c, s,
                           a);
}

template <bool AGG>
static void launch_bucket_pass(const BucketArgs& a, int step, bool det, size_t lds, size_t lds_c, hipStream_t s)
{
    const dim3 grid(a.NB), block(kBucketThreads);
    switch (step) {
    case 0:
        if (a.balance)
            hipLaunchKernelGGL(k_bucket_split, dim3(1), dim3(1024), 0, s, a);
        if (a.reserve)
            break;  // reserved placement counts nothing
        if (a.U == 4)
            hipLaunchKernelGGL((k_bucket_hist<4, AGG>), grid, block, lds, s, a);
        else
            hipLaunchKernelGGL((k_bucket_hist<8, AGG>), grid, block, lds, s, a);
        break;
    case 2: {
        // placement counters aggregate by their own flag
        const bool pl = a.sd.lat != nullptr;
        if (a.AGP) {
            if (det)
                a.U == 4 ? launch_place<true, 4, true>(a, pl, lds_c, s) : launch_place<true, 8, true>(a, pl, lds_c, s);
            else
                a.U == 4 ? launch_place<false, 4, true>(a, pl, lds_c, s) : launch_place<false, 8, true>(a, pl, lds_c, s);
        } else {
            if (det)
                a.U == 4 ? launch_place<true, 4, false>(a, pl, lds_c, s) : launch_place<true, 8, false>(a, pl, lds_c, s);
            else
                a.U == 4 ? launch_place<false, 4, false>(a, pl, lds_c, s) : launch_place<false, 8, false>(a, pl, lds_c, s);
        }
        break;
    }
    default:
        if (a.direct)
            break;  // the placement wrote the tile buckets
        if (a.reserve) {
            const dim3 g((a.NBK >> a.gt_shift) * a.RS);
            if (a.ncls)
                hipLaunchKernelGGL(k_bucket_refine_res<true>, g, dim3(kResThreads), 0, s, a);
            else
                hipLaunchKernelGGL(k_bucket_refine_res<false>, g, dim3(kResThreads), 0, s, a);
            break;
        }
        if (a.ncls)
            hipLaunchKernelGGL((k_bucket_refine<AGG, true>), dim3((a.NBK >> a.gt_shift) * a.RS), dim3(kRefineThreads),
                               0, s, a);
        else
            hipLaunchKernelGGL((k_bucket_refine<AGG, false>), dim3((a.NBK >> a.gt_shift) * a.RS),
                               dim3(kRefineThreads), 0, s, a);
        break;
    }
}

hipError_t launch_bucket_step(const BucketArgs& a, int step, bool det, hipStream_t s)
{
    hipError_t e = bucket_lds_attributes(a);
    if (e != hipSuccess)
        return e;
    const size_t lds = (size_t)a.NBK * sizeof(uint32_t);
    // reserved placement: each counter with its region's end and start and
    // its spill region
    const size_t lds_c = (size_t)a.n_cnt * sizeof(uint32_t) * (a.reserve ? kResCounterWords : 1);
    switch (step) {
    case 1:
        if (a.reserve)
            break;
        hipLaunchKernelGGL(k_bucket_colscan,
                           dim3((a.direct ? 0u : cdiv(a.NBK >> a.gt_shift, kScanCols)) + cdiv(a.NBK, kScanCols)),
                           dim3(kScanCols * kScanGroups), 0, s, a);
        hipLaunchKernelGGL(k_bucket_starts, dim3(1), dim3(1024), lds, s, a);
        break;
    case 0:
    case 2:
    case 3:
        if (a.AG)
            launch_bucket_pass<true>(a, step, det, lds, lds_c, s);
        else
            launch_bucket_pass<false>(a, step, det, lds, lds_c, s);
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_bucket_scrub(const UpdateArgs& a, uint32_t j_begin, uint32_t j_end, hipStream_t s)
{
    if (j_end <= j_begin)
        return hipSuccess;
    const uint32_t t0 = j_begin >> a.tw_shift, t1 = (j_end - 1) >> a.tw_shift;
    hipLaunchKernelGGL(k_bucket_scrub, dim3(t1 - t0 + 1, a.R * a.M), dim3(kBlock), 0, s, a, j_begin, j_end);
    return hipGetLastError();
}

hipError_t launch_scatter(const ScatterArgs& a, bool det, bool plastic, uint32_t grid_x,
                          uint32_t chunks, hipStream_t s)
{
    const dim3 grid(grid_x, chunks);
    if (det) {
        if (plastic)
            hipLaunchKernelGGL((k_scatter<true, true>), grid, dim3(kBlock), 0, s, a);
        else
            hipLaunchKernelGGL((k_scatter<true, false>), grid, dim3(kBlock), 0, s, a);
    } else {
        if (plastic)
            hipLaunchKernelGGL((k_scatter<false, true>), grid, dim3(kBlock), 0, s, a);
        else
            hipLaunchKernelGGL((k_scatter<false, false>), grid, dim3(kBlock), 0, s, a);
    }
    return hipGetLastError();
}

hipError_t launch_check_finite(const UpdateArgs& a, uint32_t* flags, uint32_t n_check, hipStream_t s)
{
    if (a.N == 0)
        return hipSuccess;
    if (n_check == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_check_finite, dim3(cdiv(n_check, kBlock)), dim3(kBlock), 0, s, a, flags);
    return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t* p, unsigned long long n, uint32_t v, hipStream_t s)
{
    if (n == 0)
        return hipSuccess;
    const uint32_t grid = (uint32_t)std::min<unsigned long long>(4096, (n + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_fill_u32, dim3(grid), dim3(kBlock), 0, s, p, n, v);
    return hipGetLastError();
}

hipError_t launch_clock_advance(DevState* st, uint32_t b, hipStream_t s)
{
    hipLaunchKernelGGL(k_clock_advance, dim3(1), dim3(kWave), 0, s, st, b);
    return hipGetLastError();
}

hipError_t launch_flush(const FlushArgs& a, uint32_t grid, hipStream_t s)
{
    hipLaunchKernelGGL(k_flush, dim3(grid), dim3(kBlock), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_conn(const ConnArgs& c, uint32_t rule, bool place, uint32_t* deg,
                       unsigned long long* cursor, uint32_t* col, float* w, hipStream_t s)
{
    if (rule == NAER_RULE_FIXED_INDEGREE) {
        const dim3 grid(c.n_dst < 65536u ? (c.n_dst ? c.n_dst : 1u) : 65536u);
        if (place)
            hipLaunchKernelGGL((k_conn_fixed<true>), grid, dim3(kBlock), 0, s, c, deg, cursor, col, w);
        else
            hipLaunchKernelGGL((k_conn_fixed<false>), grid, dim3(kBlock), 0, s, c, deg, cursor, col, w);
    } else {
        const dim3 grid(c.n_src < 65536u ? (c.n_src ? c.n_src : 1u) : 65536u);
        if (place)
            hipLaunchKernelGGL((k_conn_bernoulli<true>), grid, dim3(kBlock), 0, s, c, deg, cursor, col, w);
        else
            hipLaunchKernelGGL((k_conn_bernoulli<false>), grid, dim3(kBlock), 0, s, c, deg, cursor, col, w);
    }
    return hipGetLastError();
}

// Rows [row_begin, row_end) are one network's neurons: row_begin is its base.
hipError_t launch_sort_rows(const unsigned long long* row_ptr, uint32_t* col, float* w, uint32_t row_begin,
                            uint32_t row_end, uint32_t max_row, hipStream_t s)
{
    if (row_end <= row_begin || max_row <= 1)
        return hipSuccess;
    uint32_t L = 1;
    while (L < max_row)
        L <<= 1;
    const size_t lds = (size_t)L * sizeof(unsigned long long);
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute((const void*)k_sort_rows,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess)
            return e;
    }
    const uint32_t rows = row_end - row_begin;
    const dim3 grid(rows < 65536u ? rows : 65536u);
    hipLaunchKernelGGL(k_sort_rows, grid, dim3(512), lds, s, row_ptr, col, w, row_begin, row_end, row_begin);
    return hipGetLastError();
}

hipError_t launch_count_plastic(const uint32_t* col, unsigned long long begin, unsigned long long end,
                                unsigned long long* count, hipStream_t s)
{
    const uint32_t grid = std::max<uint32_t>(1, std::min<uint64_t>(4096, cdiv(end - begin, kBlock)));
    hipLaunchKernelGGL(k_count_plastic, dim3(grid), dim3(kBlock), 0, s, col, begin, end, count);
    return hipGetLastError();
}

hipError_t launch_syn_stats(const uint32_t* col, unsigned long long S, uint32_t dmin, uint32_t nd, uint32_t ng,
                            unsigned long long* out, hipStream_t s)
{
    if (nd * ng > kSynStatsBins || S == 0)
        return hipErrorInvalidValue;
    const uint32_t grid = std::max<uint32_t>(1, std::min<uint64_t>(2048, cdiv(S, kBlock * 64)));
    hipLaunchKernelGGL(k_syn_stats, dim3(grid), dim3(kBlock), nd * ng * sizeof(uint32_t), s, col, S, dmin, nd, ng, out);
    return hipGetLastError();
}

hipError_t launch_init_state(const UpdateArgs& a, bool det, uint32_t j_begin, uint32_t j_end, hipStream_t s)
{
    if (j_end <= j_begin)
        return hipSuccess;
    const dim3 grid(cdiv(j_end - j_begin, kBlock));
    if (det)
        hipLaunchKernelGGL((k_init_state<true>), grid, dim3(kBlock), 0, s, a, j_begin, j_end);
    else
        hipLaunchKernelGGL((k_init_state<false>), grid, dim3(kBlock), 0, s, a, j_begin, j_end);
    return hipGetLastError();
}

hipError_t launch_philox(const uint32_t* ctr, const uint32_t* key, uint32_t* out,
                         unsigned long long n, hipStream_t s)
{
    hipLaunchKernelGGL(k_philox, dim3(cdiv(n, kBlock)), dim3(kBlock), 0, s, ctr, key, out, n);
    return hipGetLastError();
}

hipError_t launch_ref_poisson_epoch(uint32_t n, uint32_t n_pois, uint32_t seed, double neg_rate,
                                    uint32_t t0, uint32_t t1, uint32_t* times, uint32_t* count,
                                    hipStream_t s)
{
    hipLaunchKernelGGL(k_ref_poisson_epoch, dim3(cdiv(n, kBlock)), dim3(kBlock), 0, s, n, n_pois,
                       seed, neg_rate, t0, t1, times, count);
    return hipGetLastError();
}

}  // namespace naer
