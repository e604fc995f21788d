// naer_device.hpp — device-side building blocks and the kernel argument
// records shared by the kernels (kernels.hip) and the host runtime
// (naer_runtime.hip).  gfx950 only.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/naer/naer.h"

namespace naer {

constexpr uint32_t kStagePops = 4;     // drive tables of a block's first populations staged in LDS
constexpr uint32_t kMaxBatch = 16;     // steps per update launch
constexpr uint32_t kMaxBatchLargeTile = 9;  // ... with 1024-neuron tiles
constexpr uint32_t kBlock = 256;       // threads per block of the step kernels
constexpr uint32_t kWave = 64;         // CDNA wavefront
constexpr uint32_t kLdsSpikes = 256;   // per-block spike staging in LDS
constexpr uint32_t kChunkUnroll = 4;   // synapses per lane per scatter task
constexpr uint32_t kChunk = kWave * kChunkUnroll;
constexpr uint32_t kBucketUnroll = 8;  // default synapses per lane per task of the bucketed passes
constexpr uint32_t kTargetMask = 0x00FFFFFFu;
constexpr uint32_t kPlasticBit = 0x80000000u;
constexpr uint32_t kEpochUs = 1500u;   // reference epoch, naer_types.h:4
constexpr uint32_t kMaxPoissonPerEpoch = 10u;  // input_neuron_group.cl:30

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon et al., SC'11), the generator behind the reference's
// uniform_rand (random.cl:8-16).  Round: (x0,x1,x2,x3) ->
// (hi(M1 x2) ^ x1 ^ k0, lo(M1 x2), hi(M0 x0) ^ x3 ^ k1, lo(M0 x0)).
// ---------------------------------------------------------------------------
struct U4 {
    uint32_t x, y, z, w;
};

__host__ __device__ __forceinline__ uint32_t mulhi32(uint32_t a, uint32_t b)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __umulhi(a, b);
#else
    return (uint32_t)(((uint64_t)a * b) >> 32);
#endif
}

__host__ __device__ __forceinline__ U4 philox(U4 c, uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        // one 32x32->64 product per multiplier (v_mad_u64_u32 on the device)
        const uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
        const uint32_t h0 = (uint32_t)(p0 >> 32), l0 = (uint32_t)p0;
        const uint32_t h1 = (uint32_t)(p1 >> 32), l1 = (uint32_t)p1;
        c = U4{ h1 ^ c.y ^ k0, l1, h0 ^ c.w ^ k1, l0 };
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return c;
}

__host__ __device__ __forceinline__ uint32_t word_of(const U4& r, uint32_t i)
{
    return i == 0 ? r.x : (i == 1 ? r.y : (i == 2 ? r.z : r.w));
}

// fp32 uniform on the 2^-24 grid in [0,1)
__host__ __device__ __forceinline__ float unit24(uint32_t w) { return (float)(w >> 8) * 0x1p-24f; }
// index in [0,n) by multiply-shift
__host__ __device__ __forceinline__ uint32_t below(uint32_t w, uint32_t n) { return mulhi32(w, n); }

// (uint)exprnd_fast(rate, seed, iteration), random.cl:6-44 with the cast of
// input_neuron_group.cl:29,37: third Philox word over UINT_MAX in fp64,
// log(1-u) / -(rate/1e6), truncated; out-of-range values saturate.
__device__ __forceinline__ uint32_t ref_isi_us(double neg_rate_per_us, uint32_t gid, uint32_t seed,
                                               uint32_t iteration)
{
    const U4 r = philox(U4{ iteration, 0u, 0u, 0u }, gid, seed);
    const double u = (double)r.z / 4294967295.0;
    const double isi = log(1.0 - u) / neg_rate_per_us;
    return isi < 4294967296.0 ? (uint32_t)isi : 0xFFFFFFFFu;
}

// get_global_id(0) of POISSON neuron idx under the reference launch shape
// (128 work-items per 2048 neurons, Network.cpp:181-184).
__host__ __device__ __forceinline__ uint32_t ref_poisson_gid(uint32_t idx)
{
    return (idx / 2048u) * 128u + idx % 128u;
}

// ---------------------------------------------------------------------------
// Records in device memory
// ---------------------------------------------------------------------------
struct PopParams {
    uint32_t begin, end;  // neuron range [begin, end) in the pack
    uint32_t kind;        // NAER_POP_*
    uint32_t pois_base;   // POISSON: index of `begin` among the POISSON neurons of its net
    float P;              // exp(-dt/tau_m)
    float v_th, v_reset, v_rest;
    float v0_min, v0_max;  // initial potential range
    float w_ext;          // drive jump
    int32_t n_ref;        // refractory steps
    uint32_t has_drive;
    uint32_t cdf_offset;  // drive CDF table offset (floats)
    double neg_rate_per_us;  // POISSON: -(rate/1e6)
    // the network the population belongs to: random draws and clocks are
    // local to it, so a network gives the same results alone or packed
    uint32_t net_base;    // first neuron of the net in the pack
    uint32_t n_pois;      // POISSON neurons of the net
    uint32_t seed_lo, seed_hi;
    uint32_t pois_seed;
    uint32_t _pad;
    uint32_t t_start_lo, t_start_hi;  // pack step at which the net's local time is 0
};

// Step bookkeeping that lives on the device so that launches (and captured
// graphs) need no per-step host arguments.
struct DevState {
    unsigned long long clock;   // first step of the next batch
    unsigned long long t0_cur;  // first step of the batch in flight (set by the update)
    unsigned long long rec_n;   // spikes offered to the record buffer since the last drain
    uint32_t batch;             // launches done (the update writes the lists of aligned batch A to list[A & 1])
    uint32_t cur;               // list of the batch in flight
    uint32_t overflow;          // spikes lost to a full list segment
    uint32_t window;            // plastic nets: steps within which no neuron fires kSpikeHist times
    unsigned long long t_flush; // plastic nets: step of the last flush (flushes come every window / 2 steps)
};

// Pair STDP with potentiation applied lazily (as update_synapse does at the
// next pre-synaptic input, stdp.cl:19-26): a plastic synapse i -> j is
// touched only when i fires, and then first takes the potentiations of j's
// spikes since i's previous spike (or the last flush), in time order, then
// its depression.  Every neuron keeps its last kSpikeHist spikes; a flush
// (all plastic synapses) runs before any neuron can have fired kSpikeHist - 1
// times since the last one.  Traces in closed form: the trace of neuron j at
// step t (spikes before t) is (a_s + 1) * dec^(t - s) for its last spike s < t,
// a_s = its trace at s; dec^k from a table (0 once below 2^-40).
constexpr uint32_t kSpikeHist = 32;

// Flag in the per-tile first-population table: no population overlaps the tile.
constexpr uint32_t kTileIdle = 0x80000000u;

constexpr uint32_t kNoSpike = 0x80000000u;  // step of an absent spike

struct StdpArgs {
    // per neuron, its newest spike s1 and the one before, s0 (kNoSpike if absent)
    uint4* lat;          // [N] pre side: {spikes so far, s1, pre trace a_s(s1) bits, pre trace a_s(s0) bits}
    uint4* pst;          // [N][2] post side: {s1, post trace a_s(s1) bits, s0, post trace a_s(s0) bits},
                         // {s2, s3 (the two spikes before s0), spikes so far, 0}: one 32-byte line part
    uint32_t* hist;      // [N][kSpikeHist] step of spike k at k % kSpikeHist
    uint32_t* flushed;   // [N] row i's potentiation is applied for every post spike before flushed[i]
    const float* tp;     // [n_tp] exp(-k dt / tau_plus)
    const float* tm;     // [n_tm] exp(-k dt / tau_minus)
    uint32_t n_tp, n_tm;
    float a_plus, a_minus, w_min, w_max;
};

// Spike lists are segmented by update block: block b of the update kernel
// owns list[b * seg_cap, (b+1) * seg_cap) and seg_count[b], so compaction
// needs no global atomics.  Per-block spike and event totals accumulate in
// spk_part / ev_part (each block owns its entry).
struct UpdateArgs {
    float* V;
    int32_t* ref;
    int32_t* last;
    StdpArgs sd;                       // plastic nets
    uint32_t* pnext;
    uint32_t* pcnt;
    void* ring;
    const unsigned long long* row_ptr;
    naer_spike* list0;
    naer_spike* list1;
    uint32_t* seg_count;               // [2][n_seg]
    unsigned long long* seg_ev;        // [2][n_seg] synaptic events of the batch's spikes per segment
    unsigned long long* spk_part;      // [n_seg]
    unsigned long long* ev_part;       // [n_seg]
    naer_spike* rec;
    DevState* st;
    const float* cdf;                  // NAER_DRIVE_TABLE floats per population
    const PopParams* pops;             // [n_pops]
    const uint32_t* blk_pop;           // first population of each update block (| kTileIdle)
    float* probe;
    unsigned long long rec_cap;
    unsigned long long probe_t0;
    uint32_t probe_n;
    uint32_t seg_cap;
    uint32_t n_seg;
    uint32_t n_act;  // tiles launched: [0, n_act) holds every neuron of a population (<= n_seg)
    uint32_t N, D;
    uint32_t B;
    uint32_t dt_us;
    uint32_t n_pops;
    uint32_t tw_shift;   // update block = one tile of 1 << tw_shift neurons
    uint32_t Bfull;      // steps of an aligned batch
    uint32_t defer;      // 1: launches within an aligned batch append to its spike lists, which are
                         // delivered once the batch is complete (bucketed static nets)
    // bucketed delivery (see BucketArgs): events of earlier batches sorted by
    // (arrival batch offset m, target tile)
    const uint2* ev;          // [R][e_cap]
    const uint32_t* bs;       // [R][NBK + 1] bucket starts
    uint32_t* flags;          // [R] bit 0: that batch's events are in the buckets; bit 1 (reserved
                              // placement): some of them went through the ring
    unsigned long long e_cap;
    uint32_t R, M, NT, NBK;
    // reserved placement (see BucketArgs): the tile ranges of each part of a
    // group bucket, for batches whose flags have bit 2 set; the update clears
    // the flags of the batch it starts when `reserve`
    const uint32_t* tb;       // [R][groups][parts][tiles per group + 1]
    uint32_t reserve, parts, gt_shift;
    // prefetch (reserved placement): the rows of the batch's spikes are
    // loaded into the caches as the spikes happen (the placement reads them
    // next)
    const uint32_t* col;
    uint32_t prefetch;
    // neurons of source-defined models (NAER_POP_MODEL): their synaptic input
    // and drive per step, [B][N], read by the model's own kernel
    float2* cin;
    // class-weight delivery (4-byte records, see kRec4*): the weight of class
    // c (the source's population) in 32.32 fixed point
    const unsigned long long* cls_fx;
    uint32_t n_cls;
};

// Class-weight delivery.  When every population's outgoing synapses carry
// one constant weight (all static networks of the benchmark workloads), a
// bucketed event needs no weight bits: a 4-byte record
//   target offset (bits 0-13) | arrival slot (14-17) | class (18-31)
// names the source's population, whose weight the update looks up.  The
// target offset is within the 2^14-neuron group in the staging buffer and
// within the tile in the event buffers.  The placement then reads no
// weights, and every pass moves half the bytes of the 8-byte records
// {offset | slot << 16, weight bits}.
constexpr uint32_t kRec4SlotShift = 14;
constexpr uint32_t kRec4ClassShift = 18;
constexpr uint32_t kRec4OffsetMask = (1u << kRec4SlotShift) - 1u;
constexpr uint32_t kRec4MaxClasses = 1u << (32 - kRec4ClassShift);
constexpr uint32_t kRec4StagedClasses = 64;  // class weights the update keeps in LDS

// Event-driven mode (kernels_event.hip): one 1500 us window per epoch.
constexpr uint32_t kEvLo = 4096;    // entries of a population's L decay table
constexpr uint32_t kEvSrcBits = 21; // an input key: time offset in the window (< 2^11) << 21 | source (< 2^21)
constexpr uint32_t kEvHiCap = 8192; // most entries of its H table
enum { kEvPoisson = 0, kEvRowLen, kEvExpand, kEvProcess };
// stdp.cl / stdp_group.cl / basic_stdp.cl constants of the event-driven STDP
constexpr uint32_t kEvHist = 10;          // NAER_HIST_LENGTH
constexpr uint32_t kEvHistUnit = 4;       // NAER_HIST_UNIT (us)
constexpr float kEvStdpTau = 20000.0f;    // NAER_T_STDP (us)
constexpr float kEvStdpAlpha = 2.02f;     // NAER_ALPHA
constexpr double kEvStdpLambda = 0.01;    // NAER_LAMBDA (a double literal there)
constexpr float kEvStdpWMax = 0.0003f;    // NAER_W_MAX
constexpr float kEvStdpConstDt = 100.0f;  // NAER_CONST_DT (us)

struct EvArgs {
    const PopParams* pops;
    const uint32_t* blk_pop;
    uint32_t n_pops, tw_shift, N;
    uint32_t t0;     // window start (us)
    uint32_t epoch;
    uint32_t in_cap; // capacity of an input list
    const unsigned long long* row_ptr;
    const uint32_t* col;
    const float* w;
    const uint32_t* in_keys;            // the window's inputs (time offset << kEvSrcBits | source), sorted
    uint32_t* out_keys;                 // input list being appended to (offsets from its window's start)
    uint32_t* n_out;                    // its count
    uint32_t* lost;                     // inputs beyond in_cap
    unsigned long long* row_len;        // [in_cap + 1]
    const unsigned long long* row_off;  // [in_cap + 1] exclusive scan of row_len
    uint32_t* ev_keys;                  // target << 11 | time offset (N < 2^21)
    float* ev_w;
    uint32_t n_ev;                      // events of the window
    uint32_t* ev_begin;                 // [N] each target's run of sorted events (zero when it has none)
    uint32_t* ev_end;
    // plastic nets (stdp.cl): every input carries the slot of its list entry
    // (sorted along with its key), whose payload a plastic neuron's spike
    // writes: its previous spike time and post trace; every event carries its
    // synapse (index | kPlasticBit for an STDP -> STDP synapse) and its input's slot
    const uint32_t* in_vals;            // slots of the sorted inputs
    uint32_t* out_vals;                 // slots of the list being appended to
    const uint32_t* in_prev;            // payload of the window's list, by slot
    const float* in_trace;
    uint32_t* out_prev;                 // payload of the list being appended to
    float* out_trace;
    uint32_t* ev_syn;
    uint32_t* ev_slot;
    float* w_mut;                       // the CSR weights the plastic events update
    float* trace;                       // [N] post trace (stdp_traces.stdp)
    uint32_t* hist;                     // [N][kEvHist / 2] inter-spike history, 16 bits per entry
    const uint32_t* pop_plastic;        // [n_pops] 1: a population with a plastic connection
    float* V;
    uint32_t* t_last;                   // [N] time of the last input taken (us)
    const double* dec_lo;               // [n_pops][kEvLo]
    const double* dec_hi;               // [n_pops][kEvHiCap]
    const uint32_t* n_hi;               // [n_pops]
    const uint32_t* t_ref;              // [n_pops] refractory period (us)
    naer_spike* rec;
    unsigned long long rec_cap;
    unsigned long long* rec_n;
    unsigned long long* spikes;
};

// Bucketed delivery of one aligned batch E (steps [E*Bfull, (E+1)*Bfull)):
// every synaptic event (target j, arrival step a = t + d) ends up in bucket
// k = (m - 1) * NT + T, m = a/Bfull - E in [1, M], T = the target's tile, as
// {j - T*TW | (a mod Bfull) << 16, weight bits}.  A group is 2^14 neurons
// (1 << gt_shift tiles); NT is the tile count rounded up to whole groups, so
// the buckets of a group (m, g) are adjacent.  Five kernels:
//   hist     per block: counts per bucket (LDS), written as a row of hist_f
//            and, summed per group, as a row of hist_c
//   colscan  hist_c columns: exclusive scan over blocks; hist_f columns: totals
//   starts   exclusive scan of the bucket totals -> bs[E % R], flags[E % R]
//   place    events to their group's range of the staging buffer (runs of a
//            block's events per group are contiguous, so stores coalesce)
//   refine   RS blocks per group: staging -> ev[E % R] in bucket order
// The update of batch E+m reads bucket (m, T) into LDS.
constexpr uint32_t kGroupNeuronShift = 14;  // neurons per group: 16384
constexpr uint32_t kMaxGroupTiles = 64;      // 16384 / the smallest tile (256)

struct BucketArgs {
    const naer_spike* list0;
    const naer_spike* list1;
    const uint32_t* seg_count;
    const unsigned long long* seg_ev;  // [2][n_seg]
    uint32_t* blk_seg;                 // [NB + 1]: block b takes segments [blk_seg[b], blk_seg[b + 1])
    uint32_t seg_cap, n_seg;
    uint32_t n_act;    // segments (tiles) [0, n_act) hold every neuron; the rest are idle (an arena's
                       // free ranges): no block walks them, their count columns stay zero
    DevState* st;
    const unsigned long long* row_ptr;
    const uint32_t* col;
    float* w;
    StdpArgs sd;       // plastic nets: lazy pair STDP in the placement pass
    uint32_t* hist;    // hist_f [NB][NBK]
    uint32_t* hist_c;  // [NB][NBK >> gt_shift]
    uint32_t* totals;  // [NBK]
    uint32_t* bs;      // [R][NBK + 1]
    uint32_t* gcur;    // [NBK] refine cursors of the batch's buckets
    uint32_t* flags;   // [R]
    uint2* ev;         // [R][e_cap]
    uint2* stage;      // [e_cap] events of the batch in group order
    void* ring;        // fallback when the batch exceeds e_cap
    unsigned long long e_cap;
    uint32_t N, D, Bfull, B, R, M, NT, NBK, tw_shift;
    uint32_t gt_shift;  // tiles per group = 1 << gt_shift = 2^14 >> tw_shift
    uint32_t G;        // mean list segments per histogram/placement block
    uint32_t NB;       // histogram/placement blocks
    uint32_t C;        // (64 U)-synapse tasks of the longest row
    uint32_t U;        // synapses per lane per task: 4 or 8
    uint32_t RS;       // refine blocks per group bucket (on average: they split the batch evenly)
    uint32_t direct;   // 1: groups are single tiles (gt_shift 0) and the placement writes each event
                       // straight into its tile bucket of ev[E % R]; no refine pass; the column scan
                       // leaves the per-block prefixes in the tile-level rows (no group rows)
    uint32_t n_cnt;    // placement LDS counters in use: keys [0, n_cnt) (group buckets, or with direct
                       // placement the tile buckets up to the last active tile of the last arrival batch)
    uint32_t balance;  // 1: k_bucket_split sets blk_seg per batch; 0: fixed ranges of G segments
    uint32_t AG;       // 1: histogram/refine LDS counters take one atomic per run of equal keys in a wave
    uint32_t AGP;      // 1: the same for the placement counters
    uint32_t is_last;
    // class-weight delivery (4-byte records): class and weight of each neuron's population
    const uint32_t* ncls;   // [N]
    const float* cls_w;     // [classes]
    // reserved placement (reserve = 1; single nets): no histogram passes.
    // Placement block b writes the events of group bucket kc into its region
    // of the staging buffer, whose start is
    //   off(b, kc) = res_base(Cc[kc], E) + (c[kc] * P_b >> 24) + kResPad * b
    // (P_b: events of the blocks before b, E: of the batch; c[kc]: the share of
    // kc reserved per event, 2^-24 units, from the network's synapses; Cc: its
    // prefix), so a region holds its expected events with a margin; events
    // beyond it go to the bucket's spill region after off(NB, kc), and beyond
    // that through the ring.  Refine: RS parts per group bucket, each a
    // counting sort by tile of its blocks' regions into the part's span of
    // ev[E % R], the tile ranges in tb.
    uint32_t reserve;
    const uint32_t* rc;               // [groups] c
    const unsigned long long* rcc;    // [groups + 1] Cc
    uint32_t* rcnt;                   // [groups][NB] events each block's region holds
    uint32_t* spill;                  // [groups] spill cursors (zeroed by the refine that reads them)
    unsigned long long* pb;           // [NB + 1] P_b, and E at [NB]
    uint32_t* tb;                     // [R][groups][parts][tiles per group + 1]
    uint32_t parts;                   // part slots per group bucket in tb (RS <= parts; the refine
                                      // empties slots [RS, parts))
};

constexpr uint32_t kResPad = 8;        // records every region holds beyond its share
constexpr uint32_t kResSpillPad = 64;   // and every spill region beyond a quarter of its bucket's share
constexpr uint32_t kResShift = 24;      // c in 2^-24 units
constexpr uint32_t kResMaxShare = 16u << kResShift;  // sum of c: at most 16 records reserved per event
constexpr uint32_t kResMaxParts = 8;    // RS bound
constexpr uint32_t kResMaxRanges = 128; // M * RS bound: the update's ranges per tile
constexpr uint32_t kResMaxPartRegions = 1024;  // regions of a part (ceil(NB / RS) + 1)
constexpr uint32_t kResCounterWords = 5;       // LDS words per placement counter: cursor, region end and start,
                                               // spill region start and end
constexpr uint32_t kResCntPerThread = 2;       // placement counters per 1024-thread block: 5 words each in 32 KB

// start of group bucket kc's regions for a batch of E events (Cc: its prefix);
// with kc = groups, the extent of the staging layout
__host__ __device__ __forceinline__ unsigned long long res_base(unsigned long long Cc, unsigned long long E, uint32_t kc,
                                                               uint32_t NB)
{
    const unsigned long long x = Cc * E;
    return (x >> kResShift) + (x >> (kResShift + 2)) + (unsigned long long)kc * (kResPad * NB + kResSpillPad);
}

constexpr uint32_t kMaxSegPerBlock = 64;  // segments a histogram/placement block stages at once
constexpr uint32_t kMaxSegSchedule = 4096;  // bound of the schedule's mean segments per block
constexpr uint32_t kMaxArrivalBatches = 64;  // M bound: (B - 1 + 127) / B for B >= 2

struct ScatterArgs {
    const naer_spike* list0;
    const naer_spike* list1;
    const uint32_t* seg_count;
    uint32_t seg_cap, n_seg;
    DevState* st;
    const unsigned long long* row_ptr;
    const uint32_t* col;
    float* w;
    void* ring;
    StdpArgs sd;
    uint32_t N, D;
    uint32_t B;
    uint32_t is_last;
    uint32_t* flags;   // bucketed delivery: the batch is marked as not bucketed
    uint32_t R, Bfull;
};

// Flush of pending potentiation over every plastic synapse (rows of the LIF
// neurons of the active networks; the row pointers of neurons outside them,
// such as an arena's free ranges, are not walked); `force` flushes
// regardless of the window.
struct FlushArgs {
    DevState* st;
    const unsigned long long* row_ptr;
    const uint32_t* col;
    float* w;
    StdpArgs sd;
    const PopParams* pops;   // [n_pops]
    const uint32_t* blk_pop; // first population of each update tile
    uint32_t n_pops, tw_shift;
    uint32_t N;
    uint32_t B;        // steps of the batch that follows
    uint32_t force;
};

// One connection of the device connectivity build.
struct ConnArgs {
    uint32_t src_begin, n_src;
    uint32_t dst_begin, n_dst;
    uint32_t K;            // FIXED_INDEGREE in-degree
    uint32_t conn_index;   // position in the description (stream word)
    unsigned long long seed;
    double p;              // BERNOULLI probability
    float weight, wmin, wmax;
    uint32_t dmin, dq, dchoices;
    uint32_t plastic_bit;
    uint32_t _pad;
};

}  // namespace naer
