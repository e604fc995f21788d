// naer_runtime.hip — host side of the C ABI (include/naer/naer.h): context,
// network/pack/arena objects, the device connectivity build (Network::compile,
// src/Network.cpp:267-569 of the reference) and the step loop (Network::step,
// src/Network.cpp:122-256).
//
// One object type serves the three forms of the ABI:
//   net    one network (a pack of one)
//   pack   several networks fixed at creation, laid out back to back
//   arena  preallocated capacity; networks (jobs) are admitted and retired
//          between steps, each built in place in free neuron and synapse
//          ranges
// Every network keeps its own seeds, local neuron ids and local clock, so its
// results do not depend on the form or on its neighbours.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <initializer_list>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "naer_device.hpp"
#include "naer_kernels.hpp"
#include "naer_model_abi.hpp"

using namespace naer;

namespace naer {
NAER_MODEL_ARGS;
}

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...)
{
    char buf[768];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIP_TRY(expr)                                                                                  \
    do {                                                                                               \
        hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess)                                                                          \
            return fail(e_ == hipErrorOutOfMemory ? NAER_ERR_NOMEM : NAER_ERR_HIP, "%s: %s (%s:%d)", #expr, \
                        hipGetErrorString(e_), __FILE__, __LINE__);                                    \
    } while (0)

// First-fit allocator of [0, capacity) ranges with coalescing.
class RangeAlloc {
public:
    void reset(uint64_t capacity)
    {
        free_.clear();
        if (capacity)
            free_[0] = capacity;
    }
    bool take(uint64_t size, uint64_t align, uint64_t* begin)
    {
        if (size == 0) {
            *begin = 0;
            return true;
        }
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            const uint64_t b = (it->first + align - 1) / align * align;
            const uint64_t end = it->first + it->second;
            if (b + size <= end) {
                const uint64_t lo = it->first;
                free_.erase(it);
                if (b > lo)
                    free_[lo] = b - lo;
                if (b + size < end)
                    free_[b + size] = end - (b + size);
                *begin = b;
                return true;
            }
        }
        return false;
    }
    void give(uint64_t begin, uint64_t size)
    {
        if (size == 0)
            return;
        auto it = free_.emplace(begin, size).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }

private:
    std::map<uint64_t, uint64_t> free_;
};

}  // namespace

struct naer_ctx {
    int device;
};

// A neuron model compiled from source (naer_model_create).
struct naer_model {
    std::vector<char> code;  // gfx950 code object
    uint32_t n_state = 0;
};

// A population of a source-defined model in a built net: the module loaded
// from the model's code object (owned by the net, so the model object may be
// destroyed once the net is built) and the population's device state.
struct ModelPop {
    hipModule_t mod = nullptr;
    hipFunction_t fn = nullptr;
    uint32_t begin = 0, end = 0, n_state = 0;
    uint64_t t_start = 0;
    float* state = nullptr;   // [n_state][end - begin]
    float* params = nullptr;  // [NAER_MODEL_PARAMS]
};

struct PendingTimer {
    int kind;
    hipEvent_t a, b;
};

// One network of a pack or arena.
struct NetInfo {
    naer_net_desc desc{};
    std::vector<naer_pop_desc> pops;
    std::vector<naer_conn_desc> conns;  // population indices local to the network
    uint32_t neuron_begin = 0, n_neurons = 0;
    uint64_t syn_begin = 0, n_syn = 0;
    uint64_t n_plastic = 0;  // plastic synapses
    uint32_t max_row = 0;
    uint32_t dmin = 0xFFFFFFFFu, dmax = 0;
    bool plastic = false;
    bool active = false;
    uint64_t t_start = 0;
};

struct naer_net {
    naer_ctx* ctx = nullptr;
    bool arena = false;
    bool built = false;
    bool poisson_ok = true;  // arena: jobs may contain POISSON sources

    // settings shared by every network
    float dt_ms = 0.0f;
    uint32_t dt_us = 0, mode = 0;
    bool det = false;
    bool plastic = false;  // plasticity present (pack) or allowed (arena)
    bool stdp_set = false;
    naer_stdp_desc stdp{};
    uint32_t D = 0, B = 0, dmin = 0, dmax = 0;
    uint32_t N = 0;      // neuron capacity
    uint64_t S_cap = 0;  // synapse capacity
    uint32_t max_row = 0;

    std::vector<NetInfo> nets;
    RangeAlloc neuron_alloc, syn_alloc;

    // population table of the active networks
    std::vector<PopParams> pp;
    std::vector<float> cdf;
    std::vector<uint32_t> blk_pop_h;
    uint32_t pop_cap = 0;

    hipStream_t stream = nullptr;
    hipStream_t build_stream = nullptr;  // arenas: admissions build their connectivity here
    std::vector<void*> allocs;
    uint64_t dev_bytes = 0;

    float* V = nullptr;
    // lazy pair STDP (plastic nets): spike history and the traces' dec^k tables
    uint4 *lat = nullptr, *pst = nullptr;
    uint32_t* shist = nullptr;
    uint32_t* flushed = nullptr;  // per row: potentiation applied for every post spike before this step
    float *d_tp = nullptr, *d_tm = nullptr;
    std::vector<float> h_tp, h_tm;
    uint32_t nref_min = 0;  // fewest refractory steps of any LIF population
    int32_t *ref = nullptr, *last = nullptr;
    uint32_t *pnext = nullptr, *pcnt = nullptr;
    void* ring = nullptr;
    unsigned long long* row_ptr = nullptr;
    uint32_t* col = nullptr;
    float* w = nullptr;
    unsigned long long* count = nullptr;   // build scratch, 1 entry
    uint32_t* deg = nullptr;               // build scratch, N entries
    unsigned long long* cursor = nullptr;  // build scratch, N entries
    naer_spike *list0 = nullptr, *list1 = nullptr, *rec = nullptr;
    uint32_t seg_cap = 0, n_seg = 0;
    uint32_t n_act = 1;  // tiles [0, n_act) hold every neuron of a population (arenas: the occupied prefix)
    uint32_t* seg_count = nullptr;
    unsigned long long* seg_ev = nullptr;
    uint32_t* blk_seg = nullptr;
    unsigned long long *spk_part = nullptr, *ev_part = nullptr;
    uint64_t rec_cap = 0, dropped = 0;
    DevState* st = nullptr;
    float* d_cdf = nullptr;
    PopParams* d_pops = nullptr;
    uint32_t* d_blk_pop = nullptr;
    float* probe = nullptr;
    uint64_t probe_t0 = 0;
    uint32_t probe_n = 0;

    // source-defined neuron models: staged input [B][N] and their populations
    float2* cin = nullptr;
    std::vector<ModelPop> mpops;

    // event-driven mode (NAER_MODE_EVENT): 1500 us windows
    bool event = false;
    uint32_t ev_in_cap = 0;
    uint64_t ev_e_cap = 0, ev_epoch = 0, ev_events = 0;
    uint32_t *ev_list[2] = { nullptr, nullptr }, *ev_sorted = nullptr;
    unsigned long long *ev_len = nullptr, *ev_off = nullptr;
    uint32_t *ev_keys = nullptr, *ev_keys_s = nullptr;
    float *ev_w = nullptr, *ev_w_s = nullptr;
    uint32_t* ev_tlast = nullptr;
    uint32_t *ev_begin = nullptr, *ev_end = nullptr;  // each target's run of sorted events
    double *ev_dlo = nullptr, *ev_dhi = nullptr;
    uint32_t *ev_nhi = nullptr, *ev_tref = nullptr;
    // [0], [1]: counts of the two input lists, [2]: inputs lost; spikes at [4] (u64)
    uint32_t* ev_ctr = nullptr;
    // plastic nets: input slots and payloads, event synapses and slots,
    // per-neuron trace and history, plastic populations
    uint32_t *ev_vals[2] = { nullptr, nullptr }, *ev_vals_s = nullptr;
    uint32_t* ev_prev[2] = { nullptr, nullptr };
    float* ev_ptr[2] = { nullptr, nullptr };
    uint32_t *ev_syn = nullptr, *ev_syn_s = nullptr, *ev_slot = nullptr, *ev_slot_s = nullptr;
    float* ev_trace = nullptr;
    uint32_t *ev_hist = nullptr, *ev_pop_plastic = nullptr;
    uint32_t* ev_sort_ws = nullptr;            // radix sort counts
    unsigned long long* ev_scan_ws = nullptr;  // scan partial sums

    uint32_t scatter_chunks = 1;

    // spike raster file (naer_net_raster_*)
    FILE* raster = nullptr;
    naer_spike* raster_stage = nullptr;  // pinned host copy of the record buffer
    uint64_t raster_records = 0;
    uint64_t steps_enqueued = 0;
    double tuned_us = 0.0;  // delivery time per batch of the schedule chosen by naer_net_autotune

    // update tiles and bucketed delivery
    uint32_t tw_shift = 8;   // update block = tile of 1 << tw_shift neurons
    uint32_t gt_shift = 6;   // tiles per bucket group
    bool bucketed = false;   // full aligned batches are delivered through buckets
    uint32_t R = 0, M = 0, NT = 0, NBK = 0, NB = 0, G = 1;
    uint32_t U = kBucketUnroll;  // schedule: synapses per lane per task
    uint32_t RS = 1;             // schedule: refine blocks per group bucket; 0: direct placement
    uint32_t AG = 1;             // schedule: run-aggregated LDS counters
    bool balance = false;        // histogram/placement blocks split each batch by events
    uint64_t e_cap = 0;       // records per event buffer
    uint32_t rec_bytes = 8;   // 4: class-weight delivery (kRec4* records)
    uint2* ev = nullptr;      // [R][e_cap] records of rec_bytes
    uint2* stage = nullptr;   // [e_cap] records of rec_bytes
    uint32_t* finite_flags = nullptr;  // naer_net_check_finite: per 256-neuron chunk
    // class-weight delivery: class (population) of every neuron, weight of every class
    uint32_t* ncls = nullptr;
    float* cls_w = nullptr;
    unsigned long long* cls_fx = nullptr;
    uint32_t cls_cap = 0;
    uint32_t *bs = nullptr, *flags = nullptr, *hist = nullptr, *hist_c = nullptr, *totals = nullptr;
    uint32_t* gcur = nullptr;
    // reserved placement (single nets; see BucketArgs)
    bool res_ok = false;    // shares known: the schedule may ask for it
    uint32_t RES = 0;       // schedule: 1 reserved placement
    uint32_t parts = 1;     // part slots per group bucket in res_tb
    uint32_t* res_c = nullptr;
    unsigned long long* res_cc = nullptr;
    uint32_t* res_cnt = nullptr;
    uint32_t* res_spill = nullptr;
    unsigned long long* res_pb = nullptr;
    uint32_t* res_tb = nullptr;

    bool profiling = false;
    uint32_t graph_chunk = 0;  // steps per captured graph (0 = plain launches)
    hipGraphExec_t graph_exec = nullptr;
    std::vector<PendingTimer> pending;
    std::vector<hipEvent_t> event_pool;
    double kernel_ms[8] = {};
    uint64_t kernel_calls[8] = {};

    ~naer_net()
    {
        if (stream)
            (void)hipStreamSynchronize(stream);
        if (graph_exec)
            (void)hipGraphExecDestroy(graph_exec);
        for (auto& p : pending) {
            (void)hipEventDestroy(p.a);
            (void)hipEventDestroy(p.b);
        }
        for (auto e : event_pool)
            (void)hipEventDestroy(e);
        for (void* p : allocs)
            (void)hipFree(p);
        for (auto& m : mpops)
            if (m.mod)
                (void)hipModuleUnload(m.mod);
        if (raster)
            std::fclose(raster);  // unfinished: the header keeps 0 records
        if (raster_stage)
            (void)hipHostFree(raster_stage);
        if (build_stream)
            (void)hipStreamDestroy(build_stream);
        if (stream)
            (void)hipStreamDestroy(stream);
    }

    template <typename T>
    int alloc(T** p, size_t count)
    {
        size_t bytes = count * sizeof(T);
        if (bytes == 0)
            bytes = sizeof(T);
        void* q = nullptr;
        hipError_t e = hipMalloc(&q, bytes);
        if (e != hipSuccess)
            return fail(NAER_ERR_NOMEM, "hipMalloc of %zu bytes failed: %s", bytes, hipGetErrorString(e));
        allocs.push_back(q);
        dev_bytes += bytes;
        *p = (T*)q;
        return NAER_OK;
    }

    void release(void* p, size_t bytes)
    {
        auto it = std::find(allocs.begin(), allocs.end(), p);
        if (it != allocs.end()) {
            allocs.erase(it);
            (void)hipFree(p);
            dev_bytes -= bytes;
        }
    }

    hipEvent_t take_event()
    {
        if (!event_pool.empty()) {
            hipEvent_t e = event_pool.back();
            event_pool.pop_back();
            return e;
        }
        hipEvent_t e = nullptr;
        (void)hipEventCreate(&e);
        return e;
    }

    int collect_timers()
    {
        for (auto& p : pending) {
            float ms = 0.0f;
            HIP_TRY(hipEventElapsedTime(&ms, p.a, p.b));
            kernel_ms[p.kind] += ms;
            if (p.kind >= NAER_K_BUCKET_HIST && p.kind <= NAER_K_BUCKET_REFINE)
                kernel_ms[NAER_K_SCATTER] += ms;
            event_pool.push_back(p.a);
            event_pool.push_back(p.b);
        }
        pending.clear();
        return NAER_OK;
    }

    int drop_graph()
    {
        if (graph_exec) {
            HIP_TRY(hipGraphExecDestroy(graph_exec));
            graph_exec = nullptr;
        }
        return NAER_OK;
    }
};

namespace {

// ---------------------------------------------------------------------------
// description checks
// ---------------------------------------------------------------------------
int check_settings(naer_net* n, float dt_ms, uint32_t mode)
{
    const double dtus = (double)dt_ms * 1000.0;
    const uint32_t dt_us = (uint32_t)std::llround(dtus);
    if (!(dt_ms > 0.0f) || std::fabs(dtus - (double)dt_us) > 1e-3 || dt_us == 0)
        return fail(NAER_ERR_INVALID, "dt must be a positive whole number of microseconds");
    if (mode != NAER_MODE_FAST && mode != NAER_MODE_DETERMINISTIC && mode != NAER_MODE_EVENT)
        return fail(NAER_ERR_INVALID, "unknown mode %u", mode);
    n->dt_ms = dt_ms;
    n->dt_us = dt_us;
    n->mode = mode;
    n->det = mode == NAER_MODE_DETERMINISTIC;
    n->event = mode == NAER_MODE_EVENT;
    return NAER_OK;
}

// Validates one network and fills its NetInfo (ranges excluded).
int check_net(naer_net* n, const naer_net_desc& d, NetInfo* info)
{
    if (d.n_pops == 0 || !d.pops)
        return fail(NAER_ERR_INVALID, "a network needs at least one population");
    if (d.n_conns && !d.conns)
        return fail(NAER_ERR_INVALID, "conns is NULL");
    if (d.dt_ms != n->dt_ms || d.mode != n->mode)
        return fail(NAER_ERR_INVALID, "all networks of a pack or arena share dt and mode");
    info->desc = d;
    info->pops.assign(d.pops, d.pops + d.n_pops);
    info->conns.assign(d.conns, d.conns + d.n_conns);
    info->desc.pops = nullptr;
    info->desc.conns = nullptr;
    uint64_t nn = 0;
    auto finite = [](std::initializer_list<double> v) {
        for (double x : v)
            if (!std::isfinite(x))
                return false;
        return true;
    };
    if (!finite({ d.dt_ms }) || !(d.dt_ms > 0.0f))
        return fail(NAER_ERR_INVALID, "dt must be finite and > 0");
    for (uint32_t p = 0; p < d.n_pops; ++p) {
        const naer_pop_desc& q = d.pops[p];
        nn += q.size;
        if (!finite({ q.tau_m_ms, q.v_th, q.v_reset, q.v_rest, q.t_ref_ms, q.v0_min, q.v0_max, q.ext_rate_hz,
                      q.ext_weight, q.src_rate_hz }))
            return fail(NAER_ERR_INVALID, "population %u: parameters must be finite", p);
        for (uint32_t k = 0; q.kind == NAER_POP_MODEL && k < NAER_MODEL_PARAMS; ++k)
            if (!std::isfinite(q.params[k]))
                return fail(NAER_ERR_INVALID, "population %u: model parameter %u is not finite", p, k);
        if (q.kind == NAER_POP_POISSON) {
            if (kEpochUs % n->dt_us != 0)
                return fail(NAER_ERR_INVALID, "POISSON sources need dt dividing the 1500 us epoch");
        } else if (q.kind == NAER_POP_LIF || q.kind == NAER_POP_MODEL) {
            if (q.kind == NAER_POP_LIF && !(q.tau_m_ms > 0.0f))
                return fail(NAER_ERR_INVALID, "population %u: tau_m must be > 0", p);
            if (q.kind == NAER_POP_MODEL && !q.model)
                return fail(NAER_ERR_INVALID, "population %u: a MODEL population needs a model", p);
            const double lam = (double)q.ext_rate_hz * (double)d.dt_ms * 1e-3;
            if (!(lam >= 0.0 && lam <= 20.0))
                return fail(NAER_ERR_INVALID, "population %u: drive rate*dt must be in [0, 20]", p);
        } else {
            return fail(NAER_ERR_INVALID, "population %u: unknown kind %u", p, q.kind);
        }
    }
    if (nn == 0 || nn > (1ull << 24))
        return fail(NAER_ERR_CAPACITY, "neuron count must be in [1, 2^24] (got %llu)", (unsigned long long)nn);
    info->n_neurons = (uint32_t)nn;
    for (uint32_t c = 0; c < d.n_conns; ++c) {
        const naer_conn_desc& k = d.conns[c];
        if (k.src_pop >= d.n_pops || k.dst_pop >= d.n_pops)
            return fail(NAER_ERR_INVALID, "connection %u: population index out of range", c);
        if (d.pops[k.dst_pop].kind == NAER_POP_POISSON)
            return fail(NAER_ERR_INVALID, "connection %u: target population must be LIF or MODEL", c);
        const uint32_t q = k.delay_quantum ? k.delay_quantum : 1;
        if (k.delay_min < 1 || k.delay_max < k.delay_min || k.delay_max > 127 || (k.delay_max - k.delay_min) % q)
            return fail(NAER_ERR_INVALID,
                        "connection %u: delays must satisfy 1 <= min <= max <= 127 on the quantum grid", c);
        if (k.plastic && (d.pops[k.src_pop].kind != NAER_POP_LIF || d.pops[k.dst_pop].kind != NAER_POP_LIF))
            return fail(NAER_ERR_INVALID, "connection %u: plastic connections need LIF source and target", c);
        if (k.rule != NAER_RULE_BERNOULLI && k.rule != NAER_RULE_FIXED_INDEGREE)
            return fail(NAER_ERR_INVALID, "connection %u: unknown rule %u", c, k.rule);
        if (k.rule == NAER_RULE_BERNOULLI && !(k.probability >= 0.0 && k.probability <= 1.0))
            return fail(NAER_ERR_INVALID, "connection %u: probability must be in [0, 1]", c);
        if (!finite({ k.weight, k.wmin, k.wmax }))
            return fail(NAER_ERR_INVALID, "connection %u: weights must be finite", c);
        info->dmin = std::min(info->dmin, k.delay_min);
        info->dmax = std::max(info->dmax, k.delay_max);
        if (k.plastic) {
            const naer_stdp_desc& s = d.stdp;
            if (!(s.tau_plus_ms > 0.0f) || !(s.tau_minus_ms > 0.0f) || !(s.w_max >= s.w_min) ||
                !finite({ s.tau_plus_ms, s.tau_minus_ms, s.a_plus, s.a_minus, s.w_min, s.w_max }))
                return fail(NAER_ERR_INVALID, "STDP needs finite parameters, tau_plus, tau_minus > 0 and w_max >= w_min");
            info->plastic = true;
        }
    }
    if (d.mode == NAER_MODE_EVENT) {
        // the reference's event-driven step: static LIF populations without
        // drive, POISSON sources
        for (uint32_t p = 0; p < d.n_pops; ++p) {
            const naer_pop_desc& q = d.pops[p];
            if (q.kind == NAER_POP_MODEL || (q.kind == NAER_POP_LIF && q.ext_rate_hz > 0.0f))
                return fail(NAER_ERR_INVALID,
                            "population %u: the event-driven mode takes LIF populations without drive and POISSON "
                            "sources", p);
            if (q.kind == NAER_POP_LIF && (!(q.t_ref_ms >= 0.0f) || q.tau_m_ms * 1000.0f < 1.0f))
                return fail(NAER_ERR_INVALID, "population %u: tau_m must be >= 1 us and t_ref >= 0", p);
        }
        // stdp.cl's update reads the weight the window's expansion carried:
        // a plastic synapse's source fires at most once per window
        for (uint32_t c = 0; c < d.n_conns; ++c) {
            const naer_conn_desc& k = d.conns[c];
            if (k.plastic && !(d.pops[k.src_pop].t_ref_ms * 1000.0f >= (float)kEpochUs))
                return fail(NAER_ERR_INVALID,
                            "connection %u: plastic synapses in the event-driven mode need a source refractory period "
                            "of at least 1.5 ms",
                            c);
        }
    }
    return NAER_OK;
}

// ---------------------------------------------------------------------------
// population table of the active networks (sorted by neuron range)
// ---------------------------------------------------------------------------
void fill_pop(const naer_net* n, const NetInfo& net, uint32_t p_local, uint32_t begin, uint32_t pois_base,
              uint32_t n_pois, PopParams& r, float* cdf)
{
    const naer_pop_desc& q = net.pops[p_local];
    std::memset(&r, 0, sizeof r);
    r.begin = begin;
    r.end = begin + q.size;
    r.kind = q.kind;
    r.net_base = net.neuron_begin;
    r.seed_lo = (uint32_t)net.desc.seed;
    r.seed_hi = (uint32_t)(net.desc.seed >> 32);
    r.pois_seed = net.desc.poisson_seed;
    r.t_start_lo = (uint32_t)net.t_start;
    r.t_start_hi = (uint32_t)(net.t_start >> 32);
    r.n_pois = n_pois;
    for (uint32_t k = 0; k < NAER_DRIVE_TABLE; ++k)
        cdf[k] = 2.0f;
    if (q.kind == NAER_POP_POISSON) {
        r.pois_base = pois_base;
        r.neg_rate_per_us = -((double)q.src_rate_hz / 1000000.0);
        return;
    }
    r.P = (float)std::exp(-(double)n->dt_ms / (double)q.tau_m_ms);
    r.v_th = q.v_th;
    r.v_reset = q.v_reset;
    r.v_rest = q.v_rest;
    r.v0_min = q.v0_min;
    r.v0_max = q.v0_max;
    r.w_ext = q.ext_weight;
    r.n_ref = (int32_t)std::llround((double)q.t_ref_ms / (double)n->dt_ms);
    r.has_drive = q.ext_rate_hz > 0.0f ? 1u : 0u;
    // inversion table, see NAER_DRIVE_TABLE in naer.h
    const double lam = (double)q.ext_rate_hz * (double)n->dt_ms * 1e-3;
    double pk = std::exp(-lam), cum = pk;
    for (uint32_t k = 0;; ++k) {
        const float f = (float)cum;
        if (f >= 1.0f || k == NAER_DRIVE_TABLE - 1) {
            cdf[k] = 2.0f;
            break;
        }
        cdf[k] = f;
        pk = pk * lam / (double)(k + 1);
        cum = cum + pk;
    }
}

int upload_table(naer_net* n)
{
    std::vector<uint32_t> order;
    for (uint32_t i = 0; i < n->nets.size(); ++i)
        if (n->nets[i].active)
            order.push_back(i);
    std::sort(order.begin(), order.end(),
              [&](uint32_t a, uint32_t b) { return n->nets[a].neuron_begin < n->nets[b].neuron_begin; });
    n->pp.clear();
    n->cdf.clear();
    for (uint32_t i : order) {
        const NetInfo& net = n->nets[i];
        uint32_t n_pois = 0;
        for (auto& q : net.pops)
            if (q.kind == NAER_POP_POISSON)
                n_pois += q.size;
        uint32_t begin = net.neuron_begin, pois = 0;
        for (uint32_t p = 0; p < net.pops.size(); ++p) {
            if (net.pops[p].size == 0)
                continue;
            PopParams r;
            n->cdf.resize(n->cdf.size() + NAER_DRIVE_TABLE);
            fill_pop(n, net, p, begin, pois, n_pois, r, &n->cdf[n->cdf.size() - NAER_DRIVE_TABLE]);
            r.cdf_offset = (uint32_t)(n->cdf.size() - NAER_DRIVE_TABLE);
            if (net.pops[p].kind == NAER_POP_POISSON)
                pois += net.pops[p].size;
            begin += net.pops[p].size;
            n->pp.push_back(r);
        }
    }
    const uint32_t P = (uint32_t)n->pp.size();
    if (P > n->pop_cap) {
        if (n->d_pops)
            n->release(n->d_pops, (size_t)n->pop_cap * sizeof(PopParams));
        if (n->d_cdf)
            n->release(n->d_cdf, (size_t)n->pop_cap * NAER_DRIVE_TABLE * sizeof(float));
        n->pop_cap = std::max<uint32_t>(P, 2 * n->pop_cap);
        int rc;
        if ((rc = n->alloc(&n->d_pops, n->pop_cap)) ||
            (rc = n->alloc(&n->d_cdf, (size_t)n->pop_cap * NAER_DRIVE_TABLE)))
            return rc;
    }
    // first population of every update block: the last one starting at or
    // before the block's first neuron
    // (kTileIdle: no population overlaps the tile, e.g. an arena's free
    // ranges; the update skips it)
    n->blk_pop_h.assign(n->n_seg, 0);
    uint32_t p = 0;
    const uint32_t TW = 1u << n->tw_shift;
    for (uint32_t b = 0; b < n->n_seg; ++b) {
        const uint32_t j = b << n->tw_shift;
        while (p + 1 < P && n->pp[p + 1].begin <= j)
            ++p;
        bool active = false;
        for (uint32_t q = p; q < P && n->pp[q].begin < j + TW && !active; ++q)
            active = n->pp[q].end > j;
        n->blk_pop_h[b] = p | (active ? 0u : kTileIdle);
    }
    // the update, the scatter and the bucketed passes cover the tiles up to
    // the last one with a population (first-fit keeps an arena's jobs low)
    n->n_act = 1;
    for (uint32_t b = n->n_seg; b > 0; --b)
        if (!(n->blk_pop_h[b - 1] & kTileIdle)) {
            n->n_act = b;
            break;
        }
    std::vector<float> cls_w;
    std::vector<unsigned long long> cls_fx;
    std::vector<uint32_t> ncls;
    if (n->rec_bytes == 4) {
        // class c = population c of the table: its neurons and its outgoing weight
        ncls.assign(n->N, 0u);
        cls_w.assign(P, 0.0f);
        for (uint32_t i : order) {
            const NetInfo& net = n->nets[i];
            uint32_t q = 0;  // index of the net's first population in the table
            while (q < P && n->pp[q].begin < net.neuron_begin)
                ++q;
            for (uint32_t p = 0; p < net.pops.size(); ++p) {
                if (net.pops[p].size == 0)
                    continue;
                for (const auto& c : net.conns)
                    if (c.src_pop == p)
                        cls_w[q] = c.weight;
                for (uint32_t j = n->pp[q].begin; j < n->pp[q].end; ++j)
                    ncls[j] = q;
                ++q;
            }
        }
        cls_fx.resize(P);
        for (uint32_t c = 0; c < P; ++c)
            cls_fx[c] = (unsigned long long)std::llrint((double)cls_w[c] * 4294967296.0);
        if (P > n->cls_cap) {
            if (n->cls_w)
                n->release(n->cls_w, (size_t)n->cls_cap * sizeof(float));
            if (n->cls_fx)
                n->release(n->cls_fx, (size_t)n->cls_cap * sizeof(unsigned long long));
            n->cls_cap = P;
            int rc;
            if ((rc = n->alloc(&n->cls_w, P)) || (rc = n->alloc(&n->cls_fx, P)))
                return rc;
        }
        HIP_TRY(hipMemcpyAsync(n->cls_w, cls_w.data(), P * sizeof(float), hipMemcpyHostToDevice, n->stream));
        HIP_TRY(hipMemcpyAsync(n->cls_fx, cls_fx.data(), P * sizeof(unsigned long long), hipMemcpyHostToDevice,
                               n->stream));
        HIP_TRY(hipMemcpyAsync(n->ncls, ncls.data(), (size_t)n->N * sizeof(uint32_t), hipMemcpyHostToDevice,
                               n->stream));
    }
    if (P) {
        HIP_TRY(hipMemcpyAsync(n->d_pops, n->pp.data(), P * sizeof(PopParams), hipMemcpyHostToDevice, n->stream));
        HIP_TRY(hipMemcpyAsync(n->d_cdf, n->cdf.data(), n->cdf.size() * sizeof(float), hipMemcpyHostToDevice,
                               n->stream));
    }
    HIP_TRY(hipMemcpyAsync(n->d_blk_pop, n->blk_pop_h.data(), n->n_seg * sizeof(uint32_t), hipMemcpyHostToDevice,
                           n->stream));
    // the copies read host vectors that may change at the next upload
    HIP_TRY(hipStreamSynchronize(n->stream));
    // captured graphs hold the population count in their arguments
    return n->drop_graph();
}

// dec^k = exp(-k dt / tau) for k = 0, 1, ... while it is >= 2^-40 (at most
// kMaxTraceSteps entries), each rounded to fp32; beyond the table the
// traces take dec^k = 0.  The oracle builds the same table.
constexpr uint32_t kMaxTraceSteps = 1u << 22;

std::vector<float> trace_table(float dt_ms, float tau_ms)
{
    std::vector<float> t;
    for (uint32_t k = 0; k < kMaxTraceSteps; ++k) {
        const double v = std::exp(-(double)k * (double)dt_ms / (double)tau_ms);
        if (v < 0x1p-40)
            break;
        t.push_back((float)v);
    }
    return t;
}

StdpArgs stdp_args(const naer_net* n)
{
    StdpArgs a;
    std::memset(&a, 0, sizeof a);
    if (!n->plastic)
        return a;
    a.lat = n->lat;
    a.pst = n->pst;
    a.hist = n->shist;
    a.flushed = n->flushed;
    a.tp = n->d_tp;
    a.tm = n->d_tm;
    a.n_tp = (uint32_t)n->h_tp.size();
    a.n_tm = (uint32_t)n->h_tm.size();
    a.a_plus = n->stdp.a_plus;
    a.a_minus = n->stdp.a_minus;
    a.w_min = n->stdp.w_min;
    a.w_max = n->stdp.w_max;
    return a;
}

// Steps within which no neuron can fire kSpikeHist times (its spikes are >=
// nref_min + 1 steps apart): a row's pending potentiation must be applied
// (at the pre neuron's spike or by a flush) within this many steps of its
// start.  Flushes run every half window and take only the rows whose pending
// range would otherwise outgrow the window before the next one.
uint32_t flush_window(const naer_net* n)
{
    return (kSpikeHist - 1) * (n->nref_min + 1);
}

UpdateArgs make_update_args(naer_net* n)
{
    UpdateArgs a;
    std::memset(&a, 0, sizeof a);
    a.V = n->V;
    a.ref = n->ref;
    a.last = n->last;
    a.sd = stdp_args(n);
    a.pnext = n->pnext;
    a.pcnt = n->pcnt;
    a.ring = n->ring;
    a.row_ptr = n->row_ptr;
    a.list0 = n->list0;
    a.list1 = n->list1;
    a.seg_count = n->seg_count;
    a.seg_ev = n->seg_ev;
    a.spk_part = n->spk_part;
    a.ev_part = n->ev_part;
    a.rec = n->rec;
    a.st = n->st;
    a.cdf = n->d_cdf;
    a.pops = n->d_pops;
    a.blk_pop = n->d_blk_pop;
    a.probe = n->probe;
    a.rec_cap = n->rec_cap;
    a.probe_t0 = n->probe_t0;
    a.probe_n = n->probe_n;
    a.seg_cap = n->seg_cap;
    a.n_seg = n->n_seg;
    a.n_act = n->n_act;
    a.N = n->N;
    a.D = n->D;
    a.B = n->B;
    a.dt_us = n->dt_us;
    a.n_pops = (uint32_t)n->pp.size();
    a.tw_shift = n->tw_shift;
    a.Bfull = n->B;
    a.ev = n->ev;
    a.bs = n->bs;
    a.flags = n->flags;
    a.e_cap = n->e_cap;
    a.R = n->R;
    a.M = n->M;
    a.NT = n->NT;
    a.NBK = n->NBK;
    a.cin = n->cin;
    a.cls_fx = n->cls_fx;
    a.n_cls = n->rec_bytes == 4 ? (uint32_t)n->pp.size() : 0u;
    a.tb = n->res_tb;
    a.parts = n->parts;
    a.gt_shift = n->gt_shift;
    a.col = n->col;
    a.prefetch = 0;
    return a;
}

ConnArgs make_conn_args(const NetInfo& net, uint32_t c)
{
    const naer_conn_desc& k = net.conns[c];
    // population starts within the network (populations are contiguous, in order)
    std::vector<uint32_t> pb(net.pops.size());
    uint32_t acc = net.neuron_begin;
    for (size_t p = 0; p < net.pops.size(); ++p) {
        pb[p] = acc;
        acc += net.pops[p].size;
    }
    ConnArgs a;
    std::memset(&a, 0, sizeof a);
    a.src_begin = pb[k.src_pop];
    a.n_src = net.pops[k.src_pop].size;
    a.dst_begin = pb[k.dst_pop];
    a.n_dst = net.pops[k.dst_pop].size;
    a.K = k.indegree;
    a.conn_index = c;
    a.seed = k.seed;
    a.p = k.probability;
    a.weight = k.weight;
    a.wmin = k.wmin;
    a.wmax = k.wmax;
    a.dmin = k.delay_min;
    a.dq = k.delay_quantum ? k.delay_quantum : 1;
    a.dchoices = (k.delay_max - k.delay_min) / a.dq + 1;
    a.plastic_bit = k.plastic ? kPlasticBit : 0u;
    return a;
}

// ---------------------------------------------------------------------------
// device build of one network
// ---------------------------------------------------------------------------
int count_net(naer_net* n, const NetInfo& net)
{
    HIP_TRY(hipMemsetAsync(n->deg + net.neuron_begin, 0, (size_t)net.n_neurons * sizeof(uint32_t), n->stream));
    for (uint32_t c = 0; c < net.conns.size(); ++c) {
        ConnArgs a = make_conn_args(net, c);
        HIP_TRY(launch_conn(a, net.conns[c].rule, false, n->deg, nullptr, nullptr, nullptr, n->stream));
    }
    return NAER_OK;
}

// Reads the counts of [begin, begin+count) and returns their prefix sums.
int read_prefix(naer_net* n, uint32_t begin, uint32_t count, std::vector<unsigned long long>* prefix,
                uint32_t* max_len)
{
    std::vector<uint32_t> d(count);
    if (count)
        HIP_TRY(hipMemcpyAsync(d.data(), n->deg + begin, count * sizeof(uint32_t), hipMemcpyDeviceToHost, n->stream));
    HIP_TRY(hipStreamSynchronize(n->stream));
    prefix->resize((size_t)count + 1);
    unsigned long long acc = 0;
    uint32_t mx = 0;
    for (uint32_t j = 0; j < count; ++j) {
        (*prefix)[j] = acc;
        acc += d[j];
        mx = std::max(mx, d[j]);
    }
    (*prefix)[count] = acc;
    if (max_len)
        *max_len = mx;
    return NAER_OK;
}

// Writes ptr[begin .. begin+count] = base + prefix.
int write_ptr(naer_net* n, unsigned long long* d_ptr, uint32_t begin, uint64_t base,
              std::vector<unsigned long long>& prefix)
{
    for (auto& v : prefix)
        v += base;
    HIP_TRY(hipMemcpyAsync(d_ptr + begin, prefix.data(), prefix.size() * sizeof(unsigned long long),
                           hipMemcpyHostToDevice, n->stream));
    HIP_TRY(hipStreamSynchronize(n->stream));
    return NAER_OK;
}

int place_net(naer_net* n, const NetInfo& net)
{
    HIP_TRY(hipMemcpyAsync(n->cursor + net.neuron_begin, n->row_ptr + net.neuron_begin,
                           (size_t)net.n_neurons * sizeof(unsigned long long), hipMemcpyDeviceToDevice, n->stream));
    for (uint32_t c = 0; c < net.conns.size(); ++c) {
        ConnArgs a = make_conn_args(net, c);
        HIP_TRY(launch_conn(a, net.conns[c].rule, true, nullptr, n->cursor, n->col, n->w, n->stream));
    }
    HIP_TRY(launch_sort_rows(n->row_ptr, n->col, n->w, net.neuron_begin, net.neuron_begin + net.n_neurons,
                             net.max_row, n->stream));
    return NAER_OK;
}

// Counts the plastic synapses of a network (statistics).
int count_plastic(naer_net* n, NetInfo& net)
{
    HIP_TRY(hipMemsetAsync(n->count, 0, sizeof(unsigned long long), n->stream));
    HIP_TRY(launch_count_plastic(n->col, net.syn_begin, net.syn_begin + net.n_syn, n->count, n->stream));
    unsigned long long k = 0;
    HIP_TRY(hipMemcpyAsync(&k, n->count, sizeof k, hipMemcpyDeviceToHost, n->stream));
    HIP_TRY(hipStreamSynchronize(n->stream));
    net.n_plastic = k;
    return NAER_OK;
}

int init_net(naer_net* n, const NetInfo& net)
{
    UpdateArgs ua = make_update_args(n);
    HIP_TRY(launch_init_state(ua, n->det, net.neuron_begin, net.neuron_begin + net.n_neurons, n->stream));
    if (n->bucketed)
        HIP_TRY(launch_bucket_scrub(ua, net.neuron_begin, net.neuron_begin + net.n_neurons, n->stream));
    return NAER_OK;
}

// Steps per update launch and update tile width.  The inputs of a batch
// were all delivered before it starts when B <= the minimum delay; in plastic
// nets B <= the fewest refractory steps + 1, so that a neuron fires at most
// once per batch (lazy STDP).  Large nets (N >= 2^18) whose minimum delay
// allows fewer than kLongBatch steps per batch use 1024-neuron tiles with B
// <= kMaxBatchLargeTile, which keeps the update's [B][1024] 64-bit LDS
// accumulators of two tiles within one CU (measured: 1M neurons, d_min 10:
// 9 steps and two tiles per CU beat 10 steps and one).  Longer batches keep
// their length (fewer batches, less fixed cost) on 256-neuron tiles
// (measured on the sweep's packs of 1.4-1.6M neurons, B = 15: 256-neuron
// tiles 0.63 / 1.02 ms per step (static / plastic pack), 512: 0.63 / 1.05,
// 1024: 0.69 / 1.22).
constexpr uint32_t kLongBatch = 12;

uint32_t batch_steps(bool plastic, uint32_t dmin, uint64_t N, uint32_t nref_min)
{
    uint32_t b = std::min<uint32_t>(kMaxBatch, dmin == 0xFFFFFFFFu ? kMaxBatch : dmin);
    if (plastic)  // lazy STDP: every neuron fires at most once per batch
        b = std::min(b, nref_min + 1);
    if (N >= (1u << 18) && b < kLongBatch)
        b = std::min(b, kMaxBatchLargeTile);
    return std::max(b, 1u);
}

// Histogram LDS holds 4 bytes per bucket: at most kMaxBuckets (128 KB).
constexpr uint32_t kMaxBuckets = 32768;

uint32_t tile_shift(uint64_t N, uint32_t B, uint32_t dmax)
{
    if (N >= (1u << 18) && B < kLongBatch)
        return 10u;
    // the smallest tile whose buckets (arrival batches x tiles, whole groups
    // of 2^14 neurons) fit the histogram, e.g. for an arena of 2e7 neurons
    const uint64_t M = (B - 1 + dmax) / std::max(B, 1u);
    const uint64_t groups = (N + (1u << kGroupNeuronShift) - 1) >> kGroupNeuronShift;
    for (uint32_t shift = 8; shift < 10; ++shift)
        if (M * (groups << (kGroupNeuronShift - shift)) <= kMaxBuckets)
            return shift;
    return 10u;
}

// Allocation of everything sized by the neuron capacity.

int alloc_neurons(naer_net* n, bool any_poisson)
{
    const uint32_t N = n->N;
    int rc;
    if ((rc = n->alloc(&n->deg, N)) || (rc = n->alloc(&n->cursor, N)) || (rc = n->alloc(&n->count, 1)) ||
        (rc = n->alloc(&n->row_ptr, (size_t)N + 1)))
        return rc;
    HIP_TRY(hipMemsetAsync(n->row_ptr, 0, ((size_t)N + 1) * sizeof(unsigned long long), n->stream));
    if (n->plastic) {
        n->h_tp = trace_table(n->dt_ms, n->stdp.tau_plus_ms);
        n->h_tm = trace_table(n->dt_ms, n->stdp.tau_minus_ms);
        if ((rc = n->alloc(&n->lat, N)) ||
            (rc = n->alloc(&n->pst, 2 * (size_t)N)) || (rc = n->alloc(&n->shist, (size_t)N * kSpikeHist)) ||
            (rc = n->alloc(&n->flushed, N)) ||
            (rc = n->alloc(&n->d_tp, std::max<size_t>(1, n->h_tp.size()))) ||
            (rc = n->alloc(&n->d_tm, std::max<size_t>(1, n->h_tm.size()))))
            return rc;
        HIP_TRY(hipMemsetAsync(n->lat, 0, (size_t)N * sizeof(uint4), n->stream));
        HIP_TRY(hipMemsetAsync(n->flushed, 0, (size_t)N * sizeof(uint32_t), n->stream));
        // every word kNoSpike: steps absent (the trace words are never read then)
        HIP_TRY(hipMemsetD32Async((hipDeviceptr_t)n->pst, (int)kNoSpike, (size_t)N * 8, n->stream));
        HIP_TRY(hipMemcpy(n->d_tp, n->h_tp.data(), n->h_tp.size() * sizeof(float), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(n->d_tm, n->h_tm.data(), n->h_tm.size() * sizeof(float), hipMemcpyHostToDevice));
    }
    if ((rc = n->alloc(&n->V, N)) || (rc = n->alloc(&n->ref, N)) || (rc = n->alloc(&n->last, N)) ||
        (rc = n->alloc(&n->pnext, N)) || (rc = n->alloc(&n->pcnt, N)))
        return rc;
    const size_t ring_bytes = (size_t)n->D * N * (n->det ? sizeof(long long) : sizeof(float));
    if ((rc = n->alloc((char**)&n->ring, ring_bytes)))
        return rc;
    HIP_TRY(hipMemsetAsync(n->ring, 0, ring_bytes, n->stream));
    // list segment of one update block: every LIF neuron fires at most once
    // per step, a POISSON source at most 10 times per 1500 us epoch
    const uint64_t epochs = ((uint64_t)n->B * n->dt_us + kEpochUs - 1) / kEpochUs + 1;
    const uint64_t per_neuron = any_poisson ? std::max<uint64_t>(n->B, kMaxPoissonPerEpoch * epochs) : n->B;
    // update tiles of 256, 512 or 1024 neurons (tile_shift)
    n->tw_shift = tile_shift(N, n->B, n->dmax);
    const uint32_t TW = 1u << n->tw_shift;
    n->seg_cap = (uint32_t)(TW * per_neuron);
    n->n_seg = (N + TW - 1) / TW;
    const size_t list_elems = (size_t)n->seg_cap * n->n_seg;
    if ((rc = n->alloc(&n->list0, list_elems)) || (rc = n->alloc(&n->list1, list_elems)) ||
        (rc = n->alloc(&n->seg_count, 2 * (size_t)n->n_seg)) || (rc = n->alloc(&n->spk_part, n->n_seg)) ||
        (rc = n->alloc(&n->ev_part, n->n_seg)) || (rc = n->alloc(&n->d_blk_pop, n->n_seg)) ||
        (rc = n->alloc(&n->seg_ev, 2 * (size_t)n->n_seg)) || (rc = n->alloc(&n->blk_seg, (size_t)n->n_seg + 1)))
        return rc;
    HIP_TRY(hipMemsetAsync(n->seg_count, 0, 2 * (size_t)n->n_seg * sizeof(uint32_t), n->stream));
    HIP_TRY(hipMemsetAsync(n->seg_ev, 0, 2 * (size_t)n->n_seg * sizeof(unsigned long long), n->stream));
    HIP_TRY(hipMemsetAsync(n->spk_part, 0, n->n_seg * sizeof(unsigned long long), n->stream));
    HIP_TRY(hipMemsetAsync(n->ev_part, 0, n->n_seg * sizeof(unsigned long long), n->stream));
    if ((rc = n->alloc(&n->rec, n->rec_cap ? n->rec_cap : 1)) || (rc = n->alloc(&n->st, 1)))
        return rc;
    DevState st0;
    std::memset(&st0, 0, sizeof st0);
    st0.window = flush_window(n);
    HIP_TRY(hipMemcpy(n->st, &st0, sizeof st0, hipMemcpyHostToDevice));
    return NAER_OK;
}

// Buffers of the bucketed delivery: R = M + 1 batches of events in flight
// (a batch's events are read during the M batches that follow it), e_cap
// events each, sized for a mean rate over all synapses of the refractory
// limit 1 / ((n_ref + 1) dt), kept within [kBucketRateMinHz, kBucketRateMaxHz]
// (and within a quarter of the free memory): a batch that does not fit goes
// through the ring's device atomics, several times slower (measured: the
// sweep's plastic networks at 155 Hz overflowed a 200 Hz sizing in bursts).
constexpr double kBucketRateMinHz = 200.0;
constexpr double kBucketRateMaxHz = 500.0;

// Default of the run-aggregated LDS counters (naer_schedule.run_aggregate):
// bit 0 (histogram and refine counters, keyed by tile) is on when a wave of
// 64 row synapses is expected to hold runs of >= 4 equal keys, i.e. a row's
// synapses per (tile, arrival batch) of its own network; bit 1 (placement
// counters, keyed by 2^14-neuron group; rows are ordered by group, then
// delay) when runs per (group, arrival batch) reach 8 (on the 1M network's
// runs of ~4 the aggregated placement measured 20 % slower).  Weighted by synapses
// over the networks of a pack.  Arenas (networks unknown at build) default
// on; the kernels fall back to lane-by-lane atomics on waves of short runs.
uint32_t run_aggregation(const naer_net* n)
{
    double wt = 0.0, wg = 0.0, sum = 0.0;
    for (const auto& net : n->nets) {
        if (net.n_neurons == 0 || net.n_syn == 0)
            continue;
        const double tiles = std::ceil((double)net.n_neurons / (double)(1u << n->tw_shift));
        const double groups = std::ceil((double)net.n_neurons / (double)(1u << kGroupNeuronShift));
        const double batches = (double)(net.dmax - std::min(net.dmin, net.dmax)) / n->B + 1.0;
        const double row = (double)net.n_syn / net.n_neurons;
        wt += (double)net.n_syn * row / (tiles * batches);
        wg += (double)net.n_syn * row / (groups * batches);
        sum += (double)net.n_syn;
    }
    if (sum == 0.0)
        return 3u;
    return (wt / sum >= 4.0 ? 1u : 0u) | (wg / sum >= 8.0 ? 2u : 0u);
}

// Class-weight delivery (kRec4*): possible when no synapse is plastic,
// the net is no arena (its jobs are not known at build), every population's
// outgoing connections share one constant weight, and the populations fit
// the record's class bits.
bool class_weights(const naer_net* n)
{
    if (n->arena || n->plastic)
        return false;
    uint64_t pops = 0;
    for (const auto& net : n->nets) {
        pops += net.pops.size();
        for (uint32_t p = 0; p < net.pops.size(); ++p) {
            bool have = false;
            float w = 0.0f;
            for (const auto& c : net.conns) {
                if (c.src_pop != p)
                    continue;
                if (c.weight == 0.0f || c.plastic || (have && c.weight != w))
                    return false;  // random, plastic or unequal weights
                have = true;
                w = c.weight;
            }
        }
    }
    return pops < kRec4MaxClasses;
}

// Segment ranges of the histogram/placement blocks.  Packs and arenas
// (networks of unequal size and rate side by side) split every batch by its
// events on the device (k_bucket_split); a single network uses fixed ranges
// of G segments, set here.
int set_block_ranges(naer_net* n)
{
    if (n->balance)
        return NAER_OK;
    std::vector<uint32_t> h(n->NB + 1);
    for (uint32_t b = 0; b <= n->NB; ++b)
        h[b] = (uint32_t)std::min<uint64_t>((uint64_t)b * n->G, n->n_seg);
    HIP_TRY(hipMemcpy(n->blk_seg, h.data(), h.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    return NAER_OK;
}

// The placement keeps one LDS counter per group bucket: at most 8192 (32 KB).
constexpr uint32_t kMaxPlaceCounters = 8192;

// Direct placement (schedule refine_split 0): groups of a single tile, so the
// placement writes every event straight into its tile bucket and no refine
// pass runs.  It pays where a placement block holds many events per tile
// bucket (dense packs, arenas: runs of hundreds of records), and needs one
// placement counter per tile bucket up to the last active tile of the last
// arrival batch: (M - 1) NT + n_act of them.  An arena's occupied prefix
// changes with its jobs: a batch whose counters do not fit is placed in two
// levels.
uint32_t direct_counters(const naer_net* n, uint32_t tiles)
{
    return (n->M - 1) * n->NT + tiles;
}

bool direct_ok(const naer_net* n)
{
    return direct_counters(n, n->n_act) <= kMaxPlaceCounters;
}

// refine blocks per group bucket of the two-level placement by default:
// about 2048 refine blocks in all
uint32_t default_refine_split(const naer_net* n)
{
    return std::max<uint32_t>(1, std::min<uint32_t>(64, 2048 / std::max<uint32_t>(1, n->NBK >> n->gt_shift)));
}

// Reserved placement (see BucketArgs): no counting passes; each placement
// block writes its events into regions of the staging buffer reserved from
// the batch's event counts.  The share of events reserved for group bucket
// (m, g) is 1.25x the largest fraction of the network's synapses that can
// arrive there (delay d from a spike at step r of the batch arrives in batch
// m = (r + d) / B; the largest over r), measured once over the built CSR.
constexpr double kResMargin = 1.25;  // measured on the 1M network: 1.2 beat 1.5 by ~1 %, 2.0 lost ~0.5 %

// its parts per group bucket: refine blocks of about 4k events or fewer
uint32_t default_res_parts(const naer_net* n)
{
    return std::max<uint32_t>(1, std::min<uint32_t>(n->parts, 4));
}

// Reserved placement is the default where a batch has at least this many
// group buckets: its refine has one block per part of a group bucket, so a
// net with few groups (up to ~10^5 neurons) keeps the counted two-level
// placement, whose refine splits the whole batch evenly (measured: 1M
// neurons, 558 group buckets: +6 % events/s; 10^5 neurons with 70 group
// buckets: -8 %)
constexpr uint32_t kResDefaultGroups = 512;

bool res_fits(const naer_net* n, uint32_t G, uint32_t RS)
{
    const uint32_t NB = (n->n_seg + G - 1) / G;
    return n->res_ok && RS >= 1 && RS <= n->parts && (NB + RS - 1) / RS + 1 <= kResMaxPartRegions;
}

int setup_reserve(naer_net* n)
{
    n->parts = 1;
    n->res_ok = false;
    if (n->arena || n->nets.size() != 1 || n->S_cap == 0 || n->dmax < n->dmin)
        return NAER_OK;
    const uint32_t ng = n->NT >> n->gt_shift, ngb = n->NBK >> n->gt_shift, gt = 1u << n->gt_shift;
    const uint32_t nd = n->dmax - n->dmin + 1;
    // the update's ranges per tile, and the placement's LDS (kResCounterWords words per counter)
    const uint32_t parts = std::min(kResMaxParts, kResMaxRanges / n->M);
    if ((uint64_t)nd * ng > kSynStatsBins || (uint64_t)kResCounterWords * ngb * sizeof(uint32_t) > 32768 || parts == 0)
        return NAER_OK;
    int rc;
    unsigned long long* d_h = nullptr;
    const size_t hb = (size_t)nd * ng * sizeof(unsigned long long);
    HIP_TRY(hipMalloc(&d_h, hb));
    std::vector<unsigned long long> h((size_t)nd * ng);
    hipError_t e = hipMemsetAsync(d_h, 0, hb, n->stream);
    if (e == hipSuccess)
        e = launch_syn_stats(n->col, n->S_cap, n->dmin, nd, ng, d_h, n->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(h.data(), d_h, hb, hipMemcpyDeviceToHost, n->stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(n->stream);
    (void)hipFree(d_h);
    if (e != hipSuccess)
        return fail(NAER_ERR_HIP, "synapse statistics failed: %s", hipGetErrorString(e));
    std::vector<uint32_t> c(ngb, 0);
    std::vector<unsigned long long> cc(ngb + 1, 0);
    const double S = (double)n->S_cap;
    // NAER_RESERVE_MARGIN overrides the margin (tests use small ones to send
    // events into the spill regions and the ring)
    double margin = kResMargin;
    if (const char* env = std::getenv("NAER_RESERVE_MARGIN"))
        margin = std::max(0.0, std::strtod(env, nullptr));
    for (uint32_t m = 1; m <= n->M; ++m)
        for (uint32_t g = 0; g < ng; ++g) {
            unsigned long long most = 0;
            for (uint32_t r = 0; r < n->B; ++r) {
                unsigned long long f = 0;
                for (uint32_t i = 0; i < nd; ++i)
                    if ((r + n->dmin + i) / n->B == m)
                        f += h[(size_t)i * ng + g];
                most = std::max(most, f);
            }
            if (most)
                c[(m - 1) * ng + g] = (uint32_t)std::ceil(margin * (double)most / S * (double)(1u << kResShift));
        }
    for (uint32_t k = 0; k < ngb; ++k)
        cc[k + 1] = cc[k] + c[k];
    if (cc[ngb] > kResMaxShare)
        return NAER_OK;
    if ((rc = n->alloc(&n->res_c, ngb)) || (rc = n->alloc(&n->res_cc, ngb + 1)) ||
        (rc = n->alloc(&n->res_cnt, (size_t)n->n_seg * ngb)) || (rc = n->alloc(&n->res_spill, ngb)) ||
        (rc = n->alloc(&n->res_pb, (size_t)n->n_seg + 1)) ||
        (rc = n->alloc(&n->res_tb, (size_t)n->R * ngb * parts * (gt + 1))))
        return rc;
    HIP_TRY(hipMemcpy(n->res_c, c.data(), ngb * sizeof(uint32_t), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(n->res_cc, cc.data(), (ngb + 1) * sizeof(unsigned long long), hipMemcpyHostToDevice));
    HIP_TRY(hipMemset(n->res_spill, 0, ngb * sizeof(uint32_t)));
    HIP_TRY(hipMemset(n->res_tb, 0, (size_t)n->R * ngb * parts * (gt + 1) * sizeof(uint32_t)));
    n->parts = parts;
    n->res_ok = true;
    // the default only where the reserved layout of a batch of a quarter of
    // the event buffer's capacity still fits it (the layout takes about twice
    // a batch's events; a batch that does not fit goes through the ring)
    if (ngb >= kResDefaultGroups && res_fits(n, n->G, default_res_parts(n)) &&
        res_base(cc[ngb], n->e_cap / 4, ngb, n->NB) <= n->e_cap) {
        n->RES = 1;
        n->RS = default_res_parts(n);
    }
    return NAER_OK;
}

int alloc_buckets(naer_net* n)
{
    n->bucketed = false;
    if (n->B < 2 || n->S_cap == 0 || n->event)
        return NAER_OK;
    n->M = (n->B - 1 + n->dmax) / n->B;
    n->gt_shift = kGroupNeuronShift - n->tw_shift;
    n->NT = (n->n_seg + (1u << n->gt_shift) - 1) >> n->gt_shift << n->gt_shift;
    n->NBK = n->M * n->NT;
    if (n->NBK > kMaxBuckets)
        return NAER_OK;
    n->R = n->M + 1;
    const double rate = std::min(kBucketRateMaxHz,
                                 std::max(kBucketRateMinHz, 1e3 / ((n->nref_min + 1.0) * (double)n->dt_ms)));
    const double expect = (double)n->S_cap * n->B * n->dt_ms * 1e-3 * rate;
    // the R + 1 event buffers take at most a quarter of the free device memory
    size_t free_b = 0, total_b = 0;
    HIP_TRY(hipMemGetInfo(&free_b, &total_b));
    const double fit = (double)free_b / 4.0 / (8.0 * (n->R + 1));
    n->e_cap = (uint64_t)std::min(std::min(std::max(expect, (double)(1u << 20)), fit), (double)0x7FFFFFFFu);
    // NAER_BUCKET_EVENTS overrides the capacity (tests use a small one to
    // drive batches through the ring fallback)
    if (const char* env = std::getenv("NAER_BUCKET_EVENTS")) {
        const unsigned long long v = std::strtoull(env, nullptr, 10);
        if (v > 0)
            n->e_cap = std::min<unsigned long long>(v, 0x7FFFFFFFull);
    }
    // default schedule (naer_net_autotune may change it): about 256
    // histogram/placement blocks, 512-synapse tasks, ~2048 refine blocks
    n->G = std::max<uint32_t>(1, (n->n_seg + 255) / 256);
    n->NB = (n->n_seg + n->G - 1) / n->G;
    n->U = kBucketUnroll;
    // packs and arenas place directly whenever their (occupied) tiles fit the
    // counters: they hold many events per (placement block, tile bucket)
    // (measured: the sweep's packs alone 13-25 % faster per batch, serving at
    // 10-20 jobs/s 2.6-5 % more jobs/s); a single net keeps two levels
    const bool pack_fits = n->nets.size() > 1 && direct_counters(n, n->n_seg) <= kMaxPlaceCounters;
    n->RS = (n->arena || pack_fits) ? 0u : default_refine_split(n);
    n->AG = run_aggregation(n);
    // placement counters: one per group bucket in LDS (<= 32 KB without raising the limit)
    if (n->M > kMaxArrivalBatches || (n->NBK >> n->gt_shift) > kMaxPlaceCounters)
        return NAER_OK;
    int rc;
    n->rec_bytes = class_weights(n) ? 4u : 8u;
    // 4-byte records: buffers of uint2 holding twice as many records
    const size_t rec_pairs = n->rec_bytes == 4 ? (size_t)(n->e_cap + 1) / 2 : (size_t)n->e_cap;
    if (n->rec_bytes == 4 && (rc = n->alloc(&n->ncls, n->N)))
        return rc;
    if ((rc = n->alloc(&n->ev, (size_t)n->R * rec_pairs)) || (rc = n->alloc(&n->stage, rec_pairs)) ||
        (rc = n->alloc(&n->bs, (size_t)n->R * (n->NBK + 1))) || (rc = n->alloc(&n->flags, n->R)) ||
        (rc = n->alloc(&n->hist, (size_t)n->n_seg * n->NBK)) ||  // rows for any G (G = 1: n_seg blocks)
        (rc = n->alloc(&n->hist_c, (size_t)n->n_seg * (n->NBK >> n->gt_shift))) ||
        (rc = n->alloc(&n->totals, n->NBK)) ||
        (rc = n->alloc(&n->gcur, n->NBK)))
        return rc;
    HIP_TRY(hipMemsetAsync(n->flags, 0, n->R * sizeof(uint32_t), n->stream));
    n->balance = n->arena || n->nets.size() > 1;
    if ((rc = set_block_ranges(n)) || (rc = setup_reserve(n)))
        return rc;
    n->bucketed = true;
    return NAER_OK;
}

void set_scatter_shape(naer_net* n)
{
    n->scatter_chunks = std::max<uint32_t>(1, (n->max_row + kChunk - 1) / kChunk);
}

int launch_timed(naer_net* n, int kind, hipError_t (*fn)(naer_net*, void*), void* arg, bool timed)
{
    hipEvent_t a = nullptr, b = nullptr;
    timed = timed && n->profiling;
    if (timed) {
        a = n->take_event();
        b = n->take_event();
        HIP_TRY(hipEventRecord(a, n->stream));
    }
    hipError_t e = fn(n, arg);
    if (e != hipSuccess)
        return fail(NAER_ERR_HIP, "kernel launch failed: %s", hipGetErrorString(e));
    n->kernel_calls[kind] += 1;
    if (timed) {
        HIP_TRY(hipEventRecord(b, n->stream));
        n->pending.push_back(PendingTimer{ kind, a, b });
    }
    return NAER_OK;
}

int make_stream(naer_net* n)
{
    hipError_t e = hipSetDevice(n->ctx->device);
    if (e == hipSuccess)
        e = hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking);
    if (e != hipSuccess)
        return fail(NAER_ERR_HIP, "stream creation failed: %s", hipGetErrorString(e));
    return NAER_OK;
}

}  // namespace

extern "C" {

const char* naer_last_error(void) { return g_err.c_str(); }

const char* naer_version(void) { return "naer-mi355x 0.2 (gfx950)"; }

// Static scratch of naer_ref_poisson_epoch_noalloc: staging for its
// outputs, one copy per GPU (held until the process ends).  Host code uses
// only the device address hipGetSymbolAddress gives for the current device,
// never the array's name.
constexpr uint64_t kScratchWords = 1u << 20;  // 4 MiB
static __device__ __align__(256) uint32_t scratch[kScratchWords];
static std::mutex g_scratch_mu;
static std::map<int, uint32_t*> g_scratch;  // device -> its copy of `scratch`

// The device address of `scratch` on `device` (current), looked up once per device.
static int scratch_of(int device, uint32_t** out)
{
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    auto it = g_scratch.find(device);
    if (it == g_scratch.end()) {
        uint32_t* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(NAER_ERR_HIP, "no address of the static scratch on device %d: %s", device,
                        hipGetErrorString(e));
        it = g_scratch.emplace(device, ptr).first;
    }
    *out = it->second;
    return NAER_OK;
}

int naer_ctx_create(int device, naer_ctx** out)
{
    if (!out)
        return fail(NAER_ERR_INVALID, "out is NULL");
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0)
        return fail(NAER_ERR_NODEVICE, "no HIP device available (%s)", hipGetErrorString(e));
    if (device < 0 || device >= count)
        return fail(NAER_ERR_NODEVICE, "device %d out of range (%d devices)", device, count);
    HIP_TRY(hipSetDevice(device));
    // the static scratch's address on this device, with the device current
    uint32_t* sp = nullptr;
    const int rc = scratch_of(device, &sp);
    if (rc)
        return rc;
    naer_ctx* c = new naer_ctx;
    c->device = device;
    *out = c;
    return NAER_OK;
}

int naer_ctx_destroy(naer_ctx* ctx)
{
    delete ctx;
    return NAER_OK;
}

// Fewest refractory steps of the LIF populations of the active networks.
uint32_t min_refractory(const naer_net* n)
{
    int64_t m = -1;
    for (const auto& net : n->nets) {
        if (!net.active)
            continue;
        for (const auto& q : net.pops) {
            if (q.kind != NAER_POP_LIF)
                continue;
            const int64_t r = std::max<int64_t>(0, std::llround((double)q.t_ref_ms / (double)n->dt_ms));
            m = m < 0 ? r : std::min(m, r);
        }
    }
    return m < 0 ? 0u : (uint32_t)std::min<int64_t>(m, 1 << 20);
}

int naer_pack_create(naer_ctx* ctx, const naer_net_desc* descs, uint32_t n_nets, naer_net** out)
{
    if (!ctx || !descs || !out || n_nets == 0)
        return fail(NAER_ERR_INVALID, "NULL argument or empty pack");
    naer_net* n = new naer_net;
    n->ctx = ctx;
    int rc = check_settings(n, descs[0].dt_ms, descs[0].mode);
    if (rc == NAER_OK && n->event && n_nets > 1)
        rc = fail(NAER_ERR_INVALID, "the event-driven mode runs one network");
    uint64_t N = 0, rec = 0;
    n->dmin = 0xFFFFFFFFu;
    n->dmax = 1;
    for (uint32_t i = 0; rc == NAER_OK && i < n_nets; ++i) {
        NetInfo info;
        if ((rc = check_net(n, descs[i], &info)))
            break;
        if (info.plastic) {
            if (n->stdp_set && std::memcmp(&n->stdp, &descs[i].stdp, sizeof n->stdp) != 0) {
                rc = fail(NAER_ERR_INVALID, "plastic networks of a pack must share the STDP parameters");
                break;
            }
            n->stdp = descs[i].stdp;
            n->stdp_set = true;
            n->plastic = true;
        }
        info.neuron_begin = (uint32_t)N;
        info.active = true;
        N += info.n_neurons;
        n->dmin = std::min(n->dmin, info.dmin);
        n->dmax = std::max(n->dmax, info.dmax);
        rec += descs[i].record_capacity;
        n->nets.push_back(std::move(info));
        if (N > 0xFFFFFF00ull)
            rc = fail(NAER_ERR_CAPACITY, "a pack is limited to 2^32 neurons");
    }
    if (rc == NAER_OK) {
        n->N = (uint32_t)N;
        n->D = n->dmax + 1;
        n->nref_min = min_refractory(n);
        n->B = batch_steps(n->plastic, n->dmin, n->N, n->nref_min);
        n->rec_cap = rec;
        rc = make_stream(n);
    }
    if (rc != NAER_OK) {
        delete n;
        return rc;
    }
    *out = n;
    return NAER_OK;
}

int naer_net_create(naer_ctx* ctx, const naer_net_desc* desc, naer_net** out)
{
    if (!desc)
        return fail(NAER_ERR_INVALID, "NULL argument");
    return naer_pack_create(ctx, desc, 1, out);
}

int naer_pack_layout(naer_net* n, uint32_t cap, uint32_t* neuron_begin, uint32_t* n_neurons, uint64_t* t_start,
                     uint32_t* n_slots)
{
    if (!n)
        return fail(NAER_ERR_INVALID, "net is NULL");
    if (n_slots)
        *n_slots = (uint32_t)n->nets.size();
    for (size_t i = 0; i < std::min<size_t>(cap, n->nets.size()); ++i) {
        if (neuron_begin)
            neuron_begin[i] = n->nets[i].neuron_begin;
        if (n_neurons)
            n_neurons[i] = n->nets[i].active ? n->nets[i].n_neurons : 0u;
        if (t_start)
            t_start[i] = n->nets[i].t_start;
    }
    return NAER_OK;
}

// Event-driven mode: input lists, event buffers and the decay tables of every
// population (the oracle builds the same tables, orc_ev_tables).
static int setup_events(naer_net* n)
{
    if (!n->event)
        return NAER_OK;
    // event keys hold target << 11 and input keys the source in 21 bits
    if (n->N >= (1u << kEvSrcBits))
        return fail(NAER_ERR_CAPACITY, "event-driven mode: fewer than 2^21 neurons");
    // plastic nets: an event's synapse index is 31 bits
    if (n->plastic && n->S_cap >= 0x80000000ull)
        return fail(NAER_ERR_CAPACITY, "event-driven mode: plastic nets of fewer than 2^31 synapses");
    int rc;
    const uint32_t np = (uint32_t)n->pp.size();
    // a POISSON source emits at most 10 spikes per window, a LIF neuron one
    // per refractory period
    uint64_t in_cap = 1024;
    for (auto& net : n->nets)
        for (auto& q : net.pops) {
            const double tr = std::max(1.0, (double)q.t_ref_ms * 1000.0);
            in_cap += q.kind == NAER_POP_POISSON ? (uint64_t)q.size * kMaxPoissonPerEpoch
                                                 : (uint64_t)q.size * (uint64_t)(kEpochUs / tr + 1.0);
        }
    if (in_cap >= 0x7FFFFFFFull)
        return fail(NAER_ERR_CAPACITY, "event-driven mode: too many inputs per window");
    n->ev_in_cap = (uint32_t)in_cap;
    size_t free_b = 0, total_b = 0;
    HIP_TRY(hipMemGetInfo(&free_b, &total_b));
    // a quarter of the free memory for the events (key and weight, plastic
    // nets also synapse and slot; sorted and not); event indices are 32-bit
    const double fit = (double)free_b / 4.0 / (n->plastic ? 32.0 : 16.0);
    n->ev_e_cap = (uint64_t)std::min(std::min(std::max((double)n->S_cap, (double)(1u << 22)), fit), 2147483647.0);
    // the step kernel reads a target's run in aligned quads: up to 3 entries past the last event
    const uint64_t ev_pad = n->ev_e_cap + 4;
    if ((rc = n->alloc(&n->ev_list[0], in_cap)) || (rc = n->alloc(&n->ev_list[1], in_cap)) ||
        (rc = n->alloc(&n->ev_sorted, in_cap)) || (rc = n->alloc(&n->ev_len, in_cap + 1)) ||
        (rc = n->alloc(&n->ev_off, in_cap + 1)) || (rc = n->alloc(&n->ev_keys, ev_pad)) ||
        (rc = n->alloc(&n->ev_keys_s, ev_pad)) || (rc = n->alloc(&n->ev_w, ev_pad)) ||
        (rc = n->alloc(&n->ev_w_s, ev_pad)) || (rc = n->alloc(&n->ev_tlast, n->N)) ||
        (rc = n->alloc(&n->ev_begin, n->N)) || (rc = n->alloc(&n->ev_end, n->N)) ||
        (rc = n->alloc(&n->ev_dlo, (size_t)std::max(1u, np) * kEvLo)) ||
        (rc = n->alloc(&n->ev_dhi, (size_t)std::max(1u, np) * kEvHiCap)) || (rc = n->alloc(&n->ev_nhi, np)) ||
        (rc = n->alloc(&n->ev_tref, np)) || (rc = n->alloc(&n->ev_ctr, 8)))
        return rc;
    if (n->plastic) {
        if ((rc = n->alloc(&n->ev_vals[0], in_cap)) || (rc = n->alloc(&n->ev_vals[1], in_cap)) ||
            (rc = n->alloc(&n->ev_vals_s, in_cap)) || (rc = n->alloc(&n->ev_prev[0], in_cap)) ||
            (rc = n->alloc(&n->ev_prev[1], in_cap)) || (rc = n->alloc(&n->ev_ptr[0], in_cap)) ||
            (rc = n->alloc(&n->ev_ptr[1], in_cap)) || (rc = n->alloc(&n->ev_syn, ev_pad)) ||
            (rc = n->alloc(&n->ev_syn_s, ev_pad)) || (rc = n->alloc(&n->ev_slot, ev_pad)) ||
            (rc = n->alloc(&n->ev_slot_s, ev_pad)) || (rc = n->alloc(&n->ev_trace, n->N)) ||
            (rc = n->alloc(&n->ev_hist, (size_t)n->N * (kEvHist / 2))) ||
            (rc = n->alloc(&n->ev_pop_plastic, std::max(1u, np))))
            return rc;
        // populations with a plastic connection (stdp.cl's neurons), in the
        // population table's order (non-empty populations)
        std::vector<uint32_t> pl(np, 0);
        uint32_t base = 0;
        for (auto& net : n->nets) {
            std::vector<uint32_t> at(net.pops.size(), 0);
            uint32_t q = base;
            for (size_t k = 0; k < net.pops.size(); ++k) {
                at[k] = q;
                q += net.pops[k].size ? 1u : 0u;
            }
            for (auto& k : net.conns)
                if (k.plastic && net.pops[k.src_pop].size && net.pops[k.dst_pop].size)
                    pl[at[k.src_pop]] = pl[at[k.dst_pop]] = 1;
            base = q;
        }
        HIP_TRY(hipMemcpy(n->ev_pop_plastic, pl.data(), np * sizeof(uint32_t), hipMemcpyHostToDevice));
        HIP_TRY(hipMemset(n->ev_trace, 0, n->N * sizeof(float)));
        HIP_TRY(hipMemset(n->ev_hist, 0, (size_t)n->N * (kEvHist / 2) * sizeof(uint32_t)));
    }
    std::vector<double> lo((size_t)np * kEvLo, 0.0), hi((size_t)np * kEvHiCap, 0.0);
    std::vector<uint32_t> nhi(np, 0), tref(np, 0);
    uint32_t p = 0;
    for (auto& net : n->nets)
        for (auto& q : net.pops) {
            if (q.size == 0)
                continue;
            if (q.kind == NAER_POP_LIF) {
                const double tau = (double)q.tau_m_ms * 1000.0;
                for (uint32_t l = 0; l < kEvLo; ++l)
                    lo[(size_t)p * kEvLo + l] = std::exp(-(double)l / tau);
                uint32_t h = 0;
                for (; h < kEvHiCap; ++h) {
                    const double v = std::exp(-(double)(h * kEvLo) / tau);
                    if (v == 0.0)
                        break;
                    hi[(size_t)p * kEvHiCap + h] = v;
                }
                nhi[p] = h;
                tref[p] = (uint32_t)std::llround((double)q.t_ref_ms * 1000.0);
            }
            ++p;
        }
    HIP_TRY(hipMemcpy(n->ev_dlo, lo.data(), lo.size() * sizeof(double), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(n->ev_dhi, hi.data(), hi.size() * sizeof(double), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(n->ev_nhi, nhi.data(), np * sizeof(uint32_t), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(n->ev_tref, tref.data(), np * sizeof(uint32_t), hipMemcpyHostToDevice));
    HIP_TRY(hipMemset(n->ev_tlast, 0, n->N * sizeof(uint32_t)));
    HIP_TRY(hipMemset(n->ev_begin, 0, n->N * sizeof(uint32_t)));
    HIP_TRY(hipMemset(n->ev_end, 0, n->N * sizeof(uint32_t)));
    HIP_TRY(hipMemset(n->ev_ctr, 0, 8 * sizeof(uint32_t)));
    // workspaces of the sorts (inputs and events) and of the row scan
    const uint32_t sort_max = (uint32_t)std::max<uint64_t>(n->ev_in_cap, n->ev_e_cap);
    if ((rc = n->alloc(&n->ev_sort_ws, ev_sort_space(sort_max))) ||
        (rc = n->alloc(&n->ev_scan_ws, ev_scan_space(n->ev_in_cap + 1))))
        return rc;
    return NAER_OK;
}

static EvArgs ev_args(const naer_net* n)
{
    EvArgs a;
    std::memset(&a, 0, sizeof a);
    a.pops = n->d_pops;
    a.blk_pop = n->d_blk_pop;
    a.n_pops = (uint32_t)n->pp.size();
    a.tw_shift = n->tw_shift;
    a.N = n->N;
    a.t0 = (uint32_t)(n->ev_epoch * kEpochUs);
    a.epoch = (uint32_t)n->ev_epoch;
    a.in_cap = n->ev_in_cap;
    a.row_ptr = n->row_ptr;
    a.col = n->col;
    a.w = n->w;
    a.lost = n->ev_ctr + 2;
    a.row_len = n->ev_len;
    a.row_off = n->ev_off;
    a.V = n->V;
    a.t_last = n->ev_tlast;
    a.ev_begin = n->ev_begin;
    a.ev_end = n->ev_end;
    a.dec_lo = n->ev_dlo;
    a.dec_hi = n->ev_dhi;
    a.n_hi = n->ev_nhi;
    a.t_ref = n->ev_tref;
    a.rec = n->rec;
    a.rec_cap = n->rec_cap;
    a.rec_n = (unsigned long long*)((char*)n->st + offsetof(DevState, rec_n));
    a.spikes = (unsigned long long*)(n->ev_ctr + 4);
    if (n->plastic) {
        a.w_mut = n->w;
        a.trace = n->ev_trace;
        a.hist = n->ev_hist;
        a.pop_plastic = n->ev_pop_plastic;
    }
    return a;
}

// Runs `epochs` windows; two host reads per window (the input and event
// counts size the sorts).  Inputs are sorted on all 32 key bits (time offset,
// source); events on their target bits only, stably, so that each target's
// events keep the expansion order (time, source, row position).
static int ev_step(naer_net* n, uint32_t epochs)
{
    uint32_t tgt_bits = 1;
    while (tgt_bits < 21 && (n->N - 1) >> tgt_bits)
        ++tgt_bits;
    for (uint32_t k = 0; k < epochs; ++k) {
        if (n->ev_epoch * kEpochUs + 2 * kEpochUs > 0xFFFFFFFFull)
            return fail(NAER_ERR_CAPACITY, "event-driven mode: microsecond clock exhausted");
        const uint32_t cur = (uint32_t)(n->ev_epoch & 1u);
        EvArgs a = ev_args(n);
        a.out_keys = n->ev_list[cur];
        a.n_out = n->ev_ctr + cur;
        if (n->plastic)
            a.out_vals = n->ev_vals[cur];  // POISSON inputs carry no payload
        HIP_TRY(ev_launch(a, kEvPoisson, 0, n->stream));
        uint32_t c[3];
        HIP_TRY(hipMemcpyAsync(c, n->ev_ctr, sizeof c, hipMemcpyDeviceToHost, n->stream));
        HIP_TRY(hipStreamSynchronize(n->stream));
        if (c[2])
            return fail(NAER_ERR_CAPACITY, "event-driven mode: %u inputs beyond the list capacity %u", c[2],
                        n->ev_in_cap);
        const uint32_t n_in = c[cur];
        bool alt = false;
        uint32_t* const in_v[1] = { n->ev_vals[cur] };
        uint32_t* const in_v_alt[1] = { n->ev_vals_s };
        HIP_TRY(ev_radix_sort(n->ev_list[cur], n->ev_sorted, in_v, in_v_alt, n->plastic ? 1u : 0u, n_in, 0, 32,
                              n->ev_sort_ws, &alt, n->stream));
        a.in_keys = alt ? n->ev_sorted : n->ev_list[cur];
        a.out_vals = nullptr;
        if (n->plastic) {
            a.in_vals = alt ? n->ev_vals_s : n->ev_vals[cur];
            a.in_prev = n->ev_prev[cur];
            a.in_trace = n->ev_ptr[cur];
        }
        HIP_TRY(ev_launch(a, kEvRowLen, n_in, n->stream));
        HIP_TRY(ev_scan_u64(n->ev_len, n->ev_off, n_in + 1, n->ev_scan_ws, n->stream));
        unsigned long long n_ev = 0;
        HIP_TRY(hipMemcpyAsync(&n_ev, n->ev_off + n_in, sizeof n_ev, hipMemcpyDeviceToHost, n->stream));
        HIP_TRY(hipStreamSynchronize(n->stream));
        if (n_ev > n->ev_e_cap)
            return fail(NAER_ERR_CAPACITY, "event-driven mode: %llu events in a window (capacity %llu)", n_ev,
                        (unsigned long long)n->ev_e_cap);
        a.ev_keys = n->ev_keys;
        a.ev_w = n->ev_w;
        a.ev_syn = n->ev_syn;
        a.ev_slot = n->ev_slot;
        HIP_TRY(ev_launch(a, kEvExpand, n_in, n->stream));
        uint32_t* const ev_v[3] = { reinterpret_cast<uint32_t*>(n->ev_w), n->ev_syn, n->ev_slot };
        uint32_t* const ev_v_alt[3] = { reinterpret_cast<uint32_t*>(n->ev_w_s), n->ev_syn_s, n->ev_slot_s };
        HIP_TRY(ev_radix_sort(n->ev_keys, n->ev_keys_s, ev_v, ev_v_alt, n->plastic ? 3u : 1u, (uint32_t)n_ev, 11,
                              11 + tgt_bits, n->ev_sort_ws, &alt, n->stream));
        EvArgs b = a;
        b.ev_keys = alt ? n->ev_keys_s : n->ev_keys;
        b.ev_w = alt ? n->ev_w_s : n->ev_w;
        b.ev_syn = alt ? n->ev_syn_s : n->ev_syn;
        b.ev_slot = alt ? n->ev_slot_s : n->ev_slot;
        if (n->plastic) {
            b.out_vals = n->ev_vals[cur ^ 1u];
            b.out_prev = n->ev_prev[cur ^ 1u];
            b.out_trace = n->ev_ptr[cur ^ 1u];
        }
        b.out_keys = n->ev_list[cur ^ 1u];
        b.n_out = n->ev_ctr + (cur ^ 1u);
        b.n_ev = (uint32_t)n_ev;
        HIP_TRY(hipMemsetAsync(n->ev_ctr + (cur ^ 1u), 0, sizeof(uint32_t), n->stream));
        HIP_TRY(ev_launch(b, kEvProcess, 0, n->stream));
        n->ev_events += n_ev;
        n->ev_epoch += 1;
    }
    return NAER_OK;
}

// Populations of source-defined models: staged-input buffer, per-population
// state (s[0] = the initial potential init_state drew) and loaded modules.
static int setup_models(naer_net* n)
{
    bool any = false;
    for (auto& net : n->nets)
        for (auto& q : net.pops)
            any = any || (q.kind == NAER_POP_MODEL && q.size);
    if (!any)
        return NAER_OK;
    int rc;
    if ((rc = n->alloc(&n->cin, (size_t)n->B * n->N)))
        return rc;
    for (auto& net : n->nets) {
        uint32_t begin = net.neuron_begin;
        for (auto& q : net.pops) {
            if (q.kind == NAER_POP_MODEL && q.size) {
                ModelPop m;
                m.begin = begin;
                m.end = begin + q.size;
                m.n_state = q.model->n_state;
                m.t_start = net.t_start;
                if ((rc = n->alloc(&m.state, (size_t)m.n_state * q.size)) ||
                    (rc = n->alloc(&m.params, NAER_MODEL_PARAMS)))
                    return rc;
                HIP_TRY(hipMemcpyAsync(m.params, q.params, sizeof q.params, hipMemcpyHostToDevice, n->stream));
                HIP_TRY(hipMemsetAsync(m.state, 0, sizeof(float) * m.n_state * q.size, n->stream));
                HIP_TRY(hipMemcpyAsync(m.state, n->V + begin, sizeof(float) * q.size, hipMemcpyDeviceToDevice,
                                       n->stream));
                HIP_TRY(hipModuleLoadData(&m.mod, q.model->code.data()));
                n->mpops.push_back(m);  // the destructor unloads the module
                HIP_TRY(hipModuleGetFunction(&n->mpops.back().fn, n->mpops.back().mod, "naer_model_kernel"));
            }
            begin += q.size;
        }
    }
    return NAER_OK;
}

int naer_net_build(naer_net* n)
{
    if (!n)
        return fail(NAER_ERR_INVALID, "net is NULL");
    if (n->built || n->arena)
        return fail(NAER_ERR_STATE, "net already built");
    HIP_TRY(hipSetDevice(n->ctx->device));
    bool any_pois = false;
    for (auto& net : n->nets)
        for (auto& q : net.pops)
            any_pois = any_pois || q.kind == NAER_POP_POISSON;
    int rc;
    n->poisson_ok = any_pois;
    if ((rc = alloc_neurons(n, any_pois)))
        return rc;
    // 1. out-degrees of every network, rows back to back
    std::vector<std::vector<unsigned long long>> prefix(n->nets.size());
    uint64_t S = 0;
    for (size_t i = 0; i < n->nets.size(); ++i) {
        NetInfo& net = n->nets[i];
        if ((rc = count_net(n, net)) ||
            (rc = read_prefix(n, net.neuron_begin, net.n_neurons, &prefix[i], &net.max_row)))
            return rc;
        if (net.max_row > kMaxSortedRow)
            return fail(NAER_ERR_CAPACITY, "a source has %u synapses; rows are limited to %u", net.max_row,
                        kMaxSortedRow);
        net.syn_begin = S;
        net.n_syn = prefix[i].back();
        S += net.n_syn;
        n->max_row = std::max(n->max_row, net.max_row);
    }
    n->S_cap = S;
    if ((rc = n->alloc(&n->col, S)) || (rc = n->alloc(&n->w, S)))
        return rc;
    // 2. placement and canonical row order
    for (size_t i = 0; i < n->nets.size(); ++i) {
        NetInfo& net = n->nets[i];
        if ((rc = write_ptr(n, n->row_ptr, net.neuron_begin, net.syn_begin, prefix[i])) ||
            (rc = place_net(n, net)))
            return rc;
    }
    // 3. plastic synapse counts
    if (n->plastic)
        for (auto& net : n->nets)
            if (net.plastic && (rc = count_plastic(n, net)))
                return rc;
    // 4. delivery buffers, neuron state
    if ((rc = alloc_buckets(n)) || (rc = upload_table(n)))
        return rc;
    for (auto& net : n->nets)
        if ((rc = init_net(n, net)))
            return rc;
    if ((rc = setup_models(n)) || (rc = setup_events(n)))
        return rc;
    HIP_TRY(hipStreamSynchronize(n->stream));
    set_scatter_shape(n);
    n->built = true;
    return NAER_OK;
}

int naer_arena_create(naer_ctx* ctx, const naer_arena_desc* a, naer_net** out)
{
    if (!ctx || !a || !out)
        return fail(NAER_ERR_INVALID, "NULL argument");
    naer_net* n = new naer_net;
    n->ctx = ctx;
    n->arena = true;
    int rc = check_settings(n, a->dt_ms, a->mode);
    if (rc == NAER_OK && n->event)
        rc = fail(NAER_ERR_INVALID, "arenas do not run the event-driven mode");
    if (rc == NAER_OK && (a->max_neurons == 0 || a->max_neurons > 0xFFFFFF00ull))
        rc = fail(NAER_ERR_CAPACITY, "arena neuron capacity must be in [1, 2^32)");
    if (rc == NAER_OK && (a->delay_min < 1 || a->delay_max < a->delay_min || a->delay_max > 127))
        rc = fail(NAER_ERR_INVALID, "arena delays must satisfy 1 <= min <= max <= 127");
    if (rc == NAER_OK && (a->max_row == 0 || a->max_row > kMaxSortedRow))
        rc = fail(NAER_ERR_CAPACITY, "arena max_row must be in [1, %u]", kMaxSortedRow);
    if (rc == NAER_OK && a->plastic &&
        (!(a->stdp.tau_plus_ms > 0.0f) || !(a->stdp.tau_minus_ms > 0.0f) || !(a->stdp.w_max >= a->stdp.w_min)))
        rc = fail(NAER_ERR_INVALID, "STDP needs tau_plus, tau_minus > 0 and w_max >= w_min");
    if (rc == NAER_OK) {
        n->N = (uint32_t)((a->max_neurons + kBlock - 1) / kBlock * kBlock);
        n->S_cap = a->max_synapses;
        n->plastic = a->plastic != 0;
        n->stdp = a->stdp;
        n->stdp_set = n->plastic;
        n->dmin = a->delay_min;
        n->dmax = a->delay_max;
        n->D = n->dmax + 1;
        // jobs arrive later: plastic arenas batch (and flush) for the
        // fewest refractory steps a job may have
        n->nref_min = n->plastic ? a->refractory_min : 0;
        n->B = batch_steps(n->plastic, n->dmin, n->N, n->nref_min);
        n->max_row = a->max_row;
        n->rec_cap = a->record_capacity;
        n->neuron_alloc.reset(n->N);
        n->syn_alloc.reset(n->S_cap);
        rc = make_stream(n);
    }
    n->poisson_ok = a->poisson != 0;
    if (rc == NAER_OK)
        rc = alloc_neurons(n, n->poisson_ok);
    if (rc == NAER_OK)
        rc = n->alloc(&n->col, n->S_cap);
    if (rc == NAER_OK)
        rc = n->alloc(&n->w, n->S_cap);
    if (rc == NAER_OK)
        rc = alloc_buckets(n);
    if (rc == NAER_OK)
        rc = upload_table(n);
    if (rc == NAER_OK) {
        set_scatter_shape(n);
        n->built = true;
        hipError_t e = hipStreamSynchronize(n->stream);
        if (e != hipSuccess)
            rc = fail(NAER_ERR_HIP, "arena set-up failed: %s", hipGetErrorString(e));
    }
    if (rc != NAER_OK) {
        delete n;
        return rc;
    }
    *out = n;
    return NAER_OK;
}

int naer_arena_admit(naer_net* n, const naer_net_desc* desc, uint32_t* slot)
{
    if (!n || !desc || !slot)
        return fail(NAER_ERR_INVALID, "NULL argument");
    if (!n->arena)
        return fail(NAER_ERR_STATE, "not an arena");
    NetInfo net;
    int rc;
    if ((rc = check_net(n, *desc, &net)))
        return rc;
    for (auto& q : net.pops)
        if (q.kind == NAER_POP_MODEL)
            return fail(NAER_ERR_INVALID, "arenas do not admit populations of source-defined models");
    if (net.dmin < n->dmin || net.dmax > n->dmax)
        return fail(NAER_ERR_INVALID, "job delays [%u, %u] outside the arena's [%u, %u]", net.dmin, net.dmax,
                    n->dmin, n->dmax);
    if (net.plastic && (!n->plastic || std::memcmp(&n->stdp, &desc->stdp, sizeof n->stdp) != 0))
        return fail(NAER_ERR_INVALID, "job needs STDP parameters the arena was not created with");
    for (auto& q : net.pops) {
        if (q.kind == NAER_POP_POISSON && !n->poisson_ok)
            return fail(NAER_ERR_INVALID, "the arena was created without POISSON sources");
        if (q.kind == NAER_POP_LIF && n->plastic &&
            std::llround((double)q.t_ref_ms / (double)n->dt_ms) < (long long)n->nref_min)
            return fail(NAER_ERR_INVALID, "job refractory period below the arena's %u steps", n->nref_min);
    }
    // a job larger than the whole arena can never be admitted: that is an
    // invalid job for this arena, not a capacity that retiring others frees
    const uint64_t footprint = ((uint64_t)net.n_neurons + 1 + kBlock - 1) / kBlock * kBlock;
    if (footprint > n->N)
        return fail(NAER_ERR_INVALID, "a job of %u neurons (%llu with its alignment) exceeds the arena's %u",
                    net.n_neurons, (unsigned long long)footprint, n->N);
    // The connectivity is built on a second stream, while steps already
    // enqueued on the arena's stream run: the build writes only the job's
    // free ranges (row_ptr, col, w) and the build scratch, which no step
    // reads.  The population table and the job's state are installed on the
    // arena's stream once those steps are done.
    if (!n->build_stream) {
        const hipError_t e = hipStreamCreateWithFlags(&n->build_stream, hipStreamNonBlocking);
        if (e != hipSuccess)
            return fail(NAER_ERR_HIP, "stream creation failed: %s", hipGetErrorString(e));
    }
    struct BuildStream {  // the build helpers enqueue on n->stream
        naer_net* n;
        hipStream_t main;
        explicit BuildStream(naer_net* net) : n(net), main(net->stream) { n->stream = n->build_stream; }
        ~BuildStream() { n->stream = main; }
    };
    // one extra entry: row_ptr[begin + n] (the end of the job's last
    // row) must not be shared with a job placed right behind it
    uint64_t nb = 0;
    if (!n->neuron_alloc.take((uint64_t)net.n_neurons + 1, kBlock, &nb))
        return fail(NAER_ERR_CAPACITY, "no free range of %u neurons", net.n_neurons);
    net.neuron_begin = (uint32_t)nb;
    // a job starts at the next step of the arena clock
    net.t_start = n->steps_enqueued;
    net.active = true;
    bool syn_taken = false;
    auto undo = [&](int code) {
        n->neuron_alloc.give(net.neuron_begin, (uint64_t)net.n_neurons + 1);
        if (syn_taken)
            n->syn_alloc.give(net.syn_begin, net.n_syn);
        return code;
    };
    std::vector<unsigned long long> prefix;
    {
        BuildStream on_build(n);
        rc = count_net(n, net);
        if (!rc)
            rc = read_prefix(n, net.neuron_begin, net.n_neurons, &prefix, &net.max_row);
    }
    if (rc)
        return undo(rc);
    if (net.max_row > n->max_row)
        return undo(fail(NAER_ERR_INVALID, "a source has %u synapses; the arena admits rows up to %u", net.max_row,
                         n->max_row));
    net.n_syn = prefix.back();
    if (net.n_syn > n->S_cap)
        return undo(fail(NAER_ERR_INVALID, "a job of %llu synapses exceeds the arena's %llu",
                         (unsigned long long)net.n_syn, (unsigned long long)n->S_cap));
    if (!n->syn_alloc.take(net.n_syn, 64, &net.syn_begin))
        return undo(fail(NAER_ERR_CAPACITY, "no free range of %llu synapses", (unsigned long long)net.n_syn));
    syn_taken = true;
    {
        BuildStream on_build(n);
        rc = write_ptr(n, n->row_ptr, net.neuron_begin, net.syn_begin, prefix);
        if (!rc)
            rc = place_net(n, net);
        if (!rc && net.plastic)
            rc = count_plastic(n, net);
        if (!rc && hipStreamSynchronize(n->build_stream) != hipSuccess)
            rc = fail(NAER_ERR_HIP, "connectivity build failed");
    }
    if (rc)
        return undo(rc);
    // the steps enqueued before this admission are done before the table changes
    if ((rc = naer_net_sync(n)))
        return undo(rc);
    // slot: reuse a retired one
    uint32_t s = (uint32_t)n->nets.size();
    for (uint32_t i = 0; i < n->nets.size(); ++i)
        if (!n->nets[i].active) {
            s = i;
            break;
        }
    if (s == n->nets.size())
        n->nets.push_back(net);
    else
        n->nets[s] = net;
    if ((rc = upload_table(n)) || (rc = init_net(n, n->nets[s]))) {
        n->nets[s].active = false;
        (void)upload_table(n);
        return undo(rc);
    }
    HIP_TRY(hipStreamSynchronize(n->stream));
    *slot = s;
    return NAER_OK;
}

int naer_net_check_finite(naer_net* n, uint32_t* bad_slots, uint32_t cap, uint32_t* n_bad)
{
    if (!n || !n_bad || (cap && !bad_slots))
        return fail(NAER_ERR_INVALID, "NULL argument");
    if (!n->built)
        return fail(NAER_ERR_STATE, "net not built");
    if (n->event)
        return fail(NAER_ERR_INVALID, "event-driven nets are not checked");
    int rc = naer_net_sync(n);
    if (rc)
        return rc;
    // one flag per 256-neuron chunk: arena jobs start on 256-neuron
    // boundaries, so a chunk belongs to one job at most (the networks of a
    // pack lie back to back, so a chunk may flag two neighbours)
    const uint32_t chunks = (n->N + kBlock - 1) / kBlock;
    if (!n->finite_flags && (rc = n->alloc(&n->finite_flags, chunks)))
        return rc;
    // every active neuron lies in the tiles [0, n_act): only those chunks are
    // checked and read back
    const uint32_t act = (uint32_t)std::min<uint64_t>(n->N, (uint64_t)n->n_act << n->tw_shift);
    const uint32_t act_chunks = (act + kBlock - 1) / kBlock;
    const UpdateArgs ua = make_update_args(n);
    HIP_TRY(launch_check_finite(ua, n->finite_flags, act, n->stream));
    std::vector<uint32_t> flags(chunks, 0u);
    HIP_TRY(hipMemcpyAsync(flags.data(), n->finite_flags, act_chunks * sizeof(uint32_t), hipMemcpyDeviceToHost,
                           n->stream));
    HIP_TRY(hipStreamSynchronize(n->stream));
    uint32_t k = 0;
    for (uint32_t s = 0; s < n->nets.size(); ++s) {
        const NetInfo& net = n->nets[s];
        if (!net.active || net.n_neurons == 0)
            continue;
        bool bad = false;
        for (uint32_t c = net.neuron_begin / kBlock; c <= (net.neuron_begin + net.n_neurons - 1) / kBlock && !bad; ++c)
            bad = flags[c] != 0;
        if (bad) {
            if (k < cap)
                bad_slots[k] = s;
            ++k;
        }
    }
    *n_bad = k;
    return NAER_OK;
}

int naer_arena_retire(naer_net* n, uint32_t slot)
{
    if (!n)
        return fail(NAER_ERR_INVALID, "net is NULL");
    if (!n->arena)
        return fail(NAER_ERR_STATE, "not an arena");
    if (slot >= n->nets.size() || !n->nets[slot].active)
        return fail(NAER_ERR_INVALID, "slot %u holds no job", slot);
    int rc;
    if ((rc = naer_net_sync(n)))
        return rc;
    NetInfo& net = n->nets[slot];
    net.active = false;
    n->neuron_alloc.give(net.neuron_begin, (uint64_t)net.n_neurons + 1);
    n->syn_alloc.give(net.syn_begin, net.n_syn);
    net.n_syn = net.n_plastic = 0;
    return upload_table(n);
}

// Deferred delivery (enqueue_steps): bucketed static nets other than arenas
// (an arena admits and retires jobs between any two steps, and plastic
// synapses take their STDP when their spikes are delivered).
static bool deferred_delivery(const naer_net* n)
{
    return n->bucketed && !n->plastic && !n->arena;
}

static hipError_t do_update(naer_net* n, void* arg)
{
    UpdateArgs* a = (UpdateArgs*)arg;
    // the input probe is compiled into the EXTRA variants only (it costs the
    // plain variant's step loop scalar registers)
    return launch_update(*a, n->det, n->plastic, n->bucketed,
                         n->poisson_ok || !n->mpops.empty() || a->probe != nullptr, n->rec_bytes == 4, n->stream);
}

static hipError_t do_scatter(naer_net* n, void* arg)
{
    ScatterArgs* a = (ScatterArgs*)arg;
    return launch_scatter(*a, n->det, n->plastic, n->n_act, n->scatter_chunks, n->stream);
}

FlushArgs flush_args(const naer_net* n, bool force)
{
    FlushArgs f;
    std::memset(&f, 0, sizeof f);
    f.st = n->st;
    f.row_ptr = n->row_ptr;
    f.col = n->col;
    f.w = n->w;
    f.sd = stdp_args(n);
    f.pops = n->d_pops;
    f.blk_pop = n->d_blk_pop;
    f.n_pops = (uint32_t)n->pp.size();
    f.tw_shift = n->tw_shift;
    f.N = n->N;
    f.B = n->B;
    f.force = force ? 1u : 0u;
    return f;
}

// grid-stride: about 8 rows per wave, at most 4096 blocks
static uint32_t flush_grid(const naer_net* n)
{
    const uint32_t wpb = kBlock / kWave;
    return std::max<uint32_t>(1, std::min<uint64_t>(4096, ((uint64_t)n->N + 8 * wpb - 1) / (8 * wpb)));
}

static hipError_t do_flush(naer_net* n, void* arg)
{
    return launch_flush(*(const FlushArgs*)arg, flush_grid(n), n->stream);
}

// Applies all pending potentiation now (before weights are read or saved)
// and records the flush step on the device.
int flush_now(naer_net* n)
{
    if (!n->plastic || !n->built)
        return NAER_OK;
    int rc = naer_net_sync(n);
    if (rc)
        return rc;
    const FlushArgs fa = flush_args(n, true);
    HIP_TRY(launch_flush(fa, flush_grid(n), n->stream));
    HIP_TRY(hipStreamSynchronize(n->stream));
    DevState st;
    HIP_TRY(hipMemcpy(&st, n->st, sizeof st, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy((char*)n->st + offsetof(DevState, t_flush), &st.clock, sizeof st.clock, hipMemcpyHostToDevice));
    return NAER_OK;
}

struct ModelStep {
    const ModelPop* m;
    const ModelArgs* a;
};

// one block of one tile's width per tile the population overlaps
static hipError_t do_model(naer_net* n, void* arg)
{
    const ModelStep* ms = (const ModelStep*)arg;
    void* params[] = { (void*)ms->a };
    const uint32_t tiles = ((ms->m->end - 1) >> n->tw_shift) - (ms->m->begin >> n->tw_shift) + 1;
    return hipModuleLaunchKernel(ms->m->fn, tiles, 1, 1, 1u << n->tw_shift, 1, 1, 0, n->stream, params, nullptr);
}

static ModelArgs model_args(const naer_net* n, const ModelPop& m)
{
    ModelArgs a;
    std::memset(&a, 0, sizeof a);
    char* st = (char*)n->st;
    a.cin = n->cin;
    a.state = m.state;
    a.V = n->V;
    a.row_ptr = n->row_ptr;
    a.list0 = (uint4*)n->list0;
    a.list1 = (uint4*)n->list1;
    a.seg_count = n->seg_count;
    a.seg_ev = n->seg_ev;
    a.spk_part = n->spk_part;
    a.ev_part = n->ev_part;
    a.rec = (uint4*)n->rec;
    a.rec_n = (unsigned long long*)(st + offsetof(DevState, rec_n));
    a.overflow = (unsigned int*)(st + offsetof(DevState, overflow));
    a.cur = (const unsigned int*)(st + offsetof(DevState, cur));
    a.t0_cur = (const unsigned long long*)(st + offsetof(DevState, t0_cur));
    a.params = m.params;
    a.rec_cap = n->rec_cap;
    a.t_start = m.t_start;
    a.begin = m.begin;
    a.end = m.end;
    a.tile0 = m.begin >> n->tw_shift;
    a.seg_cap = n->seg_cap;
    a.n_seg = n->n_seg;
    a.tw_shift = n->tw_shift;
    a.N = n->N;
    a.B = n->B;
    a.dt_us = n->dt_us;
    return a;
}

struct BucketStep {
    const BucketArgs* a;
    int step;
};

static hipError_t do_bucket_step(naer_net* n, void* arg)
{
    const BucketStep* b = (const BucketStep*)arg;
    return launch_bucket_step(*b->a, b->step, n->det, n->stream);
}

static hipError_t do_clock_advance(naer_net* n, void* arg)
{
    return launch_clock_advance(n->st, *(const uint32_t*)arg, n->stream);
}

// Enqueues steps batch by batch.  Batches are aligned to multiples of B
// steps.  A bucketed static net (no arena) defers the delivery of a launch
// that ends inside its aligned batch: its spikes stay in the batch's lists,
// and the launch that completes the batch delivers all of them through the
// buckets, so any run length costs one bucketed delivery per aligned batch.
// Other nets deliver every launch: full aligned batches of bucketed nets
// through the buckets, partial ones through the ring.
static int enqueue_steps(naer_net* n, uint32_t n_steps, bool timed)
{
    UpdateArgs ua = make_update_args(n);
    ua.defer = deferred_delivery(n) ? 1u : 0u;
    ScatterArgs sa;
    std::memset(&sa, 0, sizeof sa);
    sa.list0 = n->list0;
    sa.list1 = n->list1;
    sa.seg_count = n->seg_count;
    sa.seg_cap = n->seg_cap;
    sa.n_seg = n->n_seg;
    sa.st = n->st;
    sa.row_ptr = n->row_ptr;
    sa.col = n->col;
    sa.w = n->w;
    sa.ring = n->ring;
    sa.sd = stdp_args(n);
    sa.N = n->N;
    sa.D = n->D;
    sa.is_last = 1;
    sa.flags = n->bucketed ? n->flags : nullptr;
    sa.R = n->R;
    sa.Bfull = n->B;
    BucketArgs ba;
    std::memset(&ba, 0, sizeof ba);
    ba.list0 = n->list0;
    ba.list1 = n->list1;
    ba.seg_count = n->seg_count;
    ba.seg_ev = n->seg_ev;
    ba.blk_seg = n->blk_seg;
    ba.seg_cap = n->seg_cap;
    ba.n_seg = n->n_seg;
    ba.n_act = n->n_act;
    ba.st = n->st;
    ba.row_ptr = n->row_ptr;
    ba.col = n->col;
    ba.w = n->w;
    ba.sd = stdp_args(n);
    ba.hist = n->hist;
    ba.hist_c = n->hist_c;
    ba.stage = n->stage;
    ba.totals = n->totals;
    ba.bs = n->bs;
    ba.flags = n->flags;
    ba.ev = n->ev;
    ba.ring = n->ring;
    ba.e_cap = n->e_cap;
    ba.N = n->N;
    ba.D = n->D;
    ba.Bfull = n->B;
    ba.R = n->R;
    ba.M = n->M;
    ba.NT = n->NT;
    ba.NBK = n->NBK;
    ba.tw_shift = n->tw_shift;
    // direct placement (one tile per group) when asked for and its counters
    // fit the occupied prefix; two levels otherwise
    const bool direct = n->RS == 0 && direct_ok(n);
    ba.gt_shift = direct ? 0u : n->gt_shift;
    ba.direct = direct ? 1u : 0u;
    ba.n_cnt = direct ? direct_counters(n, n->n_act) : n->NBK >> n->gt_shift;
    ba.G = n->G;
    ba.NB = n->NB;
    ba.U = n->U;
    ba.C = std::max<uint32_t>(1, (n->max_row + kWave * n->U - 1) / (kWave * n->U));
    ba.gcur = n->gcur;
    // about 2048 refine blocks over the group buckets
    ba.RS = n->RS ? n->RS : default_refine_split(n);
    ba.AG = n->AG & 1u;
    ba.AGP = (n->AG >> 1) & 1u;
    ba.balance = n->balance ? 1u : 0u;
    ba.is_last = 1;
    ba.ncls = n->rec_bytes == 4 ? n->ncls : nullptr;
    ba.cls_w = n->cls_w;
    // reserved placement (set_schedule admits it only where it fits)
    if (!direct && n->RES && res_fits(n, n->G, n->RS)) {
        ba.reserve = 1;
        ba.rc = n->res_c;
        ba.rcc = n->res_cc;
        ba.rcnt = n->res_cnt;
        ba.spill = n->res_spill;
        ba.pb = n->res_pb;
        ba.tb = n->res_tb;
        ua.reserve = 1;
        // the placement reads the rows of the batch's spikes first (no
        // histogram pass has brought them into the caches): the update loads
        // them as the spikes happen
        ua.prefetch = 1;
    }
    ba.parts = n->parts;
    FlushArgs fa = flush_args(n, false);
    std::vector<ModelArgs> ma;
    for (const auto& m : n->mpops)
        ma.push_back(model_args(n, m));
    int rc;
    while (n_steps > 0) {
        const uint32_t off = (uint32_t)(n->steps_enqueued % n->B);
        const uint32_t b = std::min(n_steps, n->B - off);
        ua.B = b;
        sa.B = b;
        ba.B = b;
        fa.B = b;
        // plastic nets: every half window (the device's condition) the rows
        // whose pending potentiation would outgrow the window are flushed
        if (n->plastic && (rc = launch_timed(n, NAER_K_POTENTIATE, do_flush, &fa, timed)))
            return rc;
        if ((rc = launch_timed(n, NAER_K_UPDATE, do_update, &ua, timed)))
            return rc;
        for (size_t i = 0; i < ma.size(); ++i) {
            ma[i].B = b;
            ModelStep ms{ &n->mpops[i], &ma[i] };
            if ((rc = launch_timed(n, NAER_K_MODEL, do_model, &ms, timed)))
                return rc;
        }
        const bool completes = off + b == n->B;
        if (n->bucketed && completes && (ua.defer || off == 0)) {
            // timed per part; the parts add up to NAER_K_SCATTER
            for (int k = 0; k < kBucketSteps; ++k) {
                if (ba.reserve && k < 2)
                    continue;  // reserved placement: no counting passes
                BucketStep bst{ &ba, k };
                if ((rc = launch_timed(n, NAER_K_BUCKET_HIST + k, do_bucket_step, &bst, timed)))
                    return rc;
            }
            n->kernel_calls[NAER_K_SCATTER] += 1;
        } else if (ua.defer) {
            uint32_t steps = b;
            if ((rc = launch_timed(n, NAER_K_SCATTER, do_clock_advance, &steps, false)))
                return rc;
            n->kernel_calls[NAER_K_SCATTER] -= 1;  // not a delivery
        } else if ((rc = launch_timed(n, NAER_K_SCATTER, do_scatter, &sa, timed))) {
            return rc;
        }
        n_steps -= b;
        n->steps_enqueued += b;
    }
    return NAER_OK;
}

int naer_net_step(naer_net* n, uint32_t n_steps)
{
    if (!n)
        return fail(NAER_ERR_INVALID, "net is NULL");
    if (!n->built)
        return fail(NAER_ERR_STATE, "net not built");
    if (n->event)
        return ev_step(n, n_steps);
    if (n->graph_chunk) {
        // a captured chunk is a whole number of aligned batches, so every
        // replay sees the same batch pattern: align plainly first
        const uint32_t misalign = (uint32_t)((n->B - n->steps_enqueued % n->B) % n->B);
        if (misalign && n_steps > misalign) {
            int rc = enqueue_steps(n, misalign, true);
            if (rc)
                return rc;
            n_steps -= misalign;
        }
    }
    if (n->graph_chunk && n->steps_enqueued % n->B == 0) {
        if (!n->graph_exec && n_steps >= n->graph_chunk) {
            // capture one chunk; its kernels read the clock from the device,
            // so the same graph replays for every chunk
            hipGraph_t g = nullptr;
            HIP_TRY(hipStreamBeginCapture(n->stream, hipStreamCaptureModeThreadLocal));
            const uint64_t before = n->steps_enqueued;
            int rc = enqueue_steps(n, n->graph_chunk, false);
            hipError_t e = hipStreamEndCapture(n->stream, &g);
            n->steps_enqueued = before;
            if (rc)
                return rc;
            if (e != hipSuccess)
                return fail(NAER_ERR_HIP, "graph capture failed: %s", hipGetErrorString(e));
            e = hipGraphInstantiate(&n->graph_exec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess)
                return fail(NAER_ERR_HIP, "graph instantiation failed: %s", hipGetErrorString(e));
        }
        while (n->graph_exec && n_steps >= n->graph_chunk) {
            HIP_TRY(hipGraphLaunch(n->graph_exec, n->stream));
            n_steps -= n->graph_chunk;
            n->steps_enqueued += n->graph_chunk;
        }
    }
    return n_steps ? enqueue_steps(n, n_steps, true) : NAER_OK;
}

int naer_net_set_graph(naer_net* n, uint32_t chunk_steps)
{
    if (!n)
        return fail(NAER_ERR_INVALID, "net is NULL");
    if (n->event && chunk_steps)
        return fail(NAER_ERR_INVALID, "the event-driven mode is not captured in graphs");
    int rc = naer_net_sync(n);
    if (rc)
        return rc;
    if ((rc = n->drop_graph()))
        return rc;
    // whole aligned batches per chunk
    n->graph_chunk = chunk_steps ? (chunk_steps + n->B - 1) / n->B * n->B : 0;
    return NAER_OK;
}

int naer_net_get_schedule(naer_net* n, naer_schedule* out)
{
    if (!n || !out)
        return fail(NAER_ERR_INVALID, "NULL argument");
    std::memset(out, 0, sizeof *out);
    out->bucketed = n->bucketed ? 1u : 0u;
    out->task_synapses = kWave * n->U;
    out->segments_per_block = n->G;
    out->refine_split = n->RS;
    out->run_aggregate = n->AG;
    out->reserve = n->RES;
    out->parts_max = n->res_ok ? n->parts : 0u;
    out->record_bytes = n->bucketed ? n->rec_bytes : 0u;
    out->delivery_us = n->tuned_us;
    return NAER_OK;
}

int naer_net_set_schedule(naer_net* n, const naer_schedule* s)
{
    if (!n || !s)
        return fail(NAER_ERR_INVALID, "NULL argument");
    if (s->task_synapses != 4 * kWave && s->task_synapses != 8 * kWave)
        return fail(NAER_ERR_INVALID, "task_synapses must be %u or %u", 4 * kWave, 8 * kWave);
    if (s->segments_per_block == 0 || s->segments_per_block > kMaxSegSchedule)
        return fail(NAER_ERR_INVALID, "segments_per_block must be in [1, %u]", kMaxSegSchedule);
    if (s->refine_split > 64)
        return fail(NAER_ERR_INVALID, "refine_split must be in [0, 64]");
    // an arena's occupied prefix changes with its jobs (direct placement
    // then applies to the batches it fits); a net's whole tile range must fit
    if (s->refine_split == 0 && n->bucketed && !n->arena && direct_counters(n, n->n_seg) > kMaxPlaceCounters)
        return fail(NAER_ERR_INVALID, "refine_split 0 (direct placement) needs at most %u placement counters "
                    "(the net has %u)", kMaxPlaceCounters, direct_counters(n, n->n_seg));
    if (s->run_aggregate > 3)
        return fail(NAER_ERR_INVALID, "run_aggregate must be in [0, 3]");
    if (s->reserve > 1)
        return fail(NAER_ERR_INVALID, "reserve must be 0 or 1");
    if (s->reserve && !n->res_ok)
        return fail(NAER_ERR_INVALID, "reserved placement is not available for this net");
    if (s->reserve && !res_fits(n, s->segments_per_block, s->refine_split))
        return fail(NAER_ERR_INVALID, "reserved placement needs refine_split in [1, %u] and at most %u "
                    "placement blocks per part", n->parts, kResMaxPartRegions - 1);
    int rc = naer_net_sync(n);
    if (rc || (rc = n->drop_graph()))  // captured launches carry the old shapes
        return rc;
    n->U = s->task_synapses / kWave;
    n->G = s->segments_per_block;
    n->NB = (n->n_seg + n->G - 1) / n->G;
    n->RS = s->refine_split;
    n->AG = s->run_aggregate;
    n->RES = s->reserve;
    return n->bucketed ? set_block_ranges(n) : NAER_OK;
}

// Tries schedules of the bucketed delivery on the running simulation: each
// candidate delivers `batches` aligned batches (real steps; in deterministic
// mode the results do not depend on the schedule) and is scored by its
// delivery time per synaptic event, measured with HIP events.  Three stages,
// each starting from the previous stage's winner: the placement (two-level
// or direct, where the net allows both), the shape (task size x segments per
// block x, two-level only, refine blocks per group bucket, each halved or
// doubled), then each counter mode flipped.  Every candidate of a stage is
// tried twice, keeping its better score, so that a drift of the activity
// during the trials does not decide.
int naer_net_autotune(naer_net* n, uint32_t batches, naer_schedule* chosen)
{
    if (!n)
        return fail(NAER_ERR_INVALID, "net is NULL");
    if (!n->built)
        return fail(NAER_ERR_STATE, "net not built");
    if (!n->bucketed || batches == 0)
        return chosen ? naer_net_get_schedule(n, chosen) : NAER_OK;
    int rc = naer_net_sync(n);
    if (rc || (rc = n->drop_graph()))
        return rc;
    const uint32_t graph = n->graph_chunk;
    n->graph_chunk = 0;
    // align to a batch boundary
    const uint32_t mis = (uint32_t)((n->B - n->steps_enqueued % n->B) % n->B);
    if (mis && (rc = naer_net_step(n, mis)))
        return rc;
    const bool prof = n->profiling;
    n->profiling = true;
    // scores (ms per event) and delivery times per batch of a stage's candidates
    auto trial = [&](const std::vector<naer_schedule>& cand, std::vector<double>& score,
                     std::vector<double>& per_batch) -> int {
        score.assign(cand.size(), 1e300);
        per_batch.assign(cand.size(), 0.0);
        for (int pass = 0; pass < 2; ++pass) {
            for (size_t i = 0; i < cand.size(); ++i) {
                int r;
                naer_stats s0, s1;
                double ms0 = 0, ms1 = 0;
                if ((r = naer_net_set_schedule(n, &cand[i])) || (r = naer_net_query(n, &s0)))
                    return r;
                for (int k = NAER_K_BUCKET_HIST; k <= NAER_K_BUCKET_REFINE; ++k)
                    ms0 += n->kernel_ms[k];
                if ((r = naer_net_step(n, batches * n->B)) || (r = naer_net_sync(n)) ||
                    (r = naer_net_query(n, &s1)))
                    return r;
                for (int k = NAER_K_BUCKET_HIST; k <= NAER_K_BUCKET_REFINE; ++k)
                    ms1 += n->kernel_ms[k];
                const double sc = (ms1 - ms0) / std::max(1.0, (double)(s1.events - s0.events));
                if (sc < score[i]) {
                    score[i] = sc;
                    per_batch[i] = (ms1 - ms0) * 1e3 / batches;
                }
            }
        }
        return NAER_OK;
    };
    auto argmin = [](const std::vector<double>& v) {
        size_t b = 0;
        for (size_t i = 1; i < v.size(); ++i)
            if (v[i] < v[b])
                b = i;
        return b;
    };
    std::vector<double> score, per_batch;
    naer_schedule best;
    naer_net_get_schedule(n, &best);
    double best_us = 0.0;
    // stage 1: the placement (two-level counted, direct, reserved, where the
    // net allows them); another one replaces the net's default only when it
    // is clearly faster (trials of a few batches, possibly beside other packs
    // tuning on the same device, are noisy)
    {
        std::vector<naer_schedule> cand{ best };  // the default first
        naer_schedule two = best, one = best, res = best;
        two.reserve = 0;
        if (two.refine_split == 0 || best.reserve)
            two.refine_split = default_refine_split(n);
        one.reserve = 0;
        one.refine_split = 0;
        res.reserve = 1;
        res.refine_split = default_res_parts(n);
        if (best.reserve || best.refine_split == 0)
            cand.push_back(two);
        if (direct_ok(n) && best.refine_split != 0)
            cand.push_back(one);
        if (!best.reserve && res_fits(n, res.segments_per_block, res.refine_split))
            cand.push_back(res);
        if (cand.size() > 1 && (rc = trial(cand, score, per_batch)) == NAER_OK) {
            size_t i = 0;
            for (size_t k = 1; k < cand.size(); ++k)
                if (score[k] < 0.9 * score[0] && score[k] < score[i])
                    i = k;
            best = cand[i];
            best_us = per_batch[i];
        }
    }
    // stage 2: the shape
    if (rc == NAER_OK) {
        std::vector<naer_schedule> cand;
        std::vector<uint32_t> splits;
        const uint32_t top = best.reserve ? n->parts : 64u;
        if (best.refine_split == 0)
            splits = { 0u };
        else
            splits = { std::max<uint32_t>(1, best.refine_split / 2), best.refine_split,
                       std::min<uint32_t>(top, best.refine_split * 2) };
        for (uint32_t task : { 4 * kWave, 8 * kWave })
            for (uint32_t g : { std::max<uint32_t>(1, best.segments_per_block / 2), best.segments_per_block,
                                std::min<uint32_t>(kMaxSegSchedule, best.segments_per_block * 2) })
                for (uint32_t r : splits) {
                    naer_schedule c = best;
                    c.task_synapses = task;
                    c.segments_per_block = g;
                    c.refine_split = r;
                    if (c.reserve && !res_fits(n, g, r))
                        continue;
                    bool dup = false;
                    for (auto& x : cand)
                        dup = dup || (x.task_synapses == task && x.segments_per_block == g && x.refine_split == r);
                    if (!dup)
                        cand.push_back(c);
                }
        if ((rc = trial(cand, score, per_batch)) == NAER_OK) {
            const size_t i = argmin(score);
            best = cand[i];
            best_us = per_batch[i];
        }
    }
    // stage 3: the counter modes, each flipped on the winner; it takes each
    // flip that did better than the winner itself
    if (rc == NAER_OK) {
        std::vector<naer_schedule> cand{ best, best, best };
        cand[1].run_aggregate ^= 1u;
        cand[2].run_aggregate ^= 2u;
        if ((rc = trial(cand, score, per_batch)) == NAER_OK) {
            naer_schedule pick = best;
            double us = per_batch[0];
            for (size_t f = 0; f < 2; ++f)
                if (score[1 + f] < score[0]) {
                    pick.run_aggregate ^= 1u << f;
                    us = std::min(us, per_batch[1 + f]);
                }
            best = pick;
            best_us = us;
        }
    }
    n->profiling = prof;
    n->graph_chunk = graph;
    if (rc)
        return rc;
    best.delivery_us = best_us;
    if ((rc = naer_net_set_schedule(n, &best)))
        return rc;
    n->tuned_us = best_us;
    if (chosen)
        return naer_net_get_schedule(n, chosen);
    return NAER_OK;
}

int naer_net_idle(naer_net* n, int* idle)
{
    if (!n || !idle)
        return fail(NAER_ERR_INVALID, "NULL argument");
    const hipError_t e = hipStreamQuery(n->stream);
    if (e == hipErrorNotReady) {
        *idle = 0;
        return NAER_OK;
    }
    if (e != hipSuccess)
        return fail(NAER_ERR_HIP, "stream query failed: %s", hipGetErrorString(e));
    *idle = 1;
    return NAER_OK;
}

int naer_net_sync(naer_net* n)
{
    if (!n)
        return fail(NAER_ERR_INVALID, "net is NULL");
    HIP_TRY(hipStreamSynchronize(n->stream));
    return n->collect_timers();
}

int naer_net_set_profiling(naer_net* n, int enable)
{
    if (!n)
        return fail(NAER_ERR_INVALID, "net is NULL");
    n->profiling = enable != 0;
    return NAER_OK;
}

int naer_net_query(naer_net* n, naer_stats* out)
{
    if (!n || !out)
        return fail(NAER_ERR_INVALID, "NULL argument");
    int rc = naer_net_sync(n);
    if (rc)
        return rc;
    std::memset(out, 0, sizeof *out);
    for (auto& net : n->nets)
        if (net.active) {
            out->n_neurons += net.n_neurons;
            out->n_synapses += net.n_syn;
            out->n_plastic += net.n_plastic;
        }
    out->ring_depth = n->D;
    out->batch = n->B;
    out->max_row = n->max_row;
    out->device_bytes = n->dev_bytes;
    out->kernels = 8;
    for (int k = 0; k < 8; ++k) {
        out->kernel_ms[k] = n->kernel_ms[k];
        out->kernel_calls[k] = n->kernel_calls[k];
    }
    if (n->built) {
        DevState s;
        HIP_TRY(hipMemcpy(&s, n->st, sizeof s, hipMemcpyDeviceToHost));
        std::vector<unsigned long long> part(n->n_seg);
        HIP_TRY(hipMemcpy(part.data(), n->spk_part, n->n_seg * sizeof(unsigned long long), hipMemcpyDeviceToHost));
        for (auto v : part)
            out->spikes += v;
        HIP_TRY(hipMemcpy(part.data(), n->ev_part, n->n_seg * sizeof(unsigned long long), hipMemcpyDeviceToHost));
        for (auto v : part)
            out->events += v;
        out->step = s.clock;
        out->time_us = s.clock * n->dt_us;
        if (n->event) {
            unsigned long long sp = 0;
            HIP_TRY(hipMemcpy(&sp, n->ev_ctr + 4, sizeof sp, hipMemcpyDeviceToHost));
            out->spikes = sp;
            out->events = n->ev_events;
            out->step = n->ev_epoch;
            out->time_us = n->ev_epoch * kEpochUs;
        }
        out->dropped = n->dropped + (s.rec_n > n->rec_cap ? s.rec_n - n->rec_cap : 0) + s.overflow;
    }
    return NAER_OK;
}

int naer_net_read_spikes(naer_net* n, naer_spike* out, uint64_t cap, uint64_t* n_out)
{
    if (!n || !n_out || (cap && !out))
        return fail(NAER_ERR_INVALID, "NULL argument");
    if (!n->built)
        return fail(NAER_ERR_STATE, "net not built");
    int rc = naer_net_sync(n);
    if (rc)
        return rc;
    DevState s;
    HIP_TRY(hipMemcpy(&s, n->st, sizeof s, hipMemcpyDeviceToHost));
    const uint64_t pend = s.rec_n;
    const uint64_t avail = std::min<uint64_t>(pend, n->rec_cap);
    const uint64_t k = std::min<uint64_t>(avail, cap);
    if (k)
        HIP_TRY(hipMemcpy(out, n->rec, k * sizeof(naer_spike), hipMemcpyDeviceToHost));
    // the buffer is drained: spikes not returned (beyond the record capacity
    // or beyond `cap`) are counted as dropped
    n->dropped += pend - k;
    const unsigned long long zero = 0;
    HIP_TRY(hipMemcpy(&n->st->rec_n, &zero, sizeof zero, hipMemcpyHostToDevice));
    *n_out = k;
    return NAER_OK;
}

int naer_net_raster_open(naer_net* n, const char* path)
{
    if (!n || !path)
        return fail(NAER_ERR_INVALID, "NULL argument");
    if (!n->built)
        return fail(NAER_ERR_STATE, "net not built");
    if (n->raster)
        return fail(NAER_ERR_STATE, "a raster file is already open");
    if (n->rec_cap == 0)
        return fail(NAER_ERR_STATE, "the net records no spikes (record_capacity 0)");
    if (!n->raster_stage)
        HIP_TRY(hipHostMalloc((void**)&n->raster_stage, n->rec_cap * sizeof(naer_spike), hipHostMallocDefault));
    FILE* f = std::fopen(path, "wb");
    if (!f)
        return fail(NAER_ERR_INVALID, "cannot open %s for writing", path);
    unsigned char h[NAER_RASTER_HEADER_BYTES] = {};
    std::memcpy(h, "NAERRST1", 8);
    const uint32_t version = 1, rec_bytes = 8;
    const double dt = n->dt_ms;
    const uint64_t neurons = n->N, records = 0;
    std::memcpy(h + 8, &version, 4);
    std::memcpy(h + 12, &rec_bytes, 4);
    std::memcpy(h + 16, &dt, 8);
    std::memcpy(h + 24, &neurons, 8);
    std::memcpy(h + 32, &records, 8);
    if (std::fwrite(h, 1, sizeof h, f) != sizeof h) {
        std::fclose(f);
        return fail(NAER_ERR_INVALID, "cannot write %s", path);
    }
    n->raster = f;
    n->raster_records = 0;
    return NAER_OK;
}

int naer_net_raster_drain(naer_net* n, uint64_t* n_written)
{
    if (!n)
        return fail(NAER_ERR_INVALID, "net is NULL");
    if (!n->raster)
        return fail(NAER_ERR_STATE, "no raster file is open");
    uint64_t k = 0;
    int rc = naer_net_read_spikes(n, n->raster_stage, n->rec_cap, &k);
    if (rc)
        return rc;
    std::vector<uint64_t> keys(k);
    for (uint64_t i = 0; i < k; ++i)
        keys[i] = (uint64_t)n->raster_stage[i].step << 32 | n->raster_stage[i].neuron;
    std::sort(keys.begin(), keys.end());
    std::vector<uint32_t> rec(2 * k);
    for (uint64_t i = 0; i < k; ++i) {
        rec[2 * i] = (uint32_t)(keys[i] >> 32);
        rec[2 * i + 1] = (uint32_t)keys[i];
    }
    if (k && std::fwrite(rec.data(), 8, k, n->raster) != k)
        return fail(NAER_ERR_INVALID, "raster write failed");
    n->raster_records += k;
    if (n_written)
        *n_written = k;
    return NAER_OK;
}

int naer_net_raster_close(naer_net* n, uint64_t* n_total)
{
    if (!n)
        return fail(NAER_ERR_INVALID, "net is NULL");
    if (!n->raster)
        return fail(NAER_ERR_STATE, "no raster file is open");
    int rc = naer_net_raster_drain(n, nullptr);
    FILE* f = n->raster;
    n->raster = nullptr;
    bool ok = rc == NAER_OK && std::fseek(f, 32, SEEK_SET) == 0 &&
              std::fwrite(&n->raster_records, 8, 1, f) == 1;
    ok = (std::fclose(f) == 0) && ok;
    if (rc)
        return rc;
    if (!ok)
        return fail(NAER_ERR_INVALID, "raster file could not be finished");
    if (n_total)
        *n_total = n->raster_records;
    return NAER_OK;
}

// --- checkpoint / resume ---------------------------------------------------
extern "C++" {
namespace {

constexpr char kCheckpointMagic[8] = { 'N', 'A', 'E', 'R', 'C', 'K', 'P', '5' };

// FNV-1a over the description of every network (populations, connections,
// STDP parameters, step, mode, seeds): a checkpoint loads only into a net
// built from the same description.
uint64_t description_hash(const naer_net* n)
{
    uint64_t h = 0xcbf29ce484222325ull;
    auto mix = [&](const void* p, size_t bytes) {
        const unsigned char* c = (const unsigned char*)p;
        for (size_t i = 0; i < bytes; ++i)
            h = (h ^ c[i]) * 0x100000001b3ull;
    };
    for (const auto& net : n->nets) {
        const uint32_t np = (uint32_t)net.pops.size(), nc = (uint32_t)net.conns.size();
        mix(&np, 4);
        for (naer_pop_desc q : net.pops) {
            q.model = nullptr;  // an address, not part of the description
            mix(&q, sizeof q);
        }
        mix(&nc, 4);
        for (const auto& c : net.conns)
            mix(&c, sizeof c);
        mix(&net.desc.stdp, sizeof net.desc.stdp);
        mix(&net.desc.dt_ms, sizeof net.desc.dt_ms);
        mix(&net.desc.mode, sizeof net.desc.mode);
        mix(&net.desc.seed, sizeof net.desc.seed);
        mix(&net.desc.poisson_seed, sizeof net.desc.poisson_seed);
        mix(&net.neuron_begin, sizeof net.neuron_begin);
    }
    return h;
}

// Structure of a net that a checkpoint must match to be loaded.
struct CheckpointShape {
    uint64_t N, S_cap, D, B, R, NBK, e_cap, n_seg, rec_cap, seg_cap, desc_hash;
    uint32_t det, plastic, poisson, bucketed, rec_bytes, defer;
    uint32_t parts, _pad;  // part slots of the reserved placement's tile ranges
};

CheckpointShape shape_of(const naer_net* n)
{
    CheckpointShape c;
    std::memset(&c, 0, sizeof c);
    c.N = n->N;
    c.S_cap = n->S_cap;
    c.D = n->D;
    c.B = n->B;
    c.R = n->R;
    c.NBK = n->NBK;
    c.e_cap = n->e_cap;
    c.n_seg = n->n_seg;
    c.rec_cap = n->rec_cap;
    c.seg_cap = n->seg_cap;
    c.desc_hash = description_hash(n);
    c.det = n->det;
    c.plastic = n->plastic;
    c.poisson = n->poisson_ok;
    c.bucketed = n->bucketed;
    c.rec_bytes = n->rec_bytes;
    c.defer = deferred_delivery(n) ? 1u : 0u;
    c.parts = n->res_tb ? n->parts : 0u;
    return c;
}

// Device <-> file copies through a bounded host buffer.
struct CheckpointIO {
    FILE* f;
    bool writing;
    bool dry = false;  // loading: read and validate the file, copy nothing to the device
    std::vector<char> buf = std::vector<char>(64u << 20);
    int copy(void* dev, uint64_t bytes)
    {
        if (dry) {
            const long at = std::ftell(f);
            if (std::fseek(f, 0, SEEK_END) != 0)
                return fail(NAER_ERR_INVALID, "checkpoint unreadable");
            const long size = std::ftell(f);
            if (at < 0 || size < 0 || (uint64_t)(size - at) < bytes ||
                std::fseek(f, at + (long)bytes, SEEK_SET) != 0)
                return fail(NAER_ERR_INVALID, "checkpoint truncated");
            return NAER_OK;
        }
        for (uint64_t off = 0; off < bytes;) {
            const size_t k = (size_t)std::min<uint64_t>(buf.size(), bytes - off);
            if (writing) {
                HIP_TRY(hipMemcpy(buf.data(), (char*)dev + off, k, hipMemcpyDeviceToHost));
                if (std::fwrite(buf.data(), 1, k, f) != k)
                    return fail(NAER_ERR_INVALID, "checkpoint write failed");
            } else {
                if (std::fread(buf.data(), 1, k, f) != k)
                    return fail(NAER_ERR_INVALID, "checkpoint truncated");
                HIP_TRY(hipMemcpy((char*)dev + off, buf.data(), k, hipMemcpyHostToDevice));
            }
            off += k;
        }
        return NAER_OK;
    }
    int scalar(void* host, size_t bytes)
    {
        const size_t k = writing ? std::fwrite(host, 1, bytes, f) : std::fread(host, 1, bytes, f);
        return k == bytes ? NAER_OK : fail(NAER_ERR_INVALID, "checkpoint %s failed", writing ? "write" : "read");
    }
    // A section the kernels index with: saved from the device; on load read
    // into host memory, checked by `ok` (which sets the error), and only
    // then copied to the device.
    template <typename T, typename Ok>
    int checked(T* dev, size_t count, std::vector<T>* host, Ok ok)
    {
        host->resize(count);
        if (writing) {
            HIP_TRY(hipMemcpy(host->data(), dev, count * sizeof(T), hipMemcpyDeviceToHost));
            return scalar(host->data(), count * sizeof(T));
        }
        int rc = scalar(host->data(), count * sizeof(T));
        if (rc || (rc = ok(*host)) || dry)
            return rc;
        HIP_TRY(hipMemcpy(dev, host->data(), count * sizeof(T), hipMemcpyHostToDevice));
        return NAER_OK;
    }
};

int corrupt(const char* what) { return fail(NAER_ERR_INVALID, "corrupt checkpoint: %s", what); }

// The dynamic state in a fixed order; used for both directions.  Every value
// the kernels use as an index or a count is validated before it reaches the
// device.
int checkpoint_body(naer_net* n, CheckpointIO& io)
{
    const uint64_t N = n->N;
    int rc;
    uint64_t steps = n->steps_enqueued, dropped = n->dropped;
    if ((rc = io.scalar(&steps, 8)) || (rc = io.scalar(&dropped, 8)))
        return rc;
    // device clock and bookkeeping
    DevState mine;
    HIP_TRY(hipMemcpy(&mine, n->st, sizeof mine, hipMemcpyDeviceToHost));
    std::vector<DevState> st;
    rc = io.checked(n->st, 1, &st, [&](const std::vector<DevState>& v) {
        const DevState& s = v[0];
        const uint64_t B = std::max<uint32_t>(n->B, 1u);
        if (s.clock != steps)
            return corrupt("device clock differs from the step count");
        if (s.clock > 0 && (s.t0_cur >= s.clock || s.cur != (uint32_t)((s.t0_cur / B) & 1u)))
            return corrupt("batch in flight");
        if (s.window != mine.window || s.t_flush > s.clock)
            return corrupt("flush window");
        return NAER_OK;
    });
    if (rc)
        return rc;
    if (!io.dry) {
        n->steps_enqueued = steps;
        n->dropped = dropped;
    }
    if ((rc = io.copy(n->V, N * 4)) || (rc = io.copy(n->ref, N * 4)) || (rc = io.copy(n->last, N * 4)))
        return rc;
    if (n->plastic && ((rc = io.copy(n->lat, N * sizeof(uint4))) || (rc = io.copy(n->pst, 2 * N * sizeof(uint4))) ||
                       (rc = io.copy(n->shist, N * kSpikeHist * sizeof(uint32_t))) ||
                       (rc = io.copy(n->flushed, N * sizeof(uint32_t))) ||
                       (rc = io.copy(n->w, n->S_cap * 4))))
        return rc;
    if (n->pnext && ((rc = io.copy(n->pnext, N * 4)) || (rc = io.copy(n->pcnt, N * 4))))
        return rc;
    if ((rc = io.copy(n->ring, (uint64_t)n->D * N * (n->det ? 8 : 4))) ||
        (rc = io.copy(n->spk_part, (uint64_t)n->n_seg * 8)) || (rc = io.copy(n->ev_part, (uint64_t)n->n_seg * 8)))
        return rc;
    // undrained recorded spikes
    if ((rc = io.copy(n->rec, std::min<uint64_t>(st[0].rec_n, n->rec_cap) * sizeof(naer_spike))))
        return rc;
    if (n->bucketed) {
        // flags and starts of every buffer (bit 0: bucketed, bit 1: some
        // events in the ring, bit 2: the reserved placement's tile ranges),
        // the tile ranges, then the events of the batches in flight
        std::vector<uint32_t> flags, bs, tb;
        rc = io.checked(n->flags, n->R, &flags, [&](const std::vector<uint32_t>& v) {
            for (uint32_t f : v)
                if (f > 7 || ((f & 6u) && !(f & 1u)) || ((f & 4u) && !n->res_tb))
                    return corrupt("bucket flags");
            return NAER_OK;
        });
        if (rc)
            return rc;
        const size_t row = (size_t)n->NBK + 1;
        std::vector<uint64_t> used(n->R, 0);
        rc = io.checked(n->bs, n->R * row, &bs, [&](const std::vector<uint32_t>& v) {
            for (uint32_t r = 0; r < n->R; ++r) {
                if (flags[r] != 1u && flags[r] != 3u)
                    continue;
                const uint32_t* b = v.data() + r * row;
                if (b[0] != 0 || b[n->NBK] > n->e_cap)
                    return corrupt("bucket starts");
                for (uint32_t k = 0; k < n->NBK; ++k)
                    if (b[k] > b[k + 1])
                        return corrupt("bucket starts");
            }
            return NAER_OK;
        });
        if (rc)
            return rc;
        for (uint32_t r = 0; r < n->R; ++r)
            if (flags[r] == 1u || flags[r] == 3u)
                used[r] = bs[r * row + n->NBK];
        if (n->res_tb) {
            const size_t gt1 = ((size_t)1 << n->gt_shift) + 1;
            const size_t per = (size_t)(n->NBK >> n->gt_shift) * n->parts;  // ranges per buffer
            rc = io.checked(n->res_tb, n->R * per * gt1, &tb, [&](const std::vector<uint32_t>& v) {
                for (uint32_t r = 0; r < n->R; ++r) {
                    if (!(flags[r] & 4u))
                        continue;
                    for (size_t k = 0; k < per; ++k) {
                        const uint32_t* t = v.data() + (r * per + k) * gt1;
                        for (size_t i = 0; i + 1 < gt1; ++i)
                            if (t[i] > t[i + 1])
                                return corrupt("tile ranges");
                        if (t[gt1 - 1] > n->e_cap)
                            return corrupt("tile ranges");
                    }
                }
                return NAER_OK;
            });
            if (rc)
                return rc;
            for (uint32_t r = 0; r < n->R; ++r)
                if (flags[r] & 4u)
                    for (size_t k = 0; k < per; ++k)
                        used[r] = std::max<uint64_t>(used[r], tb[(r * per + k) * gt1 + gt1 - 1]);
        }
        for (uint32_t r = 0; r < n->R; ++r)
            if ((rc = io.copy((char*)n->ev + (size_t)r * n->e_cap * n->rec_bytes, used[r] * n->rec_bytes)))
                return rc;
    }
    if (deferred_delivery(n)) {
        // the spikes of an aligned batch that is not complete wait in its lists
        std::vector<uint32_t> cnt;
        std::vector<unsigned long long> sev;
        rc = io.checked(n->seg_count, 2 * (size_t)n->n_seg, &cnt, [&](const std::vector<uint32_t>& v) {
            for (uint32_t c : v)
                if (c > n->seg_cap)
                    return corrupt("spike list counts");
            return NAER_OK;
        });
        if (rc || (rc = io.checked(n->seg_ev, 2 * (size_t)n->n_seg, &sev,
                                   [](const std::vector<unsigned long long>&) { return NAER_OK; })))
            return rc;
        if (steps % n->B != 0) {
            const uint32_t pl = (uint32_t)((steps / n->B) & 1u);
            naer_spike* list = pl ? n->list1 : n->list0;
            for (uint32_t g = 0; g < n->n_seg; ++g)
                if ((rc = io.copy(list + (size_t)g * n->seg_cap, (uint64_t)cnt[pl * n->n_seg + g] * sizeof(naer_spike))))
                    return rc;
        }
    }
    return NAER_OK;
}

}  // namespace
}  // extern "C++"

int naer_net_save(naer_net* n, const char* path)
{
    if (!n || !path)
        return fail(NAER_ERR_INVALID, "NULL argument");
    if (!n->built)
        return fail(NAER_ERR_STATE, "net not built");
    if (n->arena)
        return fail(NAER_ERR_STATE, "arenas are not checkpointed");
    if (!n->mpops.empty() || n->event)
        return fail(NAER_ERR_INVALID,
                    "nets with populations of source-defined models and event-driven nets are not checkpointed");
    int rc = naer_net_sync(n);
    if (rc)
        return rc;
    FILE* f = std::fopen(path, "wb");
    if (!f)
        return fail(NAER_ERR_INVALID, "cannot open %s for writing", path);
    CheckpointIO io{ f, true };
    CheckpointShape c = shape_of(n);
    rc = io.scalar((void*)kCheckpointMagic, 8);
    if (!rc)
        rc = io.scalar(&c, sizeof c);
    if (!rc)
        rc = checkpoint_body(n, io);
    if (std::fclose(f) != 0 && !rc)
        rc = fail(NAER_ERR_INVALID, "cannot finish %s", path);
    return rc;
}

int naer_net_load(naer_net* n, const char* path)
{
    if (!n || !path)
        return fail(NAER_ERR_INVALID, "NULL argument");
    if (!n->built)
        return fail(NAER_ERR_STATE, "net not built");
    if (n->arena)
        return fail(NAER_ERR_STATE, "arenas are not checkpointed");
    if (!n->mpops.empty() || n->event)
        return fail(NAER_ERR_INVALID,
                    "nets with populations of source-defined models and event-driven nets are not checkpointed");
    int rc = naer_net_sync(n);
    if (rc)
        return rc;
    FILE* f = std::fopen(path, "rb");
    if (!f)
        return fail(NAER_ERR_INVALID, "cannot open %s", path);
    CheckpointIO io{ f, false };
    char magic[8];
    CheckpointShape c, mine = shape_of(n);
    rc = io.scalar(magic, 8);
    if (!rc && std::memcmp(magic, kCheckpointMagic, 8) != 0)
        rc = fail(NAER_ERR_INVALID, "%s is not a checkpoint", path);
    if (!rc)
        rc = io.scalar(&c, sizeof c);
    if (!rc && c.desc_hash != mine.desc_hash)
        rc = fail(NAER_ERR_INVALID, "the checkpoint was written by a net of another description");
    if (!rc && std::memcmp(&c, &mine, sizeof c) != 0)
        rc = fail(NAER_ERR_INVALID, "the checkpoint was written by a net of another shape");
    // a dry pass validates the whole file before anything reaches the device
    const long body = rc ? -1 : std::ftell(f);
    if (!rc) {
        io.dry = true;
        rc = checkpoint_body(n, io);
        io.dry = false;
    }
    if (!rc && (body < 0 || std::fseek(f, body, SEEK_SET) != 0))
        rc = fail(NAER_ERR_INVALID, "cannot reread %s", path);
    if (!rc && (rc = n->drop_graph()) == NAER_OK)
        rc = checkpoint_body(n, io);
    std::fclose(f);
    return rc;
}

int naer_net_read_state(naer_net* n, float* v, int32_t* refr, float* pre, float* post)
{
    if (!n)
        return fail(NAER_ERR_INVALID, "net is NULL");
    if (!n->built)
        return fail(NAER_ERR_STATE, "net not built");
    int rc = naer_net_sync(n);
    if (rc)
        return rc;
    const size_t N = n->N;
    if (v)
        HIP_TRY(hipMemcpy(v, n->V, N * sizeof(float), hipMemcpyDeviceToHost));
    if (refr)
        HIP_TRY(hipMemcpy(refr, n->ref, N * sizeof(int32_t), hipMemcpyDeviceToHost));
    if (pre)
        std::memset(pre, 0, N * sizeof(float));
    if (post)
        std::memset(post, 0, N * sizeof(float));
    if (post && n->plastic && n->event) {
        HIP_TRY(hipMemcpy(post, n->ev_trace, N * sizeof(float), hipMemcpyDeviceToHost));
    } else if ((pre || post) && n->plastic) {
        // the traces the next step sees: (a_s + 1) dec^(clock - s) from the last spike s
        std::vector<uint4> lat(N), pst(2 * N);
        DevState st;
        HIP_TRY(hipMemcpy(lat.data(), n->lat, N * sizeof(uint4), hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(pst.data(), n->pst, 2 * N * sizeof(uint4), hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(&st, n->st, sizeof st, hipMemcpyDeviceToHost));
        auto dec = [](const std::vector<float>& t, uint64_t k) { return k < t.size() ? t[k] : 0.0f; };
        for (size_t j = 0; j < N; ++j) {
            if (!lat[j].x)
                continue;
            const uint64_t k = st.clock - lat[j].y;
            float a;
            if (pre) {
                std::memcpy(&a, &lat[j].z, 4);
                const float b = a + 1.0f;
                pre[j] = b * dec(n->h_tp, k);
            }
            if (post) {
                std::memcpy(&a, &pst[2 * j].y, 4);
                const float b = a + 1.0f;
                post[j] = b * dec(n->h_tm, k);
            }
        }
    }
    return NAER_OK;
}

int naer_net_read_csr(naer_net* n, uint64_t* row_ptr, uint32_t* col, float* w)
{
    if (!n)
        return fail(NAER_ERR_INVALID, "net is NULL");
    if (!n->built)
        return fail(NAER_ERR_STATE, "net not built");
    int rc = naer_net_sync(n);
    if (rc)
        return rc;
    if (row_ptr)
        HIP_TRY(hipMemcpy(row_ptr, n->row_ptr, ((size_t)n->N + 1) * sizeof(uint64_t), hipMemcpyDeviceToHost));
    if (col && n->S_cap)
        HIP_TRY(hipMemcpy(col, n->col, n->S_cap * sizeof(uint32_t), hipMemcpyDeviceToHost));
    // plastic weights take their pending potentiation first
    if (w && n->S_cap && (rc = flush_now(n)))
        return rc;
    if (w && n->S_cap)
        HIP_TRY(hipMemcpy(w, n->w, n->S_cap * sizeof(float), hipMemcpyDeviceToHost));
    return NAER_OK;
}

int naer_net_set_probe(naer_net* n, uint64_t t_begin, uint32_t n_steps)
{
    if (!n)
        return fail(NAER_ERR_INVALID, "net is NULL");
    if (!n->built)
        return fail(NAER_ERR_STATE, "net not built");
    int rc = naer_net_sync(n);
    if (rc)
        return rc;
    // captured graphs hold the probe pointer in their arguments
    if ((rc = n->drop_graph()))
        return rc;
    if (n->probe) {
        n->release(n->probe, (size_t)n->probe_n * n->N * sizeof(float));
        n->probe = nullptr;
    }
    n->probe_t0 = t_begin;
    n->probe_n = n_steps;
    if (n_steps) {
        if ((rc = n->alloc(&n->probe, (size_t)n_steps * n->N)))
            return rc;
        HIP_TRY(hipMemset(n->probe, 0, (size_t)n_steps * n->N * sizeof(float)));
    }
    return NAER_OK;
}

int naer_net_read_probe(naer_net* n, float* out)
{
    if (!n || !out)
        return fail(NAER_ERR_INVALID, "NULL argument");
    if (!n->probe)
        return fail(NAER_ERR_STATE, "no probe set");
    int rc = naer_net_sync(n);
    if (rc)
        return rc;
    HIP_TRY(hipMemcpy(out, n->probe, (size_t)n->probe_n * n->N * sizeof(float), hipMemcpyDeviceToHost));
    return NAER_OK;
}

int naer_net_destroy(naer_net* n)
{
    delete n;
    return NAER_OK;
}

int naer_model_create(const char* source, uint32_t n_state, naer_model** out)
{
    if (!source || !out)
        return fail(NAER_ERR_INVALID, "NULL argument");
    if (n_state < 1 || n_state > NAER_MODEL_STATE)
        return fail(NAER_ERR_INVALID, "n_state must be in [1, %d]", NAER_MODEL_STATE);
    naer_model* m = new naer_model;
    m->n_state = n_state;
    std::string log;
    if (compile_model(source, n_state, &m->code, &log)) {
        delete m;
        return fail(NAER_ERR_INVALID, "model source does not compile:\n%s", log.c_str());
    }
    *out = m;
    return NAER_OK;
}

int naer_model_destroy(naer_model* m)
{
    delete m;
    return NAER_OK;
}

int naer_model_source(const char* name, const char** out)
{
    if (!name || !out)
        return fail(NAER_ERR_INVALID, "NULL argument");
    *out = model_source(name);
    if (!*out)
        return fail(NAER_ERR_INVALID, "no model named '%s' (the library ships lif and cuba_lif)", name);
    return NAER_OK;
}

int naer_philox(naer_ctx* ctx, const uint32_t* ctr, const uint32_t* key, uint32_t* out, uint64_t n)
{
    if (!ctx || (n && (!ctr || !key || !out)))
        return fail(NAER_ERR_INVALID, "NULL argument");
    if (n == 0)
        return NAER_OK;
    HIP_TRY(hipSetDevice(ctx->device));
    uint32_t *dc = nullptr, *dk = nullptr, *dout = nullptr;
    HIP_TRY(hipMalloc(&dc, n * 16));
    hipError_t e = hipMalloc(&dk, n * 8);
    if (e == hipSuccess)
        e = hipMalloc(&dout, n * 16);
    if (e == hipSuccess)
        e = hipMemcpy(dc, ctr, n * 16, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(dk, key, n * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = launch_philox(dc, dk, dout, n, nullptr);
    if (e == hipSuccess)
        e = hipMemcpy(out, dout, n * 16, hipMemcpyDeviceToHost);
    (void)hipFree(dc);
    (void)hipFree(dk);
    (void)hipFree(dout);
    if (e != hipSuccess)
        return fail(NAER_ERR_HIP, "naer_philox: %s", hipGetErrorString(e));
    return NAER_OK;
}

int naer_ref_poisson_epoch_noalloc(naer_ctx* ctx, uint32_t n, uint32_t n_pois_total, uint32_t seed,
                                   float rate_hz, uint32_t t0, uint32_t t1, uint32_t* times, uint32_t* count)
{
    if (!ctx || (n && (!times || !count)))
        return fail(NAER_ERR_INVALID, "NULL argument");
    if (n == 0)
        return NAER_OK;
    const uint64_t words = (uint64_t)n * (kMaxPoissonPerEpoch + 1);
    if (words > kScratchWords)
        return fail(NAER_ERR_CAPACITY,
                    "naer_ref_poisson_epoch_noalloc: %u neurons need %llu words of its static scratch, which holds "
                    "%llu (at most %llu neurons)",
                    n, (unsigned long long)words, (unsigned long long)kScratchWords,
                    (unsigned long long)(kScratchWords / (kMaxPoissonPerEpoch + 1)));
    HIP_TRY(hipSetDevice(ctx->device));
    uint32_t* base = nullptr;
    int rc = scratch_of(ctx->device, &base);
    if (rc)
        return rc;
    uint32_t* dt = base;
    uint32_t* dc = base + (size_t)n * kMaxPoissonPerEpoch;
    // the scratch keeps the previous call's contents: clear the times
    hipError_t e = launch_fill_u32(dt, (unsigned long long)n * kMaxPoissonPerEpoch, 0u, nullptr);
    if (e == hipSuccess)
        e = launch_ref_poisson_epoch(n, n_pois_total, seed, -((double)rate_hz / 1000000.0), t0, t1, dt, dc, nullptr);
    if (e == hipSuccess)
        e = hipMemcpy(times, dt, (size_t)n * kMaxPoissonPerEpoch * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess)
        e = hipMemcpy(count, dc, (size_t)n * 4, hipMemcpyDeviceToHost);
    if (e != hipSuccess)
        return fail(NAER_ERR_HIP, "naer_ref_poisson_epoch_noalloc: %s", hipGetErrorString(e));
    return NAER_OK;
}

int naer_ref_poisson_epoch(naer_ctx* ctx, uint32_t n, uint32_t n_pois_total, uint32_t seed, float rate_hz,
                           uint32_t t0, uint32_t t1, uint32_t* times, uint32_t* count)
{
    if (!ctx || (n && (!times || !count)))
        return fail(NAER_ERR_INVALID, "NULL argument");
    if (n == 0)
        return NAER_OK;
    HIP_TRY(hipSetDevice(ctx->device));
    uint32_t *dt = nullptr, *dc = nullptr;
    HIP_TRY(hipMalloc(&dt, (size_t)n * kMaxPoissonPerEpoch * 4));
    hipError_t e = hipMalloc(&dc, (size_t)n * 4);
    if (e == hipSuccess)
        e = hipMemset(dt, 0, (size_t)n * kMaxPoissonPerEpoch * 4);
    if (e == hipSuccess)
        e = launch_ref_poisson_epoch(n, n_pois_total, seed, -((double)rate_hz / 1000000.0), t0, t1, dt, dc, nullptr);
    if (e == hipSuccess)
        e = hipMemcpy(times, dt, (size_t)n * kMaxPoissonPerEpoch * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess)
        e = hipMemcpy(count, dc, (size_t)n * 4, hipMemcpyDeviceToHost);
    (void)hipFree(dt);
    (void)hipFree(dc);
    if (e != hipSuccess)
        return fail(NAER_ERR_HIP, "naer_ref_poisson_epoch: %s", hipGetErrorString(e));
    return NAER_OK;
}

}  // extern "C"
