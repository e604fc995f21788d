// naer_runtime.hip — host side of the C ABI (include/naer/naer.h): context
// and network objects, the device connectivity build (Network::compile,
// src/Network.cpp:267-569 of the reference) and the step loop
// (Network::step, src/Network.cpp:122-256).
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "naer_device.hpp"
#include "naer_kernels.hpp"

using namespace naer;

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

#define HIP_TRY(expr)                                                                        \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            return fail(e_ == hipErrorOutOfMemory ? NAER_ERR_NOMEM : NAER_ERR_HIP, "%s: %s (%s:%d)", \
                        #expr, hipGetErrorString(e_), __FILE__, __LINE__);                   \
    } while (0)

}  // namespace

struct naer_ctx {
    int device;
};

struct PendingTimer {
    int kind;
    hipEvent_t a, b;
};

// One network of a pack.
struct NetInfo {
    uint32_t neuron_begin = 0, n_neurons = 0;
    uint32_t pop_begin = 0, n_pops = 0;
    uint32_t conn_begin = 0, n_conns = 0;
    uint64_t t_start = 0;  // pack step at which its local time is 0
};

struct naer_net {
    naer_ctx* ctx = nullptr;
    naer_net_desc desc{};               // pack-wide settings (dt, mode, STDP) of the first net
    std::vector<naer_net_desc> net_descs;
    std::vector<NetInfo> nets;
    std::vector<naer_pop_desc> pops;    // all populations, pack order
    std::vector<naer_conn_desc> conns;  // all connections, population indices into `pops`
    std::vector<uint32_t> conn_local;   // index of each connection within its net
    std::vector<uint32_t> pop_net;      // net of each population
    std::vector<uint32_t> blk_pop_h;    // first population of each update block
    uint32_t N = 0, n_pois = 0, dt_us = 0, D = 0, B = 0, dmin = 0, dmax = 0;
    bool plastic = false, det = false, built = false;
    std::vector<PopParams> pp;
    std::vector<float> cdf;
    hipStream_t stream = nullptr;

    std::vector<void*> allocs;
    uint64_t dev_bytes = 0;

    float *V = nullptr, *xh = nullptr, *yh = nullptr;
    int32_t *ref = nullptr, *last = nullptr;
    uint32_t *pnext = nullptr, *pcnt = nullptr;
    void* ring = nullptr;
    unsigned long long* row_ptr = nullptr;
    uint32_t* col = nullptr;
    float* w = nullptr;
    uint64_t S = 0;
    uint32_t max_row = 0;
    unsigned long long* rev_ptr = nullptr;
    uint32_t *rev_pos = nullptr, *rev_src = nullptr;
    uint64_t n_plastic = 0;
    naer_spike *list0 = nullptr, *list1 = nullptr, *rec = nullptr;
    uint32_t seg_cap = 0, n_seg = 0;
    uint32_t* seg_count = nullptr;
    unsigned long long *spk_part = nullptr, *ev_part = nullptr;
    uint64_t rec_cap = 0, dropped = 0;
    DevState* st = nullptr;
    float* d_cdf = nullptr;
    PopParams* d_pops = nullptr;
    uint32_t* d_blk_pop = nullptr;
    float* probe = nullptr;
    uint64_t probe_t0 = 0;
    uint32_t probe_n = 0;

    uint32_t scatter_chunks = 1, scatter_grid = 1;
    uint64_t steps_enqueued = 0;

    bool profiling = false;
    uint32_t graph_chunk = 0;         // steps per captured graph (0 = plain launches)
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
            event_pool.push_back(p.a);
            event_pool.push_back(p.b);
        }
        pending.clear();
        return NAER_OK;
    }
};

namespace {

int validate_and_derive(naer_net* n)
{
    if (n->net_descs.empty())
        return fail(NAER_ERR_INVALID, "a pack needs at least one network");
    n->desc = n->net_descs[0];
    const naer_net_desc& d0 = n->desc;
    const double dtus = (double)d0.dt_ms * 1000.0;
    n->dt_us = (uint32_t)std::llround(dtus);
    if (!(d0.dt_ms > 0.0f) || std::fabs(dtus - (double)n->dt_us) > 1e-3 || n->dt_us == 0)
        return fail(NAER_ERR_INVALID, "dt must be a positive whole number of microseconds");
    if (d0.mode != NAER_MODE_FAST && d0.mode != NAER_MODE_DETERMINISTIC)
        return fail(NAER_ERR_INVALID, "unknown mode %u", d0.mode);
    n->det = d0.mode == NAER_MODE_DETERMINISTIC;

    uint64_t N = 0;
    n->dmin = 0xFFFFFFFFu;
    n->dmax = 1;
    bool stdp_seen = false;
    naer_stdp_desc stdp{};
    uint64_t rec = 0;
    for (uint32_t ni = 0; ni < n->net_descs.size(); ++ni) {
        const naer_net_desc& d = n->net_descs[ni];
        if (d.n_pops == 0 || !d.pops)
            return fail(NAER_ERR_INVALID, "network %u has no populations", ni);
        if (d.n_conns && !d.conns)
            return fail(NAER_ERR_INVALID, "network %u: conns is NULL", ni);
        if (d.dt_ms != d0.dt_ms || d.mode != d0.mode)
            return fail(NAER_ERR_INVALID, "network %u: all networks of a pack share dt and mode", ni);
        NetInfo info;
        info.neuron_begin = (uint32_t)N;
        info.pop_begin = (uint32_t)n->pops.size();
        info.n_pops = d.n_pops;
        info.conn_begin = (uint32_t)n->conns.size();
        info.n_conns = d.n_conns;
        uint64_t nn = 0;
        for (uint32_t p = 0; p < d.n_pops; ++p) {
            n->pops.push_back(d.pops[p]);
            n->pop_net.push_back(ni);
            nn += d.pops[p].size;
        }
        if (nn == 0 || nn > (1ull << 24))
            return fail(NAER_ERR_CAPACITY, "network %u: neuron count must be in [1, 2^24] (got %llu)", ni,
                        (unsigned long long)nn);
        info.n_neurons = (uint32_t)nn;
        N += nn;
        for (uint32_t c = 0; c < d.n_conns; ++c) {
            naer_conn_desc k = d.conns[c];
            if (k.src_pop >= d.n_pops || k.dst_pop >= d.n_pops)
                return fail(NAER_ERR_INVALID, "network %u connection %u: population index out of range", ni, c);
            if (d.pops[k.dst_pop].kind != NAER_POP_LIF)
                return fail(NAER_ERR_INVALID, "network %u connection %u: target population must be LIF", ni, c);
            const uint32_t q = k.delay_quantum ? k.delay_quantum : 1;
            if (k.delay_min < 1 || k.delay_max < k.delay_min || k.delay_max > 127 ||
                (k.delay_max - k.delay_min) % q)
                return fail(NAER_ERR_INVALID,
                            "network %u connection %u: delays must satisfy 1 <= min <= max <= 127 on the quantum grid",
                            ni, c);
            if (k.plastic && d.pops[k.src_pop].kind != NAER_POP_LIF)
                return fail(NAER_ERR_INVALID, "network %u connection %u: plastic connections need a LIF source",
                            ni, c);
            if (k.rule != NAER_RULE_BERNOULLI && k.rule != NAER_RULE_FIXED_INDEGREE)
                return fail(NAER_ERR_INVALID, "network %u connection %u: unknown rule %u", ni, c, k.rule);
            if (k.rule == NAER_RULE_BERNOULLI && !(k.probability >= 0.0 && k.probability <= 1.0))
                return fail(NAER_ERR_INVALID, "network %u connection %u: probability must be in [0, 1]", ni, c);
            if (k.plastic) {
                if (stdp_seen && std::memcmp(&stdp, &d.stdp, sizeof stdp) != 0)
                    return fail(NAER_ERR_INVALID, "plastic networks of a pack must share the STDP parameters");
                stdp = d.stdp;
                stdp_seen = true;
                n->plastic = true;
            }
            n->dmin = std::min(n->dmin, k.delay_min);
            n->dmax = std::max(n->dmax, k.delay_max);
            k.src_pop += info.pop_begin;
            k.dst_pop += info.pop_begin;
            n->conns.push_back(k);
            n->conn_local.push_back(c);
        }
        rec += d.record_capacity;
        n->nets.push_back(info);
    }
    if (N > 0xFFFFFFF0ull)
        return fail(NAER_ERR_CAPACITY, "a pack is limited to 2^32 neurons");
    n->N = (uint32_t)N;
    if (stdp_seen) {
        n->desc.stdp = stdp;
        if (!(stdp.tau_plus_ms > 0.0f) || !(stdp.tau_minus_ms > 0.0f) || !(stdp.w_max >= stdp.w_min))
            return fail(NAER_ERR_INVALID, "STDP needs tau_plus, tau_minus > 0 and w_max >= w_min");
    }
    n->desc.record_capacity = (uint32_t)std::min<uint64_t>(rec, 0xFFFFFFFFull);

    const uint32_t P = (uint32_t)n->pops.size();
    n->pp.resize(P);
    n->cdf.assign((size_t)P * NAER_DRIVE_TABLE, 2.0f);
    std::vector<uint32_t> pois_in_net(n->nets.size(), 0);
    uint64_t acc = 0;
    for (uint32_t p = 0; p < P; ++p) {
        const naer_pop_desc& q = n->pops[p];
        const uint32_t ni = n->pop_net[p];
        const naer_net_desc& d = n->net_descs[ni];
        PopParams& r = n->pp[p];
        std::memset(&r, 0, sizeof r);
        r.begin = (uint32_t)acc;
        acc += q.size;
        r.end = (uint32_t)acc;
        r.kind = q.kind;
        r.cdf_offset = p * NAER_DRIVE_TABLE;
        r.net_base = n->nets[ni].neuron_begin;
        r.seed_lo = (uint32_t)d.seed;
        r.seed_hi = (uint32_t)(d.seed >> 32);
        r.pois_seed = d.poisson_seed;
        r.t_start = n->nets[ni].t_start;
        if (q.kind == NAER_POP_POISSON) {
            if (kEpochUs % n->dt_us != 0)
                return fail(NAER_ERR_INVALID, "POISSON sources need dt dividing the 1500 us epoch");
            r.pois_base = pois_in_net[ni];
            pois_in_net[ni] += q.size;
            r.neg_rate_per_us = -((double)q.src_rate_hz / 1000000.0);
        } else if (q.kind == NAER_POP_LIF) {
            if (!(q.tau_m_ms > 0.0f))
                return fail(NAER_ERR_INVALID, "population %u: tau_m must be > 0", p);
            r.P = (float)std::exp(-(double)d0.dt_ms / (double)q.tau_m_ms);
            r.v_th = q.v_th;
            r.v_reset = q.v_reset;
            r.v_rest = q.v_rest;
            r.v0_min = q.v0_min;
            r.v0_max = q.v0_max;
            r.w_ext = q.ext_weight;
            r.n_ref = (int32_t)std::llround((double)q.t_ref_ms / (double)d0.dt_ms);
            r.has_drive = q.ext_rate_hz > 0.0f ? 1u : 0u;
            const double lam = (double)q.ext_rate_hz * (double)d0.dt_ms * 1e-3;
            if (!(lam >= 0.0 && lam <= 20.0))
                return fail(NAER_ERR_INVALID, "population %u: drive rate*dt must be in [0, 20]", p);
            // inversion table, see NAER_DRIVE_TABLE in naer.h
            float* c = &n->cdf[(size_t)p * NAER_DRIVE_TABLE];
            double pk = std::exp(-lam), cum = pk;
            for (uint32_t k = 0;; ++k) {
                const float f = (float)cum;
                if (f >= 1.0f || k == NAER_DRIVE_TABLE - 1) {
                    c[k] = 2.0f;
                    break;
                }
                c[k] = f;
                pk = pk * lam / (double)(k + 1);
                cum = cum + pk;
            }
        } else {
            return fail(NAER_ERR_INVALID, "population %u: unknown kind %u", p, q.kind);
        }
    }
    for (uint32_t p = 0; p < P; ++p)
        n->pp[p].n_pois = pois_in_net[n->pop_net[p]];
    for (uint32_t p = 0; p < P; ++p)
        if (n->pops[p].kind == NAER_POP_POISSON)
            n->n_pois += n->pops[p].size;
    // first population of every update block
    const uint32_t nblk = (n->N + kBlock - 1) / kBlock;
    n->blk_pop_h.resize(nblk);
    uint32_t p = 0;
    for (uint32_t b = 0; b < nblk; ++b) {
        const uint32_t j = b * kBlock;
        while (p + 1 < P && j >= n->pp[p + 1].begin)
            ++p;
        // skip empty populations that start at the same neuron
        n->blk_pop_h[b] = p;
    }
    n->D = n->dmax + 1;
    // Steps per update launch: inputs of a batch were all scattered before
    // it starts when B <= minimum delay.  Plastic nets interleave scatter and
    // potentiation every step.
    n->B = n->plastic ? 1u : std::min<uint32_t>(kMaxBatch, n->dmin == 0xFFFFFFFFu ? kMaxBatch : n->dmin);
    return NAER_OK;
}

UpdateArgs make_update_args(naer_net* n)
{
    UpdateArgs a;
    std::memset(&a, 0, sizeof a);
    a.V = n->V;
    a.ref = n->ref;
    a.last = n->last;
    a.xh = n->xh;
    a.yh = n->yh;
    a.pnext = n->pnext;
    a.pcnt = n->pcnt;
    a.ring = n->ring;
    a.row_ptr = n->row_ptr;
    a.list0 = n->list0;
    a.list1 = n->list1;
    a.seg_count = n->seg_count;
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
    a.N = n->N;
    a.D = n->D;
    a.B = n->B;
    a.dt_us = n->dt_us;
    a.n_pops = (uint32_t)n->pp.size();
    if (n->plastic) {
        a.dec_p = (float)std::exp(-(double)n->desc.dt_ms / (double)n->desc.stdp.tau_plus_ms);
        a.dec_m = (float)std::exp(-(double)n->desc.dt_ms / (double)n->desc.stdp.tau_minus_ms);
    }
    return a;
}

ConnArgs make_conn_args(naer_net* n, uint32_t c)
{
    const naer_conn_desc& k = n->conns[c];
    ConnArgs a;
    std::memset(&a, 0, sizeof a);
    a.src_begin = n->pp[k.src_pop].begin;
    a.n_src = n->pops[k.src_pop].size;
    a.dst_begin = n->pp[k.dst_pop].begin;
    a.n_dst = n->pops[k.dst_pop].size;
    a.K = k.indegree;
    a.conn_index = n->conn_local[c];
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

// exclusive scan of per-row counts into a device row_ptr
int scan_rows(naer_net* n, const uint32_t* d_deg, unsigned long long* d_ptr, uint64_t* total,
              uint32_t* max_len)
{
    std::vector<uint32_t> deg(n->N);
    HIP_TRY(hipMemcpyAsync(deg.data(), d_deg, n->N * sizeof(uint32_t), hipMemcpyDeviceToHost, n->stream));
    HIP_TRY(hipStreamSynchronize(n->stream));
    std::vector<unsigned long long> ptr(n->N + 1);
    unsigned long long acc = 0;
    uint32_t mx = 0;
    for (uint32_t j = 0; j < n->N; ++j) {
        ptr[j] = acc;
        acc += deg[j];
        mx = std::max(mx, deg[j]);
    }
    ptr[n->N] = acc;
    HIP_TRY(hipMemcpyAsync(d_ptr, ptr.data(), (n->N + 1) * sizeof(unsigned long long),
                           hipMemcpyHostToDevice, n->stream));
    HIP_TRY(hipStreamSynchronize(n->stream));
    *total = acc;
    if (max_len)
        *max_len = mx;
    return NAER_OK;
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

}  // namespace

extern "C" {

const char* naer_last_error(void) { return g_err.c_str(); }

const char* naer_version(void) { return "naer-mi355x 0.1 (gfx950)"; }

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

int naer_pack_create(naer_ctx* ctx, const naer_net_desc* descs, uint32_t n_nets, naer_net** out)
{
    if (!ctx || !descs || !out || n_nets == 0)
        return fail(NAER_ERR_INVALID, "NULL argument or empty pack");
    naer_net* n = new naer_net;
    n->ctx = ctx;
    n->net_descs.assign(descs, descs + n_nets);
    int rc = validate_and_derive(n);
    // the caller's arrays are not referenced after creation
    for (auto& d : n->net_descs) {
        d.pops = nullptr;
        d.conns = nullptr;
    }
    n->desc.pops = n->pops.data();
    n->desc.conns = n->conns.data();
    n->desc.n_pops = (uint32_t)n->pops.size();
    n->desc.n_conns = (uint32_t)n->conns.size();
    if (rc != NAER_OK) {
        delete n;
        return rc;
    }
    hipError_t e = hipSetDevice(ctx->device);
    if (e == hipSuccess)
        e = hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete n;
        return fail(NAER_ERR_HIP, "stream creation failed: %s", hipGetErrorString(e));
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

int naer_pack_layout(naer_net* n, uint32_t* neuron_begin, uint32_t* n_neurons, uint64_t* t_start)
{
    if (!n)
        return fail(NAER_ERR_INVALID, "net is NULL");
    for (size_t i = 0; i < n->nets.size(); ++i) {
        if (neuron_begin)
            neuron_begin[i] = n->nets[i].neuron_begin;
        if (n_neurons)
            n_neurons[i] = n->nets[i].n_neurons;
        if (t_start)
            t_start[i] = n->nets[i].t_start;
    }
    return NAER_OK;
}

int naer_net_build(naer_net* n)
{
    if (!n)
        return fail(NAER_ERR_INVALID, "net is NULL");
    if (n->built)
        return fail(NAER_ERR_STATE, "net already built");
    HIP_TRY(hipSetDevice(n->ctx->device));
    const uint32_t N = n->N;
    int rc;

    // 1. out-degree of every source over all connections
    uint32_t* deg = nullptr;
    unsigned long long* cursor = nullptr;
    if ((rc = n->alloc(&deg, N)) || (rc = n->alloc(&n->row_ptr, (size_t)N + 1)))
        return rc;
    HIP_TRY(hipMemsetAsync(deg, 0, N * sizeof(uint32_t), n->stream));
    for (uint32_t c = 0; c < n->conns.size(); ++c) {
        ConnArgs a = make_conn_args(n, c);
        HIP_TRY(launch_conn(a, n->conns[c].rule, false, deg, nullptr, nullptr, nullptr, n->stream));
    }
    // 2. row pointers
    if ((rc = scan_rows(n, deg, n->row_ptr, &n->S, &n->max_row)))
        return rc;
    if (n->max_row > kMaxSortedRow)
        return fail(NAER_ERR_CAPACITY, "a source has %u synapses; rows are limited to %u",
                    n->max_row, kMaxSortedRow);
    if (n->plastic && n->S >= 0xFFFFFFFFull)
        return fail(NAER_ERR_CAPACITY, "plastic nets are limited to 2^32-1 synapses");
    // 3. placement, then canonical row order
    if ((rc = n->alloc(&n->col, n->S)) || (rc = n->alloc(&n->w, n->S)) || (rc = n->alloc(&cursor, N)))
        return rc;
    HIP_TRY(hipMemcpyAsync(cursor, n->row_ptr, N * sizeof(unsigned long long), hipMemcpyDeviceToDevice,
                           n->stream));
    for (uint32_t c = 0; c < n->conns.size(); ++c) {
        ConnArgs a = make_conn_args(n, c);
        HIP_TRY(launch_conn(a, n->conns[c].rule, true, nullptr, cursor, n->col, n->w, n->stream));
    }
    if (n->S)
        HIP_TRY(launch_sort_rows(n->row_ptr, n->col, n->w, N, n->max_row, n->stream));
    // 4. reverse index of plastic synapses
    if (n->plastic) {
        HIP_TRY(hipMemsetAsync(deg, 0, N * sizeof(uint32_t), n->stream));
        HIP_TRY(launch_reverse(false, n->row_ptr, n->col, N, deg, nullptr, nullptr, nullptr, n->stream));
        if ((rc = n->alloc(&n->rev_ptr, (size_t)N + 1)))
            return rc;
        if ((rc = scan_rows(n, deg, n->rev_ptr, &n->n_plastic, nullptr)))
            return rc;
        if ((rc = n->alloc(&n->rev_pos, n->n_plastic)) || (rc = n->alloc(&n->rev_src, n->n_plastic)))
            return rc;
        HIP_TRY(hipMemcpyAsync(cursor, n->rev_ptr, N * sizeof(unsigned long long),
                               hipMemcpyDeviceToDevice, n->stream));
        HIP_TRY(launch_reverse(true, n->row_ptr, n->col, N, nullptr, cursor, n->rev_pos, n->rev_src,
                               n->stream));
    }
    HIP_TRY(hipStreamSynchronize(n->stream));
    n->release(deg, N * sizeof(uint32_t));
    n->release(cursor, N * sizeof(unsigned long long));

    // 5. neuron state, delay ring, spike lists
    if ((rc = n->alloc(&n->V, N)) || (rc = n->alloc(&n->ref, N)) || (rc = n->alloc(&n->last, N)) ||
        (rc = n->alloc(&n->pnext, N)) || (rc = n->alloc(&n->pcnt, N)))
        return rc;
    if (n->plastic && ((rc = n->alloc(&n->xh, N)) || (rc = n->alloc(&n->yh, N))))
        return rc;
    const size_t ring_elems = (size_t)n->D * N;
    const size_t ring_bytes = ring_elems * (n->det ? sizeof(long long) : sizeof(float));
    if ((rc = n->alloc((char**)&n->ring, ring_bytes)))
        return rc;
    HIP_TRY(hipMemsetAsync(n->ring, 0, ring_bytes, n->stream));
    // list segment of one update block: every LIF neuron fires at most once
    // per step, a POISSON source at most 10 times per 1500 us epoch
    const uint64_t epochs = ((uint64_t)n->B * n->dt_us + kEpochUs - 1) / kEpochUs + 1;
    const uint64_t per_neuron = n->n_pois ? std::max<uint64_t>(n->B, kMaxPoissonPerEpoch * epochs) : n->B;
    n->seg_cap = (uint32_t)(kBlock * per_neuron);
    n->n_seg = (N + kBlock - 1) / kBlock;
    const size_t list_elems = (size_t)n->seg_cap * n->n_seg;
    if ((rc = n->alloc(&n->list0, list_elems)) || (rc = n->alloc(&n->list1, list_elems)) ||
        (rc = n->alloc(&n->seg_count, 2 * (size_t)n->n_seg)) || (rc = n->alloc(&n->spk_part, n->n_seg)) ||
        (rc = n->alloc(&n->ev_part, n->n_seg)))
        return rc;
    HIP_TRY(hipMemsetAsync(n->seg_count, 0, 2 * (size_t)n->n_seg * sizeof(uint32_t), n->stream));
    HIP_TRY(hipMemsetAsync(n->spk_part, 0, n->n_seg * sizeof(unsigned long long), n->stream));
    HIP_TRY(hipMemsetAsync(n->ev_part, 0, n->n_seg * sizeof(unsigned long long), n->stream));
    n->rec_cap = n->desc.record_capacity;
    if ((rc = n->alloc(&n->rec, n->rec_cap ? n->rec_cap : 1)))
        return rc;
    if ((rc = n->alloc(&n->st, 1)) || (rc = n->alloc(&n->d_cdf, n->cdf.size())) ||
        (rc = n->alloc(&n->d_pops, n->pp.size())) || (rc = n->alloc(&n->d_blk_pop, n->blk_pop_h.size())))
        return rc;
    HIP_TRY(hipMemcpyAsync(n->d_pops, n->pp.data(), n->pp.size() * sizeof(PopParams), hipMemcpyHostToDevice,
                           n->stream));
    HIP_TRY(hipMemcpyAsync(n->d_blk_pop, n->blk_pop_h.data(), n->blk_pop_h.size() * sizeof(uint32_t),
                           hipMemcpyHostToDevice, n->stream));
    HIP_TRY(hipMemsetAsync(n->st, 0, sizeof(DevState), n->stream));
    HIP_TRY(hipMemcpyAsync(n->d_cdf, n->cdf.data(), n->cdf.size() * sizeof(float), hipMemcpyHostToDevice,
                           n->stream));
    UpdateArgs ua = make_update_args(n);
    HIP_TRY(launch_init_state(ua, n->stream));
    HIP_TRY(hipStreamSynchronize(n->stream));

    n->scatter_chunks = std::max<uint32_t>(1, (n->max_row + kChunk - 1) / kChunk);
    n->scatter_grid = std::max<uint32_t>(4, std::min<uint32_t>(1024, 2048 / n->scatter_chunks));
    n->built = true;
    return NAER_OK;
}

static hipError_t do_update(naer_net* n, void* arg)
{
    UpdateArgs* a = (UpdateArgs*)arg;
    return launch_update(*a, n->det, n->plastic, n->stream);
}

static hipError_t do_scatter(naer_net* n, void* arg)
{
    ScatterArgs* a = (ScatterArgs*)arg;
    return launch_scatter(*a, n->det, n->plastic, n->scatter_grid, n->scatter_chunks, n->stream);
}

static hipError_t do_potentiate(naer_net* n, void* arg)
{
    PotentiateArgs* a = (PotentiateArgs*)arg;
    return launch_potentiate(*a, 256, n->stream);
}

static int enqueue_steps(naer_net* n, uint32_t n_steps, bool timed)
{
    UpdateArgs ua = make_update_args(n);
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
    sa.yh = n->yh;
    sa.a_minus = n->desc.stdp.a_minus;
    sa.w_min = n->desc.stdp.w_min;
    sa.w_max = n->desc.stdp.w_max;
    sa.N = n->N;
    sa.D = n->D;
    sa.is_last = n->plastic ? 0u : 1u;
    PotentiateArgs pa;
    std::memset(&pa, 0, sizeof pa);
    pa.list0 = n->list0;
    pa.list1 = n->list1;
    pa.seg_count = n->seg_count;
    pa.seg_cap = n->seg_cap;
    pa.n_seg = n->n_seg;
    pa.st = n->st;
    pa.rev_ptr = n->rev_ptr;
    pa.rev_pos = n->rev_pos;
    pa.rev_src = n->rev_src;
    pa.w = n->w;
    pa.xh = n->xh;
    pa.a_plus = n->desc.stdp.a_plus;
    pa.w_min = n->desc.stdp.w_min;
    pa.w_max = n->desc.stdp.w_max;
    int rc;
    while (n_steps > 0) {
        const uint32_t b = std::min(n_steps, n->B);
        ua.B = b;
        sa.B = b;
        pa.B = b;
        if ((rc = launch_timed(n, NAER_K_UPDATE, do_update, &ua, timed)))
            return rc;
        if ((rc = launch_timed(n, NAER_K_SCATTER, do_scatter, &sa, timed)))
            return rc;
        if (n->plastic && (rc = launch_timed(n, NAER_K_POTENTIATE, do_potentiate, &pa, timed)))
            return rc;
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
    if (n->graph_chunk) {
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
    int rc = naer_net_sync(n);
    if (rc)
        return rc;
    if (n->graph_exec) {
        HIP_TRY(hipGraphExecDestroy(n->graph_exec));
        n->graph_exec = nullptr;
    }
    n->graph_chunk = chunk_steps;
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
    out->n_neurons = n->N;
    out->n_synapses = n->S;
    out->n_plastic = n->n_plastic;
    out->ring_depth = n->D;
    out->batch = n->B;
    out->max_row = n->max_row;
    out->device_bytes = n->dev_bytes;
    out->kernels = 4;
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
    if (pre) {
        if (n->xh)
            HIP_TRY(hipMemcpy(pre, n->xh, N * sizeof(float), hipMemcpyDeviceToHost));
        else
            std::memset(pre, 0, N * sizeof(float));
    }
    if (post) {
        if (n->yh)
            HIP_TRY(hipMemcpy(post, n->yh, N * sizeof(float), hipMemcpyDeviceToHost));
        else
            std::memset(post, 0, N * sizeof(float));
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
    if (col && n->S)
        HIP_TRY(hipMemcpy(col, n->col, n->S * sizeof(uint32_t), hipMemcpyDeviceToHost));
    if (w && n->S)
        HIP_TRY(hipMemcpy(w, n->w, n->S * sizeof(float), hipMemcpyDeviceToHost));
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
    if (n->graph_exec) {
        HIP_TRY(hipGraphExecDestroy(n->graph_exec));
        n->graph_exec = nullptr;
    }
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

int naer_ref_poisson_epoch(naer_ctx* ctx, uint32_t n, uint32_t n_pois_total, uint32_t seed,
                           float rate_hz, uint32_t t0, uint32_t t1, uint32_t* times, uint32_t* count)
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
        e = launch_ref_poisson_epoch(n, n_pois_total, seed, -((double)rate_hz / 1000000.0), t0, t1, dt,
                                     dc, nullptr);
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
