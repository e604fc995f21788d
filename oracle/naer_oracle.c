/*
 * naer_oracle.c — CPU oracle of the simulation step.  TEST INFRASTRUCTURE:
 * only tests/, the smoke check and bench.py's cpu_baseline leg use it, never
 * the product path.  Written independently of naer_amd/csrc.
 *
 * Parity status of each part (how it is pinned is stated in the tests):
 *   orc_philox                 Random123 known-answer vectors
 *   orc_ref_uniform/_isi_us    reference values computed from the reference
 *                              map + closed-form exponential-ISI statistics
 *   orc_ref_poisson_epoch      reference first-draw values, rate/ISI statistics
 *   orc_ref_bernoulli_row      reference (src, dst) lists and synapse counts
 *   net step (LIF/ring/STDP)   closed forms (propagator, f-I curve, STDP pair
 *                              window), brute force across ORC_FIXED / ORC_FLOAT
 *                              / ORC_DOUBLE
 */
#include "naer_oracle.h"

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* errors                                                                   */
/* ------------------------------------------------------------------------ */
static __thread char g_err[512];
static void set_err(const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}
const char* orc_error(void) { return g_err; }

/* ------------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as  */
/* easy as 1, 2, 3", SC'11): 10 rounds of                                  */
/*   (x0,x1,x2,x3) -> (hi(M1*x2)^x1^k0, lo(M1*x2), hi(M0*x0)^x3^k1, lo(M0*x0))*/
/* with the key bumped by the Weyl constants between rounds.                */
/* Used by the reference through Random123 (random.cl:4,8-16).              */
/* ------------------------------------------------------------------------ */
#define PHILOX_M0 0xD2511F53u
#define PHILOX_M1 0xCD9E8D57u
#define PHILOX_W0 0x9E3779B9u
#define PHILOX_W1 0xBB67AE85u

void orc_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t x0 = ctr[0], x1 = ctr[1], x2 = ctr[2], x3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)PHILOX_M0 * x0;
        uint64_t p1 = (uint64_t)PHILOX_M1 * x2;
        uint32_t y0 = (uint32_t)(p1 >> 32) ^ x1 ^ k0;
        uint32_t y1 = (uint32_t)p1;
        uint32_t y2 = (uint32_t)(p0 >> 32) ^ x3 ^ k1;
        uint32_t y3 = (uint32_t)p0;
        x0 = y0;
        x1 = y1;
        x2 = y2;
        x3 = y3;
        k0 += PHILOX_W0;
        k1 += PHILOX_W1;
    }
    out[0] = x0;
    out[1] = x1;
    out[2] = x2;
    out[3] = x3;
}

void orc_philox_n(const uint32_t* ctr, const uint32_t* key, uint32_t* out, uint64_t n)
{
    for (uint64_t i = 0; i < n; ++i)
        orc_philox(ctr + 4 * i, key + 2 * i, out + 4 * i);
}

static inline void philox_seeded(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                 uint64_t seed, uint32_t out[4])
{
    uint32_t c[4] = { c0, c1, c2, c3 };
    uint32_t k[2] = { (uint32_t)seed, (uint32_t)(seed >> 32) };
    orc_philox(c, k, out);
}

static inline float u24(uint32_t w) { return (float)(w >> 8) * 0x1p-24f; }
static inline uint32_t bounded(uint32_t w, uint32_t n) { return (uint32_t)(((uint64_t)w * n) >> 32); }

/* ------------------------------------------------------------------------ */
/* reference RNG maps                                                       */
/* ------------------------------------------------------------------------ */
/* uniform_rand (random.cl:6-24): key {get_global_id(0), seed}, counter
 * {iteration,0,0,0}, third output word divided by UINT_MAX in fp64. */
double orc_ref_uniform(uint32_t gid, uint32_t seed, uint32_t iteration)
{
    uint32_t c[4] = { iteration, 0, 0, 0 };
    uint32_t k[2] = { gid, seed };
    uint32_t o[4];
    orc_philox(c, k, o);
    return (double)o[2] / 4294967295.0;
}

/* (uint)exprnd_fast(rate, seed, iteration) (random.cl:40-44 and the cast at
 * input_neuron_group.cl:29,37): log(1-u) / -(rate/1e6) in fp64, truncated to
 * an unsigned 32-bit microsecond count.  The OpenCL cast of an out-of-range
 * value is undefined; it saturates here (u == 1 gives +inf). */
uint32_t orc_ref_isi_us(float rate_hz, uint32_t gid, uint32_t seed, uint32_t iteration)
{
    double u = orc_ref_uniform(gid, seed, iteration);
    double r = (double)rate_hz / 1000000.0;
    double isi = log(1.0 - u) / -r;
    if (!(isi < 4294967296.0))
        return 0xFFFFFFFFu;
    return (uint32_t)isi;
}

/* get_global_id(0) of POISSON neuron idx under the reference launch of
 * 128 work-items per 2048-neuron group (Network.cpp:181-184,
 * input_neuron_group.cl:17,26). */
uint32_t orc_ref_poisson_gid(uint32_t idx) { return (idx / 2048u) * 128u + idx % 128u; }

/* One epoch [t0, t1) of the train of POISSON neuron idx
 * (input_neuron_group.cl:26-39): first spike at t0 + ISI drawn with counter
 * t0*n_pois + idx, each next one ISI later with counter time*n_pois + idx
 * (uint32 arithmetic), at most 10 spikes. */
int orc_ref_poisson_epoch(uint32_t idx, uint32_t n_pois, uint32_t seed, float rate_hz,
                          uint32_t t0, uint32_t t1, uint32_t* times)
{
    uint32_t gid = orc_ref_poisson_gid(idx);
    uint32_t time = t0 + orc_ref_isi_us(rate_hz, gid, seed, t0 * n_pois + idx);
    int n = 0;
    for (int m = 0; m < 10; ++m) {
        if (time < t1) {
            times[n++] = time;
            time += orc_ref_isi_us(rate_hz, gid, seed, time * n_pois + idx);
        }
    }
    return n;
}

/* count_connections / generate_connections (network.cl:34-41, 88-95): the
 * work-group of source src draws for destination dst with global id
 * src*32 + dst%32 and iteration dst. */
uint32_t orc_ref_bernoulli_row(uint32_t src, uint32_t n_dst, double p, uint32_t seed32,
                               uint32_t* out, uint32_t cap)
{
    uint32_t n = 0;
    for (uint32_t dst = 0; dst < n_dst; ++dst) {
        if (orc_ref_uniform(src * 32u + dst % 32u, seed32, dst) < p) {
            if (n < cap)
                out[n] = dst;
            ++n;
        }
    }
    return n;
}

/* ------------------------------------------------------------------------ */
/* network                                                                  */
/* ------------------------------------------------------------------------ */
struct orc_net {
    naer_net_desc desc;
    naer_pop_desc* pops;
    naer_conn_desc* conns;
    int prec;
    int threads;
    int built;

    uint32_t N;
    uint64_t S;
    uint32_t* pop_base;
    uint16_t* pop_of;
    uint32_t* pois_idx;
    uint32_t n_pois;
    uint32_t dt_us;

    /* per-population derived constants */
    float* P;
    double* Pd;
    int32_t* nref;
    float* cdf; /* NAER_DRIVE_TABLE per pop */

    uint64_t* row_ptr;
    uint32_t* col;
    float* w;
    double* wd;
    uint64_t n_plastic;
    uint64_t* rev_ptr;
    uint64_t* rev_pos;
    uint32_t* rev_src;

    float* V;
    double* Vd;
    int32_t* ref;
    int32_t* last;
    /* pair STDP: a neuron's traces at step t count its spikes before t,
     * x(t) = (a_s + 1) dec^(t - s) for its last spike s < t, where a_s is the
     * trace at s; xh/yh (xhd/yhd) hold a_s of the last spike (pre / post). */
    float *xh, *yh;
    double *xhd, *yhd;
    uint32_t* pnext;
    uint32_t* pcnt;
    int has_plastic;
    float *tp, *tm; /* dec^k in fp32 while exp(-k dt / tau) >= 2^-40 */
    uint32_t n_tp, n_tm;
    double dt_d, tau_pd, tau_md;

    uint32_t D;
    int64_t* ring_fx;
    float* ring_f;
    double* ring_d;
    double* ring_ex;
    double* ring_abs;

    uint64_t t;
    uint64_t spikes, events;

    naer_spike* rec;
    uint64_t rec_n, rec_cap;

    uint64_t probe_t0;
    uint32_t probe_n;
    float* probe_used;
    double* probe_exact;
    double* probe_abs;

    /* spikes of the current step */
    uint32_t* sp;
    uint32_t* sp_time;
    uint64_t sp_n, sp_cap;
    /* per-thread spike buffers of the parallel update */
    uint32_t** tsp;
    uint32_t** tsp_time;
    uint64_t* tsp_n;
};

static float* dec_table(float dt_ms, float tau_ms, uint32_t* len);
static float trace_f(const orc_net* n, int pre, uint32_t j, uint64_t t);
static double trace_d(const orc_net* n, int pre, uint32_t j, uint64_t t);

static void* xcalloc(size_t n, size_t sz)
{
    if (n == 0)
        n = 1;
    return calloc(n, sz);
}

void orc_net_destroy(orc_net* n)
{
    if (!n)
        return;
    free(n->pops);
    free(n->conns);
    free(n->pop_base);
    free(n->pop_of);
    free(n->pois_idx);
    free(n->P);
    free(n->Pd);
    free(n->nref);
    free(n->cdf);
    free(n->row_ptr);
    free(n->col);
    free(n->w);
    free(n->wd);
    free(n->rev_ptr);
    free(n->rev_pos);
    free(n->rev_src);
    free(n->V);
    free(n->Vd);
    free(n->ref);
    free(n->last);
    free(n->xh);
    free(n->yh);
    free(n->tp);
    free(n->tm);
    free(n->xhd);
    free(n->yhd);
    free(n->pnext);
    free(n->pcnt);
    free(n->ring_fx);
    free(n->ring_f);
    free(n->ring_d);
    free(n->ring_ex);
    free(n->ring_abs);
    free(n->rec);
    free(n->probe_used);
    free(n->probe_exact);
    free(n->probe_abs);
    free(n->sp);
    free(n->sp_time);
    if (n->tsp) {
        for (int i = 0; i < n->threads; ++i) {
            free(n->tsp[i]);
            free(n->tsp_time[i]);
        }
    }
    free(n->tsp);
    free(n->tsp_time);
    free(n->tsp_n);
    free(n);
}

static void drive_table(double lam, float* cdf)
{
    double p = exp(-lam);
    double c = p;
    int k = 0;
    for (;;) {
        float f = (float)c;
        if (f >= 1.0f || k == NAER_DRIVE_TABLE - 1) {
            cdf[k] = 2.0f;
            for (int j = k + 1; j < NAER_DRIVE_TABLE; ++j)
                cdf[j] = 2.0f;
            return;
        }
        cdf[k] = f;
        ++k;
        p = p * lam / (double)k;
        c = c + p;
    }
}

orc_net* orc_net_create(const naer_net_desc* d, int precision, int threads)
{
    if (!d || d->n_pops == 0 || !d->pops) {
        set_err("net needs at least one population");
        return NULL;
    }
    if (precision < ORC_FIXED || precision > ORC_DOUBLE) {
        set_err("bad precision");
        return NULL;
    }
    orc_net* n = (orc_net*)calloc(1, sizeof *n);
    n->desc = *d;
    n->prec = precision;
    n->threads = threads < 1 ? 1 : threads;
    n->pops = (naer_pop_desc*)xcalloc(d->n_pops, sizeof(naer_pop_desc));
    memcpy(n->pops, d->pops, d->n_pops * sizeof(naer_pop_desc));
    n->conns = (naer_conn_desc*)xcalloc(d->n_conns, sizeof(naer_conn_desc));
    if (d->n_conns)
        memcpy(n->conns, d->conns, d->n_conns * sizeof(naer_conn_desc));
    n->desc.pops = n->pops;
    n->desc.conns = n->conns;

    double dtus = (double)d->dt_ms * 1000.0;
    n->dt_us = (uint32_t)llround(dtus);
    if (!(d->dt_ms > 0.0f) || fabs(dtus - (double)n->dt_us) > 1e-3 || n->dt_us == 0) {
        set_err("dt must be a positive whole number of microseconds");
        orc_net_destroy(n);
        return NULL;
    }

    uint64_t N = 0;
    n->pop_base = (uint32_t*)xcalloc(d->n_pops + 1, sizeof(uint32_t));
    for (uint32_t p = 0; p < d->n_pops; ++p) {
        n->pop_base[p] = (uint32_t)N;
        N += n->pops[p].size;
    }
    if (N == 0 || N > (1u << 24)) {
        set_err("neuron count must be in [1, 2^24]");
        orc_net_destroy(n);
        return NULL;
    }
    n->pop_base[d->n_pops] = (uint32_t)N;
    n->N = (uint32_t)N;
    n->pop_of = (uint16_t*)xcalloc(N, sizeof(uint16_t));
    n->pois_idx = (uint32_t*)xcalloc(N, sizeof(uint32_t));
    uint32_t pi = 0;
    for (uint32_t p = 0; p < d->n_pops; ++p) {
        for (uint32_t j = n->pop_base[p]; j < n->pop_base[p + 1]; ++j) {
            n->pop_of[j] = (uint16_t)p;
            n->pois_idx[j] = n->pops[p].kind == NAER_POP_POISSON ? pi++ : 0xFFFFFFFFu;
        }
    }
    n->n_pois = pi;
    if (pi && 1500u % n->dt_us != 0) {
        set_err("POISSON sources need dt dividing the 1500 us epoch");
        orc_net_destroy(n);
        return NULL;
    }

    n->P = (float*)xcalloc(d->n_pops, sizeof(float));
    n->Pd = (double*)xcalloc(d->n_pops, sizeof(double));
    n->nref = (int32_t*)xcalloc(d->n_pops, sizeof(int32_t));
    n->cdf = (float*)xcalloc((size_t)d->n_pops * NAER_DRIVE_TABLE, sizeof(float));
    for (uint32_t p = 0; p < d->n_pops; ++p) {
        const naer_pop_desc* q = &n->pops[p];
        if (q->kind == NAER_POP_LIF) {
            if (!(q->tau_m_ms > 0.0f)) {
                set_err("population %u: tau_m must be > 0", p);
                orc_net_destroy(n);
                return NULL;
            }
            n->Pd[p] = exp(-(double)d->dt_ms / (double)q->tau_m_ms);
            n->P[p] = (float)n->Pd[p];
            n->nref[p] = (int32_t)llround((double)q->t_ref_ms / (double)d->dt_ms);
            double lam = (double)q->ext_rate_hz * (double)d->dt_ms * 1e-3;
            if (lam > 20.0 || lam < 0.0) {
                set_err("population %u: drive rate*dt must be in [0, 20]", p);
                orc_net_destroy(n);
                return NULL;
            }
            drive_table(lam, n->cdf + (size_t)p * NAER_DRIVE_TABLE);
        } else if (q->kind != NAER_POP_POISSON) {
            set_err("population %u: unknown kind", p);
            orc_net_destroy(n);
            return NULL;
        }
    }

    uint32_t dmax = 1;
    for (uint32_t c = 0; c < d->n_conns; ++c) {
        const naer_conn_desc* k = &n->conns[c];
        if (k->src_pop >= d->n_pops || k->dst_pop >= d->n_pops) {
            set_err("connection %u: population index out of range", c);
            orc_net_destroy(n);
            return NULL;
        }
        if (n->pops[k->dst_pop].kind != NAER_POP_LIF) {
            set_err("connection %u: target population must be LIF", c);
            orc_net_destroy(n);
            return NULL;
        }
        uint32_t q = k->delay_quantum ? k->delay_quantum : 1;
        if (k->delay_min < 1 || k->delay_max < k->delay_min || k->delay_max > 127 ||
            (k->delay_max - k->delay_min) % q) {
            set_err("connection %u: delays must satisfy 1 <= min <= max <= 127 on the quantum grid", c);
            orc_net_destroy(n);
            return NULL;
        }
        if (k->plastic && n->pops[k->src_pop].kind != NAER_POP_LIF) {
            set_err("connection %u: plastic connections need a LIF source", c);
            orc_net_destroy(n);
            return NULL;
        }
        if (k->rule != NAER_RULE_BERNOULLI && k->rule != NAER_RULE_FIXED_INDEGREE) {
            set_err("connection %u: unknown rule", c);
            orc_net_destroy(n);
            return NULL;
        }
        if (k->delay_max > dmax)
            dmax = k->delay_max;
        if (k->plastic)
            n->has_plastic = 1;
    }
    n->D = dmax + 1;
    if (n->has_plastic) {
        const naer_stdp_desc* s = &d->stdp;
        n->dt_d = (double)d->dt_ms;
        n->tau_pd = (double)s->tau_plus_ms;
        n->tau_md = (double)s->tau_minus_ms;
        n->tp = dec_table(d->dt_ms, s->tau_plus_ms, &n->n_tp);
        n->tm = dec_table(d->dt_ms, s->tau_minus_ms, &n->n_tm);
    }
    return n;
}

/* dec^k = exp(-k dt / tau) rounded to fp32, for k = 0, 1, ... while the
 * exact value is >= 2^-40 (at most 2^22 entries); past the table the fp32
 * traces take dec^k = 0. */
static float* dec_table(float dt_ms, float tau_ms, uint32_t* len)
{
    uint32_t cap = 1024, n = 0;
    float* t = (float*)malloc(cap * sizeof(float));
    for (uint32_t k = 0; k < (1u << 22); ++k) {
        double v = exp(-(double)k * (double)dt_ms / (double)tau_ms);
        if (v < 0x1p-40)
            break;
        if (n == cap) {
            cap *= 2;
            t = (float*)realloc(t, cap * sizeof(float));
        }
        t[n++] = (float)v;
    }
    *len = n;
    return t;
}

/* ---------------- connectivity ------------------------------------------ */
typedef struct {
    uint32_t src; /* global */
    uint32_t word;
    float w;
} syn_t;

typedef void (*syn_cb)(void* ctx, uint32_t src, uint32_t word, float w);

static inline uint32_t delay_of(const naer_conn_desc* k, uint32_t word)
{
    uint32_t q = k->delay_quantum ? k->delay_quantum : 1;
    uint32_t choices = (k->delay_max - k->delay_min) / q + 1;
    return k->delay_min + q * bounded(word, choices);
}

/* Test hook: plants one known mistake in the fixed in-degree rule, so the
 * connectivity tests can show that their checks catch it.  0 (the default)
 * is the rule as specified; the other values are listed in naer_oracle.h. */
static uint32_t g_conn_mutation = ORC_MUT_NONE;
void orc_set_conn_mutation(uint32_t m) { g_conn_mutation = m; }

/* Enumerates the synapses of connection c, calling cb for each.
 * FIXED_INDEGREE: target j (index within the target population) takes
 * indegree draws k; draw k is Philox block {j, k, NAER_STREAM_CONN, c} under
 * the connection's 64-bit seed: word 0 -> source (uniform over the source
 * population), word 1 -> delay class, word 2 -> weight when weight == 0. */
static void gen_conn(const orc_net* n, uint32_t c, syn_cb cb, void* ctx)
{
    const naer_conn_desc* k = &n->conns[c];
    uint32_t sb = n->pop_base[k->src_pop], ns = n->pops[k->src_pop].size;
    uint32_t db = n->pop_base[k->dst_pop], nd = n->pops[k->dst_pop].size;
    uint32_t pl = k->plastic ? 0x80000000u : 0u;
    if (k->rule == NAER_RULE_FIXED_INDEGREE) {
        const uint32_t mut = g_conn_mutation;
        const uint32_t s_base = mut == ORC_MUT_SRC_BASE ? db : sb;
        const uint32_t s_range = mut == ORC_MUT_SRC_RANGE ? nd : ns;
        for (uint32_t j = 0; j < nd; ++j) {
            for (uint32_t kk = 0; kk < k->indegree; ++kk) {
                uint32_t r[4];
                philox_seeded(j, kk, NAER_STREAM_CONN, c, k->seed, r);
                uint32_t w_src = r[0], w_del = r[1], w_wt = r[2];
                if (mut == ORC_MUT_WORDS_SWAPPED) {
                    w_src = r[1];
                    w_del = r[0];
                } else if (mut == ORC_MUT_DELAY_REUSE) {
                    w_del = r[0];
                } else if (mut == ORC_MUT_WEIGHT_REUSE) {
                    w_wt = r[0];
                }
                uint32_t src = s_base + bounded(w_src, s_range);
                if (mut != ORC_MUT_NONE)
                    src %= n->N;  /* a planted mistake still names a neuron of the net */
                uint32_t del = delay_of(k, w_del);
                float w = k->weight;
                if (k->weight == 0.0f) {
                    float span = mut == ORC_MUT_WEIGHT_SPAN ? k->wmax : k->wmax - k->wmin;
                    float x = span * u24(w_wt);
                    w = k->wmin + x;
                }
                cb(ctx, src, (db + j) | (del << 24) | pl, w);
            }
        }
    } else {
        uint32_t seed32 = (uint32_t)k->seed;
        for (uint32_t s = 0; s < ns; ++s) {
            for (uint32_t j = 0; j < nd; ++j) {
                uint32_t cc[4] = { j, 0, 0, 0 };
                uint32_t kk[2] = { s * 32u + j % 32u, seed32 };
                uint32_t r[4];
                orc_philox(cc, kk, r);
                if ((double)r[2] / 4294967295.0 < k->probability) {
                    uint32_t del = delay_of(k, r[1]);
                    float w = k->weight;
                    if (k->weight == 0.0f) {
                        float span = k->wmax - k->wmin;
                        float x = span * u24(r[3]);
                        w = k->wmin + x;
                    }
                    cb(ctx, sb + s, (db + j) | (del << 24) | pl, w);
                }
            }
        }
    }
}

static void cb_count(void* ctx, uint32_t src, uint32_t word, float w)
{
    (void)word;
    (void)w;
    ((uint64_t*)ctx)[src] += 1;
}

typedef struct {
    uint64_t* cursor;
    uint64_t* keys;
} place_ctx;

/* Canonical row order: ascending (row_key(word), weight bits).  row_key puts
 * the target's 2^14-neuron group first, then the word itself (plastic bit,
 * delay, target); a bijection of the 32-bit word for targets below 2^24. */
#define ORC_GROUP_SHIFT 14u
static inline uint32_t row_key(uint32_t word)
{
    uint32_t t = word & 0x00FFFFFFu;
    return ((t >> ORC_GROUP_SHIFT) << 22) | ((word >> 31) << 21) | (((word >> 24) & 0x7Fu) << 14) |
           (t & ((1u << ORC_GROUP_SHIFT) - 1u));
}

static inline uint32_t row_word(uint32_t key)
{
    uint32_t t = ((key >> 22) << ORC_GROUP_SHIFT) | (key & ((1u << ORC_GROUP_SHIFT) - 1u));
    return t | (((key >> 14) & 0x7Fu) << 24) | (((key >> 21) & 1u) << 31);
}

static void cb_place(void* ctx, uint32_t src, uint32_t word, float w)
{
    place_ctx* p = (place_ctx*)ctx;
    uint32_t wb;
    memcpy(&wb, &w, 4);
    p->keys[p->cursor[src]++] = ((uint64_t)row_key(word) << 32) | wb;
}

static int cmp_u64(const void* a, const void* b)
{
    uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
    return x < y ? -1 : (x > y);
}

int orc_net_build(orc_net* n)
{
    if (n->built) {
        set_err("already built");
        return -1;
    }
    uint32_t N = n->N;
    uint64_t* deg = (uint64_t*)xcalloc(N, sizeof(uint64_t));
    for (uint32_t c = 0; c < n->desc.n_conns; ++c)
        gen_conn(n, c, cb_count, deg);
    n->row_ptr = (uint64_t*)xcalloc(N + 1, sizeof(uint64_t));
    uint64_t S = 0;
    for (uint32_t j = 0; j < N; ++j) {
        n->row_ptr[j] = S;
        S += deg[j];
    }
    n->row_ptr[N] = S;
    n->S = S;
    uint64_t* keys = (uint64_t*)xcalloc(S, sizeof(uint64_t));
    memcpy(deg, n->row_ptr, N * sizeof(uint64_t));
    place_ctx pc = { deg, keys };
    for (uint32_t c = 0; c < n->desc.n_conns; ++c)
        gen_conn(n, c, cb_place, &pc);
    free(deg);
#pragma omp parallel for schedule(dynamic, 64) num_threads(n->threads)
    for (int64_t j = 0; j < (int64_t)N; ++j)
        qsort(keys + n->row_ptr[j], n->row_ptr[j + 1] - n->row_ptr[j], sizeof(uint64_t), cmp_u64);
    n->col = (uint32_t*)xcalloc(S, sizeof(uint32_t));
    n->w = (float*)xcalloc(S, sizeof(float));
    uint64_t np = 0;
    for (uint64_t i = 0; i < S; ++i) {
        uint32_t wb = (uint32_t)keys[i];
        n->col[i] = row_word((uint32_t)(keys[i] >> 32));
        memcpy(&n->w[i], &wb, 4);
        np += n->col[i] >> 31;
    }
    free(keys);
    n->n_plastic = np;
    if (n->prec == ORC_DOUBLE) {
        n->wd = (double*)xcalloc(S, sizeof(double));
        for (uint64_t i = 0; i < S; ++i)
            n->wd[i] = n->w[i];
    }

    /* reverse index of plastic synapses: for each target, (position, source) */
    if (n->has_plastic) {
        uint64_t* rdeg = (uint64_t*)xcalloc(N + 1, sizeof(uint64_t));
        for (uint32_t s = 0; s < N; ++s)
            for (uint64_t i = n->row_ptr[s]; i < n->row_ptr[s + 1]; ++i)
                if (n->col[i] >> 31)
                    rdeg[n->col[i] & 0xFFFFFFu] += 1;
        n->rev_ptr = (uint64_t*)xcalloc(N + 1, sizeof(uint64_t));
        uint64_t acc = 0;
        for (uint32_t j = 0; j < N; ++j) {
            n->rev_ptr[j] = acc;
            acc += rdeg[j];
        }
        n->rev_ptr[N] = acc;
        memcpy(rdeg, n->rev_ptr, N * sizeof(uint64_t));
        n->rev_pos = (uint64_t*)xcalloc(acc, sizeof(uint64_t));
        n->rev_src = (uint32_t*)xcalloc(acc, sizeof(uint32_t));
        for (uint32_t s = 0; s < N; ++s)
            for (uint64_t i = n->row_ptr[s]; i < n->row_ptr[s + 1]; ++i)
                if (n->col[i] >> 31) {
                    uint64_t r = rdeg[n->col[i] & 0xFFFFFFu]++;
                    n->rev_pos[r] = i;
                    n->rev_src[r] = s;
                }
        free(rdeg);
    }

    /* state */
    const naer_net_desc* d = &n->desc;
    n->V = (float*)xcalloc(N, sizeof(float));
    n->Vd = (double*)xcalloc(N, sizeof(double));
    n->ref = (int32_t*)xcalloc(N, sizeof(int32_t));
    n->last = (int32_t*)xcalloc(N, sizeof(int32_t));
    n->xh = (float*)xcalloc(N, sizeof(float));
    n->yh = (float*)xcalloc(N, sizeof(float));
    n->xhd = (double*)xcalloc(N, sizeof(double));
    n->yhd = (double*)xcalloc(N, sizeof(double));
    n->pnext = (uint32_t*)xcalloc(N, sizeof(uint32_t));
    n->pcnt = (uint32_t*)xcalloc(N, sizeof(uint32_t));
    for (uint32_t j = 0; j < N; ++j) {
        const naer_pop_desc* q = &n->pops[n->pop_of[j]];
        n->last[j] = -1;
        if (q->kind != NAER_POP_LIF)
            continue;
        uint32_t r[4];
        philox_seeded(j, 0, NAER_STREAM_V0, 0, d->seed, r);
        float span = q->v0_max - q->v0_min;
        float x = span * u24(r[0]);
        n->V[j] = q->v0_min + x;
        n->Vd[j] = n->V[j];
    }
    size_t ring = (size_t)n->D * N;
    if (n->prec == ORC_FIXED) {
        n->ring_fx = (int64_t*)xcalloc(ring, sizeof(int64_t));
        n->ring_ex = (double*)xcalloc(ring, sizeof(double));
        n->ring_abs = (double*)xcalloc(ring, sizeof(double));
    } else if (n->prec == ORC_FLOAT) {
        n->ring_f = (float*)xcalloc(ring, sizeof(float));
        n->ring_ex = (double*)xcalloc(ring, sizeof(double));
        n->ring_abs = (double*)xcalloc(ring, sizeof(double));
    } else {
        n->ring_d = (double*)xcalloc(ring, sizeof(double));
        n->ring_abs = (double*)xcalloc(ring, sizeof(double));
    }
    n->sp_cap = (uint64_t)N * 10 + 16;
    n->sp = (uint32_t*)xcalloc(n->sp_cap, sizeof(uint32_t));
    n->sp_time = (uint32_t*)xcalloc(n->sp_cap, sizeof(uint32_t));
    n->tsp = (uint32_t**)xcalloc(n->threads, sizeof(uint32_t*));
    n->tsp_time = (uint32_t**)xcalloc(n->threads, sizeof(uint32_t*));
    n->tsp_n = (uint64_t*)xcalloc(n->threads, sizeof(uint64_t));
    for (int i = 0; i < n->threads; ++i) {
        n->tsp[i] = (uint32_t*)xcalloc(n->sp_cap, sizeof(uint32_t));
        n->tsp_time[i] = (uint32_t*)xcalloc(n->sp_cap, sizeof(uint32_t));
    }
    n->rec_cap = d->record_capacity;
    n->rec = (naer_spike*)xcalloc(n->rec_cap, sizeof(naer_spike));
    n->built = 1;
    return 0;
}

/* ---------------- step -------------------------------------------------- */
static inline float drive_count_f(const float* cdf, uint32_t word)
{
    float u = u24(word);
    int k = 0;
    while (u >= cdf[k])
        ++k;
    return (float)k;
}

static inline uint32_t drive_word(uint64_t seed, uint32_t j, uint64_t t)
{
    uint32_t r[4];
    philox_seeded(j, (uint32_t)(t >> 2), NAER_STREAM_DRIVE, 0, seed, r);
    return r[t & 3];
}

static inline int64_t to_fixed(float w) { return llrint((double)w * 4294967296.0); }
static inline float from_fixed(int64_t a) { return (float)((double)a * 0x1p-32); }

/* Update of neuron j at step t; appends its spikes to (list, times, *cnt). */
static void update_neuron(orc_net* n, uint32_t j, uint64_t t, uint32_t* list, uint32_t* times,
                          uint64_t* cnt)
{
    const naer_net_desc* d = &n->desc;
    uint32_t p = n->pop_of[j];
    const naer_pop_desc* q = &n->pops[p];
    size_t slot = (size_t)(t % n->D) * n->N + j;
    uint32_t T = (uint32_t)(t * n->dt_us);

    if (q->kind == NAER_POP_POISSON) {
        uint32_t idx = n->pois_idx[j];
        uint32_t gid = orc_ref_poisson_gid(idx);
        uint32_t ep = T - T % 1500u;
        if (T == ep) {
            n->pnext[j] = ep + orc_ref_isi_us(q->src_rate_hz, gid, d->poisson_seed, ep * n->n_pois + idx);
            n->pcnt[j] = 0;
        }
        uint32_t end = T + n->dt_us;
        while (n->pcnt[j] < 10 && n->pnext[j] < end) {
            uint32_t tm = n->pnext[j];
            list[*cnt] = j;
            times[*cnt] = tm;
            ++*cnt;
            n->pcnt[j] += 1;
            n->pnext[j] = tm + orc_ref_isi_us(q->src_rate_hz, gid, d->poisson_seed, tm * n->n_pois + idx);
        }
        return;
    }

    /* synaptic input of this step */
    float I = 0.0f;
    double Id = 0.0;
    if (n->prec == ORC_FIXED) {
        I = from_fixed(n->ring_fx[slot]);
        n->ring_fx[slot] = 0;
    } else if (n->prec == ORC_FLOAT) {
        I = n->ring_f[slot];
        n->ring_f[slot] = 0.0f;
    } else {
        Id = n->ring_d[slot];
        n->ring_d[slot] = 0.0;
    }
    if (n->probe_n && t >= n->probe_t0 && t < n->probe_t0 + n->probe_n) {
        size_t o = (size_t)(t - n->probe_t0) * n->N + j;
        n->probe_used[o] = n->prec == ORC_DOUBLE ? (float)Id : I;
        n->probe_exact[o] = n->prec == ORC_DOUBLE ? Id : n->ring_ex[slot];
        n->probe_abs[o] = n->ring_abs[slot];
    }
    if (n->ring_ex)
        n->ring_ex[slot] = 0.0;
    n->ring_abs[slot] = 0.0;


    if (n->ref[j] > 0) {
        n->ref[j] -= 1;
        return;
    }
    int fired;
    const float* cdf = n->cdf + (size_t)p * NAER_DRIVE_TABLE;
    if (n->prec == ORC_DOUBLE) {
        double v = (n->Vd[j] - (double)q->v_rest) * n->Pd[p] + (double)q->v_rest;
        v = v + Id;
        if (q->ext_rate_hz > 0.0f)
            v = v + (double)drive_count_f(cdf, drive_word(d->seed, j, t)) * (double)q->ext_weight;
        fired = v > (double)q->v_th;
        n->Vd[j] = fired ? (double)q->v_reset : v;
    } else {
        float a = n->V[j] - q->v_rest;
        a = a * n->P[p];
        float v = a + q->v_rest;
        v = v + I;
        if (q->ext_rate_hz > 0.0f) {
            float e = drive_count_f(cdf, drive_word(d->seed, j, t)) * q->ext_weight;
            v = v + e;
        }
        fired = v > q->v_th;
        n->V[j] = fired ? q->v_reset : v;
    }
    if (fired) {
        if (n->has_plastic) {
            /* a_s of this spike: the traces at t from the previous spike */
            if (n->prec == ORC_DOUBLE) {
                double x = trace_d(n, 1, j, t), y = trace_d(n, 0, j, t);
                n->xhd[j] = x;
                n->yhd[j] = y;
            } else {
                float x = trace_f(n, 1, j, t), y = trace_f(n, 0, j, t);
                n->xh[j] = x;
                n->yh[j] = y;
            }
        }
        n->ref[j] = n->nref[p];
        n->last[j] = (int32_t)t;
        list[*cnt] = j;
        times[*cnt] = T;
        ++*cnt;
    }
}

/* Trace of neuron j at step t (its spikes before t; a spike at t itself not
 * counted): (a_s + 1) dec^(t - s) for the last spike s < t, or a_s when s =
 * t; 0 before the first spike.  pre = 1: tau_plus (x), 0: tau_minus (y).
 * fp32 with dec^k from the table (0 past it); fp64 with exp directly. */
static float trace_f(const orc_net* n, int pre, uint32_t j, uint64_t t)
{
    int32_t s = n->last[j];
    if (s < 0)
        return 0.0f;
    float a = pre ? n->xh[j] : n->yh[j];
    if ((uint64_t)s == t)
        return a;
    uint64_t k = t - (uint64_t)s;
    const float* tab = pre ? n->tp : n->tm;
    uint32_t len = pre ? n->n_tp : n->n_tm;
    float b = a + 1.0f;
    return b * (k < len ? tab[k] : 0.0f);
}

static double trace_d(const orc_net* n, int pre, uint32_t j, uint64_t t)
{
    int32_t s = n->last[j];
    if (s < 0)
        return 0.0;
    double a = pre ? n->xhd[j] : n->yhd[j];
    if ((uint64_t)s == t)
        return a;
    double k = (double)(t - (uint64_t)s);
    return (a + 1.0) * exp(-k * n->dt_d / (pre ? n->tau_pd : n->tau_md));
}

static inline float clipf(float w, float lo, float hi) { return w < lo ? lo : (w > hi ? hi : w); }
static inline double clipd(double w, double lo, double hi) { return w < lo ? lo : (w > hi ? hi : w); }

/* Delivery of the synapses of one spike of source s at step t. */
static void scatter_row(orc_net* n, uint32_t s, uint64_t t)
{
    const naer_stdp_desc* st = &n->desc.stdp;
    uint32_t N = n->N;
    uint32_t tD = (uint32_t)(t % n->D);
    for (uint64_t i = n->row_ptr[s]; i < n->row_ptr[s + 1]; ++i) {
        uint32_t word = n->col[i];
        uint32_t tgt = word & 0xFFFFFFu;
        uint32_t del = (word >> 24) & 0x7Fu;
        uint32_t slot = tD + del;
        if (slot >= n->D)
            slot -= n->D;
        size_t o = (size_t)slot * N + tgt;
        if (n->prec == ORC_DOUBLE) {
            double w = n->wd[i];
            if (word >> 31) {
                w = clipd(w - (double)st->a_minus * trace_d(n, 0, tgt, t), st->w_min, st->w_max);
                n->wd[i] = w;
            }
            n->ring_d[o] += w;
            n->ring_abs[o] += fabs(w);
        } else {
            float w = n->w[i];
            if (word >> 31) {
                float dw = st->a_minus * trace_f(n, 0, tgt, t);
                w = clipf(w - dw, st->w_min, st->w_max);
                n->w[i] = w;
            }
            if (n->prec == ORC_FIXED) {
                int64_t q = to_fixed(w);
#pragma omp atomic
                n->ring_fx[o] += q;
            } else {
                n->ring_f[o] += w;
            }
#pragma omp atomic
            n->ring_ex[o] += (double)w;
#pragma omp atomic
            n->ring_abs[o] += fabs((double)w);
        }
    }
}

static void potentiate(orc_net* n, uint32_t post, uint64_t t)
{
    const naer_stdp_desc* st = &n->desc.stdp;
    for (uint64_t r = n->rev_ptr[post]; r < n->rev_ptr[post + 1]; ++r) {
        uint64_t i = n->rev_pos[r];
        uint32_t src = n->rev_src[r];
        if (n->prec == ORC_DOUBLE) {
            n->wd[i] = clipd(n->wd[i] + (double)st->a_plus * trace_d(n, 1, src, t), st->w_min, st->w_max);
        } else {
            float dw = st->a_plus * trace_f(n, 1, src, t);
            n->w[i] = clipf(n->w[i] + dw, st->w_min, st->w_max);
        }
    }
}

static void one_step(orc_net* n)
{
    uint64_t t = n->t;
    uint32_t N = n->N;
    int nt = n->threads;
    /* update: neurons split in contiguous chunks so that concatenating the
     * per-thread lists gives neuron order */
#pragma omp parallel num_threads(nt)
    {
#ifdef _OPENMP
        int tid = omp_get_thread_num();
        int cnt = omp_get_num_threads();
#else
        int tid = 0, cnt = 1;
#endif
        uint64_t lo = (uint64_t)N * tid / cnt, hi = (uint64_t)N * (tid + 1) / cnt;
        uint64_t k = 0;
        for (uint64_t j = lo; j < hi; ++j)
            update_neuron(n, (uint32_t)j, t, n->tsp[tid], n->tsp_time[tid], &k);
        n->tsp_n[tid] = k;
        for (int i = cnt; i < nt && tid == 0; ++i)
            n->tsp_n[i] = 0;
    }
    n->sp_n = 0;
    for (int i = 0; i < nt; ++i) {
        memcpy(n->sp + n->sp_n, n->tsp[i], n->tsp_n[i] * sizeof(uint32_t));
        memcpy(n->sp_time + n->sp_n, n->tsp_time[i], n->tsp_n[i] * sizeof(uint32_t));
        n->sp_n += n->tsp_n[i];
    }
    /* record */
    for (uint64_t i = 0; i < n->sp_n; ++i) {
        uint32_t s = n->sp[i];
        n->events += n->row_ptr[s + 1] - n->row_ptr[s];
        if (n->rec_n < n->rec_cap) {
            naer_spike* r = &n->rec[n->rec_n++];
            r->step = (uint32_t)t;
            r->neuron = s;
            r->time_us = n->sp_time[i];
            r->_pad = 0;
        }
    }
    n->spikes += n->sp_n;
    /* scatter: fixed-point and fp64 shadow sums are order independent, so
     * the fixed variant may run in parallel; fp32/fp64 rings are summed in
     * canonical (spike list, CSR) order */
    if (n->prec == ORC_FIXED && nt > 1) {
#pragma omp parallel for schedule(dynamic, 4) num_threads(nt)
        for (int64_t i = 0; i < (int64_t)n->sp_n; ++i)
            scatter_row(n, n->sp[i], t);
    } else {
        for (uint64_t i = 0; i < n->sp_n; ++i)
            scatter_row(n, n->sp[i], t);
    }
    if (n->has_plastic) {
#pragma omp parallel for schedule(dynamic, 4) num_threads(nt)
        for (int64_t i = 0; i < (int64_t)n->sp_n; ++i)
            if (n->pops[n->pop_of[n->sp[i]]].kind == NAER_POP_LIF)
                potentiate(n, n->sp[i], t);
    }
    n->t = t + 1;
}

int orc_net_run(orc_net* n, uint32_t steps)
{
    if (!n->built) {
        set_err("net not built");
        return -1;
    }
    for (uint32_t i = 0; i < steps; ++i)
        one_step(n);
    return 0;
}

uint64_t orc_net_n_neurons(const orc_net* n) { return n->N; }
uint64_t orc_net_n_synapses(const orc_net* n) { return n->S; }
uint64_t orc_net_step_count(const orc_net* n) { return n->t; }
uint64_t orc_net_spike_count(const orc_net* n) { return n->spikes; }
uint64_t orc_net_event_count(const orc_net* n) { return n->events; }
uint32_t orc_net_ring_depth(const orc_net* n) { return n->D; }

void orc_net_read_csr(const orc_net* n, uint64_t* row_ptr, uint32_t* col, float* w)
{
    if (row_ptr)
        memcpy(row_ptr, n->row_ptr, (n->N + 1) * sizeof(uint64_t));
    if (col)
        memcpy(col, n->col, n->S * sizeof(uint32_t));
    if (w) {
        if (n->prec == ORC_DOUBLE)
            for (uint64_t i = 0; i < n->S; ++i)
                w[i] = (float)n->wd[i];
        else
            memcpy(w, n->w, n->S * sizeof(float));
    }
}

void orc_net_read_state(const orc_net* n, float* v, int32_t* ref, float* pre, float* post)
{
    for (uint32_t j = 0; j < n->N; ++j) {
        if (v)
            v[j] = n->prec == ORC_DOUBLE ? (float)n->Vd[j] : n->V[j];
        if (ref)
            ref[j] = n->ref[j];
        /* the traces the next step sees */
        if (pre)
            pre[j] = !n->has_plastic ? 0.0f : n->prec == ORC_DOUBLE ? (float)trace_d(n, 1, j, n->t) : trace_f(n, 1, j, n->t);
        if (post)
            post[j] = !n->has_plastic ? 0.0f : n->prec == ORC_DOUBLE ? (float)trace_d(n, 0, j, n->t) : trace_f(n, 0, j, n->t);
    }
}

void orc_net_read_state_f64(const orc_net* n, double* v)
{
    for (uint32_t j = 0; j < n->N; ++j)
        v[j] = n->prec == ORC_DOUBLE ? n->Vd[j] : (double)n->V[j];
}

uint64_t orc_net_read_spikes(orc_net* n, naer_spike* out, uint64_t cap)
{
    uint64_t k = n->rec_n < cap ? n->rec_n : cap;
    memcpy(out, n->rec, k * sizeof(naer_spike));
    memmove(n->rec, n->rec + k, (n->rec_n - k) * sizeof(naer_spike));
    n->rec_n -= k;
    return k;
}

int orc_net_set_probe(orc_net* n, uint64_t t_begin, uint32_t n_steps)
{
    free(n->probe_used);
    free(n->probe_exact);
    free(n->probe_abs);
    size_t sz = (size_t)n_steps * n->N;
    n->probe_used = (float*)xcalloc(sz, sizeof(float));
    n->probe_exact = (double*)xcalloc(sz, sizeof(double));
    n->probe_abs = (double*)xcalloc(sz, sizeof(double));
    n->probe_t0 = t_begin;
    n->probe_n = n_steps;
    return 0;
}

void orc_net_read_probe(const orc_net* n, float* used, double* exact, double* absw)
{
    size_t sz = (size_t)n->probe_n * n->N;
    if (used)
        memcpy(used, n->probe_used, sz * sizeof(float));
    if (exact)
        memcpy(exact, n->probe_exact, sz * sizeof(double));
    if (absw)
        memcpy(absw, n->probe_abs, sz * sizeof(double));
}

/* ---- event-driven mode -----------------------------------------------------
 * The reference's epoch (src/Network.cpp:122-256) with its neuron step
 * (lif.cl:15-34): epoch k covers the window [1500k, 1500k + 1500) us.  Its
 * inputs are the POISSON spikes timed in the window (input_neuron_group.cl:
 * 26-39) and the LIF spikes of the previous window, which arrive 1500 us
 * after they were fired (NAER_AXONAL_DELAY, naer_types.h:4; the synapses'
 * own delays are not used).  Every input is applied to its targets in time
 * order; inputs of one target at the same microsecond in the order (source,
 * position in the source's row).  Per input of weight w at time t:
 *   dt = t - t_last; if dt < 0 the input falls into the refractory period and
 *   is dropped; else V = v_rest + (V - v_rest) * e^(-dt/tau_m) (the product in
 *   double, rounded to float), V += w, t_last = t; if V > v_th the neuron
 *   fires at t: V = v_reset, t_last = t + t_ref.
 * e^(-dt/tau_m) = H[dt >> 12] * L[dt & 4095] with H[h] = exp(-4096 h / T),
 * L[l] = exp(-l / T), T = tau_m in us (H = 0 past the last nonzero entry); the
 * device uses the same tables.  Recorded spikes carry step = the epoch and
 * time_us = the firing (or POISSON emission) time. */

#define EV_WINDOW 1500u
#define EV_LO 4096u

/* Plastic connections (STDP population onto STDP population) follow
 * stdp.cl:8-140 with basic_stdp.cl:4-27's rule and constants.  A neuron of a
 * population with a plastic connection keeps a post trace (stdp_traces.stdp,
 * from 0) and ten 16-bit inter-spike entries in 4 us units (stdp_group.cl:
 * 4-16, from 0).  Per input at t (in the order above), with
 * last = t_last - 4 hist[0] (uint):
 *   if the synapse is plastic (update_synapse, stdp.cl:8-30, in its types):
 *     post = (int)last; for h = 1..9 while post > (int)prev and hist[h] > 0:
 *       pre = e^((float)(uint)(prev - post) / 20000) * pre_trace + 1,
 *       w += 0.01 (1 - w) e^(-pre / 100) (in double, rounded to float),
 *       post -= 4 hist[h];
 *     trace = e^(-(float)(uint)(t - last) / 20000) trace when that float is > 0;
 *     w -= (double)2.02f * 0.01 * w * e^(-trace / 100) (double, to float);
 *     w = max(min(w, 0.0003), -0.0003); the synapse keeps w
 *   the neuron step above with w;
 *   hist[0] = (uint16)((t - last) / 4);
 *   on a spike: hist[h] = hist[h - 1] for h = 9..1, hist[0] = 0, trace += 1,
 *   and the spike carries prev = last and its trace, which the target side
 *   reads as an integer (stdp.cl:84 holds it in a uint) for pre_trace.
 * prev and pre_trace belong to the input spike.  The exponentials are
 * orc_expf's.  (For times below 2^31 us the unsigned difference of a later
 * post time wraps, e^(...) overflows to +inf and the LTP increment is 0.) */
#define EV_HIST 10u
#define EV_HIST_UNIT 4u

typedef struct {
    uint32_t src, time;
    uint32_t prev; /* plastic source: its last spike time before this spike */
    float trace;   /* plastic source: its trace after the spike */
} ev_input;

typedef struct {
    uint32_t tgt, time;
    uint64_t seq;
    float w;
    uint64_t r; /* synapse */
    uint32_t prev;
    float trace;
} ev_event;

/* e^x in single precision with IEEE single operations only: x = k ln2 + r,
 * k = floor(x log2(e) + 1/2), ln2 as the 9-bit head 0.693359375 plus its tail
 * (Cody and Waite), e^r by its degree-7 Taylor polynomial (Horner), scaled by
 * 2^k (ldexpf).  NaN propagates, x > 88.72283935546875 gives +inf, x <
 * -103.972084045410156 gives 0.  Within 2 ulp of e^x. */
float orc_expf(float x)
{
    if (x != x)
        return x;
    if (x > 88.72283935546875f)
        return INFINITY;
    if (x < -103.972084045410156f)
        return 0.0f;
    const float kf = floorf(x * 1.44269502162933349609375f + 0.5f);
    const float r = (x - kf * 0.693359375f) - kf * -2.12194440e-4f;
    float p = 1.0f / 5040.0f;
    p = p * r + 1.0f / 720.0f;
    p = p * r + 1.0f / 120.0f;
    p = p * r + 1.0f / 24.0f;
    p = p * r + 1.0f / 6.0f;
    p = p * r + 0.5f;
    p = p * r + 1.0f;
    p = p * r + 1.0f;
    return ldexpf(p, (int)kf);
}

/* update_synapse (stdp.cl:8-30) of one plastic input */
static float ev_stdp_update(float w, const uint16_t* hist, uint32_t last, uint32_t t, uint32_t prev,
                            float pre_trace, float* trace)
{
    int32_t post = (int32_t)last;
    for (uint32_t h = 1; h < EV_HIST && post > (int32_t)prev && hist[h] > 0; ++h) {
        const float pre = orc_expf((float)(prev - (uint32_t)post) / 20000.0f) * pre_trace + 1.0f;
        const double inc = 0.01 * (double)(1.0f - w) * (double)orc_expf(-pre / 100.0f);
        w = (float)((double)w + inc);
        post -= (int32_t)(hist[h] * EV_HIST_UNIT);
    }
    const float dt = (float)(t - last);
    *trace = dt > 0.0f ? orc_expf(-dt / 20000.0f) * *trace : *trace;
    const double dec = (double)2.02f * 0.01 * (double)w * (double)orc_expf(-*trace / 100.0f);
    w = (float)((double)w - dec);
    return fmaxf(fminf(w, 0.0003f), -0.0003f);
}

static int cmp_ev_input(const void* a, const void* b)
{
    const ev_input* x = (const ev_input*)a;
    const ev_input* y = (const ev_input*)b;
    if (x->src != y->src)
        return x->src < y->src ? -1 : 1;
    return x->time < y->time ? -1 : (x->time > y->time);
}

static int cmp_ev_event(const void* a, const void* b)
{
    const ev_event* x = (const ev_event*)a;
    const ev_event* y = (const ev_event*)b;
    if (x->tgt != y->tgt)
        return x->tgt < y->tgt ? -1 : 1;
    if (x->time != y->time)
        return x->time < y->time ? -1 : 1;
    return x->seq < y->seq ? -1 : (x->seq > y->seq);
}

/* the decay tables of one membrane time constant (us); *n_hi entries of H */
void orc_ev_tables(double tau_us, double* lo, double* hi, uint32_t cap_hi, uint32_t* n_hi)
{
    for (uint32_t l = 0; l < EV_LO; ++l)
        lo[l] = exp(-(double)l / tau_us);
    uint32_t h = 0;
    for (; h < cap_hi; ++h) {
        hi[h] = exp(-(double)(h * EV_LO) / tau_us);
        if (hi[h] == 0.0)
            break;
    }
    *n_hi = h;
}

typedef struct {
    float* V;
    uint32_t* t_last;
    float* trace;      /* plastic neurons: post trace */
    uint16_t* hist;    /* [N][EV_HIST] */
    uint8_t* plastic;  /* per population: has a plastic connection */
    ev_input* in;
    uint64_t n_in, cap_in;
    double* lo; /* EV_LO per population */
    double* hi; /* 8192 per population */
    uint32_t* n_hi;
    uint32_t epoch;
} ev_state;

#define EV_HI_CAP 8192u

void orc_ev_free(void* p)
{
    ev_state* e = (ev_state*)p;
    if (!e)
        return;
    free(e->V);
    free(e->t_last);
    free(e->trace);
    free(e->hist);
    free(e->plastic);
    free(e->in);
    free(e->lo);
    free(e->hi);
    free(e->n_hi);
    free(e);
}

static void ev_push(ev_state* e, uint32_t src, uint32_t time, uint32_t prev, float trace)
{
    if (e->n_in == e->cap_in) {
        e->cap_in = e->cap_in ? 2 * e->cap_in : 1024;
        e->in = (ev_input*)realloc(e->in, e->cap_in * sizeof(ev_input));
    }
    e->in[e->n_in].src = src;
    e->in[e->n_in].time = time;
    e->in[e->n_in].prev = prev;
    e->in[e->n_in].trace = trace;
    e->n_in++;
}

static void ev_record(orc_net* n, uint32_t epoch, uint32_t j, uint32_t time)
{
    n->spikes++;
    if (n->rec_n < n->rec_cap) {
        naer_spike* r = &n->rec[n->rec_n++];
        r->step = epoch;
        r->neuron = j;
        r->time_us = time;
        r->_pad = 0;
    }
}

/* Runs `epochs` windows of the event-driven mode on a built net (its
 * connectivity and initial potentials); the net must not also be stepped by
 * orc_net_run. */
int orc_ev_run(orc_net* n, void** state, uint32_t epochs)
{
    if (!n->built) {
        set_err("net not built");
        return -1;
    }
    ev_state* e = (ev_state*)*state;
    if (!e) {
        e = (ev_state*)xcalloc(1, sizeof(ev_state));
        e->V = (float*)xcalloc(n->N, sizeof(float));
        e->t_last = (uint32_t*)xcalloc(n->N, sizeof(uint32_t));
        e->trace = (float*)xcalloc(n->N, sizeof(float));
        e->hist = (uint16_t*)xcalloc((size_t)n->N * EV_HIST, sizeof(uint16_t));
        e->plastic = (uint8_t*)xcalloc(n->desc.n_pops ? n->desc.n_pops : 1, 1);
        for (uint32_t c = 0; c < n->desc.n_conns; ++c)
            if (n->conns[c].plastic)
                e->plastic[n->conns[c].src_pop] = e->plastic[n->conns[c].dst_pop] = 1;
        memcpy(e->V, n->V, n->N * sizeof(float));
        const uint32_t np = n->desc.n_pops;
        e->lo = (double*)xcalloc((size_t)np * EV_LO, sizeof(double));
        e->hi = (double*)xcalloc((size_t)np * EV_HI_CAP, sizeof(double));
        e->n_hi = (uint32_t*)xcalloc(np, sizeof(uint32_t));
        for (uint32_t p = 0; p < np; ++p)
            orc_ev_tables((double)n->pops[p].tau_m_ms * 1000.0, e->lo + (size_t)p * EV_LO,
                          e->hi + (size_t)p * EV_HI_CAP, EV_HI_CAP, &e->n_hi[p]);
        *state = e;
    }
    uint32_t times[16];
    for (uint32_t k = 0; k < epochs; ++k, ++e->epoch) {
        const uint32_t t0 = e->epoch * EV_WINDOW, t1 = t0 + EV_WINDOW;
        /* POISSON spikes of the window */
        for (uint32_t j = 0; j < n->N; ++j) {
            const naer_pop_desc* q = &n->pops[n->pop_of[j]];
            if (q->kind != NAER_POP_POISSON)
                continue;
            int m = orc_ref_poisson_epoch(n->pois_idx[j], n->n_pois, n->desc.poisson_seed, q->src_rate_hz, t0, t1,
                                          times);
            for (int i = 0; i < m; ++i) {
                ev_push(e, j, times[i], 0, 0.0f);
                ev_record(n, e->epoch, j, times[i]);
            }
        }
        qsort(e->in, e->n_in, sizeof(ev_input), cmp_ev_input);
        uint64_t n_ev = 0;
        for (uint64_t i = 0; i < e->n_in; ++i)
            n_ev += n->row_ptr[e->in[i].src + 1] - n->row_ptr[e->in[i].src];
        ev_event* ev = (ev_event*)xcalloc(n_ev ? n_ev : 1, sizeof(ev_event));
        uint64_t c = 0;
        for (uint64_t i = 0; i < e->n_in; ++i) {
            const uint32_t s = e->in[i].src;
            for (uint64_t r = n->row_ptr[s]; r < n->row_ptr[s + 1]; ++r, ++c) {
                ev[c].tgt = n->col[r] & 0xFFFFFFu;
                ev[c].time = e->in[i].time;
                ev[c].seq = c;
                ev[c].w = n->w[r];
                ev[c].r = r;
                ev[c].prev = e->in[i].prev;
                ev[c].trace = e->in[i].trace;
            }
        }
        qsort(ev, n_ev, sizeof(ev_event), cmp_ev_event);
        n->events += n_ev;
        e->n_in = 0;
        for (uint64_t i = 0; i < n_ev; ++i) {
            const uint32_t j = ev[i].tgt, t = ev[i].time;
            const uint32_t p = n->pop_of[j];
            const naer_pop_desc* q = &n->pops[p];
            const int pl = e->plastic[p];
            uint16_t* hist = e->hist + (size_t)j * EV_HIST;
            float w = ev[i].w;
            uint32_t last = 0;
            if (pl) {
                last = e->t_last[j] - (uint32_t)hist[0] * EV_HIST_UNIT;
                if (n->col[ev[i].r] & 0x80000000u) {
                    /* stdp.cl:84: the pre trace passes through a uint */
                    const float pre_trace = (float)(uint32_t)ev[i].trace;
                    w = ev_stdp_update(w, hist, last, t, ev[i].prev, pre_trace, &e->trace[j]);
                    n->w[ev[i].r] = w;
                }
            }
            int spiked = 0;
            const int64_t dt = (int64_t)t - (int64_t)e->t_last[j];
            if (dt >= 0) {
                const uint64_t h = (uint64_t)dt >> 12;
                const double d = h < e->n_hi[p]
                                     ? e->hi[(size_t)p * EV_HI_CAP + h] * e->lo[(size_t)p * EV_LO + (dt & 4095)]
                                     : 0.0;
                float x = e->V[j] - q->v_rest;
                x = (float)((double)x * d);
                float v = x + q->v_rest;
                v = v + w;
                e->t_last[j] = t;
                if (v > q->v_th) {
                    v = q->v_reset;
                    e->t_last[j] = t + (uint32_t)llround((double)q->t_ref_ms * 1000.0);
                    spiked = 1;
                }
                e->V[j] = v;
            }
            if (pl) {
                hist[0] = (uint16_t)((t - last) / EV_HIST_UNIT);
                if (spiked) {
                    for (uint32_t h = EV_HIST - 1; h > 0; --h)
                        hist[h] = hist[h - 1];
                    hist[0] = 0;
                    e->trace[j] += 1.0f;
                }
            }
            if (spiked) {
                ev_push(e, j, t + EV_WINDOW, last, pl ? e->trace[j] : 0.0f);
                ev_record(n, e->epoch, j, t);
            }
        }
        free(ev);
    }
    return 0;
}

/* plastic neurons' post traces and inter-spike histories ([N][10]) */
void orc_ev_read_stdp(void* state, uint32_t n_neurons, float* trace, uint16_t* hist)
{
    ev_state* e = (ev_state*)state;
    if (trace)
        memcpy(trace, e->trace, n_neurons * sizeof(float));
    if (hist)
        memcpy(hist, e->hist, (size_t)n_neurons * EV_HIST * sizeof(uint16_t));
}

void orc_ev_read_state(void* state, uint32_t n_neurons, float* v, uint32_t* t_last)
{
    ev_state* e = (ev_state*)state;
    if (v)
        memcpy(v, e->V, n_neurons * sizeof(float));
    if (t_last)
        memcpy(t_last, e->t_last, n_neurons * sizeof(uint32_t));
}
