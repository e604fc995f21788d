/*
 * naer_oracle.h — CPU oracle of the simulation step (TEST INFRASTRUCTURE).
 *
 * This library is the checker the HIP implementation is compared against.  It
 * is not part of the product: only tests/, the smoke check of
 * __graft_entry__.py and the cpu_baseline leg of bench.py load it.
 *
 * It restates, in plain C and independently of the HIP sources:
 *   - Philox4x32-10 (Salmon et al., SC'11), used by the reference through the
 *     Random123 submodule (random.cl:4,8-16), pinned by the Random123 known
 *     answer vectors;
 *   - the reference's uniform_rand / exprnd maps (random.cl:6-44);
 *   - the reference's Poisson source trains (input_neuron_group.cl:6-45);
 *   - the reference's pairwise-Bernoulli connectivity (network.cl:14-107);
 *   - the clock-driven LIF / delay-ring / STDP step documented in
 *     include/naer/naer.h, in three arithmetic variants:
 *       ORC_FIXED  fp32 state, 32.32 fixed-point ring (bit-exact target of
 *                  NAER_MODE_DETERMINISTIC), plus an fp64 shadow of every
 *                  ring sum for relative-error checks;
 *       ORC_FLOAT  fp32 state, fp32 ring summed in canonical event order;
 *       ORC_DOUBLE fp64 state and ring throughout.
 */
#ifndef NAER_ORACLE_H
#define NAER_ORACLE_H

#include <stdint.h>

#include "../include/naer/naer.h"

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_FIXED 0
#define ORC_FLOAT 1
#define ORC_DOUBLE 2

void orc_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void orc_philox_n(const uint32_t* ctr, const uint32_t* key, uint32_t* out, uint64_t n);

double orc_ref_uniform(uint32_t gid, uint32_t seed, uint32_t iteration);
uint32_t orc_ref_isi_us(float rate_hz, uint32_t gid, uint32_t seed, uint32_t iteration);
uint32_t orc_ref_poisson_gid(uint32_t idx);
int orc_ref_poisson_epoch(uint32_t idx, uint32_t n_pois, uint32_t seed, float rate_hz,
                          uint32_t t0, uint32_t t1, uint32_t* times);
/* Destinations (relative) connected from relative source src in a Bernoulli
 * connection; returns the count, writes at most cap entries. */
uint32_t orc_ref_bernoulli_row(uint32_t src, uint32_t n_dst, double p, uint32_t seed32,
                               uint32_t* out, uint32_t cap);

/* Planted mistakes of the fixed in-degree rule (orc_set_conn_mutation), for
 * tests that must fail on them: sources drawn over the target population's
 * size, sources offset by the target population's base, Philox words 0 and 1
 * swapped, the delay or the weight drawn from the source's word, random
 * weights spanning [wmin, wmin + wmax). */
enum {
    ORC_MUT_NONE = 0,
    ORC_MUT_SRC_RANGE = 1,
    ORC_MUT_SRC_BASE = 2,
    ORC_MUT_WORDS_SWAPPED = 3,
    ORC_MUT_DELAY_REUSE = 4,
    ORC_MUT_WEIGHT_REUSE = 5,
    ORC_MUT_WEIGHT_SPAN = 6,
};
void orc_set_conn_mutation(uint32_t m);

typedef struct orc_net orc_net;
orc_net* orc_net_create(const naer_net_desc* desc, int precision, int threads);
const char* orc_error(void);
int orc_net_build(orc_net* net);
int orc_net_run(orc_net* net, uint32_t n_steps);
void orc_net_destroy(orc_net* net);

uint64_t orc_net_n_neurons(const orc_net* net);
uint64_t orc_net_n_synapses(const orc_net* net);
uint64_t orc_net_step_count(const orc_net* net);
uint64_t orc_net_spike_count(const orc_net* net);
uint64_t orc_net_event_count(const orc_net* net);
uint32_t orc_net_ring_depth(const orc_net* net);
void orc_net_read_csr(const orc_net* net, uint64_t* row_ptr, uint32_t* col, float* w);
/* fp32 variants report state as float, ORC_DOUBLE rounds its doubles to float
 * here; orc_net_read_state_f64 gives ORC_DOUBLE's exact values. */
void orc_net_read_state(const orc_net* net, float* v, int32_t* ref, float* pre, float* post);
void orc_net_read_state_f64(const orc_net* net, double* v);
/* Spikes recorded since the last read, in emission order. */
uint64_t orc_net_read_spikes(orc_net* net, naer_spike* out, uint64_t cap);
/* Probe: inputs of every neuron for steps [t_begin, t_begin+n): used (as fed to
 * the update), exact (fp64 sum of the same events) and abs (sum of |w|). */
int orc_net_set_probe(orc_net* net, uint64_t t_begin, uint32_t n_steps);
void orc_net_read_probe(const orc_net* net, float* used, double* exact, double* absw);

#ifdef __cplusplus
}
#endif

#endif
