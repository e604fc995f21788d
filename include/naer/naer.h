/*
 * naer.h — C ABI of the MI355X (gfx950) spiking-network simulation step.
 *
 * The reference has no FFI layer: its simulation step sits behind the C++
 * class Naer::Network (include/naer/Network.h:15-56 of the reference) and the
 * positional OpenCL kernel argument lists bound in Network::prepareKernels
 * (src/Network.cpp:571-652) and Network::step (src/Network.cpp:122-256).
 * This header is the drop-in boundary that replaces those: plain C types,
 * plain pointers and sizes, no HIP or C++ types in any signature.  The C++
 * Naer::Network in Network.h is implemented on top of it.
 *
 * Conventions
 *   - every function returns NAER_OK (0) or a negative NAER_ERR_* code; the
 *     message of the last failure on the calling thread is naer_last_error().
 *   - a naer_net is used by one host thread at a time.  Work is enqueued on the
 *     net's own HIP stream; naer_net_step() is asynchronous, naer_net_sync()
 *     is the synchronisation point (reference Network::finish,
 *     src/Network.cpp:258-265).
 *   - the library owns all device memory of a net and frees it in
 *     naer_net_destroy() (the reference never releases its buffers,
 *     src/Network.cpp:49-75).
 *
 * Simulation scheme (clock-driven, fixed step dt)
 *   step t covers [t*dt, (t+1)*dt).  Per step:
 *     1. LIF update of every neuron: read and clear the delay-ring slot t,
 *        add the counter-based Poisson drive, exact propagator, refractory
 *        countdown, threshold and reset; fired neurons are compacted into
 *        the step's spike list.  POISSON source neurons emit the reference's
 *        exponential-ISI trains instead (input_neuron_group.cl:6-45).
 *     2. spike scatter: the CSR row of every spiking source is streamed and
 *        each synapse (target, delay d, weight w) adds w into ring slot t+d of
 *        its target.  NAER_MODE_FAST accumulates with fp32 hardware atomics;
 *        NAER_MODE_DETERMINISTIC accumulates the weights as 32.32 fixed-point
 *        int64 values (w rounded once to a multiple of 2^-32 mV), whose sum is
 *        exact and therefore independent of event order.
 *     3. STDP (plastic synapses only): pair-based additive STDP with per-neuron
 *        pre/post traces; depression at pre-synaptic spikes (the depressed
 *        weight is the one delivered), potentiation at post-synaptic spikes.
 *        The device applies a synapse's potentiations when its pre-synaptic
 *        neuron next fires (update_synapse, stdp.cl:19-26), in the same order
 *        as the eager statement, or at a flush of all plastic synapses; the
 *        results are those of eager pair STDP.
 *
 * Synapse storage: outgoing CSR.  row_ptr is uint64[N+1]; each synapse has a
 * uint32 word  target | delay_steps << 24 | plastic << 31  and an fp32 weight.
 * Rows are sorted ascending by (g, word, weight bits), g = the target's index
 * within its network divided by 2^14, so the layout is a pure function of the
 * network description and seeds (not of its place in a pack or arena).
 * Limits: N <= 2^24 neurons per net, delays 1..127 steps.
 */
#ifndef NAER_H
#define NAER_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes --------------------------------------------------------- */
#define NAER_OK 0
#define NAER_ERR_INVALID (-1)  /* bad argument / description                 */
#define NAER_ERR_HIP (-2)      /* HIP runtime failure                          */
#define NAER_ERR_NOMEM (-3)    /* device allocation failed                     */
#define NAER_ERR_CAPACITY (-4) /* a documented size limit was exceeded         */
#define NAER_ERR_STATE (-5)    /* call not valid in the object's current state */
#define NAER_ERR_NODEVICE (-6) /* no HIP device / device index out of range    */

/* ---- random streams --------------------------------------------------------
 * Every random quantity is a pure function of (seed, index), drawn from
 * Philox4x32-10 with key = {low 32 bits, high 32 bits} of the seed.  j is the
 * neuron index within its network and t the network's own step count (both
 * independent of how networks are packed, see naer_pack_create):
 *   initial potential of neuron j:        ctr = {j, 0, NAER_STREAM_V0, 0}, word 0
 *   Poisson drive of neuron j at step t:   ctr = {j, t >> 2, NAER_STREAM_DRIVE, 0},
 *                                          word t & 3
 *   draw k of target j, FIXED_INDEGREE connection c (connection seed; here
 *   j is the target's index within the target population and c the
 *   connection's position in the description):
 *                                          ctr = {j, k, NAER_STREAM_CONN, c};
 *                                          word 0 -> source, 1 -> delay, 2 -> weight
 * A uniform u in [0,1) is (word >> 8) * 2^-24 in fp32; an index in [0,n) is
 * (word * n) >> 32.  BERNOULLI connections use the reference's own map
 * instead (random.cl:6-24, network.cl:34-41): key = {src*32 + dst%32, seed32},
 * ctr = {dst,0,0,0}, connected iff word 2 / 4294967295.0 < p; word 1 draws the
 * delay and word 3 the random weight. */
#define NAER_STREAM_V0 1u
#define NAER_STREAM_DRIVE 2u
#define NAER_STREAM_CONN 3u

/* The drive count k of one step is found by inversion of a per-population
 * CDF table: c_0 = e^-lambda, c_k = c_{k-1} + e^-lambda lambda^k / k! summed
 * in fp64 and rounded to fp32; the first entry that rounds to >= 1.0f (or entry
 * NAER_DRIVE_TABLE-1) is replaced by 2.0f; k = min{k : u < cdf[k]}.
 * lambda = ext_rate_hz * dt_ms * 1e-3 must be <= 20. */
#define NAER_DRIVE_TABLE 64

/* ---- description enums ---------------------------------------------------- */
/* Population kinds.  The reference's NeuronType::NOP and NeuronType::STDP are
 * both leaky integrate-and-fire neurons (nop.cl, stdp.cl); plasticity is a
 * property of connections here.  NeuronType::POISSON is a spike source. */
#define NAER_POP_LIF 0u
#define NAER_POP_POISSON 1u
/* A neuron model given as source (naer_model_create): the reference's
 * NeuronType plugin path, where a model's dynamics come as source strings
 * compiled at run time (NeuronType.h:14, processing_neuron_group.cl:8-82). */
#define NAER_POP_MODEL 2u

/* Most state floats per neuron and parameter floats per population of a
 * source-defined model. */
#define NAER_MODEL_STATE 8
#define NAER_MODEL_PARAMS 16

typedef struct naer_model naer_model;

/* Connection rules. */
#define NAER_RULE_BERNOULLI 0u      /* pairwise Bernoulli(p), network.cl:14-107 */
#define NAER_RULE_FIXED_INDEGREE 1u /* K sources per target, drawn uniformly     */

/* Accumulation modes of the spike scatter. */
#define NAER_MODE_FAST 0u
#define NAER_MODE_DETERMINISTIC 1u
/* The reference's event-driven step at microsecond resolution
 * (src/Network.cpp:122-256, lif.cl:15-34): naer_net_step advances whole
 * 1500 us windows (the reference's step()).  A window's inputs are the
 * POISSON spikes timed in it and the LIF spikes of the previous window,
 * delivered 1500 us after firing (NAER_AXONAL_DELAY; connection delays are
 * not used).  Each target takes its inputs in time order (equal times: by
 * source, then row position): dt = t - t_last; dt < 0 (refractory) drops
 * the input; else V = v_rest + (V - v_rest) e^(-dt/tau_m) (product in double),
 * V += w, t_last = t; V > v_th fires: V = v_reset, t_last = t + t_ref.
 * e^(-dt/tau_m) comes from two tables, exp(-4096 h / tau) * exp(-l / tau)
 * with dt = 4096 h + l.  Populations: LIF without drive, POISSON; no models,
 * no graphs.  Recorded spikes carry step = the window and time_us = the
 * firing time.
 * Plastic connections follow stdp.cl:8-140 with basic_stdp.cl's rule and
 * constants (T = 20 ms, alpha = 2.02, lambda = 0.01, bounds +-0.3 mV (in V),
 * the 100 us constant, ten 4 us inter-spike entries); naer_stdp_desc is not
 * used.  A neuron of a population with a plastic connection keeps a post
 * trace and the inter-spike history; per input at t, last = t_last -
 * 4 hist[0]: a plastic input first updates its synapse (LTP over the history
 * in the reference's integer types, the trace decayed from last, LTD, hard
 * bounds; the CSR weight keeps the result), then the neuron steps with that
 * weight; hist[0] = (t - last) / 4 in 16 bits; a spike shifts the history,
 * adds 1 to the trace and carries {last, trace} to its targets.  The
 * exponentials are one single-precision e^x built from IEEE operations
 * (identical on host and device).  The source population of a plastic
 * connection needs t_ref >= 1.5 ms (it fires at most once per window, as
 * with the reference's 2 ms), and a plastic net fewer than 2^31 synapses. */
#define NAER_MODE_EVENT 2u

/* ---- network description -------------------------------------------------- */
typedef struct naer_pop_desc {
    uint32_t kind; /* NAER_POP_LIF or NAER_POP_POISSON */
    uint32_t size; /* number of neurons                 */
    /* LIF parameters (ignored for POISSON) */
    float tau_m_ms;   /* membrane time constant                     */
    float v_th;       /* threshold (mV); a neuron fires when V > v_th */
    float v_reset;    /* reset potential (mV)                       */
    float v_rest;     /* resting potential E_L (mV)                 */
    float t_ref_ms;   /* absolute refractory period                 */
    float v0_min;     /* initial V drawn uniformly in [v0_min, v0_max) */
    float v0_max;
    float ext_rate_hz; /* independent Poisson drive per neuron (0 = none) */
    float ext_weight;  /* jump per drive event (mV)                        */
    /* POISSON source parameter: rate of the exponential-ISI train (Hz);
     * the reference hard-codes 20 Hz (input_neuron_group.cl:4). */
    float src_rate_hz;
    /* NAER_POP_MODEL: the compiled model (it must live until naer_net_build
     * returns) and the population's parameters, passed to every step.
     * v0_min/v0_max, ext_rate_hz and ext_weight keep their LIF meaning. */
    const naer_model* model;
    float params[NAER_MODEL_PARAMS];
} naer_pop_desc;

typedef struct naer_conn_desc {
    uint32_t src_pop; /* index into naer_net_desc.pops */
    uint32_t dst_pop;
    uint32_t rule;     /* NAER_RULE_*                                        */
    uint32_t indegree; /* FIXED_INDEGREE: sources per target                 */
    double probability; /* BERNOULLI: connection probability                 */
    /* weights: `weight` for every synapse, or, when weight == 0.0f (the
     * reference's convention, network.cl:97-100), uniform in [wmin, wmax). */
    float weight;
    float wmin;
    float wmax;
    /* delays in steps: delay_min + m * delay_quantum, m uniform in
     * {0, ..., (delay_max - delay_min) / delay_quantum}. */
    uint32_t delay_min;
    uint32_t delay_max;
    uint32_t delay_quantum;
    uint32_t plastic; /* 1: STDP on this connection (source and target LIF) */
    uint32_t _pad;
    /* The reference truncates the 64-bit seed to 32 bits
     * (src/Network.cpp:447,532); BERNOULLI uses the low 32 bits exactly as
     * the reference does, FIXED_INDEGREE uses all 64. */
    uint64_t seed;
} naer_conn_desc;

typedef struct naer_stdp_desc {
    float tau_plus_ms;  /* pre-synaptic trace time constant  */
    float tau_minus_ms; /* post-synaptic trace time constant */
    float a_plus;       /* potentiation step (mV)            */
    float a_minus;      /* depression step (mV)              */
    float w_min;        /* hard bounds (mV)                  */
    float w_max;
} naer_stdp_desc;

typedef struct naer_net_desc {
    const naer_pop_desc* pops;
    uint32_t n_pops;
    uint32_t n_conns;
    const naer_conn_desc* conns;
    naer_stdp_desc stdp;
    float dt_ms;   /* step; dt_ms*1000 must be an integer number of us     */
    uint32_t mode; /* NAER_MODE_FAST or NAER_MODE_DETERMINISTIC           */
    uint64_t seed; /* initial potentials and Poisson drive                */
    /* seed of the POISSON source trains; the reference passes
     * (uint32_t)0xbf785e3215ac7ebd (src/Network.cpp:124). */
    uint32_t poisson_seed;
    /* spikes kept on the device for naer_net_read_spikes (0 = no recording) */
    uint32_t record_capacity;
} naer_net_desc;

/* One recorded spike.  For POISSON sources `time_us` is the exact train time
 * (integer microseconds, input_neuron_group.cl:29-37); for LIF neurons it is
 * the start of the step in which the neuron fired. */
typedef struct naer_spike {
    uint32_t step;
    uint32_t neuron;
    uint32_t time_us;
    uint32_t _pad;
} naer_spike;

typedef struct naer_stats {
    uint64_t n_neurons;
    uint64_t n_synapses;
    uint64_t n_plastic;
    uint64_t step;       /* steps simulated so far                  */
    uint64_t time_us;    /* simulated time (reference getTime())    */
    uint64_t spikes;     /* spikes emitted (recording independent)  */
    uint64_t events;     /* synaptic events delivered (sum of out-degrees of emitted spikes) */
    uint64_t dropped;    /* spikes lost to a full record buffer     */
    uint32_t ring_depth; /* delay-ring slots                        */
    uint32_t batch;      /* steps per update launch                 */
    uint32_t max_row;    /* longest CSR row                         */
    uint32_t kernels;    /* number of entries in kernel_ms/kernel_calls */
    /* per-kernel device time accumulated while profiling is on (ms) and
     * launch counts; index with NAER_K_* */
    double kernel_ms[8];
    uint64_t kernel_calls[8];
    uint64_t device_bytes; /* device memory held by the net */
} naer_stats;

#define NAER_K_UPDATE 0
#define NAER_K_SCATTER 1       /* spike delivery of a batch, all its kernels */
#define NAER_K_POTENTIATE 2     /* plastic nets: flush of pending potentiation */
/* the parts of NAER_K_SCATTER when a batch is delivered through buckets */
#define NAER_K_BUCKET_HIST 3
#define NAER_K_BUCKET_SCAN 4
#define NAER_K_BUCKET_PLACE 5
#define NAER_K_BUCKET_REFINE 6
#define NAER_K_MODEL 7          /* kernels of source-defined neuron models */

/* ---- context ---------------------------------------------------------------
 * Replaces the OpenCL platform/device/context/queue set-up of
 * Network::Network (src/Network.cpp:14-47). */
typedef struct naer_ctx naer_ctx;
int naer_ctx_create(int device, naer_ctx** out);
int naer_ctx_destroy(naer_ctx* ctx);

/* ---- neuron models from source ----------------------------------------------
 * Replaces the reference's NeuronType source strings (NeuronType.h:14), whose
 * neuron_dynamics function a processing kernel calls per input
 * (processing_neuron_group.cl:8-82).  `source` is HIP device code defining
 *
 *   __device__ bool naer_model_step(float* s, float I, float ext,
 *                                   const float* p, uint32_t t);
 *
 * called once per neuron and step, in step order, with
 *   s    the neuron's n_state floats (1..NAER_MODEL_STATE).  s[0] is the
 *        membrane potential: drawn uniformly in [v0_min, v0_max) at build
 *        exactly as for a LIF neuron, and returned by naer_net_read_state;
 *        the other entries start at 0;
 *   I    the step's synaptic input (the weights arriving at that step, summed
 *        as for LIF neurons: in 32.32 fixed point in deterministic mode);
 *   ext  the step's Poisson drive k * ext_weight, drawn every step from the
 *        same counter-based stream a LIF neuron uses (a LIF neuron skips the
 *        draws of its refractory steps);
 *   p    the population's params (NAER_MODEL_PARAMS floats);
 *   t    the network-local step.
 * It returns true when the neuron fires at step t; the spike is delivered,
 * recorded and counted like a LIF spike.  Refractoriness, if any, is part of
 * the model's state.  The source is compiled with hipRTC for gfx950 with the
 * library's floating-point flags (no contraction into FMAs, no fast math).
 * naer_model_create needs no device; a source that does not compile gives
 * NAER_ERR_INVALID with the compiler's messages in naer_last_error().
 * Model populations may be targets and sources of static connections; they
 * are not plastic, are not admitted into arenas, and nets holding them are
 * not checkpointed (naer_net_save/naer_net_load give NAER_ERR_INVALID). */
int naer_model_create(const char* source, uint32_t n_state, naer_model** out);
int naer_model_destroy(naer_model* m);

/* Source text (static storage) of a model the library ships:
 *   "lif"       the built-in LIF step; n_state 2, s = {V, refractory steps
 *               left}, p = {P = exp(-dt/tau_m), v_th, v_reset, v_rest, n_ref};
 *   "cuba_lif"  current-based LIF: a spike of weight w adds w to a synaptic
 *               current c that decays with tau_s and charges V by w in total;
 *               n_state 3, s = {V, refractory steps left, c}, p = {exp(-dt/
 *               tau_m), v_th, v_reset, v_rest, n_ref, exp(-dt/tau_s),
 *               tau_m (p[5] - p[0]) / (tau_s - tau_m)}.
 * The reference names a CubaLIF neuron type for its vogels model
 * (src/main.cpp:52-69) without defining it. */
int naer_model_source(const char* name, const char** out);

/* ---- network ---------------------------------------------------------------
 * naer_net_create copies the description (pops/conns arrays may be freed
 * afterwards).  It replaces addPopulation/addConnection bookkeeping
 * (src/Network.cpp:76-104). */
typedef struct naer_net naer_net;
int naer_net_create(naer_ctx* ctx, const naer_net_desc* desc, naer_net** out);

/* A pack: n_nets independent networks simulated by the same launches (the
 * batched parameter sweep).  Networks share dt and mode (and the STDP
 * parameters when plastic); each keeps its own seeds, local neuron ids and
 * local clock, so its results do not depend on what it is packed with
 * (bit for bit in NAER_MODE_DETERMINISTIC).  Neurons of network i occupy
 * [neuron_begin[i], neuron_begin[i] + n_neurons[i]) of the pack; recorded
 * spikes carry pack neuron ids and pack steps.  All naer_net_* calls accept
 * a pack; naer_net_create is naer_pack_create with one network.  No
 * counterpart in the reference (one network per process, src/main.cpp). */
int naer_pack_create(naer_ctx* ctx, const naer_net_desc* descs, uint32_t n_nets, naer_net** out);
/* Per network (pack) or slot (arena): first neuron, neuron count (0 for a
 * retired slot) and the pack step at which its clock started.  Writes the
 * first min(cap, slots) entries of each non-NULL array and sets *n_slots
 * (if non-NULL) to the number of slots. */
int naer_pack_layout(naer_net* pack, uint32_t cap, uint32_t* neuron_begin, uint32_t* n_neurons,
                     uint64_t* t_start, uint32_t* n_slots);

/* An arena: preallocated capacity in which networks (serving jobs) are
 * admitted and retired between steps; the form used for continuous batching.
 * Each admitted job is built on the device in free neuron and synapse ranges,
 * starts at the arena's next step with its own clock at 0, and gives the same
 * results as when simulated alone (bit for bit in NAER_MODE_DETERMINISTIC).
 * Jobs must use dt and mode of the arena, delays within [delay_min,
 * delay_max] steps, rows of at most max_row synapses and, when plastic, the
 * arena's STDP parameters and LIF refractory periods of at least
 * refractory_min steps.  No counterpart in the reference. */
typedef struct naer_arena_desc {
    uint64_t max_neurons;  /* neuron capacity (each job is rounded up to 256) */
    uint64_t max_synapses; /* synapse capacity                               */
    uint32_t refractory_min; /* plastic arenas: fewest refractory steps of a job's LIF
                                populations; batches take min(delay_min, refractory_min + 1,
                                16) steps */
    float dt_ms;
    uint32_t mode;
    uint32_t delay_min; /* steps */
    uint32_t delay_max;
    uint32_t max_row;   /* longest CSR row a job may have (<= 8192)        */
    uint32_t plastic;   /* 1: jobs may carry STDP */
    naer_stdp_desc stdp;
    uint32_t record_capacity;
    uint32_t poisson;   /* 1: jobs may contain POISSON sources (larger spike lists) */
} naer_arena_desc;

int naer_arena_create(naer_ctx* ctx, const naer_arena_desc* desc, naer_net** out);
/* NAER_ERR_CAPACITY when the job does not fit now (the caller may retry
 * after retiring jobs); NAER_ERR_INVALID when it never can (more neurons or
 * synapses than the arena holds, rows longer than max_row, settings the
 * arena was not created with).  *slot identifies the job in naer_pack_layout
 * and naer_arena_retire.  Synchronises the arena. */
int naer_arena_admit(naer_net* arena, const naer_net_desc* job, uint32_t* slot);
int naer_arena_retire(naer_net* arena, uint32_t slot);
/* Watchdog (no counterpart in the reference): the slots (networks of a
 * pack, jobs of an arena; 0 for a single net) holding a neuron whose
 * potential is not finite, e.g. a job whose parameters drive its state out of
 * fp32 range.  Writes the first min(cap, *n_bad) slots.  Synchronises the net. */
int naer_net_check_finite(naer_net* net, uint32_t* bad_slots, uint32_t cap, uint32_t* n_bad);

/* Device connectivity build + state initialisation; replaces
 * Network::compile (src/Network.cpp:267-308) and allocateConnections
 * (src/Network.cpp:421-569).  Synchronous. */
int naer_net_build(naer_net* net);

/* Enqueue n_steps simulation steps on the net's stream (asynchronous).
 * Network::step (src/Network.cpp:122-256) advances one 1.5 ms epoch; the C++
 * wrapper calls this with n_steps = 1500 us / dt. */
int naer_net_step(naer_net* net, uint32_t n_steps);

/* Enqueue steps as replays of one captured HIP graph of chunk_steps steps
 * (0 = plain launches, the default).  The kernels read the step counter
 * from device memory, so one graph serves every chunk; steps that do not
 * fill a chunk are launched plainly.  Kernel timers do not cover graph
 * replays.  No counterpart in the reference. */
int naer_net_set_graph(naer_net* net, uint32_t chunk_steps);

/* Wait for all enqueued work (Network::finish, src/Network.cpp:258-265). */
int naer_net_sync(naer_net* net);

/* Whether all enqueued work is done, without waiting: *idle = 1 or 0.  No
 * counterpart in the reference (one net, one loop); a server polls its arenas
 * with it and steps each again as soon as its last chunk is done. */
int naer_net_idle(naer_net* net, int* idle);

/* Counters and timers (getTime/getSynapses/getNeurons and the public
 * *_time fields, include/naer/Network.h:44-55).  Synchronises the net. */
int naer_net_query(naer_net* net, naer_stats* out);

/* Per-kernel HIP-event timing on the net's stream (off by default). */
int naer_net_set_profiling(naer_net* net, int enable);

/* Copy out and drain the recorded spikes (unsorted).  *n_out receives the
 * number written (<= cap).  Spikes not returned (beyond record_capacity or
 * beyond cap) are discarded and counted in naer_stats.dropped.  Synchronises
 * the net. */
int naer_net_read_spikes(naer_net* net, naer_spike* out, uint64_t cap, uint64_t* n_out);

/* Spike raster files.  No counterpart in the reference (it records nothing);
 * the format:
 *   header (40 bytes, little endian): "NAERRST1", uint32 version (1),
 *   uint32 bytes per record (8), float64 dt_ms, uint64 neurons,
 *   uint64 records (written by naer_net_raster_close);
 *   records: {uint32 step, uint32 neuron}, ordered by (step, neuron).
 * Steps and neuron ids are those of the net (pack steps and pack ids for a
 * pack or an arena).  naer_net_raster_drain appends the spikes recorded since
 * the previous drain (the same spikes naer_net_read_spikes would return)
 * through a pinned host buffer; drains happen between steps, so the file stays
 * ordered.  Each synchronises the net. */
#define NAER_RASTER_HEADER_BYTES 40

/* Schedule of the bucketed delivery (no counterpart in the reference): how
 * the histogram/placement kernels cut the batch's spike rows into wave tasks
 * and blocks, how many blocks refine each group bucket, and how the kernels'
 * LDS counters are incremented.  Any schedule
 * gives the same results.  naer_net_autotune tries candidates on the running
 * simulation (each for `batches` aligned batches of real steps, timed with HIP
 * events, scored per synaptic event) and keeps the fastest. */
typedef struct naer_schedule {
    uint32_t bucketed;           /* out: 1 if the net delivers through buckets */
    uint32_t task_synapses;      /* synapses of a row per wave task: 256 or 512 */
    uint32_t segments_per_block; /* mean spike-list segments per histogram/placement block, 1..4096 */
    uint32_t refine_split;       /* refine blocks per group bucket, 1..64; 0: direct placement,
                                    every event written straight into its tile bucket with no
                                    refine pass, for batches whose tile buckets up to the last
                                    occupied tile number at most 8192 (the default of arenas;
                                    a net whose whole tile range exceeds that refuses 0) */
    uint32_t run_aggregate;      /* bit 0: the histogram and refine counters (keyed by target
                                    tile), bit 1: the placement counters (keyed by target group):
                                    one LDS counter atomic per run of equal keys in a wave instead
                                    of one per synapse */
    uint32_t reserve;            /* 1: reserved placement (single nets): no histogram, column scan
                                    or bucket-start passes; each placement block writes into
                                    regions of the staging buffer sized from the batch's event
                                    count and the net's synapse shares per (arrival batch, group),
                                    events beyond them into spill regions (beyond those, the ring);
                                    refine_split is then the refine parts per group bucket,
                                    1..parts_max.  Refused where the shares are unknown (packs,
                                    arenas, nets whose (delay, group) table exceeds 8192 entries or
                                    with more than 1638 group buckets) */
    uint32_t parts_max;          /* out: largest refine_split of reserved placement (0: unavailable) */
    uint32_t record_bytes;       /* out: bytes per bucketed event: 4 when every population's
                                    outgoing synapses share one constant weight (the record
                                    names the source's population), else 8 (the record carries
                                    the weight) */
    double delivery_us;          /* out: delivery time per batch measured by the autotuner */
} naer_schedule;
int naer_net_get_schedule(naer_net* net, naer_schedule* out);
int naer_net_set_schedule(naer_net* net, const naer_schedule* s);
int naer_net_autotune(naer_net* net, uint32_t batches, naer_schedule* chosen);

/* Checkpoint and resume (no counterpart in the reference).  naer_net_save
 * writes the dynamic state of a built net at a step boundary (potentials,
 * refractory counters, traces, Poisson source state, plastic weights, the
 * delay ring, the in-flight bucketed events, the clock, counters and
 * undrained recorded spikes); naer_net_load restores it into a net built from
 * the same description, which then continues bit for bit as the saved one
 * would have (deterministic mode).  Arenas are not checkpointed.  Both
 * synchronise the net. */
int naer_net_save(naer_net* net, const char* path);
int naer_net_load(naer_net* net, const char* path);
int naer_net_raster_open(naer_net* net, const char* path);
int naer_net_raster_drain(naer_net* net, uint64_t* n_written);
int naer_net_raster_close(naer_net* net, uint64_t* n_total);

/* Copy neuron state.  Any pointer may be NULL.  Arrays have n_neurons
 * entries; POISSON neurons report v = 0 and refractory = 0.  pre/post are the
 * STDP traces the next step sees (zeros when the net has no plastic
 * connection); in the event-driven mode post is stdp.cl's post trace and pre
 * is zero. */
int naer_net_read_state(naer_net* net, float* v, int32_t* refractory, float* pre_trace,
                        float* post_trace);

/* Copy the CSR (after build).  row_ptr has n_neurons+1 entries, col and
 * weight n_synapses entries; any pointer may be NULL.  Plastic weights are
 * read after all pending potentiation has been applied (a flush). */
int naer_net_read_csr(naer_net* net, uint64_t* row_ptr, uint32_t* col, float* weight);

/* Record the synaptic input I_syn (mV, as used by the update) of every neuron
 * for steps [t_begin, t_begin + n_steps) into a device buffer; read it with
 * naer_net_read_probe (n_steps * n_neurons floats, step-major). */
int naer_net_set_probe(naer_net* net, uint64_t t_begin, uint32_t n_steps);
int naer_net_read_probe(naer_net* net, float* out);

int naer_net_destroy(naer_net* net);

/* ---- generators exposed for parity checks ----------------------------------
 * Run on the device with the same device functions the simulation uses. */

/* Philox4x32-10 blocks: out[4*i..4*i+3] = philox(ctr[4*i..], key[2*i..]). */
int naer_philox(naer_ctx* ctx, const uint32_t* ctr, const uint32_t* key, uint32_t* out,
                uint64_t n);

/* The reference's Poisson train of one 1500 us epoch
 * (simulate_poisson_neurons, input_neuron_group.cl:6-45) for POISSON neurons
 * idx in [0, n): times[10*idx + m], count[idx] (<= 10). */
int naer_ref_poisson_epoch(naer_ctx* ctx, uint32_t n, uint32_t n_pois_total, uint32_t seed,
                           float rate_hz, uint32_t t0, uint32_t t1, uint32_t* times,
                           uint32_t* count);

/* naer_ref_poisson_epoch without any device allocation: the same kernel and
 * results, staged through a static device array of 4 MiB per GPU instead of
 * buffers allocated per call.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (tight serving loops, harnesses
 * that forbid hipMalloc).  n may be at most 95,325 (n * 11 words of the
 * array), else NAER_ERR_CAPACITY.  Every call on a GPU shares that GPU's
 * array: calls that overlap on one GPU (from several host threads) race. */
int naer_ref_poisson_epoch_noalloc(naer_ctx* ctx, uint32_t n, uint32_t n_pois_total, uint32_t seed,
                                   float rate_hz, uint32_t t0, uint32_t t1, uint32_t* times,
                                   uint32_t* count);

/* Message of the last failure on the calling thread ("" if none). */
const char* naer_last_error(void);

/* Library version string. */
const char* naer_version(void);

#ifdef __cplusplus
}
#endif

#endif /* NAER_H */
