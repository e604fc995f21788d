// naer_model_abi.hpp — argument block of the kernel that steps a population
// of a source-defined neuron model (naer_model_create).  The same text is
// compiled into the library and, as a string, into every model's hipRTC
// program, so both sides agree on the layout.  Plain C types only: the
// hipRTC side has no library headers.
#pragma once

#define NAER_MODEL_ARGS                                                                      \
    struct ModelArgs {                                                                       \
        const float2* cin;                 /* [B][N] (synaptic input, drive) per step */    \
        float* state;                      /* [n_state][end - begin] */                     \
        float* V;                          /* [N] potentials: state[0] is mirrored here */  \
        const unsigned long long* row_ptr; /* [N + 1] outgoing rows */                      \
        uint4* list0;                      /* spike list segments of even batches */        \
        uint4* list1;                      /* and of odd batches */                         \
        unsigned int* seg_count;           /* [2][n_seg] */                                 \
        unsigned long long* seg_ev;        /* [2][n_seg] */                                 \
        unsigned long long* spk_part;      /* [n_seg] spikes per update tile */             \
        unsigned long long* ev_part;       /* [n_seg] events per update tile */             \
        uint4* rec;                        /* record buffer, rec_cap spikes */              \
        unsigned long long* rec_n;         /* spikes offered to the record buffer */        \
        unsigned int* overflow;            /* spikes lost to a full segment */              \
        const unsigned int* cur;           /* parity of the batch in flight */              \
        const unsigned long long* t0_cur;  /* first step of the batch in flight */          \
        const float* params;               /* the population's parameters */                \
        unsigned long long rec_cap;                                                          \
        unsigned long long t_start;        /* pack step of the network's local step 0 */    \
        unsigned int begin, end;           /* neuron range of the population */            \
        unsigned int tile0;                /* tile of `begin` */                            \
        unsigned int seg_cap, n_seg, tw_shift;                                               \
        unsigned int N, B, dt_us, _pad;                                                      \
    }

#define NAER_MODEL_STR_(...) #__VA_ARGS__
#define NAER_MODEL_STR(...) NAER_MODEL_STR_(__VA_ARGS__)
