// naer_kernels.hpp — host launch wrappers of the kernels in kernels.hip.
#pragma once

#include <string>
#include <vector>

#include "naer_device.hpp"

namespace naer {

// extra: the net may contain POISSON sources or source-defined models
// (otherwise their code is left out)
hipError_t launch_update(const UpdateArgs& a, bool det, bool plastic, bool bucket, bool extra, bool rec4,
                         hipStream_t s);
// the bucketed delivery of one aligned batch: steps 0 split + histogram,
// 1 scans, 2 place, 3 refine (reserved placement: 2 and 3 only)
constexpr int kBucketSteps = 4;
hipError_t launch_bucket_step(const BucketArgs& a, int step, bool det, hipStream_t s);
hipError_t launch_bucket_scrub(const UpdateArgs& a, uint32_t j_begin, uint32_t j_end, hipStream_t s);
hipError_t launch_scatter(const ScatterArgs& a, bool det, bool plastic, uint32_t grid_x,
                          uint32_t chunks, hipStream_t s);
hipError_t launch_flush(const FlushArgs& a, uint32_t grid, hipStream_t s);
hipError_t launch_clock_advance(DevState* st, uint32_t b, hipStream_t s);
hipError_t launch_fill_u32(uint32_t* p, unsigned long long n, uint32_t v, hipStream_t s);
hipError_t launch_check_finite(const UpdateArgs& a, uint32_t* flags, uint32_t n_check, hipStream_t s);
hipError_t launch_conn(const ConnArgs& c, uint32_t rule, bool place, uint32_t* deg,
                       unsigned long long* cursor, uint32_t* col, float* w, hipStream_t s);
hipError_t launch_sort_rows(const unsigned long long* row_ptr, uint32_t* col, float* w, uint32_t row_begin,
                            uint32_t row_end, uint32_t max_row, hipStream_t s);
hipError_t launch_count_plastic(const uint32_t* col, unsigned long long begin, unsigned long long end,
                                unsigned long long* count, hipStream_t s);
// Synapses of [0, S) per (delay - dmin, target group of 2^14 neurons): out
// has nd * ng <= kSynStatsBins zeroed entries, to which the counts are added.
constexpr uint32_t kSynStatsBins = 8192;
hipError_t launch_syn_stats(const uint32_t* col, unsigned long long S, uint32_t dmin, uint32_t nd, uint32_t ng,
                            unsigned long long* out, hipStream_t s);
hipError_t launch_init_state(const UpdateArgs& a, bool det, uint32_t j_begin, uint32_t j_end, hipStream_t s);
hipError_t launch_philox(const uint32_t* ctr, const uint32_t* key, uint32_t* out,
                         unsigned long long n, hipStream_t s);
hipError_t launch_ref_poisson_epoch(uint32_t n, uint32_t n_pois, uint32_t seed, double neg_rate,
                                    uint32_t t0, uint32_t t1, uint32_t* times, uint32_t* count,
                                    hipStream_t s);

// Compiles a source-defined neuron model (naer_model.hip) into a gfx950 code
// object whose kernel naer_model_kernel takes one ModelArgs; on failure `log`
// holds the compiler's messages.
int compile_model(const char* source, uint32_t n_state, std::vector<char>* code, std::string* log);
// Source of a model the library ships ("lif", "cuba_lif"), or NULL.
const char* model_source(const char* name);

// Event-driven mode (kernels_event.hip): kernel `which` (kEv*), with n the
// input or event count where the kernel needs one.
hipError_t ev_launch(const EvArgs& a, int which, uint32_t n, hipStream_t s);
// Stable LSD radix sort of n 32-bit keys on bits [lo_bit, hi_bit), with nv
// (0, 1 or 3) arrays of 32-bit values moved along; `space` holds
// ev_sort_space(n) words; the result is in keys_alt/vals_alt when *in_alt,
// else in keys/vals.
uint64_t ev_sort_space(uint32_t n_max);
hipError_t ev_radix_sort(uint32_t* keys, uint32_t* keys_alt, uint32_t* const* vals, uint32_t* const* vals_alt,
                         uint32_t nv, uint32_t n, uint32_t lo_bit, uint32_t hi_bit, uint32_t* space, bool* in_alt,
                         hipStream_t s);
// Exclusive scan of n 64-bit values; `space` holds ev_scan_space(n) entries.
uint64_t ev_scan_space(uint32_t n_max);
hipError_t ev_scan_u64(const unsigned long long* in, unsigned long long* out, uint32_t n, unsigned long long* space,
                       hipStream_t s);

// Longest CSR row the LDS row sort accepts.
constexpr uint32_t kMaxSortedRow = 8192;

}  // namespace naer
