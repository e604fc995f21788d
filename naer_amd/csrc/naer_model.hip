// naer_model.hip — neuron models given as source and compiled at run time
// with hipRTC: the reference's NeuronType source-string path
// (NeuronType.h:14; the dynamics function its processing kernel calls,
// processing_neuron_group.cl:8-82).
//
// A model's kernel steps one population through one batch after the update
// kernel has staged every neuron's synaptic input and drive per step (cin).
// One block per update tile: spikes go to the tile's list segment of the
// batch, exactly where the update kernel puts LIF spikes, so delivery,
// recording and statistics treat them alike.
#include <hip/hiprtc.h>

#include <cstdio>
#include <string>
#include <vector>

#include "naer/naer.h"
#include "naer_model_abi.hpp"

namespace naer {

namespace {

const char kPrelude[] =
    "typedef unsigned int uint32_t;\n"
    "typedef unsigned long long uint64_t;\n"
    "#define NAER_MODEL_PARAMS 16\n"
    NAER_MODEL_STR(NAER_MODEL_ARGS) ";\n";

const char kKernel[] = R"NAERSRC(
extern "C" __global__ void __launch_bounds__(1024) naer_model_kernel(const ModelArgs a)
{
    __shared__ unsigned int s_spk;
    __shared__ unsigned long long s_ev;
    if (threadIdx.x == 0) {
        s_spk = 0u;
        s_ev = 0ull;
    }
    __syncthreads();
    const unsigned int tile = a.tile0 + blockIdx.x;
    const unsigned int j = (tile << a.tw_shift) + threadIdx.x;
    const unsigned int pl = *a.cur;
    const unsigned long long t0 = *a.t0_cur;
    const unsigned int t0l = (unsigned int)(t0 - a.t_start);
    unsigned int* cnt = a.seg_count + pl * a.n_seg + tile;
    if (j >= a.begin && j < a.end) {
        const unsigned int n = a.end - a.begin, i = j - a.begin;
        float s[NAER_N_STATE];
#pragma unroll
        for (int k = 0; k < NAER_N_STATE; ++k)
            s[k] = a.state[(unsigned long long)k * n + i];
        float p[NAER_MODEL_PARAMS];
#pragma unroll
        for (int k = 0; k < NAER_MODEL_PARAMS; ++k)
            p[k] = a.params[k];
        const unsigned int row = (unsigned int)(a.row_ptr[j + 1] - a.row_ptr[j]);
        uint4* seg = (pl ? a.list1 : a.list0) + (unsigned long long)tile * a.seg_cap;
        unsigned int fired = 0;
        for (unsigned int st = 0; st < a.B; ++st) {
            const float2 c = a.cin[(unsigned long long)st * a.N + j];
            if (naer_model_step(s, c.x, c.y, p, t0l + st)) {
                ++fired;
                const uint4 sp = make_uint4((unsigned int)(t0 + st), j, (t0l + st) * a.dt_us, 0u);
                const unsigned int k = atomicAdd(cnt, 1u);
                if (k < a.seg_cap) {
                    seg[k] = sp;
                    if (a.rec_cap) {
                        const unsigned long long r = atomicAdd(a.rec_n, 1ull);
                        if (r < a.rec_cap)
                            a.rec[r] = sp;
                    }
                }
            }
        }
#pragma unroll
        for (int k = 0; k < NAER_N_STATE; ++k)
            a.state[(unsigned long long)k * n + i] = s[k];
        a.V[j] = s[0];
        if (fired) {
            atomicAdd(&s_spk, fired);
            atomicAdd(&s_ev, (unsigned long long)fired * row);
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        // spikes past the segment's end were not stored: count them lost and
        // leave the segment full
        const unsigned int c = atomicAdd(cnt, 0u);
        if (c > a.seg_cap) {
            atomicAdd(a.overflow, c - a.seg_cap);
            atomicExch(cnt, a.seg_cap);
        }
        if (s_spk) {
            a.seg_ev[pl * a.n_seg + tile] += s_ev;
            a.spk_part[tile] += s_spk;
            a.ev_part[tile] += s_ev;
        }
    }
}
)NAERSRC";

// Models the library ships as source (naer_model_source).
const char kLifSource[] = R"NAERSRC(// LIF as the built-in update computes it.
// s = {V, refractory steps left}; p = {P = exp(-dt/tau_m), v_th, v_reset, v_rest, n_ref}
__device__ bool naer_model_step(float* s, float I, float ext, const float* p, uint32_t t)
{
    if (s[1] > 0.0f) {
        s[1] = s[1] - 1.0f;
        return false;
    }
    const float x = s[0] - p[3];
    const float xp = x * p[0];
    float v = xp + p[3];
    v = v + I;
    v = v + ext;
    const bool f = v > p[1];
    s[0] = f ? p[2] : v;
    if (f)
        s[1] = p[4];
    return f;
}
)NAERSRC";

const char kCubaLifSource[] = R"NAERSRC(// Current-based LIF: a spike of weight w adds w to
// the synaptic current c, which decays with tau_s and charges V by w in total
// (dV/dt = -(V - v_rest)/tau_m + c/tau_s), integrated exactly over each step.
// s = {V, refractory steps left, c};
// p = {P22 = exp(-dt/tau_m), v_th, v_reset, v_rest, n_ref, P11 = exp(-dt/tau_s),
//      P21 = tau_m (P11 - P22) / (tau_s - tau_m)}
__device__ bool naer_model_step(float* s, float I, float ext, const float* p, uint32_t t)
{
    const float c = s[2];
    const float cd = c * p[5];
    s[2] = cd + I;
    if (s[1] > 0.0f) {
        s[1] = s[1] - 1.0f;
        return false;
    }
    const float x = s[0] - p[3];
    const float xp = x * p[0];
    float v = xp + p[3];
    const float q = c * p[6];
    v = v + q;
    v = v + ext;
    const bool f = v > p[1];
    s[0] = f ? p[2] : v;
    if (f)
        s[1] = p[4];
    return f;
}
)NAERSRC";

}  // namespace

const char* model_source(const char* name)
{
    const std::string n = name;
    if (n == "lif")
        return kLifSource;
    if (n == "cuba_lif")
        return kCubaLifSource;
    return nullptr;
}

// Compiles a model's source into a gfx950 code object; on failure `log`
// holds the compiler's messages.
int compile_model(const char* source, uint32_t n_state, std::vector<char>* code, std::string* log)
{
    std::string src = kPrelude;
    src += "#define NAER_N_STATE " + std::to_string(n_state) + "\n";
    src += "#line 1 \"model\"\n";
    src += source;
    src += "\n#line 1 \"naer_model_kernel\"\n";
    src += kKernel;
    hiprtcProgram prog = nullptr;
    if (hiprtcCreateProgram(&prog, src.c_str(), "naer_model.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS) {
        *log = "hiprtcCreateProgram failed";
        return -1;
    }
    // the library's own floating-point flags: every operation rounded on its
    // own, in the order written
    const char* opts[] = { "--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
                           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops" };
    const hiprtcResult r = hiprtcCompileProgram(prog, (int)(sizeof opts / sizeof opts[0]), opts);
    size_t n_log = 0;
    if (hiprtcGetProgramLogSize(prog, &n_log) == HIPRTC_SUCCESS && n_log > 1) {
        std::vector<char> buf(n_log + 1, '\0');
        if (hiprtcGetProgramLog(prog, buf.data()) == HIPRTC_SUCCESS)
            *log = buf.data();
    }
    int rc = -1;
    if (r == HIPRTC_SUCCESS) {
        size_t n_code = 0;
        if (hiprtcGetCodeSize(prog, &n_code) == HIPRTC_SUCCESS && n_code > 0) {
            code->assign(n_code, 0);
            if (hiprtcGetCode(prog, code->data()) == HIPRTC_SUCCESS)
                rc = 0;
        }
        if (rc && log->empty())
            *log = "hiprtc produced no code object";
    } else if (log->empty()) {
        *log = hiprtcGetErrorString(r);
    }
    (void)hiprtcDestroyProgram(&prog);
    return rc;
}

}  // namespace naer
