// naer_expf.hpp — e^x in single precision from IEEE-754 single operations
// only (no fused multiply-add, no library call), so that it gives the same
// bits wherever it is compiled with contraction off.  The event-driven STDP
// (basic_stdp.cl:10-27) takes its exponentials from here.
//
// x = k ln2 + r, k = floor(x log2(e) + 1/2); ln2 is split into a 9-bit head
// (0.693359375) and the tail ln2 - head, so that k * head is exact for every
// k in range (Cody and Waite's reduction); |r| <= 0.347.  e^r is its Taylor
// polynomial of degree 7 (truncation below 6e-9 relative) by Horner's rule,
// and the result is scaled by 2^k with ldexpf (one rounding, also for
// subnormal results).  NaN propagates; x > 88.72283935546875 gives +inf and
// x < -103.972084045410156 gives 0.
#pragma once

#include <hip/hip_runtime.h>

namespace naer {

__device__ __forceinline__ float naer_expf(float x)
{
    if (x != x)
        return x;
    if (x > 88.72283935546875f)
        return __builtin_huge_valf();
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

}  // namespace naer
