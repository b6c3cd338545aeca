// Device helpers shared by the element kernels (k_element.hip, k_q1p0.hip).
#pragma once

#include <hip/hip_runtime.h>

namespace tg {

// Sum of 8 lanes' 24 partials, scattered: lane q ends with node q's 3 sums.
__device__ __forceinline__ void reduce_scatter8(const double r[24], int q, double out[3])
{
    double s12[12];
#pragma unroll
    for (int k = 0; k < 12; k++) {
        const bool hi = q & 4;
        const double keep = hi ? r[12 + k] : r[k];
        const double send = hi ? r[k] : r[12 + k];
        s12[k] = keep + __shfl_xor(send, 4);
    }
    double s6[6];
#pragma unroll
    for (int k = 0; k < 6; k++) {
        const bool hi = q & 2;
        const double keep = hi ? s12[6 + k] : s12[k];
        const double send = hi ? s12[k] : s12[6 + k];
        s6[k] = keep + __shfl_xor(send, 2);
    }
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const bool hi = q & 1;
        const double keep = hi ? s6[3 + k] : s6[k];
        const double send = hi ? s6[k] : s6[3 + k];
        out[k] = keep + __shfl_xor(send, 1);
    }
}

// Position of (m, n), m <= n < 24, in the packed upper triangle.
__device__ __forceinline__ int upper_index(int m, int n) { return m * 24 - (m * (m - 1)) / 2 + (n - m); }

// The element slot map holds, per element upper entry, the MSR value index
// s, or -s-2 where that entry is the first contribution to s in the
// stiffness launch order (the values are cleared before the passes, so the
// read of a first contribution would return 0), or -1 for no entry.
__device__ __forceinline__ int slot_target(int s) { return s < -1 ? -s - 2 : s; }

}  // namespace tg
