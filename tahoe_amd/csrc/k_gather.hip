// Element-buffer assembly of the stiffness into the upper MSR values.
//
// The element kernels write the 300 upper entries (m <= n, packed row-major)
// of each element's K_e to ebuf[p][300], p = the element's position in the
// stiffness launch order.  Each MSR value s lists its contributions
// (p * 300 + entry) in launch order; the gather pass sums them in that order
// and writes s once.  The sums are the scatter path's sums in the same order
// (0 + v, then + v per later contribution), so both paths give the same
// values bitwise; the gather writes every value (no clearing pass) and reads
// no old values.
#include <hipcub/hipcub.hpp>

#include "tg_internal.h"
#include "tg_element.h"

namespace tg {

__global__ void k_gather_count(int nel, const int* __restrict__ order, const int* __restrict__ slot,
                               int* __restrict__ cnt)
{
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= 300L * nel) return;
    const int pos = (int)(idx / 300), k = (int)(idx % 300);
    const int s = slot_target(slot[300L * order[pos] + k]);
    if (s >= 0) atomicAdd(&cnt[s], 1);
}

__global__ void k_gather_fill(int nel, const int* __restrict__ order, const int* __restrict__ slot,
                              int* __restrict__ cursor, int* __restrict__ list)
{
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= 300L * nel) return;
    const int pos = (int)(idx / 300), k = (int)(idx % 300);
    const int s = slot_target(slot[300L * order[pos] + k]);
    if (s >= 0) list[atomicAdd(&cursor[s], 1)] = (int)idx;
}

// each value's contributions into launch order (the fill order is arbitrary)
__global__ void k_gather_sort(long nval, const int* __restrict__ ptr, int* __restrict__ list)
{
    const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nval) return;
    const int b = ptr[s], e = ptr[s + 1];
    for (int i = b + 1; i < e; i++) {
        const int x = list[i];
        int j = i - 1;
        while (j >= b && list[j] > x) {
            list[j + 1] = list[j];
            j--;
        }
        list[j + 1] = x;
    }
}

template <class VT>
__global__ void __launch_bounds__(256) k_gather(long nval, const int* __restrict__ ptr, const int* __restrict__ list,
                                                const VT* __restrict__ ebuf, VT* __restrict__ val)
{
    const long s = (long)blockIdx.x * 256 + threadIdx.x;
    if (s >= nval) return;
    const int b = ptr[s], e = ptr[s + 1];
    VT acc = 0;
    for (int j = b; j < e; j++) acc += ebuf[list[j]];
    val[s] = acc;
}

void build_gather_lists(tg_ctx* c, const int* dorder)
{
    const long nval = c->msr_len;
    const long n = 300L * c->nel;
    const unsigned blocks = (unsigned)((n + 255) / 256);
    DBuf<int> cnt;
    cnt.alloc((size_t)nval);
    TG_HIP(hipMemsetAsync(cnt.p, 0, sizeof(int) * nval, c->stream));
    hipLaunchKernelGGL(k_gather_count, dim3(blocks), dim3(256), 0, c->stream, c->nel, dorder, c->dslot.p, cnt.p);
    TG_HIP(hipGetLastError());
    c->dgptr.alloc((size_t)nval + 1);
    TG_HIP(hipMemsetAsync(c->dgptr.p, 0, sizeof(int), c->stream));
    size_t tmp_bytes = 0;
    TG_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, tmp_bytes, cnt.p, c->dgptr.p + 1, (int)nval, c->stream));
    DBuf<char> tmp;
    tmp.alloc(tmp_bytes);
    TG_HIP(hipcub::DeviceScan::InclusiveSum(tmp.p, tmp_bytes, cnt.p, c->dgptr.p + 1, (int)nval, c->stream));
    int total = 0;
    TG_HIP(hipMemcpyAsync(&total, c->dgptr.p + nval, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    TG_HIP(hipStreamSynchronize(c->stream));
    c->dglist.alloc((size_t)std::max(total, 1));
    // the counts become the fill cursors
    TG_HIP(hipMemcpyAsync(cnt.p, c->dgptr.p, sizeof(int) * nval, hipMemcpyDeviceToDevice, c->stream));
    hipLaunchKernelGGL(k_gather_fill, dim3(blocks), dim3(256), 0, c->stream, c->nel, dorder, c->dslot.p, cnt.p,
                       c->dglist.p);
    hipLaunchKernelGGL(k_gather_sort, dim3((unsigned)((nval + 255) / 256)), dim3(256), 0, c->stream, nval,
                       c->dgptr.p, c->dglist.p);
    TG_HIP(hipGetLastError());
    TG_HIP(hipStreamSynchronize(c->stream));  // cnt and tmp outlive their use
}

void launch_gather(tg_ctx* c)
{
    const long nval = c->msr_len;
    const unsigned blocks = (unsigned)((nval + 255) / 256);
    if (mixed_stiffness(c))
        hipLaunchKernelGGL(k_gather<float>, dim3(blocks), dim3(256), 0, c->stream, nval, c->dgptr.p, c->dglist.p,
                           c->debuf32.p, c->dval32.p);
    else
        hipLaunchKernelGGL(k_gather<double>, dim3(blocks), dim3(256), 0, c->stream, nval, c->dgptr.p, c->dglist.p,
                           c->debuf.p, c->dval.p);
    TG_HIP(hipGetLastError());
}

}  // namespace tg
