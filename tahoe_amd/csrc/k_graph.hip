// Structure of the equation system built on the device: the node graph
// (element lists per node, then each node's sorted unique neighbours), the
// solver operator's sliced layout and block columns, and the row pointers
// of the upper MSR pattern.  The reference builds the same structure on the
// host in FEManagerT::SetEquationSystem (tahoe/src/main/FEManagerT.cpp:
// 2248-2313) through GraphT::MakeGraph (toolbox/src/graph/GraphT.cpp:376-456)
// and MSRBuilderT::GenerateMSR (toolbox/src/graph/MSRBuilderT.cpp:134-181).
// Every output is a function of the mesh and the equation numbers alone
// (sorted lists, counts, prefix sums), so the atomics used to gather the
// element lists leave no trace in the results.
#include <hipcub/hipcub.hpp>

#include "tg_internal.h"

namespace tg {

namespace {

constexpr int GB = 64;  // threads per workgroup of the neighbour kernels

// incidence k of the connectivity (element k / 8, node conn[k]): count per
// node into cnt[node + 1]
__global__ void k_ne_count(long n8, const int* __restrict__ conn, int* __restrict__ cnt)
{
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n8) atomicAdd(&cnt[conn[k] + 1], 1);
}

// element of incidence k into its node's list (order within a list is
// arbitrary; only sets are read from it)
__global__ void k_ne_fill(long n8, const int* __restrict__ conn, int* __restrict__ fill, int* __restrict__ list)
{
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n8) list[atomicAdd(&fill[conn[k]], 1)] = (int)(k >> 3);
}

// candidate t of a node whose element list starts at p0: node t % 8 of the
// (t / 8)-th element
__device__ __forceinline__ int candidate(const int* conn, const int* list, int p0, int t)
{
    return conn[8L * list[p0 + (t >> 3)] + (t & 7)];
}

// FILL = false: deg[a] = the number of distinct candidates of node a (its
// neighbours, itself included).  FILL = true: those neighbours ascending
// into nlist[nptr[a] ..).  A node of at most 8 elements (64 candidates)
// keeps its candidates in its own LDS column and marks first occurrences in
// a 64-bit mask; a neighbour's place is the number of distinct candidates
// below it.  A node of 9..64 elements reads its candidates from memory and
// inserts each new one into its sorted output range.  More than 64
// elements on a node raise the flag.
template <bool FILL>
__global__ void __launch_bounds__(GB) k_neighbours(int nnd, const int* __restrict__ ne_ptr,
                                                  const int* __restrict__ list, const int* __restrict__ conn,
                                                  long long* __restrict__ deg, const long long* __restrict__ nptr,
                                                  int* __restrict__ nlist, int* __restrict__ flag)
{
    __shared__ int cand[64][GB];
    const int tid = threadIdx.x;
    const int a = blockIdx.x * GB + tid;
    if (a >= nnd) return;  // no barrier below: each thread works in its own column
    const int p0 = ne_ptr[a], ne = ne_ptr[a + 1] - p0;
    if (ne > 64) {
        if (!FILL) {
            atomicOr(flag, 1);
            deg[a] = 0;
        }
        return;
    }
    const int n = 8 * ne;
    if (n <= 64) {
        for (int t = 0; t < n; t++) cand[t][tid] = candidate(conn, list, p0, t);
        unsigned long long first = 0;
        for (int t = 0; t < n; t++) {
            const int x = cand[t][tid];
            bool f = true;
            for (int u = 0; u < t; u++) f &= cand[u][tid] != x;
            if (f) first |= 1ull << t;
        }
        if (!FILL) {
            deg[a] = __popcll(first);
            return;
        }
        const long long o = nptr[a];
        for (int t = 0; t < n; t++) {
            if (!((first >> t) & 1ull)) continue;
            const int x = cand[t][tid];
            int r = 0;
            for (int u = 0; u < n; u++) r += (int)((first >> u) & 1ull) & (cand[u][tid] < x ? 1 : 0);
            nlist[o + r] = x;
        }
        return;
    }
    if (!FILL) {
        int m = 0;
        for (int t = 0; t < n; t++) {
            const int x = candidate(conn, list, p0, t);
            bool f = true;
            for (int u = 0; u < t && f; u++) f = candidate(conn, list, p0, u) != x;
            m += f ? 1 : 0;
        }
        deg[a] = m;
        return;
    }
    const long long o = nptr[a];
    const int cap = (int)(nptr[a + 1] - o);
    int m = 0;
    for (int t = 0; t < n; t++) {
        const int x = candidate(conn, list, p0, t);
        int lo = 0, hi = m;  // first entry >= x
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (nlist[o + mid] < x)
                lo = mid + 1;
            else
                hi = mid;
        }
        if (lo < m && nlist[o + lo] == x) continue;
        if (m >= cap) break;  // the count pass saw the same distinct set: never taken
        for (int v = m; v > lo; v--) nlist[o + v] = nlist[o + v - 1];
        nlist[o + lo] = x;
        m++;
    }
}

// widest row of each slice of 64 nodes
__global__ void k_slice_width(int nnd, int nslices, const long long* __restrict__ deg, int* __restrict__ swidth)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nslices) return;
    long long w = 0;
    for (int a = 64 * s; a < 64 * s + 64 && a < nnd; a++) w = deg[a] > w ? deg[a] : w;
    swidth[s] = (int)w;
}

// block columns: row node a (slice s, lane l) lists its neighbours in its
// first deg[a] slots, -1 after them and in rows past the last node
__global__ void k_block_columns(int nnd, int nslices, const int* __restrict__ sbase, const int* __restrict__ swidth,
                                const long long* __restrict__ nptr, const int* __restrict__ nlist,
                                int* __restrict__ col)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    const int s = a >> 6, l = a & 63;
    if (s >= nslices) return;
    const long long o = a < nnd ? nptr[a] : 0;
    const int d = a < nnd ? (int)(nptr[a + 1] - o) : 0;
    for (int k = 0; k < swidth[s]; k++)
        col[((long)sbase[s] + k) * 64 + l] = k < d ? nlist[o + k] : -1;
}

// MSR row lengths (equation numbers grow with (node, dof)): row (a, i)
// holds the active dofs of a after i and every active dof of each
// neighbour b > a; a node that no element uses has no neighbours, itself
// included, and empty rows.  rl[0] = n_eq + 1 and rl[r + 1] = the length of row r,
// so that the inclusive sum of rl is bindx[0 .. n_eq].
__global__ void k_msr_rowlen(int nnd, int n_eq, const long long* __restrict__ nptr, const int* __restrict__ nlist,
                             const int* __restrict__ eqn, long long* __restrict__ rl)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a == 0) rl[0] = n_eq + 1;
    if (a >= nnd) return;
    if (nptr[a + 1] == nptr[a]) return;  // a node of no element couples nothing: empty rows
    int later = 0;
    for (long long k = nptr[a]; k < nptr[a + 1]; k++) {
        const int b = nlist[k];
        if (b > a) later += (eqn[3L * b] >= 0) + (eqn[3L * b + 1] >= 0) + (eqn[3L * b + 2] >= 0);
    }
    for (int i = 0; i < 3; i++) {
        const int r = eqn[3L * a + i];
        if (r < 0) continue;
        int own = 0;
        for (int j = i + 1; j < 3; j++) own += eqn[3L * a + j] >= 0;
        rl[r + 1] = later + own;
    }
}

__global__ void k_narrow(long n, const long long* __restrict__ in, int* __restrict__ out)
{
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) out[k] = (int)in[k];
}

// MSR value ranges of the rows of each half slice (32 nodes): off-diagonal
// [bindx[r0], bindx[r1]) and diagonal [r0, r1) for the active rows r0 .. r1 - 1
// of its nodes (node-major equations: consecutive), (0, 0, 0, 0) without
// any; most[h] = the values the expansion stages for it.
__global__ void k_half_slice_ranges(int nnd, int nhalf, const int* __restrict__ eqn, const int* __restrict__ bindx,
                                    int4* __restrict__ rng, long long* __restrict__ most)
{
    const int h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= nhalf) return;
    int r0 = -1, r1 = -1;
    for (int a = 32 * h; a < 32 * h + 32 && a < nnd; a++)
        for (int i = 0; i < 3; i++) {
            const int e = eqn[3L * a + i];
            if (e < 0) continue;
            if (r0 < 0) r0 = e;
            r1 = e + 1;
        }
    if (r0 < 0) {
        rng[h] = make_int4(0, 0, 0, 0);
        most[h] = 0;
        return;
    }
    rng[h] = make_int4(bindx[r0], bindx[r1], r0, r1);
    most[h] = (long long)(bindx[r1] - bindx[r0]) + (r1 - r0);
}

unsigned grid(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// out[k] = in[0] + .. + in[k], k < n (temporary storage in c->g_scan)
template <class T>
void inclusive_sum(tg_ctx* c, const T* in, T* out, long n)
{
    size_t bytes = 0;
    TG_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, bytes, in, out, (int)n, c->stream));
    if (c->g_scan.n < bytes) c->g_scan.alloc(bytes);
    TG_HIP(hipcub::DeviceScan::InclusiveSum(c->g_scan.p, bytes, in, out, (int)n, c->stream));
}

}  // namespace

// Node graph on the device: g_neptr/g_nelist (elements per node), g_deg and
// g_nptr/g_nlist (sorted unique neighbours per node).
void device_node_graph(tg_ctx* c)
{
    const int nnd = c->nnd;
    const long n8 = 8L * c->nel;
    if (n8 > 0x7fffffffL) fail(TG_OUT_OF_RANGE, "connectivity exceeds 32-bit indexing");
    c->g_neptr.alloc((size_t)nnd + 1);
    c->g_nelist.alloc((size_t)n8);
    c->g_deg.alloc((size_t)nnd);
    c->g_nptr.alloc((size_t)nnd + 1);
    c->g_flag.alloc(1);
    c->g_fill.alloc((size_t)nnd + 1);
    TG_HIP(hipMemsetAsync(c->g_fill.p, 0, sizeof(int) * ((size_t)nnd + 1), c->stream));
    TG_HIP(hipMemsetAsync(c->g_flag.p, 0, sizeof(int), c->stream));
    // counts into fill[a + 1], their sums into neptr, then fill = the list starts
    hipLaunchKernelGGL(k_ne_count, dim3(grid(n8, 256)), dim3(256), 0, c->stream, n8, c->dconn.p, c->g_fill.p);
    inclusive_sum(c, c->g_fill.p, c->g_neptr.p, (long)nnd + 1);
    TG_HIP(hipMemcpyAsync(c->g_fill.p, c->g_neptr.p, sizeof(int) * nnd, hipMemcpyDeviceToDevice, c->stream));
    hipLaunchKernelGGL(k_ne_fill, dim3(grid(n8, 256)), dim3(256), 0, c->stream, n8, c->dconn.p, c->g_fill.p,
                       c->g_nelist.p);
    hipLaunchKernelGGL(k_neighbours<false>, dim3(grid(nnd, GB)), dim3(GB), 0, c->stream, nnd, c->g_neptr.p,
                       c->g_nelist.p, c->dconn.p, c->g_deg.p, nullptr, nullptr, c->g_flag.p);
    TG_HIP(hipGetLastError());
    int flag = 0;
    TG_HIP(hipMemcpyAsync(&flag, c->g_flag.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    // nptr[0] = 0, nptr[a + 1] = deg[0] + .. + deg[a]
    TG_HIP(hipMemsetAsync(c->g_nptr.p, 0, sizeof(long long), c->stream));
    inclusive_sum(c, c->g_deg.p, c->g_nptr.p + 1, (long)nnd);
    long long total = 0;
    TG_HIP(hipMemcpyAsync(&total, c->g_nptr.p + nnd, sizeof(long long), hipMemcpyDeviceToHost, c->stream));
    TG_HIP(hipStreamSynchronize(c->stream));
    if (flag) fail(TG_OUT_OF_RANGE, "a node is shared by more than 64 elements");
    if (total > 0x7fffffffLL) fail(TG_OUT_OF_RANGE, "node graph exceeds 32-bit indexing");
    c->g_nlist.alloc((size_t)total);
    hipLaunchKernelGGL(k_neighbours<true>, dim3(grid(nnd, GB)), dim3(GB), 0, c->stream, nnd, c->g_neptr.p,
                       c->g_nelist.p, c->dconn.p, c->g_deg.p, c->g_nptr.p, c->g_nlist.p, c->g_flag.p);
    TG_HIP(hipGetLastError());
}

// The solver operator's layout (slice widths and bases, block columns) and
// the MSR row pointers from the device node graph; the MSR columns, the
// expansion's value map, the lower blocks' partners and the half-slice
// value ranges follow.  Allocates the MSR value arrays (their length is
// known only here).
void device_structure(tg_ctx* c)
{
    Bsell& B = c->op;
    const int nnd = c->nnd;
    B.nslices = (nnd + 63) / 64;
    B.d_slice_width.alloc(B.nslices);
    B.d_slice_base.alloc(B.nslices);
    hipLaunchKernelGGL(k_slice_width, dim3(grid(B.nslices, 256)), dim3(256), 0, c->stream, nnd, B.nslices,
                       c->g_deg.p, B.d_slice_width.p);
    TG_HIP(hipGetLastError());
    B.slice_width.resize((size_t)B.nslices);
    TG_HIP(hipMemcpyAsync(B.slice_width.data(), B.d_slice_width.p, sizeof(int) * B.nslices, hipMemcpyDeviceToHost,
                          c->stream));
    TG_HIP(hipStreamSynchronize(c->stream));
    B.slice_base.resize((size_t)B.nslices);
    long base = 0;
    for (int s = 0; s < B.nslices; s++) {
        B.slice_base[s] = (int)base;
        base += B.slice_width[s];
    }
    B.nblocks_padded = base * 64;
    if (B.nblocks_padded * 9 > 0x7fffffffL) fail(TG_OUT_OF_RANGE, "block operator exceeds 32-bit indexing");
    TG_HIP(hipMemcpyAsync(B.d_slice_base.p, B.slice_base.data(), sizeof(int) * B.nslices, hipMemcpyHostToDevice,
                          c->stream));
    B.d_col.alloc((size_t)B.nblocks_padded);
    B.d_src.alloc((size_t)B.nblocks_padded * 9);
    B.d_tpos.alloc((size_t)B.nblocks_padded);
    B.d_val.alloc((size_t)B.nblocks_padded * 9);
    hipLaunchKernelGGL(k_block_columns, dim3(grid(64L * B.nslices, 256)), dim3(256), 0, c->stream, nnd, B.nslices,
                       B.d_slice_base.p, B.d_slice_width.p, c->g_nptr.p, c->g_nlist.p, B.d_col.p);
    // MSR row pointers, summed in 64 bits
    const int n_eq = c->n_eq;
    c->g_rowlen.alloc((size_t)n_eq + 1);
    c->g_rows.alloc((size_t)n_eq + 1);
    TG_HIP(hipMemsetAsync(c->g_rowlen.p, 0, sizeof(long long) * ((size_t)n_eq + 1), c->stream));
    hipLaunchKernelGGL(k_msr_rowlen, dim3(grid(nnd, 256)), dim3(256), 0, c->stream, nnd, n_eq, c->g_nptr.p,
                       c->g_nlist.p, c->deqn.p, c->g_rowlen.p);
    TG_HIP(hipGetLastError());
    inclusive_sum(c, c->g_rowlen.p, c->g_rows.p, (long)n_eq + 1);
    long long len = 0;
    TG_HIP(hipMemcpyAsync(&len, c->g_rows.p + n_eq, sizeof(long long), hipMemcpyDeviceToHost, c->stream));
    TG_HIP(hipStreamSynchronize(c->stream));
    if (len > 0x7fffffffLL) fail(TG_OUT_OF_RANGE, "MSR pattern exceeds 32-bit indexing");
    c->msr_len = (long)len;
    c->dbindx.alloc((size_t)c->msr_len);
    c->dval.alloc((size_t)c->msr_len);
    c->dval32.alloc(mixed_stiffness(c) ? (size_t)c->msr_len : 0);
    hipLaunchKernelGGL(k_narrow, dim3(grid((long)n_eq + 1, 256)), dim3(256), 0, c->stream, (long)n_eq + 1,
                       c->g_rows.p, c->dbindx.p);
    TG_HIP(hipGetLastError());
    launch_msr_cols(c);  // the MSR columns from the block columns, before anything reads them
    launch_build_src(c);
    // half-slice value ranges of the expansion
    const int nhalf = 2 * B.nslices;
    B.d_srange.alloc((size_t)nhalf);
    c->g_most.alloc((size_t)nhalf);
    hipLaunchKernelGGL(k_half_slice_ranges, dim3(grid(nhalf, 256)), dim3(256), 0, c->stream, nnd, nhalf, c->deqn.p,
                       c->dbindx.p, B.d_srange.p, c->g_most.p);
    TG_HIP(hipGetLastError());
    std::vector<long long> most((size_t)nhalf);
    TG_HIP(hipMemcpyAsync(most.data(), c->g_most.p, sizeof(long long) * nhalf, hipMemcpyDeviceToHost, c->stream));
    TG_HIP(hipStreamSynchronize(c->stream));
    B.lds_values = most.empty() ? 0 : (long)*std::max_element(most.begin(), most.end());
}

}  // namespace tg
