// Solver kernels: MSR -> block-operator expansion, block SpMV, deterministic
// dot products and the Jacobi-preconditioned conjugate-gradient loop.
//
// All vectors live in node space (3 entries per node, prescribed dofs held at
// zero), so the SpMV needs no equation lookups.  Reductions are two-level:
// every block writes a partial sum and a one-block finalize kernel adds the
// partials in block order (deterministic, no atomics) and updates the
// scalars of the iteration on the device.  The host only polls a "done" flag
// every few iterations; kernels enqueued after convergence return at once.
#include "tg_internal.h"

namespace tg {


constexpr int VB = 256;      // threads per block of the vector kernels
constexpr int VGRID = 1024;  // blocks of the grid-stride vector kernels

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// Block sum of NV values; thread 0 gets the result.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV])
{
    __shared__ double red[NV][VB / 64];
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
#pragma unroll
    for (int k = 0; k < NV; k++) {
        v[k] = wave_sum(v[k]);
        if (l == 0) red[k][w] = v[k];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; k++) {
            double s = 0.0;
            for (int i = 0; i < (int)(blockDim.x >> 6); i++) s += red[k][i];
            v[k] = s;
        }
    }
}

// Writes this block's NV partial sums to part[k * gridDim.x + blockIdx.x].
template <int NV>
__device__ __forceinline__ void block_partials(double (&v)[NV], double* part)
{
    block_sum<NV>(v);
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < NV; k++) part[(long)k * gridDim.x + blockIdx.x] = v[k];
}

// Ordered sum of nparts partials per value, by one block of VB threads:
// thread i sums partials i, i+VB, ... in order, then thread 0 adds the VB
// thread sums in order.  Deterministic for a fixed grid.
template <int NV>
__device__ __forceinline__ void ordered_total(const double* part, int nparts, double (&tot)[NV])
{
    __shared__ double red[NV][VB];
#pragma unroll
    for (int k = 0; k < NV; k++) {
        double x = 0.0;
        for (int b = threadIdx.x; b < nparts; b += VB) x += part[(long)k * nparts + b];
        red[k][threadIdx.x] = x;
    }
    __syncthreads();
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < NV; k++) {
            double x = 0.0;
            for (int i = 0; i < VB; i++) x += red[k][i];
            tot[k] = x;
        }
}

// ---------------------------------------------------------------------------
// operator
// ---------------------------------------------------------------------------

// pass 1: one workgroup per half slice (32 block rows, lanes 32 h .. 32 h +
// 31 of slice s).  Those rows own one contiguous range of off-diagonal MSR
// values and one of diagonal values (node-major equations); STAGED copies
// both into LDS with coalesced reads, so every MSR value is read from HBM
// once.  Then 8 groups of 32 threads take the block slots in turn, and a
// lane of a slot with column >= row (or padding) fetches its 9 components
// through the MSR map and writes them (32 lanes of one component are
// consecutive).  Without STAGED (rows too long for LDS) the components are
// gathered from HBM.
template <class VT, bool STAGED>
__global__ void __launch_bounds__(256) k_expand(const int* __restrict__ sbase, const int* __restrict__ swidth,
                                                const int4* __restrict__ srange, const int* __restrict__ src,
                                                const int* __restrict__ tpos, const VT* __restrict__ val,
                                                VT* __restrict__ out)
{
    extern __shared__ unsigned char lds_raw[];
    VT* buf = reinterpret_cast<VT*>(lds_raw);
    const int s = blockIdx.x >> 1, h = blockIdx.x & 1;
    const int4 rg = srange[blockIdx.x];  // off-diagonal [x, y), diagonal [z, w)
    const int ne = rg.y - rg.x, nd = rg.w - rg.z;
    if (STAGED) {
        for (int x = threadIdx.x; x < ne; x += 256) buf[x] = val[(long)rg.x + x];
        for (int x = threadIdx.x; x < nd; x += 256) buf[ne + x] = val[(long)rg.z + x];
        __syncthreads();
    }
    const long b0 = sbase[s];
    const int width = swidth[s];
    const int l = 32 * h + (threadIdx.x & 31);
    for (int k = threadIdx.x >> 5; k < width; k += 8) {
        const long blk = b0 + k;
        if (tpos[blk * 64 + l] >= 0) continue;
#pragma unroll
        for (int c = 0; c < 9; c++) {
            const long o = (blk * 9 + c) * 64 + l;
            const int sv = src[o];
            VT v = VT(0);
            if (sv >= 0) {
                if (STAGED && sv >= rg.x && sv < rg.y)
                    v = buf[sv - rg.x];
                else if (STAGED && sv >= rg.z && sv < rg.w)
                    v = buf[ne + (sv - rg.z)];
                else
                    v = val[sv];
            }
            out[o] = v;
        }
    }
}

// the expansion's pass 1 for values of type VT
template <class VT>
static void launch_expand_pass1(tg_ctx* c, const VT* val, VT* out)
{
    const size_t lds = (size_t)c->op.lds_values * sizeof(VT);
    const unsigned grid = 2u * (unsigned)c->op.nslices;
    if (lds > 0 && lds <= 65536)
        hipLaunchKernelGGL((k_expand<VT, true>), dim3(grid), dim3(256), lds, c->stream, c->op.d_slice_base.p,
                           c->op.d_slice_width.p, c->op.d_srange.p, c->op.d_src.p, c->op.d_tpos.p, val, out);
    else
        hipLaunchKernelGGL((k_expand<VT, false>), dim3(grid), dim3(256), 0, c->stream, c->op.d_slice_base.p,
                           c->op.d_slice_width.p, c->op.d_srange.p, c->op.d_src.p, c->op.d_tpos.p, val, out);
}

// pass 2: lower blocks as the transposes of their upper partners, written
// by pass 1 (consecutive rows read consecutive lanes of the partner slot)
template <class VT>
__global__ void k_expand_lower(long nb, const int* __restrict__ tpos, VT* __restrict__ out)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < nb; i += (long)gridDim.x * blockDim.x) {
        const int t = tpos[i];
        if (t < 0) continue;
        const long blk = i >> 6, l = i & 63;
#pragma unroll
        for (int r = 0; r < 3; r++)
#pragma unroll
            for (int q = 0; q < 3; q++) out[(blk * 9 + 3 * r + q) * 64 + l] = out[t + (3 * q + r) * 64];
    }
}

// inverse diagonal in node space (Jacobi; DiagonalMatrixT::Factorize/BackSubstitute,
// DiagonalMatrixT.cpp:267-324)
template <class VT>
__global__ void k_dinv(long n, const int* __restrict__ eqn, const VT* __restrict__ val, double* __restrict__ dinv)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const int e = eqn[i];
        const double d = e >= 0 ? (double)val[e] : 0.0;
        dinv[i] = d != 0.0 ? 1.0 / d : 0.0;
    }
}

// inverse diagonal from the operator's diagonal blocks (the MSR diagonal
// when the stiffness went straight into the operator)
template <class VT>
__global__ void k_dinv_op(int nnd, const int* __restrict__ eqn, const int* __restrict__ diag,
                          const VT* __restrict__ val, double* __restrict__ dinv)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= nnd) return;
    const int dg = diag[a];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const double d = (eqn[3L * a + i] >= 0 && dg >= 0) ? (double)val[dg + 4 * i * 64] : 0.0;
        dinv[3L * a + i] = d != 0.0 ? 1.0 / d : 0.0;
    }
}

// 3x3 nodal block Jacobi: inverse of each row's diagonal block over its
// active dofs (prescribed rows and columns of the operator are zero; they
// stay zero in the inverse).  Singular active blocks are left zero, the
// block analogue of DiagonalMatrixT's skipped zero pivots.
template <class VT>
__global__ void k_binv_op(int nnd, const int* __restrict__ eqn, const int* __restrict__ diag,
                          const VT* __restrict__ val, double* __restrict__ binv)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= nnd) return;
    const int dg = diag[a];
    bool act[3];
    double M[9];
#pragma unroll
    for (int i = 0; i < 3; i++) act[i] = eqn[3L * a + i] >= 0 && dg >= 0;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++)
            M[3 * i + j] = (act[i] && act[j]) ? (double)val[dg + (3 * i + j) * 64] : (i == j ? 1.0 : 0.0);
    const double det = M[0] * (M[4] * M[8] - M[5] * M[7]) - M[1] * (M[3] * M[8] - M[5] * M[6]) +
                       M[2] * (M[3] * M[7] - M[4] * M[6]);
    double Mi[9];
    if (det != 0.0) {
        const double id = 1.0 / det;
        Mi[0] = (M[4] * M[8] - M[5] * M[7]) * id;
        Mi[1] = (M[2] * M[7] - M[1] * M[8]) * id;
        Mi[2] = (M[1] * M[5] - M[2] * M[4]) * id;
        Mi[3] = (M[5] * M[6] - M[3] * M[8]) * id;
        Mi[4] = (M[0] * M[8] - M[2] * M[6]) * id;
        Mi[5] = (M[2] * M[3] - M[0] * M[5]) * id;
        Mi[6] = (M[3] * M[7] - M[4] * M[6]) * id;
        Mi[7] = (M[1] * M[6] - M[0] * M[7]) * id;
        Mi[8] = (M[0] * M[4] - M[1] * M[3]) * id;
    } else {
#pragma unroll
        for (int k = 0; k < 9; k++) Mi[k] = 0.0;
    }
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) binv[9L * a + 3 * i + j] = (act[i] && act[j]) ? Mi[3 * i + j] : 0.0;
}

// y = A x for block rows of one thread each; with PCG also the block
// partials of p.q (x = p, y = q) for k_fin_alpha.
template <bool PCG, class VT>
__global__ void __launch_bounds__(VB) k_spmv(int nnd, const int* __restrict__ sbase, const int* __restrict__ swidth,
                                             const int* __restrict__ bcol, const VT* __restrict__ bval,
                                             const VT* __restrict__ x, VT* __restrict__ y,
                                             double* __restrict__ part, const double* __restrict__ scal)
{
    if (PCG && scal[S_DONE] != 0.0) return;
    const long tid = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int s = (int)(tid >> 6), l = (int)(tid & 63);
    const int a = (int)tid;
    double dotv[1] = {0.0};
    if (a < nnd) {
        const long b0 = sbase[s];
        const int w = swidth[s];
        VT y0 = 0, y1 = 0, y2 = 0;
        for (int k = 0; k < w; k++) {
            const long blk = b0 + k;
            const int c = bcol[blk * 64 + l];
            if (c < 0) continue;
            const VT x0 = x[3L * c], x1 = x[3L * c + 1], x2 = x[3L * c + 2];
            const VT* v = bval + blk * 9 * 64 + l;
            y0 += v[0 * 64] * x0 + v[1 * 64] * x1 + v[2 * 64] * x2;
            y1 += v[3 * 64] * x0 + v[4 * 64] * x1 + v[5 * 64] * x2;
            y2 += v[6 * 64] * x0 + v[7 * 64] * x1 + v[8 * 64] * x2;
        }
        y[3L * a] = y0;
        y[3L * a + 1] = y1;
        y[3L * a + 2] = y2;
        if (PCG)
            dotv[0] = (double)y0 * (double)x[3L * a] + (double)y1 * (double)x[3L * a + 1] +
                      (double)y2 * (double)x[3L * a + 2];
    }
    if (PCG) block_partials<1>(dotv, part);
}

__global__ void __launch_bounds__(VB) k_dot(long n, const double* __restrict__ a, const double* __restrict__ b,
                                            double* __restrict__ part)
{
    double v[1] = {0.0};
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        v[0] += a[i] * b[i];
    block_partials<1>(v, part);
}

__global__ void __launch_bounds__(VB) k_fin_dot(const double* __restrict__ part, int nparts, double* __restrict__ scal)
{
    double tot[1];
    ordered_total<1>(part, nparts, tot);
    if (threadIdx.x == 0) scal[S_DOT] = tot[0];
}

template <class TO, class TI>
__global__ void k_convert(long n, const TI* __restrict__ src, TO* __restrict__ dst)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        dst[i] = (TO)src[i];
}

__global__ void k_axpy(long n, double* __restrict__ y, const double* __restrict__ x, double a)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        y[i] += a * x[i];
}

__global__ void k_eq_to_node(long n, const int* __restrict__ eqn, const double* __restrict__ src,
                             double* __restrict__ dst)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const int e = eqn[i];
        dst[i] = e >= 0 ? src[e] : 0.0;
    }
}

__global__ void k_node_to_eq(long n, const int* __restrict__ eqn, const double* __restrict__ src,
                             double* __restrict__ dst)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const int e = eqn[i];
        if (e >= 0) dst[e] = src[i];
    }
}

__global__ void k_apply_kbc(int n, const int* __restrict__ idx, const double* __restrict__ val, double t,
                            double* __restrict__ u)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) u[idx[i]] = val[i] * t;
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------

static unsigned grid_for(long n)
{
    long g = (n + VB - 1) / VB;
    return (unsigned)std::max(1L, std::min(g, (long)VGRID));
}

void launch_expand(tg_ctx* c)
{
    TimedRegion tr(c, T_EXPAND);
    const long n = c->op.nblocks_padded;
    launch_expand_pass1<double>(c, c->dval.p, c->op.d_val.p);
    hipLaunchKernelGGL(k_expand_lower<double>, dim3(grid_for(n) * 4), dim3(VB), 0, c->stream, n, c->op.d_tpos.p,
                       c->op.d_val.p);
    const long nn = 3L * c->nnd;
    hipLaunchKernelGGL(k_dinv<double>, dim3(grid_for(nn)), dim3(VB), 0, c->stream, nn, c->deqn.p, c->dval.p,
                       c->vdinv.p);
    TG_HIP(hipGetLastError());
}

void launch_expand32(tg_ctx* c)
{
    TimedRegion tr(c, T_EXPAND);
    const long n = c->op.nblocks_padded;
    launch_expand_pass1<float>(c, c->dval32.p, c->op.d_val32.p);
    hipLaunchKernelGGL(k_expand_lower<float>, dim3(grid_for(n) * 4), dim3(VB), 0, c->stream, n, c->op.d_tpos.p,
                       c->op.d_val32.p);
    const long nn = 3L * c->nnd;
    hipLaunchKernelGGL(k_dinv<float>, dim3(grid_for(nn)), dim3(VB), 0, c->stream, nn, c->deqn.p, c->dval32.p,
                       c->vdinv.p);
    TG_HIP(hipGetLastError());
}

void launch_operator_dinv(tg_ctx* c)
{
    TimedRegion tr(c, T_EXPAND);
    const unsigned blocks = (unsigned)((c->nnd + VB - 1) / VB);
    if (c->precision == TG_MIXED)
        hipLaunchKernelGGL(k_dinv_op<float>, dim3(blocks), dim3(VB), 0, c->stream, c->nnd, c->deqn.p,
                           c->op.d_diag.p, c->op.d_val32.p, c->vdinv.p);
    else
        hipLaunchKernelGGL(k_dinv_op<double>, dim3(blocks), dim3(VB), 0, c->stream, c->nnd, c->deqn.p,
                           c->op.d_diag.p, c->op.d_val.p, c->vdinv.p);
    TG_HIP(hipGetLastError());
}

void launch_block_inverse(tg_ctx* c)
{
    if (c->precond != TG_PC_BLOCK_JACOBI) return;
    TimedRegion tr(c, T_EXPAND);
    const unsigned blocks = (unsigned)((c->nnd + VB - 1) / VB);
    const bool f32 = mixed_operator(c);
    if (f32)
        hipLaunchKernelGGL(k_binv_op<float>, dim3(blocks), dim3(VB), 0, c->stream, c->nnd, c->deqn.p,
                           c->op.d_diag.p, c->op.d_val32.p, c->vbinv.p);
    else
        hipLaunchKernelGGL(k_binv_op<double>, dim3(blocks), dim3(VB), 0, c->stream, c->nnd, c->deqn.p,
                           c->op.d_diag.p, c->op.d_val.p, c->vbinv.p);
    TG_HIP(hipGetLastError());
}

void widen_values(tg_ctx* c)
{
    const long n = c->msr_len;
    hipLaunchKernelGGL((k_convert<double, float>), dim3(grid_for(n) * 4), dim3(VB), 0, c->stream, n, c->dval32.p,
                       c->dval.p);
    TG_HIP(hipGetLastError());
}

// FP32 copy of an FP64 operator (the multigrid smoother's operator in FP64 runs)
void launch_narrow_operator(tg_ctx* c)
{
    TimedRegion tr(c, T_PRECOND);
    const long n = c->op.nblocks_padded * 9;
    hipLaunchKernelGGL((k_convert<float, double>), dim3(grid_for(n) * 4), dim3(VB), 0, c->stream, n, c->op.d_val.p,
                       c->op.d_val32.p);
    TG_HIP(hipGetLastError());
}

void launch_spmv(tg_ctx* c, const double* x, double* y)
{
    TimedRegion tr(c, T_SPMV);
    const unsigned blocks = (unsigned)((c->nnd + VB - 1) / VB);
    hipLaunchKernelGGL((k_spmv<false, double>), dim3(blocks), dim3(VB), 0, c->stream, c->nnd, c->op.d_slice_base.p,
                       c->op.d_slice_width.p, c->op.d_col.p, c->op.d_val.p, x, y, c->dpart.p, c->dscal.p);
    TG_HIP(hipGetLastError());
}

double device_dot(tg_ctx* c, const double* a, const double* b, long n)
{
    {
        TimedRegion tr(c, T_VECTOR);
        const unsigned g = grid_for(n);
        hipLaunchKernelGGL(k_dot, dim3(g), dim3(VB), 0, c->stream, n, a, b, c->dpart.p);
        hipLaunchKernelGGL(k_fin_dot, dim3(1), dim3(VB), 0, c->stream, c->dpart.p, (int)g, c->dscal.p);
        TG_HIP(hipGetLastError());
    }
    TG_HIP(hipMemcpyAsync(c->hscal, c->dscal.p + S_DOT, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    TG_HIP(hipStreamSynchronize(c->stream));
    return c->hscal[0];
}

void launch_axpy(tg_ctx* c, double* y, const double* x, double a, long n)
{
    TimedRegion tr(c, T_VECTOR);
    hipLaunchKernelGGL(k_axpy, dim3(grid_for(n)), dim3(VB), 0, c->stream, n, y, x, a);
    TG_HIP(hipGetLastError());
}

void eq_to_node(tg_ctx* c, const double* src_eq, double* dst_node)
{
    const long n = 3L * c->nnd;
    hipLaunchKernelGGL(k_eq_to_node, dim3(grid_for(n)), dim3(VB), 0, c->stream, n, c->deqn.p, src_eq, dst_node);
    TG_HIP(hipGetLastError());
}

void node_to_eq(tg_ctx* c, const double* src_node, double* dst_eq)
{
    const long n = 3L * c->nnd;
    hipLaunchKernelGGL(k_node_to_eq, dim3(grid_for(n)), dim3(VB), 0, c->stream, n, c->deqn.p, src_node, dst_eq);
    TG_HIP(hipGetLastError());
}

void launch_apply_kbc(tg_ctx* c, double t)
{
    const int n = (int)c->kbc_nodes.size();
    if (n == 0) return;
    hipLaunchKernelGGL(k_apply_kbc, dim3((n + 255) / 256), dim3(256), 0, c->stream, n, c->dkbc_idx.p, c->dkbc_val.p, t,
                       c->du.p);
    TG_HIP(hipGetLastError());
}

void pcg_solve(tg_ctx* c, const double* b, double* x, double rtol, int maxit, tg_lin_stats* st)
{
    pcg_run<double>(c, c->op.d_val.p, c->vdinv.p, b, x, c->vr.p, c->vp.p, c->vq.p, rtol, maxit, c->seg.d_active.p,
                    c->seg.nseg, st);
}

// Mixed precision inner solve: the operator values are the FP32 stiffness;
// the Krylov vectors, products and dot products stay FP64.  (CG with FP32
// vectors stalls on these systems: at 128^3 the condition number is ~1e7,
// the order of 1/eps_32, and the FP32 recurrence loses the residual.)
void pcg_solve_mixed(tg_ctx* c, const double* b, double* x, double rtol, int maxit, tg_lin_stats* st)
{
    pcg_run<float>(c, c->op.d_val32.p, c->vdinv.p, b, x, c->vr.p, c->vp.p, c->vq.p, rtol, maxit, c->seg.d_active.p,
                   c->seg.nseg, st);
}

}  // namespace tg

namespace tg {

__global__ void __launch_bounds__(256) k_copy_calib(long n2, const double2* __restrict__ src, double2* __restrict__ dst)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n2; i += (long)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

}  // namespace tg

extern "C" int tg_calib_copy(int device, long bytes, int reps, double* ms)
{
    using namespace tg;
    if (bytes < 16 || reps < 1 || !ms) return TG_BAD_INPUT;
    if (hipSetDevice(device) != hipSuccess) return TG_OUT_OF_RANGE;
    const long n2 = bytes / 16;
    double2 *a = nullptr, *b = nullptr;
    if (hipMalloc((void**)&a, n2 * 16) != hipSuccess) return TG_OUT_OF_MEMORY;
    if (hipMalloc((void**)&b, n2 * 16) != hipSuccess) {
        (void)hipFree(a);
        return TG_OUT_OF_MEMORY;
    }
    int rc = TG_OK;
    hipEvent_t e0, e1;
    if (hipMemset(a, 0, n2 * 16) != hipSuccess || hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
        rc = TG_GENERAL_FAIL;
    if (rc == TG_OK) {
        const unsigned g = 4096;
        hipLaunchKernelGGL(k_copy_calib, dim3(g), dim3(256), 0, 0, n2, a, b);  // warm-up
        (void)hipEventRecord(e0, 0);
        for (int r = 0; r < reps; r++) hipLaunchKernelGGL(k_copy_calib, dim3(g), dim3(256), 0, 0, n2, a, b);
        (void)hipEventRecord(e1, 0);
        float t = 0.0f;
        if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&t, e0, e1) != hipSuccess)
            rc = TG_GENERAL_FAIL;
        *ms = (double)t / reps;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }
    (void)hipFree(a);
    (void)hipFree(b);
    return rc;
}
