// Solver kernels: MSR -> block-operator expansion, block SpMV, deterministic
// dot products and the Jacobi-preconditioned conjugate-gradient loop.
//
// All vectors live in node space (3 entries per node, prescribed dofs held at
// zero), so the SpMV needs no equation lookups.  Reductions are two-level:
// every block writes a partial sum, the last block to finish adds the
// partials in block order (deterministic, no floating-point atomics) and
// updates the scalars of the iteration on the device.  The host only polls a
// "done" flag every few iterations; kernels enqueued after convergence
// return immediately.
#include "tg_internal.h"

namespace tg {

// device scalar slots
enum { S_RZ = 0, S_PQ, S_ALPHA, S_RZNEW, S_RR, S_BETA, S_BB, S_DONE, S_ITER, S_THRESH, S_MAXIT, S_DOT, S_NSCAL };

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

// Writes this block's partials; returns true in thread 0 of the last block,
// whose tot[] then holds the ordered sum over all blocks.
template <int NV>
__device__ __forceinline__ bool grid_sum(double (&v)[NV], double* part, unsigned* counter, double (&tot)[NV])
{
    __shared__ bool last;
    block_sum<NV>(v);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; k++) part[(long)k * gridDim.x + blockIdx.x] = v[k];
        __threadfence();
        const unsigned ticket = atomicAdd(counter, 1u);
        last = (ticket == gridDim.x - 1);
    }
    __syncthreads();
    if (!last) return false;
    __threadfence();
    double s[NV];
#pragma unroll
    for (int k = 0; k < NV; k++) s[k] = 0.0;
    for (unsigned b = threadIdx.x; b < gridDim.x; b += blockDim.x)
#pragma unroll
        for (int k = 0; k < NV; k++) s[k] += ((volatile double*)part)[(long)k * gridDim.x + b];
    // ordered combine: per-thread strided sums are added in thread order
    __shared__ double red[NV][VB];
#pragma unroll
    for (int k = 0; k < NV; k++) red[k][threadIdx.x] = s[k];
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; k++) {
            double x = 0.0;
            for (int i = 0; i < (int)blockDim.x; i++) x += red[k][i];
            tot[k] = x;
        }
        *counter = 0;
        return true;
    }
    return false;
}

// ---------------------------------------------------------------------------
// operator
// ---------------------------------------------------------------------------

__global__ void k_expand(long n, const int* __restrict__ src, const double* __restrict__ val, double* __restrict__ out)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const int s = src[i];
        out[i] = s >= 0 ? val[s] : 0.0;
    }
}

// inverse diagonal in node space (Jacobi; DiagonalMatrixT::Factorize/BackSubstitute,
// DiagonalMatrixT.cpp:267-324)
__global__ void k_dinv(long n, const int* __restrict__ eqn, const double* __restrict__ val, double* __restrict__ dinv)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const int e = eqn[i];
        const double d = e >= 0 ? val[e] : 0.0;
        dinv[i] = d != 0.0 ? 1.0 / d : 0.0;
    }
}

// y = A x for block rows of one thread each; optionally partial dot(y, w).
// The last block then sets alpha = rz / (p.q) when `pcg` is set.
template <bool PCG>
__global__ void __launch_bounds__(VB) k_spmv(int nnd, const int* __restrict__ sbase, const int* __restrict__ swidth,
                                             const int* __restrict__ bcol, const double* __restrict__ bval,
                                             const double* __restrict__ x, double* __restrict__ y,
                                             double* __restrict__ part, unsigned* __restrict__ counter,
                                             double* __restrict__ scal)
{
    if (PCG && scal[S_DONE] != 0.0) return;
    const long tid = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int s = (int)(tid >> 6), l = (int)(tid & 63);
    const int a = (int)tid;
    double dotv[1] = {0.0};
    if (a < nnd) {
        const long b0 = sbase[s];
        const int w = swidth[s];
        double y0 = 0.0, y1 = 0.0, y2 = 0.0;
        for (int k = 0; k < w; k++) {
            const long blk = b0 + k;
            const int c = bcol[blk * 64 + l];
            if (c < 0) continue;
            const double x0 = x[3L * c], x1 = x[3L * c + 1], x2 = x[3L * c + 2];
            const double* v = bval + blk * 9 * 64 + l;
            y0 += v[0 * 64] * x0 + v[1 * 64] * x1 + v[2 * 64] * x2;
            y1 += v[3 * 64] * x0 + v[4 * 64] * x1 + v[5 * 64] * x2;
            y2 += v[6 * 64] * x0 + v[7 * 64] * x1 + v[8 * 64] * x2;
        }
        y[3L * a] = y0;
        y[3L * a + 1] = y1;
        y[3L * a + 2] = y2;
        if (PCG) dotv[0] = y0 * x[3L * a] + y1 * x[3L * a + 1] + y2 * x[3L * a + 2];
    }
    if (PCG) {
        double tot[1];
        if (grid_sum<1>(dotv, part, counter, tot)) {
            scal[S_PQ] = tot[0];
            scal[S_ALPHA] = scal[S_RZ] / tot[0];
        }
    }
}

// ---------------------------------------------------------------------------
// PCG vector steps
// ---------------------------------------------------------------------------

// x = 0, r = b, p = D^-1 r; rz = r.D^-1 r, bb = b.b
__global__ void __launch_bounds__(VB) k_pcg_init(long n, const double* __restrict__ b, const double* __restrict__ dinv,
                                                 double* __restrict__ x, double* __restrict__ r, double* __restrict__ p,
                                                 double* __restrict__ part, unsigned* __restrict__ counter,
                                                 double* __restrict__ scal, double rtol, int maxit)
{
    double v[2] = {0.0, 0.0};
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const double bi = b[i], zi = dinv[i] * bi;
        x[i] = 0.0;
        r[i] = bi;
        p[i] = zi;
        v[0] += bi * zi;
        v[1] += bi * bi;
    }
    double tot[2];
    if (grid_sum<2>(v, part, counter, tot)) {
        scal[S_RZ] = tot[0];
        scal[S_BB] = tot[1];
        scal[S_RR] = tot[1];
        scal[S_ITER] = 0.0;
        scal[S_THRESH] = rtol * rtol * tot[1];
        scal[S_MAXIT] = (double)maxit;
        scal[S_DONE] = (tot[1] == 0.0 || maxit <= 0) ? 1.0 : 0.0;
    }
}

// x += alpha p, r -= alpha q; rz' = r.D^-1 r, rr = r.r; beta = rz'/rz
__global__ void __launch_bounds__(VB) k_pcg_update(long n, const double* __restrict__ p, const double* __restrict__ q,
                                                   const double* __restrict__ dinv, double* __restrict__ x,
                                                   double* __restrict__ r, double* __restrict__ part,
                                                   unsigned* __restrict__ counter, double* __restrict__ scal)
{
    if (scal[S_DONE] != 0.0) return;
    const double alpha = scal[S_ALPHA];
    double v[2] = {0.0, 0.0};
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        x[i] += alpha * p[i];
        const double ri = r[i] - alpha * q[i];
        r[i] = ri;
        v[0] += ri * dinv[i] * ri;
        v[1] += ri * ri;
    }
    double tot[2];
    if (grid_sum<2>(v, part, counter, tot)) {
        const double it = scal[S_ITER] + 1.0;
        scal[S_ITER] = it;
        scal[S_BETA] = tot[0] / scal[S_RZ];
        scal[S_RZ] = tot[0];
        scal[S_RR] = tot[1];
        if (tot[1] <= scal[S_THRESH] || it >= scal[S_MAXIT]) scal[S_DONE] = 1.0;
    }
}

// p = D^-1 r + beta p
__global__ void __launch_bounds__(VB) k_pcg_direction(long n, const double* __restrict__ r,
                                                      const double* __restrict__ dinv, double* __restrict__ p,
                                                      const double* __restrict__ scal)
{
    if (scal[S_DONE] != 0.0) return;
    const double beta = scal[S_BETA];
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        p[i] = dinv[i] * r[i] + beta * p[i];
}

__global__ void __launch_bounds__(VB) k_dot(long n, const double* __restrict__ a, const double* __restrict__ b,
                                            double* __restrict__ part, unsigned* __restrict__ counter,
                                            double* __restrict__ scal)
{
    double v[1] = {0.0};
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        v[0] += a[i] * b[i];
    double tot[1];
    if (grid_sum<1>(v, part, counter, tot)) scal[S_DOT] = tot[0];
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

static unsigned* counter_of(tg_ctx* c) { return reinterpret_cast<unsigned*>(c->dstatus.p + 1); }

void launch_expand(tg_ctx* c)
{
    TimedRegion tr(c, T_EXPAND);
    const long n = c->op.nblocks_padded * 9;
    hipLaunchKernelGGL(k_expand, dim3(grid_for(n) * 4), dim3(VB), 0, c->stream, n, c->op.d_src.p, c->dval.p,
                       c->op.d_val.p);
    const long nn = 3L * c->nnd;
    hipLaunchKernelGGL(k_dinv, dim3(grid_for(nn)), dim3(VB), 0, c->stream, nn, c->deqn.p, c->dval.p, c->vdinv.p);
    TG_HIP(hipGetLastError());
}

void launch_spmv(tg_ctx* c, const double* x, double* y, double*, const double*)
{
    TimedRegion tr(c, T_SPMV);
    const unsigned blocks = (unsigned)((c->nnd + VB - 1) / VB);
    hipLaunchKernelGGL(k_spmv<false>, dim3(blocks), dim3(VB), 0, c->stream, c->nnd, c->op.d_slice_base.p,
                       c->op.d_slice_width.p, c->op.d_col.p, c->op.d_val.p, x, y, c->dpart.p, counter_of(c),
                       c->dscal.p);
    TG_HIP(hipGetLastError());
}

double device_dot(tg_ctx* c, const double* a, const double* b, long n)
{
    {
        TimedRegion tr(c, T_VECTOR);
        hipLaunchKernelGGL(k_dot, dim3(grid_for(n)), dim3(VB), 0, c->stream, n, a, b, c->dpart.p, counter_of(c),
                           c->dscal.p);
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

// Jacobi-PCG on the expanded operator: b, x in node space.  The reference has
// no linear CG; this is the textbook method (Saad, Iterative Methods for
// Sparse Linear Systems, Alg. 9.1) on the reference's building blocks:
// Multx (MSRMatrixT.cpp:385-420), Jacobi scaling (DiagonalMatrixT.cpp:267-324),
// InnerProduct (SolverT.cpp:854-860).  Stops at ||r|| <= rtol ||b||.
void pcg_solve(tg_ctx* c, const double* b, double* x, double rtol, int maxit, tg_lin_stats* st)
{
    const long n = 3L * c->nnd;
    const unsigned g = grid_for(n);
    const unsigned sblocks = (unsigned)((c->nnd + VB - 1) / VB);
    double* r = c->vr.p;
    double* p = c->vp.p;
    double* q = c->vq.p;
    {
        TimedRegion tr(c, T_VECTOR);
        hipLaunchKernelGGL(k_pcg_init, dim3(g), dim3(VB), 0, c->stream, n, b, c->vdinv.p, x, r, p, c->dpart.p,
                           counter_of(c), c->dscal.p, rtol, maxit);
    }
    const int CHUNK = 16;
    int issued = 0;
    for (;;) {
        for (int k = 0; k < CHUNK; k++) {
            {
                TimedRegion tr(c, T_SPMV);
                hipLaunchKernelGGL(k_spmv<true>, dim3(sblocks), dim3(VB), 0, c->stream, c->nnd, c->op.d_slice_base.p,
                                   c->op.d_slice_width.p, c->op.d_col.p, c->op.d_val.p, p, q, c->dpart.p,
                                   counter_of(c), c->dscal.p);
            }
            {
                TimedRegion tr(c, T_VECTOR);
                hipLaunchKernelGGL(k_pcg_update, dim3(g), dim3(VB), 0, c->stream, n, p, q, c->vdinv.p, x, r,
                                   c->dpart.p, counter_of(c), c->dscal.p);
                hipLaunchKernelGGL(k_pcg_direction, dim3(g), dim3(VB), 0, c->stream, n, r, c->vdinv.p, p, c->dscal.p);
            }
        }
        TG_HIP(hipGetLastError());
        issued += CHUNK;
        TG_HIP(hipMemcpyAsync(c->hscal, c->dscal.p, sizeof(double) * S_NSCAL, hipMemcpyDeviceToHost, c->stream));
        TG_HIP(hipStreamSynchronize(c->stream));
        if (c->hscal[S_DONE] != 0.0 || issued >= maxit) break;
    }
    if (st) {
        st->iterations = (int)c->hscal[S_ITER];
        const double bb = c->hscal[S_BB];
        st->rel_residual = bb > 0.0 ? sqrt(c->hscal[S_RR] / bb) : 0.0;
        st->converged = (c->hscal[S_RR] <= c->hscal[S_THRESH]) ? 1 : 0;
    }
}

}  // namespace tg
