// Restarted GMRES(m) with left preconditioning on the device: the linear
// solve of the Newton-Krylov solver (newton_krylov_solver, NLSolver_NK::
// Iterate/GMRES, NLSolver_NK.cpp:76-238).
//
// Same iteration as the reference: each restart cycle starts from
// r = M^-1 (b - A x), beta = |r|, and stops when the Givens-rotated residual
// |g_{j+1}| < rtol * beta of that cycle; Givens rotations [c s; -s c]; back
// substitution skipping |H_ii| <= 1e-14; at most maxit Arnoldi steps (3 n_eq
// when maxit <= 0).  The Arnoldi orthogonalisation is classical Gram-Schmidt
// applied twice (all j+1 projections of a pass in one multi-dot kernel and
// one update kernel) instead of the reference's modified Gram-Schmidt, which
// needs one reduction per basis vector; two CGS passes are as stable as MGS.
// M is the context's preconditioner (Jacobi as in the reference, or block
// Jacobi / multigrid).  The small Hessenberg work runs on the host.  Vectors
// live in node space [nnd][3] (prescribed dofs zero throughout).
#include <cmath>
#include <cstring>

#include "tg_internal.h"

namespace tg {

namespace {

constexpr int GB = 256;   // threads per block
constexpr int GNB = 512;  // blocks of the partial-sum kernels

inline unsigned nblocks(long n) { return (unsigned)((n + GB - 1) / GB); }

// y = A x on the block operator (storage MT, FP64 sums)
template <class MT>
__global__ void __launch_bounds__(GB) k_gm_spmv(int nnd, const int* __restrict__ sbase, const int* __restrict__ swidth,
                                                const int* __restrict__ col, const MT* __restrict__ val,
                                                const double* __restrict__ x, double* __restrict__ y)
{
    const int a = blockIdx.x * GB + threadIdx.x;
    if (a >= nnd) return;
    const int s = a >> 6, l = a & 63;
    const long b0 = sbase[s];
    const int w = swidth[s];
    double y0 = 0.0, y1 = 0.0, y2 = 0.0;
    for (int q = 0; q < w; q++) {
        const long blk = b0 + q;
        const int c = col[blk * 64 + l];
        if (c < 0) continue;
        const double x0 = x[3L * c], x1 = x[3L * c + 1], x2 = x[3L * c + 2];
        const MT* v = val + blk * 576 + l;
        y0 += (double)v[0] * x0 + (double)v[64] * x1 + (double)v[128] * x2;
        y1 += (double)v[192] * x0 + (double)v[256] * x1 + (double)v[320] * x2;
        y2 += (double)v[384] * x0 + (double)v[448] * x1 + (double)v[512] * x2;
    }
    y[3L * a] = y0;
    y[3L * a + 1] = y1;
    y[3L * a + 2] = y2;
}

// z = M^-1 y for the point and block Jacobi preconditioners
__global__ void __launch_bounds__(GB) k_gm_jacobi(long n3, const double* __restrict__ dinv, const double* __restrict__ y,
                                                  double* __restrict__ z)
{
    const long k = (long)blockIdx.x * GB + threadIdx.x;
    if (k < n3) z[k] = dinv[k] * y[k];
}

__global__ void __launch_bounds__(GB) k_gm_block(int nnd, const double* __restrict__ binv, const double* __restrict__ y,
                                                 double* __restrict__ z)
{
    const int a = blockIdx.x * GB + threadIdx.x;
    if (a >= nnd) return;
    const double* b = binv + 9L * a;
    const double r0 = y[3L * a], r1 = y[3L * a + 1], r2 = y[3L * a + 2];
#pragma unroll
    for (int i = 0; i < 3; i++) z[3L * a + i] = b[3 * i] * r0 + b[3 * i + 1] * r1 + b[3 * i + 2] * r2;
}

// w = b - w
__global__ void __launch_bounds__(GB) k_gm_bminus(long n3, const double* __restrict__ b, double* __restrict__ w)
{
    const long k = (long)blockIdx.x * GB + threadIdx.x;
    if (k < n3) w[k] = b[k] - w[k];
}

// part[i][blk] = partial of V_i . w, i < j1 (V_i = V + i * n3)
__global__ void __launch_bounds__(GB) k_gm_mdot(long n3, const double* __restrict__ V, int j1,
                                                const double* __restrict__ w, double* __restrict__ part)
{
    __shared__ double red[GB / 64];
    for (int i = 0; i < j1; i++) {
        const double* v = V + (long)i * n3;
        double s = 0.0;
        for (long k = (long)blockIdx.x * GB + threadIdx.x; k < n3; k += (long)gridDim.x * GB) s += v[k] * w[k];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
        __syncthreads();
        if (threadIdx.x == 0) {
            double t = 0.0;
            for (int q = 0; q < GB / 64; q++) t += red[q];
            part[(long)i * gridDim.x + blockIdx.x] = t;
        }
        __syncthreads();
    }
}

// out[i] = sum of part[i][0..np) (one block per i, fixed order)
__global__ void __launch_bounds__(GB) k_gm_fin(const double* __restrict__ part, int np, double* __restrict__ out)
{
    __shared__ double red[GB / 64];
    const int i = blockIdx.x;
    double s = 0.0;
    for (int k = threadIdx.x; k < np; k += GB) s += part[(long)i * np + k];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int q = 0; q < GB / 64; q++) t += red[q];
        out[i] = t;
    }
}

// w -= sum_i h[i] V_i  (sign = -1), or x += sum_i h[i] V_i (sign = +1)
__global__ void __launch_bounds__(GB) k_gm_combine(long n3, const double* __restrict__ V, int j1,
                                                   const double* __restrict__ h, double sign, double* __restrict__ w)
{
    const long k = (long)blockIdx.x * GB + threadIdx.x;
    if (k >= n3) return;
    double s = 0.0;
    for (int i = 0; i < j1; i++) s += h[i] * V[(long)i * n3 + k];
    w[k] += sign * s;
}

// v = w * s
__global__ void __launch_bounds__(GB) k_gm_scale(long n3, const double* __restrict__ w, double s, double* __restrict__ v)
{
    const long k = (long)blockIdx.x * GB + threadIdx.x;
    if (k < n3) v[k] = s * w[k];
}

}  // namespace

template <class MT>
static void gmres_t(tg_ctx* c, const MT* val, const double* b, double* x, double rtol, int maxit, tg_lin_stats* st)
{
    const int nnd = c->nnd;
    const long n3 = 3L * nnd;
    const int m = c->gmres_restart;
    Gmres& G = c->gm;
    if (G.m != m || G.n3 != n3) {
        G.V.alloc((size_t)(m + 1) * n3);
        G.w.alloc(n3);
        G.q.alloc(n3);
        G.part.alloc((size_t)(m + 1) * GNB);
        G.h.alloc(m + 1);
        G.m = m;
        G.n3 = n3;
    }
    if (maxit <= 0) maxit = 3 * c->n_eq;
    const unsigned vb = nblocks(n3), nb = (unsigned)((nnd + GB - 1) / GB);
    const unsigned pb = (unsigned)std::min<long>(GNB, std::max<long>(1, (n3 + GB - 1) / GB));
    auto spmv = [&](const double* in, double* out) {
        TimedRegion tr(c, T_SPMV);
        hipLaunchKernelGGL(k_gm_spmv<MT>, dim3(nb), dim3(GB), 0, c->stream, nnd, c->op.d_slice_base.p,
                           c->op.d_slice_width.p, c->op.d_col.p, val, in, out);
    };
    auto precondition = [&](const double* in, double* out) {
        TimedRegion tr(c, c->precond == TG_PC_AMG ? T_PRECOND : T_VECTOR);
        if (c->precond == TG_PC_AMG)
            amg_vcycle(c, in, out, nullptr);
        else if (c->precond == TG_PC_BLOCK_JACOBI)
            hipLaunchKernelGGL(k_gm_block, dim3(nb), dim3(GB), 0, c->stream, nnd, c->vbinv.p, in, out);
        else
            hipLaunchKernelGGL(k_gm_jacobi, dim3(vb), dim3(GB), 0, c->stream, n3, c->vdinv.p, in, out);
    };
    // V_0..V_{j1-1} . w -> host h[0..j1)
    std::vector<double> hv((size_t)m + 1);
    auto project = [&](int j1, const double* w, double* host) {
        {
            TimedRegion tr(c, T_VECTOR);
            hipLaunchKernelGGL(k_gm_mdot, dim3(pb), dim3(GB), 0, c->stream, n3, G.V.p, j1, w, G.part.p);
            hipLaunchKernelGGL(k_gm_fin, dim3(j1), dim3(GB), 0, c->stream, G.part.p, (int)pb, G.h.p);
        }
        TG_HIP(hipMemcpyAsync(host, G.h.p, sizeof(double) * j1, hipMemcpyDeviceToHost, c->stream));
        TG_HIP(hipStreamSynchronize(c->stream));
    };
    auto norm = [&](const double* v) {
        double r = 0.0;
        {
            TimedRegion tr(c, T_VECTOR);
            hipLaunchKernelGGL(k_gm_mdot, dim3(pb), dim3(GB), 0, c->stream, n3, v, 1, v, G.part.p);
            hipLaunchKernelGGL(k_gm_fin, dim3(1), dim3(GB), 0, c->stream, G.part.p, (int)pb, G.h.p);
        }
        TG_HIP(hipMemcpyAsync(&r, G.h.p, sizeof(double), hipMemcpyDeviceToHost, c->stream));
        TG_HIP(hipStreamSynchronize(c->stream));
        return std::sqrt(r);
    };
    std::vector<double> H((size_t)(m + 1) * m), g((size_t)m + 1), cs((size_t)m), sn((size_t)m), y((size_t)m);
    auto HH = [&](int i, int j) -> double& { return H[(size_t)i * m + j]; };
    TG_HIP(hipMemsetAsync(x, 0, sizeof(double) * n3, c->stream));
    int total = 0, steps = 0;
    bool done = false;
    double beta0 = -1.0, last = 0.0;
    while (!done && total < maxit) {
        // r = M^-1 (b - A x)
        spmv(x, G.q.p);
        {
            TimedRegion tr(c, T_VECTOR);
            hipLaunchKernelGGL(k_gm_bminus, dim3(vb), dim3(GB), 0, c->stream, n3, b, G.q.p);
        }
        precondition(G.q.p, G.w.p);
        const double beta = norm(G.w.p);
        if (beta0 < 0.0) beta0 = beta;
        last = beta;
        if (beta == 0.0) {
            done = true;
            break;
        }
        const double tol_abs = rtol * beta;
        if (tol_abs == 0.0) {
            done = true;
            break;
        }
        hipLaunchKernelGGL(k_gm_scale, dim3(vb), dim3(GB), 0, c->stream, n3, G.w.p, 1.0 / beta, G.V.p);
        std::fill(g.begin(), g.end(), 0.0);
        g[0] = beta;
        int j_end = 0;
        for (int j = 0; j < m && total < maxit; j++, total++) {
            j_end = j;
            steps++;
            spmv(G.V.p + (long)j * n3, G.q.p);
            precondition(G.q.p, G.w.p);
            // two classical Gram-Schmidt passes
            project(j + 1, G.w.p, hv.data());
            for (int i = 0; i <= j; i++) HH(i, j) = hv[i];
            {
                TimedRegion tr(c, T_VECTOR);
                hipLaunchKernelGGL(k_gm_combine, dim3(vb), dim3(GB), 0, c->stream, n3, G.V.p, j + 1, G.h.p, -1.0,
                                   G.w.p);
            }
            project(j + 1, G.w.p, hv.data());
            for (int i = 0; i <= j; i++) HH(i, j) += hv[i];
            {
                TimedRegion tr(c, T_VECTOR);
                hipLaunchKernelGGL(k_gm_combine, dim3(vb), dim3(GB), 0, c->stream, n3, G.V.p, j + 1, G.h.p, -1.0,
                                   G.w.p);
            }
            const double hn = norm(G.w.p);
            HH(j + 1, j) = hn;
            if (hn > 0.0)
                hipLaunchKernelGGL(k_gm_scale, dim3(vb), dim3(GB), 0, c->stream, n3, G.w.p, 1.0 / hn,
                                   G.V.p + (long)(j + 1) * n3);
            for (int i = 0; i < j; i++) {
                const double t = cs[i] * HH(i, j) + sn[i] * HH(i + 1, j);
                HH(i + 1, j) = -sn[i] * HH(i, j) + cs[i] * HH(i + 1, j);
                HH(i, j) = t;
            }
            const double den = std::sqrt(HH(j, j) * HH(j, j) + HH(j + 1, j) * HH(j + 1, j));
            cs[j] = den > 0.0 ? HH(j, j) / den : 1.0;
            sn[j] = den > 0.0 ? HH(j + 1, j) / den : 0.0;
            double t = cs[j] * HH(j, j) + sn[j] * HH(j + 1, j);
            HH(j + 1, j) = -sn[j] * HH(j, j) + cs[j] * HH(j + 1, j);
            HH(j, j) = t;
            t = cs[j] * g[j] + sn[j] * g[j + 1];
            g[j + 1] = -sn[j] * g[j] + cs[j] * g[j + 1];
            g[j] = t;
            last = std::fabs(g[j + 1]);
            if (last < tol_abs) {
                j_end = j;
                done = true;
                break;
            }
        }
        for (int i = j_end; i >= 0; i--) {
            double sum = g[i];
            for (int l = i + 1; l <= j_end; l++) sum -= HH(i, l) * y[l];
            y[i] = std::fabs(HH(i, i)) > 1.0e-14 ? sum / HH(i, i) : 0.0;
        }
        TG_HIP(hipMemcpyAsync(G.h.p, y.data(), sizeof(double) * (j_end + 1), hipMemcpyHostToDevice, c->stream));
        {
            TimedRegion tr(c, T_VECTOR);
            hipLaunchKernelGGL(k_gm_combine, dim3(vb), dim3(GB), 0, c->stream, n3, G.V.p, j_end + 1, G.h.p, 1.0, x);
        }
        TG_HIP(hipStreamSynchronize(c->stream));  // y outlives its copy
    }
    TG_HIP(hipGetLastError());
    if (st) {
        st->iterations = steps;
        st->converged = done ? 1 : 0;
        st->rel_residual = beta0 > 0.0 ? last / beta0 : 0.0;
    }
}

void gmres_solve(tg_ctx* c, const double* b, double* x, double rtol, int maxit, tg_lin_stats* st)
{
    if (c->seg.nseg != 1) fail(TG_BAD_INPUT, "GMRES needs a single problem (no segments)");
    if (mixed_operator(c))
        gmres_t<float>(c, c->op.d_val32.p, b, x, rtol, maxit, st);
    else
        gmres_t<double>(c, c->op.d_val.p, b, x, rtol, maxit, st);
}

}  // namespace tg
