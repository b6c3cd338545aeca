// Aggregation multigrid preconditioner: coarse operators, smoothers and the
// V-cycle on the device (the structure is built by tg_amg.cpp).
//
// Smoother: Chebyshev polynomial in D^-1 A (Adams, Brezina, Hu and
// Tuminaro, J. Comput. Phys. 188 (2003) 593-610; the three-term recurrence
// of Saad, Iterative Methods for Sparse Linear Systems, Alg. 12.1) on
// [lmax / 30, lmax], with lmax a power-iteration estimate of the largest
// eigenvalue of D^-1 A raised by 10 %.  Pre- and post-smoothing use the same
// polynomial and the coarse correction is a Galerkin product, so the V-cycle
// is a symmetric positive definite preconditioner for CG.  Matrices below
// the fine level are FP32 (a preconditioner needs no more); all vectors and
// sums are FP64.
#include <cmath>
#include <cstring>

#include "tg_internal.h"

namespace tg {

namespace {

constexpr int AB = 256;          // threads per block of the vector/level kernels
constexpr double CHEB_RATIO = 30.0;
constexpr double CHEB_BOOST = 1.1;
constexpr int POWER_ITS = 12;

inline unsigned blocks_for(long n) { return (unsigned)((n + AB - 1) / AB); }

__device__ __forceinline__ bool skip(const double* done) { return done && *done != 0.0; }

// Level 0: the solver's sliced-ELL operator (3x3 blocks), one thread per node.
template <class MT>
struct Op0 {
    static constexpr int NC = 3;
    int n;
    const int* sbase;
    const int* swidth;
    const int* col;
    const MT* val;
    __host__ __device__ long units() const { return n; }
    __device__ __forceinline__ void apply(long t, const double* __restrict__ x, double (&y)[3], long (&k)[3]) const
    {
        const int s = (int)(t >> 6), l = (int)(t & 63);
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
        y[0] = y0;
        y[1] = y1;
        y[2] = y2;
        k[0] = 3 * t;
        k[1] = 3 * t + 1;
        k[2] = 3 * t + 2;
    }
};

// Levels >= 1: sliced 6x6 FP32 blocks (slot e of a slice: its 36 values at
// val[(e / 64) * 2304 + comp * 64 + e % 64], coalesced over the slice's 64
// rows), one thread per node and 6 outputs.
struct OpC {
    static constexpr int NC = 6;
    int n;
    const int* sbase;
    const int* swidth;
    const int* col;
    const float* val;
    __host__ __device__ long units() const { return n; }
    __device__ __forceinline__ void apply(long t, const double* __restrict__ x, double (&y)[6], long (&k)[6]) const
    {
        const int s = (int)(t >> 6), l = (int)(t & 63);
        const long b0 = sbase[s];
        const int w = swidth[s];
#pragma unroll
        for (int r = 0; r < 6; r++) y[r] = 0.0;
        for (int q = 0; q < w; q++) {
            const long slot = (b0 + q) * 64 + l;
            const int c = col[slot];
            if (c < 0) continue;
            double xc[6];
#pragma unroll
            for (int j = 0; j < 6; j++) xc[j] = x[6L * c + j];
            const float* v = val + (b0 + q) * 2304 + l;
#pragma unroll
            for (int r = 0; r < 6; r++)
#pragma unroll
                for (int j = 0; j < 6; j++) y[r] += (double)v[(6 * r + j) * 64] * xc[j];
        }
#pragma unroll
        for (int r = 0; r < 6; r++) k[r] = 6 * t + r;
    }
};

// y = A x
template <class Op>
__global__ void __launch_bounds__(AB) k_amg_spmv(Op op, const double* __restrict__ x, double* __restrict__ y)
{
    const long t = (long)blockIdx.x * AB + threadIdx.x;
    if (t >= op.units()) return;
    double q[Op::NC];
    long k[Op::NC];
    op.apply(t, x, q, k);
#pragma unroll
    for (int i = 0; i < Op::NC; i++) y[k[i]] = q[i];
}

// Chebyshev step s >= 1: r -= A d_prev; d_next = c1 d_prev + c2 D^-1 r;
// x += d_next (+ d_prev when the pass started from a nonzero x)
template <class Op>
__global__ void __launch_bounds__(AB) k_amg_step(Op op, const double* __restrict__ dinv,
                                                 const double* __restrict__ dprev, double* __restrict__ dnext,
                                                 double* __restrict__ r, double* __restrict__ x,
                                                 const double* __restrict__ coef, int ci, int add_prev,
                                                 const double* done)
{
    if (skip(done)) return;
    const long t = (long)blockIdx.x * AB + threadIdx.x;
    if (t >= op.units()) return;
    double q[Op::NC];
    long k[Op::NC];
    op.apply(t, dprev, q, k);
    const double c1 = coef[ci], c2 = coef[ci + 1];
#pragma unroll
    for (int i = 0; i < Op::NC; i++) {
        const long j = k[i];
        const double ri = r[j] - q[i];
        r[j] = ri;
        const double dn = c1 * dprev[j] + c2 * dinv[j] * ri;
        dnext[j] = dn;
        x[j] += (add_prev ? dprev[j] : 0.0) + dn;
    }
}

// r -= A d
template <class Op>
__global__ void __launch_bounds__(AB) k_amg_resid(Op op, const double* __restrict__ d, double* __restrict__ r,
                                                  const double* done)
{
    if (skip(done)) return;
    const long t = (long)blockIdx.x * AB + threadIdx.x;
    if (t >= op.units()) return;
    double q[Op::NC];
    long k[Op::NC];
    op.apply(t, d, q, k);
#pragma unroll
    for (int i = 0; i < Op::NC; i++) r[k[i]] -= q[i];
}

// start of post-smoothing: r = f - A x; d0 = D^-1 r / theta
template <class Op>
__global__ void __launch_bounds__(AB) k_amg_post(Op op, const double* __restrict__ dinv, const double* __restrict__ f,
                                                 const double* __restrict__ x, double* __restrict__ r,
                                                 double* __restrict__ d0, const double* __restrict__ coef,
                                                 const double* done)
{
    if (skip(done)) return;
    const long t = (long)blockIdx.x * AB + threadIdx.x;
    if (t >= op.units()) return;
    double q[Op::NC];
    long k[Op::NC];
    op.apply(t, x, q, k);
    const double ith = coef[0];
#pragma unroll
    for (int i = 0; i < Op::NC; i++) {
        const long j = k[i];
        const double ri = f[j] - q[i];
        r[j] = ri;
        d0[j] = ith * dinv[j] * ri;
    }
}

// start of pre-smoothing from x = 0: r = f; d0 = x = D^-1 f / theta
__global__ void __launch_bounds__(AB) k_amg_pre(long nd, const double* __restrict__ dinv, const double* __restrict__ f,
                                                double* __restrict__ r, double* __restrict__ d0,
                                                double* __restrict__ x, const double* __restrict__ coef,
                                                const double* done)
{
    if (skip(done)) return;
    const long j = (long)blockIdx.x * AB + threadIdx.x;
    if (j >= nd) return;
    const double fj = f[j];
    const double d = coef[0] * dinv[j] * fj;
    r[j] = fj;
    d0[j] = d;
    x[j] = d;
}

// fc[A][c] = sum over members i of A of P_i^T r_i
__global__ void __launch_bounds__(AB) k_amg_restrict(int nc, int bs, const int* __restrict__ mptr,
                                                     const int* __restrict__ mem, const float* __restrict__ P,
                                                     const double* __restrict__ r, double* __restrict__ fc,
                                                     const double* done)
{
    if (skip(done)) return;
    const long t = (long)blockIdx.x * AB + threadIdx.x;
    if (t >= 6L * nc) return;
    const int a = (int)(t / 6), c = (int)(t - 6L * a);
    double s = 0.0;
    for (int m = mptr[a]; m < mptr[a + 1]; m++) {
        const int i = mem[m];
        const float* p = P + (long)i * bs * 6 + c;
        const double* ri = r + (long)i * bs;
        for (int q = 0; q < bs; q++) s += (double)p[6 * q] * ri[q];
    }
    fc[t] = s;
}

// x_i += w P_i xc[agg(i)]
__global__ void __launch_bounds__(AB) k_amg_prolong(int n, int bs, const int* __restrict__ agg,
                                                    const float* __restrict__ P, const double* __restrict__ xc,
                                                    double w, double* __restrict__ x, const double* done)
{
    if (skip(done)) return;
    const long t = (long)blockIdx.x * AB + threadIdx.x;
    if (t >= (long)n * bs) return;
    const int i = (int)(t / bs), q = (int)(t - (long)i * bs);
    const int a = agg[i];
    if (a < 0) return;
    const float* p = P + (long)i * bs * 6 + 6 * q;
    const double* e = xc + 6L * a;
    x[t] += w * ((double)p[0] * e[0] + (double)p[1] * e[1] + (double)p[2] * e[2] + (double)p[3] * e[3] +
                 (double)p[4] * e[4] + (double)p[5] * e[5]);
}

// coarsest level: x = C^-1 f (C^-1 dense, symmetric, column-major)
__global__ void __launch_bounds__(AB) k_amg_dense(int n, const double* __restrict__ cinv, const double* __restrict__ f,
                                                  double* __restrict__ x, const double* done)
{
    if (skip(done)) return;
    const int r = blockIdx.x * AB + threadIdx.x;
    if (r >= n) return;
    double s = 0.0;
    for (int c = 0; c < n; c++) s += cinv[(long)c * n + r] * f[c];
    x[r] = s;
}

// Galerkin product, fine level = the sliced-ELL operator: one wave per
// upper coarse block, lane (r, c) < 36 sums P_i^T A_ij P_j over the block's
// gather list in order; the block and its transpose partner are written.
template <class MT>
__global__ void __launch_bounds__(AB) k_amg_rap0(long nup, const int* __restrict__ up, const long* __restrict__ gptr,
                                                 const int* __restrict__ gl, const int* __restrict__ fcol,
                                                 const MT* __restrict__ fval, const float* __restrict__ P,
                                                 float* __restrict__ cval)
{
    const long u = (long)blockIdx.x * (AB / 64) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (u >= nup || lane >= 36) return;
    const int r = lane / 6, c = lane - 6 * (lane / 6);
    double s = 0.0;
    for (long e = gptr[u]; e < gptr[u + 1]; e++) {
        const int fb = gl[2 * e], i = gl[2 * e + 1];
        const int j = fcol[fb];
        const MT* v = fval + (long)(fb >> 6) * 576 + (fb & 63);
        const float* pi = P + (long)i * 18 + r;
        const float* pj = P + (long)j * 18 + c;
        double t = 0.0;
#pragma unroll
        for (int p = 0; p < 3; p++) {
            const double ap = (double)v[(3 * p) * 64] * pj[0] + (double)v[(3 * p + 1) * 64] * pj[6] +
                              (double)v[(3 * p + 2) * 64] * pj[12];
            t += (double)pi[6 * p] * ap;
        }
        s += t;
    }
    const int ab = up[2 * u], ba = up[2 * u + 1];
    cval[(long)(ab >> 6) * 2304 + (ab & 63) + (6 * r + c) * 64] = (float)s;
    if (ba >= 0) cval[(long)(ba >> 6) * 2304 + (ba & 63) + (6 * c + r) * 64] = (float)s;
}

// Galerkin product, fine level = a sliced 6x6 level
__global__ void __launch_bounds__(AB) k_amg_rapc(long nup, const int* __restrict__ up, const long* __restrict__ gptr,
                                                 const int* __restrict__ gl, const int* __restrict__ fcol,
                                                 const float* __restrict__ fval, const float* __restrict__ P,
                                                 float* __restrict__ cval)
{
    const long u = (long)blockIdx.x * (AB / 64) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (u >= nup || lane >= 36) return;
    const int r = lane / 6, c = lane - 6 * (lane / 6);
    double s = 0.0;
    for (long e = gptr[u]; e < gptr[u + 1]; e++) {
        const int fb = gl[2 * e], i = gl[2 * e + 1];
        const int j = fcol[fb];
        const float* v = fval + (long)(fb >> 6) * 2304 + (fb & 63);
        const float* pi = P + (long)i * 36 + r;
        const float* pj = P + (long)j * 36 + c;
        double t = 0.0;
#pragma unroll
        for (int p = 0; p < 6; p++) {
            double ap = 0.0;
#pragma unroll
            for (int q = 0; q < 6; q++) ap += (double)v[(6 * p + q) * 64] * pj[6 * q];
            t += (double)pi[6 * p] * ap;
        }
        s += t;
    }
    const int ab = up[2 * u], ba = up[2 * u + 1];
    cval[(long)(ab >> 6) * 2304 + (ab & 63) + (6 * r + c) * 64] = (float)s;
    if (ba >= 0) cval[(long)(ba >> 6) * 2304 + (ba & 63) + (6 * c + r) * 64] = (float)s;
}

// inverse diagonal of a sliced 6x6 level (0 for inactive dofs)
__global__ void __launch_bounds__(AB) k_amg_dinv(int n, const int* __restrict__ diagblk, const float* __restrict__ val,
                                                 double* __restrict__ dinv)
{
    const long t = (long)blockIdx.x * AB + threadIdx.x;
    if (t >= 6L * n) return;
    const int a = (int)(t / 6), r = (int)(t - 6L * a);
    const int e = diagblk[a];
    const double d = (double)val[(long)(e >> 6) * 2304 + (e & 63) + 7 * r * 64];
    dinv[t] = d > 0.0 ? 1.0 / d : 0.0;
}

// ---- power iteration for the largest eigenvalue of D^-1 A -----------------

__global__ void __launch_bounds__(AB) k_amg_rand(long nd, const double* __restrict__ dinv, double* __restrict__ v)
{
    const long j = (long)blockIdx.x * AB + threadIdx.x;
    if (j >= nd) return;
    // SplitMix64 of the index: a fixed start vector with no structure
    unsigned long long z = (unsigned long long)j + 0x9e3779b97f4a7c15ULL;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    z ^= z >> 31;
    const double u = (double)(z >> 11) * (1.0 / 9007199254740992.0);
    v[j] = dinv[j] != 0.0 ? 2.0 * u - 1.0 : 0.0;
}

// part[b] = partial sums of a.b (KIND 0) or a.a / dinv over dinv != 0 (KIND 1)
template <int KIND>
__global__ void __launch_bounds__(AB) k_amg_part(long nd, const double* __restrict__ a, const double* __restrict__ b,
                                                 double* __restrict__ part)
{
    __shared__ double red[AB / 64];
    double s = 0.0;
    for (long j = (long)blockIdx.x * AB + threadIdx.x; j < nd; j += (long)gridDim.x * AB) {
        if (KIND == 0)
            s += a[j] * b[j];
        else if (b[j] != 0.0)
            s += a[j] * a[j] / b[j];
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < AB / 64; w++) t += red[w];
        part[blockIdx.x] = t;
    }
}

__global__ void __launch_bounds__(AB) k_amg_fin(const double* __restrict__ part, int np, double* __restrict__ out)
{
    __shared__ double red[AB / 64];
    double s = 0.0;
    for (int j = threadIdx.x; j < np; j += AB) s += part[j];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < AB / 64; w++) t += red[w];
        *out = t;
    }
}

// v = D^-1 w / sqrt(nrm2) with nrm2 = |D^-1 w|^2 computed beforehand into *nrm2
__global__ void __launch_bounds__(AB) k_amg_scale(long nd, const double* __restrict__ w,
                                                  const double* __restrict__ nrm2, double* __restrict__ v)
{
    const long j = (long)blockIdx.x * AB + threadIdx.x;
    if (j >= nd) return;
    const double s = *nrm2 > 0.0 ? 1.0 / sqrt(*nrm2) : 0.0;
    v[j] = w[j] * s;
}

// w = D^-1 w
__global__ void __launch_bounds__(AB) k_amg_jac(long nd, const double* __restrict__ dinv, double* __restrict__ w)
{
    const long j = (long)blockIdx.x * AB + threadIdx.x;
    if (j >= nd) return;
    w[j] *= dinv[j];
}

constexpr int NPART = 1024;

}  // namespace

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

template <class MT>
static Op0<MT> op0(tg_ctx* c, const MT* val)
{
    return Op0<MT>{c->nnd, c->op.d_slice_base.p, c->op.d_slice_width.p, c->op.d_col.p, val};
}

static OpC opc(const AmgLevel& L) { return OpC{L.n, L.d_sbase.p, L.d_swidth.p, L.d_col.p, L.d_val.p}; }

static double* level_dinv(tg_ctx* c, int l) { return l == 0 ? c->vdinv.p : c->amg.lv[l]->dinv; }

static void reduce(tg_ctx* c, int kind, long nd, const double* a, const double* b, double* out)
{
    Amg& G = c->amg;
    const unsigned nb = std::min<long>(NPART, std::max<long>(1, blocks_for(nd)));
    if (kind == 0)
        hipLaunchKernelGGL(k_amg_part<0>, dim3(nb), dim3(AB), 0, c->stream, nd, a, b, G.d_part.p);
    else
        hipLaunchKernelGGL(k_amg_part<1>, dim3(nb), dim3(AB), 0, c->stream, nd, a, b, G.d_part.p);
    hipLaunchKernelGGL(k_amg_fin, dim3(1), dim3(AB), 0, c->stream, G.d_part.p, (int)nb, out);
}

// Largest eigenvalue of D^-1 A on level l by power iteration: the Rayleigh
// quotient (v.Av)/(v.Dv) is left in red[0] / red[1].
template <class Op>
static void power_iteration(tg_ctx* c, const Op& op, long nd, const double* dinv, double* v, double* w, double* red)
{
    const unsigned nb = blocks_for(op.units()), vb = blocks_for(nd);
    hipLaunchKernelGGL(k_amg_rand, dim3(vb), dim3(AB), 0, c->stream, nd, dinv, v);
    for (int it = 0; it < POWER_ITS; it++) {
        hipLaunchKernelGGL(k_amg_spmv<Op>, dim3(nb), dim3(AB), 0, c->stream, op, v, w);
        hipLaunchKernelGGL(k_amg_jac, dim3(vb), dim3(AB), 0, c->stream, nd, dinv, w);
        reduce(c, 0, nd, w, w, red + 2);
        hipLaunchKernelGGL(k_amg_scale, dim3(vb), dim3(AB), 0, c->stream, nd, w, red + 2, v);
    }
    hipLaunchKernelGGL(k_amg_spmv<Op>, dim3(nb), dim3(AB), 0, c->stream, op, v, w);
    reduce(c, 0, nd, v, w, red);
    reduce(c, 1, nd, v, dinv, red + 1);
}

// Dense inverse of the coarsest level by LDL^T: inactive dofs (zero
// diagonal) and non-positive pivots get zero rows and columns.
static void coarse_inverse(const AmgLevel& L, const std::vector<float>& val, std::vector<double>& inv)
{
    const int n = 6 * L.n;
    std::vector<double> A((size_t)n * n, 0.0);
    for (int a = 0; a < L.n; a++)
        for (int q = L.rowptr[a]; q < L.rowptr[a + 1]; q++) {
            const int b = L.col[q], e = L.slot[q];
            const size_t v0 = (size_t)(e >> 6) * 2304 + (e & 63);
            for (int r = 0; r < 6; r++)
                for (int s = 0; s < 6; s++) A[(size_t)(6 * a + r) * n + 6 * b + s] = val[v0 + (6 * r + s) * 64];
        }
    double dmax = 0.0;
    for (int i = 0; i < n; i++) dmax = std::max(dmax, A[(size_t)i * n + i]);
    // in-place LDL^T (lower triangle holds L, diagonal holds D)
    std::vector<char> live((size_t)n, 1);
    for (int j = 0; j < n; j++) {
        double d = A[(size_t)j * n + j];
        for (int k = 0; k < j; k++)
            if (live[k]) d -= A[(size_t)j * n + k] * A[(size_t)j * n + k] * A[(size_t)k * n + k];
        if (!(d > 1e-12 * dmax)) {
            live[j] = 0;
            A[(size_t)j * n + j] = 0.0;
            for (int i = j + 1; i < n; i++) A[(size_t)i * n + j] = 0.0;
            continue;
        }
        A[(size_t)j * n + j] = d;
        for (int i = j + 1; i < n; i++) {
            double s = A[(size_t)i * n + j];
            for (int k = 0; k < j; k++)
                if (live[k]) s -= A[(size_t)i * n + k] * A[(size_t)j * n + k] * A[(size_t)k * n + k];
            A[(size_t)i * n + j] = s / d;
        }
    }
    inv.assign((size_t)n * n, 0.0);
    std::vector<double> y((size_t)n);
    for (int e = 0; e < n; e++) {
        if (!live[e]) continue;
        std::fill(y.begin(), y.end(), 0.0);
        y[e] = 1.0;
        for (int i = 0; i < n; i++) {  // L y = e
            if (!live[i]) {
                y[i] = 0.0;
                continue;
            }
            double s = y[i];
            for (int k = 0; k < i; k++) s -= A[(size_t)i * n + k] * y[k];
            y[i] = s;
        }
        for (int i = 0; i < n; i++) y[i] = live[i] ? y[i] / A[(size_t)i * n + i] : 0.0;
        for (int i = n - 1; i >= 0; i--) {  // L^T x = y
            if (!live[i]) continue;
            double s = y[i];
            for (int k = i + 1; k < n; k++) s -= A[(size_t)k * n + i] * y[k];
            y[i] = s;
        }
        for (int i = 0; i < n; i++) inv[(size_t)e * n + i] = live[i] ? y[i] : 0.0;
    }
}

template <class MT>
static void amg_numeric_t(tg_ctx* c, const MT* val0)
{
    Amg& G = c->amg;
    const int L = (int)G.lv.size();
    TimedRegion tr(c, T_PRECOND);
    // coarse operators, finest first
    for (int l = 0; l + 1 < L; l++) {
        AmgLevel& F = *G.lv[l];
        AmgLevel& C = *G.lv[l + 1];
        const unsigned nb = (unsigned)((F.nup + AB / 64 - 1) / (AB / 64));
        if (nb > 0) {
            if (l == 0)
                hipLaunchKernelGGL(k_amg_rap0<MT>, dim3(nb), dim3(AB), 0, c->stream, F.nup, F.d_up.p, F.d_gptr.p,
                                   F.d_glist.p, c->op.d_col.p, val0, F.d_P.p, C.d_val.p);
            else
                hipLaunchKernelGGL(k_amg_rapc, dim3(nb), dim3(AB), 0, c->stream, F.nup, F.d_up.p, F.d_gptr.p,
                                   F.d_glist.p, F.d_col.p, F.d_val.p, F.d_P.p, C.d_val.p);
        }
        hipLaunchKernelGGL(k_amg_dinv, dim3(blocks_for(6L * C.n)), dim3(AB), 0, c->stream, C.n, C.d_diagblk.p,
                           C.d_val.p, C.dinv);
        TG_HIP(hipGetLastError());
    }
    // smoother intervals
    for (int l = 0; l + 1 < L; l++) {
        AmgLevel& A = *G.lv[l];
        const long nd = (long)A.n * A.bs;
        double* red = G.d_red.p + 16 * l;
        if (l == 0)
            power_iteration(c, op0<MT>(c, val0), nd, level_dinv(c, 0), A.d_d0.p, A.d_d1.p, red);
        else
            power_iteration(c, opc(A), nd, level_dinv(c, l), A.d_d0.p, A.d_d1.p, red);
        TG_HIP(hipGetLastError());
    }
    std::vector<double> red(16 * (size_t)L);
    TG_HIP(hipMemcpyAsync(red.data(), G.d_red.p, sizeof(double) * red.size(), hipMemcpyDeviceToHost, c->stream));
    const AmgLevel& Cst = *G.lv[L - 1];
    std::vector<float> cval((size_t)Cst.nslots * 36);
    TG_HIP(hipMemcpyAsync(cval.data(), Cst.d_val.p, sizeof(float) * cval.size(), hipMemcpyDeviceToHost, c->stream));
    TG_HIP(hipStreamSynchronize(c->stream));
    std::vector<double> coef(8 * (size_t)L, 0.0);
    for (int l = 0; l + 1 < L; l++) {
        const double num = red[16 * l], den = red[16 * l + 1];
        double lmax = den > 0.0 ? num / den : 1.0;
        if (!(lmax > 0.0) || !std::isfinite(lmax)) lmax = 1.0;
        lmax *= CHEB_BOOST;
        G.lv[l]->lmax = lmax;
        const double lmin = lmax / CHEB_RATIO;
        const double theta = 0.5 * (lmax + lmin), delta = 0.5 * (lmax - lmin), sigma = theta / delta;
        double* cf = coef.data() + 8 * l;
        cf[0] = 1.0 / theta;
        double rho = 1.0 / sigma;
        for (int s = 1; s < G.degree; s++) {
            const double rn = 1.0 / (2.0 * sigma - rho);
            cf[2 * s - 1] = rn * rho;
            cf[2 * s] = 2.0 * rn / delta;
            rho = rn;
        }
    }
    std::vector<double> inv;
    coarse_inverse(Cst, cval, inv);
    TG_HIP(hipMemcpyAsync(G.d_coef.p, coef.data(), sizeof(double) * coef.size(), hipMemcpyHostToDevice, c->stream));
    TG_HIP(hipMemcpyAsync(G.d_cinv.p, inv.data(), sizeof(double) * inv.size(), hipMemcpyHostToDevice, c->stream));
    TG_HIP(hipStreamSynchronize(c->stream));  // host vectors outlive their copies
}

// The multigrid works on the FP32 copy of the fine operator in either
// precision (the solver's own operator when it is FP32; refreshed from the
// FP64 one by the caller otherwise): a preconditioner needs no more, and the
// smoother's four fine-level products per cycle move half the bytes.
// The coarse levels' numeric set-up (Galerkin products, smoother intervals,
// coarsest inverse) for the current fine operator.  With reuse = k > 0 one
// set-up serves the next k tangents of the same Newton solve too: the
// V-cycle then smooths with the current fine operator and corrects with the
// coarse operators of an earlier tangent of that solve (Galerkin products of
// an SPD matrix, so the V-cycle stays an SPD preconditioner; only the PCG
// count may change).  A new load, state or step (tg_apply_kbc,
// tg_set_displacement, tg_close_step, tg_reset_step) clears `since`: the
// tangents of a failed attempt never serve the next one.
void amg_numeric(tg_ctx* c)
{
    Amg& G = c->amg;
    if (G.lv.empty()) amg_setup(c);
    if (G.reuse > 0 && G.since >= 0 && G.since < G.reuse) {
        G.since++;
        return;
    }
    amg_numeric_t<float>(c, c->op.d_val32.p);
    G.since = 0;
}

template <class Op>
static void smooth_down(tg_ctx* c, const Op& op, long nd, const double* dinv, const double* f, double* x, AmgLevel& A,
                        const double* cf, const double* done)
{
    const unsigned nb = blocks_for(op.units());
    hipLaunchKernelGGL(k_amg_pre, dim3(blocks_for(nd)), dim3(AB), 0, c->stream, nd, dinv, f, A.d_r.p, A.d_d0.p, x, cf,
                       done);
    double *dp = A.d_d0.p, *dn = A.d_d1.p;
    for (int s = 1; s < c->amg.degree; s++) {
        hipLaunchKernelGGL(k_amg_step<Op>, dim3(nb), dim3(AB), 0, c->stream, op, dinv, dp, dn, A.d_r.p, x, cf,
                           2 * s - 1, 0, done);
        std::swap(dp, dn);
    }
    hipLaunchKernelGGL(k_amg_resid<Op>, dim3(nb), dim3(AB), 0, c->stream, op, dp, A.d_r.p, done);
}

template <class Op>
static void smooth_up(tg_ctx* c, const Op& op, const double* dinv, const double* f, double* x, AmgLevel& A,
                      const double* cf, const double* done)
{
    const unsigned nb = blocks_for(op.units());
    hipLaunchKernelGGL(k_amg_post<Op>, dim3(nb), dim3(AB), 0, c->stream, op, dinv, f, x, A.d_r.p, A.d_d0.p, cf, done);
    double *dp = A.d_d0.p, *dn = A.d_d1.p;
    for (int s = 1; s < c->amg.degree; s++) {
        hipLaunchKernelGGL(k_amg_step<Op>, dim3(nb), dim3(AB), 0, c->stream, op, dinv, dp, dn, A.d_r.p, x, cf,
                           2 * s - 1, s == 1 ? 1 : 0, done);
        std::swap(dp, dn);
    }
}

// z = M^-1 f by one V-cycle; every kernel is a no-op once *done != 0 (the
// solve has converged), so a captured chunk of iterations stays valid.
void amg_vcycle(tg_ctx* c, const double* f0, double* z, const double* done)
{
    using MT = float;
    const MT* val0 = c->op.d_val32.p;
    Amg& G = c->amg;
    const int L = (int)G.lv.size();
    for (int l = 0; l + 1 < L; l++) {
        AmgLevel& A = *G.lv[l];
        AmgLevel& C = *G.lv[l + 1];
        const double* f = l == 0 ? f0 : A.d_f.p;
        double* x = l == 0 ? z : A.d_x.p;
        const double* cf = G.d_coef.p + 8 * l;
        const long nd = (long)A.n * A.bs;
        if (l == 0)
            smooth_down(c, op0<MT>(c, val0), nd, level_dinv(c, 0), f, x, A, cf, done);
        else
            smooth_down(c, opc(A), nd, level_dinv(c, l), f, x, A, cf, done);
        hipLaunchKernelGGL(k_amg_restrict, dim3(blocks_for(6L * C.n)), dim3(AB), 0, c->stream, C.n, A.bs, A.d_mptr.p,
                           A.d_mem.p, A.d_P.p, A.d_r.p, C.d_f.p, done);
    }
    AmgLevel& Cst = *G.lv[L - 1];
    hipLaunchKernelGGL(k_amg_dense, dim3(blocks_for(G.ncd)), dim3(AB), 0, c->stream, G.ncd, G.d_cinv.p, Cst.d_f.p,
                       Cst.d_x.p, done);
    for (int l = L - 2; l >= 0; l--) {
        AmgLevel& A = *G.lv[l];
        AmgLevel& C = *G.lv[l + 1];
        const double* f = l == 0 ? f0 : A.d_f.p;
        double* x = l == 0 ? z : A.d_x.p;
        const double* cf = G.d_coef.p + 8 * l;
        hipLaunchKernelGGL(k_amg_prolong, dim3(blocks_for((long)A.n * A.bs)), dim3(AB), 0, c->stream, A.n, A.bs,
                           A.d_agg.p, A.d_P.p, C.d_x.p, 0.01 * G.weight, x, done);
        if (l == 0)
            smooth_up(c, op0<MT>(c, val0), level_dinv(c, 0), f, x, A, cf, done);
        else
            smooth_up(c, opc(A), level_dinv(c, l), f, x, A, cf, done);
    }
    TG_HIP(hipGetLastError());
}


}  // namespace tg
