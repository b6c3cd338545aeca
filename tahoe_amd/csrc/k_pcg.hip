// Segmented Jacobi-preconditioned conjugate gradients.
//
// A context holds one problem or a batch of independent problems packed as
// a disjoint union ("segments"), each starting at a multiple of 64 nodes.
// Every segment runs its own CG: its own dot products, alpha, beta and
// stopping test, all kept on the device.  Reductions are per slice of 64
// nodes (one wave, never shared by two segments) and then per segment in
// slice order, so a segment's arithmetic -- and its result -- is the same
// alone and inside any batch.  A single problem is the one-segment case.
//
// This file is compiled without FMA contraction (tahoe_amd/build.py): each
// expression rounds as written, so the iteration's variants below, which
// share their expressions, agree bit for bit.
//
// The reference has no linear CG; this is the textbook method (Saad,
// Iterative Methods for Sparse Linear Systems, Alg. 9.1) on the reference's
// building blocks: Multx (MSRMatrixT.cpp:385-420), Jacobi scaling
// (DiagonalMatrixT.cpp:267-324) and InnerProduct (SolverT.cpp:854-860).
#include <cstring>

#include <type_traits>

#include "tg_internal.h"

namespace tg {

constexpr int PB = 256;  // threads (= nodes) per block, 4 slices
constexpr int FB = 1024; // threads of the per-segment finalize blocks

__device__ __forceinline__ double wsum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

__device__ __forceinline__ double* sf(double* s, int field, int nseg, int seg) { return s + (long)field * nseg + seg; }

// z = M^-1 r at node a.  PC = 0: Jacobi, m = 1/diag [nnd][3]
// (DiagonalMatrixT); PC = 1: 3x3 nodal block Jacobi, m = inverse diagonal
// blocks [nnd][9] (row-major, prescribed rows and columns zero); PC = 2:
// multigrid, m = z itself, computed beforehand by a V-cycle.
template <int PC>
__device__ __forceinline__ void precondition(const double* __restrict__ m, long a, const double (&r)[3], double (&z)[3])
{
    if (PC == 2) {
#pragma unroll
        for (int i = 0; i < 3; i++) z[i] = m[3 * a + i];
    } else if (PC == 0) {
#pragma unroll
        for (int i = 0; i < 3; i++) z[i] = m[3 * a + i] * r[i];
    } else {
        const double* b = m + 9 * a;
#pragma unroll
        for (int i = 0; i < 3; i++) z[i] = b[3 * i] * r[0] + b[3 * i + 1] * r[1] + b[3 * i + 2] * r[2];
    }
}

// ---- per-slice partials -> per-segment totals ------------------------------

// Sum of part[k][sb..se) by one block of FB threads: thread i adds slices
// sb+i, sb+i+FB, ... in order, then a fixed tree (waves by butterfly, then
// the FB/64 wave sums in order).  The order depends only on the segment's
// own slice count, never on other segments.  The result is in thread 0.
template <int NV>
__device__ __forceinline__ void seg_total(const double* part, long nslices, int sb, int se, double (&tot)[NV])
{
    __shared__ double red[NV][FB / 64];
    const int t = threadIdx.x;
#pragma unroll
    for (int k = 0; k < NV; k++) {
        double x = 0.0;
        for (int s = sb + t; s < se; s += FB) x += part[k * nslices + s];
        x = wsum(x);
        if ((t & 63) == 0) red[k][t >> 6] = x;
    }
    __syncthreads();
    if (t == 0)
#pragma unroll
        for (int k = 0; k < NV; k++) {
            double x = 0.0;
            for (int w = 0; w < FB / 64; w++) x += red[k][w];
            tot[k] = x;
        }
}

// x = 0, r = b, p = M^-1 b; per segment rz = b.M^-1 b, bb = b.b
template <int PC>
__global__ void __launch_bounds__(PB) k_seg_init(int nnd, const double* __restrict__ b,
                                                 const double* __restrict__ m, double* __restrict__ x,
                                                 double* __restrict__ r, double* __restrict__ p,
                                                 double* __restrict__ part, long nslices)
{
    const int a = blockIdx.x * PB + threadIdx.x;
    double v0 = 0.0, v1 = 0.0;
    if (a < nnd) {
        double bi[3], zi[3];
#pragma unroll
        for (int i = 0; i < 3; i++) bi[i] = b[3L * a + i];
        precondition<PC>(m, a, bi, zi);
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const long k = 3L * a + i;
            x[k] = 0.0;
            r[k] = bi[i];
            p[k] = zi[i];
            v0 += bi[i] * zi[i];
            v1 += bi[i] * bi[i];
        }
    }
    v0 = wsum(v0);
    v1 = wsum(v1);
    if ((threadIdx.x & 63) == 0) {
        const long s = a >> 6;
        if (s < nslices) {
            part[s] = v0;
            part[nslices + s] = v1;
        }
    }
}

__global__ void __launch_bounds__(FB) k_seg_fin_init(const double* __restrict__ part, long nslices,
                                                     const int* __restrict__ seg_slice, const int* __restrict__ active,
                                                     int nseg, double* __restrict__ sc, double rtol, int maxit,
                                                     int* __restrict__ ndone)
{
    const int g = blockIdx.x;
    double tot[2];
    seg_total<2>(part, nslices, seg_slice[g], seg_slice[g + 1], tot);
    if (threadIdx.x == 0) {
        *sf(sc, S_RZ, nseg, g) = tot[0];
        *sf(sc, S_BB, nseg, g) = tot[1];
        *sf(sc, S_RR, nseg, g) = tot[1];
        *sf(sc, S_ITER, nseg, g) = 0.0;
        *sf(sc, S_BETA, nseg, g) = 0.0;  // the fused iteration forms p_0 = z_0 + 0 p
        *sf(sc, S_THRESH, nseg, g) = rtol * rtol * tot[1];
        *sf(sc, S_MAXIT, nseg, g) = (double)maxit;
        const bool done = !active[g] || tot[1] == 0.0 || maxit <= 0;
        *sf(sc, S_DONE, nseg, g) = done ? 1.0 : 0.0;
        if (done) atomicAdd(ndone, 1);
    }
}

// Workgroup -> rows.  Workgroups are dealt to the 8 XCDs round-robin, and
// each XCD has its own L2; with XCD = true workgroup w takes the row range
// (w % 8) * (G / 8) + w / 8 (G a multiple of 8), so each XCD sweeps one
// contiguous eighth of the rows and the x entries its rows gather (the
// neighbouring node layers) stay in its L2.
template <bool XCD>
__device__ __forceinline__ int row_block()
{
    if (!XCD) return blockIdx.x;
    const int per = gridDim.x >> 3;
    return (blockIdx.x & 7) * per + (blockIdx.x >> 3);
}

// The 9 values of block slot blk of lane l, row-major (v[3 r + q]).  With
// SYM a lower block (column node below the row node) is read as the
// transpose of its upper partner, whose first value index tpos holds: the
// operator is symmetric by construction (a lower block is a copy of its
// partner), so the values are the same, while the partner's row, just
// streamed by a nearby workgroup, serves them from L2 or the last-level
// cache instead of HBM.  Only the lanes of lower blocks read tpos.
template <class MT, bool SYM>
__device__ __forceinline__ void block_values(const MT* __restrict__ bval, const int* __restrict__ tpos, long blk,
                                             int l, bool lower, double v[9])
{
    long base = blk * 9 * 64 + l;
    const bool tr = SYM && lower;
    if (tr) base = tpos[blk * 64 + l];
    double u[9];
#pragma unroll
    for (int m = 0; m < 9; m++) u[m] = (double)bval[base + m * 64];
    v[0] = u[0];
    v[4] = u[4];
    v[8] = u[8];
    v[1] = tr ? u[3] : u[1];
    v[3] = tr ? u[1] : u[3];
    v[2] = tr ? u[6] : u[2];
    v[6] = tr ? u[2] : u[6];
    v[5] = tr ? u[7] : u[5];
    v[7] = tr ? u[5] : u[7];
}

// q = A p on the block operator, with per-slice partials of p.q.  MT is the
// storage type of the operator values (double, or float in mixed
// precision); products and sums are FP64.
template <class MT, bool XCD, bool UNR, bool SYM>
__global__ void __launch_bounds__(PB) k_seg_spmv(int nnd, const int* __restrict__ slice_seg,
                                                 const int* __restrict__ sbase, const int* __restrict__ swidth,
                                                 const int* __restrict__ bcol, const MT* __restrict__ bval,
                                                 const int* __restrict__ tpos, const double* __restrict__ x,
                                                 double* __restrict__ y,
                                                 double* __restrict__ part, long nslices, int nseg,
                                                 const double* __restrict__ sc)
{
    const int a = row_block<XCD>() * PB + threadIdx.x;
    const int s = a >> 6, l = a & 63;
    if (a >= 64L * nslices) return;
    const int g = slice_seg[s];  // wave-uniform
    if (*sf((double*)sc, S_DONE, nseg, g) != 0.0) return;
    double dv = 0.0;
    if (a < nnd) {
        const long b0 = sbase[s];
        const int w = swidth[s];
        double y0 = 0, y1 = 0, y2 = 0;
        if (UNR) {
            // padding slots (column -1) hold zero values: they multiply the
            // row's own x, which keeps the loop branch-free so its loads can
            // be issued ahead; adding the zero products leaves the sums
            // unchanged
#pragma unroll 3
            for (int k = 0; k < w; k++) {
                const long blk = b0 + k;
                const int cr = bcol[blk * 64 + l];
                const int c = cr < 0 ? a : cr;
                const double x0 = x[3L * c], x1 = x[3L * c + 1], x2 = x[3L * c + 2];
                double v[9];
                block_values<MT, SYM>(bval, tpos, blk, l, c < a, v);
                y0 += v[0] * x0 + v[1] * x1 + v[2] * x2;
                y1 += v[3] * x0 + v[4] * x1 + v[5] * x2;
                y2 += v[6] * x0 + v[7] * x1 + v[8] * x2;
            }
        } else {
            for (int k = 0; k < w; k++) {
                const long blk = b0 + k;
                const int c = bcol[blk * 64 + l];
                if (c < 0) continue;
                const double x0 = x[3L * c], x1 = x[3L * c + 1], x2 = x[3L * c + 2];
                double v[9];
                block_values<MT, SYM>(bval, tpos, blk, l, c < a, v);
                y0 += v[0] * x0 + v[1] * x1 + v[2] * x2;
                y1 += v[3] * x0 + v[4] * x1 + v[5] * x2;
                y2 += v[6] * x0 + v[7] * x1 + v[8] * x2;
            }
        }
        y[3L * a] = y0;
        y[3L * a + 1] = y1;
        y[3L * a + 2] = y2;
        dv = y0 * x[3L * a] + y1 * x[3L * a + 1] + y2 * x[3L * a + 2];
    }
    dv = wsum(dv);
    if (l == 0) part[s] = dv;
}

__global__ void __launch_bounds__(FB) k_seg_fin_alpha(const double* __restrict__ part, long nslices,
                                                      const int* __restrict__ seg_slice, int nseg,
                                                      double* __restrict__ sc)
{
    const int g = blockIdx.x;
    if (*sf(sc, S_DONE, nseg, g) != 0.0) return;
    double tot[1];
    seg_total<1>(part, nslices, seg_slice[g], seg_slice[g + 1], tot);
    if (threadIdx.x == 0) {
        *sf(sc, S_PQ, nseg, g) = tot[0];
        *sf(sc, S_ALPHA, nseg, g) = *sf(sc, S_RZ, nseg, g) / tot[0];
    }
}

// x += alpha p, r -= alpha q; per-slice partials of r.M^-1 r and r.r (with
// the multigrid, r.r only: r.z follows the V-cycle)
template <int PC>
__global__ void __launch_bounds__(PB) k_seg_update(int nnd, const int* __restrict__ slice_seg,
                                                   const double* __restrict__ p, const double* __restrict__ q,
                                                   const double* __restrict__ m, double* __restrict__ x,
                                                   double* __restrict__ r, double* __restrict__ part, long nslices,
                                                   int nseg, const double* __restrict__ sc)
{
    const int a = blockIdx.x * PB + threadIdx.x;
    const int s = a >> 6;
    if (a >= 64L * nslices) return;
    const int g = slice_seg[s];
    if (*sf((double*)sc, S_DONE, nseg, g) != 0.0) return;
    const double alpha = *sf((double*)sc, S_ALPHA, nseg, g);
    double v0 = 0.0, v1 = 0.0;
    if (a < nnd) {
        double ri[3], zi[3];
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const long k = 3L * a + i;
            x[k] += alpha * p[k];
            ri[i] = r[k] - alpha * q[k];
            r[k] = ri[i];
        }
        if (PC != 2) precondition<PC>(m, a, ri, zi);
#pragma unroll
        for (int i = 0; i < 3; i++) {
            if (PC != 2) v0 += ri[i] * zi[i];
            v1 += ri[i] * ri[i];
        }
    }
    v0 = wsum(v0);
    v1 = wsum(v1);
    if ((a & 63) == 0) {
        if (PC != 2) part[s] = v0;
        part[nslices + s] = v1;
    }
}

// beta = rz'/rz; stop on ||r||^2 <= rtol^2 ||b||^2 or the iteration cap
__global__ void __launch_bounds__(FB) k_seg_fin_beta(const double* __restrict__ part, long nslices,
                                                     const int* __restrict__ seg_slice, int nseg,
                                                     double* __restrict__ sc, int* __restrict__ ndone)
{
    const int g = blockIdx.x;
    if (*sf(sc, S_DONE, nseg, g) != 0.0) return;
    double tot[2];
    seg_total<2>(part, nslices, seg_slice[g], seg_slice[g + 1], tot);
    if (threadIdx.x == 0) {
        const double it = *sf(sc, S_ITER, nseg, g) + 1.0;
        *sf(sc, S_ITER, nseg, g) = it;
        *sf(sc, S_BETA, nseg, g) = tot[0] / *sf(sc, S_RZ, nseg, g);
        *sf(sc, S_RZ, nseg, g) = tot[0];
        *sf(sc, S_RR, nseg, g) = tot[1];
        if (tot[1] <= *sf(sc, S_THRESH, nseg, g) || it >= *sf(sc, S_MAXIT, nseg, g)) {
            *sf(sc, S_DONE, nseg, g) = 1.0;
            atomicAdd(ndone, 1);
        }
    }
}

// p = M^-1 r + beta p
template <int PC>
__global__ void __launch_bounds__(PB) k_seg_direction(int nnd, const int* __restrict__ slice_seg,
                                                      const double* __restrict__ r, const double* __restrict__ m,
                                                      double* __restrict__ p, long nslices, int nseg,
                                                      const double* __restrict__ sc)
{
    const int a = blockIdx.x * PB + threadIdx.x;
    if (a >= nnd) return;
    const int g = slice_seg[a >> 6];
    if (*sf((double*)sc, S_DONE, nseg, g) != 0.0) return;
    const double beta = *sf((double*)sc, S_BETA, nseg, g);
    double ri[3], zi[3];
#pragma unroll
    for (int i = 0; i < 3; i++) ri[i] = r[3L * a + i];
    precondition<PC>(m, a, ri, zi);
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const long k = 3L * a + i;
        p[k] = zi[i] + beta * p[k];
    }
}

// ---- persistent iteration: one workgroup per small segment ----------------
//
// A segment of at most PERSIST_MAX slices (PW = 1024 nodes) keeps its
// whole CG loop inside one workgroup of PW threads, thread t owning node
// 64 sb + t: its x and r stay in registers, the direction p in LDS
// (the SpMV gathers neighbours from there), the per-slice partials and the
// segment scalars in LDS and in thread 0's registers; only the operator is
// read from memory (L2) each iteration, and no kernel is launched per
// iteration.  The arithmetic is the five-kernel iteration's, in the same
// order: the row sums of k_seg_spmv (its unrolled form, padding slots
// multiplying zeros), the updates of k_seg_update and k_seg_direction, and
// each per-segment total summed as seg_total sums it with FB threads (here
// FB = PW: thread t holds slice sb + t's partial).  So the iterates are
// bitwise those of the kernel-per-step iteration.  The loop ends on the
// segment's own stopping test or iteration cap, as k_seg_fin_beta decides;
// the final x, r, p and scalars are written back as that iteration leaves
// them.
constexpr int PW = FB;
constexpr int PERSIST_MAX = PW / 64;
static_assert(PERSIST_MAX == PCG_PERSIST_MAX, "the host's threshold cap is the kernel's slice capacity");

// per-segment totals of NV partials held by threads 0..nsl-1 in lp[k][t],
// summed in seg_total's order
template <int NV>
__device__ __forceinline__ void seg_total_lds(const double (&lp)[2][PERSIST_MAX], int nsl, double (&red)[2][PW / 64],
                                              double (&tot)[NV])
{
    const int t = threadIdx.x;
#pragma unroll
    for (int k = 0; k < NV; k++) {
        double x = 0.0;
        if (t < nsl) x += lp[k][t];
        x = wsum(x);
        if ((t & 63) == 0) red[k][t >> 6] = x;
    }
    __syncthreads();
    if (t == 0)
#pragma unroll
        for (int k = 0; k < NV; k++) {
            double x = 0.0;
            for (int w = 0; w < PW / 64; w++) x += red[k][w];
            tot[k] = x;
        }
}

template <class MT, int PC>
__global__ void __launch_bounds__(PW) k_pcg_persistent(const int* __restrict__ segs, int nnd,
                                                       const int* __restrict__ seg_slice,
                                                       const int* __restrict__ sbase, const int* __restrict__ swidth,
                                                       const int* __restrict__ bcol, const MT* __restrict__ bval,
                                                       const double* __restrict__ m, double* __restrict__ x,
                                                       double* __restrict__ r, double* __restrict__ p, int nseg,
                                                       double* __restrict__ sc, int* __restrict__ ndone)
{
    __shared__ double sp[PW][3];
    __shared__ double lp[2][PERSIST_MAX];
    __shared__ double red[2][PW / 64];
    __shared__ double bc[2];
    __shared__ int bdone;
    const int g = segs[blockIdx.x];
    const int sb = seg_slice[g], nsl = seg_slice[g + 1] - sb;
    const int t = threadIdx.x, l = t & 63;
    const int s = sb + (t >> 6);          // this thread's slice
    const int base = 64 * sb;             // first node of the segment
    const int a = base + t;               // this thread's node
    const bool own = (t >> 6) < nsl && a < nnd;
    // thread 0: the segment's scalars
    double rz = 0.0, thresh = 0.0, maxit = 0.0, iter = 0.0, beta = 0.0, pq = 0.0, alpha = 0.0, rr = 0.0;
    if (t == 0) {
        bdone = *sf(sc, S_DONE, nseg, g) != 0.0;
        rz = *sf(sc, S_RZ, nseg, g);
        thresh = *sf(sc, S_THRESH, nseg, g);
        maxit = *sf(sc, S_MAXIT, nseg, g);
        iter = *sf(sc, S_ITER, nseg, g);
        rr = *sf(sc, S_RR, nseg, g);
    }
    double xi[3] = {0, 0, 0}, ri[3] = {0, 0, 0}, zi[3] = {0, 0, 0}, pi[3] = {0, 0, 0};
#pragma unroll
    for (int i = 0; i < 3; i++) {
        if (own) {
            xi[i] = x[3L * a + i];
            ri[i] = r[3L * a + i];
            pi[i] = p[3L * a + i];
        }
        sp[t][i] = pi[i];
    }
    __syncthreads();
    if (bdone) return;
    for (;;) {
        // q = A p and the slice partial of p.q (k_seg_spmv, unrolled form)
        double qi[3] = {0, 0, 0};
        {
            double dv = 0.0;
            if (own) {
                const long b0 = sbase[s];
                const int wd = swidth[s];
                double y0 = 0, y1 = 0, y2 = 0;
#pragma unroll 3
                for (int k = 0; k < wd; k++) {
                    const long blk = b0 + k;
                    const int cr = bcol[blk * 64 + l];
                    const int c = (cr < 0 ? a : cr) - base;
                    const double x0 = sp[c][0], x1 = sp[c][1], x2 = sp[c][2];
                    double v[9];
                    block_values<MT, false>(bval, nullptr, blk, l, false, v);
                    y0 += v[0] * x0 + v[1] * x1 + v[2] * x2;
                    y1 += v[3] * x0 + v[4] * x1 + v[5] * x2;
                    y2 += v[6] * x0 + v[7] * x1 + v[8] * x2;
                }
                qi[0] = y0;
                qi[1] = y1;
                qi[2] = y2;
                dv = y0 * pi[0] + y1 * pi[1] + y2 * pi[2];
            }
            dv = wsum(dv);
            if (l == 0 && (t >> 6) < nsl) lp[0][t >> 6] = dv;
        }
        __syncthreads();
        {
            double tot[1];
            seg_total_lds<1>(lp, nsl, red, tot);
            if (t == 0) {
                pq = tot[0];
                alpha = rz / tot[0];  // k_seg_fin_alpha
                bc[0] = alpha;
            }
        }
        __syncthreads();
        const double al = bc[0];
        // x += alpha p, r -= alpha q, z = M^-1 r; slice partials of r.z, r.r (k_seg_update)
        {
            double v0 = 0.0, v1 = 0.0;
            if (own) {
#pragma unroll
                for (int i = 0; i < 3; i++) {
                    xi[i] += al * pi[i];
                    ri[i] = ri[i] - al * qi[i];
                }
                precondition<PC>(m, a, ri, zi);
#pragma unroll
                for (int i = 0; i < 3; i++) {
                    v0 += ri[i] * zi[i];
                    v1 += ri[i] * ri[i];
                }
            }
            v0 = wsum(v0);
            v1 = wsum(v1);
            if (l == 0 && (t >> 6) < nsl) {
                lp[0][t >> 6] = v0;
                lp[1][t >> 6] = v1;
            }
        }
        __syncthreads();
        {
            double tot[2];
            seg_total_lds<2>(lp, nsl, red, tot);
            if (t == 0) {
                // k_seg_fin_beta
                iter = iter + 1.0;
                beta = tot[0] / rz;
                rz = tot[0];
                rr = tot[1];
                bc[1] = beta;
                bdone = tot[1] <= thresh || iter >= maxit;
            }
        }
        __syncthreads();
        if (bdone) break;
        const double be = bc[1];
        // p = M^-1 r + beta p, with z formed again here as k_seg_direction
        // forms it (the same expression, so the same contraction into FMAs)
        if (own) {
            double zd[3];
            precondition<PC>(m, a, ri, zd);
#pragma unroll
            for (int i = 0; i < 3; i++) {
                pi[i] = zd[i] + be * pi[i];
                sp[t][i] = pi[i];
            }
        }
        __syncthreads();
    }
    if (own)
#pragma unroll
        for (int i = 0; i < 3; i++) {
            x[3L * a + i] = xi[i];
            r[3L * a + i] = ri[i];
            p[3L * a + i] = pi[i];
        }
    if (t == 0) {
        *sf(sc, S_ITER, nseg, g) = iter;
        *sf(sc, S_PQ, nseg, g) = pq;
        *sf(sc, S_ALPHA, nseg, g) = alpha;
        *sf(sc, S_BETA, nseg, g) = beta;
        *sf(sc, S_RZ, nseg, g) = rz;
        *sf(sc, S_RR, nseg, g) = rr;
        *sf(sc, S_DONE, nseg, g) = 1.0;
        atomicAdd(ndone, 1);
    }
}

// ---- fused iteration: four kernels per PCG iteration -----------------------
//
// The direction update p = z + beta p_old is formed inside the SpMV -- for
// the SpMV's own rows, which it stores, and for every neighbour it gathers --
// so the separate direction pass (and its reads of r, z and p) disappears;
// p is double-buffered (the SpMV reads p_old of its neighbours while it
// writes p of its own rows).  The per-segment reductions stay in their own
// finalize kernels: a kernel boundary publishes every workgroup's partials
// at no cost, where a device-scope fence per wave would write back and
// invalidate the XCD's L2 each time.

// z at dof k: Jacobi (m = 1/diag) or multigrid (m = z)
template <int PC>
__device__ __forceinline__ double zval(const double* __restrict__ m, const double* __restrict__ r, long k)
{
    return PC == 2 ? m[k] : m[k] * r[k];
}

template <class MT, int PC, bool XCD, bool UNR, bool SYM>
__global__ void __launch_bounds__(PB) k_fused_spmv(int nnd, const int* __restrict__ slice_seg,
                                                   const int* __restrict__ sbase, const int* __restrict__ swidth,
                                                   const int* __restrict__ bcol, const MT* __restrict__ bval,
                                                   const int* __restrict__ tpos,
                                                   const double* __restrict__ m, const double* __restrict__ r,
                                                   const double* __restrict__ pold, double* __restrict__ pnew,
                                                   double* __restrict__ y, double* __restrict__ part, long nslices,
                                                   int nseg, const double* __restrict__ sc)
{
    const int a = row_block<XCD>() * PB + threadIdx.x;
    const int s = a >> 6, l = a & 63;
    if (a >= 64L * nslices) return;
    const int g = slice_seg[s];  // wave-uniform
    if (*sf((double*)sc, S_DONE, nseg, g) != 0.0) return;
    const double beta = *sf((double*)sc, S_BETA, nseg, g);
    double dv = 0.0;
    if (a < nnd) {
        const long b0 = sbase[s];
        const int w = swidth[s];
        double y0 = 0, y1 = 0, y2 = 0;
        // UNR: padding slots (column -1, zero values) multiply the row's own
        // direction, as in k_seg_spmv
#pragma unroll(UNR ? 3 : 1)
        for (int k = 0; k < w; k++) {
            const long blk = b0 + k;
            const int cr = bcol[blk * 64 + l];
            if (!UNR && cr < 0) continue;
            const int c = cr < 0 ? a : cr;
            const double x0 = zval<PC>(m, r, 3L * c) + beta * pold[3L * c];
            const double x1 = zval<PC>(m, r, 3L * c + 1) + beta * pold[3L * c + 1];
            const double x2 = zval<PC>(m, r, 3L * c + 2) + beta * pold[3L * c + 2];
            double v[9];
            block_values<MT, SYM>(bval, tpos, blk, l, c < a, v);
            y0 += v[0] * x0 + v[1] * x1 + v[2] * x2;
            y1 += v[3] * x0 + v[4] * x1 + v[5] * x2;
            y2 += v[6] * x0 + v[7] * x1 + v[8] * x2;
        }
        double p3[3];
#pragma unroll
        for (int i = 0; i < 3; i++) {
            p3[i] = zval<PC>(m, r, 3L * a + i) + beta * pold[3L * a + i];
            pnew[3L * a + i] = p3[i];
        }
        y[3L * a] = y0;
        y[3L * a + 1] = y1;
        y[3L * a + 2] = y2;
        dv = y0 * p3[0] + y1 * p3[1] + y2 * p3[2];
    }
    dv = wsum(dv);
    if (l == 0) part[s] = dv;
}

// per-segment dot product a.b into out[seg]
__global__ void __launch_bounds__(PB) k_seg_dot(int nnd, const double* __restrict__ x, const double* __restrict__ y,
                                                double* __restrict__ part, long nslices)
{
    const int a = blockIdx.x * PB + threadIdx.x;
    double v = 0.0;
    if (a < nnd)
#pragma unroll
        for (int i = 0; i < 3; i++) v += x[3L * a + i] * y[3L * a + i];
    v = wsum(v);
    if ((threadIdx.x & 63) == 0 && (a >> 6) < nslices) part[a >> 6] = v;
}

__global__ void __launch_bounds__(FB) k_seg_fin_dot(const double* __restrict__ part, long nslices,
                                                    const int* __restrict__ seg_slice, double* __restrict__ out)
{
    const int g = blockIdx.x;
    double tot[1];
    seg_total<1>(part, nslices, seg_slice[g], seg_slice[g + 1], tot);
    if (threadIdx.x == 0) out[g] = tot[0];
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

static unsigned node_blocks(const tg_ctx* c) { return (unsigned)((c->nnd + PB - 1) / PB); }

// f(std::integral_constant<int, v>) for the SpMV variant v = 0..7 (bit 0
// XCD row ranges, bit 1 unrolled loop, bit 2 symmetric read), so every
// variant is its own kernel instance.
template <class F>
static void spmv_dispatch(int v, F f)
{
    switch (v & 7) {
    case 0: f(std::integral_constant<int, 0>{}); break;
    case 1: f(std::integral_constant<int, 1>{}); break;
    case 2: f(std::integral_constant<int, 2>{}); break;
    case 3: f(std::integral_constant<int, 3>{}); break;
    case 4: f(std::integral_constant<int, 4>{}); break;
    case 5: f(std::integral_constant<int, 5>{}); break;
    case 6: f(std::integral_constant<int, 6>{}); break;
    default: f(std::integral_constant<int, 7>{}); break;
    }
}

void segment_dots(tg_ctx* c, const double* a, const double* b, double* host_out)
{
    const long ns = c->seg.nslices;
    {
        TimedRegion tr(c, T_VECTOR);
        hipLaunchKernelGGL(k_seg_dot, dim3(node_blocks(c)), dim3(PB), 0, c->stream, c->nnd, a, b, c->seg.d_part.p, ns);
        hipLaunchKernelGGL(k_seg_fin_dot, dim3(c->seg.nseg), dim3(FB), 0, c->stream, c->seg.d_part.p, ns,
                           c->seg.d_seg_slice.p, c->seg.d_out.p);
        TG_HIP(hipGetLastError());
    }
    TG_HIP(hipMemcpyAsync(host_out, c->seg.d_out.p, sizeof(double) * c->seg.nseg, hipMemcpyDeviceToHost, c->stream));
    TG_HIP(hipStreamSynchronize(c->stream));
}

template <class MT, int PC>
static void pcg_loop(tg_ctx* c, const MT* val, const double* m, const double* b, double* x, double* r, double* p,
                     double* q, double rtol, int maxit, const int* d_active, const PcgWork& W, tg_lin_stats* st);

template <class MT>
void pcg_run(tg_ctx* c, const MT* val, const double* dinv, const double* b, double* x, double* r, double* p,
             double* q, double rtol, int maxit, const int* d_active, int nactive, tg_lin_stats* st,
             const PcgWork* work)
{
    (void)nactive;
    PcgWork W;
    if (work) {
        W = *work;
    } else {
        // the context's own reductions, scalars and done count; the second
        // direction buffer of the fused iteration is allocated on first use
        if (pcg_fused(c) && c->precond != TG_PC_BLOCK_JACOBI && c->vp2.n != 3 * (size_t)c->nnd)
            c->vp2.alloc(3 * (size_t)c->nnd);
        W = PcgWork{c->seg.d_part.p, c->seg.d_scal.p, c->seg.d_ndone.p, c->vp2.p, true};
    }
    if (c->precond == TG_PC_BLOCK_JACOBI)
        pcg_loop<MT, 1>(c, val, c->vbinv.p, b, x, r, p, q, rtol, maxit, d_active, W, st);
    else if (c->precond == TG_PC_AMG)
        pcg_loop<MT, 2>(c, val, c->amg.lv[0]->d_x.p, b, x, r, p, q, rtol, maxit, d_active, W, st);
    else
        pcg_loop<MT, 0>(c, val, dinv, b, x, r, p, q, rtol, maxit, d_active, W, st);
}

template <class MT, int PC>
static void pcg_loop(tg_ctx* c, const MT* val, const double* dinv, const double* b, double* x, double* r, double* p,
                     double* q, double rtol, int maxit, const int* d_active, const PcgWork& W, tg_lin_stats* st)
{
    Segments& S = c->seg;
    const long ns = S.nslices;
    const int nseg = S.nseg;
    const unsigned nb = node_blocks(c);
    TG_HIP(hipMemsetAsync(W.ndone, 0, sizeof(int), c->stream));
    // multigrid: dinv is the z vector; *done of the single segment gates the V-cycle
    const double* done = W.scal + S_DONE * nseg;
    if (PC == 2) {
        TimedRegion tr(c, T_PRECOND);
        amg_vcycle(c, b, (double*)dinv, nullptr);
    }
    {
        TimedRegion tr(c, T_VECTOR);
        hipLaunchKernelGGL(k_seg_init<PC>, dim3(nb), dim3(PB), 0, c->stream, c->nnd, b, dinv, x, r, p, W.part, ns);
        hipLaunchKernelGGL(k_seg_fin_init, dim3(nseg), dim3(FB), 0, c->stream, W.part, ns, S.d_seg_slice.p,
                           d_active, nseg, W.scal, rtol, maxit, W.ndone);
    }
    // small segments: the whole iteration in one persistent workgroup each
    // (the list is built at set-up and when the threshold changes)
    const int npers = PC == 2 || S.persist_built != pcg_persist(c) ? 0 : (int)S.h_persist.size();
    if (npers > 0) {
        TimedRegion tr(c, T_SPMV);
        hipLaunchKernelGGL((k_pcg_persistent<MT, PC>), dim3(npers), dim3(PW), 0, c->stream, S.d_persist.p, c->nnd,
                           S.d_seg_slice.p, c->op.d_slice_base.p, c->op.d_slice_width.p, c->op.d_col.p, val, dinv, x,
                           r, p, nseg, W.scal, W.ndone);
        TG_HIP(hipGetLastError());
    }
    if (npers < nseg) {
    const int CHUNK = PC == 2 ? 8 : 16;  // even: the fused iteration alternates its two direction buffers
    const bool fused = pcg_fused(c) && PC != 1;
    double* pbuf[2] = {p, nullptr};
    if (fused) {
        if (!W.p2) fail(TG_GENERAL_FAIL, "fused PCG iteration without its second direction buffer");
        pbuf[1] = W.p2;
    }
    const int variant = spmv_variant(c);
    auto issue_fused = [&]() {
        for (int k = 0; k < CHUNK; k++) {
            double* pold = pbuf[k & 1];
            double* pnew = pbuf[(k + 1) & 1];
            {
                TimedRegion tr(c, T_SPMV);
spmv_dispatch(variant, [&](auto V) {
                    constexpr int v = decltype(V)::value;
                    hipLaunchKernelGGL((k_fused_spmv<MT, PC, (v & 1) != 0, (v & 2) != 0, (v & 4) != 0>),
                                       dim3((v & 1) ? (nb + 7) / 8 * 8 : nb), dim3(PB), 0, c->stream, c->nnd,
                                       S.d_slice_seg.p, c->op.d_slice_base.p, c->op.d_slice_width.p, c->op.d_col.p, val,
                                       c->op.d_tpos.p, dinv, r, pold, pnew, q, W.part, ns, nseg, W.scal);
                });
            }
            {
                TimedRegion tr(c, T_VECTOR);
                hipLaunchKernelGGL(k_seg_fin_alpha, dim3(nseg), dim3(FB), 0, c->stream, W.part, ns,
                                   S.d_seg_slice.p, nseg, W.scal);
                hipLaunchKernelGGL(k_seg_update<PC>, dim3(nb), dim3(PB), 0, c->stream, c->nnd, S.d_slice_seg.p, pnew,
                                   q, dinv, x, r, W.part, ns, nseg, W.scal);
            }
            if (PC == 2) {
                TimedRegion tr(c, T_PRECOND);
                amg_vcycle(c, r, (double*)dinv, done);
                hipLaunchKernelGGL(k_seg_dot, dim3(nb), dim3(PB), 0, c->stream, c->nnd, r, dinv, W.part, ns);
            }
            {
                TimedRegion tr(c, T_VECTOR);
                hipLaunchKernelGGL(k_seg_fin_beta, dim3(nseg), dim3(FB), 0, c->stream, W.part, ns,
                                   S.d_seg_slice.p, nseg, W.scal, W.ndone);
            }
        }
    };
    auto issue_chunk = [&]() {
        if (fused) {
            issue_fused();
            return;
        }
        for (int k = 0; k < CHUNK; k++) {
            {
                TimedRegion tr(c, T_SPMV);
spmv_dispatch(variant, [&](auto V) {
                    constexpr int v = decltype(V)::value;
                    hipLaunchKernelGGL((k_seg_spmv<MT, (v & 1) != 0, (v & 2) != 0, (v & 4) != 0>),
                                       dim3((v & 1) ? (nb + 7) / 8 * 8 : nb), dim3(PB), 0, c->stream, c->nnd,
                                       S.d_slice_seg.p, c->op.d_slice_base.p, c->op.d_slice_width.p, c->op.d_col.p, val,
                                       c->op.d_tpos.p, p, q, W.part, ns, nseg, W.scal);
                });
            }
            {
                TimedRegion tr(c, T_VECTOR);
                hipLaunchKernelGGL(k_seg_fin_alpha, dim3(nseg), dim3(FB), 0, c->stream, W.part, ns,
                                   S.d_seg_slice.p, nseg, W.scal);
                hipLaunchKernelGGL(k_seg_update<PC>, dim3(nb), dim3(PB), 0, c->stream, c->nnd, S.d_slice_seg.p, p, q,
                                   dinv, x, r, W.part, ns, nseg, W.scal);
            }
            if (PC == 2) {
                TimedRegion tr(c, T_PRECOND);
                amg_vcycle(c, r, (double*)dinv, done);
                hipLaunchKernelGGL(k_seg_dot, dim3(nb), dim3(PB), 0, c->stream, c->nnd, r, dinv, W.part, ns);
            }
            {
                TimedRegion tr(c, T_VECTOR);
                hipLaunchKernelGGL(k_seg_fin_beta, dim3(nseg), dim3(FB), 0, c->stream, W.part, ns,
                                   S.d_seg_slice.p, nseg, W.scal, W.ndone);
                hipLaunchKernelGGL(k_seg_direction<PC>, dim3(nb), dim3(PB), 0, c->stream, c->nnd, S.d_slice_seg.p, r,
                                   dinv, p, ns, nseg, W.scal);
            }
        }
    };
    // The chunk's kernels only ever see the same pointers, so one capture
    // serves every later solve of this context with the same vectors; the
    // iteration state lives in device memory (done flags make finished
    // segments' work a no-op).  The graph holds kernel nodes only.
    const bool graph = W.graph && c->pcg_graph && !c->timing.enabled;
    if (graph) {
        // every address and size the chunk's launches take (tg_initialize
        // also drops the graph, as buffers may come back at old addresses)
        const void* key[24] = {val, dinv, x, r, p, q, W.part, W.scal, W.ndone,
                               (const void*)(long)(PC + 4L * (variant + 8L * ((int)fused + 2L * c->amg.gen))),
                               (const void*)(long)sizeof(MT), (const void*)(long)c->nnd, (const void*)(long)nseg,
                               (const void*)ns, S.d_seg_slice.p, S.d_slice_seg.p, c->op.d_slice_base.p,
                               c->op.d_slice_width.p, c->op.d_col.p, pbuf[1], c->vbinv.p,
                               c->amg.lv.empty() ? nullptr : (const void*)c->amg.lv[0], c->op.d_tpos.p};
        if (!c->pg.exec || std::memcmp(key, c->pg.key, sizeof key) != 0) {
            drop_pcg_graph(c);
            TG_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            issue_chunk();
            TG_HIP(hipStreamEndCapture(c->stream, &c->pg.graph));
            TG_HIP(hipGraphInstantiate(&c->pg.exec, c->pg.graph, nullptr, nullptr, 0));
            std::memcpy(c->pg.key, key, sizeof key);
        }
    }
    // (tg_create made the two poll events)
    // Chunk k+1 is queued before chunk k's done count is read, so the GPU
    // never waits for the host; a chunk queued after every segment finished
    // does nothing.
    int issued = 0, chunks = 0;
    for (;;) {
        if (graph)
            TG_HIP(hipGraphLaunch(c->pg.exec, c->stream));
        else
            issue_chunk();
        TG_HIP(hipGetLastError());
        issued += CHUNK;
        const int slot = chunks & 1;
        TG_HIP(hipMemcpyAsync(c->hstatus + 2 + slot, W.ndone, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        TG_HIP(hipEventRecord(c->poll_ev[slot], c->stream));
        chunks++;
        if (issued >= maxit) break;
        if (chunks >= 2) {
            const int prev = (chunks - 2) & 1;
            TG_HIP(hipEventSynchronize(c->poll_ev[prev]));
            if (c->hstatus[2 + prev] >= nseg) break;
        }
    }
    }  // npers < nseg
    TG_HIP(hipStreamSynchronize(c->stream));
    // per-segment results
    if (S.h_scal.size() != (size_t)S_NSCAL * nseg) S.h_scal.resize((size_t)S_NSCAL * nseg);
    TG_HIP(hipMemcpyAsync(S.h_scal.data(), W.scal, sizeof(double) * S_NSCAL * nseg, hipMemcpyDeviceToHost,
                          c->stream));
    TG_HIP(hipStreamSynchronize(c->stream));
    if (st) {
        // aggregate: max iterations, worst relative residual, all converged
        st->iterations = 0;
        st->rel_residual = 0.0;
        st->converged = 1;
        for (int g = 0; g < nseg; g++) {
            const double it = S.h_scal[(size_t)S_ITER * nseg + g];
            const double bb = S.h_scal[(size_t)S_BB * nseg + g];
            const double rr = S.h_scal[(size_t)S_RR * nseg + g];
            const double th = S.h_scal[(size_t)S_THRESH * nseg + g];
            st->iterations = std::max(st->iterations, (int)it);
            if (bb > 0.0) st->rel_residual = std::max(st->rel_residual, sqrt(rr / bb));
            if (rr > th) st->converged = 0;
        }
    }
}

template void pcg_run<double>(tg_ctx*, const double*, const double*, const double*, double*, double*, double*,
                              double*, double, int, const int*, int, tg_lin_stats*, const PcgWork*);
template void pcg_run<float>(tg_ctx*, const float*, const double*, const double*, double*, double*, double*,
                             double*, double, int, const int*, int, tg_lin_stats*, const PcgWork*);

}  // namespace tg
