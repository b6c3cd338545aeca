// Q1P0 mean-dilatation hexahedron kernels (residual and forward-difference
// stiffness).
#include "tg_internal.h"
#include "tg_kinematics.h"
#include "tg_element.h"

// ---------------------------------------------------------------------------
// Q1P0 mean-dilatation hexahedron (SimoQ1P0, tahoe/src/elements/continuum/solid/SimoQ1P0.cpp)
// ---------------------------------------------------------------------------

namespace tg {

__device__ __forceinline__ double allreduce8(double v)
{
    v += __shfl_xor(v, 1);
    v += __shfl_xor(v, 2);
    v += __shfl_xor(v, 4);
    return v;
}

// Per-lane (integration point q) share of the element internal force
// f = sum_q w det_q (J_bar/J_q) B_bar_q^T sigma(F_bar_q) for the state in
// LDS (x current coordinates, u displacements), ComputeInternalForce
// (SimoQ1P0.cpp:390-439) with E = 0 and no phase field:
//   SetMeanGradient (:661-686): H = sum w det0, v = sum w det, mean = sum (w det/v) dN/dx
//   F_bar = (v/(H J))^(1/3) F, F from reference derivatives (:286-296)
//   B_bar from SolidElementT::Set_B_bar (SolidElementT.cpp:1011-1041).
// All 8 lanes of the element must call it together (wave-wide shuffles).
template <int MAT, class XF, class UF>
__device__ __forceinline__ bool q1p0_ip_force(int q, const double d[3][8], const double det0, const double g0[3][8],
                                              XF xcur, UF u, double mu, double kappa, double f[24])
{
    bool ok = true;
    double jc[9], jci[9], dnx[3][8];
    jacobian(xcur, d, jc);
    const double det = det3(jc);
    ok &= det > 0.0;
    inv3(jc, det, jci);
    spatial_derivs(jci, d, dnx);
    const double H = allreduce8(det0);  // weights 1
    const double v = allreduce8(det);
    const double wv = det / v;
    double mean[3][8];
#pragma unroll
    for (int k = 0; k < 3; k++)
#pragma unroll
        for (int a = 0; a < 8; a++) mean[k][a] = allreduce8(wv * dnx[k][a]);
    double F[9];
    deformation_gradient(u, g0, F);
    const double J = det3(F);
    ok &= J > 0.0;
    const double sc = cbrt(v / (H * J));
#pragma unroll
    for (int k = 0; k < 9; k++) F[k] *= sc;
    // the material sees F_bar (SimoQ1P0 passes it on as the mechanical F)
    const double Jbar = det3(F);
    double Jk, kin[6], sig[6];
    ok &= material_stress<MAT>(mu, kappa, F, Jbar, Jk, kin, sig);
    const double s = det * (Jbar / J);  // w det_q J_bar/J_q, weight 1
#pragma unroll
    for (int a = 0; a < 8; a++) {
        const double nx = dnx[0][a], ny = dnx[1][a], nz = dnx[2][a];
        const double fx = (mean[0][a] - nx) / 3.0, fy = (mean[1][a] - ny) / 3.0, fz = (mean[2][a] - nz) / 3.0;
        const double bx = (nx + fx) * sig[0] + fx * sig[1] + fx * sig[2] + nz * sig[4] + ny * sig[5];
        const double by = fy * sig[0] + (ny + fy) * sig[1] + fy * sig[2] + nz * sig[3] + nx * sig[5];
        const double bz = fz * sig[0] + fz * sig[1] + (nz + fz) * sig[2] + ny * sig[3] + nx * sig[4];
        f[3 * a + 0] = s * bx;
        f[3 * a + 1] = s * by;
        f[3 * a + 2] = s * bz;
    }
    if (!ok)
#pragma unroll
        for (int k = 0; k < 24; k++) f[k] = 0.0;
    return ok;
}

// Q1P0 residual R -= f (FormKd(-1), SimoQ1P0.cpp:528-582) and, when `stiff`,
// the reference's forward-difference tangent (FormStiffness, SimoQ1P0.cpp:446-525):
// for each element dof col, u_col += eps and x_col += eps, K(:,col) =
// (f(u+eps) - f(u))/eps, then both are restored by subtracting eps, eps = 1e-7.
// Upper entries (row <= col) are scattered through the slot map.
template <bool STIFF, int MAT>
__global__ void __launch_bounds__(256) k_q1p0(const int* __restrict__ elems, int count, const int* __restrict__ conn,
                                              const int* __restrict__ emat, const double2* __restrict__ props,
                                              const double* __restrict__ X, const double* __restrict__ U,
                                              const int* __restrict__ eqn, double* __restrict__ R,
                                              const int* __restrict__ slot, double* __restrict__ val,
                                              int* __restrict__ status, const int* __restrict__ elem_seg,
                                              int* __restrict__ segstat)
{
    __shared__ double sX[32][8][3], sx[32][8][3], su[32][8][3];
    const int t = threadIdx.x, le = t >> 3, q = t & 7;
    const long sl = (long)blockIdx.x * 32 + le;
    const bool valid = sl < count;
    int e = 0, node = 0;
    if (valid) {
        e = elems[sl];
        node = conn[8L * e + q];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const double Xc = X[3L * node + c], uc = U[3L * node + c];
            sX[le][q][c] = Xc;
            su[le][q][c] = uc;
            sx[le][q][c] = Xc + uc;
        }
    } else {
        // inert unit element so the shuffled element reductions stay finite
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const double Xc = c == 0 ? (RA[q] + 1.0) * 0.5 : (c == 1 ? (SA[q] + 1.0) * 0.5 : (TA[q] + 1.0) * 0.5);
            sX[le][q][c] = Xc;
            su[le][q][c] = 0.0;
            sx[le][q][c] = Xc;
        }
    }
    __syncthreads();
    double2 mk = make_double2(1.0, 1.0);
    if (valid) mk = props[emat[e]];
    double d[3][8], j0[9], ji[9], g0[3][8];
    parent_derivs(q, d);
    jacobian([&](int a, int c) { return sX[le][a][c]; }, d, j0);
    const double det0 = det3(j0);
    bool ok = det0 > 0.0;
    inv3(j0, det0, ji);
    spatial_derivs(ji, d, g0);
    auto xa = [&](int a, int c) { return sx[le][a][c]; };
    auto ua = [&](int a, int c) { return su[le][a][c]; };
    double f[24], v0[3];
    ok &= q1p0_ip_force<MAT>(q, d, det0, g0, xa, ua, mk.x, mk.y, f);
    reduce_scatter8(f, q, v0);
    if (!STIFF) {
        if (valid)
#pragma unroll
            for (int i = 0; i < 3; i++)
                if (eqn[3L * node + i] >= 0) R[3L * node + i] -= v0[i];
    } else {
        const double eps = 1.0e-7;
        const int* es = slot + 300L * e;
        for (int col = 0; col < 24; col++) {
            const int pa = col / 3, pi = col % 3;
            if (q == pa) {
                su[le][pa][pi] += eps;
                sx[le][pa][pi] += eps;
            }
            __syncthreads();
            double fp[24], vp[3];
            ok &= q1p0_ip_force<MAT>(q, d, det0, g0, xa, ua, mk.x, mk.y, fp);
            reduce_scatter8(fp, q, vp);
            __syncthreads();
            if (q == pa) {
                su[le][pa][pi] -= eps;
                sx[le][pa][pi] -= eps;
            }
            if (valid) {
#pragma unroll
                for (int i = 0; i < 3; i++) {
                    const int m = 3 * q + i;
                    if (m <= col) {
                        const int s = slot_target(es[upper_index(m, col)]);
                        if (s >= 0) val[s] += (vp[i] - v0[i]) / eps;
                    }
                }
            }
            __syncthreads();
        }
    }
    if (valid && !ok) {
        // the context's flag, and the flag of the element's problem in a batch
        atomicOr(status, 7);
        atomicOr(&segstat[elem_seg[e]], 7);
    }
}

void launch_q1p0(tg_ctx* c, const double* u, double* R_node, bool stiff)
{
    for (int col = 0; col < c->ncolors; col++) {
        const int first = c->color_ptr[col], count = c->color_ptr[col + 1] - first;
        if (count == 0) continue;
        const int blocks = (count + 31) / 32;
#define TG_Q1P0(S, M)                                                                                          \
    hipLaunchKernelGGL((k_q1p0<S, M>), dim3(blocks), dim3(256), 0, c->stream, c->dcolor_elems.p + first, count, \
                       c->dconn.p, c->dmat.p, c->dmatprops.p, c->dX.p, u, c->deqn.p, R_node, c->dslot.p, c->dval.p, \
                       c->dstatus.p, c->seg.d_elem_seg.p, c->seg.d_segstat.p)
        const bool stv = c->material_model == TG_MAT_STVENANT;
        if (stiff && stv)
            TG_Q1P0(true, MAT_STVENANT);
        else if (stiff)
            TG_Q1P0(true, MAT_SIMO);
        else if (stv)
            TG_Q1P0(false, MAT_STVENANT);
        else
            TG_Q1P0(false, MAT_SIMO);
#undef TG_Q1P0
    }
    TG_HIP(hipGetLastError());
}

}  // namespace tg
