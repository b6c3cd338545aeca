// Per-integration-point kinematics and the SimoIso3D material, shared by the
// element kernels.  Everything is FP64 and lives in registers.
#pragma once

#include <hip/hip_runtime.h>

namespace tg {

// 2x2x2 Gauss points at (+-1/sqrt3) in vertex-sign order, weights 1
// (toolbox/src/geometry/HexahedronT.cpp:1547-1554, vertex signs :23-25).
constexpr double GAUSS = 0.57735026918962576451;
constexpr double RA[8] = {-1.0, 1.0, 1.0, -1.0, -1.0, 1.0, 1.0, -1.0};
constexpr double SA[8] = {-1.0, -1.0, 1.0, 1.0, -1.0, -1.0, 1.0, 1.0};
constexpr double TA[8] = {-1.0, -1.0, -1.0, -1.0, 1.0, 1.0, 1.0, 1.0};

// vertex signs of integration point q (lane-dependent, no table lookup)
__device__ __forceinline__ double sign_r(int q) { return ((q ^ (q >> 1)) & 1) ? 1.0 : -1.0; }
__device__ __forceinline__ double sign_s(int q) { return (q & 2) ? 1.0 : -1.0; }
__device__ __forceinline__ double sign_t(int q) { return (q & 4) ? 1.0 : -1.0; }

// parent-domain derivatives dN_a/dxi_k at integration point q
// (HexahedronT::EvaluateShapeFunctions vertex branch, HexahedronT.cpp:407-418)
__device__ __forceinline__ void parent_derivs(int q, double d[3][8])
{
    const double r = GAUSS * sign_r(q), s = GAUSS * sign_s(q), t = GAUSS * sign_t(q);
#pragma unroll
    for (int a = 0; a < 8; a++) {
        const double r1 = 1.0 + RA[a] * r, s1 = 1.0 + SA[a] * s, t1 = 1.0 + TA[a] * t;
        d[0][a] = 0.125 * RA[a] * s1 * t1;
        d[1][a] = 0.125 * r1 * SA[a] * t1;
        d[2][a] = 0.125 * r1 * s1 * TA[a];
    }
}

// 3x3 column-major: m[r + 3c]
__device__ __forceinline__ double det3(const double* a)
{
    return a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) +
           a[2] * (a[3] * a[7] - a[4] * a[6]);
}

__device__ __forceinline__ void inv3(const double* a, double d, double* ai)
{
    const double id = 1.0 / d;
    ai[0] = (a[4] * a[8] - a[7] * a[5]) * id;
    ai[3] = (a[6] * a[5] - a[3] * a[8]) * id;
    ai[6] = (a[3] * a[7] - a[6] * a[4]) * id;
    ai[1] = (a[7] * a[2] - a[1] * a[8]) * id;
    ai[4] = (a[0] * a[8] - a[6] * a[2]) * id;
    ai[7] = (a[6] * a[1] - a[0] * a[7]) * id;
    ai[2] = (a[1] * a[5] - a[4] * a[2]) * id;
    ai[5] = (a[3] * a[2] - a[0] * a[5]) * id;
    ai[8] = (a[0] * a[4] - a[3] * a[1]) * id;
}

// Jacobian j(i,k) = sum_a x_a,i dN_a/dxi_k (ParentDomainT::Jacobian,
// ParentDomainT.cpp:163-186); x is read through the accessor.
template <class XF>
__device__ __forceinline__ void jacobian(XF x, const double d[3][8], double* j)
{
#pragma unroll
    for (int k = 0; k < 9; k++) j[k] = 0.0;
#pragma unroll
    for (int a = 0; a < 8; a++) {
        const double x0 = x(a, 0), x1 = x(a, 1), x2 = x(a, 2);
#pragma unroll
        for (int k = 0; k < 3; k++) {
            j[0 + 3 * k] += x0 * d[k][a];
            j[1 + 3 * k] += x1 * d[k][a];
            j[2 + 3 * k] += x2 * d[k][a];
        }
    }
}

// dN/dX_l = sum_k jinv(k,l) dN/dxi_k (ParentDomainT::ComputeDNa, ParentDomainT.cpp:466-480)
__device__ __forceinline__ void spatial_derivs(const double* ji, const double d[3][8], double g[3][8])
{
#pragma unroll
    for (int a = 0; a < 8; a++)
#pragma unroll
        for (int l = 0; l < 3; l++)
            g[l][a] = ji[0 + 3 * l] * d[0][a] + ji[1 + 3 * l] * d[1][a] + ji[2 + 3 * l] * d[2][a];
}

// F = 1 + sum_a u_a (x) g_a (FiniteStrainT::SetGlobalShape, FiniteStrainT.cpp:278-303)
template <class UF>
__device__ __forceinline__ void deformation_gradient(UF u, const double g[3][8], double* F)
{
#pragma unroll
    for (int k = 0; k < 9; k++) F[k] = 0.0;
#pragma unroll
    for (int a = 0; a < 8; a++) {
        const double u0 = u(a, 0), u1 = u(a, 1), u2 = u(a, 2);
#pragma unroll
        for (int k = 0; k < 3; k++) {
            F[0 + 3 * k] += u0 * g[k][a];
            F[1 + 3 * k] += u1 * g[k][a];
            F[2 + 3 * k] += u2 * g[k][a];
        }
    }
    F[0] += 1.0;
    F[4] += 1.0;
    F[8] += 1.0;
}

// b = F F^T (Voigt 11,22,33,23,13,12), J = sqrt(det b), b_bar = J^(-2/3) b
// (FSSolidMatT::Compute_b, FSSolidMatT.cpp:415-430; SimoIso3D.cpp:24-30).
// Returns false when det b <= 0.
__device__ __forceinline__ bool simo_kinematics(const double* f, double& J, double* bb)
{
    double b[6];
    b[0] = f[0] * f[0] + f[3] * f[3] + f[6] * f[6];
    b[1] = f[1] * f[1] + f[4] * f[4] + f[7] * f[7];
    b[2] = f[2] * f[2] + f[5] * f[5] + f[8] * f[8];
    b[3] = f[1] * f[2] + f[4] * f[5] + f[7] * f[8];
    b[4] = f[0] * f[2] + f[3] * f[5] + f[6] * f[8];
    b[5] = f[0] * f[1] + f[3] * f[4] + f[6] * f[7];
    const double detb = b[0] * (b[1] * b[2] - b[3] * b[3]) - b[5] * (b[5] * b[2] - b[3] * b[4]) +
                        b[4] * (b[5] * b[3] - b[1] * b[4]);
    if (!(detb > 0.0)) return false;
    J = sqrt(detb);
    const double r = rcbrt(J);  // J^(-1/3)
    const double s = r * r;
#pragma unroll
    for (int k = 0; k < 6; k++) bb[k] = s * b[k];
    return true;
}

// Cauchy stress sigma = (mu/J) dev(b_bar) + U'(J) 1, U' = kappa/2 (J - 1/J)
// (SimoIso3D::ComputeCauchy SimoIso3D.cpp:127-136, SimoIso3D.h:93-96)
__device__ __forceinline__ void simo_cauchy(double mu, double kappa, double J, const double* bb, double* sig)
{
    const double m = mu / J;
#pragma unroll
    for (int k = 0; k < 6; k++) sig[k] = m * bb[k];
    const double mean = (sig[0] + sig[1] + sig[2]) / 3.0;
    const double p = 0.5 * kappa * (J - 1.0 / J);
    sig[0] += p - mean;
    sig[1] += p - mean;
    sig[2] += p - mean;
}

// Spatial tangent in Voigt form, c[v][w] (SimoIso3D::ComputeModuli,
// SimoIso3D.cpp:102-125):
//   c = (U' + J U'') 1x1 - 2U' I + 2 mu_bar Dev - 4/(3J) sym(dev(mu b_bar) x 1)
__device__ __forceinline__ void simo_moduli(double mu, double kappa, double J, const double* bb, double c[6][6])
{
    const double du = 0.5 * kappa * (J - 1.0 / J);
    const double ddu = 0.5 * kappa * (1.0 + 1.0 / (J * J));
    const double mu_bar = mu * (bb[0] + bb[1] + bb[2]) / (J * 3.0);
    double s[6];
#pragma unroll
    for (int k = 0; k < 6; k++) s[k] = mu * bb[k];
    const double mean = (s[0] + s[1] + s[2]) / 3.0;
    s[0] -= mean;
    s[1] -= mean;
    s[2] -= mean;
    const double vol = du + J * ddu;
    const double f4 = -4.0 / (J * 3.0);
#pragma unroll
    for (int v = 0; v < 6; v++)
#pragma unroll
        for (int w = 0; w < 6; w++) {
            const double one_v = v < 3 ? 1.0 : 0.0, one_w = w < 3 ? 1.0 : 0.0;
            double x = vol * one_v * one_w;
            if (v == w) x += -2.0 * du * (v < 3 ? 1.0 : 0.5);
            const double dev = (v < 3 && w < 3) ? (v == w ? 2.0 / 3.0 : -1.0 / 3.0) : (v == w ? 0.5 : 0.0);
            x += 2.0 * mu_bar * dev;
            x += f4 * 0.5 * (s[v] * one_w + one_v * s[w]);
            c[v][w] = x;
        }
}

// Kirchhoff-St.Venant (large_strain_StVenant, FDKStV.cpp with
// FDHookeanMatT::s_ij/c_ijkl, FDHookeanMatT.cpp, and
// IsotropicT::ComputeModuli, IsotropicT.cpp:153-169): S = lambda tr(E) 1 +
// 2 mu E with E = (F^T F - 1)/2, pushed forward, sigma = F S F^T / J.  With
// b = F F^T (Voigt 11,22,33,23,13,12) and F E F^T = (b b - b)/2 this is
//   sigma = (lambda tr(E) b + mu (b b - b)) / J,  tr(E) = (tr b - 3)/2,
// and the pushed-forward moduli of lambda 1x1 + 2 mu I are
//   c_ijkl = (lambda b_ij b_kl + mu (b_ik b_jl + b_il b_jk)) / J.
// Material parameters arrive as (mu, kappa): lambda = kappa - 2 mu / 3.
__device__ __forceinline__ void stvenant_b(const double* f, double* b)
{
    b[0] = f[0] * f[0] + f[3] * f[3] + f[6] * f[6];
    b[1] = f[1] * f[1] + f[4] * f[4] + f[7] * f[7];
    b[2] = f[2] * f[2] + f[5] * f[5] + f[8] * f[8];
    b[3] = f[1] * f[2] + f[4] * f[5] + f[7] * f[8];
    b[4] = f[0] * f[2] + f[3] * f[5] + f[6] * f[8];
    b[5] = f[0] * f[1] + f[3] * f[4] + f[6] * f[7];
}

// full 3x3 index of Voigt component v: (row, column)
__device__ __forceinline__ int vrow(int v) { return v < 3 ? v : (v == 3 ? 1 : 0); }
__device__ __forceinline__ int vcol(int v) { return v < 3 ? v : (v == 5 ? 1 : 2); }
__device__ __forceinline__ int vidx(int i, int j)
{
    return i == j ? i : (i + j == 3 ? 3 : (i + j == 2 ? 4 : 5));
}

__device__ __forceinline__ void stvenant_cauchy(double mu, double kappa, double J, const double* b, double* sig)
{
    const double lam = kappa - 2.0 * mu / 3.0;
    const double trE = 0.5 * (b[0] + b[1] + b[2] - 3.0);
#pragma unroll
    for (int v = 0; v < 6; v++) {
        const int i = vrow(v), j = vcol(v);
        double bb = 0.0;  // (b b)_ij
#pragma unroll
        for (int k = 0; k < 3; k++) bb += b[vidx(i, k)] * b[vidx(k, j)];
        sig[v] = (lam * trE * b[v] + mu * (bb - b[v])) / J;
    }
}

__device__ __forceinline__ void stvenant_moduli(double mu, double kappa, double J, const double* b, double c[6][6])
{
    const double lam = kappa - 2.0 * mu / 3.0;
#pragma unroll
    for (int v = 0; v < 6; v++)
#pragma unroll
        for (int w = 0; w < 6; w++) {
            const int i = vrow(v), j = vcol(v), k = vrow(w), l = vcol(w);
            c[v][w] = (lam * b[v] * b[w] + mu * (b[vidx(i, k)] * b[vidx(j, l)] + b[vidx(i, l)] * b[vidx(j, k)])) / J;
        }
}

// Material models of the large-strain elements: the template parameter of
// the element kernels.
enum { MAT_SIMO = 0, MAT_STVENANT = 1 };

// Cauchy stress of model MAT at F (J = det F > 0 checked by the caller);
// kin holds what the tangent of the same model needs (J and b_bar for
// Simo, J and b for St.Venant).  Returns false when det b <= 0.
template <int MAT>
__device__ __forceinline__ bool material_stress(double mu, double kappa, const double* F, double J, double& Jk,
                                                double* kin, double* sig)
{
    if (MAT == MAT_STVENANT) {
        stvenant_b(F, kin);
        Jk = J;
        stvenant_cauchy(mu, kappa, J, kin, sig);
        return true;
    } else {
        const bool ok = simo_kinematics(F, Jk, kin);
        simo_cauchy(mu, kappa, Jk, kin, sig);
        return ok;
    }
}

template <int MAT>
__device__ __forceinline__ void material_moduli(double mu, double kappa, double Jk, const double* kin, double c[6][6])
{
    if (MAT == MAT_STVENANT)
        stvenant_moduli(mu, kappa, Jk, kin, c);
    else
        simo_moduli(mu, kappa, Jk, kin, c);
}

// B-matrix structure (SolidElementT::Set_B 3D, SolidElementT.cpp:890-918):
// B[v][3a+i] = g[bcomp(v,i)][a] or 0 when bcomp < 0.
__device__ __forceinline__ int bcomp(int v, int i)
{
    // rows v: 11 22 33 23 13 12 ; columns i: x y z
    constexpr int T[6][3] = {{0, -1, -1}, {-1, 1, -1}, {-1, -1, 2}, {-1, 2, 1}, {2, -1, 0}, {1, 0, -1}};
    return T[v][i];
}

}  // namespace tg
