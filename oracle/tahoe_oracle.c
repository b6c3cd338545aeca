/* CPU restatement of Tahoe's continuum-solid Newton step -- TEST INFRASTRUCTURE.
 *
 * The checker for the HIP path (see tahoe_oracle.h for conventions).  Only
 * tests/, the smoke check and the CPU-baseline leg of bench.py load it.
 * Citations are file:line in the Tahoe source tree.
 */
#include "tahoe_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* small dense helpers                                                       */
/* ------------------------------------------------------------------------ */

/* 3x3 matrices are stored column-major like dMatrixT: m[r + 3*c]. */
#define M3(m, r, c) ((m)[(r) + 3 * (c)])

/* dMatrixT::Det, toolbox/src/abc/matrix/dMatrixT.cpp:190-193 */
static double det3(const double* a)
{
    return a[0] * (a[4] * a[8] - a[5] * a[7])
         - a[1] * (a[3] * a[8] - a[5] * a[6])
         + a[2] * (a[3] * a[7] - a[4] * a[6]);
}

/* inverse of a 3x3 (adjugate / determinant); dMatrixT::Inverse
 * (dMatrixT.cpp:56-110) evaluates the same cofactors in another order. */
static void inv3(const double* a, double* ai)
{
    double d = det3(a);
    M3(ai, 0, 0) = (M3(a, 1, 1) * M3(a, 2, 2) - M3(a, 1, 2) * M3(a, 2, 1)) / d;
    M3(ai, 0, 1) = (M3(a, 0, 2) * M3(a, 2, 1) - M3(a, 0, 1) * M3(a, 2, 2)) / d;
    M3(ai, 0, 2) = (M3(a, 0, 1) * M3(a, 1, 2) - M3(a, 0, 2) * M3(a, 1, 1)) / d;
    M3(ai, 1, 0) = (M3(a, 1, 2) * M3(a, 2, 0) - M3(a, 1, 0) * M3(a, 2, 2)) / d;
    M3(ai, 1, 1) = (M3(a, 0, 0) * M3(a, 2, 2) - M3(a, 0, 2) * M3(a, 2, 0)) / d;
    M3(ai, 1, 2) = (M3(a, 0, 2) * M3(a, 1, 0) - M3(a, 0, 0) * M3(a, 1, 2)) / d;
    M3(ai, 2, 0) = (M3(a, 1, 0) * M3(a, 2, 1) - M3(a, 1, 1) * M3(a, 2, 0)) / d;
    M3(ai, 2, 1) = (M3(a, 0, 1) * M3(a, 2, 0) - M3(a, 0, 0) * M3(a, 2, 1)) / d;
    M3(ai, 2, 2) = (M3(a, 0, 0) * M3(a, 1, 1) - M3(a, 0, 1) * M3(a, 1, 0)) / d;
}

/* Voigt order (11,22,33,23,13,12): dSymMatrixT, toolbox/src/abc/matrix/dSymMatrixT.h */
static const int VI[6] = {0, 1, 2, 1, 0, 0};
static const int VJ[6] = {0, 1, 2, 2, 2, 1};

/* dSymMatrixT::Det, dSymMatrixT.cpp:396-399 */
static double sym_det(const double* s)
{
    return s[0] * (s[1] * s[2] - s[3] * s[3])
         - s[5] * (s[5] * s[2] - s[3] * s[4])
         + s[4] * (s[5] * s[3] - s[1] * s[4]);
}

/* ------------------------------------------------------------------------ */
/* A6: SimoIso3D (materials/Simo/SimoIso3D.{h,cpp})                          */
/* ------------------------------------------------------------------------ */

/* IsotropicT::Set_E_nu, materials/primitives/IsotropicT.cpp:32-45 */
void or_mu_kappa(double E, double nu, double* mu, double* kappa)
{
    double m = 0.5 * E / (1.0 + nu);
    double lambda = 2.0 * m * nu / (1.0 - 2.0 * nu);
    *mu = m;
    *kappa = lambda + 2.0 / 3.0 * m;
}

/* b = F F^T, FSSolidMatT::Compute_b, materials/primitives/FSSolidMatT.cpp:415-430;
 * J = sqrt(det b), b_bar = J^(-2/3) b, SimoIso3D.cpp:24-30.  Returns 7 if det b <= 0. */
static int simo_kinematics(const double* f, double* J_out, double* bb)
{
    double b[6];
    b[0] = f[0] * f[0] + f[3] * f[3] + f[6] * f[6];
    b[1] = f[1] * f[1] + f[4] * f[4] + f[7] * f[7];
    b[2] = f[2] * f[2] + f[5] * f[5] + f[8] * f[8];
    b[3] = f[1] * f[2] + f[4] * f[5] + f[7] * f[8];
    b[4] = f[0] * f[2] + f[3] * f[5] + f[6] * f[8];
    b[5] = f[0] * f[1] + f[3] * f[4] + f[6] * f[7];
    double J = sym_det(b);
    if (J <= 0.0) return 7;
    J = sqrt(J);
    double s = pow(J, -2.0 / 3.0);
    for (int k = 0; k < 6; k++) bb[k] = s * b[k];
    *J_out = J;
    return 0;
}

/* U'(J) and U''(J), SimoIso3D.h:93-101 */
static double dU(double kappa, double J) { return 0.5 * kappa * (J - 1.0 / J); }
static double ddU(double kappa, double J) { return 0.5 * kappa * (1.0 + 1.0 / (J * J)); }

/* sigma = (mu/J) dev(b_bar) + U'(J) 1, SimoIso3D::ComputeCauchy SimoIso3D.cpp:127-136 */
int or_simo_cauchy(double mu, double kappa, const double F[9], double sig[6])
{
    double J, bb[6];
    if (simo_kinematics(F, &J, bb)) return 7;
    for (int k = 0; k < 6; k++) sig[k] = (mu / J) * bb[k];
    double mean = (sig[0] + sig[1] + sig[2]) / 3.0; /* dSymMatrixT::Deviatoric */
    sig[0] -= mean; sig[1] -= mean; sig[2] -= mean;
    double p = dU(kappa, J);
    sig[0] += p; sig[1] += p; sig[2] += p;
    return 0;
}

/* spatial tangent, SimoIso3D::ComputeModuli SimoIso3D.cpp:102-125:
 *   c = (U' + J U'') 1x1 - 2 U' I + 2 mu_bar Dev - 4/(3J) sym(dev(mu b_bar) x 1),
 *   mu_bar = mu tr(b_bar)/(3J); I = ReducedIndexI, Dev = ReducedIndexDeviatoric
 *   (dMatrixT.cpp:328-387), sym = dMatrixT::Symmetrize (dMatrixT.cpp:235-258). */
int or_simo_moduli(double mu, double kappa, const double F[9], double c[36])
{
    double J, bb[6];
    if (simo_kinematics(F, &J, bb)) return 7;
    double du = dU(kappa, J), ddu = ddU(kappa, J);
    double mu_bar = mu * (bb[0] + bb[1] + bb[2]) / (J * 3.0);
    double s[6];
    for (int k = 0; k < 6; k++) s[k] = mu * bb[k];
    double mean = (s[0] + s[1] + s[2]) / 3.0;
    s[0] -= mean; s[1] -= mean; s[2] -= mean;
    static const double one[6] = {1, 1, 1, 0, 0, 0};
    for (int j = 0; j < 6; j++)
        for (int i = 0; i < 6; i++) {
            double v = 0.0;
            v += (du + J * ddu) * one[i] * one[j];
            double I4 = (i == j) ? (i < 3 ? 1.0 : 0.5) : 0.0;
            v += -2.0 * du * I4;
            double D4 = (i < 3 && j < 3) ? (i == j ? 2.0 / 3.0 : -1.0 / 3.0) : (i == j ? 0.5 : 0.0);
            v += 2.0 * mu_bar * D4;
            double r4 = 0.5 * (s[i] * one[j] + one[i] * s[j]);
            v += -4.0 / (J * 3.0) * r4;
            c[i + 6 * j] = v;
        }
    return 0;
}

/* W = U(J) + mu/2 (tr b_bar - 3), SimoIso3D.cpp:138-142, U from SimoIso3D.h:88-91 */
double or_simo_energy(double mu, double kappa, const double F[9])
{
    double J, bb[6];
    if (simo_kinematics(F, &J, bb)) return NAN;
    return 0.5 * kappa * (0.5 * (J * J - 1.0) - log(J)) + 0.5 * mu * (bb[0] + bb[1] + bb[2] - 3.0);
}

/* large_strain_StVenant (FDKStV, FDKStV.cpp): FDHookeanMatT::s_ij and c_ijkl
 * (FDHookeanMatT.cpp: Compute_E, HookeanStress, PushForward, / det F) with
 * the isotropic moduli of IsotropicT::ComputeModuli (IsotropicT.cpp:153-169),
 * lambda = kappa - 2 mu / 3 (IsotropicT::Set_E_nu gives mu and kappa).
 * Restated in index form: E_IJ = (F_kI F_kJ - delta_IJ)/2, S_IJ = C_IJKL E_KL,
 * sigma_ij = F_iI S_IJ F_jJ / J, c_ijkl = F_iI F_jJ F_kK F_lL C_IJKL / J. */
static double iso_C(double lam, double mu, int I, int J, int K, int L)
{
    return lam * (I == J) * (K == L) + mu * ((I == K) * (J == L) + (I == L) * (J == K));
}

int or_stvenant_cauchy(double mu, double kappa, const double F[9], double sig[6])
{
    const double lam = kappa - 2.0 * mu / 3.0;
    const double J = det3(F);
    if (!(J > 0.0)) return 7;
    double E[9], S[9];
    for (int I = 0; I < 3; I++)
        for (int Jx = 0; Jx < 3; Jx++) {
            double c = 0.0;
            for (int k = 0; k < 3; k++) c += M3(F, k, I) * M3(F, k, Jx);
            M3(E, I, Jx) = 0.5 * (c - (I == Jx ? 1.0 : 0.0));
        }
    for (int I = 0; I < 3; I++)
        for (int Jx = 0; Jx < 3; Jx++) {
            double v = 0.0;
            for (int K = 0; K < 3; K++)
                for (int L = 0; L < 3; L++) v += iso_C(lam, mu, I, Jx, K, L) * M3(E, K, L);
            M3(S, I, Jx) = v;
        }
    for (int v = 0; v < 6; v++) {
        const int i = VI[v], j = VJ[v];
        double x = 0.0;
        for (int I = 0; I < 3; I++)
            for (int Jx = 0; Jx < 3; Jx++) x += M3(F, i, I) * M3(S, I, Jx) * M3(F, j, Jx);
        sig[v] = x / J;
    }
    return 0;
}

int or_stvenant_moduli(double mu, double kappa, const double F[9], double c[36])
{
    const double lam = kappa - 2.0 * mu / 3.0;
    const double J = det3(F);
    if (!(J > 0.0)) return 7;
    for (int v = 0; v < 6; v++)
        for (int w = 0; w < 6; w++) {
            const int i = VI[v], j = VJ[v], k = VI[w], l = VJ[w];
            double x = 0.0;
            for (int I = 0; I < 3; I++)
                for (int Jx = 0; Jx < 3; Jx++)
                    for (int K = 0; K < 3; K++)
                        for (int L = 0; L < 3; L++)
                            x += M3(F, i, I) * M3(F, j, Jx) * M3(F, k, K) * M3(F, l, L) * iso_C(lam, mu, I, Jx, K, L);
            c[v + 6 * w] = x / J;
        }
    return 0;
}

/* W = lambda/2 tr(E)^2 + mu E:E (HookeanMatT::HookeanEnergy with the
 * isotropic moduli) */
double or_stvenant_energy(double mu, double kappa, const double F[9])
{
    const double lam = kappa - 2.0 * mu / 3.0;
    double E[9], tr = 0.0, ee = 0.0;
    for (int I = 0; I < 3; I++)
        for (int Jx = 0; Jx < 3; Jx++) {
            double c = 0.0;
            for (int k = 0; k < 3; k++) c += M3(F, k, I) * M3(F, k, Jx);
            M3(E, I, Jx) = 0.5 * (c - (I == Jx ? 1.0 : 0.0));
        }
    for (int I = 0; I < 3; I++) tr += M3(E, I, I);
    for (int q = 0; q < 9; q++) ee += E[q] * E[q];
    return 0.5 * lam * tr * tr + mu * ee;
}

typedef int (*stress_fn)(double, double, const double*, double*);

/* ------------------------------------------------------------------------ */
/* A3: 8-node hexahedron, 2x2x2 Gauss (toolbox/src/geometry/HexahedronT.cpp) */
/* ------------------------------------------------------------------------ */

static const double RA[8] = {-1.0, 1.0, 1.0, -1.0, -1.0, 1.0, 1.0, -1.0}; /* HexahedronT.cpp:23-25 */
static const double SA[8] = {-1.0, -1.0, 1.0, 1.0, -1.0, -1.0, 1.0, 1.0};
static const double TA[8] = {-1.0, -1.0, -1.0, -1.0, 1.0, 1.0, 1.0, 1.0};

/* SetLocalShape case 8 (HexahedronT.cpp:1547-1554) + EvaluateShapeFunctions
 * vertex branch (HexahedronT.cpp:407-418).  DNa[q][k][a] = dN_a/dxi_k at ip q. */
void or_hex8_parent(double Na[8][8], double DNa[8][3][8], double w[8])
{
    double g = 1.0 / sqrt(3.0);
    for (int q = 0; q < 8; q++) {
        double r = g * RA[q], s = g * SA[q], t = g * TA[q];
        for (int a = 0; a < 8; a++) {
            double r1 = 1.0 + RA[a] * r, s1 = 1.0 + SA[a] * s, t1 = 1.0 + TA[a] * t;
            Na[q][a] = 0.125 * r1 * s1 * t1;
            DNa[q][0][a] = 0.125 * RA[a] * s1 * t1;
            DNa[q][1][a] = 0.125 * r1 * SA[a] * t1;
            DNa[q][2][a] = 0.125 * r1 * s1 * TA[a];
        }
        w[q] = 1.0;
    }
}

typedef struct {
    double Na[8][8], DNa[8][3][8], w[8];
} parent_t;

static const parent_t* parent(void)
{
    static parent_t P;
    static int init = 0;
    if (!init) { or_hex8_parent(P.Na, P.DNa, P.w); init = 1; }
    return &P;
}

/* A4: ParentDomainT::Jacobian (ParentDomainT.cpp:163-186) and ComputeDNa
 * (ParentDomainT.cpp:426-485).  x is [8][3] node-major; out: det[q],
 * dN[q][k][a] = dN_a/dx_k.  Returns 7 on det <= 0 (ParentDomainT.cpp:451). */
static int shape_derivs(const double* x, double det[8], double dN[8][3][8])
{
    const parent_t* P = parent();
    for (int q = 0; q < 8; q++) {
        double jac[9] = {0};
        for (int a = 0; a < 8; a++)
            for (int k = 0; k < 3; k++) { /* j(i,k) += x_i(a) dN_a/dxi_k */
                double d = P->DNa[q][k][a];
                M3(jac, 0, k) += x[3 * a + 0] * d;
                M3(jac, 1, k) += x[3 * a + 1] * d;
                M3(jac, 2, k) += x[3 * a + 2] * d;
            }
        det[q] = det3(jac);
        if (det[q] <= 0.0) return 7;
        double ji[9];
        inv3(jac, ji);
        for (int a = 0; a < 8; a++)
            for (int l = 0; l < 3; l++)
                dN[q][l][a] = M3(ji, 0, l) * P->DNa[q][0][a] + M3(ji, 1, l) * P->DNa[q][1][a]
                            + M3(ji, 2, l) * P->DNa[q][2][a];
    }
    return 0;
}

/* A5: F = 1 + sum_a u_a (x) dN_a/dX (ShapeFunctionT::GradU + PlusIdentity,
 * FiniteStrainT.cpp:278-303). */
static void def_grad(const double* ue, const double dN[3][8], double* F)
{
    for (int i = 0; i < 3; i++)
        for (int k = 0; k < 3; k++) {
            double s = 0.0;
            for (int a = 0; a < 8; a++) s += ue[3 * a + i] * dN[k][a];
            M3(F, i, k) = s + (i == k ? 1.0 : 0.0);
        }
}

/* B (6 x 24), SolidElementT::Set_B 3D branch, SolidElementT.cpp:890-918 (Hughes 2.8.21) */
static void set_B(const double dN[3][8], double B[6][24])
{
    memset(B, 0, sizeof(double) * 6 * 24);
    for (int a = 0; a < 8; a++) {
        double nx = dN[0][a], ny = dN[1][a], nz = dN[2][a];
        B[0][3 * a + 0] = nx; B[4][3 * a + 0] = nz; B[5][3 * a + 0] = ny;
        B[1][3 * a + 1] = ny; B[3][3 * a + 1] = nz; B[5][3 * a + 1] = nx;
        B[2][3 * a + 2] = nz; B[3][3 * a + 2] = ny; B[4][3 * a + 2] = nx;
    }
}

/* ------------------------------------------------------------------------ */
/* A7/A8: TotalLagrangianT (elements/continuum/solid/TotalLagrangianT.cpp)   */
/* ------------------------------------------------------------------------ */

/* Re = FormKd(-1) (TotalLagrangianT.cpp:107-143, called with -constKd at
 * SolidElementT.cpp:1240), i.e. minus the internal force; Ke = FormStiffness(1)
 * (TotalLagrangianT.cpp:40-104) with both triangles filled.  Either output may
 * be NULL. */
static int tl_element(stress_fn cauchy, stress_fn moduli, const double Xe[24], const double ue[24], double mu,
                      double kappa, double Re[24], double Ke[576])
{
    const parent_t* P = parent();
    double det0[8], dNX[8][3][8];
    if (shape_derivs(Xe, det0, dNX)) return 7;
    if (Re) memset(Re, 0, 24 * sizeof(double));
    double K[24][24];
    double G[8][8];
    if (Ke) { memset(K, 0, sizeof K); memset(G, 0, sizeof G); }
    for (int q = 0; q < 8; q++) {
        double F[9], Fi[9], sig[6];
        def_grad(ue, dNX[q], F);
        if (cauchy(mu, kappa, F, sig)) return 7;
        double J = det3(F);
        if (J <= 0.0) return 7; /* TotalLagrangianT.cpp:127-130 */
        inv3(F, Fi);
        double S[9]; /* sigma * F^-T, full 3x3 */
        double sm[9];
        for (int v = 0; v < 6; v++) { M3(sm, VI[v], VJ[v]) = sig[v]; M3(sm, VJ[v], VI[v]) = sig[v]; }
        for (int i = 0; i < 3; i++)
            for (int k = 0; k < 3; k++) {
                double s = 0.0;
                for (int m = 0; m < 3; m++) s += M3(sm, i, m) * M3(Fi, k, m);
                M3(S, i, k) = s;
            }
        if (Re) {
            double f = J * (-1.0) * P->w[q] * det0[q];
            for (int a = 0; a < 8; a++)
                for (int i = 0; i < 3; i++) {
                    double s = 0.0;
                    for (int k = 0; k < 3; k++) s += M3(S, i, k) * dNX[q][k][a];
                    Re[3 * a + i] += f * s;
                }
        }
        if (Ke) {
            double dnx[3][8]; /* TransformDerivatives: dN/dx_j = sum_k F^-1(k,j) dN/dX_k */
            for (int a = 0; a < 8; a++)
                for (int j = 0; j < 3; j++)
                    dnx[j][a] = M3(Fi, 0, j) * dNX[q][0][a] + M3(Fi, 1, j) * dNX[q][1][a]
                              + M3(Fi, 2, j) * dNX[q][2][a];
            double scale = 1.0 * det0[q] * P->w[q] * J;
            /* geometric: G += gradN^T (scale sigma) gradN (MultQTBQ) */
            for (int a = 0; a < 8; a++)
                for (int b = 0; b < 8; b++) {
                    double s = 0.0;
                    for (int k = 0; k < 3; k++)
                        for (int l = 0; l < 3; l++) s += dnx[k][a] * (scale * M3(sm, k, l)) * dnx[l][b];
                    G[a][b] += s;
                }
            double c[36], B[6][24], DB[6][24];
            if (moduli(mu, kappa, F, c)) return 7;
            set_B(dnx, B);
            for (int i = 0; i < 6; i++)
                for (int n = 0; n < 24; n++) {
                    double s = 0.0;
                    for (int m = 0; m < 6; m++) s += scale * c[i + 6 * m] * B[m][n];
                    DB[i][n] = s;
                }
            for (int m = 0; m < 24; m++)
                for (int n = 0; n < 24; n++) {
                    double s = 0.0;
                    for (int i = 0; i < 6; i++) s += B[i][m] * DB[i][n];
                    K[m][n] += s;
                }
        }
    }
    if (Ke) {
        /* dMatrixT::Expand(G, 3) (dMatrixT.cpp:690-727) */
        for (int a = 0; a < 8; a++)
            for (int b = 0; b < 8; b++)
                for (int i = 0; i < 3; i++) K[3 * a + i][3 * b + i] += G[a][b];
        memcpy(Ke, K, sizeof K);
    }
    return 0;
}

int or_tl_element(const double Xe[24], const double ue[24], double mu, double kappa,
                  double Re[24], double Ke[576])
{
    return tl_element(or_simo_cauchy, or_simo_moduli, Xe, ue, mu, kappa, Re, Ke);
}

/* TotalLagrangianT with large_strain_StVenant */
int or_tl_stvenant_element(const double Xe[24], const double ue[24], double mu, double kappa,
                           double Re[24], double Ke[576])
{
    return tl_element(or_stvenant_cauchy, or_stvenant_moduli, Xe, ue, mu, kappa, Re, Ke);
}

/* ------------------------------------------------------------------------ */
/* A9: SimoQ1P0 (elements/continuum/solid/SimoQ1P0.cpp)                      */
/* ------------------------------------------------------------------------ */

typedef struct {
    double det0[8], dNX[8][3][8]; /* reference configuration */
    double det[8], dnx[8][3][8];  /* current configuration */
    double mean[3][8];            /* mean gradient, SetMeanGradient */
    double H, v;
    double F[8][9];               /* F_bar at each ip */
    double J[8];                  /* det F (unmodified) */
} q1p0_state_t;

/* SetGlobalShape (SimoQ1P0.cpp:222-311) / RecomputeDeformationState
 * (SimoQ1P0.cpp:359-387): current-configuration derivatives, Hughes (4.5.23)
 * mean gradient (SetMeanGradient, SimoQ1P0.cpp:661-686), and
 * F_bar = (v/(H J))^(1/3) F. */
static int q1p0_state(const double* Xe, const double* xe, const double* ue, q1p0_state_t* st, int need_ref)
{
    const parent_t* P = parent();
    if (need_ref && shape_derivs(Xe, st->det0, st->dNX)) return 7;
    if (shape_derivs(xe, st->det, st->dnx)) return 7;
    st->H = 0.0;
    st->v = 0.0;
    for (int q = 0; q < 8; q++) {
        st->H += P->w[q] * st->det0[q];
        st->v += P->w[q] * st->det[q];
    }
    memset(st->mean, 0, sizeof st->mean);
    for (int q = 0; q < 8; q++) {
        double s = P->w[q] * st->det[q] / st->v;
        for (int k = 0; k < 3; k++)
            for (int a = 0; a < 8; a++) st->mean[k][a] += s * st->dnx[q][k][a];
    }
    for (int q = 0; q < 8; q++) {
        double* F = st->F[q];
        def_grad(ue, st->dNX[q], F);
        double J = det3(F);
        st->J[q] = J;
        double s = pow(st->v / (st->H * J), 1.0 / 3.0);
        for (int k = 0; k < 9; k++) F[k] *= s;
    }
    return 0;
}

/* ComputeInternalForce / FormKd (SimoQ1P0.cpp:390-439, 528-582) with E = 0 and
 * no phase field: f = constK sum_q w det_q (J_bar/J_q) B_bar^T sigma(F_bar).
 * B_bar from SolidElementT::Set_B_bar (SolidElementT.cpp:1011-1041). */
static int q1p0_force(stress_fn cauchy, const q1p0_state_t* st, double mu, double kappa, double constK, double* f)
{
    const parent_t* P = parent();
    memset(f, 0, 24 * sizeof(double));
    for (int q = 0; q < 8; q++) {
        double sig[6];
        if (cauchy(mu, kappa, st->F[q], sig)) return 7;
        double nee[24];
        for (int a = 0; a < 8; a++) {
            double nx = st->dnx[q][0][a], ny = st->dnx[q][1][a], nz = st->dnx[q][2][a];
            double fx = (st->mean[0][a] - nx) / 3.0;
            double fy = (st->mean[1][a] - ny) / 3.0;
            double fz = (st->mean[2][a] - nz) / 3.0;
            double Bx[6] = {nx + fx, fx, fx, 0.0, nz, ny};
            double By[6] = {fy, ny + fy, fy, nz, 0.0, nx};
            double Bz[6] = {fz, fz, nz + fz, ny, nx, 0.0};
            double sx = 0, sy = 0, sz = 0;
            for (int v = 0; v < 6; v++) { sx += Bx[v] * sig[v]; sy += By[v] * sig[v]; sz += Bz[v] * sig[v]; }
            nee[3 * a + 0] = sx; nee[3 * a + 1] = sy; nee[3 * a + 2] = sz;
        }
        double Jbar = det3(st->F[q]);
        double Jc = Jbar / st->J[q];
        double s = constK * P->w[q] * st->det[q] * Jc;
        for (int k = 0; k < 24; k++) f[k] += s * nee[k];
    }
    return 0;
}

/* Re = FormKd(-1); Ke = FormStiffness(1): forward finite differences of the
 * internal force, eps = 1e-7, column by column (SimoQ1P0.cpp:446-525).  Ke is
 * the full (unsymmetric) FD matrix; assembly reads its upper triangle.  The
 * material sees F_bar through its Cauchy stress (SimoQ1P0 with any large-
 * strain material: Simo_isotropic, or large_strain_StVenant as in the
 * reference deck Q1P0.3D.a). */
static int q1p0_element(stress_fn cauchy, const double Xe[24], const double ue[24], double mu, double kappa,
                        double Re[24], double Ke[576])
{
    q1p0_state_t st;
    /* current coordinates x = X + u, formed once per node as the node
     * manager's current coordinates are */
    double xe[24], uw[24];
    for (int k = 0; k < 24; k++) {
        xe[k] = Xe[k] + ue[k];
        uw[k] = ue[k];
    }
    if (q1p0_state(Xe, xe, uw, &st, 1)) return 7;
    if (Re && q1p0_force(cauchy, &st, mu, kappa, -1.0, Re)) return 7;
    if (Ke) {
        const double eps = 1.0e-7;
        double f0[24], fp[24];
        if (q1p0_force(cauchy, &st, mu, kappa, 1.0, f0)) return 7;
        memset(Ke, 0, 576 * sizeof(double));
        for (int col = 0; col < 24; col++) {
            /* fLocDisp(a,i) += eps and fLocCurrCoords(a,i) += eps, undone
             * by subtracting eps after the column (SimoQ1P0.cpp:493-513):
             * the perturbed x is (X + u) + eps, and an entry keeps whatever
             * rounding (v + eps) - eps leaves for the later columns */
            uw[col] += eps;
            xe[col] += eps;
            q1p0_state_t sp;
            memcpy(sp.det0, st.det0, sizeof st.det0);
            memcpy(sp.dNX, st.dNX, sizeof st.dNX);
            if (q1p0_state(Xe, xe, uw, &sp, 0)) return 7;
            if (q1p0_force(cauchy, &sp, mu, kappa, 1.0, fp)) return 7;
            for (int j = 0; j < 24; j++) Ke[24 * j + col] += (fp[j] - f0[j]) / eps;
            uw[col] -= eps;
            xe[col] -= eps;
        }
    }
    return 0;
}

int or_q1p0_element(const double Xe[24], const double ue[24], double mu, double kappa, double Re[24],
                    double Ke[576])
{
    return q1p0_element(or_simo_cauchy, Xe, ue, mu, kappa, Re, Ke);
}

int or_q1p0_stvenant_element(const double Xe[24], const double ue[24], double mu, double kappa, double Re[24],
                             double Ke[576])
{
    return q1p0_element(or_stvenant_cauchy, Xe, ue, mu, kappa, Re, Ke);
}

/* ------------------------------------------------------------------------ */
/* A1: equation numbers                                                      */
/* ------------------------------------------------------------------------ */

/* NodeManagerT::SetEquationNumbers (tahoe/src/nodes/NodeManagerT.cpp:712-760)
 * -> FieldT::InitEquations/SetEquations (FieldT.cpp:635-690): dofs carrying a
 * kinematic card are marked -1, the rest numbered 1.. node-major, dof-minor. */
int or_equation_numbers(int nnd, int nkbc, const int* kbc_nodes, const int* kbc_dofs, int* eqnos)
{
    for (int k = 0; k < 3 * nnd; k++) eqnos[k] = 0;
    for (int k = 0; k < nkbc; k++) eqnos[3 * kbc_nodes[k] + kbc_dofs[k]] = -1;
    int n = 0;
    for (int k = 0; k < 3 * nnd; k++)
        if (eqnos[k] == 0) eqnos[k] = ++n;
    return n;
}

/* ------------------------------------------------------------------------ */
/* A2: upper-triangle MSR pattern                                            */
/* ------------------------------------------------------------------------ */

static int cmp_int(const void* a, const void* b)
{
    int x = *(const int*)a, y = *(const int*)b;
    return (x > y) - (x < y);
}


/* GraphT::MakeGraph(active, add_self, upper_only) (toolbox/src/graph/GraphT.cpp:376-456)
 * + MSRBuilderT::SetMSRData/GenerateMSR (MSRBuilderT.cpp:16-34, 134-181): row i
 * holds every active equation j > i sharing an element with i, ascending.
 * Written node-wise: the columns of a row are the equations of the nodes that
 * share an element with the row's node.  Returns the bindx length
 * (n_eq + 1 + #off-diagonals); fills bindx only if cap is large enough. */
/* sorted unique columns > row of the row (node nd, dof i) into cols; returns count */
static int row_columns(int nd, int row, const int* n2e_ptr, const int* n2e, const int* conn,
                       const int* eqnos, int* cols)
{
    int nc = 0;
    for (int p = n2e_ptr[nd]; p < n2e_ptr[nd + 1]; p++) {
        int e = n2e[p];
        for (int b = 0; b < 8; b++)
            for (int j = 0; j < 3; j++) {
                int col = eqnos[3 * conn[8 * e + b] + j];
                if (col > row) cols[nc++] = col - 1;
            }
    }
    qsort(cols, (size_t)nc, sizeof(int), cmp_int);
    int m = 0;
    for (int k = 0; k < nc; k++)
        if (m == 0 || cols[k] != cols[m - 1]) cols[m++] = cols[k];
    return m;
}

long or_msr_pattern(int n_eq, int nnd, int nel, const int* conn, const int* eqnos,
                    int* bindx, long cap)
{
    /* node -> elements (CSR) */
    int* ptr = (int*)calloc((size_t)nnd + 1, sizeof(int));
    for (int e = 0; e < nel; e++)
        for (int a = 0; a < 8; a++) ptr[conn[8 * e + a] + 1]++;
    for (int k = 0; k < nnd; k++) ptr[k + 1] += ptr[k];
    int* n2e = (int*)malloc(sizeof(int) * (size_t)(ptr[nnd] > 0 ? ptr[nnd] : 1));
    int* fill = (int*)calloc((size_t)nnd, sizeof(int));
    for (int e = 0; e < nel; e++)
        for (int a = 0; a < 8; a++) {
            int nd = conn[8 * e + a];
            n2e[ptr[nd] + fill[nd]++] = e;
        }
    int* cols = (int*)malloc(sizeof(int) * 8 * 24 * 64);
    int* rowlen = (int*)calloc((size_t)n_eq + 1, sizeof(int));
    long total = 0;
    for (int nd = 0; nd < nnd; nd++)
        for (int i = 0; i < 3; i++) {
            int row = eqnos[3 * nd + i];
            if (row <= 0) continue;
            rowlen[row - 1] = row_columns(nd, row, ptr, n2e, conn, eqnos, cols);
            total += rowlen[row - 1];
        }
    long len = (long)n_eq + 1 + total;
    if (bindx && cap >= len) {
        bindx[0] = n_eq + 1;
        for (int r = 0; r < n_eq; r++) bindx[r + 1] = bindx[r] + rowlen[r];
        for (int nd = 0; nd < nnd; nd++)
            for (int i = 0; i < 3; i++) {
                int row = eqnos[3 * nd + i];
                if (row <= 0) continue;
                int m = row_columns(nd, row, ptr, n2e, conn, eqnos, cols);
                memcpy(bindx + bindx[row - 1], cols, sizeof(int) * (size_t)m);
            }
    }
    free(rowlen);
    free(cols);
    free(fill);
    free(n2e);
    free(ptr);
    return len;
}

/* ------------------------------------------------------------------------ */
/* A10: assembly                                                             */
/* ------------------------------------------------------------------------ */

typedef int (*element_fn)(const double*, const double*, double, double, double*, double*);

static element_fn element_for(int type)
{
    switch (type) {
    case OR_Q1P0: return or_q1p0_element;
    case OR_TL_STVENANT: return or_tl_stvenant_element;
    case OR_Q1P0_STVENANT: return or_q1p0_stvenant_element;
    default: return or_tl_element;
    }
}

static void gather(const double* A, const int* conn, int e, double* ae)
{
    for (int a = 0; a < 8; a++)
        for (int i = 0; i < 3; i++) ae[3 * a + i] = A[3 * conn[8 * e + a] + i];
}

/* SolidElementT::ElementRHSDriver (SolidElementT.cpp:1166-1295) in element
 * order, SolverT::AssembleRHS (SolverT.cpp:446-477): R[eq-1] += Re for eq > 0. */
int or_residual(int type, int nnd, int nel, const double* X, const int* conn, const int* mat,
                const double* mu, const double* kappa, const double* u, const int* eqnos,
                int n_eq, double* R)
{
    (void)nnd;
    element_fn fn = element_for(type);
    for (int k = 0; k < n_eq; k++) R[k] = 0.0;
    for (int e = 0; e < nel; e++) {
        double Xe[24], ue[24], Re[24];
        gather(X, conn, e, Xe);
        gather(u, conn, e, ue);
        int m = mat ? mat[e] : 0;
        if (fn(Xe, ue, mu[m], kappa[m], Re, NULL)) return 7;
        for (int a = 0; a < 8; a++)
            for (int i = 0; i < 3; i++) {
                int eq = eqnos[3 * conn[8 * e + a] + i];
                if (eq > 0) R[eq - 1] += Re[3 * a + i];
            }
    }
    return 0;
}

/* MSRMatrixT::Assemble symmetric branch (MSRMatrixT.cpp:110-165): active element
 * dofs sorted by global equation; row i gets, for every j >= i in that order,
 * the element upper-triangle value K(min(dex),max(dex)), doubled when the same
 * global equation repeats; AssembleRow (MSRMatrixT.cpp:532-604) merges into
 * the diagonal val[row] or the matching off-diagonal slot. */
static int msr_assemble_element(const double* Ke, const int* eq, const int* bindx, double* val)
{
    int dex[24], act[24], na = 0;
    for (int j = 0; j < 24; j++)
        if (eq[j] > 0) { dex[na] = j; act[na] = eq[j] - 1; na++; }
    /* stable insertion sort of active rows by equation (iArrayT::SortAscending) */
    for (int i = 1; i < na; i++) {
        int d = dex[i], a = act[i], k = i - 1;
        while (k >= 0 && act[k] > a) { act[k + 1] = act[k]; dex[k + 1] = dex[k]; k--; }
        act[k + 1] = a; dex[k + 1] = d;
    }
    for (int i = 0; i < na; i++) {
        int row = act[i], di = dex[i];
        int start = bindx[row], end = bindx[row + 1];
        for (int j = i; j < na; j++) {
            int dj = dex[j], col = act[j];
            double v = (dj >= di) ? Ke[24 * di + dj] : Ke[24 * dj + di];
            if (col == row && di != dj) v *= 2.0;
            if (col == row) { val[row] += v; continue; }
            /* binary search the sorted row */
            int lo = start, hi = end - 1, hit = -1;
            while (lo <= hi) {
                int mid = (lo + hi) / 2;
                if (bindx[mid] == col) { hit = mid; break; }
                if (bindx[mid] < col) lo = mid + 1; else hi = mid - 1;
            }
            if (hit < 0) return 1; /* "assembly error" GeneralFail */
            val[hit] += v;
        }
    }
    return 0;
}

/* SolidElementT::ElementLHSDriver (SolidElementT.cpp:1100-1153) in element
 * order after GlobalMatrixT::Clear (MSRMatrixT.cpp:423). */
int or_stiffness(int type, int nnd, int nel, const double* X, const int* conn, const int* mat,
                 const double* mu, const double* kappa, const double* u, const int* eqnos,
                 int n_eq, const int* bindx, double* val)
{
    (void)nnd;
    element_fn fn = element_for(type);
    long len = bindx[n_eq];
    for (long k = 0; k < len; k++) val[k] = 0.0;
    double* Ke = (double*)malloc(576 * sizeof(double));
    int status = 0;
    for (int e = 0; e < nel && !status; e++) {
        double Xe[24], ue[24];
        int eq[24];
        gather(X, conn, e, Xe);
        gather(u, conn, e, ue);
        for (int a = 0; a < 8; a++)
            for (int i = 0; i < 3; i++) eq[3 * a + i] = eqnos[3 * conn[8 * e + a] + i];
        int m = mat ? mat[e] : 0;
        if (fn(Xe, ue, mu[m], kappa[m], NULL, Ke)) { status = 7; break; }
        status = msr_assemble_element(Ke, eq, bindx, val);
    }
    free(Ke);
    return status;
}

/* ------------------------------------------------------------------------ */
/* Multi-threaded element passes (CPU baseline only)                         */
/* ------------------------------------------------------------------------ */

/* Greedy element colouring: an element takes the lowest colour that none of
 * the elements sharing one of its nodes has taken.  Elements of one colour
 * share no node, so they write disjoint R entries and disjoint MSR rows.
 * Returns the colour count (at most 64), or -1. */
int or_color_elements(int nnd, int nel, const int* conn, int* color_ptr, int* color_elems)
{
    unsigned long long* used = (unsigned long long*)calloc((size_t)nnd, sizeof(unsigned long long));
    int* col = (int*)malloc(sizeof(int) * (size_t)(nel > 0 ? nel : 1));
    if (!used || !col) { free(used); free(col); return -1; }
    int ncol = 0;
    for (int e = 0; e < nel; e++) {
        unsigned long long m = 0;
        for (int a = 0; a < 8; a++) m |= used[conn[8 * e + a]];
        int c = 0;
        while (c < 64 && (m >> c) & 1ULL) c++;
        if (c == 64) { free(used); free(col); return -1; }
        col[e] = c;
        for (int a = 0; a < 8; a++) used[conn[8 * e + a]] |= 1ULL << c;
        if (c + 1 > ncol) ncol = c + 1;
    }
    for (int c = 0; c <= ncol; c++) color_ptr[c] = 0;
    for (int e = 0; e < nel; e++) color_ptr[col[e] + 1]++;
    for (int c = 0; c < ncol; c++) color_ptr[c + 1] += color_ptr[c];
    int* fill = (int*)malloc(sizeof(int) * (size_t)(ncol + 1));
    if (!fill) { free(used); free(col); return -1; }
    memcpy(fill, color_ptr, sizeof(int) * (size_t)ncol);
    for (int e = 0; e < nel; e++) color_elems[fill[col[e]]++] = e;
    free(fill);
    free(used);
    free(col);
    return ncol;
}

/* The residual and/or stiffness pass of or_residual / or_stiffness with the
 * elements of each colour spread over nthreads OpenMP threads (R and val may
 * be NULL to skip that pass).  Same element arithmetic; the summation order
 * into shared entries follows the colour order instead of the element
 * order. */
int or_element_passes_parallel(int type, int nel, const double* X, const int* conn, const int* mat,
                               const double* mu, const double* kappa, const double* u, const int* eqnos,
                               int n_eq, const int* bindx, double* R, double* val, int ncolors,
                               const int* color_ptr, const int* color_elems, int nthreads)
{
    (void)nel;
    element_fn fn = element_for(type);
    (void)parent();  /* its one-time set-up happens here, before the threads start */
    int status = 0;
    if (R)
        for (int k = 0; k < n_eq; k++) R[k] = 0.0;
    if (val) {
        long len = bindx[n_eq];
        for (long k = 0; k < len; k++) val[k] = 0.0;
    }
    for (int c = 0; c < ncolors; c++) {
        const int b = color_ptr[c], en = color_ptr[c + 1];
#pragma omp parallel for schedule(static) num_threads(nthreads)
        for (int i = b; i < en; i++) {
            const int e = color_elems[i];
            double Xe[24], ue[24], Re[24], Ke[576];
            int eq[24];
            gather(X, conn, e, Xe);
            gather(u, conn, e, ue);
            for (int a = 0; a < 24; a++) eq[a] = eqnos[3 * conn[8 * e + a / 3] + a % 3];
            const int m = mat ? mat[e] : 0;
            int rc = fn(Xe, ue, mu[m], kappa[m], R ? Re : NULL, val ? Ke : NULL) ? 7 : 0;
            if (!rc && R)
                for (int a = 0; a < 24; a++)
                    if (eq[a] > 0) R[eq[a] - 1] += Re[a];
            if (!rc && val) rc = msr_assemble_element(Ke, eq, bindx, val);
            if (rc) {
#pragma omp atomic write
                status = rc;
            }
        }
        if (status) break;
    }
    return status;
}

/* ------------------------------------------------------------------------ */
/* A13: SpMV, Jacobi-PCG                                                     */
/* ------------------------------------------------------------------------ */

/* MSRMatrixT::Multx symmetric (MSRMatrixT.cpp:385-420) */
void or_msr_multx(int n, const int* bindx, const double* val, const double* x, double* b)
{
    for (int r = 0; r < n; r++) b[r] = 0.0;
    for (int r = 0; r < n; r++) {
        b[r] += val[r] * x[r];
        for (int j = bindx[r]; j < bindx[r + 1]; j++) {
            int c = bindx[j];
            b[r] += val[j] * x[c];
            b[c] += val[j] * x[r];
        }
    }
}

static double dot(int n, const double* a, const double* b)
{
    double s = 0.0;
    for (int k = 0; k < n; k++) s += a[k] * b[k];
    return s;
}

/* Conjugate gradients with the Jacobi preconditioner of DiagonalMatrixT
 * (DiagonalMatrixT.cpp:267-324: z = r / diag), x0 = 0, stop on
 * ||r|| <= rtol ||b||.  The reference keeps these pieces (Multx, CopyDiagonal
 * MSRMatrixT.cpp:351-378, InnerProduct SolverT.cpp:854-860) but has no linear
 * PCG of its own; this is the textbook algorithm (Saad, Iterative Methods,
 * Alg. 9.1). */
int or_pcg_jacobi(int n, const int* bindx, const double* val, const double* b, double* x,
                  double rtol, int maxit, int* iters, double* relres)
{
    double* r = (double*)malloc(sizeof(double) * (size_t)n * 4);
    double* z = r + n;
    double* p = z + n;
    double* q = p + n;
    for (int k = 0; k < n; k++) { x[k] = 0.0; r[k] = b[k]; }
    double bn = sqrt(dot(n, b, b));
    int it = 0;
    double rn = bn;
    if (bn == 0.0) { *iters = 0; *relres = 0.0; free(r); return 0; }
    for (int k = 0; k < n; k++) z[k] = r[k] / val[k];
    for (int k = 0; k < n; k++) p[k] = z[k];
    double rz = dot(n, r, z);
    while (it < maxit) {
        or_msr_multx(n, bindx, val, p, q);
        double alpha = rz / dot(n, p, q);
        for (int k = 0; k < n; k++) { x[k] += alpha * p[k]; r[k] -= alpha * q[k]; }
        it++;
        rn = sqrt(dot(n, r, r));
        if (rn <= rtol * bn) break;
        for (int k = 0; k < n; k++) z[k] = r[k] / val[k];
        double rz1 = dot(n, r, z);
        double beta = rz1 / rz;
        rz = rz1;
        for (int k = 0; k < n; k++) p[k] = z[k] + beta * p[k];
    }
    *iters = it;
    *relres = rn / bn;
    free(r);
    return (rn <= rtol * bn) ? 0 : 1;
}

/* ------------------------------------------------------------------------ */
/* A11: static Newton time loop                                              */
/* ------------------------------------------------------------------------ */

/* Restarted GMRES(m) with left Jacobi preconditioning, NLSolver_NK::GMRES
 * (NLSolver_NK.cpp:121-238): M = diag(A) (1 where |A_ii| <= 1e-14); each
 * restart cycle starts from r = M^-1 (b - A x), beta = |r|, and stops when the
 * Givens-rotated residual |g_{j+1}| < tol * beta (tolerance relative to the
 * cycle's own start); Arnoldi by modified Gram-Schmidt; Givens rotations
 * [c s; -s c] (NLSolver_NK.cpp:241-246); back substitution skipping
 * |H_ii| <= 1e-14; at most maxit Arnoldi steps in total (the reference's
 * 3 n when max_linear_iter = 0).  matvec(ctx, x, y) computes y = A x.
 * Returns 1 when converged; *steps gets the Arnoldi steps taken. */
typedef void (*matvec_fn)(const void* ctx, const double* x, double* y);

static int gmres(int n, matvec_fn matvec, const void* ctx, const double* diag, const double* b, double* x, int m,
                 double tol, int maxit, int* steps)
{
    double* Q = (double*)calloc((size_t)(m + 1) * n, sizeof(double));
    double* H = (double*)calloc((size_t)(m + 1) * m, sizeof(double));
    double* g = (double*)calloc((size_t)m + 1, sizeof(double));
    double* cs = (double*)calloc((size_t)m, sizeof(double));
    double* sn = (double*)calloc((size_t)m, sizeof(double));
    double* y = (double*)calloc((size_t)m, sizeof(double));
    double* w = (double*)calloc((size_t)n, sizeof(double));
    double* z = (double*)calloc((size_t)n, sizeof(double));
    double* md = (double*)calloc((size_t)n, sizeof(double));
#define HH(i, j) H[(size_t)(i) * m + (j)]
    for (int i = 0; i < n; i++) md[i] = fabs(diag[i]) > 1.0e-14 ? 1.0 / diag[i] : 1.0;
    int total = 0, done = 0;
    while (!done && total < maxit) {
        matvec(ctx, x, z);
        for (int i = 0; i < n; i++) w[i] = (b[i] - z[i]) * md[i];
        double beta = sqrt(dot(n, w, w));
        if (beta == 0.0) { done = 1; break; }
        const double tol_abs = tol * beta;
        if (tol_abs == 0.0) { done = 1; break; }
        for (int i = 0; i < n; i++) Q[i] = w[i] / beta;
        for (int i = 0; i <= m; i++) g[i] = 0.0;
        g[0] = beta;
        int j_end = 0;
        for (int j = 0; j < m && total < maxit; j++, total++) {
            j_end = j;
            matvec(ctx, Q + (size_t)j * n, w);
            for (int i = 0; i < n; i++) w[i] *= md[i];
            for (int i = 0; i <= j; i++) {
                const double h = dot(n, w, Q + (size_t)i * n);
                HH(i, j) = h;
                for (int k = 0; k < n; k++) w[k] -= h * Q[(size_t)i * n + k];
            }
            const double hn = sqrt(dot(n, w, w));
            HH(j + 1, j) = hn;
            if (hn > 0.0)
                for (int k = 0; k < n; k++) Q[(size_t)(j + 1) * n + k] = w[k] / hn;
            for (int i = 0; i < j; i++) {
                const double t = cs[i] * HH(i, j) + sn[i] * HH(i + 1, j);
                HH(i + 1, j) = -sn[i] * HH(i, j) + cs[i] * HH(i + 1, j);
                HH(i, j) = t;
            }
            const double den = sqrt(HH(j, j) * HH(j, j) + HH(j + 1, j) * HH(j + 1, j));
            cs[j] = den > 0.0 ? HH(j, j) / den : 1.0;
            sn[j] = den > 0.0 ? HH(j + 1, j) / den : 0.0;
            double t = cs[j] * HH(j, j) + sn[j] * HH(j + 1, j);
            HH(j + 1, j) = -sn[j] * HH(j, j) + cs[j] * HH(j + 1, j);
            HH(j, j) = t;
            t = cs[j] * g[j] + sn[j] * g[j + 1];
            g[j + 1] = -sn[j] * g[j] + cs[j] * g[j + 1];
            g[j] = t;
            if (fabs(g[j + 1]) < tol_abs) {
                j_end = j;
                done = 1;
                break;
            }
        }
        for (int i = j_end; i >= 0; i--) {
            double sum = g[i];
            for (int l = i + 1; l <= j_end; l++) sum -= HH(i, l) * y[l];
            y[i] = fabs(HH(i, i)) > 1.0e-14 ? sum / HH(i, i) : 0.0;
        }
        for (int i = 0; i <= j_end; i++)
            for (int k = 0; k < n; k++) x[k] += y[i] * Q[(size_t)i * n + k];
    }
#undef HH
    if (steps) *steps = total;
    free(Q); free(H); free(g); free(cs); free(sn); free(y); free(w); free(z); free(md);
    return done;
}

typedef struct { int n; const double* A; } dense_t;
static void dense_mv(const void* ctx, const double* x, double* y)
{
    const dense_t* d = (const dense_t*)ctx;
    for (int i = 0; i < d->n; i++) {
        double s = 0.0;
        for (int k = 0; k < d->n; k++) s += d->A[(size_t)i * d->n + k] * x[k];
        y[i] = s;
    }
}

int or_gmres_dense(int n, const double* A, const double* b, double* x, int m, double tol, int maxit, int* steps)
{
    dense_t d = {n, A};
    double* diag = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    for (int i = 0; i < n; i++) diag[i] = A[(size_t)i * n + i];
    const int rc = gmres(n, dense_mv, &d, diag, b, x, m, tol, maxit > 0 ? maxit : 3 * n, steps);
    free(diag);
    return rc;
}

typedef struct { int n; const int* bindx; const double* val; } msr_t;
static void msr_mv(const void* ctx, const double* x, double* y)
{
    const msr_t* d = (const msr_t*)ctx;
    or_msr_multx(d->n, d->bindx, d->val, x, y);
}

int or_gmres_msr(int n, const int* bindx, const double* val, const double* b, double* x, int m, double tol,
                 int maxit, int* steps)
{
    msr_t d = {n, bindx, val};
    return gmres(n, msr_mv, &d, val, b, x, m, tol, maxit > 0 ? maxit : 3 * n, steps);
}

/* R += t * f at the active dofs of the force cards (FieldT::FormRHS,
 * FieldT.cpp:390-401, with FBC_CardT::CurrentValue, FBC_CardT.cpp:57-63,
 * on the schedule (0,0)-(1,1)). */
static void add_forces(int nf, const int* fn, const int* fd, const double* fv, double t, const int* eqnos,
                       double* R)
{
    for (int k = 0; k < nf; k++) {
        const int e = eqnos[3 * fn[k] + fd[k]];
        if (e > 0) R[e - 1] += fv[k] * t;
    }
}

/* FEManagerT::Solve (FEManagerT.cpp:138-269) with the static integrator:
 * per step t = s/num_steps, nStaticIntegrator::ConsistentKBC
 * (nStaticIntegrator.cpp:18-46) sets u = value * t on each card (schedule
 * (0,0)-(1,1)) and the force cards contribute value * t to R;
 * NLSolver::Solve (NLSolver.cpp:57-264) iterates
 * R -> ||R|| -> ExitIteration (NLSolver.cpp:675-757) -> K -> solve -> u += du
 * (nStaticIntegrator::Corrector, nStaticIntegrator.cpp:69-87). */
int or_static_solve(int type, int nnd, int nel, const double* X, const int* conn, const int* mat,
                    const double* mu, const double* kappa, int nkbc, const int* kbc_nodes,
                    const int* kbc_dofs, const double* kbc_values, int nforce, const int* f_nodes,
                    const int* f_dofs, const double* f_values, int num_steps,
                    const or_newton_opts* o, double* u, double* step_u, int* step_iters,
                    int* lin_iters_total)
{
    int* eqnos = (int*)malloc(sizeof(int) * 3 * (size_t)nnd);
    int n_eq = or_equation_numbers(nnd, nkbc, kbc_nodes, kbc_dofs, eqnos);
    long len = or_msr_pattern(n_eq, nnd, nel, conn, eqnos, NULL, 0);
    int* bindx = (int*)malloc(sizeof(int) * (size_t)len);
    or_msr_pattern(n_eq, nnd, nel, conn, eqnos, bindx, len);
    double* val = (double*)malloc(sizeof(double) * (size_t)len);
    double* R = (double*)malloc(sizeof(double) * (size_t)(n_eq > 0 ? n_eq : 1) * 2);
    double* du = R + n_eq;
    int status = 0, lin_total = 0;
    for (int s = 1; s <= num_steps && !status; s++) {
        double t = (double)s / (double)num_steps;
        for (int k = 0; k < nkbc; k++) u[3 * kbc_nodes[k] + kbc_dofs[k]] = kbc_values[k] * t;
        if (or_residual(type, nnd, nel, X, conn, mat, mu, kappa, u, eqnos, n_eq, R)) { status = 7; break; }
        add_forces(nforce, f_nodes, f_dofs, f_values, t, eqnos, R);
        double err0 = sqrt(dot(n_eq, R, R));
        int iter = -1, done = (err0 < o->abs_tol);
        while (!done) {
            if (or_stiffness(type, nnd, nel, X, conn, mat, mu, kappa, u, eqnos, n_eq, bindx, val)) { status = 7; break; }
            int its; double rr;
            or_pcg_jacobi(n_eq, bindx, val, R, du, o->lin_rtol, o->lin_maxit, &its, &rr);
            lin_total += its;
            for (int k = 0; k < 3 * nnd; k++)
                if (eqnos[k] > 0) u[k] += du[eqnos[k] - 1];
            iter++;
            if (or_residual(type, nnd, nel, X, conn, mat, mu, kappa, u, eqnos, n_eq, R)) { status = 7; break; }
            add_forces(nforce, f_nodes, f_dofs, f_values, t, eqnos, R);
            double err = sqrt(dot(n_eq, R, R));
            double rel = err / err0;
            if (rel > o->div_tol) { status = 1; break; }
            if (rel < o->rel_tol || err < o->abs_tol) done = 1;
            else if (iter >= o->max_iterations) { status = 1; break; }
        }
        if (step_iters) step_iters[s - 1] = iter + 1;
        if (step_u) memcpy(step_u + (size_t)(s - 1) * 3 * nnd, u, sizeof(double) * 3 * (size_t)nnd);
    }
    if (lin_iters_total) *lin_iters_total = lin_total;
    free(R);
    free(val);
    free(bindx);
    free(eqnos);
    return status;
}
