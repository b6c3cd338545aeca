// Element kernels: total-Lagrangian Q1 hexahedron residual and stiffness with
// the SimoIso3D material, scattered colour by colour (no atomics) into the
// node-space residual and the upper-triangle MSR values.
//
// Lane mapping: 8 consecutive lanes own one element, lane q evaluates
// integration point q.  Nodal coordinates/displacements are gathered once per
// element (lane q loads node q) into LDS and read by all 8 lanes.
#include "tg_internal.h"
#include "tg_kinematics.h"

namespace tg {

typedef double double4_t __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// residual
// ---------------------------------------------------------------------------

// Sum of 8 lanes' 24 partials, scattered: lane q ends with node q's 3 sums.
__device__ __forceinline__ void reduce_scatter8(const double r[24], int q, double out[3])
{
    double s12[12];
#pragma unroll
    for (int k = 0; k < 12; k++) {
        const bool hi = q & 4;
        const double keep = hi ? r[12 + k] : r[k];
        const double send = hi ? r[k] : r[12 + k];
        s12[k] = keep + __shfl_xor(send, 4);
    }
    double s6[6];
#pragma unroll
    for (int k = 0; k < 6; k++) {
        const bool hi = q & 2;
        const double keep = hi ? s12[6 + k] : s12[k];
        const double send = hi ? s12[k] : s12[6 + k];
        s6[k] = keep + __shfl_xor(send, 2);
    }
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const bool hi = q & 1;
        const double keep = hi ? s6[3 + k] : s6[k];
        const double send = hi ? s6[k] : s6[3 + k];
        out[k] = keep + __shfl_xor(send, 1);
    }
}

// R -= f_int for the elements of one colour.  Per integration point
// (TotalLagrangianT::FormKd, TotalLagrangianT.cpp:107-143, called with
// constK = -1 from SolidElementT::ElementRHSDriver, SolidElementT.cpp:1240):
//   R_a,i += -J w det0 sum_k (sigma F^-T)_ik dN_a/dX_k
__global__ void __launch_bounds__(256) k_residual_tl(const int* __restrict__ elems, int count,
                                                     const int* __restrict__ conn, const int* __restrict__ emat,
                                                     const double2* __restrict__ props,
                                                     const double* __restrict__ X, const double* __restrict__ U,
                                                     const int* __restrict__ eqn, double* __restrict__ R,
                                                     int* __restrict__ status)
{
    __shared__ double sxu[32][8][6];
    const int t = threadIdx.x, le = t >> 3, q = t & 7;
    const long slot = (long)blockIdx.x * 32 + le;
    const bool valid = slot < count;
    int e = 0, node = 0;
    if (valid) {
        e = elems[slot];
        node = conn[8L * e + q];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            sxu[le][q][c] = X[3L * node + c];
            sxu[le][q][3 + c] = U[3L * node + c];
        }
    }
    __syncthreads();
    double r[24];
#pragma unroll
    for (int k = 0; k < 24; k++) r[k] = 0.0;
    bool bad = false;
    if (valid) {
        const double2 mk = props[emat[e]];
        double d[3][8], j0[9], ji[9], g[3][8], F[9], Fi[9];
        parent_derivs(q, d);
        jacobian([&](int a, int c) { return sxu[le][a][c]; }, d, j0);
        const double det0 = det3(j0);
        bad |= !(det0 > 0.0);
        inv3(j0, det0, ji);
        spatial_derivs(ji, d, g);
        deformation_gradient([&](int a, int c) { return sxu[le][a][3 + c]; }, g, F);
        const double J = det3(F);
        bad |= !(J > 0.0);
        double Jb, bb[6], sig[6];
        bad |= !simo_kinematics(F, Jb, bb);
        simo_cauchy(mk.x, mk.y, Jb, bb, sig);
        inv3(F, J, Fi);
        // S = sigma F^-T ; S(i,k) = sum_m sigma(i,m) Fi(k,m)
        const double sm[9] = {sig[0], sig[5], sig[4], sig[5], sig[1], sig[3], sig[4], sig[3], sig[2]};
        double S[9];
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int k = 0; k < 3; k++)
                S[i + 3 * k] = sm[i + 0] * Fi[k + 0] + sm[i + 3] * Fi[k + 3] + sm[i + 6] * Fi[k + 6];
        const double f = -J * det0;  // weight 1
#pragma unroll
        for (int a = 0; a < 8; a++)
#pragma unroll
            for (int i = 0; i < 3; i++)
                r[3 * a + i] = f * (S[i] * g[0][a] + S[i + 3] * g[1][a] + S[i + 6] * g[2][a]);
        if (bad) {
#pragma unroll
            for (int k = 0; k < 24; k++) r[k] = 0.0;
        }
    }
    double v[3];
    reduce_scatter8(r, q, v);
    if (valid) {
#pragma unroll
        for (int i = 0; i < 3; i++)
            if (eqn[3L * node + i] >= 0) R[3L * node + i] += v[i];
    }
    if (bad) atomicOr(status, 7);
}

void launch_q1p0(tg_ctx* c, const double* u, double* R_node, bool stiff);

void launch_residual(tg_ctx* c, const double* u, double* R_node)
{
    TimedRegion tr(c, T_RESIDUAL);
    TG_HIP(hipMemsetAsync(R_node, 0, sizeof(double) * 3 * (size_t)c->nnd, c->stream));
    if (c->formulation == TG_Q1P0) {
        launch_q1p0(c, u, R_node, false);
        return;
    }
    for (int col = 0; col < c->ncolors; col++) {
        const int first = c->color_ptr[col], count = c->color_ptr[col + 1] - first;
        if (count == 0) continue;
        const int blocks = (count + 31) / 32;
        if (c->formulation == TG_TOTAL_LAGRANGIAN)
            hipLaunchKernelGGL(k_residual_tl, dim3(blocks), dim3(256), 0, c->stream, c->dcolor_elems.p + first,
                               count, c->dconn.p, c->dmat.p, c->dmatprops.p, c->dX.p, u, c->deqn.p, R_node,
                               c->dstatus.p);
    }
    TG_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// stiffness (FP64 MFMA)
// ---------------------------------------------------------------------------

// Position of (m, n), m <= n < 24, in the packed upper triangle.
__device__ __forceinline__ int upper_index(int m, int n) { return m * 24 - (m * (m - 1)) / 2 + (n - m); }

// Element stiffness (TotalLagrangianT::FormStiffness, TotalLagrangianT.cpp:40-104)
//   K = sum_q B_q^T (s_q c_q) B_q  +  Expand(sum_q gradN^T (s_q sigma_q) gradN, 3)
// with s_q = det0 w J, B from dN/dx.  The material part is one 24 x 48 x 24
// product per element, Bs^T * DBs, run as 16x16x4 FP64 MFMA tiles (rows and
// columns 0-15 / 16-31, the lower-left tile is skipped: only the upper
// triangle is assembled, MSRMatrixT.cpp:110-165).  A wave handles 8 elements:
// all 64 lanes first evaluate the 8x8 integration points, then the wave runs
// the MFMA tiles element by element and scatters through the slot map.
__global__ void __launch_bounds__(64) k_stiffness_tl(const int* __restrict__ elems, int count,
                                                     const int* __restrict__ conn, const int* __restrict__ emat,
                                                     const double2* __restrict__ props,
                                                     const double* __restrict__ X, const double* __restrict__ U,
                                                     const int* __restrict__ slot, double* __restrict__ val,
                                                     int* __restrict__ status)
{
    __shared__ double sxu[8][8][6];
    __shared__ double sg[8][8][3][8];   // dN/dx [element][ip][dir][node]
    __shared__ double sd[8][8][36];     // s_q c_q [element][ip][v*6+w]
    __shared__ double ss[8][8][6];      // s_q sigma_q
    __shared__ double sgeo[8][64];      // geometric stiffness G[a*8+b]
    const int t = threadIdx.x, le = t >> 3, q = t & 7;
    const long base = (long)blockIdx.x * 8;
    const bool valid = base + le < count;
    int e = 0;
    bool bad = false;
    if (valid) {
        e = elems[base + le];
        const int node = conn[8L * e + q];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            sxu[le][q][c] = X[3L * node + c];
            sxu[le][q][3 + c] = U[3L * node + c];
        }
    }
    __syncthreads();
    // ---- phase 1: integration point data ----
    {
        double d[3][8], j0[9], ji[9], g0[3][8], F[9], Fi[9];
        double scale = 0.0, Jb = 1.0, bb[6] = {1, 1, 1, 0, 0, 0}, sig[6] = {0, 0, 0, 0, 0, 0};
        double2 mk = make_double2(0.0, 0.0);
        if (valid) {
            mk = props[emat[e]];
            parent_derivs(q, d);
            jacobian([&](int a, int c) { return sxu[le][a][c]; }, d, j0);
            const double det0 = det3(j0);
            bad |= !(det0 > 0.0);
            inv3(j0, det0, ji);
            spatial_derivs(ji, d, g0);
            deformation_gradient([&](int a, int c) { return sxu[le][a][3 + c]; }, g0, F);
            const double J = det3(F);
            bad |= !(J > 0.0);
            bad |= !simo_kinematics(F, Jb, bb);
            simo_cauchy(mk.x, mk.y, Jb, bb, sig);
            inv3(F, J, Fi);
            scale = det0 * J;  // weight 1
            // dN/dx_j = sum_k Fi(k,j) dN/dX_k (ShapeFunctionT::TransformDerivatives)
#pragma unroll
            for (int a = 0; a < 8; a++)
#pragma unroll
                for (int j = 0; j < 3; j++)
                    sg[le][q][j][a] = Fi[0 + 3 * j] * g0[0][a] + Fi[1 + 3 * j] * g0[1][a] + Fi[2 + 3 * j] * g0[2][a];
            if (bad) scale = 0.0;
        } else {
#pragma unroll
            for (int a = 0; a < 8; a++)
#pragma unroll
                for (int j = 0; j < 3; j++) sg[le][q][j][a] = 0.0;
        }
        double cm[6][6];
        simo_moduli(mk.x, mk.y, Jb, bb, cm);
#pragma unroll
        for (int v = 0; v < 6; v++)
#pragma unroll
            for (int w = 0; w < 6; w++) sd[le][q][v * 6 + w] = scale * cm[v][w];
#pragma unroll
        for (int k = 0; k < 6; k++) ss[le][q][k] = scale * sig[k];
    }
    __syncthreads();
    // ---- geometric stiffness G_ab = sum_q gradN_a . (s sigma) gradN_b ----
    for (int task = t; task < 8 * 64; task += 64) {
        const int el = task >> 6, ab = task & 63, a = ab >> 3, b = ab & 7;
        double s = 0.0;
        for (int qq = 0; qq < 8; qq++) {
            const double* sq = ss[el][qq];
            const double ga0 = sg[el][qq][0][a], ga1 = sg[el][qq][1][a], ga2 = sg[el][qq][2][a];
            const double gb0 = sg[el][qq][0][b], gb1 = sg[el][qq][1][b], gb2 = sg[el][qq][2][b];
            // sigma Voigt (11,22,33,23,13,12)
            s += ga0 * (sq[0] * gb0 + sq[5] * gb1 + sq[4] * gb2) + ga1 * (sq[5] * gb0 + sq[1] * gb1 + sq[3] * gb2) +
                 ga2 * (sq[4] * gb0 + sq[3] * gb1 + sq[2] * gb2);
        }
        sgeo[el][ab] = s;
    }
    __syncthreads();
    // ---- phase 2: MFMA tiles per element ----
    const int lr = t & 15, lk = t >> 4;  // operand row/col within tile, k within step
    for (int el = 0; el < 8; el++) {
        if (base + el >= count) break;
        double4_t acc00 = {0, 0, 0, 0}, acc01 = {0, 0, 0, 0}, acc11 = {0, 0, 0, 0};
        // A operand rows m0 = lr (tile 0), m1 = 16 + lr (tile 1); B operand columns n likewise
        const int m0 = lr, a0 = m0 / 3, i0 = m0 % 3;
        const int m1 = 16 + lr, a1 = m1 / 3, i1 = m1 % 3;
        const bool m1ok = m1 < 24;
#pragma unroll 4
        for (int ks = 0; ks < 12; ks++) {
            const int kk = 4 * ks + lk, qq = kk / 6, v = kk % 6;
            const double* D = sd[el][qq] + v * 6;
            // A = Bs^T(m, kk) = B_q[v][m]
            const int c0 = bcomp(v, i0);
            const double A0 = c0 >= 0 ? sg[el][qq][c0][a0] : 0.0;
            const int c1 = m1ok ? bcomp(v, i1) : -1;
            const double A1 = c1 >= 0 ? sg[el][qq][c1][a1] : 0.0;
            // B = DBs(kk, n) = sum_w D[v][w] B_q[w][n]
            double B0, B1;
            {
                const int n = m0, b = a0, j = i0;
                double s = 0.0;
#pragma unroll
                for (int w = 0; w < 6; w++) {
                    const int cw = bcomp(w, j);
                    if (cw >= 0) s += D[w] * sg[el][qq][cw][b];
                }
                B0 = s;
                (void)n;
            }
            {
                const int b = a1, j = i1;
                double s = 0.0;
                if (m1ok) {
#pragma unroll
                    for (int w = 0; w < 6; w++) {
                        const int cw = bcomp(w, j);
                        if (cw >= 0) s += D[w] * sg[el][qq][cw][b];
                    }
                }
                B1 = s;
            }
            acc00 = __builtin_amdgcn_mfma_f64_16x16x4f64(A0, B0, acc00, 0, 0, 0);
            acc01 = __builtin_amdgcn_mfma_f64_16x16x4f64(A0, B1, acc01, 0, 0, 0);
            acc11 = __builtin_amdgcn_mfma_f64_16x16x4f64(A1, B1, acc11, 0, 0, 0);
        }
        // ---- geometric part and scatter ----
        const int ee = elems[base + el];
        const int* es = slot + 300L * ee;
        const double* G = sgeo[el];
#pragma unroll
        for (int tile = 0; tile < 3; tile++) {
            const double4_t acc = tile == 0 ? acc00 : (tile == 1 ? acc01 : acc11);
            const int row0 = tile == 2 ? 16 : 0, col0 = tile == 0 ? 0 : 16;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int m = row0 + 4 * r + lk, n = col0 + lr;
                if (m > n || n >= 24) continue;
                double x = acc[r];
                if (m % 3 == n % 3) x += G[(m / 3) * 8 + n / 3];
                const int s = es[upper_index(m, n)];
                if (s >= 0) val[s] += x;
            }
        }
    }
    if (bad) atomicOr(status, 7);
}

void launch_stiffness(tg_ctx* c, const double* u)
{
    TimedRegion tr(c, T_STIFFNESS);
    TG_HIP(hipMemsetAsync(c->dval.p, 0, sizeof(double) * (size_t)c->msr_len, c->stream));
    if (c->formulation == TG_Q1P0) {
        launch_q1p0(c, u, c->dR.p, true);
        return;
    }
    for (int col = 0; col < c->ncolors; col++) {
        const int first = c->color_ptr[col], count = c->color_ptr[col + 1] - first;
        if (count == 0) continue;
        const int blocks = (count + 7) / 8;
        if (c->formulation == TG_TOTAL_LAGRANGIAN)
            hipLaunchKernelGGL(k_stiffness_tl, dim3(blocks), dim3(64), 0, c->stream, c->dcolor_elems.p + first, count,
                               c->dconn.p, c->dmat.p, c->dmatprops.p, c->dX.p, u, c->dslot.p, c->dval.p,
                               c->dstatus.p);
    }
    TG_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// slot map: element upper entry (m <= n) -> MSR value index
// ---------------------------------------------------------------------------

// MSRMatrixT::Assemble symmetric branch (MSRMatrixT.cpp:110-165) sends the
// element value at local (m, n) to the global upper entry
// (min(eq_m, eq_n), max(eq_m, eq_n)); the diagonal lives in val[row], an
// off-diagonal at its position in the sorted row (AssembleRow, :532-604).
__global__ void k_build_slots(int nel, const int* __restrict__ conn, const int* __restrict__ eqn,
                              const int* __restrict__ bindx, int* __restrict__ slot)
{
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= 300L * nel) return;
    const int e = (int)(idx / 300), tt = (int)(idx % 300);
    int m = 0, rem = tt;
    while (rem >= 24 - m) {
        rem -= 24 - m;
        m++;
    }
    const int n = m + rem;
    const int rm = eqn[3L * conn[8L * e + m / 3] + m % 3];
    const int rn = eqn[3L * conn[8L * e + n / 3] + n % 3];
    int s = -1;
    if (rm >= 0 && rn >= 0) {
        const int row = rm < rn ? rm : rn, col = rm < rn ? rn : rm;
        if (row == col)
            s = row;
        else {
            int lo = bindx[row], hi = bindx[row + 1] - 1;
            while (lo <= hi) {
                const int mid = (lo + hi) >> 1;
                const int v = bindx[mid];
                if (v == col) {
                    s = mid;
                    break;
                }
                if (v < col)
                    lo = mid + 1;
                else
                    hi = mid - 1;
            }
        }
    }
    slot[idx] = s;
}

void launch_build_slots(tg_ctx* c)
{
    const long n = 300L * c->nel;
    const int bs = 256;
    hipLaunchKernelGGL(k_build_slots, dim3((unsigned)((n + bs - 1) / bs)), dim3(bs), 0, c->stream, c->nel,
                       c->dconn.p, c->deqn.p, c->dbindx.p, c->dslot.p);
    TG_HIP(hipGetLastError());
}

}  // namespace tg

namespace tg {

// One 16x16x4 FP64 MFMA on known operands; checks the operand/result lane
// layout the stiffness kernel assumes: A[i][k] from lane k*16+i,
// B[k][j] from lane k*16+j, D[4*r+l/16][l%16] in result register r.
__global__ void k_mfma_selftest(const double* A, const double* B, double* D)
{
    const int l = threadIdx.x;
    const double a = A[(l % 16) * 4 + l / 16];
    const double b = B[(l / 16) * 16 + l % 16];
    double4_t acc = {0, 0, 0, 0};
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
#pragma unroll
    for (int r = 0; r < 4; r++) D[(4 * r + l / 16) * 16 + l % 16] = acc[r];
}

}  // namespace tg

extern "C" int tg_selftest_mfma(int device, double* err)
{
    using namespace tg;
    double A[64], B[64], D[256], R[256];
    for (int i = 0; i < 64; i++) {
        A[i] = 1.0 + 0.25 * i;
        B[i] = 0.5 - 0.125 * i;
    }
    for (int i = 0; i < 16; i++)
        for (int j = 0; j < 16; j++) {
            double s = 0.0;
            for (int k = 0; k < 4; k++) s += A[i * 4 + k] * B[k * 16 + j];
            R[i * 16 + j] = s;
        }
    double *dA = nullptr, *dB = nullptr, *dD = nullptr;
    if (hipSetDevice(device) != hipSuccess) return TG_OUT_OF_RANGE;
    if (hipMalloc((void**)&dA, sizeof A) != hipSuccess) return TG_OUT_OF_MEMORY;
    if (hipMalloc((void**)&dB, sizeof B) != hipSuccess) { (void)hipFree(dA); return TG_OUT_OF_MEMORY; }
    if (hipMalloc((void**)&dD, sizeof D) != hipSuccess) { (void)hipFree(dA); (void)hipFree(dB); return TG_OUT_OF_MEMORY; }
    int rc = TG_OK;
    if (hipMemcpy(dA, A, sizeof A, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dB, B, sizeof B, hipMemcpyHostToDevice) != hipSuccess)
        rc = TG_GENERAL_FAIL;
    if (rc == TG_OK) {
        hipLaunchKernelGGL(k_mfma_selftest, dim3(1), dim3(64), 0, 0, dA, dB, dD);
        if (hipDeviceSynchronize() != hipSuccess || hipMemcpy(D, dD, sizeof D, hipMemcpyDeviceToHost) != hipSuccess)
            rc = TG_GENERAL_FAIL;
    }
    double e = 0.0;
    if (rc == TG_OK)
        for (int i = 0; i < 256; i++) e = std::max(e, std::fabs(D[i] - R[i]));
    if (err) *err = e;
    (void)hipFree(dA);
    (void)hipFree(dB);
    (void)hipFree(dD);
    if (rc == TG_OK && e > 1e-12) rc = TG_GENERAL_FAIL;
    return rc;
}

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
template <class XF, class UF>
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
    const double sc = pow(v / (H * J), 1.0 / 3.0);
#pragma unroll
    for (int k = 0; k < 9; k++) F[k] *= sc;
    double Jb, bb[6], sig[6];
    ok &= simo_kinematics(F, Jb, bb);
    simo_cauchy(mu, kappa, Jb, bb, sig);
    const double Jbar = det3(F);
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
template <bool STIFF>
__global__ void __launch_bounds__(256) k_q1p0(const int* __restrict__ elems, int count, const int* __restrict__ conn,
                                              const int* __restrict__ emat, const double2* __restrict__ props,
                                              const double* __restrict__ X, const double* __restrict__ U,
                                              const int* __restrict__ eqn, double* __restrict__ R,
                                              const int* __restrict__ slot, double* __restrict__ val,
                                              int* __restrict__ status)
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
    ok &= q1p0_ip_force(q, d, det0, g0, xa, ua, mk.x, mk.y, f);
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
            ok &= q1p0_ip_force(q, d, det0, g0, xa, ua, mk.x, mk.y, fp);
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
                        const int s = es[upper_index(m, col)];
                        if (s >= 0) val[s] += (vp[i] - v0[i]) / eps;
                    }
                }
            }
            __syncthreads();
        }
    }
    if (valid && !ok) atomicOr(status, 7);
}

void launch_q1p0(tg_ctx* c, const double* u, double* R_node, bool stiff)
{
    for (int col = 0; col < c->ncolors; col++) {
        const int first = c->color_ptr[col], count = c->color_ptr[col + 1] - first;
        if (count == 0) continue;
        const int blocks = (count + 31) / 32;
        if (stiff)
            hipLaunchKernelGGL(k_q1p0<true>, dim3(blocks), dim3(256), 0, c->stream, c->dcolor_elems.p + first, count,
                               c->dconn.p, c->dmat.p, c->dmatprops.p, c->dX.p, u, c->deqn.p, R_node, c->dslot.p,
                               c->dval.p, c->dstatus.p);
        else
            hipLaunchKernelGGL(k_q1p0<false>, dim3(blocks), dim3(256), 0, c->stream, c->dcolor_elems.p + first, count,
                               c->dconn.p, c->dmat.p, c->dmatprops.p, c->dX.p, u, c->deqn.p, R_node, c->dslot.p,
                               c->dval.p, c->dstatus.p);
    }
    TG_HIP(hipGetLastError());
}

}  // namespace tg
