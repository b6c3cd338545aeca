// Element kernels: total-Lagrangian Q1 hexahedron residual and stiffness with
// the SimoIso3D material, scattered colour by colour (no atomics) into the
// node-space residual and the upper-triangle MSR values.
//
// Lane mapping: 8 consecutive lanes own one element, lane q evaluates
// integration point q.  Nodal coordinates/displacements are gathered once per
// element (lane q loads node q) into LDS and read by all 8 lanes.
#include "tg_internal.h"
#include "tg_kinematics.h"
#include "tg_element.h"

namespace tg {

typedef double double4_t __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// residual
// ---------------------------------------------------------------------------


// R -= f_int for the elements of one colour.  Per integration point
// (TotalLagrangianT::FormKd, TotalLagrangianT.cpp:107-143, called with
// constK = -1 from SolidElementT::ElementRHSDriver, SolidElementT.cpp:1240):
//   R_a,i += -J w det0 sum_k (sigma F^-T)_ik dN_a/dX_k
template <int MAT>
__global__ void __launch_bounds__(256) k_residual_tl(const int* __restrict__ elems, int count,
                                                     const int* __restrict__ conn, const int* __restrict__ emat,
                                                     const double2* __restrict__ props,
                                                     const double* __restrict__ X, const double* __restrict__ U,
                                                     const int* __restrict__ eqn, double* __restrict__ R,
                                                     int* __restrict__ status, const int* __restrict__ elem_seg,
                                                     int* __restrict__ segstat)
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
        bad |= !material_stress<MAT>(mk.x, mk.y, F, J, Jb, bb, sig);
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
    if (bad) {
        // the context's flag, and the flag of the element's problem in a batch
        atomicOr(status, 7);
        atomicOr(&segstat[elem_seg[e]], 7);
    }
}

void launch_q1p0(tg_ctx* c, const double* u, double* R_node, bool stiff);

// R += applied nodal forces at active dofs (FieldT::FormRHS, FieldT.cpp:390-401)
__global__ void __launch_bounds__(256) k_add_forces(long n3, const int* __restrict__ eqn,
                                                    const double* __restrict__ f, double* __restrict__ R)
{
    const long k = (long)blockIdx.x * 256 + threadIdx.x;
    if (k < n3 && eqn[k] >= 0) R[k] += f[k];
}

static void element_residual(tg_ctx* c, const double* u, double* R_node);

void launch_residual(tg_ctx* c, const double* u, double* R_node)
{
    TimedRegion tr(c, T_RESIDUAL);
    TG_HIP(hipMemsetAsync(R_node, 0, sizeof(double) * 3 * (size_t)c->nnd, c->stream));
    element_residual(c, u, R_node);
    if (c->dfext.p) {
        const long n3 = 3L * c->nnd;
        hipLaunchKernelGGL(k_add_forces, dim3((unsigned)((n3 + 255) / 256)), dim3(256), 0, c->stream, n3, c->deqn.p,
                           c->dfext.p, R_node);
        TG_HIP(hipGetLastError());
    }
}

static void element_residual(tg_ctx* c, const double* u, double* R_node)
{
    if (c->formulation == TG_Q1P0) {
        launch_q1p0(c, u, R_node, false);
        return;
    }
    for (int col = 0; col < c->ncolors; col++) {
        const int first = c->color_ptr[col], count = c->color_ptr[col + 1] - first;
        if (count == 0) continue;
        const int blocks = (count + 31) / 32;
        if (c->material_model == TG_MAT_STVENANT)
            hipLaunchKernelGGL(k_residual_tl<MAT_STVENANT>, dim3(blocks), dim3(256), 0, c->stream,
                               c->dcolor_elems.p + first, count, c->dconn.p, c->dmat.p, c->dmatprops.p, c->dX.p, u,
                               c->deqn.p, R_node, c->dstatus.p, c->seg.d_elem_seg.p, c->seg.d_segstat.p);
        else
            hipLaunchKernelGGL(k_residual_tl<MAT_SIMO>, dim3(blocks), dim3(256), 0, c->stream,
                               c->dcolor_elems.p + first, count, c->dconn.p, c->dmat.p, c->dmatprops.p, c->dX.p, u,
                               c->deqn.p, R_node, c->dstatus.p, c->seg.d_elem_seg.p, c->seg.d_segstat.p);
    }
    TG_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// stiffness
// ---------------------------------------------------------------------------


// Element stiffness (TotalLagrangianT::FormStiffness, TotalLagrangianT.cpp:40-104),
// one lane per upper 3x3 node block (a <= b) of an element.  With
// D_q = s_q c_q and T_q = s_q sigma_q (s_q = det0 w J, Voigt 11,22,33,23,13,12):
//   K(ai,bj) = sum_q sum_kl g_q[k][a] g_q[l][b] ( D_q[V(i,k)][V(j,l)] + delta_ij T_q[V(k,l)] )
// i.e. B^T D B plus Expand(gradN^T T gradN, 3), accumulated over integration
// points in order.  A block of 36*EPB lanes owns EPB elements: 8 lanes per
// element first evaluate the integration points into LDS, then 36 lanes per
// element form one 3x3 block each (no divergence, 18 accumulators) and
// scatter its upper entries through the slot map.  FP64 MFMA peaks at the
// FP64 VALU rate on gfx950 while the dense 24x48x24 tile product does about
// twice the arithmetic of this sparse form, so the FP64 path uses the VALU.
//
// The block shape trades lane utilisation against occupancy: at ~120 VGPRs a
// SIMD holds 4 waves, so a 16-element block (9 waves) runs alone on its CU
// while 8-element blocks (5 waves) run three at a time and hide each other's
// barrier and scatter waits.  EPB (elements per block) and the phase-2 lane
// order are template parameters; the autotuner picks among the compiled
// shapes.
constexpr int STIFF_EPB[] = {16, 12, 8, 4, 16, 8, 16, 8, 4, 12, 8, 4, 12, 8, 4, 16, 8, 4, 8};
constexpr bool STIFF_PM[] = {false, false, false, false, true, true, false, false, false, false,
                             false, false, false, false, false, true, false, false, false};
constexpr bool STIFF_DB[] = {false, false, false, false, false, false, true, true, true, false,
                             false, false, false, false, false, false, false, false, false};
// scatter form; 3: the matrix-core kernel k_stiffness_mfma (EPB elements per
// wave), 4: the same with registers bounded for 3 waves per SIMD
constexpr int STIFF_SC[] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 2, 2, 2, 2, 3, 3, 4};
constexpr int N_STIFF_SHAPES = sizeof(STIFF_EPB) / sizeof(STIFF_EPB[0]);

// Voigt index of (i,k): (11,22,33,23,13,12)
__device__ __forceinline__ constexpr int voigt(int i, int k) { return i == k ? i : 6 - i - k; }

// packed upper-triangle index of a symmetric 6x6 (row-major, v <= w)
__device__ __forceinline__ constexpr int sym6(int v, int w)
{
    return (v < w ? v : w) * 6 - ((v < w ? v : w) * ((v < w ? v : w) - 1)) / 2 + ((v < w ? w : v) - (v < w ? v : w));
}

// VT is the arithmetic of the 3x3 block products and of the assembled
// values: double (FP64 path) or float (mixed path, integration-point
// kinematics and material stay FP64 and are rounded once to FP32).
// Integration point q of one element (TotalLagrangianT::FormStiffness's
// per-IP data): g = dN/dx (ShapeFunctionT::TransformDerivatives), s c and
// s sigma with s = det0 * J * w (w = 1), all in FP64; xu(a, c) reads node a's
// X (c < 3) or u (c >= 3).  Returns false on J <= 0 or det0 <= 0 (the data
// are then zero).  An invalid slot (no element) gives zeros.
template <int MAT, class XU>
__device__ __forceinline__ bool stiffness_ip(int q, bool valid, double2 mk, XU xu, double (&g)[3][8],
                                             double (&D)[21], double (&T)[6])
{
    bool bad = false;
    double scale = 0.0, Jb = 1.0, bb[6] = {1, 1, 1, 0, 0, 0}, sig[6] = {0, 0, 0, 0, 0, 0};
    if (valid) {
        double d[3][8], j0[9], ji[9], g0[3][8], F[9], Fi[9];
        parent_derivs(q, d);
        jacobian([&](int a, int c) { return xu(a, c); }, d, j0);
        const double det0 = det3(j0);
        bad |= !(det0 > 0.0);
        inv3(j0, det0, ji);
        spatial_derivs(ji, d, g0);
        deformation_gradient([&](int a, int c) { return xu(a, 3 + c); }, g0, F);
        const double J = det3(F);
        bad |= !(J > 0.0);
        bad |= !material_stress<MAT>(mk.x, mk.y, F, J, Jb, bb, sig);
        inv3(F, J, Fi);
        scale = bad ? 0.0 : det0 * J;  // weight 1
        // dN/dx_j = sum_k Fi(k,j) dN/dX_k
#pragma unroll
        for (int a = 0; a < 8; a++)
#pragma unroll
            for (int j = 0; j < 3; j++)
                g[j][a] = Fi[0 + 3 * j] * g0[0][a] + Fi[1 + 3 * j] * g0[1][a] + Fi[2 + 3 * j] * g0[2][a];
    } else {
#pragma unroll
        for (int a = 0; a < 8; a++)
#pragma unroll
            for (int j = 0; j < 3; j++) g[j][a] = 0.0;
    }
    double cm[6][6];
    material_moduli<MAT>(mk.x, mk.y, Jb, bb, cm);
#pragma unroll
    for (int v = 0; v < 6; v++)
#pragma unroll
        for (int w = v; w < 6; w++) D[sym6(v, w)] = scale * cm[v][w];
#pragma unroll
    for (int k = 0; k < 6; k++) T[k] = scale * sig[k];
    return !bad;
}


// Scatter of one upper 3x3 node block (a <= b, local upper index p) of
// element e: into the MSR values through the slot map (TO_OP = false,
// MSRMatrixT::Assemble) or the block and its transpose into the solver
// operator through the operator slot map (TO_OP = true, prescribed rows and
// columns left zero).
// ATOM: the additions are issued as memory-side atomic adds with no return
// value, so the lane does not wait for the old value; within one launch no
// two lanes touch one entry (colouring), so the summation order per entry is
// still the launch order.
template <class VT>
__device__ __forceinline__ void add_to(VT* p, VT v, bool atom)
{
    if (atom)
        unsafeAtomicAdd(p, v);
    else
        *p += v;
}

template <class VT, bool TO_OP, bool ATOM = false>
__device__ __forceinline__ void scatter_block(int e, int a, int b, int p, const VT (&K)[3][3],
                                              const int* __restrict__ conn, const int* __restrict__ slot,
                                              const int* __restrict__ eqn, VT* __restrict__ val)
{
    if (TO_OP) {
        const int b1 = slot[72L * e + 2 * p], b2 = slot[72L * e + 2 * p + 1];
        if (b1 < 0) return;
        const int A = conn[8L * e + a], B = conn[8L * e + b];
        bool ra[3], cb[3];
#pragma unroll
        for (int i = 0; i < 3; i++) {
            ra[i] = eqn[3L * A + i] >= 0;
            cb[i] = eqn[3L * B + i] >= 0;
        }
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) {
                if (!(ra[i] && cb[j])) continue;
                if (a == b) {
                    // both triangles of a diagonal block from the upper value
                    add_to(&val[b1 + (3 * i + j) * 64], K[i < j ? i : j][i < j ? j : i], ATOM);
                } else {
                    add_to(&val[b1 + (3 * i + j) * 64], K[i][j], ATOM);
                    if (b2 >= 0) add_to(&val[b2 + (3 * j + i) * 64], K[i][j], ATOM);
                }
            }
    } else {
        const int* es = slot + 300L * e;
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) {
                const int m = 3 * a + i, n = 3 * b + j;
                if (n < m) continue;  // lower half of a diagonal block
                const int s = slot_target(es[upper_index(m, n)]);
                if (s >= 0) add_to(&val[s], K[i][j], ATOM);
            }
    }
}

// upper block index p (row-major over the upper triangle of 8x8) -> (a, b)
__device__ __forceinline__ void upper_block(int p, int& a, int& b)
{
    a = 0;
    int rem = p;
    while (rem >= 8 - a) {
        rem -= 8 - a;
        a++;
    }
    b = a + rem;
}

// TO_OP = false: upper entries into the MSR values through the slot map
// (MSRMatrixT::Assemble).  TO_OP = true: the block and its transpose
// straight into the solver's block operator through the operator slot map,
// prescribed rows and columns left zero -- the same sums in the same order
// as expanding the assembled MSR, without the MSR round trip.
// PM selects the lane order of phase 2: false = element-major (lane
// le*36 + p), true = block-major (lane p*EPB + le), where one wave holds the
// same block p of many elements -- their scatter targets are neighbouring
// rows -- and the LDS tables are element-minor so those lanes read
// consecutive words.
// SC selects the scatter: 0 load-add-store after the block products; 1 (MSR
// path) the slot indices and old values of the lane's 9 entries are loaded
// before the block products, so the read half of the read-modify-write is in
// flight during the arithmetic (no other element of the launch touches those
// entries); 2 memory-side atomic adds with no return value; 5 (MSR path) the
// element buffer of the gather assembly instead of the MSR.
template <class VT, int EPB, bool TO_OP, bool PM, int SC, int MAT = MAT_SIMO>
__global__ void __launch_bounds__(36 * EPB) k_stiffness_tl(const int* __restrict__ elems, int count,
                                                             const int* __restrict__ conn,
                                                             const int* __restrict__ emat,
                                                             const double2* __restrict__ props,
                                                             const double* __restrict__ X,
                                                             const double* __restrict__ U,
                                                             const int* __restrict__ slot,
                                                             const int* __restrict__ eqn, VT* __restrict__ val,
                                                             int* __restrict__ status)
{
    __shared__ double sxu[EPB][8][6];
    __shared__ VT sg[EPB * 8 * 24];  // dN/dx per (element, ip, dir, node)
    __shared__ VT sd[EPB * 8 * 21];  // s c, packed symmetric Voigt
    __shared__ VT st[EPB * 8 * 6];   // s sigma, Voigt
    auto gi = [](int le, int q, int j, int a) { return PM ? ((q * 3 + j) * 8 + a) * EPB + le : ((le * 8 + q) * 3 + j) * 8 + a; };
    auto di = [](int le, int q, int k) { return PM ? (q * 21 + k) * EPB + le : (le * 8 + q) * 21 + k; };
    auto ti = [](int le, int q, int k) { return PM ? (q * 6 + k) * EPB + le : (le * 8 + q) * 6 + k; };
    const int t = threadIdx.x;
    const long base = (long)blockIdx.x * EPB;
    bool bad = false;
    if (t < 8 * EPB) {
        const int le = t >> 3, q = t & 7;
        if (base + le < count) {
            const int node = conn[8L * elems[base + le] + q];
#pragma unroll
            for (int c = 0; c < 3; c++) {
                sxu[le][q][c] = X[3L * node + c];
                sxu[le][q][3 + c] = U[3L * node + c];
            }
        }
    }
    __syncthreads();
    // ---- phase 1: integration-point data ----
    if (t < 8 * EPB) {
        const int le = t >> 3, q = t & 7;
        const bool valid = base + le < count;
        const double2 mk = valid ? props[emat[elems[base + le]]] : make_double2(0.0, 0.0);
        double g[3][8], D[21], T[6];
        bad |= !stiffness_ip<MAT>(q, valid, mk, [&](int a, int c) { return sxu[le][a][c]; }, g, D, T);
#pragma unroll
        for (int a = 0; a < 8; a++)
#pragma unroll
            for (int j = 0; j < 3; j++) sg[gi(le, q, j, a)] = (VT)g[j][a];
#pragma unroll
        for (int k = 0; k < 21; k++) sd[di(le, q, k)] = (VT)D[k];
#pragma unroll
        for (int k = 0; k < 6; k++) st[ti(le, q, k)] = (VT)T[k];
    }
    __syncthreads();
    // ---- phase 2: one 3x3 block per lane ----
    const int le = PM ? t % EPB : t / 36, p = PM ? t / EPB : t % 36;
    if (base + le < count) {
        int a, b;
        upper_block(p, a, b);
        constexpr bool PRE = SC == 1 && !TO_OP;
        int sidx[PRE ? 9 : 1];
        VT old[PRE ? 9 : 1];
        if (PRE) {
            const int* es = slot + 300L * elems[base + le];
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int j = 0; j < 3; j++) {
                    const int m = 3 * a + i, n = 3 * b + j;
                    sidx[PRE ? 3 * i + j : 0] = n < m ? -1 : es[upper_index(m, n)];
                }
            // a first contribution skips the read (the entry is still 0)
#pragma unroll
            for (int k = 0; k < (PRE ? 9 : 1); k++) old[k] = sidx[k] >= 0 ? val[sidx[k]] : VT(0);
#pragma unroll
            for (int k = 0; k < (PRE ? 9 : 1); k++) sidx[k] = slot_target(sidx[k]);
        }
        VT K[3][3];
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) K[i][j] = 0;
#pragma unroll 1
        for (int q = 0; q < 8; q++) {
            VT D[21], T[6], ga[3], gb[3];
#pragma unroll
            for (int k = 0; k < 21; k++) D[k] = sd[di(le, q, k)];
#pragma unroll
            for (int k = 0; k < 6; k++) T[k] = st[ti(le, q, k)];
#pragma unroll
            for (int k = 0; k < 3; k++) {
                ga[k] = sg[gi(le, q, k, a)];
                gb[k] = sg[gi(le, q, k, b)];
            }
            VT h[3][3], geo = 0;
#pragma unroll
            for (int k = 0; k < 3; k++)
#pragma unroll
                for (int l = 0; l < 3; l++) {
                    h[k][l] = ga[k] * gb[l];
                    geo += h[k][l] * T[voigt(k, l)];
                }
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int j = 0; j < 3; j++) {
                    VT x = 0;
#pragma unroll
                    for (int k = 0; k < 3; k++)
#pragma unroll
                        for (int l = 0; l < 3; l++) x += h[k][l] * D[sym6(voigt(i, k), voigt(j, l))];
                    K[i][j] += x;
                }
#pragma unroll
            for (int i = 0; i < 3; i++) K[i][i] += geo;
        }
        if (PRE) {
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int j = 0; j < 3; j++) {
                    const int k = PRE ? 3 * i + j : 0;
                    if (sidx[k] >= 0) val[sidx[k]] = old[k] + K[i][j];
                }
        } else if (SC == 5) {
            // element buffer: val = the row of the launch's first element
            VT* eb = val + 300L * (base + le);
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int j = 0; j < 3; j++) {
                    const int m = 3 * a + i, n = 3 * b + j;
                    if (n >= m) eb[upper_index(m, n)] = K[i][j];
                }
        } else {
            scatter_block<VT, TO_OP, SC == 2>(elems[base + le], a, b, p, K, conn, slot, eqn, val);
        }
    }
    if (bad) atomicOr(status, 7);
}

// First contributions in launch order: first[s] = the smallest launch-order
// position of an element with a contribution to MSR value s (within one
// launch no two elements share an entry, so that element's entry is the
// first to reach s); then those slots are marked -s-2.
__global__ void k_first_touch(int nel, const int* __restrict__ order, const int* __restrict__ slot,
                              int* __restrict__ first)
{
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= 300L * nel) return;
    const int pos = (int)(idx / 300);
    const int s = slot_target(slot[300L * order[pos] + idx % 300]);
    if (s >= 0) atomicMin(&first[s], pos);
}

__global__ void k_mark_first(int nel, const int* __restrict__ order, int* __restrict__ slot,
                             const int* __restrict__ first)
{
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= 300L * nel) return;
    const int pos = (int)(idx / 300);
    int* p = slot + 300L * order[pos] + idx % 300;
    const int s = slot_target(*p);
    if (s >= 0) *p = first[s] == pos ? -s - 2 : s;
}

// Launch list of the colour passes.  Within one colour no two elements share
// a node, so any split of a colour into launches is race-free.  Each
// segment's elements are split by their rank inside the segment into ranges
// of `chunk`, and all colours of range j run before range j+1: the matrix
// rows a range touches then stay resident in the memory-side cache across
// the colour passes instead of the whole matrix streaming once per colour.
// An entry's contributions are summed in (range, colour) order, which
// depends only on its own segment and the chunk, so a problem's result is
// the same alone and inside any batch, run to run.
static const std::vector<std::pair<int, int>>& stiffness_launches(tg_ctx* c)
{
    long chunk = c->stiff_chunk;
    if (chunk < 0) {
        const double vb = (c->precision == TG_MIXED ? 4.0 : 8.0) * (double)c->msr_len / std::max(1, c->nel);
        chunk = (long)(256.0 * (1 << 20) / std::max(vb, 1.0));
        chunk = std::max(chunk, 131072L);
    }
    if (chunk == 0 || chunk >= c->nel) chunk = c->nel;
    if (c->stiff_chunk_built == chunk) return c->stiff_launches;
    // range index of every element: rank inside its segment / chunk
    const int nseg = c->seg.nseg;
    std::vector<long> rank((size_t)nseg, 0);
    std::vector<int> key((size_t)c->nel);
    int nkeys = 1;
    for (int e = 0; e < c->nel; e++) {
        const int g = c->seg.elem_seg.empty() ? 0 : c->seg.elem_seg[e];
        key[e] = (int)(rank[g]++ / chunk);
        nkeys = std::max(nkeys, key[e] + 1);
    }
    // order: (range, colour, element); the colour lists are ascending
    std::vector<int> cnt((size_t)nkeys * c->ncolors + 1, 0);
    for (int col = 0; col < c->ncolors; col++)
        for (int i = c->color_ptr[col]; i < c->color_ptr[col + 1]; i++)
            cnt[(size_t)key[c->color_elems[i]] * c->ncolors + col + 1]++;
    for (size_t k = 1; k < cnt.size(); k++) cnt[k] += cnt[k - 1];
    std::vector<int> order((size_t)c->nel);
    std::vector<int> fill(cnt.begin(), cnt.end() - 1);
    for (int col = 0; col < c->ncolors; col++)
        for (int i = c->color_ptr[col]; i < c->color_ptr[col + 1]; i++) {
            const int e = c->color_elems[i];
            order[(size_t)fill[(size_t)key[e] * c->ncolors + col]++] = e;
        }
    c->stiff_launches.clear();
    for (size_t k = 0; k + 1 < cnt.size(); k++)
        if (cnt[k + 1] > cnt[k]) c->stiff_launches.emplace_back(cnt[k], cnt[k + 1] - cnt[k]);
    if (c->dstiff_elems.n != (size_t)c->nel) c->dstiff_elems.alloc((size_t)c->nel);
    TG_HIP(hipMemcpyAsync(c->dstiff_elems.p, order.data(), sizeof(int) * c->nel, hipMemcpyHostToDevice, c->stream));
    {
        // mark the first contributions of this order in the slot map
        tg::DBuf<int> first;
        first.alloc((size_t)c->msr_len);
        TG_HIP(hipMemsetAsync(first.p, 0x7f, sizeof(int) * (size_t)c->msr_len, c->stream));  // > any position
        const long n = 300L * c->nel;
        const unsigned blocks = (unsigned)((n + 255) / 256);
        hipLaunchKernelGGL(k_first_touch, dim3(blocks), dim3(256), 0, c->stream, c->nel, c->dstiff_elems.p,
                           c->dslot.p, first.p);
        hipLaunchKernelGGL(k_mark_first, dim3(blocks), dim3(256), 0, c->stream, c->nel, c->dstiff_elems.p,
                           c->dslot.p, first.p);
        TG_HIP(hipGetLastError());
        TG_HIP(hipStreamSynchronize(c->stream));  // order (host) and first outlive their use
    }
    c->stiff_chunk_built = chunk;
    return c->stiff_launches;
}

// Element stiffness, "DB form": the same K_ab, accumulated as
//   K_ab += B_a^T (D_q B_b) + (g_a . T_q g_b) 1
// with the 6x24 product D_q B (and the 8 vectors T_q g_b) formed once per
// element and integration point -- 36 lanes produce its 168 entries into
// LDS -- instead of every lane recontracting D_q with its own g_b.  Per
// element and point that is 168 x 3 + 36 x 33 FMA instead of 36 x 99, with
// fewer live registers.  Same mapping otherwise: EPB elements per
// workgroup, phase 1 on 8 lanes per element, lane (le, p) owns upper block
// p.  Voigt rows of B for node b (11,22,33,23,13,12): u1 -> (g1,0,0,0,g3,g2),
// u2 -> (0,g2,0,g3,0,g1), u3 -> (0,0,g3,g2,g1,0) (Hughes 2.8.21).
template <class VT, int EPB, bool TO_OP, int MAT = MAT_SIMO>
__global__ void __launch_bounds__(36 * EPB) k_stiffness_db(const int* __restrict__ elems, int count,
                                                             const int* __restrict__ conn,
                                                             const int* __restrict__ emat,
                                                             const double2* __restrict__ props,
                                                             const double* __restrict__ X,
                                                             const double* __restrict__ U,
                                                             const int* __restrict__ slot,
                                                             const int* __restrict__ eqn, VT* __restrict__ val,
                                                             int* __restrict__ status)
{
    constexpr int NDB = 6 * 24 + 8 * 3;  // D B (row-major 6 x 24), then T g_b (8 x 3)
    __shared__ double sxu[EPB][8][6];
    __shared__ VT sg[EPB][8][24];        // dN/dx [element][ip][dir * 8 + node]
    __shared__ VT sd[EPB][8][21];        // s c, packed symmetric Voigt
    __shared__ VT st[EPB][8][6];         // s sigma, Voigt
    __shared__ VT sdb[EPB][NDB];
    const int t = threadIdx.x;
    const long base = (long)blockIdx.x * EPB;
    bool bad = false;
    if (t < 8 * EPB) {
        const int le = t >> 3, q = t & 7;
        if (base + le < count) {
            const int node = conn[8L * elems[base + le] + q];
#pragma unroll
            for (int c = 0; c < 3; c++) {
                sxu[le][q][c] = X[3L * node + c];
                sxu[le][q][3 + c] = U[3L * node + c];
            }
        }
    }
    __syncthreads();
    if (t < 8 * EPB) {
        const int le = t >> 3, q = t & 7;
        const bool valid = base + le < count;
        const double2 mk = valid ? props[emat[elems[base + le]]] : make_double2(0.0, 0.0);
        double g[3][8], D[21], T[6];
        bad |= !stiffness_ip<MAT>(q, valid, mk, [&](int a, int c) { return sxu[le][a][c]; }, g, D, T);
#pragma unroll
        for (int j = 0; j < 3; j++)
#pragma unroll
            for (int a = 0; a < 8; a++) sg[le][q][j * 8 + a] = (VT)g[j][a];
#pragma unroll
        for (int k = 0; k < 21; k++) sd[le][q][k] = (VT)D[k];
#pragma unroll
        for (int k = 0; k < 6; k++) st[le][q][k] = (VT)T[k];
    }
    __syncthreads();
    const int le = t / 36, p = t % 36;
    int a, b;
    upper_block(p, a, b);
    VT K[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) K[i][j] = 0;
#pragma unroll 1
    for (int q = 0; q < 8; q++) {
        // D_q B and T_q g_b of element le: entries p, p+36, ...
        for (int o = p; o < NDB; o += 36) {
            VT v;
            if (o < 144) {
                const int vv = o / 24, n = o % 24, bn = n / 3, j = n % 3;
                v = sd[le][q][sym6(vv, voigt(j, 0))] * sg[le][q][bn] +
                    sd[le][q][sym6(vv, voigt(j, 1))] * sg[le][q][8 + bn] +
                    sd[le][q][sym6(vv, voigt(j, 2))] * sg[le][q][16 + bn];
            } else {
                const int m = o - 144, bn = m / 3, k = m % 3;
                v = st[le][q][voigt(k, 0)] * sg[le][q][bn] + st[le][q][voigt(k, 1)] * sg[le][q][8 + bn] +
                    st[le][q][voigt(k, 2)] * sg[le][q][16 + bn];
            }
            sdb[le][o] = v;
        }
        __syncthreads();
        VT ga[3];
#pragma unroll
        for (int k = 0; k < 3; k++) ga[k] = sg[le][q][k * 8 + a];
        VT geo = 0;
#pragma unroll
        for (int k = 0; k < 3; k++) geo += ga[k] * sdb[le][144 + 3 * b + k];
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) {
                VT x = 0;
#pragma unroll
                for (int k = 0; k < 3; k++) x += ga[k] * sdb[le][voigt(i, k) * 24 + 3 * b + j];
                K[i][j] += x;
            }
#pragma unroll
        for (int i = 0; i < 3; i++) K[i][i] += geo;
        __syncthreads();
    }
    if (base + le < count) scatter_block<VT, TO_OP>(elems[base + le], a, b, p, K, conn, slot, eqn, val);
    if (bad) atomicOr(status, 7);
}

// Element stiffness on the matrix cores: K_e = sum_q B_q^T (D_q B_q) + the
// geometric term, as 16x16x4 MFMA tiles (v_mfma_f32_16x16x4f32 in the mixed
// path, v_mfma_f64_16x16x4f64 in FP64).  One wave per workgroup, EPW
// elements per wave:
//   - phase 1 (lane = element x integration point): g = dN/dx, s c, s sigma
//     in FP64, rounded to VT into LDS, as in k_stiffness_tl;
//   - per element, the 24 element dofs are taken dof-major (p = 8 i + a), so
//     tile rows/columns 0..15 are the x and y dofs of the 8 nodes and 16..23
//     the z dofs; the three upper tiles (0,0), (0,1), (1,1) cover every
//     unordered dof pair once (136 + 128 + 36 = 300 entries).  The K
//     dimension is (Voigt row v, integration point q), 48 = 6 x 8: lane l
//     supplies A[p = l%16][k] = B_q[v][p] (a gradient entry or 0, Hughes
//     2.8.21) and B[k][p] = (D_q B_q)[v][p] (3 products, the three nonzeros
//     of a B column) for q = 4h + l/16; a lane's two dofs (l%16 and 16 +
//     l%16) belong to the same node, so it reads 3 gradient entries and D_q.
//   - the geometric term G_ab = sum_q g_q(a) . (T_q g_q(b)) is one more 16x16
//     tile over K = 24 (q, k); it seeds tile (1,1) lane for lane and tile
//     (0,0) through a lane permutation, on the blocks with equal dof.
// Operands: lane l supplies A[l%16][l/16] and B[l/16][l%16] in both
// precisions.  Results: lane l, register r holds row 4(l/16) + r (f32) or
// row l/16 + 4r (f64), column l%16 (measured on gfx950); the entries go
// through the slot map into the MSR (slot and old-value loads issued before
// the products; first contributions stored without the read).
template <class VT>
struct Mfma;
template <>
struct Mfma<float> {
    typedef float v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 mac(float a, float b, v4 c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    // tile row of result register r in lane group kq = l/16
    static __device__ __forceinline__ constexpr int row(int kq, int r) { return 4 * kq + r; }
};
template <>
struct Mfma<double> {
    typedef double v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 mac(double a, double b, v4 c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ constexpr int row(int kq, int r) { return kq + 4 * r; }
};

// local node-major index (3 a + i) of dof-major index p = 8 i + a
__device__ __forceinline__ int node_major(int p) { return 3 * (p & 7) + (p >> 3); }

// MINW > 1 asks the compiler for registers that fit MINW waves per SIMD.
// GATHER: the entries go to the element buffer, val = its row of the
// launch's first element (no slot map, no loads); else into the MSR.
template <class VT, int EPW, int MAT = MAT_SIMO, int MINW = 1, bool GATHER = false>
__global__ void __launch_bounds__(64, MINW) k_stiffness_mfma(const int* __restrict__ elems, int count,
                                                        const int* __restrict__ conn, const int* __restrict__ emat,
                                                        const double2* __restrict__ props,
                                                        const double* __restrict__ X, const double* __restrict__ U,
                                                        const int* __restrict__ slot, VT* __restrict__ val,
                                                        int* __restrict__ status)
{
    using v4 = typename Mfma<VT>::v4;
    __shared__ double sxu[EPW][8][6];
    __shared__ VT sg[EPW][8][3][8];  // dN/dx [element][ip][direction][node]
    __shared__ VT sd[EPW][8][21];    // s c, packed symmetric Voigt
    __shared__ VT st[EPW][8][6];     // s sigma, Voigt (11,22,33,23,13,12)
    const int l = threadIdx.x;
    const long base = (long)blockIdx.x * EPW;
    const int nel = (int)(count - base < EPW ? count - base : EPW);
    bool bad = false;
    if (l < 8 * EPW) {
        const int le = l >> 3, q = l & 7;
        if (le < nel) {
            const int node = conn[8L * elems[base + le] + q];
#pragma unroll
            for (int k = 0; k < 3; k++) {
                sxu[le][q][k] = X[3L * node + k];
                sxu[le][q][3 + k] = U[3L * node + k];
            }
        }
    }
    __syncthreads();
    if (l < 8 * EPW) {
        const int le = l >> 3, q = l & 7;
        if (le < nel) {
            const double2 mk = props[emat[elems[base + le]]];
            double g[3][8], D[21], T[6];
            bad |= !stiffness_ip<MAT>(q, true, mk, [&](int a, int k) { return sxu[le][a][k]; }, g, D, T);
#pragma unroll
            for (int k = 0; k < 3; k++)
#pragma unroll
                for (int a = 0; a < 8; a++) sg[le][q][k][a] = (VT)g[k][a];
#pragma unroll
            for (int k = 0; k < 21; k++) sd[le][q][k] = (VT)D[k];
#pragma unroll
            for (int k = 0; k < 6; k++) st[le][q][k] = (VT)T[k];
        }
    }
    __syncthreads();
    const int kq = l >> 4, c = l & 15;
    const int na = c & 7;        // node of both of this lane's dofs
    const int i0 = c >> 3;       // dof of p = c (x or y); p = 16 + c is the z dof (valid when c < 8)
    const bool lo8 = c < 8;
    // this lane's output entries: tile t (0: (0,0), 1: (0,1), 2: (1,1)), register r
    int ent[12];                 // packed upper index into the element's 300, or -1
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const int m = Mfma<VT>::row(kq, r);  // row of tiles (0,0), (0,1); 16 + m of tile (1,1)
        ent[r] = m <= c ? upper_index(min(node_major(m), node_major(c)), max(node_major(m), node_major(c))) : -1;
        const int a01 = node_major(m), b01 = 3 * na + 2;
        ent[4 + r] = lo8 ? upper_index(min(a01, b01), max(a01, b01)) : -1;
        // tile (1,1): z dof of node m against the z dof of node na
        ent[8 + r] = (lo8 && m < 8 && m <= c) ? upper_index(3 * m + 2, 3 * na + 2) : -1;
    }
    // the scatter's loads run one element ahead: the slot indices of element
    // le + 1 are issued before the products of element le and its old values
    // after them (elements of one launch share no entry)
    // Loads are issued unconditionally (entry 0 of the element's slot row,
    // value 0 of the MSR, both always present, stand in for a lane with
    // nothing there) and their results selected, so the loop has no branches
    // around loads.
    int nsidx[12];
    VT nold[12];
    if (!GATHER) {
        const int* es = slot + 300L * elems[base];
#pragma unroll
        for (int k = 0; k < 12; k++) {
            const int x = es[ent[k] >= 0 ? ent[k] : 0];
            nsidx[k] = ent[k] >= 0 ? x : -1;
        }
        // a first contribution skips the read (the entry is still 0)
#pragma unroll
        for (int k = 0; k < 12; k++) {
            const VT x = val[nsidx[k] >= 0 ? nsidx[k] : 0];
            nold[k] = nsidx[k] >= 0 ? x : VT(0);
        }
    }
    for (int le = 0; le < nel; le++) {
        int sidx[12];
        VT old[12];
#pragma unroll
        for (int k = 0; k < 12; k++) {
            sidx[k] = nsidx[k];
            old[k] = nold[k];
        }
        const bool more = !GATHER && le + 1 < nel;
        if (more) {
            const int* es = slot + 300L * elems[base + le + 1];
#pragma unroll
            for (int k = 0; k < 12; k++) {
                const int x = es[ent[k] >= 0 ? ent[k] : 0];
                nsidx[k] = ent[k] >= 0 ? x : -1;
            }
        }
        // geometric 8x8 tile: A[a][(q,k)] = g_qk(a), B[(q,k)][b] = (T_q g_q(b))_k
        v4 accg = {0, 0, 0, 0};
#pragma unroll
        for (int s = 0; s < 6; s++) {
            const int k = 4 * s + kq, q = k / 3, kk = k - 3 * q;
            VT ga = 0, tb = 0;
            if (lo8) {
                const VT* T = st[le][q];
                const VT tk0 = T[voigt(kk, 0)], tk1 = T[voigt(kk, 1)], tk2 = T[voigt(kk, 2)];
                ga = sg[le][q][kk][na];
                tb = tk0 * sg[le][q][0][na] + tk1 * sg[le][q][1][na] + tk2 * sg[le][q][2][na];
            }
            accg = Mfma<VT>::mac(ga, tb, accg);
        }
        // seed: tile (1,1) = G lane for lane; tile (0,0) gets G on its two
        // equal-dof diagonal blocks: its row m (dof m >> 3, node m & 7) and
        // column c take G[m & 7][c & 7], held by the lane and register whose
        // result row is m & 7 in column c & 7
        v4 acc11 = accg, acc00, acc01 = {0, 0, 0, 0};
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int m = Mfma<VT>::row(kq, r);
            constexpr bool f64 = sizeof(VT) == 8;
            // f32: row 4 kq + r -> m & 7 = 4 (kq & 1) + r; f64: row kq + 4 r -> m & 7 = kq + 4 (r & 1)
            const int src = (c & 7) | ((f64 ? kq : (kq & 1)) << 4);
            const VT gv = __shfl(accg[f64 ? (r & 1) : r], src, 64);
            acc00[r] = (m >> 3) == i0 ? gv : VT(0);
        }
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int q = 4 * h + kq;
            VT D[21];
#pragma unroll
            for (int k = 0; k < 21; k++) D[k] = sd[le][q][k];
            const VT g0 = sg[le][q][0][na], g1 = sg[le][q][1][na], g2 = sg[le][q][2][na];
#pragma unroll
            for (int v = 0; v < 6; v++) {
                // B_q[v][(node, dof)] for dofs x, y, z (Hughes 2.8.21, Voigt 11,22,33,23,13,12)
                const VT bx = v == 0 ? g0 : v == 4 ? g2 : v == 5 ? g1 : VT(0);
                const VT by = v == 1 ? g1 : v == 3 ? g2 : v == 5 ? g0 : VT(0);
                const VT bz = v == 2 ? g2 : v == 3 ? g1 : v == 4 ? g0 : VT(0);
                // (D_q B_q)[v][(node, dof)]; both x and y are formed and one
                // kept (a selected register index would become an indexed
                // register access)
                const VT dx = D[sym6(v, 0)] * g0 + D[sym6(v, 4)] * g2 + D[sym6(v, 5)] * g1;
                const VT dy = D[sym6(v, 1)] * g1 + D[sym6(v, 3)] * g2 + D[sym6(v, 5)] * g0;
                const VT dz = D[sym6(v, 2)] * g2 + D[sym6(v, 3)] * g1 + D[sym6(v, 4)] * g0;
                const VT A0 = i0 ? by : bx, B0 = i0 ? dy : dx;
                const VT A1 = lo8 ? bz : VT(0), B1 = lo8 ? dz : VT(0);
                acc00 = Mfma<VT>::mac(A0, B0, acc00);
                acc01 = Mfma<VT>::mac(A0, B1, acc01);
                acc11 = Mfma<VT>::mac(A1, B1, acc11);
            }
        }
        if (more) {
#pragma unroll
            for (int k = 0; k < 12; k++) {
                const VT x = val[nsidx[k] >= 0 ? nsidx[k] : 0];
                nold[k] = nsidx[k] >= 0 ? x : VT(0);
            }
        }
        if (GATHER) {
            VT* eb = val + 300L * (base + le);
#pragma unroll
            for (int r = 0; r < 4; r++) {
                if (ent[r] >= 0) eb[ent[r]] = acc00[r];
                if (ent[4 + r] >= 0) eb[ent[4 + r]] = acc01[r];
                if (ent[8 + r] >= 0) eb[ent[8 + r]] = acc11[r];
            }
            continue;
        }
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int s0 = slot_target(sidx[r]), s1 = slot_target(sidx[4 + r]), s2 = slot_target(sidx[8 + r]);
            if (s0 >= 0) val[s0] = old[r] + acc00[r];
            if (s1 >= 0) val[s1] = old[4 + r] + acc01[r];
            if (s2 >= 0) val[s2] = old[8 + r] + acc11[r];
        }
    }
    if (bad) atomicOr(status, 7);
}

template <class VT, int EPB, bool TO_OP, bool PM, bool DB, int SC, int MAT = MAT_SIMO>
static void launch_stiffness_v(tg_ctx* c, int first, int count, const double* u, VT* val)
{
    const int blocks = (count + EPB - 1) / EPB;
    if constexpr (SC >= 3 && !TO_OP)
        hipLaunchKernelGGL((k_stiffness_mfma<VT, EPB, MAT, SC == 4 ? 3 : 1>), dim3(blocks), dim3(64), 0, c->stream,
                           c->dstiff_elems.p + first, count, c->dconn.p, c->dmat.p, c->dmatprops.p, c->dX.p, u,
                           c->dslot.p, val, c->dstatus.p);
    else if constexpr (SC >= 3)
        // the matrix-core kernel scatters into the MSR only; the block
        // operator gets the loads-first block kernel
        launch_stiffness_v<VT, 12, TO_OP, false, false, 1, MAT>(c, first, count, u, val);
    else if (DB)
        hipLaunchKernelGGL((k_stiffness_db<VT, EPB, TO_OP, MAT>), dim3(blocks), dim3(36 * EPB), 0, c->stream,
                           c->dstiff_elems.p + first, count, c->dconn.p, c->dmat.p, c->dmatprops.p, c->dX.p, u,
                           TO_OP ? c->op.d_oslot.p : c->dslot.p, c->deqn.p, val, c->dstatus.p);
    else
        hipLaunchKernelGGL((k_stiffness_tl<VT, EPB, TO_OP, PM, SC, MAT>), dim3(blocks), dim3(36 * EPB), 0,
                           c->stream, c->dstiff_elems.p + first, count, c->dconn.p, c->dmat.p, c->dmatprops.p,
                           c->dX.p, u, TO_OP ? c->op.d_oslot.p : c->dslot.p, c->deqn.p, val, c->dstatus.p);
}

// Element-buffer pass of the gather assembly: the matrix-core shapes write
// their tiles, every other shape runs the 12-element block kernel.
template <class VT>
static void launch_stiffness_gather(tg_ctx* c, int v, int first, int count, const double* u, VT* eb)
{
    const int blocks8 = (count + 7) / 8, blocks4 = (count + 3) / 4;
    const int* el = c->dstiff_elems.p + first;
    if (c->material_model == TG_MAT_STVENANT)
        hipLaunchKernelGGL((k_stiffness_tl<VT, 12, false, false, 5, MAT_STVENANT>), dim3((count + 11) / 12),
                           dim3(36 * 12), 0, c->stream, el, count, c->dconn.p, c->dmat.p, c->dmatprops.p, c->dX.p, u,
                           c->dslot.p, c->deqn.p, eb, c->dstatus.p);
    else if (v == 16)
        hipLaunchKernelGGL((k_stiffness_mfma<VT, 8, MAT_SIMO, 1, true>), dim3(blocks8), dim3(64), 0, c->stream, el,
                           count, c->dconn.p, c->dmat.p, c->dmatprops.p, c->dX.p, u, c->dslot.p, eb, c->dstatus.p);
    else if (v == 17)
        hipLaunchKernelGGL((k_stiffness_mfma<VT, 4, MAT_SIMO, 1, true>), dim3(blocks4), dim3(64), 0, c->stream, el,
                           count, c->dconn.p, c->dmat.p, c->dmatprops.p, c->dX.p, u, c->dslot.p, eb, c->dstatus.p);
    else if (v == 18)
        hipLaunchKernelGGL((k_stiffness_mfma<VT, 8, MAT_SIMO, 3, true>), dim3(blocks8), dim3(64), 0, c->stream, el,
                           count, c->dconn.p, c->dmat.p, c->dmatprops.p, c->dX.p, u, c->dslot.p, eb, c->dstatus.p);
    else
        hipLaunchKernelGGL((k_stiffness_tl<VT, 12, false, false, 5, MAT_SIMO>), dim3((count + 11) / 12),
                           dim3(36 * 12), 0, c->stream, el, count, c->dconn.p, c->dmat.p, c->dmatprops.p, c->dX.p, u,
                           c->dslot.p, c->deqn.p, eb, c->dstatus.p);
}

template <class VT, bool TO_OP>
static void launch_stiffness_shape(tg_ctx* c, int v, int first, int count, const double* u, VT* val)
{
    if (c->material_model == TG_MAT_STVENANT) {
        // one shape (16 elements per workgroup, load-add-store scatter) for
        // the St.Venant material
        launch_stiffness_v<VT, 16, TO_OP, false, false, 0, MAT_STVENANT>(c, first, count, u, val);
        return;
    }
#define TG_SHAPE(i) \
    launch_stiffness_v<VT, STIFF_EPB[i], TO_OP, STIFF_PM[i], STIFF_DB[i], STIFF_SC[i]>(c, first, count, u, val)
    switch (v) {
    case 0: TG_SHAPE(0); break;
    case 1: TG_SHAPE(1); break;
    case 2: TG_SHAPE(2); break;
    case 3: TG_SHAPE(3); break;
    case 4: TG_SHAPE(4); break;
    case 5: TG_SHAPE(5); break;
    case 6: TG_SHAPE(6); break;
    case 7: TG_SHAPE(7); break;
    case 8: TG_SHAPE(8); break;
    case 9: TG_SHAPE(9); break;
    case 10: TG_SHAPE(10); break;
    case 11: TG_SHAPE(11); break;
    case 12: TG_SHAPE(12); break;
    case 13: TG_SHAPE(13); break;
    case 14: TG_SHAPE(14); break;
    case 15: TG_SHAPE(15); break;
    case 16: TG_SHAPE(16); break;
    case 17: TG_SHAPE(17); break;
    default: TG_SHAPE(18); break;
    }
#undef TG_SHAPE
}

// whether the element-buffer assembly runs; builds its lists and buffer for
// the current launch order on first use (positions * 300 must fit an int)
static bool gather_ready(tg_ctx* c)
{
    if (!c->stiff_gather || c->formulation != TG_TOTAL_LAGRANGIAN || 300L * c->nel >= (1L << 31)) return false;
    stiffness_launches(c);
    if (c->gather_built != c->stiff_chunk_built) {
        build_gather_lists(c, c->dstiff_elems.p);
        c->gather_built = c->stiff_chunk_built;
    }
    const size_t n = 300 * (size_t)c->nel;
    if (mixed_stiffness(c)) {
        if (c->debuf32.n != n) c->debuf32.alloc(n);
    } else if (c->debuf.n != n) {
        c->debuf.alloc(n);
    }
    return true;
}

void launch_stiffness(tg_ctx* c, const double* u)
{
    const bool gather = gather_ready(c);  // (setup on first use, outside the timed region)
    TimedRegion tr(c, T_STIFFNESS);
    const bool mixed = mixed_stiffness(c);
    const int v = std::min(std::max(stiffness_shape(c), 0), N_STIFF_SHAPES - 1);
    if (gather) {
        // every MSR value is written by the gather: no clearing pass
        for (const auto& fc : stiffness_launches(c)) {
            if (mixed)
                launch_stiffness_gather<float>(c, v, fc.first, fc.second, u, c->debuf32.p + 300L * fc.first);
            else
                launch_stiffness_gather<double>(c, v, fc.first, fc.second, u, c->debuf.p + 300L * fc.first);
        }
        TG_HIP(hipGetLastError());
        launch_gather(c);
        return;
    }
    if (mixed)
        TG_HIP(hipMemsetAsync(c->dval32.p, 0, sizeof(float) * (size_t)c->msr_len, c->stream));
    else
        TG_HIP(hipMemsetAsync(c->dval.p, 0, sizeof(double) * (size_t)c->msr_len, c->stream));
    if (c->formulation == TG_Q1P0) {
        launch_q1p0(c, u, c->dR.p, true);
        return;
    }
    for (const auto& fc : stiffness_launches(c)) {
        if (mixed)
            launch_stiffness_shape<float, false>(c, v, fc.first, fc.second, u, c->dval32.p);
        else
            launch_stiffness_shape<double, false>(c, v, fc.first, fc.second, u, c->dval.p);
    }
    TG_HIP(hipGetLastError());
}

// Newton's stiffness straight into the solver operator (total-Lagrangian
// element): cleared values, the colour-ordered passes of TO_OP kernels.
void launch_stiffness_operator(tg_ctx* c, const double* u)
{
    TimedRegion tr(c, T_STIFFNESS);
    const bool mixed = c->precision == TG_MIXED;
    const size_t nv = (size_t)c->op.nblocks_padded * 9;
    if (mixed)
        TG_HIP(hipMemsetAsync(c->op.d_val32.p, 0, sizeof(float) * nv, c->stream));
    else
        TG_HIP(hipMemsetAsync(c->op.d_val.p, 0, sizeof(double) * nv, c->stream));
    const int v = std::min(std::max(stiffness_shape(c), 0), N_STIFF_SHAPES - 1);
    for (const auto& fc : stiffness_launches(c)) {
        if (mixed)
            launch_stiffness_shape<float, true>(c, v, fc.first, fc.second, u, c->op.d_val32.p);
        else
            launch_stiffness_shape<double, true>(c, v, fc.first, fc.second, u, c->op.d_val.p);
    }
    TG_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// slot map: element upper entry (m <= n) -> MSR value index
// ---------------------------------------------------------------------------

// Position (value index of component 0) of operator block (row R, column C):
// slice R / 64, lane R % 64, the slot whose column is C.  A row lists its
// columns ascending, then padding (-1, taken as larger than any node); -1
// when C is not a neighbour of R.
__device__ __forceinline__ int block_pos(int R, int C, const int* __restrict__ sbase, const int* __restrict__ swidth,
                                         const int* __restrict__ bcol)
{
    const int s = R >> 6, l = R & 63;
    int lo = 0, hi = swidth[s] - 1;
    while (lo <= hi) {
        const int mid = (lo + hi) >> 1;
        const int x = bcol[((long)sbase[s] + mid) * 64 + l];
        if (x == C) return (int)((((long)sbase[s] + mid) * 9) * 64 + l);
        if (x >= 0 && x < C)
            lo = mid + 1;
        else
            hi = mid - 1;
    }
    return -1;
}

// Slot maps of element e, one thread per node pair p = (a, b), a <= b in
// local upper order, nodes A = conn[a], B = conn[b]:
//  - slot (MSR path): MSRMatrixT::Assemble's symmetric branch
//    (MSRMatrixT.cpp:110-165) sends the element value at local (m, n),
//    m <= n, to the upper entry (min(eq_m, eq_n), max(eq_m, eq_n)) -- the
//    diagonal val[row], an off-diagonal its place in the sorted row
//    (AssembleRow, :532-604).  That is the expansion's value map (d_src) of
//    entry (i, j) of the upper one of the blocks (A, B) and (B, A): node-
//    major equation numbers order the two nodes' equations as the nodes.
//  - oslot (operator path, when given): the positions of blocks (A, B) and
//    (B, A); the second is -1 for a diagonal pair.
// Two searches in the sorted block rows per pair replace a search in the
// MSR row per entry.
__global__ void k_element_slots(int nel, const int* __restrict__ conn, const int* __restrict__ sbase,
                                const int* __restrict__ swidth, const int* __restrict__ bcol,
                                const int* __restrict__ src, int* __restrict__ slot, int* __restrict__ oslot)
{
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= 36L * nel) return;
    const int e = (int)(idx / 36), p = (int)(idx % 36);
    int a, b;
    upper_block(p, a, b);
    const int A = conn[8L * e + a], B = conn[8L * e + b];
    const int pab = block_pos(A, B, sbase, swidth, bcol);
    const int pba = a == b ? -1 : block_pos(B, A, sbase, swidth, bcol);
    if (oslot) {
        oslot[72L * e + 2 * p] = pab;
        oslot[72L * e + 2 * p + 1] = pba;
    }
    int* es = slot + 300L * e;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const int m = 3 * a + i, n = 3 * b + j;
            if (n < m) continue;  // lower half of a diagonal block
            int v = -1;
            if (A <= B) {
                if (pab >= 0) v = src[pab + (3 * i + j) * 64];
            } else if (pba >= 0) {
                v = src[pba + (3 * j + i) * 64];
            }
            es[upper_index(m, n)] = v;
        }
}

// value index of component 0 of each row's diagonal block
__global__ void k_build_operator_diag(int nnd, const int* __restrict__ sbase, const int* __restrict__ swidth,
                                      const int* __restrict__ bcol, int* __restrict__ diag)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= nnd) return;
    diag[a] = block_pos(a, a, sbase, swidth, bcol);
}

// both slot maps (the operator's only when d_oslot is allocated: the
// total-Lagrangian element), from the block columns and the value map
void launch_element_slots(tg_ctx* c)
{
    const long n = 36L * c->nel;
    const int bs = 256;
    hipLaunchKernelGGL(k_element_slots, dim3((unsigned)((n + bs - 1) / bs)), dim3(bs), 0, c->stream, c->nel,
                       c->dconn.p, c->op.d_slice_base.p, c->op.d_slice_width.p, c->op.d_col.p, c->op.d_src.p,
                       c->dslot.p, c->op.d_oslot.n ? c->op.d_oslot.p : nullptr);
    TG_HIP(hipGetLastError());
}

void launch_build_operator_diag(tg_ctx* c)
{
    const int bs = 256;
    hipLaunchKernelGGL(k_build_operator_diag, dim3((unsigned)((c->nnd + bs - 1) / bs)), dim3(bs), 0, c->stream,
                       c->nnd, c->op.d_slice_base.p, c->op.d_slice_width.p, c->op.d_col.p, c->op.d_diag.p);
    TG_HIP(hipGetLastError());
}

// Columns of the upper MSR rows of node A (slice s, lane l), from the block
// operator's neighbour list (ascending, padding -1 at the end): row
// r = eq(A, i) takes every active eq(B, j) > r.  Equation numbers grow with
// (node, dof), so the columns come out ascending and their count is the row
// length counted beforehand (bindx[r + 1] - bindx[r], already on the device);
// the write stays inside the row's range whatever the count.
__global__ void k_msr_cols(int nnd, const int* __restrict__ sbase, const int* __restrict__ swidth,
                           const int* __restrict__ bcol, const int* __restrict__ eqn, int* __restrict__ bindx)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= nnd) return;
    const int s = a >> 6, l = a & 63;
    for (int i = 0; i < 3; i++) {
        const int r = eqn[3L * a + i];
        if (r < 0) continue;
        int pos = bindx[r];
        const int end = bindx[r + 1];
        for (int k = 0; k < swidth[s]; k++) {
            const int B = bcol[((long)sbase[s] + k) * 64 + l];
            if (B < 0) break;
            for (int j = 0; j < 3; j++) {
                const int cc = eqn[3L * B + j];
                if (cc > r && pos < end) bindx[pos++] = cc;
            }
        }
    }
}

// Operator value -> MSR value it mirrors, for the diagonal and upper blocks
// of row node A (slice s, lane l): entry (i, j) is the upper MSR entry
// (min(r, c), max(r, c)) of r = eq(A, i), c = eq(B, j); -1 when either dof is
// prescribed, for padding, and for lower blocks (column B < A), which the
// expansion copies from their upper partners.  Equation numbers grow with
// (node, dof), so row r = eq(A, i) holds the active dofs of A after i, then
// the active dofs of each neighbour B > A in ascending order: an entry's
// position is the row start plus a running count over the slots, with no
// search.
__global__ void k_build_src(int nnd, const int* __restrict__ sbase, const int* __restrict__ swidth,
                            const int* __restrict__ bcol, const int* __restrict__ eqn,
                            const int* __restrict__ bindx, int* __restrict__ src)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    const int s = a >> 6, l = a & 63;
    if (s >= (nnd + 63) / 64) return;
    int r[3], first[3];  // row of each dof of A (-1 prescribed); its first column after A's own dofs
#pragma unroll
    for (int i = 0; i < 3; i++) r[i] = a < nnd ? eqn[3L * a + i] : -1;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        int own = 0;
        for (int j = i + 1; j < 3; j++) own += r[j] >= 0;
        first[i] = r[i] >= 0 ? bindx[r[i]] + own : 0;
    }
    int run = 0;  // active dofs of the neighbours after A and before B
    for (int k = 0; k < swidth[s]; k++) {
        const long g = (long)sbase[s] + k;
        const int B = a < nnd ? bcol[g * 64 + l] : -1;
        int cb[3] = {-1, -1, -1};
        if (B > a)
#pragma unroll
            for (int j = 0; j < 3; j++) cb[j] = eqn[3L * B + j];
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) {
                int v = -1;
                if (B == a && r[i] >= 0 && r[j] >= 0) {
                    const int lo = i < j ? i : j, hi = i < j ? j : i;
                    v = i == j ? r[i] : bindx[r[lo]] + (hi - lo == 2 && r[lo + 1] >= 0 ? 1 : 0);
                } else if (B > a && r[i] >= 0 && cb[j] >= 0) {
                    v = first[i] + run + (j > 0 && cb[0] >= 0) + (j > 1 && cb[1] >= 0);
                }
                src[(g * 9 + 3 * i + j) * 64 + l] = v;
            }
        if (B > a) run += (cb[0] >= 0) + (cb[1] >= 0) + (cb[2] >= 0);
    }
}

// Lower blocks (column node B < row node A) are the transposes of upper
// blocks (B, A): the slot k' of column A in row B gives the partner's value
// index, so the expansion copies them from the operator itself instead of
// gathering them from MSR rows of other nodes.
__global__ void k_build_tpos(int nnd, const int* __restrict__ sbase, const int* __restrict__ swidth,
                             const int* __restrict__ bcol, int* __restrict__ tpos)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    const int s = a >> 6, l = a & 63;
    if (s >= (nnd + 63) / 64) return;
    for (int k = 0; k < swidth[s]; k++) {
        const long g = (long)sbase[s] + k;
        const int B = a < nnd ? bcol[g * 64 + l] : -1;
        int out = -1;
        if (B >= 0 && B < a) {
            // row B lists its neighbours ascending, then padding (-1): a
            // binary search with padding taken as larger than any node
            const int sB = B >> 6, lB = B & 63;
            int lo = 0, hi = swidth[sB] - 1;
            while (lo <= hi) {
                const int mid = (lo + hi) >> 1;
                const int x = bcol[((long)sbase[sB] + mid) * 64 + lB];
                if (x == a) {
                    out = (int)((((long)sbase[sB] + mid) * 9) * 64 + lB);
                    break;
                }
                if (x >= 0 && x < a)
                    lo = mid + 1;
                else
                    hi = mid - 1;
            }
        }
        tpos[g * 64 + l] = out;
    }
}

void launch_msr_cols(tg_ctx* c)
{
    const int bs = 256;
    if (c->nnd == 0) return;
    hipLaunchKernelGGL(k_msr_cols, dim3((unsigned)((c->nnd + bs - 1) / bs)), dim3(bs), 0, c->stream, c->nnd,
                       c->op.d_slice_base.p, c->op.d_slice_width.p, c->op.d_col.p, c->deqn.p, c->dbindx.p);
    TG_HIP(hipGetLastError());
}

void launch_build_src(tg_ctx* c)
{
    const int bs = 256;
    const long n = 64L * c->op.nslices;
    hipLaunchKernelGGL(k_build_src, dim3((unsigned)((n + bs - 1) / bs)), dim3(bs), 0, c->stream, c->nnd,
                       c->op.d_slice_base.p, c->op.d_slice_width.p, c->op.d_col.p, c->deqn.p, c->dbindx.p,
                       c->op.d_src.p);
    hipLaunchKernelGGL(k_build_tpos, dim3((unsigned)((n + bs - 1) / bs)), dim3(bs), 0, c->stream, c->nnd,
                       c->op.d_slice_base.p, c->op.d_slice_width.p, c->op.d_col.p, c->op.d_tpos.p);
    TG_HIP(hipGetLastError());
}

}  // namespace tg
