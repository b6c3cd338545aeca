// Internal declarations shared by the host setup code, the C ABI and the
// kernels of libtahoe_gpu.so.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <memory>
#include <utility>
#include <string>
#include <thread>
#include <vector>

#include "../../include/tahoe_gpu.h"

namespace tg {

// Allocator whose value-less construct leaves the element uninitialized:
// resize() of a vector<int, NoInit<int>> does not zero-fill, so the large
// index arrays of the setup (the MSR pattern, the node adjacency) are
// written once, by the parallel passes that fill them.
template <class T>
struct NoInit : std::allocator<T> {
    NoInit() = default;
    template <class U>
    NoInit(const NoInit<U>&) noexcept {}
    template <class U>
    struct rebind {
        using other = NoInit<U>;
    };
    template <class U>
    void construct(U* p) noexcept
    {
        ::new ((void*)p) U;
    }
    template <class U, class... A>
    void construct(U* p, A&&... a)
    {
        ::new ((void*)p) U(std::forward<A>(a)...);
    }
};
using ivec = std::vector<int, NoInit<int>>;

// host buffers of node_graph (capacity kept across set-ups)
struct GraphWork {
    std::vector<int> ptr, deg;
    ivec list;
    std::vector<ivec> part;
    std::vector<std::vector<int>> stamp;
};

// Status raised inside the library; converted to the C status at the ABI.
struct Error {
    int code;
    std::string msg;
};

[[noreturn]] void fail(int code, const std::string& msg);
void hip_check(hipError_t e, const char* what);

#define TG_HIP(call) ::tg::hip_check((call), #call)

// Host threads for set-up passes (at most 16).
inline int worker_count()
{
    unsigned n = std::thread::hardware_concurrency();
    if (n == 0) n = 4;
    return (int)std::min(n, 16u);
}

// f(begin, end) over [0, n) split among worker_count() threads; small
// ranges run on the calling thread.
template <class F>
void parallel_for(long n, F f)
{
    const int nt = worker_count();
    if (n < 4096 || nt == 1) {
        f(0L, n);
        return;
    }
    std::vector<std::thread> th;
    const long chunk = (n + nt - 1) / nt;
    for (int t = 0; t < nt; t++) {
        const long b = t * chunk, e = std::min(n, b + chunk);
        if (b >= e) break;
        th.emplace_back([=] { f(b, e); });
    }
    for (auto& x : th) x.join();
}

// dst = src[0, n), copied by the set-up threads
template <class V, class T>
void parallel_copy(V& dst, const T* src, long n)
{
    dst.resize((size_t)n);
    parallel_for(n, [&](long b, long e) { std::copy(src + b, src + e, dst.data() + b); });
}

// f(t, begin, end) for the t-th of `parts` contiguous ranges of [0, n), one
// thread each (t = 0 on the calling thread when parts == 1).
template <class F>
void parallel_parts(long n, int parts, F f)
{
    const long chunk = (n + parts - 1) / std::max(parts, 1);
    if (parts <= 1) {
        f(0, 0L, n);
        return;
    }
    std::vector<std::thread> th;
    for (int t = 0; t < parts; t++) {
        const long b = std::min(n, t * chunk), e = std::min(n, b + chunk);
        th.emplace_back([=] { f(t, b, e); });
    }
    for (auto& x : th) x.join();
}

// Device buffer owned by a context.  alloc keeps an allocation that is
// already large enough (n shrinks, the memory stays), so a context set up
// again for a mesh no larger than its largest so far calls no hipMalloc or
// hipFree; alloc(0) and release free it.  The contents are undefined after
// alloc either way.
template <class T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;    // elements in use
    size_t cap = 0;  // elements allocated
    void alloc(size_t count)
    {
        if (count == 0) {
            release();
            return;
        }
        if (count <= cap) {
            n = count;
            return;
        }
        release();
        TG_HIP(hipMalloc((void**)&p, count * sizeof(T)));
        n = cap = count;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = cap = 0;
    }
    ~DBuf() { release(); }
};

// Sliced-ELL layout of the full (both triangles) 3x3-block operator used by
// the PCG SpMV.  Block rows = nodes; slice s covers block rows
// [64 s, 64 s + 64); within a slice, block k of lane l lives at
// (slice_base[s] + k) * 64 + l; its 9 values are stored component-major
// over the slice: val[((slice_base[s] + k) * 9 + c) * 64 + l].
struct Bsell {
    int nslices = 0;
    long nblocks_padded = 0;               // sum over slices of width * 64
    std::vector<int> slice_base, slice_width;
    DBuf<int> d_slice_base, d_slice_width;
    DBuf<int> d_col;                       // [nblocks_padded] column node, -1 = padding
    DBuf<int> d_src;                       // [nblocks_padded * 9] MSR val index, -1 = zero (and lower blocks: copied from partners)
    // per half slice (32 block rows): the contiguous MSR ranges its rows own
    // -- off-diagonal values [e0, e1) and diagonal values [r0, r1) -- staged
    // in LDS by the expansion; lds_values = the largest (e1 - e0 + r1 - r0)
    DBuf<int4> d_srange;                   // [2 nslices] (e0, e1, r0, r1)
    long lds_values = 0;
    DBuf<int> d_tpos;                      // [nblocks_padded] for a lower block (column < row): value
                                           // index of its transpose partner, else -1
    DBuf<double> d_val;                    // [nblocks_padded * 9]
    DBuf<float> d_val32;                   // FP32 copy for the mixed-precision inner solve
    DBuf<int> d_oslot;                     // [nel][36][2] element block (a<=b) -> value index of
                                           // block (A,B) and of (B,A) (-1 when a == b)
    DBuf<int> d_diag;                      // [nnd] value index of the diagonal block of each row
};

struct Timing {
    bool enabled = false;
    tg_timing acc{};
};

// One level of the aggregation multigrid preconditioner.  Level 0 is the
// solver's block operator (3 dofs per node); every coarser level has 6 dofs
// per node (the rigid-body modes of an aggregate) and its matrix is a block
// CSR of 6x6 FP32 blocks, row-major inside a block.
struct AmgLevel {
    int n = 0, bs = 3;                       // nodes, dofs per node
    std::vector<int> rowptr, col;            // pattern (levels >= 1; sorted, diagonal included)
    // levels >= 1 are stored sliced like the fine operator: slice s holds
    // block rows [64 s, 64 s + 64); entry k of row a is slot
    // (sbase[a / 64] + k) * 64 + a % 64; the 36 values of slot e (row-major
    // inside the block) are val[(e / 64) * 2304 + comp * 64 + e % 64]
    std::vector<int> sbase, swidth, slot;    // slot: CSR entry -> slot
    long nslots = 0;
    DBuf<int> d_sbase, d_swidth, d_col, d_diagblk;  // d_col: column node per slot, -1 padding; d_diagblk: slot
    DBuf<float> d_val;                       // [nslots][36], component-major per slice
    DBuf<double> d_dinv_own;                 // [n*bs] inverse diagonal (levels >= 1)
    double* dinv = nullptr;                  // level 0: the PCG's Jacobi vector
    DBuf<double> d_f, d_x, d_r, d_d0, d_d1;  // V-cycle vectors [n*bs]
    // transfer to level + 1
    int nnext = 0;
    DBuf<int> d_agg;                         // [n] aggregate (node of level + 1), -1 none
    DBuf<int> d_mptr, d_mem;                 // members of each aggregate (CSR)
    DBuf<float> d_P;                         // [n][bs][6] tentative prolongator rows
    long nup = 0;                            // upper blocks (A <= B) of level + 1
    DBuf<int> d_up;                          // [nup][2] slot of (A,B) and of (B,A) (-1 when A == B) in level + 1
    DBuf<long> d_gptr;                       // [nup+1]
    DBuf<int> d_glist;                       // [entries][2] (slot in this level, block row)
    double lmax = 0.0;
};

struct Amg {
    int gen = 0;                             // bumped on every rebuild or knob change (captured PCG graphs key on it)
    int degree = 2;                          // Chebyshev degree per smoothing pass
    int weight = 150;                        // coarse-grid correction weight, percent
    int reuse = 6;                           // further tangents of one Newton solve served by one numeric set-up
                                             // of the coarse levels (a 128^3 Newton step: 1.4x faster, fewer PCG
                                             // iterations)
    int since = -1;                          // tangents served since the last numeric set-up (-1: none yet)
    std::vector<AmgLevel*> lv;               // owned
    int ncd = 0;                             // dofs of the coarsest level (dense solve)
    DBuf<double> d_cinv;                     // [ncd][ncd] inverse of the coarsest matrix
    DBuf<double> d_coef;                     // [levels][8] Chebyshev coefficients
    DBuf<double> d_red;                      // reduction scratch
    DBuf<double> d_part;                     // reduction partials
    void clear()
    {
        for (auto* p : lv) delete p;
        lv.clear();
        ncd = 0;
        since = -1;
    }
    ~Amg() { clear(); }
};

// GMRES work space (k_gmres.hip)
struct Gmres {
    int m = 0;
    long n3 = 0;
    DBuf<double> V;                          // [m+1][3 nnd] Krylov basis
    DBuf<double> w, q;                       // [3 nnd]
    DBuf<double> part, h;                    // reduction partials, projections
};

// device scalar slots of an iteration (per segment for the PCG)
enum { S_RZ = 0, S_PQ, S_ALPHA, S_RZNEW, S_RR, S_BETA, S_BB, S_DONE, S_ITER, S_THRESH, S_MAXIT, S_DOT, S_NSCAL };

// Independent problems packed in one context: segment g owns nodes
// [node_begin[g], node_begin[g+1]), each starting at a multiple of 64, so a
// 64-node slice (one wave) never spans two segments.
struct Segments {
    int nseg = 1;
    long nslices = 0;
    std::vector<int> node_begin;          // [nseg+1]
    std::vector<int> elem_seg;            // [nel] segment of each element (host)
    DBuf<int> d_slice_seg;                // [nslices]
    DBuf<int> d_seg_slice;                // [nseg+1] first slice of each segment
    DBuf<int> d_active;                   // [nseg] 1 = solve this segment
    DBuf<int> d_elem_seg;                 // [nel]
    DBuf<double> d_scal;                  // [S_NSCAL][nseg]
    DBuf<double> d_part;                  // [2][nslices]
    DBuf<double> d_out;                   // [nseg]
    DBuf<int> d_ndone;
    DBuf<int> d_segstat;                  // [nseg] 7 when an element of the segment inverted in a residual pass
    std::vector<double> h_scal;
    // segments solved by the persistent PCG kernel (at most persist_built
    // slices each), rebuilt when the threshold or the segments change
    long persist_built = -2;
    std::vector<int> h_persist;
    DBuf<int> d_persist;
};

}  // namespace tg

struct tg_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;

    // problem definition (host)
    int nnd = 0, nel = 0;
    std::vector<double> X;
    std::vector<int> conn, mat;
    std::vector<double> mu, kappa;
    int formulation = TG_TOTAL_LAGRANGIAN;
    int material_model = TG_MAT_SIMO_ISOTROPIC;
    int precision = TG_FP64;
    std::vector<int> fbc_idx;           // force cards: 3*node+dof
    std::vector<double> fbc_vals;
    tg::DBuf<double> dfext;             // [nnd*3] applied nodal forces (allocated once a card exists)
    std::vector<int> kbc_nodes, kbc_dofs;
    std::vector<double> kbc_vals;

    // equation system
    bool initialized = false;
    int n_eq = 0;
    std::vector<int> eqnos;   // [nnd*3] 1-based, -1 prescribed
    long msr_len = 0;
    std::vector<int> color_ptr, color_elems;
    int ncolors = 0;
    tg::GraphWork graph_work;           // host node_graph's buffers (multigrid set-up), kept across set-ups
    // device structure build (k_graph.hip), kept across set-ups
    tg::DBuf<int> g_neptr, g_nelist, g_fill, g_flag, g_nlist;
    tg::DBuf<long long> g_deg, g_nptr, g_rowlen, g_rows, g_most;
    tg::DBuf<unsigned char> g_scan;
    // stiffness launch order: every segment's elements in ranges of
    // stiff_chunk (by rank inside the segment), all colours of range j
    // before range j+1 (0: whole colours)
    long stiff_chunk = -1;              // -1: chosen from the matrix footprint
    int stiff_variant = -1;             // shape of the stiffness kernel; -1: 17 (matrix cores, 4 elements
                                        // per wave) for the FP32 stiffness, 9 (12 elements, loads first) for FP64
    int newton_assembly = 0;            // 0: Newton's K into the MSR then expanded, 1: straight into the operator
    int precond = TG_PC_JACOBI;         // preconditioner of the inner PCG
    int linear_solver = TG_LIN_PCG;     // inner linear solver of the Newton loop
    int gmres_restart = 30;
    int pcg_graph = 1;                  // replay PCG chunks as a captured HIP graph
    int spmv_xcd = -1;                  // PCG SpMV variant (bits: XCD rows, unrolled, symmetric); -1 auto
    int pcg_fused = -1;                 // PCG direction formed inside the SpMV; -1 auto
    int pcg_persist = -1;               // max slices of a segment run by the persistent PCG; 0 off, -1 auto
    tg::DBuf<double> vp2;               // second direction buffer of the fused iteration
    struct {
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        const void* key[24] = {};
    } pg;                               // the captured chunk and the arguments it was captured with
    hipEvent_t poll_ev[2] = {nullptr, nullptr};
    long stiff_chunk_built = -2;
    // stiffness assembly through the element buffer (TG_TUNE_STIFF_GATHER):
    // the element kernels write their 300 upper entries to ebuf[position in
    // launch order][300]; one gather pass sums each MSR value's
    // contributions, listed in launch order, and writes it once
    int stiff_gather = 0;
    long gather_built = -2;             // the launch order (chunk) the lists were built for
    tg::DBuf<int> dgptr;                // [msr_len + 1] first list entry of each MSR value
    tg::DBuf<int> dglist;               // [contributions] launch position * 300 + element entry
    tg::DBuf<double> debuf;             // [nel][300] element upper entries (FP64)
    tg::DBuf<float> debuf32;            // [nel][300] element upper entries (FP32, mixed precision)
    std::vector<std::pair<int, int>> stiff_launches;  // (first, count) into dstiff_elems

    // device storage
    tg::DBuf<double> dX, du, dR;        // node space [nnd*3]
    tg::DBuf<double> du_last;           // last converged displacement (CloseStep/ResetStep)
    tg::DBuf<int> dconn, dmat, deqn;    // deqn: 0-based equation or -1
    tg::DBuf<int> dcolor_elems;
    tg::DBuf<int> dstiff_elems;         // elements in stiffness launch order
    tg::DBuf<double2> dmatprops;        // (mu, kappa) per material
    tg::DBuf<int> dkbc_idx;             // 3*node+dof per card
    tg::DBuf<double> dkbc_val;
    tg::DBuf<int> dbindx;
    tg::DBuf<double> dval;
    tg::DBuf<float> dval32;             // mixed precision: FP32 stiffness values
    tg::DBuf<int> dslot;                // [nel*300] element upper entry -> val index (-1 none)
    tg::Bsell op;

    // PCG work vectors in node space
    tg::DBuf<double> vx, vr, vz, vp, vq, vdinv, vwork;
    tg::DBuf<double> vbinv;             // [nnd][9] inverse diagonal blocks (block Jacobi)
    tg::DBuf<double> dpart;             // partial reduction sums
    tg::DBuf<double> dscal;             // device scalars
    tg::DBuf<int> dstatus;              // device error flag (7 = bad Jacobian)
    double* hscal = nullptr;            // pinned host mirror of a few scalars
    int* hstatus = nullptr;

    // Q1P0 per-element history
    tg::DBuf<double> dvol, dvol_last;

    tg::Timing timing;
    tg::Segments seg;
    tg::Amg amg;                        // aggregation multigrid (precond == TG_PC_AMG)
    tg::Gmres gm;                       // GMRES basis (linear_solver == TG_LIN_GMRES)
};

namespace tg {

// Mixed precision: the total-Lagrangian stiffness is assembled in FP32; the
// Q1P0 stiffness (a forward difference of the FP64 internal force) stays
// FP64 and is narrowed when it becomes the inner solve's operator, which is
// FP32 for every element in mixed precision.
inline bool mixed_stiffness(const tg_ctx* c) { return c->precision == TG_MIXED && c->formulation == TG_TOTAL_LAGRANGIAN; }
inline bool mixed_operator(const tg_ctx* c) { return c->precision == TG_MIXED; }
// the stiffness shape in effect (the measured fastest untuned choice at 128^3)
inline int stiffness_shape(const tg_ctx* c) { return c->stiff_variant >= 0 ? c->stiff_variant : (mixed_stiffness(c) ? 17 : 9); }
// PCG variants in effect.  Untuned: a single problem of 2^17 nodes or more
// reads lower blocks through their partners (XCD row ranges, unrolled loop)
// and forms the direction in its own pass, the fastest pair at 128^3 in
// FP64 and mixed precision.  A batch of 2^17 nodes or more reads the full
// operator in index order with the unrolled loop, direction in its own pass
// (the fastest pair on 2 M-element batches of RVE cubes of edge 4-24: the
// symmetric read is slower on small cubes, whose rows are mostly boundary
// rows, and the fused direction's extra gathers cost more than the launch
// it saves).  Smaller problems and batches, which are launch-bound, read in
// hardware order with the direction formed inside the SpMV.  Every choice
// gives bitwise the same iterates.
inline bool large_single(const tg_ctx* c) { return c->seg.nseg <= 1 && c->nnd >= (1 << 17); }
inline bool large_total(const tg_ctx* c) { return c->nnd >= (1 << 17); }
inline int spmv_variant(const tg_ctx* c)
{
    return c->spmv_xcd >= 0 ? c->spmv_xcd : (large_single(c) ? 7 : large_total(c) ? 2 : 0);
}
inline bool pcg_fused(const tg_ctx* c) { return c->pcg_fused >= 0 ? c->pcg_fused != 0 : !large_total(c); }
// segments of at most this many slices run their whole PCG in one
// persistent workgroup of 1024 threads, one node each (0: none).  Default
// 12 slices (768 nodes, an 8^3-element cube): there it halves a PCG
// iteration of a batch of such problems; a segment of 16 slices is already
// faster with a kernel per step.
constexpr int PCG_PERSIST_MAX = 16;
constexpr int PCG_PERSIST_DEFAULT = 12;
inline int pcg_persist(const tg_ctx* c) { return c->pcg_persist >= 0 ? c->pcg_persist : PCG_PERSIST_DEFAULT; }

// TG_SETUP_TRACE=1 prints host-side set-up phases to stderr, each timed
// to a stream synchronisation; otherwise a no-op.
struct SetupTrace {
    tg_ctx* c;
    const char* who;
    bool on;
    std::chrono::steady_clock::time_point t;
    SetupTrace(tg_ctx* ctx, const char* name);
    void operator()(const char* what);
};

// host setup (tg_setup.cpp)
int equation_numbers(int nnd, const std::vector<int>& kn, const std::vector<int>& kd, std::vector<int>& eqnos);
void node_graph(int nnd, int nel, const int* conn, std::vector<long>& nptr, ivec& nlist, GraphWork* w = nullptr);
int color_elements(int nnd, int nel, const int* conn, std::vector<int>& ptr, std::vector<int>& elems);
// device structure (k_graph.hip): the node graph, then the operator layout,
// block columns, MSR row pointers and columns, value map and partners
void device_node_graph(tg_ctx* c);
void device_structure(tg_ctx* c);

// kernels launched from the ABI (k_element.hip, k_solver.hip)
void launch_residual(tg_ctx* c, const double* u, double* R_node);
void launch_stiffness(tg_ctx* c, const double* u);
// gather lists of the element-buffer assembly (k_gather.hip) for the
// current launch order; dorder = elements in launch order
void build_gather_lists(tg_ctx* c, const int* dorder);
void launch_gather(tg_ctx* c);
void launch_stiffness_operator(tg_ctx* c, const double* u);
void launch_element_slots(tg_ctx* c);
void launch_msr_cols(tg_ctx* c);
void launch_build_src(tg_ctx* c);
void launch_operator_dinv(tg_ctx* c);
void launch_block_inverse(tg_ctx* c);
void launch_build_operator_diag(tg_ctx* c);
void launch_expand(tg_ctx* c);
void launch_spmv(tg_ctx* c, const double* x, double* y);
void launch_expand32(tg_ctx* c);
void launch_narrow_operator(tg_ctx* c);
void widen_values(tg_ctx* c);
void pcg_solve_mixed(tg_ctx* c, const double* b_node, double* x_node, double rtol, int maxit, tg_lin_stats* st);
double device_dot(tg_ctx* c, const double* a, const double* b, long n);
void pcg_solve(tg_ctx* c, const double* b_node, double* x_node, double rtol, int maxit, tg_lin_stats* st);
// Device work space of one PCG run besides its vectors: per-slice reduction
// partials [2][nslices], per-segment scalar slots [S_NSCAL][nseg], the done
// count [1] and the second direction buffer of the fused iteration [3 nnd];
// graph = false issues every launch directly (no capture, no instantiation).
struct PcgWork {
    double* part = nullptr;
    double* scal = nullptr;
    int* ndone = nullptr;
    double* p2 = nullptr;
    bool graph = true;
};
// work = nullptr: the context's own buffers
template <class MT>
void pcg_run(tg_ctx* c, const MT* val, const double* dinv, const double* b, double* x, double* r, double* p,
             double* q, double rtol, int maxit, const int* d_active, int nactive, tg_lin_stats* st,
             const PcgWork* work = nullptr);
void segment_dots(tg_ctx* c, const double* a, const double* b, double* host_out);
void gmres_solve(tg_ctx* c, const double* b, double* x, double rtol, int maxit, tg_lin_stats* st);
// aggregation multigrid (tg_amg.cpp: structure on the host; k_amg.hip: values and cycles)
void amg_setup(tg_ctx* c);
void amg_numeric(tg_ctx* c);
void amg_vcycle(tg_ctx* c, const double* f, double* z, const double* done);
void eq_to_node(tg_ctx* c, const double* src_eq, double* dst_node);
void node_to_eq(tg_ctx* c, const double* src_node, double* dst_eq);
void launch_axpy(tg_ctx* c, double* y, const double* x, double a, long n);
void launch_apply_kbc(tg_ctx* c, double t);
// releases the captured PCG chunk (its kernels hold buffer addresses and sizes)
void drop_pcg_graph(tg_ctx* c);

// event-timed launch region
struct TimedRegion {
    tg_ctx* c;
    int which;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    TimedRegion(tg_ctx* ctx, int w);
    ~TimedRegion();
};
enum { T_RESIDUAL = 0, T_STIFFNESS, T_EXPAND, T_SPMV, T_VECTOR, T_PRECOND };

}  // namespace tg
