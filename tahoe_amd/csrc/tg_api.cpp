// C ABI of libtahoe_gpu.so (declared in include/tahoe_gpu.h).
//
// Each entry point converts internal errors into the reference's
// ExceptionT::CodeT values.  Host<->device copies happen only where the
// reference interface hands over host arrays; the Newton loop itself keeps
// every vector on the device.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <new>
#include <thread>

#include "tg_internal.h"

namespace tg {

void fail(int code, const std::string& msg) { throw Error{code, msg}; }

void hip_check(hipError_t e, const char* what)
{
    if (e == hipSuccess) return;
    const int code = (e == hipErrorOutOfMemory) ? TG_OUT_OF_MEMORY : TG_GENERAL_FAIL;
    fail(code, std::string(what) + ": " + hipGetErrorString(e));
}

// ---- event timing -----------------------------------------------------------

struct PendingEvent {
    int which;
    hipEvent_t e0, e1;
};

}  // namespace tg

// per-context list of recorded event pairs (kept outside tg_ctx to keep the
// struct layout independent of instrumentation)
#include <unordered_map>
#include <mutex>
static std::mutex g_ev_mu;
static std::unordered_map<const tg_ctx*, std::vector<tg::PendingEvent>> g_events;

namespace tg {

TimedRegion::TimedRegion(tg_ctx* ctx, int w) : c(ctx), which(w)
{
    if (!c->timing.enabled) return;
    TG_HIP(hipEventCreate(&e0));
    TG_HIP(hipEventCreate(&e1));
    TG_HIP(hipEventRecord(e0, c->stream));
}

TimedRegion::~TimedRegion()
{
    if (!e0) return;
    if (hipEventRecord(e1, c->stream) != hipSuccess) return;
    std::lock_guard<std::mutex> lk(g_ev_mu);
    g_events[c].push_back(PendingEvent{which, e0, e1});
}

static void drain_events(tg_ctx* c)
{
    std::vector<PendingEvent> ev;
    {
        std::lock_guard<std::mutex> lk(g_ev_mu);
        auto it = g_events.find(c);
        if (it == g_events.end()) return;
        ev.swap(it->second);
    }
    TG_HIP(hipStreamSynchronize(c->stream));
    for (auto& p : ev) {
        float ms = 0.0f;
        TG_HIP(hipEventElapsedTime(&ms, p.e0, p.e1));
        tg_timing& a = c->timing.acc;
        switch (p.which) {
            case T_RESIDUAL: a.residual_ms += ms; a.residual_launches++; break;
            case T_STIFFNESS: a.stiffness_ms += ms; a.stiffness_launches++; break;
            case T_EXPAND: a.expand_ms += ms; a.expand_launches++; break;
            case T_SPMV: a.spmv_ms += ms; a.spmv_launches++; break;
            case T_PRECOND: a.precond_ms += ms; a.precond_launches++; break;
            default: a.vector_ms += ms; a.vector_launches++; break;
        }
        (void)hipEventDestroy(p.e0);
        (void)hipEventDestroy(p.e1);
    }
}

static void set_device(tg_ctx* c) { TG_HIP(hipSetDevice(c->device)); }

// ---- scratch of tg_solve_noalloc ---------------------------------------------
// One array per GPU for the life of the process.  In host code `scratch`
// names a host-side stand-in, not device memory: its device address on each
// GPU comes from hipGetSymbolAddress (resolved once per device, in
// tg_create, with that device current) and only that address is handed to
// kernels and copies.
static __device__ __align__(256) unsigned char scratch[TG_NOALLOC_SCRATCH_BYTES];
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_dev[kMaxDevices] = {};

static void resolve_scratch(int device)
{
    if (device < 0 || device >= kMaxDevices) fail(TG_OUT_OF_RANGE, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_dev[device]) return;
    void* p = nullptr;
    TG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) fail(TG_GENERAL_FAIL, "no device address for the no-allocation scratch");
    g_scratch_dev[device] = p;
}

static void* scratch_of(int device)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return (device >= 0 && device < kMaxDevices) ? g_scratch_dev[device] : nullptr;
}

static void require_init(const tg_ctx* c)
{
    if (!c->initialized) fail(TG_GENERAL_FAIL, "tg_initialize has not been called");
}

static void build_persist_list(tg_ctx* c);

// Segment tables: slice -> segment, segment -> first slice, element -> segment.
static void setup_segments(tg_ctx* c)
{
    Segments& S = c->seg;
    if (S.node_begin.size() != (size_t)S.nseg + 1 || S.node_begin.back() != c->nnd) {
        S.nseg = 1;
        S.node_begin = {0, c->nnd};
    }
    S.nslices = (c->nnd + 63) / 64;
    std::vector<int> slice_seg((size_t)S.nslices), seg_slice((size_t)S.nseg + 1);
    std::vector<int> node_seg((size_t)c->nnd);
    for (int g = 0; g < S.nseg; g++) {
        seg_slice[g] = S.node_begin[g] / 64;
        for (int a = S.node_begin[g]; a < S.node_begin[g + 1]; a++) node_seg[a] = g;
    }
    seg_slice[S.nseg] = (int)S.nslices;
    for (long s = 0; s < S.nslices; s++) slice_seg[s] = node_seg[64 * s];
    S.elem_seg.resize((size_t)c->nel);
    int straddle = 0;
    parallel_for(c->nel, [&](long b, long end) {
        int bad = 0;
        for (long e = b; e < end; e++) {
            const int g = node_seg[c->conn[8 * e]];
            for (int a = 1; a < 8; a++) bad |= node_seg[c->conn[8 * e + a]] != g;
            S.elem_seg[e] = g;
        }
        if (bad) __atomic_store_n(&straddle, 1, __ATOMIC_RELAXED);
    });
    if (straddle) fail(TG_BAD_INPUT, "element connects two segments");
    S.d_slice_seg.alloc((size_t)S.nslices);
    S.d_seg_slice.alloc((size_t)S.nseg + 1);
    S.d_active.alloc((size_t)S.nseg);
    S.d_elem_seg.alloc((size_t)c->nel);
    S.d_scal.alloc((size_t)S_NSCAL * S.nseg);
    S.d_part.alloc(2 * (size_t)S.nslices);
    S.d_out.alloc((size_t)S.nseg);
    S.d_ndone.alloc(1);
    S.d_segstat.alloc((size_t)S.nseg);
    S.h_scal.assign((size_t)S_NSCAL * S.nseg, 0.0);
    S.d_persist.alloc((size_t)S.nseg);
    std::vector<int> ones((size_t)S.nseg, 1);
    auto up = [&](int* dst, const std::vector<int>& v) {
        TG_HIP(hipMemcpyAsync(dst, v.data(), sizeof(int) * v.size(), hipMemcpyHostToDevice, c->stream));
    };
    up(S.d_slice_seg.p, slice_seg);
    up(S.d_seg_slice.p, seg_slice);
    up(S.d_active.p, ones);
    up(S.d_elem_seg.p, S.elem_seg);
    TG_HIP(hipMemsetAsync(S.d_scal.p, 0, sizeof(double) * S_NSCAL * S.nseg, c->stream));
    TG_HIP(hipMemsetAsync(S.d_segstat.p, 0, sizeof(int) * S.nseg, c->stream));
    // every copy and clear of the context goes through its own stream (a
    // null-stream operation is not ordered with a non-blocking stream), and
    // the host vectors copied from must outlive the copies
    TG_HIP(hipStreamSynchronize(c->stream));
    build_persist_list(c);
}

// Segments of at most pcg_persist(c) slices, for the persistent PCG kernel.
// The device list holds every segment's room (allocated with the segment
// tables), so a rebuild only copies.
static void build_persist_list(tg_ctx* c)
{
    Segments& S = c->seg;
    const int thr = std::min(pcg_persist(c), PCG_PERSIST_MAX);
    S.h_persist.clear();
    if (thr > 0)
        for (int g = 0; g < S.nseg; g++) {
            const long sb = S.node_begin[g] / 64;
            const long se = g + 1 < S.nseg ? S.node_begin[g + 1] / 64 : S.nslices;
            if (se - sb <= thr) S.h_persist.push_back(g);
        }
    if (!S.h_persist.empty()) {
        TG_HIP(hipMemcpyAsync(S.d_persist.p, S.h_persist.data(), sizeof(int) * S.h_persist.size(),
                              hipMemcpyHostToDevice, c->stream));
        TG_HIP(hipStreamSynchronize(c->stream));  // h_persist outlives its copy
    }
    S.persist_built = pcg_persist(c);
}

void drop_pcg_graph(tg_ctx* c)
{
    if (c->pg.exec) (void)hipGraphExecDestroy(c->pg.exec);
    if (c->pg.graph) (void)hipGraphDestroy(c->pg.graph);
    c->pg.exec = nullptr;
    c->pg.graph = nullptr;
    std::memset(c->pg.key, 0, sizeof c->pg.key);
}

static void clear_status(tg_ctx* c)
{
    TG_HIP(hipMemsetAsync(c->dstatus.p, 0, sizeof(int), c->stream));
    if (c->seg.d_segstat.p) TG_HIP(hipMemsetAsync(c->seg.d_segstat.p, 0, sizeof(int) * c->seg.nseg, c->stream));
    TG_HIP(hipStreamSynchronize(c->stream));
}

static void check_status(tg_ctx* c)
{
    TG_HIP(hipMemcpyAsync(c->hstatus, c->dstatus.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    TG_HIP(hipStreamSynchronize(c->stream));
    if (c->hstatus[0] != 0) {
        const int code = c->hstatus[0];
        clear_status(c);
        fail(code, "zero or negative jacobian");
    }
}

// After a residual pass of a batch: the segments whose elements inverted
// (per-segment flags of the residual kernels) get status 7 in bad[g]; the
// flags are cleared.  Other failures still raise.
static void check_status_segments(tg_ctx* c, std::vector<int>& bad)
{
    TG_HIP(hipMemcpyAsync(c->hstatus, c->dstatus.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    TG_HIP(hipStreamSynchronize(c->stream));
    bad.assign((size_t)c->seg.nseg, 0);
    if (c->hstatus[0] == 0) return;
    if (c->hstatus[0] != TG_BAD_JACOBIAN) check_status(c);  // raises
    TG_HIP(hipMemcpyAsync(bad.data(), c->seg.d_segstat.p, sizeof(int) * bad.size(), hipMemcpyDeviceToHost,
                          c->stream));
    clear_status(c);
}

SetupTrace::SetupTrace(tg_ctx* ctx, const char* name) : c(ctx), who(name)
{
    static const bool trace = std::getenv("TG_SETUP_TRACE") != nullptr;
    on = trace;
    t = std::chrono::steady_clock::now();
}

void SetupTrace::operator()(const char* what)
{
    if (!on) return;
    TG_HIP(hipStreamSynchronize(c->stream));
    const auto now = std::chrono::steady_clock::now();
    std::fprintf(stderr, "[%s] %-26s %8.1f ms (nel %d)\n", who, what,
                 std::chrono::duration<double, std::milli>(now - t).count(), c->nel);
    t = now;
}

}  // namespace tg

using namespace tg;

template <class F>
static int guarded(tg_ctx* c, F f)
{
    try {
        if (c) set_device(c);
        f();
        return TG_OK;
    } catch (const Error& e) {
        if (c) c->err = e.msg;
        return e.code;
    } catch (const std::bad_alloc&) {
        if (c) c->err = "host allocation failed";
        return TG_OUT_OF_MEMORY;
    } catch (const std::exception& e) {
        if (c) c->err = e.what();
        return TG_GENERAL_FAIL;
    }
}

extern "C" {

int tg_create(tg_ctx** out, int device)
{
    if (!out) return TG_BAD_INPUT;
    *out = nullptr;
    tg_ctx* c = new (std::nothrow) tg_ctx();
    if (!c) return TG_OUT_OF_MEMORY;
    c->device = device;
    int rc = guarded(c, [&] {
        int n = 0;
        TG_HIP(hipGetDeviceCount(&n));
        if (device < 0 || device >= n) fail(TG_OUT_OF_RANGE, "no such device");
        TG_HIP(hipSetDevice(device));
        resolve_scratch(device);
        TG_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        TG_HIP(hipHostMalloc((void**)&c->hscal, sizeof(double) * 16, hipHostMallocDefault));
        TG_HIP(hipHostMalloc((void**)&c->hstatus, sizeof(int) * 4, hipHostMallocDefault));
        c->dstatus.alloc(4);
        TG_HIP(hipMemsetAsync(c->dstatus.p, 0, sizeof(int) * 4, c->stream));
        c->dscal.alloc(16);
        TG_HIP(hipMemsetAsync(c->dscal.p, 0, sizeof(double) * 16, c->stream));
        // the PCG loop polls its done count through these two events
        for (hipEvent_t& e : c->poll_ev) TG_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        TG_HIP(hipStreamSynchronize(c->stream));
    });
    if (rc != TG_OK) {
        tg_destroy(c);
        return rc;
    }
    *out = c;
    return TG_OK;
}

int tg_destroy(tg_ctx* c)
{
    if (!c) return TG_OK;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    {
        std::lock_guard<std::mutex> lk(g_ev_mu);
        auto it = g_events.find(c);
        if (it != g_events.end()) {
            for (auto& p : it->second) {
                (void)hipEventDestroy(p.e0);
                (void)hipEventDestroy(p.e1);
            }
            g_events.erase(it);
        }
    }
    drop_pcg_graph(c);
    for (hipEvent_t e : c->poll_ev)
        if (e) (void)hipEventDestroy(e);
    if (c->hscal) (void)hipHostFree(c->hscal);
    if (c->hstatus) (void)hipHostFree(c->hstatus);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;  // DBuf destructors free device memory
    return TG_OK;
}

const char* tg_last_error(const tg_ctx* c) { return c ? c->err.c_str() : "null handle"; }

int tg_set_mesh(tg_ctx* c, int nnd, int nel, const double* X, const int* conn, const int* mat)
{
    if (!c) return TG_BAD_INPUT;
    return guarded(c, [&] {
        if (nnd <= 0 || nel <= 0 || !X || !conn) fail(TG_BAD_INPUT, "empty mesh");
        // per element: 1 missing node, 2 repeated node, 4 negative material
        int found = 0;
        parallel_for(nel, [&](long b, long end) {
            int f = 0;
            for (long e = b; e < end; e++) {
                const int* n = conn + 8 * e;
                for (int a = 0; a < 8; a++) {
                    f |= (n[a] < 0 || n[a] >= nnd) ? 1 : 0;
                    for (int k = a + 1; k < 8; k++) f |= n[a] == n[k] ? 2 : 0;
                }
                if (mat) f |= mat[e] < 0 ? 4 : 0;
            }
            if (f) __atomic_fetch_or(&found, f, __ATOMIC_RELAXED);
        });
        if (found & 1) fail(TG_OUT_OF_RANGE, "connectivity refers to a missing node");
        if (found & 2) fail(TG_BAD_INPUT, "degenerate element (repeated node) is not supported");
        if (found & 4) fail(TG_BAD_INPUT, "negative material id");
        // every check above runs before the context changes: a rejected mesh
        // leaves the previous one (or none) in place
        c->nnd = nnd;
        c->nel = nel;
        c->seg.nseg = 1;
        c->seg.node_begin = {0, nnd};
        parallel_copy(c->X, X, 3L * nnd);
        parallel_copy(c->conn, conn, 8L * nel);
        if (mat)
            parallel_copy(c->mat, mat, nel);
        else
            c->mat.assign((size_t)nel, 0);
        c->initialized = false;
    });
}

int tg_set_material(tg_ctx* c, int id, double mu, double kappa)
{
    if (!c) return TG_BAD_INPUT;
    return guarded(c, [&] {
        if (id < 0 || id > 1023) fail(TG_OUT_OF_RANGE, "material id out of range");
        if (!(mu >= 0.0) || !(kappa >= 0.0)) fail(TG_GENERAL_FAIL, "negative modulus");  // IsotropicT.cpp:51
        if ((int)c->mu.size() <= id) {
            c->mu.resize(id + 1, -1.0);
            c->kappa.resize(id + 1, -1.0);
        }
        c->mu[id] = mu;
        c->kappa[id] = kappa;
        c->initialized = false;
    });
}

int tg_set_formulation(tg_ctx* c, int f)
{
    if (!c) return TG_BAD_INPUT;
    return guarded(c, [&] {
        if (f != TG_TOTAL_LAGRANGIAN && f != TG_Q1P0) fail(TG_BAD_INPUT, "unknown formulation");
        c->formulation = f;
        c->initialized = false;
    });
}

int tg_set_material_model(tg_ctx* c, int model)
{
    if (!c) return TG_BAD_INPUT;
    return guarded(c, [&] {
        if (model != TG_MAT_SIMO_ISOTROPIC && model != TG_MAT_STVENANT) fail(TG_BAD_INPUT, "unknown material model");
        c->material_model = model;
    });
}

int tg_set_precision(tg_ctx* c, int prec)
{
    if (!c) return TG_BAD_INPUT;
    return guarded(c, [&] {
        if (prec != TG_FP64 && prec != TG_MIXED) fail(TG_BAD_INPUT, "unknown precision");
        c->precision = prec;
        c->initialized = false;
    });
}

int tg_set_segments(tg_ctx* c, int nseg, const int* nb)
{
    if (!c) return TG_BAD_INPUT;
    return guarded(c, [&] {
        if (nseg < 1 || !nb) fail(TG_BAD_INPUT, "bad segment list");
        if (nb[0] != 0 || nb[nseg] != c->nnd) fail(TG_BAD_INPUT, "segments must cover the nodes");
        for (int g = 1; g <= nseg; g++) {
            if (nb[g] <= nb[g - 1]) fail(TG_BAD_INPUT, "segment node ranges must increase");
            if (g < nseg && nb[g] % 64 != 0) fail(TG_BAD_INPUT, "segment begins must be multiples of 64 nodes");
        }
        c->seg.node_begin.assign(nb, nb + nseg + 1);
        c->seg.nseg = nseg;
        c->initialized = false;
    });
}

int tg_set_preconditioner(tg_ctx* c, int pc)
{
    if (!c) return TG_BAD_INPUT;
    return guarded(c, [&] {
        if (pc != TG_PC_JACOBI && pc != TG_PC_BLOCK_JACOBI && pc != TG_PC_AMG) fail(TG_BAD_INPUT, "unknown preconditioner");
        if (pc == TG_PC_AMG && c->seg.nseg != 1)
            fail(TG_BAD_INPUT, "the multigrid preconditioner needs a single problem (no segments)");
        c->precond = pc;
        if (c->initialized) {
            c->vbinv.alloc(pc == TG_PC_BLOCK_JACOBI ? 9 * (size_t)c->nnd : 0);
            if (pc == TG_PC_AMG && !c->op.d_val32.p) c->op.d_val32.alloc((size_t)c->op.nblocks_padded * 9);
            if (pc == TG_PC_AMG && c->amg.lv.empty()) amg_setup(c);
        }
    });
}

int tg_set_linear_solver(tg_ctx* c, int solver, int restart)
{
    if (!c) return TG_BAD_INPUT;
    return guarded(c, [&] {
        if (solver != TG_LIN_PCG && solver != TG_LIN_GMRES) fail(TG_BAD_INPUT, "unknown linear solver");
        if (solver == TG_LIN_GMRES && (restart < 1 || restart > 1000))
            fail(TG_OUT_OF_RANGE, "GMRES restart must be 1..1000");
        if (solver == TG_LIN_GMRES && c->seg.nseg != 1)
            fail(TG_BAD_INPUT, "GMRES needs a single problem (no segments)");
        c->linear_solver = solver;
        if (solver == TG_LIN_GMRES) c->gmres_restart = restart;
    });
}

int tg_set_kbc(tg_ctx* c, int nkbc, const int* nodes, const int* dofs, const double* values)
{
    if (!c) return TG_BAD_INPUT;
    return guarded(c, [&] {
        if (nkbc < 0 || (nkbc > 0 && (!nodes || !dofs || !values))) fail(TG_BAD_INPUT, "bad KBC arrays");
        for (int k = 0; k < nkbc; k++) {
            if (nodes[k] < 0 || nodes[k] >= c->nnd) fail(TG_OUT_OF_RANGE, "KBC node out of range");
            if (dofs[k] < 0 || dofs[k] > 2) fail(TG_OUT_OF_RANGE, "KBC dof out of range");
        }
        c->kbc_nodes.assign(nodes, nodes + nkbc);
        c->kbc_dofs.assign(dofs, dofs + nkbc);
        c->kbc_vals.assign(values, values + nkbc);
        c->initialized = false;
    });
}

// Device storage of the mesh and the vectors, and the uploads that need
// neither the node graph nor the colouring (tg_initialize colours the
// elements on another host thread meanwhile).
static void upload_mesh(tg_ctx* c, int nmat)
{
    // device storage
    const size_t nn = 3 * (size_t)c->nnd;
    c->dX.alloc(nn);
    c->du.alloc(nn);
    c->du_last.alloc(nn);
    c->dR.alloc(nn);
    c->dconn.alloc(8 * (size_t)c->nel);
    c->dmat.alloc((size_t)c->nel);
    c->deqn.alloc(nn);
    c->dcolor_elems.alloc((size_t)c->nel);
    c->dmatprops.alloc((size_t)std::max(1, nmat));
    c->dslot.alloc(300 * (size_t)c->nel);
    for (auto* v : {&c->vx, &c->vr, &c->vz, &c->vp, &c->vq, &c->vdinv, &c->vwork}) v->alloc(nn);
    const size_t sblocks = ((size_t)c->nnd + 255) / 256;
    c->dpart.alloc(2 * std::max<size_t>(sblocks, 1024) + 16);
    std::vector<int> eqn0(nn);
    for (size_t k = 0; k < nn; k++) eqn0[k] = c->eqnos[k] > 0 ? c->eqnos[k] - 1 : -1;
    std::vector<double2> props((size_t)std::max(1, nmat));
    for (int m = 0; m < nmat; m++) props[m] = make_double2(c->mu[m], c->kappa[m]);
    std::vector<int> kidx(c->kbc_nodes.size());
    for (size_t k = 0; k < kidx.size(); k++) kidx[k] = 3 * c->kbc_nodes[k] + c->kbc_dofs[k];
    c->dkbc_idx.alloc(kidx.size());
    c->dkbc_val.alloc(kidx.size());
    TG_HIP(hipMemcpyAsync(c->dX.p, c->X.data(), sizeof(double) * nn, hipMemcpyHostToDevice, c->stream));
    TG_HIP(hipMemsetAsync(c->du.p, 0, sizeof(double) * nn, c->stream));
    TG_HIP(hipMemsetAsync(c->du_last.p, 0, sizeof(double) * nn, c->stream));
    TG_HIP(hipMemcpyAsync(c->dconn.p, c->conn.data(), sizeof(int) * c->conn.size(), hipMemcpyHostToDevice,
                          c->stream));
    TG_HIP(hipMemcpyAsync(c->dmat.p, c->mat.data(), sizeof(int) * c->mat.size(), hipMemcpyHostToDevice, c->stream));
    TG_HIP(hipMemcpyAsync(c->deqn.p, eqn0.data(), sizeof(int) * nn, hipMemcpyHostToDevice, c->stream));
    TG_HIP(hipMemcpyAsync(c->dmatprops.p, props.data(), sizeof(double2) * props.size(), hipMemcpyHostToDevice,
                          c->stream));
    if (!kidx.empty()) {
        TG_HIP(hipMemcpyAsync(c->dkbc_idx.p, kidx.data(), sizeof(int) * kidx.size(), hipMemcpyHostToDevice,
                              c->stream));
        TG_HIP(hipMemcpyAsync(c->dkbc_val.p, c->kbc_vals.data(), sizeof(double) * kidx.size(),
                              hipMemcpyHostToDevice, c->stream));
    }
    TG_HIP(hipStreamSynchronize(c->stream));  // the host sources above are locals
}

int tg_initialize(tg_ctx* c, int* n_eq, long* msr_len)
{
    if (!c) return TG_BAD_INPUT;
    return guarded(c, [&] {
        if (c->nnd == 0) fail(TG_GENERAL_FAIL, "no mesh");
        int nmat = 0;
        for (int m : c->mat) nmat = std::max(nmat, m + 1);
        if ((int)c->mu.size() < nmat) fail(TG_BAD_INPUT, "element refers to an undefined material");
        for (int m = 0; m < nmat; m++)
            if (c->mu[m] < 0.0) fail(TG_BAD_INPUT, "element refers to an undefined material");
        for (int k : c->fbc_idx)
            if (k >= 3 * c->nnd) fail(TG_OUT_OF_RANGE, "force card out of range");
        // kinematic cards were checked against the mesh present when they
        // were set; a mesh set afterwards may have fewer nodes
        for (int kn : c->kbc_nodes)
            if (kn < 0 || kn >= c->nnd) fail(TG_OUT_OF_RANGE, "KBC node out of range");
        c->initialized = false;
        // a captured PCG chunk refers to buffers that are about to be replaced
        drop_pcg_graph(c);
        c->dfext.release();
        SetupTrace phase(c, "tg_initialize");
        // equations; then the element colouring on a host thread of its own
        // while this thread uploads the mesh and builds the structure on the
        // device
        c->n_eq = equation_numbers(c->nnd, c->kbc_nodes, c->kbc_dofs, c->eqnos);
        phase("equation numbers");
        c->stiff_chunk_built = -2;
        c->gather_built = -2;
        c->dgptr.release();
        c->dglist.release();
        c->debuf.release();
        c->debuf32.release();
        {
            std::exception_ptr colour_err, main_err;
            using clk = std::chrono::steady_clock;
            const auto t0 = clk::now();
            clk::time_point t_colour, t_upload, t_graph;
            std::thread colour([&] {
                try {
                    c->ncolors = color_elements(c->nnd, c->nel, c->conn.data(), c->color_ptr, c->color_elems);
                } catch (...) {
                    colour_err = std::current_exception();
                }
                t_colour = clk::now();
            });
            try {
                upload_mesh(c, nmat);
                t_upload = clk::now();
                device_node_graph(c);
                device_structure(c);
                TG_HIP(hipStreamSynchronize(c->stream));
                t_graph = clk::now();
            } catch (...) {
                main_err = std::current_exception();
            }
            colour.join();
            for (auto& e : {main_err, colour_err})
                if (e) std::rethrow_exception(e);
            if (phase.on) {
                auto ms = [&](clk::time_point t) { return std::chrono::duration<double, std::milli>(t - t0).count(); };
                std::fprintf(stderr, "[tg_initialize]   upload %.1f ms, + device structure %.1f ms; colouring %.1f ms\n",
                             ms(t_upload), ms(t_graph), ms(t_colour));
            }
        }
        TG_HIP(hipMemcpyAsync(c->dcolor_elems.p, c->color_elems.data(), sizeof(int) * c->nel, hipMemcpyHostToDevice,
                              c->stream));
        const bool mixed = mixed_operator(c);
        phase("structure | colours");
        c->op.d_diag.alloc((size_t)c->nnd);
        launch_build_operator_diag(c);
        if (c->formulation == TG_TOTAL_LAGRANGIAN)
            c->op.d_oslot.alloc(72 * (size_t)c->nel);
        else
            c->op.d_oslot.release();
        launch_element_slots(c);
        c->vbinv.alloc(c->precond == TG_PC_BLOCK_JACOBI ? 9 * (size_t)c->nnd : 0);
        c->op.d_val32.alloc(mixed || c->precond == TG_PC_AMG ? (size_t)c->op.nblocks_padded * 9 : 0);
        phase("slot maps");
        if (c->precond == TG_PC_AMG) {
            amg_setup(c);
            phase("multigrid structure");
        } else {
            c->amg.clear();
        }
        setup_segments(c);
        TG_HIP(hipStreamSynchronize(c->stream));
        phase("segments");
        c->initialized = true;
        if (n_eq) *n_eq = c->n_eq;
        if (msr_len) *msr_len = c->msr_len;
    });
}

int tg_get_eqnos(const tg_ctx* c, int* eqnos)
{
    if (!c || !eqnos) return TG_BAD_INPUT;
    if (!c->initialized) return TG_GENERAL_FAIL;
    std::memcpy(eqnos, c->eqnos.data(), sizeof(int) * c->eqnos.size());
    return TG_OK;
}

int tg_get_msr_bindx(const tg_ctx* c, int* bindx)
{
    if (!c || !bindx) return TG_BAD_INPUT;
    if (!c->initialized) return TG_GENERAL_FAIL;
    // the pattern lives on the device (its columns are written there)
    tg_ctx* m = const_cast<tg_ctx*>(c);
    return guarded(m, [&] {
        TG_HIP(hipMemcpyAsync(bindx, m->dbindx.p, sizeof(int) * (size_t)m->msr_len, hipMemcpyDeviceToHost, m->stream));
        TG_HIP(hipStreamSynchronize(m->stream));
    });
}

int tg_set_displacement(tg_ctx* c, const double* u)
{
    if (!c || !u) return TG_BAD_INPUT;
    return guarded(c, [&] {
        require_init(c);
        TG_HIP(hipMemcpyAsync(c->du.p, u, sizeof(double) * 3 * (size_t)c->nnd, hipMemcpyHostToDevice, c->stream));
        TG_HIP(hipStreamSynchronize(c->stream));
        c->amg.since = -1;  // a new state: the next tangent sets the coarse levels up
    });
}

int tg_get_displacement(tg_ctx* c, double* u)
{
    if (!c || !u) return TG_BAD_INPUT;
    return guarded(c, [&] {
        require_init(c);
        TG_HIP(hipMemcpyAsync(u, c->du.p, sizeof(double) * 3 * (size_t)c->nnd, hipMemcpyDeviceToHost, c->stream));
        TG_HIP(hipStreamSynchronize(c->stream));
    });
}

int tg_set_forces(tg_ctx* c, int n, const int* nodes, const int* dofs, const double* values)
{
    if (!c || n < 0 || (n > 0 && (!nodes || !dofs || !values))) return TG_BAD_INPUT;
    return guarded(c, [&] {
        // validate every card before the context changes: a rejected list
        // leaves the previous cards in place
        for (int k = 0; k < n; k++)
            if (nodes[k] < 0 || nodes[k] >= c->nnd || dofs[k] < 0 || dofs[k] > 2)
                fail(TG_OUT_OF_RANGE, "force card out of range");
        c->fbc_idx.resize((size_t)n);
        for (int k = 0; k < n; k++) c->fbc_idx[k] = 3 * nodes[k] + dofs[k];
        c->fbc_vals.assign(values, values + n);
        c->dfext.release();
    });
}

int tg_apply_forces(tg_ctx* c, double scale)
{
    if (!c) return TG_BAD_INPUT;
    return guarded(c, [&] {
        require_init(c);
        if (c->fbc_idx.empty()) {
            c->dfext.release();
            return;
        }
        std::vector<double> f(3 * (size_t)c->nnd, 0.0);
        for (size_t k = 0; k < c->fbc_idx.size(); k++) f[c->fbc_idx[k]] += c->fbc_vals[k] * scale;
        if (!c->dfext.p) c->dfext.alloc(f.size());
        TG_HIP(hipMemcpyAsync(c->dfext.p, f.data(), sizeof(double) * f.size(), hipMemcpyHostToDevice, c->stream));
        TG_HIP(hipStreamSynchronize(c->stream));  // f outlives its copy
    });
}

int tg_apply_kbc(tg_ctx* c, double t)
{
    if (!c) return TG_BAD_INPUT;
    return guarded(c, [&] {
        require_init(c);
        launch_apply_kbc(c, t);
        c->amg.since = -1;  // a new load: the next tangent sets the coarse levels up
    });
}

// FieldT::CloseStep / ResetStep keep and restore the last converged
// displacement (tahoe/src/nodes/FieldT.cpp:469-472, 618); SimoQ1P0 additionally keeps the
// element volumes (SimoQ1P0.cpp:42-61), which the Simo material never reads
// (it needs no "last" deformation gradient), so the displacement is the
// whole state here.
int tg_close_step(tg_ctx* c)
{
    if (!c) return TG_BAD_INPUT;
    return guarded(c, [&] {
        require_init(c);
        TG_HIP(hipMemcpyAsync(c->du_last.p, c->du.p, sizeof(double) * 3 * (size_t)c->nnd, hipMemcpyDeviceToDevice,
                              c->stream));
        c->amg.since = -1;
    });
}

int tg_reset_step(tg_ctx* c)
{
    if (!c) return TG_BAD_INPUT;
    return guarded(c, [&] {
        require_init(c);
        TG_HIP(hipMemcpyAsync(c->du.p, c->du_last.p, sizeof(double) * 3 * (size_t)c->nnd, hipMemcpyDeviceToDevice,
                              c->stream));
        c->amg.since = -1;  // back to the last converged state (a failed step): no coarse levels of its tangents
    });
}

int tg_form_rhs(tg_ctx* c, double* R_out, double* norm_out)
{
    if (!c) return TG_BAD_INPUT;
    return guarded(c, [&] {
        require_init(c);
        launch_residual(c, c->du.p, c->dR.p);
        if (norm_out || R_out) check_status(c);
        if (norm_out) *norm_out = std::sqrt(device_dot(c, c->dR.p, c->dR.p, 3L * c->nnd));
        if (R_out) {
            node_to_eq(c, c->dR.p, c->vwork.p);
            TG_HIP(hipMemcpyAsync(R_out, c->vwork.p, sizeof(double) * c->n_eq, hipMemcpyDeviceToHost, c->stream));
            TG_HIP(hipStreamSynchronize(c->stream));
        }
    });
}

int tg_form_lhs(tg_ctx* c)
{
    if (!c) return TG_BAD_INPUT;
    return guarded(c, [&] {
        require_init(c);
        launch_stiffness(c, c->du.p);
    });
}

// FP64 copy of the assembled values: in mixed precision the total-Lagrangian
// stiffness is assembled in FP32 and widened here for the GlobalMatrixT
// queries.
static void fp64_values(tg_ctx* c)
{
    if (mixed_stiffness(c)) widen_values(c);
}

// the preconditioner's data for the current operator values
static void prepare_preconditioner(tg_ctx* c)
{
    launch_block_inverse(c);
    if (c->precond == TG_PC_AMG) {
        if (!mixed_operator(c)) launch_narrow_operator(c);
        amg_numeric(c);
    }
}

// operator of the inner linear solve, in the solve's precision: the FP32
// total-Lagrangian stiffness is expanded as it is; the FP64 Q1P0 stiffness
// is expanded (1/diag from the FP64 diagonal) and then narrowed
static void prepare_operator(tg_ctx* c)
{
    if (mixed_stiffness(c)) {
        launch_expand32(c);
    } else {
        launch_expand(c);
        if (mixed_operator(c)) launch_narrow_operator(c);
    }
    prepare_preconditioner(c);
}

// Newton's tangent in the inner solve's operator: straight into the block
// operator (total-Lagrangian), or assembled into the MSR and expanded.
static void newton_operator(tg_ctx* c)
{
    if (c->formulation == TG_TOTAL_LAGRANGIAN && c->newton_assembly == 1 && c->op.d_oslot.p) {
        launch_stiffness_operator(c, c->du.p);
        launch_operator_dinv(c);
        prepare_preconditioner(c);
    } else {
        launch_stiffness(c, c->du.p);
        prepare_operator(c);
    }
}

int tg_form_tangent(tg_ctx* c)
{
    if (!c) return TG_BAD_INPUT;
    return guarded(c, [&] {
        require_init(c);
        newton_operator(c);
    });
}

static void solve_linear(tg_ctx* c, const double* b, double* x, double rtol, int maxit, tg_lin_stats* st)
{
    if (c->linear_solver == TG_LIN_GMRES) {
        gmres_solve(c, b, x, rtol, maxit, st);
        return;
    }
    if (mixed_operator(c))
        pcg_solve_mixed(c, b, x, rtol, maxit, st);
    else
        pcg_solve(c, b, x, rtol, maxit, st);
}

int tg_get_msr_values(tg_ctx* c, double* val)
{
    if (!c || !val) return TG_BAD_INPUT;
    return guarded(c, [&] {
        require_init(c);
        check_status(c);
        fp64_values(c);
        TG_HIP(hipMemcpyAsync(val, c->dval.p, sizeof(double) * c->msr_len, hipMemcpyDeviceToHost, c->stream));
        TG_HIP(hipStreamSynchronize(c->stream));
    });
}

int tg_multx(tg_ctx* c, const double* x, double* b)
{
    if (!c || !x || !b) return TG_BAD_INPUT;
    return guarded(c, [&] {
        require_init(c);
        check_status(c);
        fp64_values(c);
        TG_HIP(hipMemcpyAsync(c->vwork.p, x, sizeof(double) * c->n_eq, hipMemcpyHostToDevice, c->stream));
        eq_to_node(c, c->vwork.p, c->vx.p);
        launch_expand(c);
        launch_spmv(c, c->vx.p, c->vq.p);
        node_to_eq(c, c->vq.p, c->vwork.p);
        TG_HIP(hipMemcpyAsync(b, c->vwork.p, sizeof(double) * c->n_eq, hipMemcpyDeviceToHost, c->stream));
        TG_HIP(hipStreamSynchronize(c->stream));
    });
}

int tg_copy_diagonal(tg_ctx* c, double* diag)
{
    if (!c || !diag) return TG_BAD_INPUT;
    return guarded(c, [&] {
        require_init(c);
        fp64_values(c);
        TG_HIP(hipMemcpyAsync(diag, c->dval.p, sizeof(double) * c->n_eq, hipMemcpyDeviceToHost, c->stream));
        TG_HIP(hipStreamSynchronize(c->stream));
    });
}

int tg_solve(tg_ctx* c, double* rhs, double rtol, int maxit, tg_lin_stats* st)
{
    if (!c || !rhs) return TG_BAD_INPUT;
    return guarded(c, [&] {
        require_init(c);
        check_status(c);
        TG_HIP(hipMemcpyAsync(c->vwork.p, rhs, sizeof(double) * c->n_eq, hipMemcpyHostToDevice, c->stream));
        eq_to_node(c, c->vwork.p, c->vz.p);
        prepare_operator(c);
        solve_linear(c, c->vz.p, c->vx.p, rtol, maxit, st);
        node_to_eq(c, c->vx.p, c->vwork.p);
        TG_HIP(hipMemcpyAsync(rhs, c->vwork.p, sizeof(double) * c->n_eq, hipMemcpyDeviceToHost, c->stream));
        TG_HIP(hipStreamSynchronize(c->stream));
    });
}

int tg_solve_noalloc(tg_ctx* c, const double* b_dev, double* x_dev, double rtol, int maxit, tg_lin_stats* st)
{
    if (!c || !b_dev || !x_dev) return TG_BAD_INPUT;
    return guarded(c, [&] {
        require_init(c);
        if (c->precond == TG_PC_AMG)
            fail(TG_BAD_INPUT, "tg_solve_noalloc: the multigrid preconditioner keeps its own buffers");
        if (c->linear_solver != TG_LIN_PCG) fail(TG_BAD_INPUT, "tg_solve_noalloc: PCG only");
        double* base = (double*)scratch_of(c->device);
        if (!base) fail(TG_GENERAL_FAIL, "tg_solve_noalloc: scratch not resolved for this device");
        // layout (doubles): b, x, r, p, q, p2 [3 nnd each], part [2 nslices],
        // scal [S_NSCAL nseg], then the done count (int)
        const size_t n3 = 3 * (size_t)c->nnd;
        const size_t ns = (size_t)c->seg.nslices, nseg = (size_t)c->seg.nseg;
        const size_t ndoubles = 6 * n3 + 2 * ns + (size_t)S_NSCAL * nseg;
        const size_t need = ndoubles * sizeof(double) + sizeof(int);
        if (need > (size_t)TG_NOALLOC_SCRATCH_BYTES)
            fail(TG_SIZE_MISMATCH, "tg_solve_noalloc: the problem needs " + std::to_string(need) +
                                       " bytes of scratch, more than the " +
                                       std::to_string((long)TG_NOALLOC_SCRATCH_BYTES) + " there are");
        double* vb = base;
        double* vx = vb + n3;
        double* vr = vx + n3;
        double* vp = vr + n3;
        double* vq = vp + n3;
        double* vp2 = vq + n3;
        double* part = vp2 + n3;
        double* scal = part + 2 * ns;
        int* ndone = (int*)(scal + (size_t)S_NSCAL * nseg);
        // no event timing inside (event creation is an allocation)
        struct TimingOff {
            tg_ctx* c;
            bool was;
            ~TimingOff() { c->timing.enabled = was; }
        } off{c, c->timing.enabled};
        c->timing.enabled = false;
        check_status(c);
        eq_to_node(c, b_dev, vb);
        prepare_operator(c);
        // every buffer the run reads before writing is set by the run itself
        // (x, r, p, partials and scalars by the init kernels, the done count
        // by a clear), so what an earlier call left does not matter
        const PcgWork W{part, scal, ndone, vp2, false};
        if (mixed_operator(c))
            pcg_run<float>(c, c->op.d_val32.p, c->vdinv.p, vb, vx, vr, vp, vq, rtol, maxit, c->seg.d_active.p,
                           c->seg.nseg, st, &W);
        else
            pcg_run<double>(c, c->op.d_val.p, c->vdinv.p, vb, vx, vr, vp, vq, rtol, maxit, c->seg.d_active.p,
                            c->seg.nseg, st, &W);
        check_status(c);
        node_to_eq(c, vx, x_dev);
        TG_HIP(hipStreamSynchronize(c->stream));
    });
}

// One correction: R(u) -> ||R||, K(u), PCG, u += du.  Returns ||R|| before
// the update in *err_before; leaves R(u) of the *old* u in dR.
static void correction(tg_ctx* c, const tg_newton_opts* o, double* err_before, long* lin_its)
{
    launch_residual(c, c->du.p, c->dR.p);
    check_status(c);
    *err_before = std::sqrt(device_dot(c, c->dR.p, c->dR.p, 3L * c->nnd));
    newton_operator(c);
    tg_lin_stats st{};
    solve_linear(c, c->dR.p, c->vx.p, o->lin_rtol, o->lin_maxit, &st);
    check_status(c);
    *lin_its += st.iterations;
    launch_axpy(c, c->du.p, c->vx.p, 1.0, 3L * c->nnd);
}

int tg_newton_correction(tg_ctx* c, const tg_newton_opts* o, tg_newton_stats* s)
{
    if (!c || !o) return TG_BAD_INPUT;
    return guarded(c, [&] {
        require_init(c);
        double e0 = 0.0;
        long lin = 0;
        correction(c, o, &e0, &lin);
        TG_HIP(hipStreamSynchronize(c->stream));
        if (s) {
            s->status = 0;
            s->iterations = 1;
            s->error0 = e0;
            s->error = e0;
            s->lin_iterations = lin;
        }
    });
}

// NLSolver::Solve (NLSolver.cpp:57-264) with ExitIteration (NLSolver.cpp:675-757):
// error0 = ||R(u0)||; converged at once when error0 < abs_tolerance; after
// each correction fail when error/error0 > divergence_tolerance, converge when
// error/error0 < rel_tolerance or error < abs_tolerance, fail when the
// correction count reaches max_iterations.
int tg_newton(tg_ctx* c, const tg_newton_opts* o, tg_newton_stats* s)
{
    if (!c || !o) return TG_BAD_INPUT;
    tg_newton_stats out{};
    int rc = guarded(c, [&] {
        require_init(c);
        launch_residual(c, c->du.p, c->dR.p);
        check_status(c);
        const double err0 = std::sqrt(device_dot(c, c->dR.p, c->dR.p, 3L * c->nnd));
        out.error0 = out.error = err0;
        if (err0 < o->abs_tolerance) return;
        int iter = -1;
        long lin = 0;
        for (;;) {
            newton_operator(c);
            tg_lin_stats st{};
            solve_linear(c, c->dR.p, c->vx.p, o->lin_rtol, o->lin_maxit, &st);
            check_status(c);
            lin += st.iterations;
            launch_axpy(c, c->du.p, c->vx.p, 1.0, 3L * c->nnd);
            iter++;
            launch_residual(c, c->du.p, c->dR.p);
            check_status(c);
            const double err = std::sqrt(device_dot(c, c->dR.p, c->dR.p, 3L * c->nnd));
            out.error = err;
            out.iterations = iter + 1;
            out.lin_iterations = lin;
            const double rel = err / err0;
            if (rel > o->divergence_tolerance) {
                out.status = 1;
                return;
            }
            if (rel < o->rel_tolerance || err < o->abs_tolerance) return;
            if (iter >= o->max_iterations) {
                out.status = 1;
                return;
            }
        }
    });
    if (rc == TG_BAD_JACOBIAN) out.status = TG_BAD_JACOBIAN;
    if (s) *s = out;
    return rc;
}

// Per-segment NLSolver::Solve: the tests of ExitIteration (NLSolver.cpp:675-757)
// applied to each segment's own ||R||; segments that converge or fail leave
// the active set (their PCG does nothing, so their u stays put).
int tg_newton_batch(tg_ctx* c, const tg_newton_opts* o, tg_newton_stats* stats)
{
    if (!c || !o || !stats) return TG_BAD_INPUT;
    return guarded(c, [&] {
        require_init(c);
        const int G = c->seg.nseg;
        std::vector<double> e0(G), e(G);
        std::vector<int> state(G, 0), iter(G, -1);  // 0 active, 1 converged, 2 failed
        std::vector<int> code(G, 0), bad;           // per segment: 7 once an element inverted
        std::vector<long> lin(G, 0);
        // a failed segment goes back to its last converged displacement, so
        // the element passes the rest still run over its elements see no
        // inverted element (its result is "failed" either way)
        auto restore = [&](int g) {
            const size_t b = 3 * (size_t)c->seg.node_begin[g], n = 3 * (size_t)(c->seg.node_begin[g + 1] - c->seg.node_begin[g]);
            TG_HIP(hipMemcpyAsync(c->du.p + b, c->du_last.p + b, sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream));
        };
        launch_residual(c, c->du.p, c->dR.p);
        check_status_segments(c, bad);
        segment_dots(c, c->dR.p, c->dR.p, e0.data());
        for (int g = 0; g < G; g++) {
            e0[g] = std::sqrt(e0[g]);
            e[g] = e0[g];
            if (bad[g]) {
                state[g] = 2;
                code[g] = TG_BAD_JACOBIAN;
                restore(g);
            } else if (e0[g] < o->abs_tolerance) {
                state[g] = 1;
            }
        }
        std::vector<int> active(G);
        for (;;) {
            int nact = 0;
            for (int g = 0; g < G; g++) {
                active[g] = state[g] == 0;
                nact += active[g];
            }
            if (nact == 0) break;
            TG_HIP(hipMemcpyAsync(c->seg.d_active.p, active.data(), sizeof(int) * G, hipMemcpyHostToDevice, c->stream));
            newton_operator(c);
            tg_lin_stats st{};
            solve_linear(c, c->dR.p, c->vx.p, o->lin_rtol, o->lin_maxit, &st);
            check_status(c);
            for (int g = 0; g < G; g++)
                if (active[g]) lin[g] += (long)c->seg.h_scal[(size_t)S_ITER * G + g];
            launch_axpy(c, c->du.p, c->vx.p, 1.0, 3L * c->nnd);
            launch_residual(c, c->du.p, c->dR.p);
            check_status_segments(c, bad);
            segment_dots(c, c->dR.p, c->dR.p, e.data());
            for (int g = 0; g < G; g++) {
                e[g] = std::sqrt(e[g]);
                if (!active[g]) continue;
                iter[g]++;
                const double rel = e[g] / e0[g];
                if (bad[g]) {
                    // an element of this problem inverted: it fails, the others go on
                    state[g] = 2;
                    code[g] = TG_BAD_JACOBIAN;
                } else if (rel > o->divergence_tolerance)
                    state[g] = 2;
                else if (rel < o->rel_tolerance || e[g] < o->abs_tolerance)
                    state[g] = 1;
                else if (iter[g] >= o->max_iterations)
                    state[g] = 2;
                if (state[g] == 2) restore(g);
            }
        }
        std::vector<int> ones(G, 1);
        TG_HIP(hipMemcpyAsync(c->seg.d_active.p, ones.data(), sizeof(int) * G, hipMemcpyHostToDevice, c->stream));
        TG_HIP(hipStreamSynchronize(c->stream));
        for (int g = 0; g < G; g++) {
            stats[g].status = state[g] == 1 ? 0 : (code[g] ? code[g] : 1);
            stats[g].iterations = iter[g] + 1;
            stats[g].error0 = e0[g];
            stats[g].error = e[g];
            stats[g].lin_iterations = lin[g];
        }
    });
}

int tg_synchronize(tg_ctx* c)
{
    if (!c) return TG_BAD_INPUT;
    return guarded(c, [&] { TG_HIP(hipStreamSynchronize(c->stream)); });
}

int tg_timing_enable(tg_ctx* c, int enable)
{
    if (!c) return TG_BAD_INPUT;
    return guarded(c, [&] {
        drain_events(c);
        c->timing.enabled = enable != 0;
    });
}

int tg_timing_reset(tg_ctx* c)
{
    if (!c) return TG_BAD_INPUT;
    return guarded(c, [&] {
        drain_events(c);
        c->timing.acc = tg_timing{};
    });
}

int tg_timing_get(tg_ctx* c, tg_timing* out)
{
    if (!c || !out) return TG_BAD_INPUT;
    return guarded(c, [&] {
        drain_events(c);
        *out = c->timing.acc;
    });
}

int tg_set_tuning(tg_ctx* c, int knob, long value)
{
    if (!c) return TG_BAD_INPUT;
    return guarded(c, [&] {
        if (knob == TG_TUNE_STIFF_CHUNK) {
            if (value < -1) fail(TG_BAD_INPUT, "stiffness chunk must be -1, 0 or positive");
            c->stiff_chunk = value;
        } else if (knob == TG_TUNE_STIFF_SHAPE) {
            if (value < 0 || value > 18) fail(TG_OUT_OF_RANGE, "stiffness shape must be 0..18");
            c->stiff_variant = (int)value;
        } else if (knob == TG_TUNE_NEWTON_ASSEMBLY) {
            if (value < 0 || value > 1) fail(TG_OUT_OF_RANGE, "Newton assembly must be 0 or 1");
            c->newton_assembly = (int)value;
        } else if (knob == TG_TUNE_PCG_GRAPH) {
            if (value < 0 || value > 1) fail(TG_OUT_OF_RANGE, "PCG graph must be 0 or 1");
            c->pcg_graph = (int)value;
        } else if (knob == TG_TUNE_SPMV_XCD) {
            if (value < -1 || value > 7) fail(TG_OUT_OF_RANGE, "SpMV variant must be 0..7, or -1 for the default");
            c->spmv_xcd = (int)value;
        } else if (knob == TG_TUNE_PCG_FUSED) {
            if (value < -1 || value > 1) fail(TG_OUT_OF_RANGE, "PCG fused must be 0 or 1, or -1 for the default");
            c->pcg_fused = (int)value;
        } else if (knob == TG_TUNE_AMG_REUSE) {
            if (value < 0 || value > 64) fail(TG_OUT_OF_RANGE, "multigrid reuse must be 0..64");
            c->amg.reuse = (int)value;
            c->amg.since = -1;
        } else if (knob == TG_TUNE_PCG_PERSIST) {
            if (value < -1 || value > PCG_PERSIST_MAX)
                fail(TG_OUT_OF_RANGE, "PCG persistent slices must be -1.." + std::to_string(PCG_PERSIST_MAX));
            c->pcg_persist = (int)value;
            if (c->initialized) build_persist_list(c);
        } else if (knob == TG_TUNE_STIFF_GATHER) {
            if (value < 0 || value > 1) fail(TG_OUT_OF_RANGE, "stiffness gather must be 0 or 1");
            c->stiff_gather = (int)value;
        } else if (knob == TG_TUNE_AMG_DEGREE) {
            if (value < 2 || value > 4) fail(TG_OUT_OF_RANGE, "multigrid smoother degree must be 2..4");
            c->amg.degree = (int)value;
            c->amg.gen++;
            c->amg.since = -1;
        } else if (knob == TG_TUNE_AMG_WEIGHT) {
            if (value < 100 || value > 190) fail(TG_OUT_OF_RANGE, "multigrid coarse weight must be 100..190");
            c->amg.weight = (int)value;
            c->amg.gen++;
            c->amg.since = -1;
        } else
            fail(TG_BAD_INPUT, "unknown tuning knob " + std::to_string(knob));
    });
}

int tg_get_tuning(tg_ctx* c, int knob, long* value)
{
    if (!c || !value) return TG_BAD_INPUT;
    return guarded(c, [&] {
        if (knob == TG_TUNE_STIFF_CHUNK)
            *value = c->stiff_chunk;
        else if (knob == TG_TUNE_STIFF_SHAPE)
            *value = stiffness_shape(c);
        else if (knob == TG_TUNE_NEWTON_ASSEMBLY)
            *value = c->newton_assembly;
        else if (knob == TG_TUNE_PCG_GRAPH)
            *value = c->pcg_graph;
        else if (knob == TG_TUNE_SPMV_XCD)
            *value = spmv_variant(c);
        else if (knob == TG_TUNE_PCG_FUSED)
            *value = pcg_fused(c) ? 1 : 0;
        else if (knob == TG_TUNE_STIFF_GATHER)
            *value = c->stiff_gather;
        else if (knob == TG_TUNE_PCG_PERSIST)
            *value = pcg_persist(c);
        else if (knob == TG_TUNE_AMG_REUSE)
            *value = c->amg.reuse;
        else if (knob == TG_TUNE_AMG_DEGREE)
            *value = c->amg.degree;
        else if (knob == TG_TUNE_AMG_WEIGHT)
            *value = c->amg.weight;
        else
            fail(TG_BAD_INPUT, "unknown tuning knob " + std::to_string(knob));
    });
}

}  // extern "C"
