// Host-side setup of the equation system: equation numbers, the upper MSR
// pattern, element colouring and the solver's sliced-ELL block structure.
// Runs once per mesh (the reference does the same work in
// FEManagerT::SetEquationSystem, tahoe/src/main/FEManagerT.cpp:2248-2313).
#include <algorithm>
#include <cstring>
#include <thread>

#include "tg_internal.h"

namespace tg {

// NodeManagerT::SetEquationNumbers (tahoe/src/nodes/NodeManagerT.cpp:712-760)
// -> FieldT::InitEquations/SetEquations (tahoe/src/nodes/FieldT.cpp:635-690):
// dofs with a kinematic card get -1, the others 1.. in node-major, dof-minor
// order.
int equation_numbers(int nnd, const std::vector<int>& kn, const std::vector<int>& kd, std::vector<int>& eqnos)
{
    eqnos.assign((size_t)nnd * 3, 0);
    for (size_t k = 0; k < kn.size(); k++) eqnos[(size_t)kn[k] * 3 + kd[k]] = -1;
    int n = 0;
    for (auto& e : eqnos)
        if (e == 0) e = ++n;
    return n;
}

// node -> incident elements (CSR).  Counted and filled by all workers with
// atomic increments, so the order of a node's elements varies from run to
// run; node_graph sorts what it derives from them.
static void node_elements(int nnd, int nel, const int* conn, std::vector<int>& ptr, ivec& list)
{
    ptr.assign((size_t)nnd + 1, 0);
    int* cnt = ptr.data() + 1;
    parallel_for(8L * nel, [&](long b, long e) {
        for (long k = b; k < e; k++) __atomic_fetch_add(&cnt[conn[k]], 1, __ATOMIC_RELAXED);
    });
    for (int a = 0; a < nnd; a++) ptr[a + 1] += ptr[a];
    list.resize((size_t)ptr[nnd]);
    std::vector<int> fill(ptr.begin(), ptr.end() - 1);
    parallel_for(8L * nel, [&](long b, long e) {
        for (long k = b; k < e; k++) list[(size_t)__atomic_fetch_add(&fill[conn[k]], 1, __ATOMIC_RELAXED)] = (int)(k / 8);
    });
}

// Node adjacency as CSR (sorted unique neighbours including the node itself),
// built once and shared by the MSR pattern and the block operator.  One pass:
// each worker de-duplicates the neighbours of its node range with a per-node
// stamp array, sorts them into a list of its own, and the lists are then
// concatenated in node order.  The buffers live in w (a context's, kept
// across set-ups so that a set-up of a mesh no larger than before touches no
// new host pages) or, without one, in a local workspace.
void node_graph(int nnd, int nel, const int* conn, std::vector<long>& nptr, ivec& nlist, GraphWork* w)
{
    GraphWork local;
    GraphWork& g = w ? *w : local;
    node_elements(nnd, nel, conn, g.ptr, g.list);
    const std::vector<int>& ptr = g.ptr;
    const ivec& list = g.list;
    for (int a = 0; a < nnd; a++)
        if (ptr[a + 1] - ptr[a] > 64) fail(TG_OUT_OF_RANGE, "a node is shared by more than 64 elements");
    const int parts = nnd < 4096 ? 1 : worker_count();
    if ((int)g.part.size() < parts) g.part.resize((size_t)parts);
    if ((int)g.stamp.size() < parts) g.stamp.resize((size_t)parts);
    g.deg.resize((size_t)nnd);
    parallel_parts(nnd, parts, [&](int t, long b, long e) {
        if (b >= e) return;
        std::vector<int>& stamp = g.stamp[(size_t)t];
        stamp.assign((size_t)nnd, -1);
        ivec& out = g.part[(size_t)t];
        out.resize((size_t)(8L * (ptr[e] - ptr[b])));  // the candidates bound the unique count
        int* o = out.data();
        size_t m = 0;
        for (long a = b; a < e; a++) {
            const size_t n0 = m;
            for (int p = ptr[a]; p < ptr[a + 1]; p++) {
                const int* c = conn + 8L * list[p];
                for (int k = 0; k < 8; k++)
                    if (stamp[c[k]] != (int)a) {
                        stamp[c[k]] = (int)a;
                        o[m++] = c[k];
                    }
            }
            std::sort(o + n0, o + m);
            g.deg[a] = (int)(m - n0);
        }
        out.resize(m);
    });
    nptr.resize((size_t)nnd + 1);
    nptr[0] = 0;
    for (int a = 0; a < nnd; a++) nptr[a + 1] = nptr[a] + g.deg[a];
    nlist.resize((size_t)nptr[nnd]);
    parallel_parts(nnd, parts, [&](int t, long b, long e) {
        const ivec& src = g.part[(size_t)t];
        if (b < e) std::memcpy(nlist.data() + nptr[b], src.data(), sizeof(int) * src.size());
    });
}

// Row pointers of the upper-triangle MSR pattern of GraphT::MakeGraph(active,
// add_self, upper_only) (toolbox/src/graph/GraphT.cpp:376-456) packed by
// MSRBuilderT::GenerateMSR (toolbox/src/graph/MSRBuilderT.cpp:134-181): row r
// lists every active equation > r sharing an element with r, ascending;
// bindx[0] = n_eq + 1 and bindx[r + 1] - bindx[r] is row r's length.
// Equation numbers are node-major and dof-minor (equation_numbers), so the
// columns of row (a, i) are the active dofs j > i of node a and every active
// dof of each neighbour b > a: a row's length is counted per node, and the
// columns themselves are written on the device (launch_msr_cols) from the
// block operator's neighbour lists.
void msr_rows(int n_eq, int nnd, const int* eqnos, const std::vector<long>& nptr, const ivec& nlist, ivec& bindx)
{
    std::vector<int> rowlen((size_t)n_eq, 0);
    parallel_for(nnd, [&](long b, long e) {
        for (long a = b; a < e; a++) {
            int later = 0;  // active dofs of the neighbours after a
            for (long k = nptr[a]; k < nptr[a + 1]; k++)
                if (nlist[k] > a)
                    for (int j = 0; j < 3; j++) later += eqnos[3L * nlist[k] + j] > 0;
            for (int i = 0; i < 3; i++) {
                const int row = eqnos[3 * a + i];
                if (row <= 0) continue;
                int own = 0;
                for (int j = i + 1; j < 3; j++) own += eqnos[3 * a + j] > 0;
                rowlen[row - 1] = later + own;
            }
        }
    });
    long total = 0;
    for (int r = 0; r < n_eq; r++) total += rowlen[r];
    if ((long)n_eq + 1 + total > 0x7fffffffL) fail(TG_OUT_OF_RANGE, "MSR pattern exceeds 32-bit indexing");
    bindx.resize((size_t)n_eq + 1);
    bindx[0] = n_eq + 1;
    for (int r = 0; r < n_eq; r++) bindx[r + 1] = bindx[r] + rowlen[r];
}

// First-fit greedy colouring: no two elements of one colour share a node, so
// a colour's scatter into node-space vectors and MSR values needs no atomics.
// Elements keep their mesh order inside a colour.
int color_elements(int nnd, int nel, const int* conn, std::vector<int>& cptr, std::vector<int>& elems)
{
    std::vector<uint64_t> used((size_t)nnd, 0);
    std::vector<int> color((size_t)nel);
    int ncol = 0;
    for (int e = 0; e < nel; e++) {
        uint64_t forbid = 0;
        for (int a = 0; a < 8; a++) forbid |= used[conn[8 * e + a]];
        if (~forbid == 0) fail(TG_GENERAL_FAIL, "element colouring needs more than 64 colours");
        int c = __builtin_ctzll(~forbid);
        color[e] = c;
        ncol = std::max(ncol, c + 1);
        for (int a = 0; a < 8; a++) used[conn[8 * e + a]] |= (1ull << c);
    }
    cptr.assign((size_t)ncol + 1, 0);
    for (int e = 0; e < nel; e++) cptr[color[e] + 1]++;
    for (int c = 0; c < ncol; c++) cptr[c + 1] += cptr[c];
    elems.resize((size_t)nel);
    std::vector<int> fill(cptr.begin(), cptr.end() - 1);
    for (int e = 0; e < nel; e++) elems[(size_t)fill[color[e]]++] = e;
    return ncol;
}

// Solver operator structure: block row = node, block columns = neighbour
// nodes (ascending); 64 block rows per slice, each slice padded to its
// widest row.  d_src (built on the device, launch_build_src) maps each
// stored value to the MSR value it mirrors (upper entry (min(r,c), max(r,c))),
// -1 for inactive dofs and padding; d_tpos gives each lower block its upper
// partner.
void build_bsell(tg_ctx* c)
{
    SetupTrace phase(c, "build_bsell");
    Bsell& B = c->op;
    const int nnd = c->nnd;
    const std::vector<long>& nptr = c->nptr;
    const ivec& nlist = c->nlist;
    B.nslices = (nnd + 63) / 64;
    B.slice_base.assign((size_t)B.nslices, 0);
    B.slice_width.assign((size_t)B.nslices, 0);
    parallel_for(B.nslices, [&](long sb, long se) {
        for (long s = sb; s < se; s++) {
            int w = 0;
            for (long a = 64 * s; a < std::min((long)nnd, 64 * s + 64); a++) w = std::max(w, (int)(nptr[a + 1] - nptr[a]));
            B.slice_width[s] = w;
        }
    });
    long base = 0;
    for (int s = 0; s < B.nslices; s++) {
        B.slice_base[s] = (int)base;
        base += B.slice_width[s];
    }
    B.nblocks_padded = base * 64;
    if (B.nblocks_padded * 9 > 0x7fffffffL) fail(TG_OUT_OF_RANGE, "block operator exceeds 32-bit indexing");
    ivec& col = B.h_col;  // kept by the context; each slice fills its own range, padding -1
    col.resize((size_t)B.nblocks_padded);
    parallel_for(B.nslices, [&](long sb, long se) {
        for (long s = sb; s < se; s++) {
            std::fill(col.begin() + (long)B.slice_base[s] * 64, col.begin() + ((long)B.slice_base[s] + B.slice_width[s]) * 64,
                      -1);
            for (int l = 0; l < 64; l++) {
                const long a = 64 * s + l;
                if (a >= nnd) break;
                for (long k = nptr[a]; k < nptr[a + 1]; k++)
                    col[(size_t)(((long)B.slice_base[s] + (k - nptr[a])) * 64 + l)] = nlist[k];
            }
        }
    });
    phase("slice widths + columns");
    B.d_slice_base.alloc(B.nslices);
    B.d_slice_width.alloc(B.nslices);
    B.d_col.alloc((size_t)B.nblocks_padded);
    B.d_src.alloc((size_t)B.nblocks_padded * 9);
    B.d_tpos.alloc((size_t)B.nblocks_padded);
    B.d_val.alloc((size_t)B.nblocks_padded * 9);
    TG_HIP(hipMemcpyAsync(B.d_slice_base.p, B.slice_base.data(), sizeof(int) * B.nslices, hipMemcpyHostToDevice, c->stream));
    TG_HIP(hipMemcpyAsync(B.d_slice_width.p, B.slice_width.data(), sizeof(int) * B.nslices, hipMemcpyHostToDevice, c->stream));
    TG_HIP(hipMemcpyAsync(B.d_col.p, col.data(), sizeof(int) * col.size(), hipMemcpyHostToDevice, c->stream));
    phase("allocate + upload");
    launch_msr_cols(c);  // the MSR columns from the neighbour lists, before anything reads them
    launch_build_src(c);
    phase("MSR columns, src, tpos");
    // MSR ranges of the rows of each half slice (node-major equations: the
    // active rows of 32 consecutive nodes are consecutive equations, their
    // off-diagonal values consecutive)
    std::vector<int4> rng(2 * (size_t)B.nslices);
    std::vector<long> most(2 * (size_t)B.nslices, 0);
    parallel_for(2L * B.nslices, [&](long sb, long se) {
        for (long s = sb; s < se; s++) {  // half slices of 32 nodes
            int r0 = -1, r1 = -1;
            for (long a = 32 * s; a < std::min((long)nnd, 32 * s + 32); a++)
                for (int i = 0; i < 3; i++) {
                    const int e = c->eqnos[3 * a + i];
                    if (e <= 0) continue;
                    if (r0 < 0) r0 = e - 1;
                    r1 = e;
                }
            if (r0 < 0) {
                rng[s] = make_int4(0, 0, 0, 0);
                continue;
            }
            rng[s] = make_int4(c->bindx[r0], c->bindx[r1], r0, r1);
            most[s] = (long)(c->bindx[r1] - c->bindx[r0]) + (r1 - r0);
        }
    });
    B.lds_values = most.empty() ? 0 : *std::max_element(most.begin(), most.end());
    B.d_srange.alloc(rng.size());
    TG_HIP(hipMemcpyAsync(B.d_srange.p, rng.data(), sizeof(int4) * rng.size(), hipMemcpyHostToDevice, c->stream));
    TG_HIP(hipStreamSynchronize(c->stream));  // col and rng (host) outlive their copies
    phase("row ranges");
}


}  // namespace tg
