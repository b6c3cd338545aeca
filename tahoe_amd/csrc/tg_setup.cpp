// Host-side setup of the equation system: equation numbers, element
// colouring, and the node graph the multigrid's aggregation reads (the
// operator's structure itself is built on the device, k_graph.hip).  Runs
// once per mesh (the reference does the same work in
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
// the same lists as the device build's (k_graph.hip).  One pass:
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

}  // namespace tg
