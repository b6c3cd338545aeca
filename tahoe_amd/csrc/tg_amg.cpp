// Aggregation multigrid preconditioner for the inner PCG: the structure,
// built once per mesh on the host.
//
// The reference has no multigrid (its direct and Krylov solvers use the
// assembled matrix as is); this is unsmoothed aggregation with rigid-body
// near-null-space (Vanek, Mandel and Brezina, Computing 56 (1996) 179-196,
// with the tentative prolongator of their section 3) on the node graph of
// the mesh:
//   - aggregates: every unaggregated node, visited in index order, takes its
//     7 nearest unaggregated neighbours (2 x 2 x 2 node blocks on a
//     structured hexahedral mesh); aggregates under 4 nodes are dissolved
//     into the aggregates of their members' nearest neighbours;
//   - tentative prolongator: per aggregate, the rigid-body modes of its
//     nodes (3 translations, 3 rotations about the aggregate centroid; rows
//     of prescribed dofs zero) orthonormalised by Gram-Schmidt, Q; the
//     coarse near-null-space is R, so every coarser node carries 6 dofs;
//   - coarse operators A_c = P^T A P, whose pattern is the aggregate
//     adjacency.  The values are formed on the device each time the
//     operator changes (k_amg.hip) from gather lists built here: for each
//     upper coarse block (A <= B) the fine blocks (i, j) with i in A and j
//     in B, in a fixed order, so the product is deterministic.
// Coarsening stops at <= 256 coarse dofs; that level is solved by a dense
// inverse.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "tg_internal.h"

namespace tg {

namespace {

struct Graph {
    std::vector<long> ptr;  // CSR adjacency, sorted, the node itself included
    std::vector<int> adj;
};

constexpr int AGG_TARGET = 8;
constexpr int AGG_MIN = 4;
constexpr int COARSEST_DOFS = 256;
constexpr int MAX_LEVELS = 12;

double dist2(const std::vector<double>& x, long i, long j)
{
    const double dx = x[3 * j] - x[3 * i], dy = x[3 * j + 1] - x[3 * i + 1], dz = x[3 * j + 2] - x[3 * i + 2];
    return dx * dx + dy * dy + dz * dz;
}

// Greedy aggregation of the active nodes; returns the aggregate count and
// agg[i] (-1 for inactive nodes).  Aggregates are numbered by their
// smallest member.
int aggregate(int n, const Graph& g, const std::vector<double>& xyz, const std::vector<char>& active,
              std::vector<int>& agg)
{
    agg.assign((size_t)n, -1);
    std::vector<int> size;
    std::vector<std::pair<double, int>> cand;
    int nagg = 0;
    for (int i = 0; i < n; i++) {
        if (!active[i] || agg[i] >= 0) continue;
        cand.clear();
        for (long k = g.ptr[i]; k < g.ptr[i + 1]; k++) {
            const int j = g.adj[k];
            if (j == i || !active[j] || agg[j] >= 0) continue;
            cand.push_back({dist2(xyz, i, j), j});
        }
        const int m = std::min((int)cand.size(), AGG_TARGET - 1);
        std::partial_sort(cand.begin(), cand.begin() + m, cand.end());
        agg[i] = nagg;
        for (int k = 0; k < m; k++) agg[cand[k].second] = nagg;
        size.push_back(m + 1);
        nagg++;
    }
    std::vector<int> merged(agg);
    for (int i = 0; i < n; i++) {
        if (agg[i] < 0 || size[agg[i]] >= AGG_MIN) continue;
        double best = INFINITY;
        int to = -1;
        for (long k = g.ptr[i]; k < g.ptr[i + 1]; k++) {
            const int j = g.adj[k];
            if (j == i || agg[j] < 0 || size[agg[j]] < AGG_MIN) continue;
            const double d = dist2(xyz, i, j);
            if (d < best) {
                best = d;
                to = agg[j];
            }
        }
        if (to >= 0) merged[i] = to;
    }
    std::vector<int> id((size_t)nagg, -1);
    int m = 0;
    for (int i = 0; i < n; i++) {
        if (merged[i] < 0) {
            agg[i] = -1;
            continue;
        }
        int& t = id[merged[i]];
        if (t < 0) t = m++;
        agg[i] = t;
    }
    return m;
}

// Thin QR of the rows x 6 matrix M (row-major) by modified Gram-Schmidt with
// one re-orthogonalisation pass: M becomes Q, R (6 x 6, row-major) is
// returned.  A column whose remainder is below 1e-10 of its norm is dropped
// (zero column of Q, zero row of R): that coarse dof is inactive.
void thin_qr(int rows, double* M, double* R)
{
    std::fill(R, R + 36, 0.0);
    for (int c = 0; c < 6; c++) {
        double n0 = 0.0;
        for (int r = 0; r < rows; r++) n0 += M[6 * r + c] * M[6 * r + c];
        n0 = std::sqrt(n0);
        for (int pass = 0; pass < 2; pass++)
            for (int k = 0; k < c; k++) {
                double d = 0.0;
                for (int r = 0; r < rows; r++) d += M[6 * r + k] * M[6 * r + c];
                R[6 * k + c] += d;
                for (int r = 0; r < rows; r++) M[6 * r + c] -= d * M[6 * r + k];
            }
        double nv = 0.0;
        for (int r = 0; r < rows; r++) nv += M[6 * r + c] * M[6 * r + c];
        nv = std::sqrt(nv);
        if (n0 == 0.0 || nv <= 1e-10 * n0) {
            for (int r = 0; r < rows; r++) M[6 * r + c] = 0.0;
            for (int k = 0; k < c; k++) R[6 * k + c] = 0.0;
            continue;
        }
        for (int r = 0; r < rows; r++) M[6 * r + c] /= nv;
        R[6 * c + c] = nv;
    }
}

template <class T>
void upload(tg_ctx* c, DBuf<T>& d, const std::vector<T>& h)
{
    d.alloc(h.size());
    if (!h.empty()) TG_HIP(hipMemcpyAsync(d.p, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice, c->stream));
}

}  // namespace

void amg_setup(tg_ctx* c)
{
    Amg& G = c->amg;
    G.clear();
    G.gen++;
    if (c->seg.nseg != 1) fail(TG_BAD_INPUT, "the multigrid preconditioner needs a single problem (no segments)");
    const int nnd = c->nnd;
    // level-0 graph: the node adjacency, whose order is also the block
    // operator's slot order within a row
    Graph g;
    {
        std::vector<long> ptr;
        ivec lst;
        node_graph(nnd, c->nel, c->conn.data(), ptr, lst, &c->graph_work);
        g.adj.assign(lst.begin(), lst.end());
        g.ptr.swap(ptr);
    }
    std::vector<double> xyz(c->X);
    std::vector<char> active((size_t)nnd, 0);
    for (int i = 0; i < nnd; i++)
        for (int k = 0; k < 3; k++)
            if (c->eqnos[3L * i + k] > 0) active[i] = 1;
    std::vector<double> B;  // [n][bs][6] near-null-space rows of the current level (levels >= 1)

    AmgLevel* cur = new AmgLevel;
    G.lv.push_back(cur);
    cur->n = nnd;
    cur->bs = 3;
    cur->dinv = c->vdinv.p;
    for (auto* v : {&cur->d_x, &cur->d_r, &cur->d_d0, &cur->d_d1}) v->alloc(3 * (size_t)nnd);

    std::vector<int> agg, mptr, mem, stamp;
    std::vector<double> Mx, R6;
    for (int level = 0;; level++) {
        const int n = cur->n, bs = cur->bs;
        const int nagg = aggregate(n, g, xyz, active, agg);
        if (nagg == 0) fail(TG_BAD_INPUT, "multigrid: no active dofs");
        // members of each aggregate, ascending
        mptr.assign((size_t)nagg + 1, 0);
        for (int i = 0; i < n; i++)
            if (agg[i] >= 0) mptr[agg[i] + 1]++;
        for (int a = 0; a < nagg; a++) mptr[a + 1] += mptr[a];
        mem.resize((size_t)mptr[nagg]);
        {
            std::vector<int> fill(mptr.begin(), mptr.end() - 1);
            for (int i = 0; i < n; i++)
                if (agg[i] >= 0) mem[fill[agg[i]]++] = i;
        }
        // centroids (the next level's coordinates)
        std::vector<double> cen(3 * (size_t)nagg, 0.0);
        for (int a = 0; a < nagg; a++) {
            const int m = mptr[a + 1] - mptr[a];
            for (int t = mptr[a]; t < mptr[a + 1]; t++)
                for (int k = 0; k < 3; k++) cen[3L * a + k] += xyz[3L * mem[t] + k];
            for (int k = 0; k < 3; k++) cen[3L * a + k] /= m;
        }
        // tentative prolongator rows and the next level's near-null-space
        std::vector<float> P((size_t)n * bs * 6, 0.0f);
        std::vector<double> Bn((size_t)nagg * 36, 0.0);
        for (int a = 0; a < nagg; a++) {
            const int m = mptr[a + 1] - mptr[a];
            Mx.assign((size_t)m * bs * 6, 0.0);
            for (int t = 0; t < m; t++) {
                const int i = mem[mptr[a] + t];
                double* rows = Mx.data() + (size_t)t * bs * 6;
                if (level == 0) {
                    const double dx = xyz[3L * i] - cen[3L * a], dy = xyz[3L * i + 1] - cen[3L * a + 1],
                                 dz = xyz[3L * i + 2] - cen[3L * a + 2];
                    const double rb[3][6] = {{1, 0, 0, 0, dz, -dy}, {0, 1, 0, -dz, 0, dx}, {0, 0, 1, dy, -dx, 0}};
                    for (int p = 0; p < 3; p++)
                        if (c->eqnos[3L * i + p] > 0)
                            for (int k = 0; k < 6; k++) rows[6 * p + k] = rb[p][k];
                } else {
                    std::memcpy(rows, B.data() + (size_t)i * 36, sizeof(double) * 36);
                }
            }
            R6.assign(36, 0.0);
            thin_qr(m * bs, Mx.data(), R6.data());
            for (int t = 0; t < m; t++) {
                const int i = mem[mptr[a] + t];
                for (int q = 0; q < bs * 6; q++) P[(size_t)i * bs * 6 + q] = (float)Mx[(size_t)t * bs * 6 + q];
            }
            std::memcpy(Bn.data() + (size_t)a * 36, R6.data(), sizeof(double) * 36);
        }
        // next level: pattern = aggregate adjacency (sorted, diagonal included)
        AmgLevel* nx = new AmgLevel;
        G.lv.push_back(nx);
        nx->n = nagg;
        nx->bs = 6;
        Graph gc;
        gc.ptr.assign((size_t)nagg + 1, 0);
        stamp.assign((size_t)nagg, -1);
        std::vector<int> row;
        for (int a = 0; a < nagg; a++) {
            row.clear();
            for (int t = mptr[a]; t < mptr[a + 1]; t++) {
                const int i = mem[t];
                for (long k = g.ptr[i]; k < g.ptr[i + 1]; k++) {
                    const int b = agg[g.adj[k]];
                    if (b >= 0 && stamp[b] != a) {
                        stamp[b] = a;
                        row.push_back(b);
                    }
                }
            }
            std::sort(row.begin(), row.end());
            gc.adj.insert(gc.adj.end(), row.begin(), row.end());
            gc.ptr[a + 1] = (long)gc.adj.size();
        }
        if (gc.adj.size() > 0x7fffffffL / 36) fail(TG_OUT_OF_RANGE, "multigrid level exceeds 32-bit indexing");
        nx->rowptr.assign(gc.ptr.begin(), gc.ptr.end());
        nx->col = gc.adj;
        // sliced layout of the next level
        {
            const int ns = (nagg + 63) / 64;
            nx->sbase.assign((size_t)ns, 0);
            nx->swidth.assign((size_t)ns, 0);
            long base = 0;
            for (int sl = 0; sl < ns; sl++) {
                int w = 0;
                for (int a = 64 * sl; a < std::min(nagg, 64 * sl + 64); a++)
                    w = std::max(w, (int)(gc.ptr[a + 1] - gc.ptr[a]));
                nx->sbase[sl] = (int)base;
                nx->swidth[sl] = w;
                base += w;
            }
            nx->nslots = base * 64;
            if (nx->nslots > 0x7fffffffL / 36) fail(TG_OUT_OF_RANGE, "multigrid level exceeds 32-bit indexing");
            nx->slot.assign(gc.adj.size(), 0);
            for (int a = 0; a < nagg; a++)
                for (long k = gc.ptr[a]; k < gc.ptr[a + 1]; k++)
                    nx->slot[k] = (int)(((long)nx->sbase[a >> 6] + (k - gc.ptr[a])) * 64 + (a & 63));
        }
        std::vector<int> diag((size_t)nagg);
        for (int a = 0; a < nagg; a++)
            diag[a] = nx->slot[std::lower_bound(gc.adj.begin() + gc.ptr[a], gc.adj.begin() + gc.ptr[a + 1], a) -
                               gc.adj.begin()];
        // Galerkin gather lists: upper coarse block (A, B), B >= A, gets
        // P_i^T A_ij P_j over the fine blocks (i, j) with i in A (ascending)
        // and j in B (in the fine row's block order)
        std::vector<int> up, gl;
        std::vector<long> gp(1, 0);
        std::vector<std::vector<int>> bucket;
        for (int a = 0; a < nagg; a++) {
            const int* r0 = gc.adj.data() + gc.ptr[a];
            const int* r1 = gc.adj.data() + gc.ptr[a + 1];
            const int* first = std::lower_bound(r0, r1, a);
            const int nb = (int)(r1 - first);
            if ((int)bucket.size() < nb) bucket.resize(nb);
            for (int k = 0; k < nb; k++) bucket[k].clear();
            for (int t = mptr[a]; t < mptr[a + 1]; t++) {
                const int i = mem[t];
                const long deg = level == 0 ? g.ptr[i + 1] - g.ptr[i] : cur->rowptr[i + 1] - cur->rowptr[i];
                for (long k = 0; k < deg; k++) {
                    int j, fb;
                    if (level == 0) {
                        j = g.adj[g.ptr[i] + k];
                        fb = (int)(((long)c->op.slice_base[i >> 6] + k) * 64 + (i & 63));
                    } else {
                        const int q = (int)(cur->rowptr[i] + k);
                        j = cur->col[q];
                        fb = cur->slot[q];
                    }
                    const int b = agg[j];
                    if (b < a) continue;
                    const int pos = (int)(std::lower_bound(first, r1, b) - first);
                    bucket[pos].push_back(fb);
                    bucket[pos].push_back(i);
                }
            }
            for (int k = 0; k < nb; k++) {
                const int b = first[k];
                const int ab = (int)(first - gc.adj.data()) + k;
                int ba = -1;
                if (b != a)
                    ba = (int)(std::lower_bound(gc.adj.begin() + gc.ptr[b], gc.adj.begin() + gc.ptr[b + 1], a) -
                               gc.adj.begin());
                up.push_back(nx->slot[ab]);
                up.push_back(ba < 0 ? -1 : nx->slot[ba]);
                gl.insert(gl.end(), bucket[k].begin(), bucket[k].end());
                gp.push_back((long)gl.size() / 2);
            }
        }
        // every index the device kernels will follow, checked here
        {
            const long fine_blocks = level == 0 ? c->op.nblocks_padded : cur->nslots;
            const long coarse_blocks = nx->nslots;
            for (size_t k = 0; k < gl.size(); k += 2)
                if (gl[k] < 0 || gl[k] >= fine_blocks || gl[k + 1] < 0 || gl[k + 1] >= n)
                    fail(TG_OUT_OF_RANGE, "multigrid gather list out of range");
            for (size_t k = 0; k < up.size(); k += 2)
                if (up[k] < 0 || up[k] >= coarse_blocks || up[k + 1] < -1 || up[k + 1] >= coarse_blocks)
                    fail(TG_OUT_OF_RANGE, "multigrid coarse block index out of range");
            for (int i = 0; i < n; i++)
                if (agg[i] < -1 || agg[i] >= nagg) fail(TG_OUT_OF_RANGE, "multigrid aggregate index out of range");
            for (int v : mem)
                if (v < 0 || v >= n) fail(TG_OUT_OF_RANGE, "multigrid member index out of range");
            if ((long)gp.size() != (long)up.size() / 2 + 1 || gp.back() * 2 != (long)gl.size())
                fail(TG_OUT_OF_RANGE, "multigrid gather pointers inconsistent");
        }
        cur->nnext = nagg;
        cur->nup = (long)up.size() / 2;
        upload(c, cur->d_agg, agg);
        upload(c, cur->d_mptr, mptr);
        upload(c, cur->d_mem, mem);
        upload(c, cur->d_P, P);
        upload(c, cur->d_up, up);
        upload(c, cur->d_gptr, gp);
        upload(c, cur->d_glist, gl);
        {
            std::vector<int> col_slot((size_t)nx->nslots, -1);
            for (size_t q = 0; q < nx->col.size(); q++) col_slot[(size_t)nx->slot[q]] = nx->col[q];
            upload(c, nx->d_col, col_slot);
            upload(c, nx->d_sbase, nx->sbase);
            upload(c, nx->d_swidth, nx->swidth);
            TG_HIP(hipStreamSynchronize(c->stream));  // col_slot outlives its copy
        }
        upload(c, nx->d_diagblk, diag);
        nx->d_val.alloc((size_t)nx->nslots * 36);
        nx->d_dinv_own.alloc(6 * (size_t)nagg);
        nx->dinv = nx->d_dinv_own.p;
        for (auto* v : {&nx->d_f, &nx->d_x, &nx->d_r, &nx->d_d0, &nx->d_d1}) v->alloc(6 * (size_t)nagg);
        TG_HIP(hipStreamSynchronize(c->stream));  // host vectors outlive their copies
        const bool stalled = nagg > 0.8 * n;
        if (6L * nagg <= COARSEST_DOFS || stalled || (int)G.lv.size() >= MAX_LEVELS) {
            if (6L * nagg > 4096) fail(TG_OUT_OF_RANGE, "multigrid coarsening stalled above 4096 coarse dofs");
            break;
        }
        // the next level becomes current
        g.ptr.swap(gc.ptr);
        g.adj.swap(gc.adj);
        xyz.swap(cen);
        B.swap(Bn);
        active.assign((size_t)nagg, 0);
        for (int a = 0; a < nagg; a++)
            for (int k = 0; k < 36; k++)
                if (B[(size_t)a * 36 + k] != 0.0) active[a] = 1;
        cur = nx;
    }
    G.ncd = 6 * G.lv.back()->n;
    G.d_cinv.alloc((size_t)G.ncd * G.ncd);
    G.d_coef.alloc(8 * G.lv.size());
    G.d_red.alloc(16 * G.lv.size());
    G.d_part.alloc(1024);
}

}  // namespace tg
