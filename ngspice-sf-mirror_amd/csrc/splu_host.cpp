/*
 * Symbolic analysis of one large sparse system for the whole-GPU
 * multifrontal LU (the klu_analyze replacement for circuits far beyond one
 * wave's LDS program, such as a 1024 x 1024 power-grid mesh; klu_analyze.c
 * :257-488 orders with BTF + AMD and leaves pivoting to klu_factor).
 *
 * Static pivots.  A column with no structural diagonal (the branch unknown
 * of a voltage source) is paired with a node the branch is incident to, and
 * the two rows are exchanged: the rows of B = P A then carry the source's
 * +-1 incidence entries on both diagonals, the pivots partial pivoting
 * picks for such a pair in klu_kernel.c.  Each pair is ordered as one
 * vertex, node before branch.
 *
 * Ordering.  Nested dissection on the graph of B + B^T: a breadth-first
 * level structure from a pseudo-peripheral vertex (two sweeps), the first
 * level holding half the vertices as the separator, thinned of vertices
 * with no neighbour beyond it; recursion stops at LEAF vertices.  Each
 * separator and each leaf is one dense frontal matrix whose pivots are its
 * vertices and whose boundary is every later unknown coupled to them or to
 * a child's boundary; fronts are numbered in postorder and grouped by their
 * height in the tree, the unit the numeric phase runs in parallel.
 */
#include <algorithm>
#include <cstring>
#include <numeric>
#include <vector>

#include "ngmi_internal.h"
#include "splu_types.h"

namespace ngmi {

namespace {

struct Nested {
    const std::vector<int64_t> &gp;
    const std::vector<int> &ga;
    std::vector<int> mark, lvl;
    int stamp = 0;
    std::vector<std::vector<int>> fronts;
    std::vector<int> parent;

    Nested(const std::vector<int64_t> &p, const std::vector<int> &a, int ng)
        : gp(p), ga(a), mark(ng, 0), lvl(ng, -1) {}

    int add(std::vector<int> &&v, int par)
    {
        fronts.push_back(std::move(v));
        parent.push_back(par);
        return (int)fronts.size() - 1;
    }

    /* BFS over the vertices marked st, from s; lvl[] of those vertices must
     * be -1 beforehand.  Appends the visit order to out; returns the depth. */
    int bfs(int s, int st, std::vector<int> &out)
    {
        const size_t b = out.size();
        out.push_back(s);
        lvl[s] = 0;
        int depth = 0;
        for (size_t q = b; q < out.size(); q++) {
            const int v = out[q];
            depth = lvl[v];
            for (int64_t e = gp[v]; e < gp[v + 1]; e++) {
                const int u = ga[e];
                if (mark[u] == st && lvl[u] < 0) {
                    lvl[u] = lvl[v] + 1;
                    out.push_back(u);
                }
            }
        }
        return depth;
    }

    void split(std::vector<int> verts, int par)
    {
        if ((int)verts.size() <= SPLU_LEAF) {
            add(std::move(verts), par);
            return;
        }
        const int st = ++stamp;
        for (int v : verts) mark[v] = st, lvl[v] = -1;
        std::vector<int> order;
        order.reserve(verts.size());
        bfs(verts[0], st, order);
        if (order.size() < verts.size()) {
            /* disconnected: every large component becomes its own subtree;
             * small ones are packed together into leaves of up to LEAF
             * vertices (block-diagonal dense fronts), which keeps the
             * number of children per front, and so of extend-add
             * launches, small */
            std::vector<std::vector<int>> comps;
            comps.push_back(order);
            for (int v : verts)
                if (lvl[v] < 0) {
                    std::vector<int> c;
                    bfs(v, st, c);
                    comps.push_back(std::move(c));
                }
            std::vector<int> bin;
            for (auto &c : comps) {
                if ((int)c.size() > SPLU_LEAF) {
                    split(std::move(c), par);
                    continue;
                }
                if ((int)(bin.size() + c.size()) > SPLU_LEAF) {
                    add(std::move(bin), par);
                    bin.clear();
                }
                bin.insert(bin.end(), c.begin(), c.end());
            }
            if (!bin.empty()) add(std::move(bin), par);
            return;
        }
        const int s = order.back();
        for (int v : verts) lvl[v] = -1;
        order.clear();
        const int m = bfs(s, st, order);
        if (m < 2) {
            add(std::move(verts), par);
            return;
        }
        std::vector<int64_t> cnt(m + 1, 0);
        for (int v : verts) cnt[lvl[v]]++;
        int L = 0;
        int64_t cum = 0;
        for (; L <= m; L++) {
            cum += cnt[L];
            if (2 * cum >= (int64_t)verts.size()) break;
        }
        L = std::max(1, std::min(L, m - 1));
        std::vector<int> A, B, S;
        for (int v : verts) {
            if (lvl[v] < L) A.push_back(v);
            else if (lvl[v] > L) B.push_back(v);
            else {
                bool beyond = false;
                for (int64_t e = gp[v]; e < gp[v + 1] && !beyond; e++) {
                    const int u = ga[e];
                    beyond = mark[u] == st && lvl[u] == L + 1;
                }
                (beyond ? S : A).push_back(v);
            }
        }
        const int f = add(std::move(S), par);
        split(std::move(A), f);
        split(std::move(B), f);
    }
};

}  // namespace

int splu_analyze(int n, const int64_t *Ap, const int *Ai, SpluSym &S)
{
    if (n <= 0 || !Ap || !Ai) {
        set_error("splu_analyze: empty system");
        return NGMI_E_PARMVAL;
    }
    /* the pattern checks of klu_analyze_given.c:55-95: Ap[0] == 0, column
     * pointers non-decreasing, rows in range, no row twice in a column */
    if (Ap[0] != 0) {
        set_error("splu_analyze: Ap[0] is %lld, not 0", (long long)Ap[0]);
        return NGMI_E_PARMVAL;
    }
    const int64_t nnz = Ap[n];
    for (int c = 0; c < n; c++)
        if (Ap[c + 1] < Ap[c]) {
            set_error("splu_analyze: column pointers decrease at %d", c);
            return NGMI_E_PARMVAL;
        }
    {
        std::vector<int> seen(n, -1);
        for (int c = 0; c < n; c++)
            for (int64_t e = Ap[c]; e < Ap[c + 1]; e++) {
                const int r = Ai[e];
                if (r < 0 || r >= n) {
                    set_error("splu_analyze: row index %d out of range at %lld", r, (long long)e);
                    return NGMI_E_PARMVAL;
                }
                if (seen[r] == c) {
                    set_error("splu_analyze: row %d appears twice in column %d", r, c);
                    return NGMI_E_PARMVAL;
                }
                seen[r] = c;
            }
    }
    S = SpluSym();
    S.n = n;
    S.nnz = nnz;

    /* structural diagonal */
    std::vector<char> diag(n, 0);
    for (int c = 0; c < n; c++)
        for (int64_t e = Ap[c]; e < Ap[c + 1]; e++)
            if (Ai[e] == c) diag[c] = 1;
    auto has = [&](int r, int c) {
        for (int64_t e = Ap[c]; e < Ap[c + 1]; e++)
            if (Ai[e] == r) return true;
        return false;
    };

    /* pair every column without a diagonal with a node it is incident to */
    S.rowsrc.resize(n);
    std::iota(S.rowsrc.begin(), S.rowsrc.end(), 0);
    std::vector<int> mate(n, -1);
    for (int j = 0; j < n; j++) {
        if (diag[j]) continue;
        int k = -1;
        for (int64_t e = Ap[j]; e < Ap[j + 1] && k < 0; e++) {
            const int i = Ai[e];
            if (i != j && diag[i] && mate[i] < 0 && has(j, i)) k = i;
        }
        if (k < 0) {
            set_error("splu_analyze: unknown %d has no diagonal and no node to pair with", j);
            return NGMI_E_SINGULAR;
        }
        mate[j] = k;
        mate[k] = j;
        S.rowsrc[j] = k;
        S.rowsrc[k] = j;
    }

    /* groups: a pair (node, branch) or a single unknown */
    std::vector<int> grp(n, -1), gfirst;
    for (int v = 0; v < n; v++) {
        if (grp[v] >= 0) continue;
        int node = v;
        if (mate[v] >= 0 && !diag[v]) node = mate[v];
        grp[node] = (int)gfirst.size();
        if (mate[node] >= 0) grp[mate[node]] = grp[node];
        gfirst.push_back(node);
    }
    const int ng = (int)gfirst.size();

    /* symmetric adjacency of B (rows exchanged) on unknowns, then on groups */
    std::vector<int64_t> vp(n + 1, 0);
    for (int c = 0; c < n; c++)
        for (int64_t e = Ap[c]; e < Ap[c + 1]; e++) {
            const int r = S.rowsrc[Ai[e]];
            if (r != c) vp[r + 1]++, vp[c + 1]++;
        }
    for (int i = 0; i < n; i++) vp[i + 1] += vp[i];
    std::vector<int> va(vp[n]);
    {
        std::vector<int64_t> w(vp.begin(), vp.end() - 1);
        for (int c = 0; c < n; c++)
            for (int64_t e = Ap[c]; e < Ap[c + 1]; e++) {
                const int r = S.rowsrc[Ai[e]];
                if (r != c) va[w[r]++] = c, va[w[c]++] = r;
            }
    }
    std::vector<int64_t> gp(ng + 1, 0);
    std::vector<int> ga;
    {
        std::vector<int> seen(ng, -1), tmp;
        for (int g = 0; g < ng; g++) {
            tmp.clear();
            const int node = gfirst[g];
            const int vs[2] = {node, mate[node]};
            for (int q = 0; q < 2; q++) {
                const int v = vs[q];
                if (v < 0) continue;
                for (int64_t e = vp[v]; e < vp[v + 1]; e++) {
                    const int h = grp[va[e]];
                    if (h != g && seen[h] != g) seen[h] = g, tmp.push_back(h);
                }
            }
            ga.insert(ga.end(), tmp.begin(), tmp.end());
            gp[g + 1] = (int64_t)ga.size();
        }
    }

    /* nested dissection */
    Nested nd(gp, ga, ng);
    {
        std::vector<int> all(ng);
        std::iota(all.begin(), all.end(), 0);
        nd.split(std::move(all), -1);
    }
    const int nfr = (int)nd.fronts.size();

    /* postorder */
    std::vector<std::vector<int>> ch(nfr);
    std::vector<int> roots;
    for (int f = 0; f < nfr; f++) (nd.parent[f] < 0 ? roots : ch[nd.parent[f]]).push_back(f);
    std::vector<int> post;
    post.reserve(nfr);
    {
        std::vector<std::pair<int, int>> stk;
        for (int r : roots) {
            stk.push_back({r, 0});
            while (!stk.empty()) {
                auto &t = stk.back();
                if (t.second < (int)ch[t.first].size()) {
                    const int c = ch[t.first][t.second++];
                    stk.push_back({c, 0});
                } else {
                    post.push_back(t.first);
                    stk.pop_back();
                }
            }
        }
    }
    std::vector<int> newid(nfr);
    for (int i = 0; i < nfr; i++) newid[post[i]] = i;
    S.nf = nfr;
    S.ns.assign(nfr, 0);
    S.nb.assign(nfr, 0);
    S.parent.assign(nfr, -1);
    S.height.assign(nfr, 0);
    S.kid_ptr.assign(nfr + 1, 0);
    for (int i = 0; i < nfr; i++) {
        const int f = post[i];
        S.parent[i] = nd.parent[f] < 0 ? -1 : newid[nd.parent[f]];
        S.kid_ptr[i + 1] = S.kid_ptr[i] + (int)ch[f].size();
        for (int c : ch[f]) S.kids.push_back(newid[c]);
    }

    /* elimination positions */
    std::vector<int> pos(n), fof(n);
    std::vector<int> first(nfr + 1, 0);
    {
        int p = 0;
        for (int i = 0; i < nfr; i++) {
            first[i] = p;
            for (int g : nd.fronts[post[i]]) {
                const int node = gfirst[g];
                pos[node] = p++;
                fof[node] = i;
                if (mate[node] >= 0) pos[mate[node]] = p++, fof[mate[node]] = i;
            }
            S.ns[i] = p - first[i];
        }
        first[nfr] = p;
    }
    std::vector<int> byPos(n);
    for (int v = 0; v < n; v++) byPos[pos[v]] = v;

    /* boundaries, bottom-up */
    std::vector<std::vector<int>> bnd(nfr);
    {
        std::vector<int> in(n, -1);
        for (int f = 0; f < nfr; f++) {
            const int last = first[f + 1] - 1;
            std::vector<int> &B = bnd[f];
            for (int p = first[f]; p <= last; p++) {
                const int v = byPos[p];
                for (int64_t e = vp[v]; e < vp[v + 1]; e++) {
                    const int u = va[e];
                    if (pos[u] > last && in[u] != f) in[u] = f, B.push_back(u);
                }
            }
            for (int q = S.kid_ptr[f]; q < S.kid_ptr[f + 1]; q++)
                for (int u : bnd[S.kids[q]])
                    if (pos[u] > last && in[u] != f) in[u] = f, B.push_back(u);
            std::sort(B.begin(), B.end(), [&](int a, int b) { return pos[a] < pos[b]; });
            S.nb[f] = (int)B.size();
            int h = 0;
            for (int q = S.kid_ptr[f]; q < S.kid_ptr[f + 1]; q++) h = std::max(h, S.height[S.kids[q]] + 1);
            S.height[f] = h;
        }
    }

    /* index lists, offsets, relative indices into the parent */
    S.off.resize(nfr);
    S.idx_off.resize(nfr + 1);
    S.upd_off.resize(nfr + 1);
    int64_t store = 0, ioff = 0, uoff = 0;
    for (int f = 0; f < nfr; f++) {
        const int64_t nf = (int64_t)S.ns[f] + S.nb[f];
        if (nf > SPLU_MAX_FRONT) {
            set_error("splu_analyze: front of %lld unknowns exceeds %d", (long long)nf, SPLU_MAX_FRONT);
            return NGMI_E_UNSUPP;
        }
        S.max_front = std::max(S.max_front, (int)nf);
        S.off[f] = store;
        S.idx_off[f] = ioff;
        S.upd_off[f] = uoff;
        store += nf * nf;
        ioff += nf;
        uoff += S.nb[f];
        S.lu_nnz += nf * nf - (int64_t)S.nb[f] * S.nb[f];
        for (int k = 0; k < S.ns[f]; k++) {
            const double rem = (double)(nf - k - 1);
            S.flops += rem + 2.0 * rem * rem;
        }
    }
    S.idx_off[nfr] = ioff;
    S.upd_off[nfr] = uoff;
    S.store = store;
    S.idx.resize(ioff);
    S.rel.assign(ioff, -1);
    for (int f = 0; f < nfr; f++) {
        int *id = S.idx.data() + S.idx_off[f];
        for (int k = 0; k < S.ns[f]; k++) id[k] = byPos[first[f] + k];
        for (int k = 0; k < S.nb[f]; k++) id[S.ns[f] + k] = bnd[f][k];
    }
    {
        std::vector<int> loc(n, -1);
        for (int f = 0; f < nfr; f++) {
            const int nf = S.ns[f] + S.nb[f];
            const int *id = S.idx.data() + S.idx_off[f];
            for (int k = 0; k < nf; k++) loc[id[k]] = k;
            for (int q = S.kid_ptr[f]; q < S.kid_ptr[f + 1]; q++) {
                const int c = S.kids[q];
                int *rl = S.rel.data() + S.idx_off[c] + S.ns[c];
                for (int k = 0; k < S.nb[c]; k++) {
                    const int u = bnd[c][k];
                    const int l = loc[u];
                    if (l < 0 || l >= nf || S.idx[S.idx_off[f] + l] != u) {
                        set_error("splu_analyze: child boundary unknown %d not in its parent", u);
                        return NGMI_E_PANIC;
                    }
                    rl[k] = l;
                }
            }
            for (int k = 0; k < nf; k++) loc[id[k]] = -1;
        }
    }

    /* every stored entry of B goes to the front that eliminates its earlier unknown */
    S.ent_pos.resize(nnz);
    for (int c = 0; c < n; c++)
        for (int64_t e = Ap[c]; e < Ap[c + 1]; e++) {
            const int r = S.rowsrc[Ai[e]];
            const int v = pos[r] < pos[c] ? r : c;
            const int f = fof[v];
            const int ns = S.ns[f], nf = ns + S.nb[f];
            const int *id = S.idx.data() + S.idx_off[f];
            auto local = [&](int u) -> int {
                if (pos[u] >= first[f] && pos[u] < first[f + 1]) return pos[u] - first[f];
                const int *b = id + ns, *en = id + nf;
                const int *it = std::lower_bound(b, en, u, [&](int a, int x) { return pos[a] < pos[x]; });
                return (it != en && *it == u) ? (int)(it - id) : -1;
            };
            const int lr = local(r), lc = local(c);
            if (lr < 0 || lc < 0) {
                set_error("splu_analyze: entry (%d, %d) outside its front", r, c);
                return NGMI_E_PANIC;
            }
            S.ent_pos[e] = S.off[f] + (int64_t)lc * nf + lr;
        }

    /* levels by height; extend-add lists by (height of the parent, child slot) */
    S.H = 0;
    for (int f = 0; f < nfr; f++) S.H = std::max(S.H, S.height[f]);
    S.lev_ptr.assign(S.H + 2, 0);
    for (int f = 0; f < nfr; f++) S.lev_ptr[S.height[f] + 1]++;
    for (int h = 0; h <= S.H; h++) S.lev_ptr[h + 1] += S.lev_ptr[h];
    S.lev.resize(nfr);
    {
        std::vector<int> w(S.lev_ptr.begin(), S.lev_ptr.end() - 1);
        for (int f = 0; f < nfr; f++) S.lev[w[S.height[f]]++] = f;
    }
    S.max_slots = 0;
    for (int f = 0; f < nfr; f++) S.max_slots = std::max(S.max_slots, S.kid_ptr[f + 1] - S.kid_ptr[f]);
    S.ea_ptr.assign((size_t)(S.H + 1) * S.max_slots + 1, 0);
    S.ea.clear();
    for (int h = 0; h <= S.H; h++)
        for (int s = 0; s < S.max_slots; s++) {
            for (int q = S.lev_ptr[h]; q < S.lev_ptr[h + 1]; q++) {
                const int f = S.lev[q];
                if (S.kid_ptr[f] + s < S.kid_ptr[f + 1]) S.ea.push_back(S.kids[S.kid_ptr[f] + s]);
            }
            S.ea_ptr[(size_t)h * S.max_slots + s + 1] = (int)S.ea.size();
        }
    S.max_lev = 0;
    for (int h = 0; h <= S.H; h++) S.max_lev = std::max(S.max_lev, S.lev_ptr[h + 1] - S.lev_ptr[h]);
    return NGMI_OK;
}

}  // namespace ngmi
