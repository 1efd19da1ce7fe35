/*
 * KLU's symbolic analysis, restated for the engine (host code): the
 * ordering klu_analyze computes with KLU's defaults (klu_defaults.c:20-32:
 * BTF on, AMD on the blocks, no work limit), so the engine's static LU can
 * run on KLU's own pivot sequence without the reference library.
 *
 *   btf: maximum transversal (btf_maxtrans.c:1-387: depth-first augmenting
 *        paths with a "cheap" first scan per column), completion of a
 *        structurally singular matching (btf_order.c:60-120), then Tarjan's
 *        strongly connected components, non-recursive (btf_strongcomp.c:
 *        1-593), giving a block upper triangular form.
 *   amd: approximate minimum degree on each block's C + C' (amd_order.c,
 *        amd_valid.c, amd_preprocess.c, amd_aat.c, amd_1.c, amd_2.c,
 *        amd_postorder.c, amd_post_tree.c) with the defaults dense = 10
 *        sqrt(n) (at least 16) and aggressive absorption; blocks of three
 *        or fewer keep their natural order (klu_analyze.c:125-135).
 *   composition: klu_analyze.c:14-250 (analyze_worker) and 260-440.
 *
 * Every tie is broken the way the reference breaks it (list heads, scan
 * order, the hash of a variable's adjacency), because the point of the
 * restatement is the identical permutation: P, Q and R equal Symbolic->P,
 * Symbolic->Q and Symbolic->R of the reference on the same CSC pattern.
 */
#include <cmath>
#include <cstdint>
#include <vector>
#include "ngmi_internal.h"

namespace {

constexpr int NONE = -1;
inline int flip(int i) { return -i - 2; }
inline int unflip(int i) { return i < NONE ? flip(i) : i; }

/* ---------------------------------------------------------------- BTF */

/* Row i matched to column match[i] (or NONE); returns the matching size. */
int max_transversal(int n, const int *Ap, const int *Ai, std::vector<int> &match)
{
    std::vector<int> cheap(n), visited(n, NONE), rstack(n), cstack(n), pstack(n);
    for (int j = 0; j < n; j++) cheap[j] = Ap[j];
    match.assign(n, NONE);
    int matched = 0;
    for (int k = 0; k < n; k++) {
        /* depth-first search for an augmenting path from column k */
        bool found = false;
        int top = 0, row = NONE;
        cstack[0] = k;
        while (top >= 0) {
            const int j = cstack[top];
            const int end = Ap[j + 1];
            if (visited[j] != k) {
                visited[j] = k;
                /* first visit: an unmatched row in what is left of column j */
                int p = cheap[j];
                while (p < end && !found) {
                    row = Ai[p];
                    found = match[row] == NONE;
                    p++;
                }
                cheap[j] = p;
                if (found) {
                    rstack[top] = row;
                    break;
                }
                pstack[top] = Ap[j];
            }
            /* descend into the column matched to the next row of column j
             * not yet visited in this search */
            int p = pstack[top];
            for (; p < end; p++) {
                const int i = Ai[p];
                const int j2 = match[i];
                if (visited[j2] != k) {
                    pstack[top] = p + 1;
                    rstack[top] = i;
                    cstack[++top] = j2;
                    break;
                }
            }
            if (p == end) top--;
        }
        if (found) {
            for (int q = top; q >= 0; q--) match[rstack[q]] = cstack[q];
            matched++;
        }
    }
    return matched;
}

/* strongly connected components of the graph of A(:, Q): node j has edges
 * to the rows of column unflip(Q[j]).  P lists the nodes block by block,
 * ascending within a block; Q becomes Q(P); R holds the block starts. */
int strong_components(int n, const int *Ap, const int *Ai, std::vector<int> &Q, std::vector<int> &P,
                      std::vector<int> &R)
{
    constexpr int UNSEEN = -2, OPEN = -1;
    std::vector<int> when(n, NONE), low(n, NONE), comp(n, UNSEEN), ostack(n + 1), jstack(n), pstack(n);
    int clock = 0, nblocks = 0;
    for (int root = 0; root < n; root++) {
        if (comp[root] != UNSEEN) continue;
        int otop = 0, jtop = 0;
        jstack[0] = root;
        while (jtop >= 0) {
            const int j = jstack[jtop];
            const int col = unflip(Q[j]);
            const int end = Ap[col + 1];
            if (comp[j] == UNSEEN) {
                ostack[++otop] = j;
                when[j] = low[j] = ++clock;
                comp[j] = OPEN;
                pstack[jtop] = Ap[col];
            }
            int p = pstack[jtop];
            for (; p < end; p++) {
                const int i = Ai[p];
                if (comp[i] == UNSEEN) {
                    pstack[jtop] = p + 1;
                    jstack[++jtop] = i;
                    break;
                }
                if (comp[i] == OPEN && when[i] < low[j]) low[j] = when[i];
            }
            if (p == end) {
                jtop--;
                if (low[j] == when[j]) {
                    for (;;) {
                        const int i = ostack[otop--];
                        comp[i] = nblocks;
                        if (i == j) break;
                    }
                    nblocks++;
                }
                if (jtop >= 0) {
                    const int parent = jstack[jtop];
                    if (low[j] < low[parent]) low[parent] = low[j];
                }
            }
        }
    }
    R.assign(nblocks + 1, 0);
    for (int j = 0; j < n; j++) R[comp[j] + 1]++;
    for (int b = 0; b < nblocks; b++) R[b + 1] += R[b];
    std::vector<int> next(R.begin(), R.end() - 1);
    P.assign(n, 0);
    for (int j = 0; j < n; j++) P[next[comp[j]]++] = j;
    std::vector<int> q2(n);
    for (int k = 0; k < n; k++) q2[k] = Q[P[k]];
    Q.swap(q2);
    return nblocks;
}

/* btf_order: returns the number of blocks; *rank = matching size */
int btf(int n, const int *Ap, const int *Ai, std::vector<int> &P, std::vector<int> &Q, std::vector<int> &R,
        int *rank)
{
    *rank = max_transversal(n, Ap, Ai, Q);
    if (*rank < n) {
        /* pair the unmatched rows with the unmatched columns, smallest
         * first, flagging those (zero) diagonal entries */
        std::vector<char> used(n, 0);
        for (int i = 0; i < n; i++)
            if (Q[i] != NONE) used[Q[i]] = 1;
        std::vector<int> free_cols;
        for (int j = n - 1; j >= 0; j--)
            if (!used[j]) free_cols.push_back(j);
        for (int i = 0; i < n; i++) {
            if (Q[i] == NONE && !free_cols.empty()) {
                Q[i] = flip(free_cols.back());
                free_cols.pop_back();
            }
        }
    }
    return strong_components(n, Ap, Ai, Q, P, R);
}

/* ---------------------------------------------------------------- AMD */

/* the minimum degree elimination on the pattern of A + A' in iw (variable
 * j's list at pe[j], len[j] entries) with elbow room up to iwlen; last
 * receives the permutation (last[k] = the k-th eliminated node) */
void amd_eliminate(int n, std::vector<int> &pe, std::vector<int> &iw, std::vector<int> &len, int iwlen, int pfree,
                   std::vector<int> &last)
{
    std::vector<int> nv(n, 1), next(n, NONE), head(n, NONE), elen(n, 0), degree(n), w(n, 1);
    last.assign(n, NONE);
    for (int i = 0; i < n; i++) degree[i] = len[i];

    const int wbig = INT32_MAX - n;
    auto reset_marks = [&](int wflg) {
        if (wflg < 2 || wflg >= wbig) {
            for (int x = 0; x < n; x++)
                if (w[x] != 0) w[x] = 1;
            wflg = 2;
        }
        return wflg;
    };
    /* unlink variable i from its degree list */
    auto unlink = [&](int i) {
        const int il = last[i], in = next[i];
        if (in != NONE) last[in] = il;
        if (il != NONE) next[il] = in;
        else head[degree[i]] = in;
    };

    int dense = (int)(10.0 * std::sqrt((double)n));
    if (dense < 16) dense = 16;
    if (dense > n) dense = n;
    int wflg = reset_marks(0);
    int nel = 0, mindeg = 0, lemax = 0;

    /* nodes of degree 0 are eliminated at once, dense ones ordered last,
     * the rest go on the degree lists (each list head-inserted) */
    for (int i = 0; i < n; i++) {
        const int deg = degree[i];
        if (deg == 0) {
            elen[i] = flip(1);
            nel++;
            pe[i] = NONE;
            w[i] = 0;
        } else if (deg > dense) {
            nv[i] = 0;
            elen[i] = NONE;
            nel++;
            pe[i] = NONE;
        } else {
            const int in = head[deg];
            if (in != NONE) last[in] = i;
            next[i] = in;
            head[deg] = i;
        }
    }

    while (nel < n) {
        /* the pivot: head of the lowest non-empty degree list */
        int deg = mindeg;
        int me = NONE;
        for (; deg < n; deg++) {
            me = head[deg];
            if (me != NONE) break;
        }
        mindeg = deg;
        {
            const int in = next[me];
            if (in != NONE) last[in] = NONE;
            head[deg] = in;
        }
        const int elenme = elen[me];
        int nvpiv = nv[me];
        nel += nvpiv;
        nv[me] = -nvpiv;
        int degme = 0;
        int pme1, pme2;

        /* construct the new element's variable list */
        if (elenme == 0) {
            /* in place: me's list minus eliminated and flagged entries */
            pme1 = pe[me];
            pme2 = pme1 - 1;
            for (int p = pme1; p <= pme1 + len[me] - 1; p++) {
                const int i = iw[p];
                const int nvi = nv[i];
                if (nvi > 0) {
                    degme += nvi;
                    nv[i] = -nvi;
                    iw[++pme2] = i;
                    unlink(i);
                }
            }
        } else {
            /* at the end of iw: the union of me's elements and variables */
            int p = pe[me];
            pme1 = pfree;
            const int slenme = len[me] - elenme;
            for (int k1 = 1; k1 <= elenme + 1; k1++) {
                int e, pj, ln;
                if (k1 > elenme) {
                    e = me;
                    pj = p;
                    ln = slenme;
                } else {
                    e = iw[p++];
                    pj = pe[e];
                    ln = len[e];
                }
                for (int k2 = 1; k2 <= ln; k2++) {
                    const int i = iw[pj++];
                    const int nvi = nv[i];
                    if (nvi <= 0) continue;
                    if (pfree >= iwlen) {
                        /* compress iw: lists keep their order; each list
                         * start is marked by its owner's flipped index */
                        pe[me] = p;
                        len[me] -= k1;
                        if (len[me] == 0) pe[me] = NONE;
                        pe[e] = pj;
                        len[e] = ln - k2;
                        if (len[e] == 0) pe[e] = NONE;
                        for (int j = 0; j < n; j++) {
                            const int pn = pe[j];
                            if (pn >= 0) {
                                pe[j] = iw[pn];
                                iw[pn] = flip(j);
                            }
                        }
                        int psrc = 0, pdst = 0;
                        const int pend = pme1 - 1;
                        while (psrc <= pend) {
                            const int j = flip(iw[psrc++]);
                            if (j >= 0) {
                                iw[pdst] = pe[j];
                                pe[j] = pdst++;
                                for (int k3 = 0; k3 <= len[j] - 2; k3++) iw[pdst++] = iw[psrc++];
                            }
                        }
                        const int p1 = pdst;
                        for (psrc = pme1; psrc <= pfree - 1; psrc++) iw[pdst++] = iw[psrc];
                        pme1 = p1;
                        pfree = pdst;
                        pj = pe[e];
                        p = pe[me];
                    }
                    degme += nvi;
                    nv[i] = -nvi;
                    iw[pfree++] = i;
                    unlink(i);
                }
                if (e != me) {
                    /* element e is absorbed into me */
                    pe[e] = flip(me);
                    w[e] = 0;
                }
            }
            pme2 = pfree - 1;
        }
        degree[me] = degme;
        pe[me] = pme1;
        len[me] = pme2 - pme1 + 1;
        elen[me] = flip(nvpiv + degme);
        wflg = reset_marks(wflg);

        /* |Le \ Lme| for every element e adjacent to me's variables */
        for (int pme = pme1; pme <= pme2; pme++) {
            const int i = iw[pme];
            const int eln = elen[i];
            if (eln <= 0) continue;
            const int nvi = -nv[i];
            const int wnvi = wflg - nvi;
            for (int p = pe[i]; p <= pe[i] + eln - 1; p++) {
                const int e = iw[p];
                int we = w[e];
                if (we >= wflg) we -= nvi;
                else if (we != 0) we = degree[e] + wnvi;
                w[e] = we;
            }
        }

        /* approximate degrees, aggressive absorption, hashes */
        for (int pme = pme1; pme <= pme2; pme++) {
            int i = iw[pme];
            const int p1 = pe[i];
            const int p2 = p1 + elen[i] - 1;
            int pn = p1;
            uint32_t hash = 0;
            int dg = 0;
            for (int p = p1; p <= p2; p++) {
                const int e = iw[p];
                const int we = w[e];
                if (we == 0) continue;
                const int dext = we - wflg;
                if (dext > 0) {
                    dg += dext;
                    iw[pn++] = e;
                    hash += (uint32_t)e;
                } else {
                    pe[e] = flip(me);   /* e is a subset of me */
                    w[e] = 0;
                }
            }
            elen[i] = pn - p1 + 1;
            const int p3 = pn;
            const int p4 = p1 + len[i];
            for (int p = p2 + 1; p < p4; p++) {
                const int j = iw[p];
                const int nvj = nv[j];
                if (nvj > 0) {
                    dg += nvj;
                    iw[pn++] = j;
                    hash += (uint32_t)j;
                }
            }
            if (elen[i] == 1 && p3 == pn) {
                /* mass elimination: i is adjacent to me only */
                pe[i] = flip(me);
                const int nvi = -nv[i];
                degme -= nvi;
                nvpiv += nvi;
                nel += nvi;
                nv[i] = 0;
                elen[i] = NONE;
            } else {
                if (dg < degree[i]) degree[i] = dg;
                /* me goes first in i's element list */
                iw[pn] = iw[p3];
                iw[p3] = iw[p1];
                iw[p1] = me;
                len[i] = pn - p1 + 1;
                const int hb = (int)(hash % (uint32_t)n);
                const int j = head[hb];
                if (j <= NONE) {
                    /* degree list hb empty: the hash bucket lives in head */
                    next[i] = flip(j);
                    head[hb] = flip(i);
                } else {
                    /* the bucket lives in last[] of the list's first node */
                    next[i] = last[j];
                    last[j] = i;
                }
                last[i] = hb;
            }
        }
        degree[me] = degme;
        if (degme > lemax) lemax = degme;
        wflg += lemax;
        wflg = reset_marks(wflg);

        /* supervariable detection: equal lists within a hash bucket */
        for (int pme = pme1; pme <= pme2; pme++) {
            int i = iw[pme];
            if (nv[i] >= 0) continue;
            const int hb = last[i];
            const int j0 = head[hb];
            if (j0 == NONE) {
                i = NONE;
            } else if (j0 < NONE) {
                i = flip(j0);
                head[hb] = NONE;
            } else {
                i = last[j0];
                last[j0] = NONE;
            }
            while (i != NONE && next[i] != NONE) {
                const int ln = len[i];
                const int eln = elen[i];
                for (int p = pe[i] + 1; p <= pe[i] + ln - 1; p++) w[iw[p]] = wflg;
                int jlast = i;
                int j = next[i];
                while (j != NONE) {
                    bool same = len[j] == ln && elen[j] == eln;
                    for (int p = pe[j] + 1; same && p <= pe[j] + ln - 1; p++)
                        if (w[iw[p]] != wflg) same = false;
                    if (same) {
                        /* j joins supervariable i */
                        pe[j] = flip(i);
                        nv[i] += nv[j];
                        nv[j] = 0;
                        elen[j] = NONE;
                        j = next[j];
                        next[jlast] = j;
                    } else {
                        jlast = j;
                        j = next[j];
                    }
                }
                wflg++;
                i = next[i];
            }
        }

        /* finalize degrees, restore the degree lists, compact me's list */
        int p = pme1;
        const int nleft = n - nel;
        for (int pme = pme1; pme <= pme2; pme++) {
            const int i = iw[pme];
            const int nvi = -nv[i];
            if (nvi <= 0) continue;
            nv[i] = nvi;
            int dg = degree[i] + degme - nvi;
            if (dg > nleft - nvi) dg = nleft - nvi;
            const int in = head[dg];
            if (in != NONE) last[in] = i;
            next[i] = in;
            last[i] = NONE;
            head[dg] = i;
            if (dg < mindeg) mindeg = dg;
            degree[i] = dg;
            iw[p++] = i;
        }
        nv[me] = nvpiv;
        len[me] = p - pme1;
        if (len[me] == 0) {
            pe[me] = NONE;
            w[me] = 0;
        }
        if (elenme != 0) pfree = p;
    }

    /* the assembly tree: pe[] becomes the parent of each element and of
     * each absorbed variable's principal element */
    for (int i = 0; i < n; i++) pe[i] = flip(pe[i]);
    for (int i = 0; i < n; i++) elen[i] = flip(elen[i]);
    for (int i = 0; i < n; i++) {
        if (nv[i] != 0) continue;
        int j = pe[i];
        if (j == NONE) continue;
        while (nv[j] == 0) j = pe[j];
        const int e = j;
        j = i;
        while (nv[j] == 0) {
            const int jn = pe[j];
            pe[j] = e;
            j = jn;
        }
    }

    /* postorder of the elements (amd_postorder.c): children by ascending
     * index, the child with the largest front (the last of equal ones)
     * moved to the end of its list, depth-first from each root */
    std::vector<int> child(n, NONE), sibling(n, NONE), order(n, NONE), stack(n);
    for (int j = n - 1; j >= 0; j--) {
        if (nv[j] > 0 && pe[j] != NONE) {
            sibling[j] = child[pe[j]];
            child[pe[j]] = j;
        }
    }
    for (int i = 0; i < n; i++) {
        if (nv[i] <= 0 || child[i] == NONE) continue;
        int fprev = NONE, maxfr = NONE, bigprev = NONE, big = NONE;
        for (int f = child[i]; f != NONE; f = sibling[f]) {
            if (elen[f] >= maxfr) {
                maxfr = elen[f];
                bigprev = fprev;
                big = f;
            }
            fprev = f;
        }
        const int fnext = sibling[big];
        if (fnext != NONE) {
            if (bigprev == NONE) child[i] = fnext;
            else sibling[bigprev] = fnext;
            sibling[big] = NONE;
            sibling[fprev] = big;
        }
    }
    int k = 0;
    for (int root = 0; root < n; root++) {
        if (pe[root] != NONE || nv[root] <= 0) continue;
        int top = 0;
        stack[0] = root;
        while (top >= 0) {
            const int i = stack[top];
            if (child[i] != NONE) {
                for (int f = child[i]; f != NONE; f = sibling[f]) top++;
                int h = top;
                for (int f = child[i]; f != NONE; f = sibling[f]) stack[h--] = f;
                child[i] = NONE;
            } else {
                top--;
                order[i] = k++;
            }
        }
    }

    /* elements in postorder, each followed by the variables it absorbed;
     * dense nodes last */
    std::vector<int> first(n, NONE), pos(n, NONE);
    for (int e = 0; e < n; e++)
        if (order[e] != NONE) first[order[e]] = e;
    nel = 0;
    for (int q = 0; q < n; q++) {
        const int e = first[q];
        if (e == NONE) break;
        pos[e] = nel;
        nel += nv[e];
    }
    for (int i = 0; i < n; i++) {
        if (nv[i] != 0) continue;
        const int e = pe[i];
        if (e != NONE) pos[i] = pos[e]++;
        else pos[i] = nel++;
    }
    for (int i = 0; i < n; i++) last[pos[i]] = i;
}

/* AMD on the pattern of the n x n CSC matrix (Ap, Ai), diagonal ignored;
 * perm[k] = the k-th pivot */
void amd(int n, const int *Ap, const int *Ai, std::vector<int> &perm)
{
    perm.assign(n, 0);
    if (n == 0) return;
    /* sorted, duplicate-free columns, or else the pattern is replaced by
     * its transpose with duplicates dropped (which sorts it) */
    bool jumbled = false;
    for (int j = 0; j < n && !jumbled; j++) {
        int prev = NONE;
        for (int p = Ap[j]; p < Ap[j + 1]; p++) {
            if (Ai[p] <= prev) {
                jumbled = true;
                break;
            }
            prev = Ai[p];
        }
    }
    std::vector<int> tp, ti;
    const int *cp = Ap, *ci = Ai;
    if (jumbled) {
        std::vector<int> cnt(n, 0), mark(n, NONE);
        for (int j = 0; j < n; j++)
            for (int p = Ap[j]; p < Ap[j + 1]; p++)
                if (mark[Ai[p]] != j) {
                    cnt[Ai[p]]++;
                    mark[Ai[p]] = j;
                }
        tp.assign(n + 1, 0);
        for (int i = 0; i < n; i++) tp[i + 1] = tp[i] + cnt[i];
        ti.assign(std::max(tp[n], 1), 0);
        std::vector<int> fill(tp.begin(), tp.end() - 1);
        mark.assign(n, NONE);
        for (int j = 0; j < n; j++)
            for (int p = Ap[j]; p < Ap[j + 1]; p++)
                if (mark[Ai[p]] != j) {
                    ti[fill[Ai[p]]++] = j;
                    mark[Ai[p]] = j;
                }
        cp = tp.data();
        ci = ti.data();
    }

    /* the lengths of A + A' without the diagonal (amd_aat.c): column k's
     * upper part, merged with the not yet visited lower entries of the
     * columns it points to */
    std::vector<int> len(n, 0), t(n, NONE);
    auto pass = [&](auto &&emit) {
        for (int k = 0; k < n; k++) t[k] = NONE;
        for (int k = 0; k < n; k++) {
            int p = cp[k];
            const int p2 = cp[k + 1];
            while (p < p2) {
                const int j = ci[p];
                if (j < k) {
                    emit(j, k);
                    p++;
                } else {
                    if (j == k) p++;
                    break;
                }
                int pj = t[j] == NONE ? cp[j] : t[j];
                const int pj2 = cp[j + 1];
                while (pj < pj2) {
                    const int i = ci[pj];
                    if (i < k) {
                        emit(i, j);
                        pj++;
                    } else {
                        if (i == k) pj++;
                        break;
                    }
                }
                t[j] = pj;
            }
            t[k] = p;
        }
        for (int j = 0; j < n; j++)
            for (int pj = (t[j] == NONE ? cp[j] : t[j]); pj < cp[j + 1]; pj++) emit(ci[pj], j);
    };
    pass([&](int a, int b) {
        len[a]++;
        len[b]++;
    });
    size_t nzaat = 0;
    for (int k = 0; k < n; k++) nzaat += (size_t)len[k];
    /* the reference sizes its workspace as nzaat + nzaat / 5 + 7 n, of which
     * 6 n are the other arrays; the rest is iw, the elbow room of the
     * elimination (its compressions happen at the same moments) */
    const size_t slen = nzaat + nzaat / 5 + 7 * (size_t)n;
    const int iwlen = (int)(slen - 6 * (size_t)n);
    std::vector<int> pe(n), sp(n), iw(iwlen, NONE);
    int pfree = 0;
    for (int j = 0; j < n; j++) {
        pe[j] = sp[j] = pfree;
        pfree += len[j];
    }
    /* the lists (amd_1.c): the same merge, appending both directions */
    pass([&](int a, int b) {
        iw[sp[a]++] = b;
        iw[sp[b]++] = a;
    });
    amd_eliminate(n, pe, iw, len, iwlen, pfree, perm);
}

}  // namespace

extern "C" {

int ngmi_klu_analyze(int n, const int *Ap, const int *Ai, int *P, int *Q, int *R, int *nblocks, int *nzoff,
                     int *structural_rank)
{
    if (n < 0 || (n > 0 && (!Ap || !Ai || !P || !Q || !R))) {
        ngmi::set_error("klu_analyze: bad arguments");
        return NGMI_E_PARMVAL;
    }
    if (n > 0 && Ap[0] != 0) {
        ngmi::set_error("klu_analyze: Ap[0] must be 0");
        return NGMI_E_PARMVAL;
    }
    for (int j = 0; j < n; j++) {
        if (Ap[j + 1] < Ap[j]) {
            ngmi::set_error("klu_analyze: column pointers decrease at %d", j);
            return NGMI_E_PARMVAL;
        }
        for (int p = Ap[j]; p < Ap[j + 1]; p++)
            if (Ai[p] < 0 || Ai[p] >= n) {
                ngmi::set_error("klu_analyze: row index %d out of range in column %d", Ai[p], j);
                return NGMI_E_PARMVAL;
            }
    }
    std::vector<int> pb, qb, rb;
    int rank = 0;
    const int nb = btf(n, Ap, Ai, pb, qb, rb, &rank);
    if (rank < n)
        for (int k = 0; k < n; k++) qb[k] = unflip(qb[k]);
    std::vector<int> pinv(n);
    for (int k = 0; k < n; k++) pinv[pb[k]] = k;
    int off = 0;
    std::vector<int> cp, ci, pblk;
    for (int b = 0; b < nb; b++) {
        const int k1 = rb[b], k2 = rb[b + 1], nk = k2 - k1;
        /* block b of A(P, Q); entries above it are off-diagonal */
        cp.assign(nk + 1, 0);
        ci.clear();
        for (int k = k1; k < k2; k++) {
            cp[k - k1] = (int)ci.size();
            const int col = qb[k];
            for (int p = Ap[col]; p < Ap[col + 1]; p++) {
                const int r = pinv[Ai[p]];
                if (r < k1) off++;
                else ci.push_back(r - k1);
            }
        }
        cp[nk] = (int)ci.size();
        if (nk <= 3) {
            pblk.resize(nk);
            for (int k = 0; k < nk; k++) pblk[k] = k;
        } else {
            amd(nk, cp.data(), ci.data(), pblk);
        }
        for (int k = 0; k < nk; k++) {
            Q[k1 + k] = qb[k1 + pblk[k]];
            P[k1 + k] = pb[k1 + pblk[k]];
        }
    }
    for (int b = 0; b <= nb; b++) R[b] = rb[b];
    if (nblocks) *nblocks = nb;
    if (nzoff) *nzoff = off;
    if (structural_rank) *structural_rank = rank;
    return NGMI_OK;
}

}  // extern "C"
