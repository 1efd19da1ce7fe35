/*
 * Host side of the batched transient engine: circuit construction (the
 * engine's counterpart of the netlist -> CKTcircuit build), the static
 * symbolic analysis shared by every circuit of a batch, and the driver loop
 * that launches one Newton pass at a time until every circuit has reached
 * its final time.
 *
 * Symbolic analysis (the role of SMPpreOrder / klu_analyze, klusmp.c:1255,
 * and of the pivot sequence klu_factor fixes for klu_refactor,
 * klu_refactor.c:18-482):
 *   1. unknowns 1..NN are node voltages, NN+1.. the voltage-source branch
 *      currents in source order (the order CKTmkCur gives them);
 *   2. a static row permutation puts each source's branch equation on the
 *      row of its positive node (and that node's KCL on the branch row), so
 *      both diagonals are the exact +-1 entries of the source stamp;
 *   3. a minimum-degree symmetric ordering of the permuted pattern limits
 *      fill; the remaining pivots are the (permuted) diagonal;
 *   4. the filled L+U pattern gives an entry-wise Doolittle program grouped
 *      into dependency levels, executed by one workgroup per circuit.
 * The pivot sequence is fixed for the whole run and for every circuit of the
 * batch; a numerically zero pivot is reported as E_SINGULAR for that
 * circuit's Newton iteration.
 */
#include <algorithm>
#include <tuple>
#include <array>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>
#include "ngmi_internal.h"
#include "tran_types.h"

namespace ngmi {
int launch_tran_pass(const B4Tab *tab0, const TranDev *D0, int c_begin, int c_end, const B4Tab *tabs,
                     const TranDev *Ds, const int2 *mos_blk, int n_mos_blk, const int2 *step_blk,
                     int n_step_blk, size_t lds_bytes, hipStream_t s, hipEvent_t *ev);
int set_tran_lds_limit(size_t bytes);
int launch_zero_int(int *p, hipStream_t s);
int launch_guess_apply(const TranDev &D, const float *y, const int64_t *dest, const double2 *base,
                       const int *count, int maxrows, unsigned long long *rows_total, hipStream_t s);
int mlp_forward_rows(ngmi_mlp *n, const float *X, float *y, int maxRows, const int *dcount,
                     hipStream_t s);
int mlp_guess_rows(const ngmi_mlp *n, const float *X, int *count, int maxRows, const int64_t *dest,
                   const double2 *base, double *x, float *yout, unsigned long long *rows_total, hipStream_t s);
const B4Tab *set_table(const ngmi_bsim4_set *s);
}

using namespace ngmi;

/* assembly leaves sum at most ASM_LEAF contributions unless a heavy entry
 * would need more than ASM_MAXPART partial sums */
constexpr int ASM_LEAF = 16;
constexpr int ASM_MAXPART = 48;
/* LU and substitution program records: PROG_INTS ints, of which the last
 * 4 * PROG_INLINE are the inline term groups */
constexpr int PROG_INLINE = 3;
constexpr int PROG_INTS = 4 + 4 * PROG_INLINE;
/* record kinds (bits 20-21 of the first int): a plain entry, a partial sum
 * of one chunk of a long entry's terms (into a partial slot), and the
 * combine that adds a long entry's partial sums */
constexpr int PROG_PLAIN = 0, PROG_PARTIAL = 1, PROG_COMBINE = 2;
/* entries with more terms run as partial sums over several lanes */
constexpr int SPLIT_TERMS = 24;
/* dense workspaces for KLU's pivoting: at most PV_POOL_SLOTS per topology
 * within PV_POOL_BYTES (one at least), none when one would exceed
 * PV_SLOT_MAX_BYTES */
constexpr int PV_POOL_SLOTS = 256;
constexpr int64_t PV_POOL_BYTES = 512ll << 20;
constexpr int64_t PV_SLOT_MAX_BYTES = 1ll << 30;

#include "circuit_types.h"

/* circuits of one topology inside a batch: shared symbolic analysis and
 * BSIM4 table, per-circuit device arrays */
struct TopoGroup {
    int C = 0, N = 0, NN = 0, NM = 0, NCAP = 0, NRES = 0, NVS = 0;
    std::vector<int> ids;               /* batch index of each circuit of the group */
    ngmi_bsim4_set *set = nullptr;
    TranDev D{};
    size_t lds = 0;
    std::vector<int> probe;
    /* host copies of the program statistics */
    int nslots = 0, nterms = 0, nlev_f = 0, nlev_fw = 0, nlev_bw = 0;
    std::vector<char> ic_nodes_host;
    std::vector<int> nbr_host;
    int method = 1, maxord = 2, noopiter = 0, uic = 0;
    ~TopoGroup()
    {
        if (set) ngmi_bsim4_set_free(set);
    }
};

struct ngmi_batch {
    int C = 0;
    std::vector<TopoGroup *> groups;
    std::vector<int2> where;            /* circuit -> (group, index in the group) */
    std::vector<void *> allocs;
    int64_t cap_pts = 0;
    /* one fused launch per pass: group descriptors and block -> group maps */
    TranDev *dD = nullptr;
    B4Tab *dTab = nullptr;
    int2 *mos_blk = nullptr, *step_blk = nullptr;
    int n_mos_blk = 0, n_step_blk = 0;
    size_t lds = 0;
    int *done_count = nullptr;
    /* kernel timing of the last run (events on the launch stream) */
    bool profile = false;
    double mos_ms = 0, step_ms = 0, guess_ms = 0;
    /* learned initial guess (single-topology batches) */
    ngmi_mlp *net = nullptr;
    float *gF = nullptr, *gy = nullptr;
    int64_t *gdest = nullptr;   /* per row: the node's offset in the iterate array */
    int *gcount = nullptr;      /* per half of the circuits: rows this pass, the fused kernel's finish counter */
    double2 *gbase = nullptr;   /* per row: linear extrapolation, activity */
    unsigned long long *grows = nullptr;   /* network rows evaluated, summed over passes */
    unsigned long long counters[2] = {0, 0};
    /* live data stream: watched circuits and the points already handed out */
    ngmi_data_cb data_cb = nullptr;
    void *data_user = nullptr;
    std::vector<int> watch;
    std::vector<int64_t> sent;
    ~ngmi_batch()
    {
        for (void *p : allocs) (void)hipFree(p);
        for (TopoGroup *g : groups) delete g;
    }
    const TopoGroup &group_of(int c) const { return *groups[where[c].x]; }
    const CktCtl *ctl_of(int c) const { return groups[where[c].x]->D.ctl + where[c].y; }
};

namespace {

template <class T>
int upload(ngmi_batch *b, const std::vector<T> &h, T **d)
{
    const size_t bytes = std::max<size_t>(h.size(), 1) * sizeof(T);
    void *p = nullptr;
    NGMI_HIP(hipMalloc(&p, bytes));
    b->allocs.push_back(p);
    if (!h.empty()) NGMI_HIP(hipMemcpy(p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    *d = reinterpret_cast<T *>(p);
    return NGMI_OK;
}

template <class T>
int alloc(ngmi_batch *b, size_t count, T **d)
{
    void *p = nullptr;
    NGMI_HIP(hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)));
    NGMI_HIP(hipMemset(p, 0, std::max<size_t>(count, 1) * sizeof(T)));
    b->allocs.push_back(p);
    *d = reinterpret_cast<T *>(p);
    return NGMI_OK;
}

bool same_topology(const ngmi_circuit &a, const ngmi_circuit &b)
{
    if (a.nnodes != b.nnodes || a.mos.size() != b.mos.size() || a.caps.size() != b.caps.size() ||
        a.res.size() != b.res.size() || a.vs.size() != b.vs.size())
        return false;
    for (size_t k = 0; k < a.mos.size(); k++)
        for (int r = 0; r < 4; r++)
            if (a.mos[k].node[r] != b.mos[k].node[r]) return false;
    for (size_t k = 0; k < a.caps.size(); k++)
        if (a.caps[k][0] != b.caps[k][0] || a.caps[k][1] != b.caps[k][1]) return false;
    for (size_t k = 0; k < a.res.size(); k++)
        if (a.res[k][0] != b.res[k][0] || a.res[k][1] != b.res[k][1]) return false;
    for (size_t k = 0; k < a.vs.size(); k++)
        if (a.vs[k].p != b.vs[k].p || a.vs[k].n != b.vs[k].n) return false;
    for (int u = 0; u <= a.nnodes; u++)
        if (a.ic_given[u] != b.ic_given[u]) return false;
    /* a group runs on its first circuit's options and pivot sequence */
    if (a.reltol != b.reltol || a.abstol != b.abstol || a.vntol != b.vntol || a.chgtol != b.chgtol ||
        a.trtol != b.trtol || a.gmin != b.gmin || a.method != b.method || a.maxord != b.maxord ||
        a.pivrel != b.pivrel || a.opstep != b.opstep || a.opfinal != b.opfinal || a.uic != b.uic ||
        a.noopiter != b.noopiter || a.gminsteps != b.gminsteps ||
        a.srcsteps != b.srcsteps || a.itl2 != b.itl2 || a.gminfactor != b.gminfactor || a.gshunt != b.gshunt)
        return false;
    return a.klu_order == b.klu_order && a.order_rows == b.order_rows && a.order_cols == b.order_cols;
}

/* one matrix contribution: original (row, col) 1-based, code */
struct Elt {
    int r, c;
    uint32_t code;
};

uint32_t mk(uint32_t type, uint32_t idx, bool neg) { return (type << CT_TYPE_SHIFT) | (neg ? CT_NEG : 0u) | idx; }

/* greedy minimum degree on an undirected graph (adjacency bitsets), steered
 * towards a short elimination tree.  Each node carries the level at which
 * it can be eliminated: one more than that of its latest eliminated
 * neighbour.  Among the nodes whose degree is within `slack` of the
 * minimum, the one of lowest level goes next (then lowest degree, then
 * lowest index).  The level-scheduled factor and solves take one
 * barrier-separated step per level, so a chain of nodes (a ring, a carry
 * chain) is better eliminated from many points at once than end to end,
 * which plain minimum degree does at no lower fill. */
std::vector<int> min_degree(int n, std::vector<std::vector<uint64_t>> adj, int slack)
{
    const int W = (n + 63) / 64;
    std::vector<char> gone(n, 0);
    std::vector<int> order, lev(n, 0), deg(n, 0);
    order.reserve(n);
    for (int step = 0; step < n; step++) {
        int md = 1 << 30;
        for (int v = 0; v < n; v++) {
            if (gone[v]) continue;
            int d = 0;
            for (int w = 0; w < W; w++) d += __builtin_popcountll(adj[v][w]);
            deg[v] = d;
            md = std::min(md, d);
        }
        int best = -1;
        for (int v = 0; v < n; v++) {
            if (gone[v] || deg[v] > md + slack) continue;
            if (best < 0 || lev[v] < lev[best] || (lev[v] == lev[best] && deg[v] < deg[best])) best = v;
        }
        const int v = best;
        gone[v] = 1;
        order.push_back(v);
        std::vector<int> nb;
        for (int w = 0; w < W; w++) {
            uint64_t m = adj[v][w];
            while (m) {
                const int bit = __builtin_ctzll(m);
                m &= m - 1;
                nb.push_back(w * 64 + bit);
            }
        }
        for (int a : nb) {
            lev[a] = std::max(lev[a], lev[v] + 1);
            adj[a][v >> 6] &= ~(1ull << (v & 63));
            for (int b2 : nb)
                if (b2 != a) adj[a][b2 >> 6] |= 1ull << (b2 & 63);
        }
    }
    return order;
}

/* host-side symbolic analysis of one topology, shared by every circuit of
 * a batch: contributions, pivot sequence, filled L+U pattern, the factor and
 * solve level programs, assembly plans and the neighbour table */
struct Plan {
    std::vector<char> ic_nodes;
    std::vector<int> rowp, colq;   /* step k pivots on (rowp[k], colq[k]), 1-based */
    int nslots = 0, nnz_l = 0, nterms = 0, nlev_f = 0, nlev_fw = 0, nlev_bw = 0;
    std::vector<int> f_lev, f_ent, fw_lev, fw_ent, bw_lev, bw_ent;
    std::vector<uint32_t> f_term, fw_term, bw_term;   /* packed index pairs, whole uint4 groups */
    std::vector<uint32_t> a_code, r_code;
    std::vector<int4> a_leaf, a_comb, r_leaf, r_comb;
    int n_aleaf = 0, n_acomb = 0, n_rleaf = 0, n_rcomb = 0, n_apart = 0, n_rpart = 0, n_icrow = 0;
    int n_aleaf1 = 0, n_rleaf1 = 0, n_aplain = 0, n_rplain = 0;
    int n_ppart = 0;   /* partial slots the programs use (two halves) */
    std::vector<int> a_dest, r_dest;
    std::vector<int> perm, unk_pos, ic_row, nbr;
    std::vector<int> mos_pos, rmos_pos, diag_slot;
    std::vector<int> row_start, slot_col;   /* the slots row by row (row-major) */
    int NA = 0, NR = 0;
};

int analyze(const ngmi_circuit &c0, Plan &P)
{
    auto fail = [](const char *what) {
        set_error("%s", what);
        return NGMI_E_UNSUPP;
    };
    const int NN = c0.nnodes, NVS = (int)c0.vs.size(), N = NN + NVS;
    const int NM = (int)c0.mos.size(), NCAP = (int)c0.caps.size(), NRES = (int)c0.res.size();
    /* ---- matrix and RHS contributions (original rows/cols, 1-based) */
    std::vector<Elt> elts;
    std::vector<std::vector<uint32_t>> rhs_codes(N + 1);
    static const int rr[16] = {0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 2, 2, 3, 3, 3, 3};
    static const int cc[16] = {0, 1, 2, 3, 0, 1, 2, 3, 0, 1, 2, 3, 0, 1, 2, 3};
    for (int j = 0; j < NM; j++) {
        const int *nd = c0.mos[j].node;
        for (int s = 0; s < 16; s++) {
            const int r = nd[rr[s]], col = nd[cc[s]];
            if (r && col) elts.push_back({r, col, mk(CT_MOS, (uint32_t)(s * NM + j), false)});
        }
        for (int r = 0; r < 4; r++)
            if (nd[r]) rhs_codes[nd[r]].push_back(mk(CT_MOS, (uint32_t)(r * NM + j), false));
    }
    for (int k = 0; k < NCAP; k++) {
        const int p = (int)c0.caps[k][0], q = (int)c0.caps[k][1];
        if (p) elts.push_back({p, p, mk(CT_CAP, k, false)});
        if (q) elts.push_back({q, q, mk(CT_CAP, k, false)});
        if (p && q) {
            elts.push_back({p, q, mk(CT_CAP, k, true)});
            elts.push_back({q, p, mk(CT_CAP, k, true)});
        }
        if (p) rhs_codes[p].push_back(mk(CT_CAP, k, true));
        if (q) rhs_codes[q].push_back(mk(CT_CAP, k, false));
    }
    for (int k = 0; k < NRES; k++) {
        const int p = (int)c0.res[k][0], q = (int)c0.res[k][1];
        if (p) elts.push_back({p, p, mk(CT_RES, k, false)});
        if (q) elts.push_back({q, q, mk(CT_RES, k, false)});
        if (p && q) {
            elts.push_back({p, q, mk(CT_RES, k, true)});
            elts.push_back({q, p, mk(CT_RES, k, true)});
        }
    }
    for (int k = 0; k < NVS; k++) {
        const int p = c0.vs[k].p, q = c0.vs[k].n, br = NN + 1 + k;
        if (p) { elts.push_back({p, br, mk(CT_ONE, 0, false)}); elts.push_back({br, p, mk(CT_ONE, 0, false)}); }
        if (q) { elts.push_back({q, br, mk(CT_ONE, 0, true)}); elts.push_back({br, q, mk(CT_ONE, 0, true)}); }
        rhs_codes[br].push_back(mk(CT_VSRC, k, false));
    }
    /* .ic rows: voltage columns are zeroed in the transient operating point
     * (cktload.c:146-173); the row must hold no branch-current column.
     * Under UIC the rows are never forced (the .ic values only seed the
     * initial vector), so there are none. */
    P.ic_nodes.assign(NN + 1, 0);
    for (int u = 1; u <= NN; u++) P.ic_nodes[u] = c0.uic ? 0 : c0.ic_given[u];
    for (Elt &e : elts) {
        if (e.r <= NN && P.ic_nodes[e.r]) {
            if (e.c > NN) return fail(".ic on a node with a voltage source is not supported");
            e.code |= CT_ICZ;
        }
    }

    /* ---- pivot sequence: step k eliminates row rowp[k], column colq[k]
     * (1-based unknowns).  Given by the caller (the order klu_factor chose,
     * Numeric->Pnum / Symbolic->Q), or built here: each source's branch row
     * is paired with the row of its positive node so both diagonals are the
     * exact +-1 entries of the source stamp, then a minimum-degree symmetric
     * ordering of the row-paired pattern limits fill. */
    std::vector<int> rowp(N), colq(N);
    /* step -> diagonal block of KLU's block triangular form (one block
     * unless the circuit uses KLU's ordering): the LU then factors the
     * diagonal blocks only and the solve runs block by block, as klu_factor
     * and klu_solve do (klu_factor.c:15-260, klu_solve.c:14-400) */
    std::vector<int> blk(N, 0);
    bool blocks = false;
    if (!c0.order_rows.empty()) {
        if ((int)c0.order_rows.size() != N) return fail("pivot order length differs from the unknown count");
        rowp = c0.order_rows;
        colq = c0.order_cols;
    } else if (c0.klu_order) {
        /* KLU's symbolic analysis of the pattern SMPpreOrder hands it:
         * distinct entries, 0-based, rows ascending per column */
        std::vector<std::vector<int>> colrows(N);
        for (const Elt &e : elts) colrows[e.c - 1].push_back(e.r - 1);
        std::vector<int> Ap(N + 1, 0), Ai;
        for (int j = 0; j < N; j++) {
            std::vector<int> &v = colrows[j];
            std::sort(v.begin(), v.end());
            v.erase(std::unique(v.begin(), v.end()), v.end());
            Ai.insert(Ai.end(), v.begin(), v.end());
            Ap[j + 1] = (int)Ai.size();
        }
        std::vector<int> kp(N), kq(N), kr(N + 1);
        int nb = 0;
        const int rc = ngmi_klu_analyze(N, Ap.data(), Ai.data(), kp.data(), kq.data(), kr.data(), &nb, nullptr,
                                        nullptr);
        if (rc != NGMI_OK) return rc;
        for (int k = 0; k < N; k++) {
            rowp[k] = kp[k] + 1;
            colq[k] = kq[k] + 1;
        }
        for (int b = 0; b < nb; b++)
            for (int k = kr[b]; k < kr[b + 1]; k++) blk[k] = b;
        blocks = nb > 1;
    } else {
        std::vector<int> prow(N + 1), prow_inv(N + 1);
        for (int u = 0; u <= N; u++) prow[u] = u;
        for (int k = 0; k < NVS; k++) {
            const int br = NN + 1 + k;
            const int node = c0.vs[k].p ? c0.vs[k].p : c0.vs[k].n;
            if (prow[node] != node) return fail("two voltage sources share a node row");
            prow[node] = br;
            prow[br] = node;
        }
        for (int u = 0; u <= N; u++) prow_inv[prow[u]] = u;
        std::vector<std::vector<char>> Bp(N, std::vector<char>(N, 0));   /* B = P A, 0-based */
        for (const Elt &e : elts) Bp[prow_inv[e.r] - 1][e.c - 1] = 1;
        for (int a = 0; a < N; a++) Bp[a][a] = 1;
        const int W = (N + 63) / 64;
        std::vector<std::vector<uint64_t>> adj(N, std::vector<uint64_t>(W, 0));
        for (int a = 0; a < N; a++)
            for (int c = 0; c < N; c++)
                if (a != c && (Bp[a][c] || Bp[c][a])) adj[a][c >> 6] |= 1ull << (c & 63);
        /* B index eliminated q-th */
        const std::vector<int> order = min_degree(N, adj, tune_get(TUNE_ORDER_SLACK));
        for (int q = 0; q < N; q++) {
            rowp[q] = prow[order[q] + 1];
            colq[q] = order[q] + 1;
        }
    }
    std::vector<int> rinv(N + 1, -1), cinv(N + 1, -1);   /* unknown -> step */
    for (int k = 0; k < N; k++) {
        if (rowp[k] < 1 || rowp[k] > N || colq[k] < 1 || colq[k] > N || rinv[rowp[k]] >= 0 ||
            cinv[colq[k]] >= 0)
            return fail("pivot order is not a pair of permutations of 1..N");
        rinv[rowp[k]] = k;
        cinv[colq[k]] = k;
    }
    P.rowp = rowp;
    P.colq = colq;

    /* ---- filled pattern of M = A(rowp, colq): the pivots, every element
     * position and the .ic diagonals are structural */
    std::vector<std::vector<char>> S(N, std::vector<char>(N, 0));
    for (const Elt &e : elts) S[rinv[e.r]][cinv[e.c]] = 1;
    for (int k = 0; k < N; k++) S[k][k] = 1;
    for (int u = 1; u <= NN; u++)
        if (P.ic_nodes[u]) S[rinv[u]][cinv[u]] = 1;
    /* fill stays inside the diagonal blocks (no entries lie below them) */
    for (int k = 0; k < N; k++)
        for (int i = k + 1; i < N; i++)
            if (S[i][k])
                for (int j = k + 1; j < N; j++)
                    if (S[k][j] && blk[j] == blk[k]) S[i][j] = 1;
    std::vector<int> slot((size_t)N * N, -1);
    int nslots = 0;
    P.row_start.assign(N + 1, 0);
    for (int i = 0; i < N; i++) {
        P.row_start[i] = nslots;
        for (int j = 0; j < N; j++)
            if (S[i][j]) {
                slot[(size_t)i * N + j] = nslots++;
                P.slot_col.push_back(j);
                if (i > j) P.nnz_l++;
            }
    }
    P.row_start[N] = nslots;
    P.nslots = nslots;

    /* ---- factor program: entries ordered by m = min(i, j), diagonal first */
    std::vector<int> lev(nslots, 0);
    /* an entry with more than SPLIT_TERMS terms runs as partial sums of up
     * to 4 * PROG_INLINE terms on several lanes at level plev, added into
     * the entry (and divided by its pivot) one level later */
    struct Ent { int t, piv, lev, plev; std::vector<int2> terms; };
    std::vector<Ent> ents;
    ents.reserve(nslots);
    for (int m = 0; m < N; m++) {
        std::vector<std::pair<int, int>> list;
        list.push_back({m, m});
        for (int j = m + 1; j < N; j++) if (S[m][j]) list.push_back({m, j});
        for (int i = m + 1; i < N; i++) if (S[i][m]) list.push_back({i, m});
        for (auto [i, j] : list) {
            Ent e;
            e.t = slot[(size_t)i * N + j];
            e.piv = (i > j) ? slot[(size_t)j * N + j] : -1;
            int L = 0;
            /* an entry off the diagonal blocks keeps its assembled value */
            for (int k = 0; k < m && blk[i] == blk[j]; k++) {
                if (S[i][k] && S[k][j]) {
                    const int a = slot[(size_t)i * N + k], c2 = slot[(size_t)k * N + j];
                    e.terms.push_back(make_int2(a, c2));
                    L = std::max(L, std::max(lev[a], lev[c2]) + 1);
                }
            }
            e.plev = L;
            if ((int)e.terms.size() > SPLIT_TERMS) L = std::max(L + 1, 1);
            if (e.piv >= 0) L = std::max(L, lev[e.piv] + 1);
            lev[e.t] = L;
            e.lev = L;
            ents.push_back(std::move(e));
        }
    }
    int nlev = 0;
    for (const Ent &e : ents) nlev = std::max(nlev, e.lev + 1);
    if (nslots + 1 > 65535 || N + 1 > 65535) return fail("circuit too large for 16-bit LU indices");
    /* term lists: two 16-bit indices per word, padded to whole groups of
     * four with a pair whose product is exactly zero (lu[nslots] and z[N]
     * hold 0).  An entry is a 64-byte record {target | g << 16, y, begin,
     * end} + three groups: its first g (0..3) groups inline, the rest at
     * [begin, end) in uint4 units of dst. */
    auto put_terms = [](std::vector<uint32_t> &dst, const std::vector<int2> &t, uint32_t pad,
                        std::vector<int> &rec, int target, int y, int kind = PROG_PLAIN) {
        std::vector<uint32_t> w;
        for (const int2 &x : t) w.push_back((uint32_t)x.x | ((uint32_t)x.y << 16));
        const int g = std::min(PROG_INLINE, (int)(w.size() + 3) / 4);
        while ((int)w.size() < 4 * g || w.size() % 4) w.push_back(pad);
        rec.push_back(target | (g << 16) | (kind << 20));
        rec.push_back(y);
        rec.push_back((int)(dst.size() / 4));
        if ((int)w.size() > 4 * PROG_INLINE) dst.insert(dst.end(), w.begin() + 4 * PROG_INLINE, w.end());
        rec.push_back((int)(dst.size() / 4));
        for (int k = 0; k < 4 * PROG_INLINE; k++) rec.push_back(k < (int)w.size() ? (int)w[k] : (int)pad);
    };
    /* a combine record: target, y, the range of partial slots to add */
    auto put_combine = [](std::vector<int> &rec, int target, int y, int p0, int p1, uint32_t pad) {
        rec.push_back(target | (PROG_COMBINE << 20));
        rec.push_back(y);
        rec.push_back(p0);
        rec.push_back(p1);
        for (int k = 0; k < 4 * PROG_INLINE; k++) rec.push_back((int)pad);
    };
    /* one record to emit: entry idx whole (chunk -1), its partial sum of
     * terms [chunk * CH, chunk * CH + CH) (chunk >= 0), or its combine (-2) */
    struct Item { int idx, chunk; };
    const int CH = 4 * PROG_INLINE;
    int part_half = 0;   /* partial slots per level parity, over all programs */
    auto count_parts = [&](const std::vector<std::vector<Item>> &by) {
        for (const auto &items : by) {
            int n = 0;
            for (const Item &it : items) n += it.chunk >= 0;
            part_half = std::max(part_half, n);
        }
    };
    const uint32_t pad_lu = (uint32_t)nslots | ((uint32_t)nslots << 16);
    const uint32_t pad_z = (uint32_t)nslots | ((uint32_t)N << 16);
    /* entries that change nothing (a U or diagonal entry with no terms: its
     * assembled value is final) are left out, and levels left empty with
     * them are dropped */
    /* the three programs as entries {target, y, level, partial level,
     * terms}: the factor's (no-op entries left out), the forward rows with
     * terms, every backward row */
    struct PEnt { int t, y, lev, plev; std::vector<int2> terms; };
    std::vector<PEnt> fe, we, be;
    int nterms = 0;
    for (const Ent &e : ents) {
        if (e.terms.empty() && e.piv < 0) continue;
        fe.push_back({e.t, e.piv, e.lev, e.plev, e.terms});
        nterms += (int)e.terms.size();
    }
    P.nterms = nterms;
    if (blocks) {
        /* klu_solve's order (klu_solve.c:14-400): blocks from the last; in
         * block b, y = L_b \ (b_b - A_b,later x_later), then x_b = U_b \ y.
         * One program: the forward entries (no pivot) of a block wait for
         * the back-substituted unknowns of later blocks their rows touch. */
        std::vector<int> lvf(N, -1), lvb(N, -1);
        int nb = 0;
        for (int k = 0; k < N; k++) nb = std::max(nb, blk[k] + 1);
        std::vector<int> bstart(nb + 1, N);
        for (int k = N - 1; k >= 0; k--) bstart[blk[k]] = k;
        for (int b = nb - 1; b >= 0; b--) {
            const int k1 = bstart[b], k2 = bstart[b + 1];
            for (int a = k1; a < k2; a++) {
                PEnt r{a, -1, 0, 0, {}};
                int L = 0;
                for (int k = k1; k < a; k++)
                    if (S[a][k]) {
                        r.terms.push_back(make_int2(slot[(size_t)a * N + k], k));
                        L = std::max(L, lvf[k] + 1);
                    }
                for (int j = k2; j < N; j++)
                    if (S[a][j]) {
                        r.terms.push_back(make_int2(slot[(size_t)a * N + j], j));
                        L = std::max(L, lvb[j] + 1);
                    }
                if (r.terms.empty()) continue;
                r.plev = L;
                if ((int)r.terms.size() > SPLIT_TERMS) L++;
                lvf[a] = r.lev = L;
                be.push_back(std::move(r));
            }
            for (int a = k2 - 1; a >= k1; a--) {
                PEnt r{a, slot[(size_t)a * N + a], 0, 0, {}};
                int L = lvf[a] + 1;
                for (int j = a + 1; j < k2; j++)
                    if (S[a][j]) {
                        r.terms.push_back(make_int2(slot[(size_t)a * N + j], j));
                        L = std::max(L, lvb[j] + 1);
                    }
                r.plev = L;
                if ((int)r.terms.size() > SPLIT_TERMS) L++;
                lvb[a] = r.lev = L;
                be.push_back(std::move(r));
            }
        }
    }
    if (!blocks) {
        std::vector<int> lv(N, 0);
        for (int a = 0; a < N; a++) {
            PEnt r{a, -1, 0, 0, {}};
            int L = 0;
            for (int k = 0; k < a; k++)
                if (S[a][k]) {
                    r.terms.push_back(make_int2(slot[(size_t)a * N + k], k));
                    L = std::max(L, lv[k] + 1);
                }
            r.plev = L;
            if ((int)r.terms.size() > SPLIT_TERMS) L++;
            lv[a] = r.lev = L;
            if (!r.terms.empty()) we.push_back(std::move(r));   /* rows without terms change nothing */
        }
    }
    if (!blocks) {
        std::vector<int> lv(N, 0);
        for (int a = N - 1; a >= 0; a--) {
            PEnt r{a, slot[(size_t)a * N + a], 0, 0, {}};
            int L = 0;
            for (int j = a + 1; j < N; j++)
                if (S[a][j]) {
                    r.terms.push_back(make_int2(slot[(size_t)a * N + j], j));
                    L = std::max(L, lv[j] + 1);
                }
            r.plev = L;
            if ((int)r.terms.size() > SPLIT_TERMS) L++;
            lv[a] = r.lev = L;
            be.push_back(std::move(r));
        }
        std::reverse(be.begin(), be.end());   /* ascending rows within a level */
    }
    auto items = [&](const std::vector<PEnt> &ps) {
        int nl = 0;
        for (const PEnt &e : ps) nl = std::max(nl, e.lev + 1);
        std::vector<std::vector<Item>> by(nl);
        for (int k = 0; k < (int)ps.size(); k++) {
            const PEnt &e = ps[k];
            if ((int)e.terms.size() > SPLIT_TERMS) {
                for (int q = 0; q * CH < (int)e.terms.size(); q++) by[e.plev].push_back({k, q});
                by[e.lev].push_back({k, -2});
            } else {
                by[e.lev].push_back({k, -1});
            }
        }
        return by;
    };
    const auto fby = items(fe), wby = items(we), bby = items(be);
    count_parts(fby);
    count_parts(wby);
    count_parts(bby);
    /* partial slots: executed level i uses half (i & 1), so the partials a
     * combine reads (written one level earlier) differ from those being
     * written beside it */
    auto emit = [&](const std::vector<PEnt> &ps, const std::vector<std::vector<Item>> &by, uint32_t pad,
                    std::vector<int> &lev_out, std::vector<int> &ent, std::vector<uint32_t> &term) {
        std::vector<int2> prange(ps.size(), make_int2(0, 0));
        for (const auto &its : by) {
            if (its.empty()) continue;
            const int base = ((int)lev_out.size() & 1) * part_half;
            int used = 0;
            lev_out.push_back((int)(ent.size() / PROG_INTS));
            for (const Item &it : its) {
                const PEnt &e = ps[it.idx];
                if (it.chunk == -1) {
                    put_terms(term, e.terms, pad, ent, e.t, e.y);
                } else if (it.chunk >= 0) {
                    const int q0 = it.chunk * CH, q1 = std::min((int)e.terms.size(), q0 + CH);
                    const std::vector<int2> part(e.terms.begin() + q0, e.terms.begin() + q1);
                    const int ps_slot = base + used++;
                    if (it.chunk == 0) prange[it.idx].x = ps_slot;
                    prange[it.idx].y = ps_slot + 1;
                    put_terms(term, part, pad, ent, ps_slot, -1, PROG_PARTIAL);
                } else {
                    put_combine(ent, e.t, e.y, prange[it.idx].x, prange[it.idx].y, pad);
                }
            }
        }
        const int n = (int)lev_out.size();
        lev_out.push_back((int)(ent.size() / PROG_INTS));
        return n;
    };
    std::vector<int> f_lev, f_ent, fw_lev, fw_ent, bw_lev, bw_ent;
    std::vector<uint32_t> f_term, fw_term, bw_term;
    P.nlev_f = emit(fe, fby, pad_lu, f_lev, f_ent, f_term);
    P.nlev_fw = emit(we, wby, pad_z, fw_lev, fw_ent, fw_term);
    P.nlev_bw = emit(be, bby, pad_z, bw_lev, bw_ent, bw_term);
    P.n_ppart = 2 * part_half;

    /* ---- assembly plans in permuted coordinates */
    std::vector<std::vector<uint32_t>> per_slot(nslots);
    std::vector<int> slot_flags(nslots, 0);
    for (const Elt &e : elts) {
        const int i = rinv[e.r], j = cinv[e.c];
        per_slot[slot[(size_t)i * N + j]].push_back(e.code);
    }
    for (int u = 1; u <= NN; u++) {
        if (!P.ic_nodes[u]) continue;
        const int i = rinv[u], j = cinv[u];
        slot_flags[slot[(size_t)i * N + j]] |= 1;
    }
    /* the diagonal elements the matrix has (device SMPmakeElt calls and the
     * .ic rows, CKTic): LoadGmin's targets */
    P.diag_slot.assign(N + 1, -1);
    for (const Elt &e : elts)
        if (e.r == e.c) P.diag_slot[e.r] = slot[(size_t)rinv[e.r] * N + cinv[e.c]];
    for (int u = 1; u <= NN; u++)
        if (P.ic_nodes[u]) P.diag_slot[u] = slot[(size_t)rinv[u] * N + cinv[u]];
    /* a slot or row with more than a few contributions is split into
     * leaves that write partial sums, plus one combine step (the supply
     * node's diagonal collects one contribution per connected device) */
    struct Leaf {
        int dest, flag;
        std::vector<uint32_t> w;   /* whole groups of four codes */
    };
    auto split = [](const std::vector<uint32_t> &codes, int dest, int flag, int partial_dest0,
                    int &npart, std::vector<Leaf> &leaves, std::vector<int4> &comb) {
        /* contributions in whole groups of four, padded with zero codes */
        const int n = (int)codes.size();
        auto put = [&](int d, int f, int q0, int q1) {
            Leaf lf{d, f, std::vector<uint32_t>(codes.begin() + q0, codes.begin() + q1)};
            while (lf.w.size() < 4 || lf.w.size() % 4) lf.w.push_back(CT_ZERO << CT_TYPE_SHIFT);
            leaves.push_back(std::move(lf));
        };
        if (n <= ASM_LEAF) {
            put(dest, flag, 0, n);
            return;
        }
        const int chunk = (std::max(ASM_LEAF, (n + ASM_MAXPART - 1) / ASM_MAXPART) + 3) / 4 * 4;
        const int p0 = npart;
        for (int q = 0; q < n; q += chunk) put(partial_dest0 + npart++, 0, q, std::min(n, q + chunk));
        comb.push_back(make_int4(dest, p0, npart, flag));
    };
    /* group e of the code (and contribution) array is leaf e's first group,
     * so a lane finds a leaf's first values without reading its record
     * first; the further groups follow all first groups.  A leaf is a
     * 32-byte record {dest, begin, end, flag} + a copy of its first group,
     * with [begin, end) its further groups in uint4 units of code.  Leaves
     * come in three runs: plain ones (a single group whose values are all
     * read from the contribution array as they are: no capacitor or source
     * code, no .ic handling), whose destinations also go to dest so the
     * step kernel reads 4 bytes per leaf instead of the record; other
     * single-group leaves; multi-group leaves.  Returns the run lengths
     * (plain, plain + single). */
    auto layout = [](std::vector<Leaf> leaves, std::vector<uint32_t> &code,
                     std::vector<int4> &leaf, std::vector<int> &dest, bool rhs) {
        auto kind = [rhs](const Leaf &lf) {
            if (lf.w.size() != 4) return 2;
            if (lf.flag) return 1;
            for (int k = 0; k < 4; k++) {
                const uint32_t t = lf.w[k] >> CT_TYPE_SHIFT;
                if (t == CT_CAP || (rhs && t == CT_VSRC) || (!rhs && (lf.w[k] & CT_ICZ))) return 1;
            }
            return 0;
        };
        std::stable_sort(leaves.begin(), leaves.end(),
                         [&](const Leaf &a, const Leaf &b) { return kind(a) < kind(b); });
        int plain = 0, singles = 0;
        for (const Leaf &lf : leaves) {
            plain += kind(lf) == 0;
            singles += kind(lf) <= 1;
        }
        for (int e = 0; e < plain; e++) dest.push_back(leaves[e].dest);
        if (dest.empty()) dest.push_back(0);
        for (const Leaf &lf : leaves) code.insert(code.end(), lf.w.begin(), lf.w.begin() + 4);
        for (const Leaf &lf : leaves) {
            const int b = (int)(code.size() / 4);
            code.insert(code.end(), lf.w.begin() + 4, lf.w.end());
            leaf.push_back(make_int4(lf.dest, b, (int)(code.size() / 4), lf.flag));
            leaf.push_back(make_int4((int)lf.w[0], (int)lf.w[1], (int)lf.w[2], (int)lf.w[3]));
        }
        return std::make_pair(plain, singles);
    };
    std::vector<uint32_t> a_code, r_code;
    std::vector<int4> a_leaf, a_comb, r_leaf, r_comb;
    std::vector<Leaf> a_leaves, r_leaves;
    int n_apart = 0, n_rpart = 0;
    for (int s = 0; s < nslots; s++)
        split(per_slot[s], s, slot_flags[s] & 1, nslots, n_apart, a_leaves, a_comb);
    std::vector<int> perm(2 * N), ic_row;
    for (int a = 0; a < N; a++) {
        perm[a] = colq[a];                    /* unknown solved in column a */
        perm[N + a] = rowp[a];                /* original row assembled in row a */
        split(rhs_codes[rowp[a]], a, 0, N, n_rpart, r_leaves, r_comb);
        if (perm[N + a] <= NN && P.ic_nodes[perm[N + a]]) ic_row.push_back(a);
    }
    std::tie(P.n_aplain, P.n_aleaf1) = layout(std::move(a_leaves), a_code, a_leaf, P.a_dest, false);
    std::tie(P.n_rplain, P.n_rleaf1) = layout(std::move(r_leaves), r_code, r_leaf, P.r_dest, true);
    P.n_aleaf = (int)a_leaf.size() / 2; P.n_acomb = (int)a_comb.size();
    P.n_rleaf = (int)r_leaf.size() / 2; P.n_rcomb = (int)r_comb.size(); P.n_icrow = (int)ic_row.size();
    P.n_apart = n_apart; P.n_rpart = n_rpart;
    /* keep every device array non-empty */
    if (a_comb.empty()) a_comb.push_back(make_int4(0, 0, 0, 0));
    if (r_comb.empty()) r_comb.push_back(make_int4(0, 0, 0, 0));
    if (r_leaf.empty()) r_leaf.assign(2, make_int4(0, 0, 0, 0));
    while (r_code.size() < 4) r_code.push_back(CT_ZERO << CT_TYPE_SHIFT);
    while (a_code.size() < 4) a_code.push_back(CT_ZERO << CT_TYPE_SHIFT);
    /* where the BSIM4 kernel writes each stamp and RHS value */
    P.mos_pos.assign((size_t)B4S__COUNT * NM, -1);
    P.rmos_pos.assign((size_t)B4R__COUNT * NM, -1);
    for (size_t p = 0; p < a_code.size(); p++)
        if ((a_code[p] >> CT_TYPE_SHIFT) == CT_MOS) P.mos_pos[a_code[p] & CT_INDEX_MASK] = (int)p;
    for (size_t p = 0; p < r_code.size(); p++)
        if ((r_code[p] >> CT_TYPE_SHIFT) == CT_MOS) P.rmos_pos[r_code[p] & CT_INDEX_MASK] = (int)p;
    P.NA = (int)a_code.size();
    P.NR = (int)r_code.size();
    if (ic_row.empty()) ic_row.push_back(0);

    /* ---- neighbour table for the learned guess: up to 8 other non-ground
     * nodes sharing an element with each node, ascending */
    {
        std::vector<std::vector<int>> nb(NN + 1);
        auto link = [&](int a, int c2) {
            if (a >= 1 && a <= NN && c2 >= 1 && c2 <= NN && a != c2) {
                nb[a].push_back(c2);
                nb[c2].push_back(a);
            }
        };
        for (int j = 0; j < NM; j++)
            for (int r1 = 0; r1 < 4; r1++)
                for (int r2 = r1 + 1; r2 < 4; r2++) link(c0.mos[j].node[r1], c0.mos[j].node[r2]);
        for (int k = 0; k < NCAP; k++) link((int)c0.caps[k][0], (int)c0.caps[k][1]);
        for (int k = 0; k < NRES; k++) link((int)c0.res[k][0], (int)c0.res[k][1]);
        for (int k = 0; k < NVS; k++) link(c0.vs[k].p, c0.vs[k].n);
        P.nbr.assign((size_t)(NN + 1) * 8, -1);
        for (int u = 1; u <= NN; u++) {
            std::sort(nb[u].begin(), nb[u].end());
            nb[u].erase(std::unique(nb[u].begin(), nb[u].end()), nb[u].end());
            for (int k = 0; k < 8 && k < (int)nb[u].size(); k++) P.nbr[(size_t)u * 8 + k] = nb[u][k];
        }
    }

    P.f_lev = std::move(f_lev); P.f_ent = std::move(f_ent); P.f_term = std::move(f_term);
    P.fw_lev = std::move(fw_lev); P.fw_ent = std::move(fw_ent); P.fw_term = std::move(fw_term);
    P.bw_lev = std::move(bw_lev); P.bw_ent = std::move(bw_ent); P.bw_term = std::move(bw_term);
    P.a_code = std::move(a_code); P.r_code = std::move(r_code);
    P.a_leaf = std::move(a_leaf); P.a_comb = std::move(a_comb);
    P.r_leaf = std::move(r_leaf); P.r_comb = std::move(r_comb);
    P.unk_pos.assign(N + 1, N);
    for (int a = 0; a < N; a++) P.unk_pos[perm[a]] = a;
    P.perm = std::move(perm); P.ic_row = std::move(ic_row);
    return NGMI_OK;
}

/* one topology group of a batch: symbolic analysis, BSIM4 table and the
 * device arrays of its circuits */
int build_group(ngmi_batch *b, TopoGroup *G, const std::vector<const ngmi_circuit *> &ckts,
                const int *probe, int nprobe)
{
    auto fail = [](const char *what) {
        if (what) set_error("%s", what);
        return NGMI_E_UNSUPP;
    };
    const ngmi_circuit &c0 = *ckts[0];
    const int C = (int)ckts.size();
    G->C = C;
    G->NN = c0.nnodes;
    G->NVS = (int)c0.vs.size();
    G->N = G->NN + G->NVS;
    G->NM = (int)c0.mos.size();
    G->NCAP = (int)c0.caps.size();
    G->NRES = (int)c0.res.size();
    const int N = G->N, NN = G->NN, NM = G->NM, NCAP = G->NCAP, NRES = G->NRES, NVS = G->NVS;

    Plan P;
    if (analyze(c0, P) != NGMI_OK) return fail(nullptr);
    G->ic_nodes_host = P.ic_nodes;
    G->nbr_host = P.nbr;
    G->nslots = P.nslots; G->nterms = P.nterms;
    G->nlev_f = P.nlev_f; G->nlev_fw = P.nlev_fw; G->nlev_bw = P.nlev_bw;

    /* ---- BSIM4 instances, circuit-major */
    G->set = ngmi_bsim4_set_new();
    for (int c = 0; c < C; c++) {
        const ngmi_circuit &ck = *ckts[c];
        for (int j = 0; j < NM; j++) {
            const ngmi_circuit::Mos &mo = ck.mos[j];
            std::vector<const char *> nm;
            for (const auto &s : mo.names) nm.push_back(s.c_str());
            std::vector<double> vals = mo.values;
            if (c0.uic) {
                /* BSIM4getic (b4getic.c:32-55): under UIC an instance IC not
                 * given is the .ic difference of its terminals (CKTrhs after
                 * CKTic: the .ic values, 0 elsewhere) */
                auto icv = [&](int u) { return (u > 0 && u <= NN && ck.ic_given[u]) ? ck.ic[u] : 0.0; };
                const double vs = icv(mo.node[2]);
                static const char *const icn[3] = {"icvds", "icvgs", "icvbs"};
                const double icval[3] = {icv(mo.node[0]) - vs, icv(mo.node[1]) - vs, icv(mo.node[3]) - vs};
                for (int q = 0; q < 3; q++) {
                    bool given = false;
                    for (const auto &s : mo.names) given |= (s == icn[q]);
                    if (!given) {
                        nm.push_back(icn[q]);
                        vals.push_back(icval[q]);
                    }
                }
            }
            ngmi_bsim4_model *mp = mo.model;
            const int zero = 0;
            const int rc = ngmi_bsim4_set_add(G->set, &mp, 1, &zero, 1, nm.data(), (int)nm.size(), vals.data());
            if (rc < 0) return fail(nullptr);
        }
    }
    if (ngmi_bsim4_set_finalize(G->set) != NGMI_OK) return fail(nullptr);

    /* ---- per-circuit element values */
    std::vector<double> cap_c((size_t)C * NCAP), res_g((size_t)C * NRES), vs_par((size_t)C * NVS * VSP);
    std::vector<double> cap_ic((size_t)C * NCAP, 0.0);
    std::vector<double> ic_val((size_t)C * (NN + 1), 0.0);
    for (int c = 0; c < C; c++) {
        const ngmi_circuit &ck = *ckts[c];
        for (int k = 0; k < NCAP; k++) cap_c[(size_t)c * NCAP + k] = ck.caps[k][2];
        if (c0.uic) {
            /* CAPgetic (capgetic.c:19-35) under UIC */
            auto icv = [&](int u) { return (u > 0 && u <= NN && ck.ic_given[u]) ? ck.ic[u] : 0.0; };
            for (int k = 0; k < NCAP; k++)
                cap_ic[(size_t)c * NCAP + k] = icv((int)ck.caps[k][0]) - icv((int)ck.caps[k][1]);
        }
        for (int k = 0; k < NRES; k++) res_g[(size_t)c * NRES + k] = 1.0 / ck.res[k][2];
        for (int k = 0; k < NVS; k++) {
            double *v = &vs_par[((size_t)c * NVS + k) * VSP];
            v[0] = ck.vs[k].dc;
            v[1] = ck.vs[k].npulse > 0 ? 1.0 : 0.0;
            v[2] = ck.vs[k].npulse;
            for (int q = 0; q < 7; q++) v[3 + q] = q < ck.vs[k].npulse ? ck.vs[k].pulse[q] : 0.0;
        }
        for (int u = 1; u <= NN; u++) ic_val[(size_t)c * (NN + 1) + u] = ck.ic[u];
    }
    std::vector<int> mos_nodes(4 * (size_t)NM), cap_nodes(2 * (size_t)NCAP), res_nodes(2 * (size_t)NRES),
        vs_nodes(3 * (size_t)NVS), ic_nodes(NN + 1);
    for (int j = 0; j < NM; j++)
        for (int r = 0; r < 4; r++) mos_nodes[(size_t)r * NM + j] = c0.mos[j].node[r];
    for (int k = 0; k < NCAP; k++) {
        cap_nodes[k] = (int)c0.caps[k][0];
        cap_nodes[NCAP + k] = (int)c0.caps[k][1];
    }
    for (int k = 0; k < NRES; k++) {
        res_nodes[k] = (int)c0.res[k][0];
        res_nodes[NRES + k] = (int)c0.res[k][1];
    }
    for (int k = 0; k < NVS; k++) {
        vs_nodes[k] = c0.vs[k].p;
        vs_nodes[NVS + k] = c0.vs[k].n;
        vs_nodes[2 * NVS + k] = NN + 1 + k;
    }
    for (int u = 0; u <= NN; u++) ic_nodes[u] = G->ic_nodes_host[u];
    if (probe && nprobe > 0) G->probe.assign(probe, probe + nprobe);
    else for (int u = 1; u <= N; u++) G->probe.push_back(u);
    for (int u : G->probe)
        if (u < 0 || u > N) return fail("probe outside the unknowns");
    
    TranDev &D = G->D;
    D.C = C; D.N = N; D.NN = NN; D.NM = NM; D.NCAP = NCAP; D.NRES = NRES; D.NVS = NVS;
    D.nslots = P.nslots; D.nlev_f = G->nlev_f; D.nlev_fw = G->nlev_fw; D.nlev_bw = G->nlev_bw;
    D.n_aleaf = P.n_aleaf; D.n_acomb = P.n_acomb; D.n_rleaf = P.n_rleaf; D.n_rcomb = P.n_rcomb;
    D.n_apart = P.n_apart; D.n_rpart = P.n_rpart; D.n_icrow = P.n_icrow;
    D.n_aleaf1 = P.n_aleaf1; D.n_rleaf1 = P.n_rleaf1;
    D.n_aplain = P.n_aplain; D.n_rplain = P.n_rplain;
    D.n_part = std::max(P.n_apart + P.n_rpart, P.n_ppart);
    D.nprobe = (int)G->probe.size();
    D.cap_pts = b->cap_pts;
    D.reltol = c0.reltol; D.abstol = c0.abstol; D.vntol = c0.vntol; D.chgtol = c0.chgtol;
    D.trtol = c0.trtol; D.gmin = c0.gmin;
    D.numGminSteps = c0.gminsteps; D.numSrcSteps = c0.srcsteps; D.dcTrcvMaxIter = c0.itl2;
    D.gminFactor = c0.gminfactor; D.gshunt = c0.gshunt;
    D.opstep = c0.opstep; D.opfinal = c0.opfinal;
    G->noopiter = c0.noopiter;
    G->uic = c0.uic;
    int *ip;
    double *dp;
    uint32_t *up;
    int2 *i2;
    int4 *i4;
#define UP(vec, field, T, tmp) do { if (upload<T>(b, vec, &tmp) != NGMI_OK) return fail(nullptr); D.field = tmp; } while (0)
    UP(mos_nodes, mos_nodes, int, ip);
    UP(cap_nodes, cap_nodes, int, ip);
    UP(res_nodes, res_nodes, int, ip);
    UP(vs_nodes, vs_nodes, int, ip);
    UP(ic_nodes, ic_nodes, int, ip);
    UP(cap_c, cap_c, double, dp);
    UP(cap_ic, cap_ic, double, dp);
    UP(res_g, res_g, double, dp);
    UP(vs_par, vs_par, double, dp);
    UP(ic_val, ic_val, double, dp);
    UP(P.a_code, a_code, uint32_t, up);
    UP(P.r_code, r_code, uint32_t, up);
    UP(P.a_leaf, a_leaf, int4, i4);
    UP(P.a_dest, a_dest, int, ip);
    UP(P.r_dest, r_dest, int, ip);
    UP(P.a_comb, a_comb, int4, i4);
    UP(P.r_leaf, r_leaf, int4, i4);
    UP(P.r_comb, r_comb, int4, i4);
    UP(P.ic_row, ic_row, int, ip);
    UP(P.f_lev, f_lev, int, ip);
    UP(P.f_ent, f_ent, int, ip);
    UP(P.f_term, f_term, uint32_t, up);
    UP(P.fw_lev, fw_lev, int, ip);
    UP(P.fw_ent, fw_ent, int, ip);
    UP(P.fw_term, fw_term, uint32_t, up);
    UP(P.bw_lev, bw_lev, int, ip);
    UP(P.bw_ent, bw_ent, int, ip);
    UP(P.bw_term, bw_term, uint32_t, up);
    UP(P.perm, perm_unknown, int, ip);
    UP(P.unk_pos, unk_pos, int, ip);
    UP(P.mos_pos, mos_pos, int, ip);
    UP(P.diag_slot, diag_slot, int, ip);
    UP(P.rmos_pos, rmos_pos, int, ip);
    UP(G->probe, probe, int, ip);
    UP(G->nbr_host, nbr, int, ip);
    UP(P.row_start, row_start, int, ip);
    UP(P.slot_col, slot_col, int, ip);
#undef UP
    const int64_t cnm = (int64_t)C * NM, cnc = (int64_t)C * NCAP;
    if (alloc(b, (size_t)C * NVS, &dp) != NGMI_OK) return fail(nullptr);
    D.vs_brk = dp;
    CktCtl *ctl;
    if (alloc(b, (size_t)C, &ctl) != NGMI_OK) return fail(nullptr);
    D.ctl = ctl;
    if (alloc(b, (size_t)2 * C * (N + 1), &dp) != NGMI_OK) return fail(nullptr);
    D.x = dp;
    if (alloc(b, (size_t)NLEV * B4X__COUNT * cnm, &dp) != NGMI_OK) return fail(nullptr);
    D.mstate = dp;
    if (alloc(b, (size_t)NLEV * 2 * cnc, &dp) != NGMI_OK) return fail(nullptr);
    D.cstate = dp;
    if (alloc(b, (size_t)B4C__COUNT * cnm, &dp) != NGMI_OK) return fail(nullptr);
    D.cache = dp;
    D.NA = P.NA;
    D.NR = P.NR;
    if (alloc(b, (size_t)C * (N + 1), &dp) != NGMI_OK) return fail(nullptr);
    D.bak_x = dp;
    if (alloc(b, (size_t)C * (B4X__COUNT * (size_t)NM + 2 * (size_t)NCAP), &dp) != NGMI_OK)
        return fail(nullptr);
    D.bak_s = dp;
    if (alloc(b, (size_t)C * P.NA, &dp) != NGMI_OK) return fail(nullptr);
    D.contrib = dp;
    {
        /* the values the BSIM4 kernel never writes: resistor conductances
         * and source +-1 entries, signed, and 0 at the padding codes; they
         * stay for the whole run */
        const size_t chunk = 256;
        std::vector<double> row;
        for (size_t c0 = 0; c0 < (size_t)C; c0 += chunk) {
            const size_t c1 = std::min<size_t>(C, c0 + chunk);
            row.assign((c1 - c0) * P.NA, 0.0);
            for (size_t c = c0; c < c1; c++)
                for (int p = 0; p < P.NA; p++) {
                    const uint32_t code = P.a_code[p], type = code >> CT_TYPE_SHIFT;
                    double v = 0.0;
                    if (type == CT_RES) v = res_g[c * NRES + (code & CT_INDEX_MASK)];
                    else if (type == CT_ONE) v = 1.0;
                    else continue;
                    row[(c - c0) * P.NA + p] = (code & CT_NEG) ? -v : v;
                }
            NGMI_HIP(hipMemcpy(dp + c0 * P.NA, row.data(), row.size() * sizeof(double),
                               hipMemcpyHostToDevice));
        }
    }
    if (alloc(b, (size_t)C * P.NR, &dp) != NGMI_OK) return fail(nullptr);
    D.rcontrib = dp;
    if (alloc(b, (size_t)cnm, &ip) != NGMI_OK) return fail(nullptr);
    D.noncon = ip;
    if (alloc(b, (size_t)C * b->cap_pts, &dp) != NGMI_OK) return fail(nullptr);
    D.out_t = dp;
    if (alloc(b, (size_t)C * b->cap_pts * D.nprobe, &dp) != NGMI_OK) return fail(nullptr);
    D.out_x = dp;
    D.done_count = b->done_count;
    /* KLU's pivoting: row scale factors and pivot rows per circuit, and a
     * pool of dense workspaces (the assembled matrix with its right-hand
     * side, and its elimination), one per circuit factoring with pivots at
     * the same moment, within a budget of PV_POOL_BYTES */
    D.pivrel = c0.pivrel;
    D.pv_ld = N + 1;
    D.pv_stride = 2 * (int64_t)N * D.pv_ld;
    {
        const int64_t per = D.pv_stride * (int64_t)sizeof(double);
        const int64_t fit = PV_POOL_BYTES / per;
        D.pv_slots = (int)std::min<int64_t>(std::min<int64_t>(C, PV_POOL_SLOTS), std::max<int64_t>(fit, 1));
        if (per > PV_SLOT_MAX_BYTES) D.pv_slots = 0;
    }
    if (alloc(b, (size_t)C * N, &dp) != NGMI_OK) return fail(nullptr);
    D.pv_rs = dp;
    if (alloc(b, (size_t)C * N, &ip) != NGMI_OK) return fail(nullptr);
    D.pv_row = ip;
    if (alloc(b, (size_t)std::max(D.pv_slots, 1) * D.pv_stride, &dp) != NGMI_OK) return fail(nullptr);
    D.pv_pool = dp;
    if (alloc(b, (size_t)std::max(D.pv_slots, 1), &ip) != NGMI_OK) return fail(nullptr);
    D.pv_owner = ip;
    /* LDS: lu, z, capacitor and source values, partial sums (doubles), then
     * the three level tables (ints); at least the dense factorization's
     * scratch (36 N + 12 bytes from the start) */
    G->lds = sizeof(double) * ((size_t)P.nslots + 1 + N + 1 + 2 * (size_t)NCAP + NVS +
                              std::max(P.n_apart + P.n_rpart, P.n_ppart)) +
             sizeof(int) * ((size_t)G->nlev_f + G->nlev_fw + G->nlev_bw + 3);
    G->lds = std::max<size_t>(G->lds, 36 * (size_t)N + 16);
    if (G->lds > 150 * 1024) return fail("circuit too large for the per-circuit LDS factorisation");
    G->method = c0.method;
    G->maxord = c0.maxord;
    return NGMI_OK;
}

}  // namespace

extern "C" {

ngmi_circuit *ngmi_circuit_new(int nnodes)
{
    if (nnodes < 1) {
        set_error("a circuit needs at least one node");
        return nullptr;
    }
    ngmi_circuit *c = new ngmi_circuit;
    c->nnodes = nnodes;
    c->ic.assign(nnodes + 1, 0.0);
    c->ic_given.assign(nnodes + 1, 0);
    return c;
}

void ngmi_circuit_free(ngmi_circuit *c) { delete c; }

static bool node_ok(const ngmi_circuit *c, int n) { return n >= 0 && n <= c->nnodes; }

int ngmi_circuit_add_bsim4(ngmi_circuit *c, ngmi_bsim4_model *m, int d, int g, int s, int b,
                           int nparams, const char *const *names, const double *values)
{
    if (!c || !m) return NGMI_E_PANIC;
    if (!node_ok(c, d) || !node_ok(c, g) || !node_ok(c, s) || !node_ok(c, b)) {
        set_error("MOSFET terminal outside 0..%d", c->nnodes);
        return NGMI_E_NOTERM;
    }
    ngmi_circuit::Mos mo;
    mo.model = m;
    mo.node[0] = d; mo.node[1] = g; mo.node[2] = s; mo.node[3] = b;
    for (int k = 0; k < nparams; k++) {
        mo.names.emplace_back(names[k]);
        mo.values.push_back(values[k]);
    }
    c->mos.push_back(std::move(mo));
    return NGMI_OK;
}

int ngmi_circuit_add_cap(ngmi_circuit *c, int p, int n, double cap)
{
    if (!c) return NGMI_E_PANIC;
    if (!node_ok(c, p) || !node_ok(c, n)) return NGMI_E_NOTERM;
    c->caps.push_back({(double)p, (double)n, cap});
    return NGMI_OK;
}

int ngmi_circuit_add_res(ngmi_circuit *c, int p, int n, double r)
{
    if (!c) return NGMI_E_PANIC;
    if (!node_ok(c, p) || !node_ok(c, n)) return NGMI_E_NOTERM;
    if (r == 0.0) {
        set_error("zero resistance");
        return NGMI_E_PARMVAL;
    }
    c->res.push_back({(double)p, (double)n, r});
    return NGMI_OK;
}

int ngmi_circuit_add_vsrc(ngmi_circuit *c, int p, int n, double dc, const double *pulse, int npulse)
{
    if (!c) return NGMI_E_PANIC;
    if (!node_ok(c, p) || !node_ok(c, n) || p == n) return NGMI_E_NOTERM;
    if (npulse < 0 || npulse > 7 || (npulse > 0 && npulse < 2)) return NGMI_E_PARMVAL;
    ngmi_circuit::Vs v{};
    v.p = p; v.n = n; v.dc = dc; v.npulse = pulse ? npulse : 0;
    for (int k = 0; k < v.npulse; k++) v.pulse[k] = pulse[k];
    c->vs.push_back(v);
    return NGMI_OK;
}

int ngmi_circuit_set_ic(ngmi_circuit *c, int node, double v)
{
    if (!c || node < 1 || node > c->nnodes) return NGMI_E_NOTERM;
    c->ic[node] = v;
    c->ic_given[node] = 1;
    return NGMI_OK;
}

int ngmi_circuit_set_options(ngmi_circuit *c, double reltol, double abstol, double vntol,
                             double chgtol, double trtol, double gmin, int method, int maxord)
{
    if (!c) return NGMI_E_PANIC;
    if (method != 1 && method != 2) return NGMI_E_PARMVAL;
    /* cktsopt.c:123-133 clamps maxord to 1..6.  DCtran raises the order
     * from 1 to 2 only (dctran.c:820-830), so states beyond level 3 are
     * never read and the engine's four-level ring serves any maxord. */
    maxord = std::min(std::max(maxord, 1), 6);
    c->reltol = reltol; c->abstol = abstol; c->vntol = vntol; c->chgtol = chgtol;
    c->trtol = trtol; c->gmin = gmin; c->method = method; c->maxord = maxord;
    return NGMI_OK;
}

int ngmi_circuit_set_op_options(ngmi_circuit *c, int noopiter, int gminsteps, int srcsteps,
                                double gminfactor, double gshunt, int itl2)
{
    if (!c) return NGMI_E_PANIC;
    if (gminsteps < 0 || srcsteps < 0 || itl2 < 1 || !(gminfactor > 1.0) || gshunt < 0) {
        set_error("operating-point option out of range");
        return NGMI_E_PARMVAL;
    }
    if (noopiter && !gminsteps && !srcsteps) {
        set_error("noopiter needs gmin or source stepping");
        return NGMI_E_PARMVAL;
    }
    c->noopiter = noopiter;
    c->gminsteps = gminsteps;
    c->srcsteps = srcsteps;
    c->gminfactor = gminfactor;
    c->gshunt = gshunt;
    c->itl2 = itl2;
    return NGMI_OK;
}

int ngmi_circuit_set_optran(ngmi_circuit *c, double opstep, double opfinal)
{
    if (!c) return NGMI_E_PANIC;
    if (!(opstep >= 0 && opfinal > 0 && opstep <= opfinal)) {
        set_error("optran step must lie in [0, final time]");
        return NGMI_E_PARMVAL;
    }
    /* com_optran (optran.c:168-172): at most a fiftieth of the final time */
    if (opstep > opfinal / 50.) opstep = opfinal / 50.;
    c->opstep = opstep;
    c->opfinal = opfinal;
    return NGMI_OK;
}

int ngmi_circuit_set_uic(ngmi_circuit *c, int uic)
{
    if (!c) return NGMI_E_PANIC;
    c->uic = uic ? 1 : 0;
    return NGMI_OK;
}

int ngmi_circuit_set_pivrel(ngmi_circuit *c, double pivrel)
{
    if (!c) return NGMI_E_PANIC;
    if (!(pivrel > 0.0 && pivrel <= 1.0)) {
        set_error("pivrel must lie in (0, 1]");
        return NGMI_E_PARMVAL;
    }
    c->pivrel = pivrel;
    return NGMI_OK;
}

int ngmi_circuit_set_pivot_order(ngmi_circuit *c, int n, const int *rows, const int *cols)
{
    if (!c) return NGMI_E_PANIC;
    if (n == 0 || !rows || !cols) {
        c->order_rows.clear();
        c->order_cols.clear();
        return NGMI_OK;
    }
    const int N = c->nnodes + (int)c->vs.size();
    if (n != N) {
        set_error("pivot order has %d steps, the circuit %d unknowns", n, N);
        return NGMI_E_PARMVAL;
    }
    std::vector<char> sr(N + 1, 0), sc(N + 1, 0);
    for (int k = 0; k < n; k++) {
        if (rows[k] < 1 || rows[k] > N || cols[k] < 1 || cols[k] > N || sr[rows[k]] || sc[cols[k]]) {
            set_error("pivot order is not a pair of permutations of 1..%d", N);
            return NGMI_E_PARMVAL;
        }
        sr[rows[k]] = sc[cols[k]] = 1;
    }
    c->order_rows.assign(rows, rows + n);
    c->order_cols.assign(cols, cols + n);
    return NGMI_OK;
}

int ngmi_circuit_use_klu_order(ngmi_circuit *c, int on)
{
    if (!c) return NGMI_E_PANIC;
    c->klu_order = on ? 1 : 0;
    return NGMI_OK;
}

int ngmi_circuit_analyze(const ngmi_circuit *c, int *n_unknowns, int *nslots, int *nterms,
                         int *nnz_l, int *levels, int *rows, int *cols)
{
    if (!c) return NGMI_E_PANIC;
    Plan P;
    const int rc = analyze(*c, P);
    if (rc != NGMI_OK) return rc;
    const int N = (int)P.rowp.size();
    if (n_unknowns) *n_unknowns = N;
    if (nslots) *nslots = P.nslots;
    if (nterms) *nterms = P.nterms;
    if (nnz_l) *nnz_l = P.nnz_l;
    if (levels) {
        levels[0] = P.nlev_f;
        levels[1] = P.nlev_fw;
        levels[2] = P.nlev_bw;
    }
    if (rows) std::memcpy(rows, P.rowp.data(), sizeof(int) * N);
    if (cols) std::memcpy(cols, P.colq.data(), sizeof(int) * N);
    return NGMI_OK;
}

ngmi_batch *ngmi_batch_new(ngmi_circuit *const *ckts, int C, const int *probe, int nprobe,
                           int64_t max_points)
{
    if (!ckts || C < 1) {
        set_error("empty batch");
        return nullptr;
    }
    ngmi_batch *b = new ngmi_batch;
    auto bail = [&](const char *what) -> ngmi_batch * {
        if (what) set_error("%s", what);
        delete b;
        return nullptr;
    };
    b->C = C;
    b->cap_pts = max_points > 0 ? max_points : 1;
    /* circuits sharing a topology form one group, groups in order of first
     * appearance */
    std::vector<std::vector<int>> members;
    b->where.resize(C);
    for (int c = 0; c < C; c++) {
        int g = -1;
        for (size_t q = 0; q < members.size() && g < 0; q++)
            if (same_topology(*ckts[members[q][0]], *ckts[c])) g = (int)q;
        if (g < 0) {
            g = (int)members.size();
            members.emplace_back();
        }
        b->where[c] = make_int2(g, (int)members[g].size());
        members[g].push_back(c);
    }
    if (alloc(b, 1, &b->done_count) != NGMI_OK)
        return bail(nullptr);
    std::vector<TranDev> Ds;
    std::vector<B4Tab> tabs;
    std::vector<int2> mb, sb;
    for (size_t g = 0; g < members.size(); g++) {
        TopoGroup *G = new TopoGroup;
        b->groups.push_back(G);
        G->ids = members[g];
        std::vector<const ngmi_circuit *> cs;
        for (int c : members[g]) cs.push_back(ckts[c]);
        if (build_group(b, G, cs, probe, nprobe) != NGMI_OK) return bail(nullptr);
        Ds.push_back(G->D);
        tabs.push_back(*set_table(G->set));
        /* MOSFET kernel: whole blocks per group; (group, first block) */
        const int64_t cnm = (int64_t)G->C * G->NM;
        const int first = (int)mb.size();
        for (int64_t k = 0; k < (cnm + 255) / 256; k++) mb.push_back(make_int2((int)g, first));
        /* step kernel: one block per circuit; (group, circuit) */
        for (int c = 0; c < G->C; c++) sb.push_back(make_int2((int)g, c));
        b->lds = std::max(b->lds, G->lds);
    }
    b->n_mos_blk = (int)mb.size();
    b->n_step_blk = (int)sb.size();
    if (upload<TranDev>(b, Ds, &b->dD) != NGMI_OK || upload<B4Tab>(b, tabs, &b->dTab) != NGMI_OK ||
        upload<int2>(b, mb, &b->mos_blk) != NGMI_OK || upload<int2>(b, sb, &b->step_blk) != NGMI_OK)
        return bail(nullptr);
    if (b->lds > 64 * 1024 && set_tran_lds_limit(b->lds) != NGMI_OK) return bail(nullptr);
    return b;
}

void ngmi_batch_free(ngmi_batch *b) { delete b; }

int ngmi_batch_info(const ngmi_batch *b, int *n_unknowns, int *nslots, int *nterms, int *levels)
{
    if (!b) return NGMI_E_PANIC;
    const TopoGroup &G = *b->groups[0];
    if (n_unknowns) *n_unknowns = G.N;
    if (nslots) *nslots = G.nslots;
    if (nterms) *nterms = G.nterms;
    if (levels) {
        levels[0] = G.nlev_f;
        levels[1] = G.nlev_fw;
        levels[2] = G.nlev_bw;
    }
    return NGMI_OK;
}

int ngmi_batch_groups(const ngmi_batch *b, int *ngroups)
{
    if (!b || !ngroups) return NGMI_E_PANIC;
    *ngroups = (int)b->groups.size();
    return NGMI_OK;
}

int ngmi_batch_circuit_info(const ngmi_batch *b, int c, int *group, int *n_unknowns, int *nprobe,
                            int *nslots)
{
    if (!b || c < 0 || c >= b->C) return NGMI_E_PANIC;
    const TopoGroup &G = b->group_of(c);
    if (group) *group = b->where[c].x;
    if (n_unknowns) *n_unknowns = G.N;
    if (nprobe) *nprobe = (int)G.probe.size();
    if (nslots) *nslots = G.nslots;
    return NGMI_OK;
}

/* TRANinit (traninit.c:17-40) + DCtran start (dctran.c:117-200) per circuit,
 * then passes until every circuit is done or max_passes is reached. */
/* hand every watched circuit's points recorded since the last call to the
 * data callback (the device is idle: called after the pass loop's sync) */
static int send_data(ngmi_batch *b)
{
    std::vector<double> t, x;
    std::vector<std::vector<CktCtl>> ctl(b->groups.size());   /* one copy per group */
    for (size_t w = 0; w < b->watch.size(); w++) {
        const int c = b->watch[w];
        std::vector<CktCtl> &gk = ctl[b->where[c].x];
        if (gk.empty()) {
            const TopoGroup &G = *b->groups[b->where[c].x];
            gk.resize(G.C);
            NGMI_HIP(hipMemcpy(gk.data(), G.D.ctl, sizeof(CktCtl) * G.C, hipMemcpyDeviceToHost));
        }
        const CktCtl &k = gk[b->where[c].y];
        const int64_t have = std::min<int64_t>(k.npts, b->cap_pts);
        const int64_t first = b->sent[w];
        if (have <= first) continue;
        const TranDev &D = b->group_of(c).D;
        const size_t lc = (size_t)b->where[c].y;
        const int64_t cnt = have - first;
        t.resize((size_t)cnt);
        x.resize((size_t)cnt * D.nprobe);
        NGMI_HIP(hipMemcpy(t.data(), D.out_t + lc * b->cap_pts + first, cnt * sizeof(double),
                           hipMemcpyDeviceToHost));
        NGMI_HIP(hipMemcpy(x.data(), D.out_x + (lc * b->cap_pts + first) * D.nprobe,
                           cnt * D.nprobe * sizeof(double), hipMemcpyDeviceToHost));
        b->sent[w] = have;
        (void)b->data_cb(c, first, cnt, t.data(), x.data(), D.nprobe, b->data_user);
    }
    return NGMI_OK;
}

int ngmi_batch_tran(ngmi_batch *b, double tstep, double tstop, double tmax, int64_t max_passes,
                    void *stream, int64_t *passes_out)
{
    if (!b) return NGMI_E_PANIC;
    if (!(tstep > 0) || !(tstop > 0)) return NGMI_E_PARMVAL;
    const hipStream_t s = (hipStream_t)stream;
    const int C = b->C;
    double maxStep = tmax;
    if (maxStep == 0) maxStep = (tstep < tstop / 50.0) ? tstep : tstop / 50.0;
    for (TopoGroup *G : b->groups) {
        const TranDev &D = G->D;
        std::vector<CktCtl> ctl(G->C);
        for (int c = 0; c < G->C; c++) {
            CktCtl &k = ctl[c];
            std::memset(&k, 0, sizeof(k));
            k.finalTime = tstop;
            k.step = tstep;
            k.maxStep = maxStep;
            k.delmin = 1e-11 * maxStep;
            k.minBreak = 10.0 * k.delmin;
            k.delta = std::min(tstop / 100, tstep) / 10;
            k.breaks[0] = 0;
            k.breaks[1] = tstop;
            k.nbreaks = 2;
            k.time = 0;
            k.firsttime = 1;
            k.mode = NGMI_MODETRANOP | NGMI_MODEINITJCT;
            k.order = 1;
            k.method = G->method;
            k.maxOrder = G->maxord;
            k.phase = PH_OP;
            k.pvre = 1;   /* NIreinit sets NISHOULDREORDER (nireinit.c:49) */
            for (int i = 0; i < 7; i++) k.deltaOld[i] = maxStep;
            k.learned = b->net ? 1 : 0;
            /* CKTop: direct Newton first, or straight to the stepping when
             * noopiter is set (cktop.c:39-52) */
            k.diagGmin = 0.0;
            k.gminCur = D.gmin;
            k.srcFact = 1.0;
            k.opStage = OP_DIRECT;
            if (G->uic) {
                /* .tran uic: CKTop's NIiter is one CKTload on the initial
                 * conditions (niiter.c:41-46); noopiter's stepping would
                 * reload on the load's right-hand side, so it is not taken */
                k.mode |= NGMI_MODEUIC;
            } else if (G->noopiter) {
                if (D.numGminSteps == 1) {
                    k.opStage = OP_DYN;
                    k.opFactor = D.gminFactor;
                    k.opOldGmin = 1e-2;
                    k.diagGmin = k.opOldGmin / k.opFactor;
                    k.opGtarget = std::max(k.gminCur, D.gshunt);
                } else if (D.numGminSteps > 1) {
                    /* spice3_gmin (cktop.c:297-302) */
                    k.opStage = OP_S3_GMIN;
                    k.diagGmin = (D.gshunt == 0) ? k.gminCur : D.gshunt;
                    for (int i = 0; i < D.numGminSteps; i++) k.diagGmin *= D.gminFactor;
                    k.opSub = 0;
                } else if (D.numSrcSteps > 1) {
                    k.opStage = OP_S3_SRC;   /* spice3_src, CKTsrcFact = 0 / steps */
                    k.srcFact = 0.0;
                    k.opSub = 0;
                } else {
                    k.opStage = OP_SRC_ZERO;
                    k.srcFact = 0.0;
                    k.opConvFact = 0.0;
                    k.opStartGmin = k.gminCur;
                }
            }
        }
        NGMI_HIP(hipMemcpyAsync(D.ctl, ctl.data(), sizeof(CktCtl) * G->C, hipMemcpyHostToDevice, s));
        const int64_t cnm = (int64_t)G->C * G->NM, cnc = (int64_t)G->C * G->NCAP;
        NGMI_HIP(hipMemsetAsync(D.x, 0, sizeof(double) * 2 * G->C * (size_t)(G->N + 1), s));
        if (G->uic) {
            /* CKTic (cktic.c:19-80): CKTrhs and CKTrhsOld hold the .ic
             * values at their nodes and 0 elsewhere; the UIC NIiter swaps
             * them and loads on that vector */
            NGMI_HIP(hipMemcpy2DAsync(D.x, sizeof(double) * (size_t)(G->N + 1), D.ic_val,
                                      sizeof(double) * (size_t)(G->NN + 1), sizeof(double) * (size_t)(G->NN + 1),
                                      (size_t)G->C, hipMemcpyDeviceToDevice, s));
        }
        NGMI_HIP(hipMemsetAsync(D.mstate, 0, sizeof(double) * NLEV * B4X__COUNT * (size_t)cnm, s));
        NGMI_HIP(hipMemsetAsync(D.cstate, 0, sizeof(double) * NLEV * 2 * (size_t)cnc, s));
        NGMI_HIP(hipMemsetAsync(D.cache, 0, sizeof(double) * B4C__COUNT * (size_t)cnm, s));
        NGMI_HIP(hipMemsetAsync(D.pv_owner, 0, sizeof(int) * (size_t)std::max(D.pv_slots, 1), s));
        std::vector<double> brk((size_t)G->C * G->NVS, -1.0);
        NGMI_HIP(hipMemcpyAsync(D.vs_brk, brk.data(), brk.size() * sizeof(double),
                                hipMemcpyHostToDevice, s));
        NGMI_HIP(hipStreamSynchronize(s));
    }
    NGMI_HIP(hipMemsetAsync(b->done_count, 0, sizeof(int), s));
    std::fill(b->sent.begin(), b->sent.end(), 0);
    /* the guess rows' counters start from zero (eager, before any capture) */
    if (b->net) {
        NGMI_HIP(hipMemsetAsync(b->gcount, 0, 4 * sizeof(int), s));
        NGMI_HIP(hipMemsetAsync(b->grows, 0, sizeof(unsigned long long), s));
    }
    const TopoGroup &G0 = *b->groups[0];
    const bool one = b->groups.size() == 1;
    const B4Tab *tab0 = set_table(G0.set);
    const bool fused = tune_get(TUNE_GUESS_FUSED) != 0;
    /* two streams: the circuits' halves alternate, so one half's latency-
     * bound step kernel runs beside the other half's BSIM4 kernel (single
     * topology; with the learned guess each half has its own feature rows
     * and network launch, which needs the fused kernel, and profiling then
     * stays on one stream) */
    const bool split = one && C >= 2 && tune_get(TUNE_STREAMS) >= 2 &&
                       (!b->net || (fused && !b->profile));
    const int half = C / 2;
    b->groups[0]->D.gsplit = split ? half : C;
    /* streams 3 or 4 (plain runs without profiling): the circuits in that
     * many parts, parts 2.. each on a stream of its own; part 1 ends at
     * cut2 */
    const int nparts = !split ? 1 : (b->net || b->profile) ? 2 : std::min(tune_get(TUNE_STREAMS), C);
    std::vector<int> cut(nparts + 1);
    for (int p = 0; p <= nparts; p++) cut[p] = (int)((int64_t)C * p / nparts);
    const int cut2 = nparts > 2 ? cut[2] : C;   /* with 2 parts cut[1] = half, as the guess rows assume */
    hipStream_t s2 = nullptr;
    hipEvent_t join = nullptr;
    std::vector<hipStream_t> sx(nparts > 2 ? nparts - 2 : 0, nullptr);
    std::vector<hipEvent_t> jx(sx.size(), nullptr);
    if (split) {
        NGMI_HIP(hipStreamCreateWithFlags(&s2, hipStreamNonBlocking));
        NGMI_HIP(hipEventCreateWithFlags(&join, hipEventDisableTiming));
        for (size_t k = 0; k < sx.size(); k++) {
            NGMI_HIP(hipStreamCreateWithFlags(&sx[k], hipStreamNonBlocking));
            NGMI_HIP(hipEventCreateWithFlags(&jx[k], hipEventDisableTiming));
        }
    }
    if (split) {   /* the other streams start after the resets above */
        NGMI_HIP(hipEventRecord(join, s));
        NGMI_HIP(hipStreamWaitEvent(s2, join, 0));
        for (hipStream_t t : sx) NGMI_HIP(hipStreamWaitEvent(t, join, 0));
    }
    int64_t pass = 0;
    int done = 0;
    const int check_every = 32;
    std::vector<hipEvent_t> ev, ev2, gev;
    if (b->profile) {
        ev.resize(3 * check_every);
        for (auto &e : ev) NGMI_HIP(hipEventCreate(&e));
        if (split) {
            ev2.resize(3 * check_every);
            for (auto &e : ev2) NGMI_HIP(hipEventCreate(&e));
        }
        if (b->net) {
            gev.resize(2 * check_every);
            for (auto &e : gev) NGMI_HIP(hipEventCreate(&e));
        }
    }
    b->mos_ms = b->step_ms = b->guess_ms = 0;
    int rc = NGMI_OK;
    /* Whole chunks of check_every passes are captured once into a HIP graph
     * and replayed: every pass launches the same kernels with the same
     * arguments (all per-circuit state lives in device memory), so a chunk
     * is one graph launch instead of 2-5 launches per pass.  The first
     * chunk runs eagerly (it makes the lazy allocations of the guess
     * network); profiling, which times each launch with events, runs
     * eagerly throughout. */
    const bool use_graph = tune_get(TUNE_GRAPH) && !b->profile;
    hipStream_t cap = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t fork = nullptr;
    /* One pass's learned guess for the circuits of half h (all circuits
     * when not split): the network over the feature rows the step kernels
     * of the previous pass appended (circuits starting a step from the
     * guess), whose outputs become those nodes' starting values.  The fused
     * kernel clears the row counter itself; the layer-by-layer path clears
     * it with a kernel, not a memset node, since the chunk may be captured
     * into a graph. */
    auto guess_pass = [&](hipStream_t sa, int h) -> int {
        const int64_t r0 = h ? (int64_t)G0.D.gsplit * G0.N : 0;   /* the half's first row */
        const int maxrows = (int)(h ? (int64_t)(C - G0.D.gsplit) * G0.N : (int64_t)G0.D.gsplit * G0.N);
        int *count = b->gcount + 2 * h;
        if (fused)
            return mlp_guess_rows(b->net, b->gF + 16 * r0, count, maxrows, b->gdest + r0, b->gbase + r0, G0.D.x,
                                  nullptr, b->grows, sa);
        int r = mlp_forward_rows(b->net, b->gF, b->gy, maxrows, b->gcount, sa);
        if (r == NGMI_OK) r = launch_guess_apply(G0.D, b->gy, b->gdest, b->gbase, b->gcount, maxrows, b->grows, sa);
        if (r == NGMI_OK) r = launch_zero_int(b->gcount, sa);
        return r;
    };
    auto launch_chunk = [&](hipStream_t sa, int count, bool capturing) -> int {
        int r = NGMI_OK;
        if (split && capturing) {
            NGMI_HIP(hipEventRecord(fork, sa));
            NGMI_HIP(hipStreamWaitEvent(s2, fork, 0));
            for (hipStream_t t : sx) NGMI_HIP(hipStreamWaitEvent(t, fork, 0));
        }
        for (int q = 0; q < count && r == NGMI_OK; q++) {
            if (b->net) r = guess_pass(sa, 0);
            if (r != NGMI_OK) break;
            if (split) {
                r = launch_tran_pass(tab0, &G0.D, 0, cut[1], nullptr, nullptr, nullptr, 1, nullptr, cut[1], b->lds, sa,
                                     nullptr);
                if (r == NGMI_OK && b->net) r = guess_pass(s2, 1);
                if (r == NGMI_OK)
                    r = launch_tran_pass(tab0, &G0.D, cut[1], cut2, nullptr, nullptr, nullptr, 1, nullptr,
                                         cut2 - cut[1], b->lds, s2, nullptr);
                for (size_t t = 0; t < sx.size() && r == NGMI_OK; t++)
                    r = launch_tran_pass(tab0, &G0.D, cut[t + 2], cut[t + 3], nullptr, nullptr, nullptr, 1, nullptr,
                                         cut[t + 3] - cut[t + 2], b->lds, sx[t], nullptr);
            } else {
                r = launch_tran_pass(one ? tab0 : nullptr, one ? &G0.D : nullptr, 0, C, b->dTab, b->dD, b->mos_blk,
                                     b->n_mos_blk, b->step_blk, b->n_step_blk, b->lds, sa, nullptr);
            }
        }
        if (split && capturing && r == NGMI_OK) {
            NGMI_HIP(hipEventRecord(join, s2));
            NGMI_HIP(hipStreamWaitEvent(sa, join, 0));
            for (size_t t = 0; t < sx.size(); t++) {
                NGMI_HIP(hipEventRecord(jx[t], sx[t]));
                NGMI_HIP(hipStreamWaitEvent(sa, jx[t], 0));
            }
        }
        return r;
    };
    while (pass < max_passes && rc == NGMI_OK) {
        int k = 0;
        if (use_graph && pass > 0 && max_passes - pass >= check_every) {
            if (!exec) {
                if (hipStreamCreateWithFlags(&cap, hipStreamNonBlocking) != hipSuccess ||
                    (split && hipEventCreateWithFlags(&fork, hipEventDisableTiming) != hipSuccess) ||
                    hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal) != hipSuccess) {
                    set_error("graph capture could not start: %s", hipGetErrorString(hipGetLastError()));
                    rc = NGMI_E_DEVICE;
                    break;
                }
                const int lrc = launch_chunk(cap, check_every, true);
                const hipError_t e = hipStreamEndCapture(cap, &graph);
                if (lrc != NGMI_OK || e != hipSuccess ||
                    hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) {
                    set_error("graph capture failed: %s", hipGetErrorString(e != hipSuccess ? e : hipGetLastError()));
                    rc = lrc != NGMI_OK ? lrc : NGMI_E_DEVICE;
                    break;
                }
            }
            if (hipGraphLaunch(exec, s) != hipSuccess) {
                set_error("graph launch failed: %s", hipGetErrorString(hipGetLastError()));
                rc = NGMI_E_DEVICE;
                break;
            }
            k = check_every;
            pass += check_every;
        } else if (!b->profile) {
            k = (int)std::min<int64_t>(check_every, max_passes - pass);
            rc = launch_chunk(s, k, false);
            pass += k;
            if (rc != NGMI_OK) break;
            if (split && (hipEventRecord(join, s2) != hipSuccess || hipStreamWaitEvent(s, join, 0) != hipSuccess)) {
                set_error("stream join failed: %s", hipGetErrorString(hipGetLastError()));
                rc = NGMI_E_DEVICE;
                break;
            }
            for (size_t t = 0; t < sx.size() && rc == NGMI_OK; t++)
                if (hipEventRecord(jx[t], sx[t]) != hipSuccess || hipStreamWaitEvent(s, jx[t], 0) != hipSuccess) {
                    set_error("stream join failed: %s", hipGetErrorString(hipGetLastError()));
                    rc = NGMI_E_DEVICE;
                }
            if (rc != NGMI_OK) break;
        }
        for (; b->profile && k < check_every && pass < max_passes; k++, pass++) {
            if (b->net) {
                if (b->profile) NGMI_HIP(hipEventRecord(gev[2 * k], s));
                rc = guess_pass(s, 0);
                if (rc != NGMI_OK) break;
                if (b->profile) NGMI_HIP(hipEventRecord(gev[2 * k + 1], s));
            }
            if (split) {
                rc = launch_tran_pass(tab0, &G0.D, 0, half, nullptr, nullptr, nullptr, 1, nullptr,
                                      half, b->lds, s, b->profile ? &ev[3 * k] : nullptr);
                if (rc == NGMI_OK)
                    rc = launch_tran_pass(tab0, &G0.D, half, C, nullptr, nullptr, nullptr, 1, nullptr,
                                          C - half, b->lds, s2, b->profile ? &ev2[3 * k] : nullptr);
            } else {
                rc = launch_tran_pass(one ? tab0 : nullptr, one ? &G0.D : nullptr, 0, C, b->dTab,
                                      b->dD, b->mos_blk, b->n_mos_blk, b->step_blk, b->n_step_blk,
                                      b->lds, s, b->profile ? &ev[3 * k] : nullptr);
            }
            if (rc != NGMI_OK) break;
        }
        if (b->profile && split &&
            (hipEventRecord(join, s2) != hipSuccess || hipStreamWaitEvent(s, join, 0) != hipSuccess)) {
            set_error("stream join failed: %s", hipGetErrorString(hipGetLastError()));
            rc = NGMI_E_DEVICE;
            break;
        }
        if (hipMemcpyAsync(&done, b->done_count, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess) {
            set_error("transient pass failed: %s", hipGetErrorString(hipGetLastError()));
            rc = NGMI_E_DEVICE;
        }
        if (b->profile && rc == NGMI_OK) {
            for (int q = 0; q < k; q++) {
                float a = 0, c2 = 0;
                (void)hipEventElapsedTime(&a, ev[3 * q], ev[3 * q + 1]);
                (void)hipEventElapsedTime(&c2, ev[3 * q + 1], ev[3 * q + 2]);
                b->mos_ms += a;
                b->step_ms += c2;
                if (split) {
                    (void)hipEventElapsedTime(&a, ev2[3 * q], ev2[3 * q + 1]);
                    (void)hipEventElapsedTime(&c2, ev2[3 * q + 1], ev2[3 * q + 2]);
                    b->mos_ms += a;
                    b->step_ms += c2;
                }
                if (b->net) {
                    float g2 = 0;
                    (void)hipEventElapsedTime(&g2, gev[2 * q], gev[2 * q + 1]);
                    b->guess_ms += g2;
                }
            }
        }
        if (rc == NGMI_OK && b->data_cb) rc = send_data(b);
        if (done >= C) break;
    }
    for (auto &e : ev) (void)hipEventDestroy(e);
    for (auto &e : ev2) (void)hipEventDestroy(e);
    for (auto &e : gev) (void)hipEventDestroy(e);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (cap) (void)hipStreamDestroy(cap);
    if (fork) (void)hipEventDestroy(fork);
    if (split) {
        (void)hipStreamSynchronize(s2);
        (void)hipEventDestroy(join);
        (void)hipStreamDestroy(s2);
        for (size_t t = 0; t < sx.size(); t++) {
            (void)hipStreamSynchronize(sx[t]);
            (void)hipEventDestroy(jx[t]);
            (void)hipStreamDestroy(sx[t]);
        }
    }
    /* evaluations and passes from the circuits' own counters */
    b->counters[0] = b->counters[1] = 0;
    for (TopoGroup *G : b->groups) {
        std::vector<CktCtl> k(G->C);
        NGMI_HIP(hipMemcpy(k.data(), G->D.ctl, sizeof(CktCtl) * G->C, hipMemcpyDeviceToHost));
        for (const CktCtl &x : k) {
            b->counters[0] += (unsigned long long)x.passes * G->NM;
            b->counters[1] += (unsigned long long)x.passes;
        }
    }
    if (passes_out) *passes_out = pass;
    if (rc != NGMI_OK) return rc;
    return done >= C ? NGMI_OK : NGMI_E_ITERLIM;
}

int ngmi_batch_set_data_callback(ngmi_batch *b, ngmi_data_cb cb, const int *circuits, int ncircuits,
                                 void *user)
{
    if (!b) return NGMI_E_PANIC;
    b->data_cb = cb;
    b->data_user = user;
    b->watch.clear();
    if (cb) {
        if (!circuits || ncircuits <= 0) {
            for (int c = 0; c < b->C; c++) b->watch.push_back(c);
        } else {
            for (int i = 0; i < ncircuits; i++) {
                if (circuits[i] < 0 || circuits[i] >= b->C) {
                    set_error("data callback: no circuit %d", circuits[i]);
                    b->data_cb = nullptr;
                    b->watch.clear();
                    return NGMI_E_PARMVAL;
                }
                b->watch.push_back(circuits[i]);
            }
        }
    }
    b->sent.assign(b->watch.size(), 0);
    return NGMI_OK;
}

int ngmi_batch_set_guess(ngmi_batch *b, ngmi_mlp *net)
{
    if (!b) return NGMI_E_PANIC;
    if (net && b->groups.size() != 1) {
        set_error("the learned guess needs a batch of one topology");
        return NGMI_E_UNSUPP;
    }
    b->net = net;
    if (!net) return NGMI_OK;
    TopoGroup &G = *b->groups[0];
    TranDev &D = G.D;
    const size_t rows = (size_t)G.C * G.N;
    if (!b->gF) {
        if (alloc(b, rows * 16, &b->gF) || alloc(b, rows, &b->gy) || alloc(b, rows, &b->gdest) ||
            alloc(b, rows, &b->gbase) || alloc(b, 4, &b->gcount) || alloc(b, 1, &b->grows))
            return NGMI_E_DEVICE;
        double *h;
        if (alloc(b, (size_t)4 * G.C * (G.N + 1), &h)) return NGMI_E_DEVICE;
        D.hist = h;
        D.gF = b->gF;
        D.gdest = b->gdest;
        D.gbase = b->gbase;
        D.gcount = b->gcount;
        NGMI_HIP(hipMemcpy(b->dD, &D, sizeof(TranDev), hipMemcpyHostToDevice));
    }
    return NGMI_OK;
}

int ngmi_batch_guess_rows(const ngmi_batch *b, int64_t *rows)
{
    if (!b || !rows) return NGMI_E_PANIC;
    unsigned long long r = 0;
    if (b->grows) NGMI_HIP(hipMemcpy(&r, b->grows, sizeof(r), hipMemcpyDeviceToHost));
    *rows = (int64_t)r;
    return NGMI_OK;
}

int ngmi_batch_neighbours(const ngmi_batch *b, int *out)
{
    if (!b || !out) return NGMI_E_PANIC;
    const TopoGroup &G = *b->groups[0];
    std::memcpy(out, G.nbr_host.data(), G.nbr_host.size() * sizeof(int));
    return NGMI_OK;
}

int ngmi_batch_guess_stats(const ngmi_batch *b, int c, int64_t *guessed, int64_t *fallbacks,
                           int64_t *paused, double *guess_ms)
{
    if (!b || c < 0 || c >= b->C) return NGMI_E_PANIC;
    CktCtl k;
    NGMI_HIP(hipMemcpy(&k, b->ctl_of(c), sizeof(k), hipMemcpyDeviceToHost));
    if (guessed) *guessed = k.guessed;
    if (fallbacks) *fallbacks = k.fallbacks;
    if (paused) *paused = k.gpaused;
    if (guess_ms) *guess_ms = b->guess_ms;
    return NGMI_OK;
}

int ngmi_batch_table_cols(const ngmi_batch *b, int *inst_cols, int *model_cols)
{
    if (!b) return NGMI_E_PANIC;
    return ngmi_bsim4_set_layout(b->groups[0]->set, inst_cols, model_cols, nullptr);
}

int ngmi_batch_set_profile(ngmi_batch *b, int enable)
{
    if (!b) return NGMI_E_PANIC;
    b->profile = enable != 0;
    return NGMI_OK;
}

int ngmi_batch_kernel_times(const ngmi_batch *b, double *mos_ms, double *step_ms,
                            unsigned long long *mos_evals, unsigned long long *newton_passes)
{
    if (!b) return NGMI_E_PANIC;
    if (mos_ms) *mos_ms = b->mos_ms;
    if (step_ms) *step_ms = b->step_ms;
    if (mos_evals) *mos_evals = b->counters[0];
    if (newton_passes) *newton_passes = b->counters[1];
    return NGMI_OK;
}

int ngmi_circuit_add_bsim4_many(ngmi_circuit *c, ngmi_bsim4_model *const *models,
                                const int *model_idx, const int *nodes, int n, int nparams,
                                const char *const *names, const double *values)
{
    if (!c || !models || (n > 0 && (!model_idx || !nodes))) return NGMI_E_PANIC;
    for (int k = 0; k < n; k++) {
        std::vector<double> v(nparams);
        for (int p = 0; p < nparams; p++) v[p] = values[(size_t)p * n + k];
        const int rc = ngmi_circuit_add_bsim4(c, models[model_idx[k]], nodes[4 * k], nodes[4 * k + 1],
                                              nodes[4 * k + 2], nodes[4 * k + 3], nparams, names,
                                              v.data());
        if (rc != NGMI_OK) return rc;
    }
    return NGMI_OK;
}

/* per-circuit results: status (0 ok, else error number), counts */
int ngmi_batch_stats(const ngmi_batch *b, int c, int *status, int64_t *npts, int64_t *accepted,
                     int64_t *rejected, int64_t *iterations)
{
    if (!b || c < 0 || c >= b->C) return NGMI_E_PANIC;
    CktCtl k;
    NGMI_HIP(hipMemcpy(&k, b->ctl_of(c), sizeof(k), hipMemcpyDeviceToHost));
    if (status) *status = k.phase == PH_DONE ? 0 : (k.phase == PH_FAIL ? k.status : -1);
    if (npts) *npts = k.npts;
    if (accepted) *accepted = k.accepted;
    if (rejected) *rejected = k.rejected;
    if (iterations) *iterations = k.iters;
    return NGMI_OK;
}

/* KLU's pivoting in circuit c: pv = 1 while it refactors on pivot rows of
 * its own (then rows[k] = the unknown whose row is the pivot of the column
 * solved k-th, cols[k] = that column's unknown, 1-based; rows and cols may
 * be NULL), and the counts of pivoted factorizations and refactors */
int ngmi_batch_pivot_stats(const ngmi_batch *b, int c, int *pv, int64_t *factors, int64_t *refactors, int *rows,
                           int *cols)
{
    if (!b || c < 0 || c >= b->C) return NGMI_E_PANIC;
    CktCtl k;
    NGMI_HIP(hipMemcpy(&k, b->ctl_of(c), sizeof(k), hipMemcpyDeviceToHost));
    const TopoGroup &G = b->group_of(c);
    const int N = G.N;
    if (pv) *pv = k.pv;
    if (factors) *factors = k.pvfull;
    if (refactors) *refactors = k.pvrefac;
    if (rows || cols) {
        std::vector<int> pr(N), perm(2 * (size_t)N);
        NGMI_HIP(hipMemcpy(pr.data(), G.D.pv_row + (size_t)b->where[c].y * N, N * sizeof(int),
                           hipMemcpyDeviceToHost));
        NGMI_HIP(hipMemcpy(perm.data(), G.D.perm_unknown, 2 * (size_t)N * sizeof(int), hipMemcpyDeviceToHost));
        for (int q = 0; q < N; q++) {
            if (rows) rows[q] = perm[N + (k.pv ? pr[q] : q)];
            if (cols) cols[q] = perm[q];
        }
    }
    return NGMI_OK;
}

/* waveform of circuit c: t[npts], x[npts][nprobe] (host buffers, up to maxpts) */
int ngmi_batch_waveform(const ngmi_batch *b, int c, double *t, double *x, int64_t maxpts,
                        int64_t *npts)
{
    if (!b || c < 0 || c >= b->C) return NGMI_E_PANIC;
    int64_t n = 0;
    CktCtl k;
    NGMI_HIP(hipMemcpy(&k, b->ctl_of(c), sizeof(k), hipMemcpyDeviceToHost));
    const TranDev &D = b->group_of(c).D;
    const size_t lc = (size_t)b->where[c].y;
    n = std::min<int64_t>(std::min<int64_t>(k.npts, b->cap_pts), maxpts);
    if (n > 0) {
        NGMI_HIP(hipMemcpy(t, D.out_t + lc * b->cap_pts, n * sizeof(double), hipMemcpyDeviceToHost));
        NGMI_HIP(hipMemcpy(x, D.out_x + lc * b->cap_pts * D.nprobe, n * D.nprobe * sizeof(double),
                           hipMemcpyDeviceToHost));
    }
    if (npts) *npts = n;
    return NGMI_OK;
}

}  // extern "C"
