/*
 * Host-side structure of the whole-GPU multifrontal LU: the static pivot
 * sequence, the frontal tree and every index the numeric kernels use.
 */
#ifndef NGMI_SPLU_TYPES_H
#define NGMI_SPLU_TYPES_H

#include <cstdint>
#include <vector>

namespace ngmi {

constexpr int SPLU_LEAF = 64;          /* nested dissection stops at this many vertices */
constexpr int SPLU_NB = 32;            /* panel width of the front factorization */
constexpr int SPLU_TILE = 32;          /* trailing-update tile edge (16 x 16 threads, 2 x 2 each) */
constexpr int SPLU_MAX_FRONT = 8000;   /* a front's right-hand side and the solve kernels'
                                          static LDS fit 64 KB */

struct SpluSym {
    int n = 0;
    int64_t nnz = 0;
    std::vector<int> rowsrc;             /* row r of B = P A is row rowsrc[r] of A */
    int nf = 0;                          /* fronts, in postorder */
    std::vector<int> ns, nb, parent, height;
    std::vector<int64_t> off;            /* front storage offset (doubles), nf x nf column-major */
    std::vector<int64_t> idx_off;        /* index list offset, ns + nb entries */
    std::vector<int64_t> upd_off;        /* forward-solve update vector offset, nb entries */
    std::vector<int> idx;                /* unknowns: pivots, then boundary by elimination order */
    std::vector<int> rel;                /* at a boundary slot: its index in the parent's list */
    std::vector<int> kid_ptr, kids;
    std::vector<int64_t> ent_pos;        /* CSC entry -> front storage position */
    int H = 0;                           /* largest height */
    std::vector<int> lev_ptr, lev;       /* fronts by height */
    int max_slots = 0;                   /* most children of one front */
    std::vector<int> ea_ptr, ea;         /* children by (parent height, child slot) */
    int max_lev = 0, max_front = 0;
    int64_t store = 0, lu_nnz = 0;
    double flops = 0;
};

int splu_analyze(int n, const int64_t *Ap, const int *Ai, SpluSym &S);

}  // namespace ngmi

#endif
