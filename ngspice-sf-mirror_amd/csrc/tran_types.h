/*
 * Data shared between the host side of the batched transient engine and its
 * kernels: per-circuit control blocks, the static LU program and the
 * assembly plans.  All device arrays are circuit-major: quantity q of
 * element j of circuit c lives at [q][c * count + j].
 */
#ifndef NGMI_TRAN_TYPES_H
#define NGMI_TRAN_TYPES_H

#include <cstdint>
#include <hip/hip_runtime.h>

namespace ngmi {

constexpr int MAXBRK = 24;       /* breakpoint table per circuit */
constexpr int NLEV = 4;          /* state levels kept: max order 2 + 2 */
constexpr int VSP = 10;          /* doubles per voltage source */

/* circuit phases */
enum : int {
    PH_OP = 0,        /* operating point Newton (CKTop) */
    PH_TRAN = 1,      /* transient Newton (NIiter under DCtran) */
    PH_DONE = 2,      /* reached the final time */
    PH_FAIL = 3       /* error; see status */
};

/* per-circuit control block (scalars of CKTcircuit used by DCtran/NIiter) */
struct CktCtl {
    double time, delta, saveDelta, olddelta;
    double deltaOld[7];
    double ag[7];
    double xfact;
    double breaks[MAXBRK];
    double finalTime, step, maxStep, delmin, minBreak;
    int mode, order, method, maxOrder;
    int phase, status, firsttime, iterno;
    int nbreaks, rb, xsel, pad;
    int64_t npts, accepted, rejected, iters;
    int64_t passes;   /* Newton passes this circuit took part in */
    /* learned initial guess: learned = enabled for this circuit; guess = 0
     * none, 1 features requested for the step being started, 2 guess applied
     * (first Newton iteration runs from it); history ring of the last 4
     * accepted points (times here, node vectors in TranDev::hist) */
    int learned, guess, nhist, hpos;
    double ht[4];
    int64_t guessed, fallbacks;
    /* back-off: gstrike = guessed steps in a row that needed more Newton
     * iterations than the floor of 2 (or failed); gpause = steps still to
     * start from the predictor; gbackoff = the next pause's length */
    int gstrike, gpause, gbackoff, gpad;
    int64_t gpaused;   /* steps started from the predictor during pauses */
    /* CKTop's stepping (cktop.c:161-670): CKTdiagGmin, the devices' CKTgmin,
     * CKTsrcFact, and the scalars of the stage running */
    double diagGmin, gminCur, srcFact;
    double opFactor, opOldGmin, opGtarget, opRaise, opConvFact, opStartGmin;
    int opStage, opSub;
    int64_t opIter0;
    /* KLU's numeric pivoting (niiter.c:107-195 with klu_factor /
     * klu_refactor): pv = 1 while the circuit's pivot rows
     * (TranDev::pv_row) differ from the static sequence, which the step
     * kernel then refactors in dense form; pvre is NIiter's NISHOULDREORDER
     * (the next iteration factors afresh); pvhold = 1 while an iteration
     * whose static pivots failed waits a pass for its dense factorization
     * (its loads kept); pvfull counts pivoted
     * factorizations, pvrefac refactors on their rows */
    int pv, pvre, pvhold, pvpad;
    int64_t pvfull, pvrefac;
    /* OPtran (optran.c:285-851): optran = 1 while the operating point runs
     * as a transient; the transient's final time, step and largest step
     * are kept in trFinal, trStep, trMax meanwhile */
    int optran, oppad;
    double trFinal, trStep, trMax;
};

/* CKTop stages: direct Newton, dynamic_gmin and its final solve, new_gmin
 * and its final solve, gillespie_src's zero-source solve, its gmin ladder and
 * its source ramp */
enum : int {
    OP_DIRECT = 0, OP_DYN = 1, OP_DYN_FINAL = 2, OP_NEW = 3, OP_NEW_FINAL = 4,
    OP_SRC_ZERO = 5, OP_SRC_GMIN = 6, OP_SRC_STEP = 7,
    OP_S3_GMIN = 8, OP_S3_GMIN_FINAL = 9, OP_S3_SRC = 10
};

/* contribution codes of the assembly plans */
enum : uint32_t {
    CT_MOS = 0u, CT_CAP = 1u, CT_RES = 2u, CT_ONE = 3u, CT_VSRC = 4u, CT_ZERO = 5u,
    CT_TYPE_SHIFT = 28u, CT_NEG = 1u << 27, CT_ICZ = 1u << 26,
    CT_INDEX_MASK = (1u << 26) - 1u
};

/* one batch (all circuits share the topology) */
struct TranDev {
    int C;            /* circuits */
    int N;            /* unknowns (1..N, 0 = ground) */
    int NN;           /* node unknowns (1..NN); N - NN branch unknowns */
    int NM, NCAP, NRES, NVS;
    int nslots;       /* nonzeros of L + U */
    int nlev_f, nlev_fw, nlev_bw;
    int nprobe;
    int64_t cap_pts;  /* waveform capacity per circuit */
    double reltol, abstol, vntol, chgtol, trtol, gmin;
    /* CKTop options: gminsteps, srcsteps (0 or 1), itl2, gminfactor, gshunt */
    int numGminSteps, numSrcSteps, dcTrcvMaxIter;
    double gminFactor, gshunt;
    double opstep, opfinal;   /* optran's step and final time; opstep 0: deselected */
    const int *diag_slot;     /* [N+1] L+U slot of A(u, u), -1 where A has no such element */
    double *bak_x;            /* [C][N+1] CKTop's saved CKTrhsOld */
    double *bak_s;            /* [C][29 NM + 2 NCAP] CKTop's saved CKTstate0 */
    /* topology (shared) */
    const int *mos_nodes;     /* [4][NM] d g s b */
    const int *cap_nodes;     /* [2][NCAP] */
    const int *res_nodes;     /* [2][NRES] */
    const int *vs_nodes;      /* [3][NVS] p n branch */
    const int *ic_nodes;      /* [NN+1] 1 where the node has an .ic */
    /* per-circuit element values */
    const double *cap_c;      /* [C][NCAP] */
    const double *cap_ic;     /* [C][NCAP] CAPinitCond: under UIC the .ic difference, else 0 */
    const double *res_g;      /* [C][NRES] */
    const double *vs_par;     /* [C][NVS][VSP]: dc, pulse flag, coefficient count, 7 coefficients */
    double *vs_brk;           /* [C][NVS] break_time */
    const double *ic_val;     /* [C][NN+1] */
    /* LU program */
    /* assembly: leaf e is a 32-byte record {dest, begin, end, flag} + a
     * copy of its first group of four contribution codes, which is group e
     * of a_code; [begin, end) are its further groups (uint4 units of
     * a_code, after all first groups; padded with zero codes).  It sums
     * into L+U slot dest (< nslots) or partial dest - nslots; combines (slot,
     * pbegin, pend, flag) add partials into a slot; likewise for the RHS
     * rows (dest < N a row, else partial n_apart + dest - N).  flag bit 0:
     * the slot is an .ic diagonal.  MOSFET values come
     * from the circuit's contribution arrays, which hold, at the position of
     * each MOSFET code in a_code / r_code, the value the BSIM4 kernel wrote
     * there (mos_pos / rmos_pos: position of stamp s / RHS row r of MOSFET j
     * at [s * NM + j], -1 where the row or column is ground). */
    const int4 *a_leaf, *a_comb, *r_leaf, *r_comb;
    int n_aleaf, n_acomb, n_rleaf, n_rcomb, n_apart, n_rpart;
    int n_aleaf1, n_rleaf1;   /* leading leaves with a single group */
    int n_aplain, n_rplain;   /* of those, the leading plain ones: values used as they are */
    int n_part;               /* LDS partial sums: the assembly's, or the programs' two halves */
    const int *a_dest, *r_dest;   /* destinations of the plain leaves */
    const uint32_t *a_code;   /* contributions */
    const uint32_t *r_code;
    const int *ic_row;        /* permuted rows of .ic nodes */
    int n_icrow;
    /* LU program.  Terms are 16-bit index pairs packed in one word (low |
     * high << 16), grouped four to a uint4 and padded with pairs whose
     * product is exactly zero (LDS lu[nslots] = z[N] = 0).  An entry is a
     * 64-byte record: four ints {target | g << 16, y, begin, end}, then
     * three term groups of which the first g (0..3) are its own; the term
     * range [begin, end) holds the further groups, in uint4 units.  Bits
     * 20-21 of the first int give the kind: 0 an entry, 1 a partial sum of
     * some of a long entry's terms into partial slot target, 2 the combine
     * that adds partial slots [begin, end) into target (then divides by the
     * pivot y when y >= 0). */
    const int *f_lev;         /* [nlev_f+1] -> entries */
    const int *f_ent;         /* [ne][16]: y = pivot slot (-1 for U) */
    const uint32_t *f_term;   /* (a, b) slot pairs: lu[t] -= lu[a] lu[b] */
    const int *fw_lev, *fw_ent;   /* forward solve: [ne][16] target row, y = -1 */
    const uint32_t *fw_term;      /* (L slot, z index) */
    const int *bw_lev, *bw_ent;   /* backward solve: [ne][16] target row, y = diag slot */
    const uint32_t *bw_term;      /* (U slot, z index) */
    const int *unk_pos;       /* [N+1]: permuted column of unknown u (1-based; [0] = N) */
    const int *perm_unknown;  /* [2N]: [a] unknown (1-based) solved in permuted column a;
                                 [N + a] original row assembled into permuted row a */
    const int *probe;         /* [nprobe] unknowns written at each accepted point */
    /* state */
    CktCtl *ctl;              /* [C] */
    double *x;                /* [2][C][N+1] iterate buffers */
    double *mstate;           /* [NLEV][29][C*NM] */
    double *cstate;           /* [NLEV][2][C*NCAP] */
    double *cache;            /* [38][C*NM] */
    int NA, NR;               /* a_code / r_code entries (contribution array lengths) */
    const int *mos_pos;       /* [16 * NM] */
    const int *rmos_pos;      /* [4 * NM] */
    double *contrib;          /* [C][NA] */
    double *rcontrib;         /* [C][NR] */
    int *noncon;              /* [C*NM] */
    double *out_t;            /* [C][cap_pts] */
    double *out_x;            /* [C][cap_pts][nprobe] */
    int *done_count;          /* circuits finished or failed */
    double *hist;             /* [4][C][N+1] accepted node vectors (learned guess) */
    const int *nbr;           /* [NN+1][8] neighbour nodes (0 = none) */
    /* the guess network's rows, appended by the step kernel of a circuit
     * that requests a guess and consumed at the start of the next pass:
     * features [rows][16], the node's offset in x, (extrapolation,
     * activity).  Circuits below gsplit use rows from 0 and the counters
     * gcount[0] (rows) and gcount[1] (the network kernel's finish counter);
     * the others (the second stream's half) rows from gsplit * N and
     * gcount[2], gcount[3]. */
    float *gF;
    int64_t *gdest;
    double2 *gbase;
    int *gcount;
    int gsplit;
    /* pivoting.  The L+U slots are row-major over the permuted matrix: row
     * a holds slots [row_start[a], row_start[a+1]), slot s in permuted
     * column slot_col[s] (ascending).  pivrel is KLU's pivot tolerance
     * (CKTpivotRelTol).  A circuit that needs pivoting takes one of
     * pv_slots dense workspaces (pv_owner[k]: 0 free, else circuit + 1) of
     * pv_stride doubles: the assembled matrix with its right-hand side as
     * column N, twice (a pristine copy and the working one), rows pv_ld
     * doubles apart; pv_slots = 0 when the matrices are too large for
     * dense workspaces (a singular static pivot is then E_SINGULAR). */
    const int *row_start, *slot_col;
    double pivrel;
    double *pv_rs;            /* [C][N] KLU's row scale factors (max |a_ij|, 0 -> 1) */
    int *pv_row;              /* [C][N] permuted pivot row of each permuted column */
    double *pv_pool;
    int *pv_owner;
    int pv_slots, pv_ld;
    int64_t pv_stride;
};

}  // namespace ngmi

#endif
