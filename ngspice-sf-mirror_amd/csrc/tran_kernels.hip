/*
 * Batched transient engine kernels.  One pass = one Newton iteration
 * (NIiter body) for every circuit still running:
 *
 *   k_tran_mos   one thread per (circuit, MOSFET): BSIM4load for the
 *                circuit's current mode, writing 16 Jacobian and 4 RHS
 *                entries, the charge states and the convergence vote;
 *   k_tran_step  one workgroup per circuit: capacitor / resistor / source
 *                loads, assembly into the static L+U slots in LDS, LU
 *                factorisation and forward/back substitution by the
 *                level-scheduled program, NIconvTest, the NIiter INITF state
 *                machine and, when NIiter returns, the DCtran step control
 *                (LTE, order, breakpoints, acceptance, output).
 *
 * Restated control flow: dctran.c:66-971 (XSPICE build without event
 * instances), niiter.c:28-367, nicomcof.c:17-130, niconv.c:21-90,
 * cktterr.c:13-78, ckttrunc.c:20-56, capload.c:19-86, vsrcload.c:40-170,
 * vsrcacct.c:20-160, cktsetbk.c, cktclrbk.c, equality.c:22-60,
 * cktload.c:146-173 (.ic rows in the transient operating point).
 */
#include <algorithm>
#include <climits>
#include "ngmi_internal.h"
#include "bsim4_eval.h"
#include "tran_types.h"
#include "guess_types.h"

namespace ngmi {

#define MODEDC_MASK (M_DCOP | M_TRANOP | M_DCTRANCURVE)
#define INITF_MASK 0x3f00

__device__ __forceinline__ int lvl(int rb, int l) { return (rb + l) & (NLEV - 1); }

/* ---------------- MOSFET evaluation ---------------- */

struct TranPrm {
    const B4Tab &t;
    int64_t i;
    int mi;
    __device__ __forceinline__ double operator()(int f) const
    {
        const int c = t.code[f];
        return c >= 0 ? t.itab[(int64_t)c * t.n + i] : t.mtab[(int64_t)mi * t.mcols + (~c)];
    }
};

struct TranState {
    double *base;
    int64_t cnm, i;
    int rb;
    __device__ __forceinline__ double *p(int l, int k) const
    {
        return base + ((int64_t)lvl(rb, l) * B4X__COUNT + k) * cnm + i;
    }
    __device__ __forceinline__ double s(int l, int k) const { return *p(l, k); }
    __device__ __forceinline__ void set(int l, int k, double v) { *p(l, k) = v; }
};

struct TranCache {
    double *c_;
    int64_t cnm, i;
    __device__ __forceinline__ double c(int k) const { return c_[k * cnm + i]; }
    __device__ __forceinline__ void setc(int k, double v) { c_[k * cnm + i] = v; }
};

/* one thread per (circuit, MOSFET) of a topology group, circuit-major
 * instances i = c * NM + j */
template <int FL>
__device__ __forceinline__ void tran_mos_body(const B4Tab &tab, const TranDev &D, int64_t i)
{
    const int64_t cnm = (int64_t)D.C * D.NM;
    if (i >= cnm) return;
    const int c = (int)(i / D.NM);
    const int j = (int)(i - (int64_t)c * D.NM);
    const CktCtl *K = D.ctl + c;
    /* pvhold: the circuit's iteration continues with the loads it has */
    if (K->phase >= PH_DONE || K->pvhold) return;
    B4Ckt ck;
    /* a circuit whose step starts from the learned guess (set by
     * k_guess_features this pass) limits against its last accepted point */
    ck.mode = K->mode | (K->guess == 1 ? M_GUESSED : 0);
    ck.order = K->order;
    ck.method = K->method;
#pragma unroll
    for (int k = 0; k < 7; k++) ck.ag[k] = K->ag[k];
    ck.xfact = K->delta / K->deltaOld[1];
    ck.gmin = K->gminCur;
    ck.reltol = D.reltol;
    ck.abstol = D.abstol;
    const double *xo = D.x + ((int64_t)K->xsel * D.C + c) * (D.N + 1);
    double v[4];
#pragma unroll
    for (int r = 0; r < 4; r++) v[r] = xo[D.mos_nodes[r * D.NM + j]];
    TranPrm p{tab, i, tab.model_of[i]};
    TranState st{D.mstate, cnm, i, K->rb};
    TranCache ca{D.cache, cnm, i};
    double J[B4S__COUNT], R[B4R__COUNT];
    const int nc = b4_load<FL>(p, st, ca, ck, v, J, R);
    /* straight into the circuit's assembly-ordered contribution arrays */
    double *ca_ = D.contrib + (int64_t)c * D.NA;
    double *cr_ = D.rcontrib + (int64_t)c * D.NR;
#pragma unroll
    for (int k = 0; k < B4S__COUNT; k++) {
        const int p = D.mos_pos[k * D.NM + j];
        if (p >= 0) ca_[p] = J[k];
    }
#pragma unroll
    for (int k = 0; k < B4R__COUNT; k++) {
        const int p = D.rmos_pos[k * D.NM + j];
        if (p >= 0) cr_[p] = R[k];
    }
    D.noncon[i] = nc;
}

/* one topology: the descriptor is a kernel argument, so every table pointer
 * is known to address global memory */
template <int WAVES, int FL>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WAVES, WAVES)))
k_tran_mos(B4Tab tab, TranDev D, int64_t i0, int64_t i1)
{
    const int64_t i = i0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < i1) tran_mos_body<FL>(tab, D, i);
}

/* several topologies in one launch: blk[b] = (group, first block of the
 * group) */
template <int WAVES>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WAVES, WAVES)))
k_tran_mos_groups(const B4Tab *__restrict__ tabs, const TranDev *__restrict__ Ds,
                  const int2 *__restrict__ blk)
{
    const int2 bg = blk[blockIdx.x];
    tran_mos_body<-1>(tabs[bg.x], Ds[bg.x], (int64_t)(blockIdx.x - bg.y) * blockDim.x + threadIdx.x);
}

/* ---------------- scalar helpers (thread 0 of a circuit) ---------------- */

__device__ int almost_equal_ulps(double A, double B, int maxUlps)
{   /* equality.c:22-60 */
    if (A == B) return 1;
    int64_t a = __double_as_longlong(A), b = __double_as_longlong(B);
    if (a < 0) a = (int64_t)0x8000000000000000LL - a;
    if (b < 0) b = (int64_t)0x8000000000000000LL - b;
    int64_t d = a - b;
    if (d < 0) d = -d;
    return d <= maxUlps;
}

__device__ int set_break(CktCtl *K, double time)
{   /* cktsetbk.c:19-100 */
    if (almost_equal_ulps(time, K->time, 3)) return 0;
    if (K->time > time) return -1;
    for (int i = 0; i < K->nbreaks; i++) {
        if (K->breaks[i] > time) {
            if ((K->breaks[i] - time) <= K->minBreak) {
                K->breaks[i] = time;
                return 0;
            }
            if (i > 0 && time - K->breaks[i - 1] <= K->minBreak) return 0;
            if (K->nbreaks >= MAXBRK) return -2;
            for (int j = K->nbreaks; j > i; j--) K->breaks[j] = K->breaks[j - 1];
            K->breaks[i] = time;
            K->nbreaks++;
            return 0;
        }
    }
    if (K->nbreaks && time - K->breaks[K->nbreaks - 1] <= K->minBreak) return 0;
    if (K->nbreaks >= MAXBRK) return -2;
    K->breaks[K->nbreaks++] = time;
    return 0;
}

__device__ void clr_break(CktCtl *K)
{   /* cktclrbk.c:15-40 */
    if (K->nbreaks > 2) {
        for (int j = 1; j < K->nbreaks; j++) K->breaks[j - 1] = K->breaks[j];
        K->nbreaks--;
    } else {
        K->breaks[0] = K->breaks[1];
        K->breaks[1] = K->finalTime;
    }
}

struct Pulse {
    double V1, V2, TD, TR, TF, PW, PER;
};

__device__ Pulse pulse_params(const double *vp, double step)
{   /* vsrcload.c:96-125 */
    Pulse P;
    const int order = (int)vp[2];
    const double *c = vp + 3;
    P.V1 = c[0];
    P.V2 = c[1];
    P.TD = order > 2 ? c[2] : 0.0;
    P.TR = order > 3 && c[3] > 0.0 ? c[3] : step;
    P.TF = order > 4 && c[4] > 0.0 ? c[4] : step;
    P.PW = order > 5 && c[5] >= 0.0 ? c[5] : 0.0;
    P.PER = order > 6 && c[6] > 0.0 ? c[6] : P.TR + P.TF + P.PW;
    return P;
}

__device__ double vsrc_value(const double *vp, int mode, double tnow, double step)
{   /* vsrcload.c:78-170 (DC, PULSE) */
    if (vp[1] == 0.0) return vp[0];
    const Pulse P = pulse_params(vp, step);
    double time = (mode & MODEDC_MASK) ? 0.0 : tnow;
    time -= P.TD;
    if (time > P.PER) {
        const double basetime = P.PER * floor(time / P.PER);
        time -= basetime;
    }
    if (time <= 0 || time >= P.TR + P.PW + P.TF) return P.V1;
    if (time >= P.TR && time <= P.TR + P.PW) return P.V2;
    if (time > 0 && time < P.TR) return P.V1 + (P.V2 - P.V1) * (time) / P.TR;
    return P.V2 + (P.V1 - P.V2) * (time - (P.TR + P.PW)) / P.TF;
}

__device__ __forceinline__ int vsrc_accept(const TranDev &D, int c, CktCtl *K)
{   /* vsrcacct.c:20-160 (PULSE corners) */
    if (!(K->mode & (M_TRAN | M_TRANOP))) return 0;
    for (int k = 0; k < D.NVS; k++) {
        const double *vp = D.vs_par + ((int64_t)c * D.NVS + k) * VSP;
        if (vp[1] == 0.0) continue;
        double *brk = D.vs_brk + (int64_t)c * D.NVS + k;
        const Pulse P = pulse_params(vp, K->step);
        double time = K->time - P.TD;
        if (K->time >= *brk) {
            double wait;
            if (time >= P.PER) {
                const double basetime = P.PER * floor(time / P.PER);
                time -= basetime;
            }
            const double atime = time + K->minBreak;
            if (atime < 0.0) wait = -time;
            else if (atime < P.TR) wait = P.TR - time;
            else if (atime < P.TR + P.PW) wait = P.TR + P.PW - time;
            else if (atime < P.TR + P.PW + P.TF) wait = P.TR + P.PW + P.TF - time;
            else wait = P.PER - time;
            *brk = K->time + wait;
            if (set_break(K, *brk)) return -1;
            *brk -= K->minBreak;
        }
    }
    return 0;
}

__device__ int com_cof(CktCtl *K)
{   /* nicomcof.c:17-130 for orders 1 and 2 (the engine keeps order <= 2);
     * the Gear branch is the reference's small LU written out in its own
     * operation order */
    if (K->method == 1) {
        if (K->order == 1) {
            K->ag[0] = 1 / K->delta;
            K->ag[1] = -1 / K->delta;
        } else if (K->order == 2) {
            const double xmu = 0.5;
            K->ag[0] = 1.0 / K->delta / (1.0 - xmu);
            K->ag[1] = xmu / (1.0 - xmu);
        } else return 104;
        return 0;
    }
    const int ord = K->order;
    if (ord < 1 || ord > 2) return 104;
    for (int k = 0; k < 7; k++) K->ag[k] = 0.0;
    double ag0 = 0.0, ag1 = -1 / K->delta, ag2 = 0.0;
    const double arg_1 = K->deltaOld[0];
    const double m11 = arg_1 / K->delta;
    if (ord == 1) {
        ag1 /= m11;
        ag0 = ag0 - 1.0 * ag1;
        ag0 /= 1.0;
    } else {
        double m21 = m11 * (arg_1 / K->delta);
        const double arg_2 = arg_1 + K->deltaOld[1];
        const double m12 = arg_2 / K->delta;
        double m22 = m12 * (arg_2 / K->delta);
        m21 /= m11;
        m22 -= m21 * m12;
        ag2 = ag2 - m21 * ag1;
        ag2 /= m22;
        ag1 = ag1 - m12 * ag2;
        ag1 /= m11;
        ag0 = ag0 - 1.0 * ag1;
        ag0 = ag0 - 1.0 * ag2;
        ag0 /= 1.0;
    }
    K->ag[0] = ag0;
    K->ag[1] = ag1;
    K->ag[2] = ag2;
    return 0;
}

/* CKTterr (cktterr.c:13-78) on one charge state, orders 1 and 2 (the
 * divided-difference loop of the reference unrolled in its own order);
 * q0..q3 = charge at levels 0..3, c0, c1 = its current at levels 0, 1 */
__device__ double terr_one(double q0, double q1, double q2, double q3, double c0, double c1,
                           const CktCtl *K, const TranDev &D)
{
    const int order = K->order;
    const double volttol = D.abstol + D.reltol * fmax(fabs(c0), fabs(c1));
    double chargetol = fmax(fabs(q0), fabs(q1));
    chargetol = D.reltol * fmax(chargetol, D.chgtol) / K->delta;
    const double tol = fmax(volttol, chargetol);
    const double d0 = K->deltaOld[0], d1 = K->deltaOld[1], d2 = K->deltaOld[2];
    double diff0, factor;
    if (order == 1) {
        double a = (q0 - q1) / d0;
        const double b = (q1 - q2) / d1;
        const double t0 = d1 + d0;
        diff0 = (a - b) / t0;
        factor = .5;                      /* trapCoeff[0] = gearCoeff[0] */
    } else {
        double a = (q0 - q1) / d0, b = (q1 - q2) / d1;
        const double c = (q2 - q3) / d2;
        const double t0 = d1 + d0, t1 = d2 + d1;
        a = (a - b) / t0;
        b = (b - c) / t1;
        const double u0 = t1 + d0;
        diff0 = (a - b) / u0;
        factor = (K->method == 2) ? .2222222222 : .08333333333;
    }
    double del = D.trtol * tol / fmax(D.abstol, factor * fabs(diff0));
    if (order == 2) del = sqrt(del);
    return del;
}

/* ---------------- per-circuit step kernel ---------------- */

enum : int { ACT_NONE = 0, ACT_ACCEPT = 1, ACT_LTE1 = 2, ACT_RETRY = 3, ACT_STOP = 4,
             ACT_OPMEM = 100 };

/* ---------------- CKTop stepping (cktop.c:26-670) ---------------- */

/* results of op_next: the operating point converged, failed (status set), or
 * a stage continues after the workgroup does nothing / zeroes / saves /
 * restores CKTrhsOld and CKTstate0 */
enum : int { OPA_OK = 0, OPA_FAIL = 1, OPA_NONE = 2, OPA_ZERO = 3, OPA_SAVE = 4, OPA_RESTORE = 5 };

constexpr int OP_FIRSTMODE = M_TRANOP | M_INITJCT;
constexpr int OP_CONTINUEMODE = M_TRANOP | M_INITFLOAT;

/* start of one NIiter call of a stage */
__device__ __forceinline__ void op_niiter(CktCtl *K, int stage)
{
    K->opStage = stage;
    K->iterno = 0;
    K->opIter0 = K->iters;
}

__device__ __forceinline__ int op_src_init(const TranDev &D, CktCtl *K)
{   /* gillespie_src start (cktop.c:480-503) */
    K->mode = (K->mode & M_UIC) | OP_FIRSTMODE;
    K->srcFact = 0;
    K->opConvFact = 0;
    K->opStartGmin = K->gminCur;
    op_niiter(K, OP_SRC_ZERO);
    return OPA_ZERO;
}

__device__ __forceinline__ int op_s3gmin_init(const TranDev &D, CktCtl *K)
{   /* spice3_gmin start (cktop.c:289-303) */
    K->mode = (K->mode & M_UIC) | OP_FIRSTMODE;
    K->diagGmin = (D.gshunt == 0) ? K->gminCur : D.gshunt;
    for (int i = 0; i < D.numGminSteps; i++) K->diagGmin *= D.gminFactor;
    K->opSub = 0;
    op_niiter(K, OP_S3_GMIN);
    return OPA_NONE;
}

__device__ __forceinline__ int op_s3src_init(const TranDev &D, CktCtl *K)
{   /* spice3_src start (cktop.c:672-686) */
    K->mode = (K->mode & M_UIC) | OP_FIRSTMODE;
    K->opSub = 0;
    K->srcFact = 0.0;
    op_niiter(K, OP_S3_SRC);
    return OPA_NONE;
}

/* source stepping after the gmin stepping failed or was not asked for
 * (cktop.c:78-86) */
__device__ __forceinline__ int op_src_any(const TranDev &D, CktCtl *K, int rc)
{
    if (D.numSrcSteps == 1) return op_src_init(D, K);
    if (D.numSrcSteps > 1) return op_s3src_init(D, K);
    K->status = rc;
    return OPA_FAIL;
}

__device__ int op_src_ramp_init(CktCtl *K)
{   /* the source ramp's start (cktop.c:559-575) */
    K->opRaise = 0.001;
    K->srcFact = K->opConvFact + K->opRaise;
    op_niiter(K, OP_SRC_STEP);
    return OPA_SAVE;
}

__device__ int op_src_end(CktCtl *K, bool ramped)
{   /* gillespie_src end (cktop.c:645-668) */
    if (ramped) K->diagGmin = K->gminCur = K->opStartGmin;
    K->srcFact = 1;
    if (K->opConvFact != 1) {
        K->status = 103;
        return OPA_FAIL;
    }
    return OPA_OK;
}

/* the next step of CKTop after an NIiter call of the current stage returned
 * rc, in the order of cktop.c; thread 0 */
__device__ __forceinline__ int op_next(const TranDev &D, CktCtl *K, int rc)
{
    const int iters = (int)(K->iters - K->opIter0);
    const int uic = K->mode & M_UIC;
    const int lo = D.dcTrcvMaxIter / 4, hi = 3 * D.dcTrcvMaxIter / 4;
    switch (K->opStage) {
    case OP_DIRECT:
        if (rc == 0) return OPA_OK;
        if (D.numGminSteps > 1) return op_s3gmin_init(D, K);
        if (D.numGminSteps == 1) {
            /* dynamic_gmin (cktop.c:161-196) */
            K->mode = uic | OP_FIRSTMODE;
            K->opFactor = D.gminFactor;
            K->opOldGmin = 1e-2;
            K->diagGmin = K->opOldGmin / K->opFactor;
            K->opGtarget = fmax(K->gminCur, D.gshunt);
            op_niiter(K, OP_DYN);
            return OPA_ZERO;
        }
        return op_src_any(D, K, rc);
    case OP_DYN:
    case OP_NEW: {
        /* the loops of dynamic_gmin (cktop.c:197-270) and new_gmin
         * (cktop.c:388-458): the former steps CKTdiagGmin, the latter the
         * devices' CKTgmin */
        const bool dyn = K->opStage == OP_DYN;
        double g = dyn ? K->diagGmin : K->gminCur;
        int mem = OPA_NONE;
        bool final_ = false;
        if (rc == 0) {
            K->mode = uic | OP_CONTINUEMODE;
            if (g <= K->opGtarget) {
                final_ = true;
            } else {
                mem = OPA_SAVE;
                double f = K->opFactor;
                if (iters <= lo) {
                    f *= sqrt(f);
                    if (f > D.gminFactor) f = D.gminFactor;
                }
                if (iters > hi) f = fmax(sqrt(f), dyn ? 1.00005 : 3.0);
                K->opOldGmin = g;
                if (g < f * K->opGtarget) {
                    f = g / K->opGtarget;
                    g = K->opGtarget;
                } else {
                    g /= f;
                }
                K->opFactor = f;
            }
        } else {
            if (K->opFactor < 1.00005) {
                final_ = true;
            } else {
                K->opFactor = sqrt(sqrt(K->opFactor));
                g = K->opOldGmin / K->opFactor;
                mem = OPA_RESTORE;
            }
        }
        if (final_) {
            if (dyn) {
                K->diagGmin = D.gshunt;
                op_niiter(K, OP_DYN_FINAL);
            } else {
                K->gminCur = fmax(K->opStartGmin, D.gshunt);
                op_niiter(K, OP_NEW_FINAL);
            }
            return OPA_NONE;
        }
        if (dyn) K->diagGmin = g;
        else K->gminCur = g;
        op_niiter(K, K->opStage);
        return mem;
    }
    case OP_DYN_FINAL:
        if (rc == 0) return OPA_OK;
        /* new_gmin (cktop.c:349-386) */
        K->mode = uic | OP_FIRSTMODE;
        K->opStartGmin = K->gminCur;
        K->opFactor = D.gminFactor;
        K->opOldGmin = 1e-2;
        K->gminCur = K->opOldGmin / K->opFactor;
        K->opGtarget = fmax(K->opStartGmin, D.gshunt);
        op_niiter(K, OP_NEW);
        return OPA_ZERO;
    case OP_NEW_FINAL:
    case OP_S3_GMIN_FINAL:
        if (rc == 0) return OPA_OK;
        return op_src_any(D, K, rc);
    case OP_S3_GMIN:
        /* the spice3_gmin loop (cktop.c:304-324) */
        if (rc == 0) {
            K->diagGmin /= D.gminFactor;
            K->mode = uic | OP_CONTINUEMODE;
            if (++K->opSub <= D.numGminSteps) {
                op_niiter(K, OP_S3_GMIN);
                return OPA_NONE;
            }
        }
        K->diagGmin = D.gshunt;
        op_niiter(K, OP_S3_GMIN_FINAL);
        return OPA_NONE;
    case OP_S3_SRC:
        /* the spice3_src loop (cktop.c:688-712) */
        K->mode = uic | OP_CONTINUEMODE;
        if (rc != 0) {
            K->srcFact = 1.0;
            K->status = rc;
            return OPA_FAIL;
        }
        if (++K->opSub <= D.numSrcSteps) {
            K->srcFact = ((double)K->opSub) / ((double)D.numSrcSteps);
            op_niiter(K, OP_S3_SRC);
            return OPA_NONE;
        }
        K->srcFact = 1.0;
        return OPA_OK;
    case OP_SRC_ZERO:
        if (rc == 0) return op_src_ramp_init(K);
        /* gmin ladder for the zero-source solution (cktop.c:517-550) */
        K->diagGmin = (D.gshunt <= 0) ? K->gminCur : D.gshunt;
        for (int i = 0; i < 10; i++) K->diagGmin *= 10;
        K->opSub = 0;
        op_niiter(K, OP_SRC_GMIN);
        return OPA_NONE;
    case OP_SRC_GMIN:
        if (rc != 0) {
            K->diagGmin = D.gshunt;
            return op_src_end(K, false);
        }
        K->diagGmin /= 10;
        K->mode = uic | OP_CONTINUEMODE;
        if (++K->opSub <= 10) {
            op_niiter(K, OP_SRC_GMIN);
            return OPA_NONE;
        }
        K->diagGmin = D.gshunt;
        return op_src_ramp_init(K);
    case OP_SRC_STEP: {
        /* the source ramp (cktop.c:577-640) */
        int mem = OPA_NONE;
        K->mode = uic | OP_CONTINUEMODE;
        if (rc == 0) {
            K->opConvFact = K->srcFact;
            mem = OPA_SAVE;
            K->srcFact = K->opConvFact + K->opRaise;
            if (iters <= lo) K->opRaise *= 1.5;
            if (iters > hi) K->opRaise *= 0.5;
        } else {
            if (K->srcFact - K->opConvFact < 1e-8) return op_src_end(K, true);
            K->opRaise /= 10;
            if (K->opRaise > 0.01) K->opRaise = 0.01;
            K->srcFact = K->opConvFact;
            mem = OPA_RESTORE;
        }
        if (K->srcFact > 1) K->srcFact = 1;
        /* the loop ends: the saved copy is no longer needed */
        if (!((K->opRaise >= 1e-7) && (K->opConvFact < 1))) return op_src_end(K, true);
        op_niiter(K, OP_SRC_STEP);
        return mem;
    }
    default:
        K->status = 1;
        return OPA_FAIL;
    }
}

/* the workgroup's part of a stepping transition: zero, save or restore
 * CKTrhsOld and CKTstate0 of circuit c */
__device__ __forceinline__ void op_memory(const TranDev &D, int c, const CktCtl *K, int what)
{
    if (what == OPA_NONE) return;
    const int tid = threadIdx.x, nt = blockDim.x, N = D.N;
    const int64_t cnm = (int64_t)D.C * D.NM, cnc = (int64_t)D.C * D.NCAP;
    double *x = D.x + ((int64_t)K->xsel * D.C + c) * (N + 1);
    double *bx = D.bak_x + (int64_t)c * (N + 1);
    double *bs = D.bak_s + (int64_t)c * (B4X__COUNT * D.NM + 2 * D.NCAP);
    const int l0 = lvl(K->rb, 0);
    for (int u = tid; u <= N; u += nt) {
        if (what == OPA_ZERO) x[u] = 0.0;
        else if (what == OPA_SAVE) bx[u] = x[u];
        else x[u] = bx[u];
    }
    for (int q = tid; q < B4X__COUNT * D.NM; q += nt) {
        const int k = q / D.NM, j = q - k * D.NM;
        double *p = D.mstate + ((int64_t)l0 * B4X__COUNT + k) * cnm + (int64_t)c * D.NM + j;
        if (what == OPA_ZERO) *p = 0.0;
        else if (what == OPA_SAVE) bs[q] = *p;
        else *p = bs[q];
    }
    for (int q = tid; q < 2 * D.NCAP; q += nt) {
        const int k = q / D.NCAP, j = q - k * D.NCAP;
        double *p = D.cstate + ((int64_t)l0 * 2 + k) * cnc + (int64_t)c * D.NCAP + j;
        double *b = bs + B4X__COUNT * D.NM + q;
        if (what == OPA_ZERO) *p = 0.0;
        else if (what == OPA_SAVE) *b = *p;
        else *p = *b;
    }
}

/* minimum of one value per lane over the step kernel's workgroup, which
 * is one wavefront: register exchanges, no LDS (min is exact in any order) */
__device__ double wave_min(double v)
{
#pragma unroll
    for (int s = 32; s > 0; s >>= 1) v = fmin(v, __shfl_xor(v, s));
    return v;
}

/* LTE over every charge state of the circuit (CAPtrunc + BSIM4trunc) */
__device__ __forceinline__ double circuit_terr(const TranDev &D, const CktCtl *K, int c)
{
    const int rb = K->rb, order = K->order;
    const int64_t cnm = (int64_t)D.C * D.NM, cnc = (int64_t)D.C * D.NCAP;
    double m = 1e308;
    for (int k = threadIdx.x; k < D.NCAP; k += blockDim.x) {
        const int64_t ci = (int64_t)c * D.NCAP + k;
        const double *Q = D.cstate + ci;
        const double q0 = Q[(int64_t)lvl(rb, 0) * 2 * cnc], q1 = Q[(int64_t)lvl(rb, 1) * 2 * cnc];
        const double q2 = Q[(int64_t)lvl(rb, 2) * 2 * cnc];
        const double q3 = order > 1 ? Q[(int64_t)lvl(rb, 3) * 2 * cnc] : 0.0;
        const double c0 = Q[((int64_t)lvl(rb, 0) * 2 + 1) * cnc];
        const double c1 = Q[((int64_t)lvl(rb, 1) * 2 + 1) * cnc];
        m = fmin(m, terr_one(q0, q1, q2, q3, c0, c1, K, D));
    }
    for (int t = threadIdx.x; t < 3 * D.NM; t += blockDim.x) {
        const int j = t / 3, r = t - 3 * j;
        const int which = r == 0 ? B4X_qb : (r == 1 ? B4X_qg : B4X_qd);
        const int64_t mi = (int64_t)c * D.NM + j;
        const double *Q = D.mstate + (int64_t)which * cnm + mi;
        const int64_t lv = (int64_t)B4X__COUNT * cnm;
        const double q0 = Q[lvl(rb, 0) * lv], q1 = Q[lvl(rb, 1) * lv], q2 = Q[lvl(rb, 2) * lv];
        const double q3 = order > 1 ? Q[lvl(rb, 3) * lv] : 0.0;
        const double c0 = Q[lvl(rb, 0) * lv + cnm], c1 = Q[lvl(rb, 1) * lv + cnm];
        m = fmin(m, terr_one(q0, q1, q2, q3, c0, c1, K, D));
    }
    return wave_min(m);
}

/* copy state level `from` to level `to` for every state of the circuit */
__device__ __forceinline__ void copy_level(const TranDev &D, int c, int rb, int from, int to)
{
    const int64_t cnm = (int64_t)D.C * D.NM, cnc = (int64_t)D.C * D.NCAP;
    const int pf = lvl(rb, from), pt = lvl(rb, to);
    for (int t = threadIdx.x; t < B4X__COUNT * D.NM; t += blockDim.x) {
        const int k = t / D.NM, j = t - k * D.NM;
        const int64_t mi = (int64_t)c * D.NM + j;
        D.mstate[((int64_t)pt * B4X__COUNT + k) * cnm + mi] =
            D.mstate[((int64_t)pf * B4X__COUNT + k) * cnm + mi];
    }
    for (int t = threadIdx.x; t < 2 * D.NCAP; t += blockDim.x) {
        const int k = t / D.NCAP, j = t - k * D.NCAP;
        const int64_t ci = (int64_t)c * D.NCAP + j;
        D.cstate[((int64_t)pt * 2 + k) * cnc + ci] = D.cstate[((int64_t)pf * 2 + k) * cnc + ci];
    }
}

__device__ __forceinline__ void fail(const TranDev &D, CktCtl *K, int code)
{
    K->phase = PH_FAIL;
    K->status = code;
    atomicAdd(D.done_count, 1);
}

/* OPtran's set-up (optran.c:297-420): the transient's time scalars are
 * kept aside; CKTstep = CKTmaxStep = opstep, final time opfinal, its own
 * breakpoints {0, opfinal}, the first step's delta, INITTRAN */
__device__ void optran_start(const TranDev &D, CktCtl *K)
{
    K->optran = 1;
    K->phase = PH_TRAN;
    K->trFinal = K->finalTime;
    K->trStep = K->step;
    K->trMax = K->maxStep;
    K->finalTime = D.opfinal;
    K->step = K->maxStep = D.opstep;
    K->delta = fmin(D.opfinal / 100, K->step) / 10;
    K->time = 0;
    K->breaks[0] = 0;
    K->breaks[1] = D.opfinal;
    K->nbreaks = 2;
    K->firsttime = 1;
    K->order = 1;
    for (int i = 0; i < 7; i++) K->deltaOld[i] = K->maxStep;
    K->saveDelta = D.opfinal / 50;
    K->mode = (K->mode & M_UIC) | M_TRAN | M_INITTRAN;
    K->ag[0] = K->ag[1] = 0;
}

/* OPtran reached its final time: CKTmaxStep and CKTstep restored, and
 * DCtran goes on as after a converged CKTop (dctran.c:268-330) with the
 * scalars it set before it (time 0, the first delta, breakpoints
 * {0, final time}) */
__device__ void optran_end(CktCtl *K)
{
    K->optran = 0;
    K->finalTime = K->trFinal;
    K->step = K->trStep;
    K->maxStep = K->trMax;
    K->time = 0;
    K->delta = fmin(K->finalTime / 100, K->step) / 10;
    K->breaks[0] = 0;
    K->breaks[1] = K->finalTime;
    K->nbreaks = 2;
    K->firsttime = 1;
    K->order = 1;
    for (int i = 0; i < 7; i++) K->deltaOld[i] = K->maxStep;
    K->saveDelta = K->finalTime / 50;
    K->mode = (K->mode & M_UIC) | M_TRAN | M_INITTRAN;
    K->ag[0] = K->ag[1] = 0;
}

/* DCtran "resume:" block (dctran.c:459-600, XSPICE branch) and state
 * rotation.  OPtran's (optran.c:490-600) has no 1.9-step rule, and its loop
 * that throws out passed breakpoints has no final-time guard: once the table
 * holds only the final time and the time is within CKTminBreak of it,
 * OPclrBreak changes nothing and the reference would loop; here it ends. */
__device__ void resume_block(CktCtl *K)
{
    K->delta = fmin(K->delta, K->maxStep);
    if (almost_equal_ulps(K->time, K->breaks[0], 100) || K->breaks[0] - K->time <= K->delmin) {
        K->order = 1;
        K->delta = fmin(K->delta, .1 * fmin(K->saveDelta, K->breaks[1] - K->breaks[0]));
        if (K->firsttime) {
            /* dctran.c:509-512: a breakpoint at the first step under UIC;
             * OPtran never runs under UIC (its operating point succeeds) */
            if ((K->mode & M_UIC) && !K->optran) set_break(K, K->step);
            K->delta /= 10;
        }
    }
    if (K->optran) {
        while ((almost_equal_ulps(K->breaks[0], K->time, 100) || K->breaks[0] <= K->time + K->minBreak) &&
               !(K->nbreaks == 2 && K->breaks[0] == K->finalTime && K->breaks[1] == K->finalTime))
            clr_break(K);
        if (K->time + K->delta > K->breaks[0]) {
            K->saveDelta = K->delta;
            K->delta = K->breaks[0] - K->time;
        }
    } else {
        while ((K->breaks[0] <= K->time + K->minBreak || almost_equal_ulps(K->breaks[0], K->time, 100)) &&
               K->breaks[0] < K->finalTime)
            clr_break(K);
        if (K->time + K->delta > K->breaks[0]) {
            K->saveDelta = K->delta;
            K->delta = K->breaks[0] - K->time;
        } else if (K->time + 1.9 * K->delta > K->breaks[0]) {
            K->saveDelta = K->delta;
            K->delta = (K->breaks[0] - K->time) / 2.;
        }
    }
    for (int i = 5; i >= 0; i--) K->deltaOld[i + 1] = K->deltaOld[i];
    K->deltaOld[0] = K->delta;
    K->rb = (K->rb + NLEV - 1) & (NLEV - 1);   /* states[i+1] = states[i], states[0] = old last */
}

/* top of the DCtran inner loop: advance time, integration coefficients */
__device__ int loop_top(CktCtl *K)
{
    K->olddelta = K->delta;
    K->time += K->delta;
    K->deltaOld[0] = K->delta;
    K->iterno = 0;
    return com_cof(K);
}

/* The learned guess's back-off.  NIiter cannot stop before its second
 * iteration, so a step that starts from the guess and needs more than two
 * has cost more than the guess can ever save.  After GUESS_STRIKES such
 * steps in a row (a failed run counts too) the circuit starts its next
 * gbackoff steps from the predictor, and each further pause is twice as
 * long (up to GUESS_PAUSE_MAX): a circuit on which the network is poor
 * falls back to the plain method instead of holding the whole batch. */
constexpr int GUESS_STRIKES = 3;
constexpr int GUESS_PAUSE_MIN = 32;
constexpr int GUESS_PAUSE_MAX = 4096;

__device__ void guess_outcome(CktCtl *K, bool slow)
{
    if (!slow) {
        K->gstrike = 0;
        return;
    }
    if (++K->gstrike < GUESS_STRIKES) return;
    K->gstrike = 0;
    const int b = K->gbackoff < GUESS_PAUSE_MIN ? GUESS_PAUSE_MIN : K->gbackoff;
    K->gpause = b;
    K->gbackoff = b < GUESS_PAUSE_MAX ? 2 * b : GUESS_PAUSE_MAX;
}

/* start the next Newton run from the learned guess when the circuit has
 * four accepted points of history and is not pausing; true when it will */
__device__ bool request_guess(CktCtl *K)
{
    if (K->learned && K->nhist >= 4 && !K->firsttime && K->phase == PH_TRAN && !K->optran) {
        if (K->gpause > 0) {
            K->gpause--;
            K->gpaused++;
            return false;
        }
        K->guess = 1;
        K->mode = (K->mode & M_UIC) | M_TRAN | M_INITFLOAT;
        return true;
    }
    return false;
}

/* a read-only table load through the global address space (the tables'
 * pointers come from the group descriptor, so the compiler cannot infer
 * the space and would emit flat loads) */
template <class T>
__device__ __forceinline__ T gload(const T *p)
{
    return *(const __attribute__((address_space(1))) T *)p;
}

__device__ __forceinline__ int4 gload(const int4 *p)
{
    typedef int v4 __attribute__((ext_vector_type(4)));
    const v4 v = *(const __attribute__((address_space(1))) v4 *)p;
    return make_int4(v.x, v.y, v.z, v.w);
}

__device__ __forceinline__ double4 gload4(const double4 *p)
{
    typedef double v2 __attribute__((ext_vector_type(2)));
    const __attribute__((address_space(1))) v2 *q = (const __attribute__((address_space(1))) v2 *)p;
    const v2 a = q[0], b = q[1];
    return make_double4(a.x, a.y, b.x, b.y);
}

__device__ __forceinline__ uint4 gload(const uint4 *p)
{
    typedef unsigned int v4 __attribute__((ext_vector_type(4)));
    const v4 v = *(const __attribute__((address_space(1))) v4 *)p;
    return make_uint4(v.x, v.y, v.z, v.w);
}

/* a 32-byte program record: four ints, then one group of four packed
 * words (terms or contribution codes) */
struct Rec {
    int4 h;
    uint4 g;
};

__device__ __forceinline__ Rec load_rec(const int4 *base, int e)
{
    Rec r;
    r.h = gload(base + 2 * e);
    r.g = gload(reinterpret_cast<const uint4 *>(base + 2 * e + 1));
    return r;
}

/* acc - x[a] y[b] for the four packed index pairs of w, in order */
__device__ __forceinline__ double sub_terms(double acc, uint4 w, const double *x, const double *y)
{
    acc -= x[w.x & 0xffffu] * y[w.x >> 16];
    acc -= x[w.y & 0xffffu] * y[w.y >> 16];
    acc -= x[w.z & 0xffffu] * y[w.z >> 16];
    acc -= x[w.w & 0xffffu] * y[w.w >> 16];
    return acc;
}

/* a 64-byte LU or substitution record: {target | g << 16, y, begin, end},
 * then three term groups of which the first g (0..3) are the entry's */
struct Prog {
    int4 h;
    uint4 g0, g1, g2;
};

__device__ __forceinline__ Prog load_prog(const int4 *base, int e)
{
    Prog r;
    r.h = gload(base + 4 * e);
    r.g0 = gload(reinterpret_cast<const uint4 *>(base + 4 * e + 1));
    r.g1 = gload(reinterpret_cast<const uint4 *>(base + 4 * e + 2));
    r.g2 = gload(reinterpret_cast<const uint4 *>(base + 4 * e + 3));
    return r;
}

/* further term groups loaded per batch by the one lane that owns a long
 * row (a supply node's row has ~45 groups on the adder) */
constexpr int TERM_BATCH = 8;

/* acc minus the record's term groups, in order: the inline ones, then the
 * further groups [h.z, h.w), TERM_BATCH at a time, loaded unconditionally
 * (clamped to the last group) so that the wait counts stay exact */
__device__ __forceinline__ double lu_row_terms(double acc, const Prog &r, const uint4 *T,
                                               const double *x, const double *y)
{
    const int g = (r.h.x >> 16) & 3;
    if (g > 0) acc = sub_terms(acc, r.g0, x, y);
    if (g > 1) acc = sub_terms(acc, r.g1, x, y);
    if (g > 2) acc = sub_terms(acc, r.g2, x, y);
    for (int q = r.h.z; q < r.h.w; q += TERM_BATCH) {
        uint4 w[TERM_BATCH];
#pragma unroll
        for (int k = 0; k < TERM_BATCH; k++) w[k] = gload(T + min(q + k, r.h.w - 1));
#pragma unroll
        for (int k = 0; k < TERM_BATCH; k++)
            if (q + k < r.h.w) acc = sub_terms(acc, w[k], x, y);
    }
    return acc;
}

/* record kinds (bits 20-21 of the header's first int) */
constexpr int PROG_PARTIAL = 1, PROG_COMBINE = 2;

/* acc plus the partial sums [h.z, h.w) of a combine record, in order */
__device__ __forceinline__ double add_parts(double acc, const Prog &r, const double *part)
{
    for (int p = r.h.z; p < r.h.w; p++) acc += part[p];
    return acc;
}

/* Level-scheduled program: entries of level L (lev[L] .. lev[L+1]) are
 * independent; a barrier separates levels.  Lane t runs the entries
 * lev[0] + t + k nt, whatever their level, and keeps the records of its
 * next two entries in flight.  Where levels are narrow (a chain of a few
 * entries each) those records belong to levels far ahead, so their loads
 * overlap many levels' work. */
template <class F>
__device__ __forceinline__ void run_levels(const int *lev, const int4 *ent, int nlev, int tid, int nt,
                                           F body)
{   /* lev: the level table staged in LDS */
    if (nlev <= 0) return;
    const int total = lev[nlev];
    if (total <= lev[0]) return;
    int e = lev[0] + tid;
    /* unconditional prefetches (clamped to the last entry), so that the
     * compiler's wait counts stay exact */
    Prog p0 = load_prog(ent, min(e, total - 1));
    Prog p1 = load_prog(ent, min(e + nt, total - 1));
    for (int L = 0; L < nlev; L++) {
        const int end = lev[L + 1];
        while (e < end) {
            const Prog cur = p0;
            p0 = p1;
            p1 = load_prog(ent, min(e + 2 * nt, total - 1));
            body(cur);
            e += nt;
        }
        __syncthreads();
    }
}

/* Plain assembly leaves 0 .. n-1 (one group, values used as they are):
 * lane t sums leaves t + k nt from their destination (4 bytes) and their
 * four values, U leaves per batch with the next batch's loads in flight;
 * past-the-end loads are clamped to the last leaf so that the wait counts
 * stay exact. */
template <int U, class F>
__device__ __forceinline__ void run_plain(const int *dest, const double4 *val, int n, int tid, int nt,
                                          F body)
{
    if (n <= 0) return;
    /* two half-batches of U/2 leaves each, held in scalar registers */
    constexpr int H = U / 2;
    int da[H], db[H];
    double4 va[H], vb[H];
#pragma unroll
    for (int k = 0; k < H; k++) {
        const int e = min(tid + k * nt, n - 1), f = min(tid + (H + k) * nt, n - 1);
        da[k] = gload(dest + e);
        va[k] = gload4(val + e);
        db[k] = gload(dest + f);
        vb[k] = gload4(val + f);
    }
    for (int e0 = tid; e0 < n; e0 += U * nt) {
        int dc[H], dd[H];
        double4 vc[H], vd[H];
#pragma unroll
        for (int k = 0; k < H; k++) {
            dc[k] = da[k];
            vc[k] = va[k];
            dd[k] = db[k];
            vd[k] = vb[k];
        }
#pragma unroll
        for (int k = 0; k < H; k++) {
            const int e = min(e0 + (U + k) * nt, n - 1), f = min(e0 + (U + H + k) * nt, n - 1);
            da[k] = gload(dest + e);
            va[k] = gload4(val + e);
            db[k] = gload(dest + f);
            vb[k] = gload4(val + f);
        }
#pragma unroll
        for (int k = 0; k < H; k++)
            if (e0 + k * nt < n) body(dc[k], vc[k]);
#pragma unroll
        for (int k = 0; k < H; k++)
            if (e0 + (H + k) * nt < n) body(dd[k], vd[k]);
    }
}

/* Assembly leaves e_begin .. n-1: lane t sums leaves e_begin + t + k nt.
 * Leaf e's first values are group e of the contribution array, so the next
 * U leaves' records and first values load together, while the current U
 * leaves are summed.  The single-group leaves come first and run with
 * MULTI = false: nothing else is loaded, so each wait is for loads issued
 * one batch earlier.  With MULTI, the second group of each current leaf
 * that has one is requested before the next leaves' loads, so waiting for
 * it does not wait for them (loads complete in order). */
template <int U, bool MULTI, class F>
__device__ __forceinline__ void run_leaves(const int4 *leaf, const double4 *val, const uint4 *code,
                                           int e_begin, int n, int tid, int nt, F body)
{
    if (n <= e_begin) return;
    /* prefetch loads are unconditional (past-the-end indices clamped to the
     * last leaf): the compiler can then count exactly how many loads are
     * outstanding and wait only for the ones a leaf needs */
    Rec rb[U];
    double4 vb[U];
#pragma unroll
    for (int k = 0; k < U; k++) {
        const int e = min(e_begin + tid + k * nt, n - 1);
        rb[k] = load_rec(leaf, e);
        vb[k] = gload4(val + e);
    }
    for (int e0 = e_begin + tid; e0 < n; e0 += U * nt) {
        Rec rc[U];
        double4 vc[U];
        uint4 xw[U];
        double4 xv[U];
#pragma unroll
        for (int k = 0; k < U; k++) {
            rc[k] = rb[k];
            vc[k] = vb[k];
            if (MULTI && e0 + k * nt < n && rc[k].h.y < rc[k].h.z) {
                xw[k] = gload(code + rc[k].h.y);
                xv[k] = gload4(val + rc[k].h.y);
            }
        }
#pragma unroll
        for (int k = 0; k < U; k++) {
            const int e = min(e0 + (U + k) * nt, n - 1);
            rb[k] = load_rec(leaf, e);
            vb[k] = gload4(val + e);
        }
#pragma unroll
        for (int k = 0; k < U; k++)
            if (e0 + k * nt < n) body(rc[k], vc[k], xw[k], xv[k]);
    }
}

/* value of one matrix contribution code: mv, the value at the code's
 * position in the contribution array (the MOS stamp the BSIM4 kernel wrote
 * there, or the signed resistor conductance, source +-1 or padding 0 put
 * there when the batch was built), except a capacitor companion
 * conductance, which the step kernel computes each pass.  No global load
 * here: one in a branch would make every use wait for all outstanding
 * loads, the prefetched leaves' included. */
__device__ __forceinline__ double mat_value(uint32_t code, double mv, const double *capg, bool icz)
{
    if ((code & CT_ICZ) && icz) return 0.0;
    if ((code >> CT_TYPE_SHIFT) != CT_CAP) return mv;
    const double v = capg[code & CT_INDEX_MASK];
    return (code & CT_NEG) ? -v : v;
}

/* value of one RHS contribution code (MOS current = mv as above,
 * capacitor companion current, source value, padding 0) */
__device__ __forceinline__ double rhs_value(uint32_t code, double mv, const double *capq,
                                            const double *vsv)
{
    const uint32_t idx = code & CT_INDEX_MASK;
    double v;
    switch (code >> CT_TYPE_SHIFT) {
    case CT_MOS: v = mv; break;
    case CT_CAP: v = capq[idx]; break;
    case CT_VSRC: v = vsv[idx]; break;
    default: v = 0.0; break;
    }
    return (code & CT_NEG) ? -v : v;
}

/* one wavefront per circuit: the level-scheduled LU program has a few
 * dozen entries per level, and a one-wave workgroup makes each level's
 * barrier nearly free while many circuits share a compute unit */
constexpr int STEP_THREADS = 64;
static_assert(STEP_THREADS == 64, "the step kernel's reductions assume one wavefront");
/* assembly leaves whose loads are issued together, per lane */
constexpr int ASM_UNROLL = 4;
constexpr int PLAIN_UNROLL = 8;
/* unknowns (and device votes) a lane loads together in the convergence test */
constexpr int CONV_CHUNK = 8;

/* ---------------- learned initial guess ---------------- */

constexpr double GUESS_RATIO_MAX = 4.0;       /* step ratios are clamped to [0, 4] */
constexpr double GUESS_MIN_ACTIVITY = 1e-5;   /* V: quieter nodes keep the linear extrapolation */

__device__ __forceinline__ double clamp_ratio(double r)
{
    return r != r ? 0.0 : fmin(fmax(r, 0.0), GUESS_RATIO_MAX);
}

/* One feature row of the guess network for a node whose last four accepted
 * values v0..v3 sit at t0 > t1 > t2 > t3, for the step t0 -> t0 + h;
 * nbv[k] is neighbour k's value at t0 where bit k of nbmask is set.  Writes
 * the linear extrapolation and the activity; returns false (f untouched)
 * when the network cannot matter: the node is quieter than
 * GUESS_MIN_ACTIVITY, or the largest correction the network may make,
 * GUESS_CLAMP times the activity, stays inside the node's Newton tolerance
 * reltol |v0| + vntol (niconv.c:50-60), so the step converges as from the
 * extrapolation.  The training rows are built by the same formula on the
 * host (products and sums rounded separately, never fused). */
__device__ __forceinline__ bool guess_feature_row(double v0, double v1, double v2, double v3, double t0,
                                                  double t1, double t2, double t3, double h,
                                                  const double *nbv, unsigned nbmask, double reltol,
                                                  double vntol, float *f, double *lin_out, double *act_out)
{
    const double r0 = clamp_ratio(h / (t0 - t1));
    const double r1 = clamp_ratio((t0 - t1) / (t1 - t2));
    const double r2 = clamp_ratio((t1 - t2) / (t2 - t3));
    const double d0 = v0 - v1, d1 = v1 - v2, d2 = v2 - v3;
    const double act = fabs(d0) + fabs(d1) + fabs(d2);
    *lin_out = v0 + r0 * d0;
    *act_out = act;
    if (!(act >= GUESS_MIN_ACTIVITY)) return false;
    if (!(__dmul_rn(GUESS_CLAMP, act) > __dadd_rn(__dmul_rn(reltol, fabs(v0)), vntol))) return false;
    if (!f) return true;   /* the selection only */
    f[0] = (float)v0;
    f[1] = (float)(d0 / act);
    f[2] = (float)(d1 / act);
    f[3] = (float)(d2 / act);
    f[4] = (float)r0;
    f[5] = (float)r1;
    f[6] = (float)r2;
    for (int k = 0; k < 8; k++) f[7 + k] = (nbmask >> k) & 1u ? (float)(nbv[k] - v0) : 0.0f;
    f[15] = (float)(log10(act) / 5.0);
    return true;
}


/*
 * Features of node u of circuit c for the step the circuit is starting
 * (t_n -> t_n + h), from the last four accepted points v0..v3 at t0 > t1 >
 * t2 > t3, with the node's activity a = |v0 - v1| + |v1 - v2| + |v2 - v3|:
 *   v0, (v0 - v1) / a, (v1 - v2) / a, (v2 - v3) / a, h / (t0 - t1),
 *   (t0 - t1) / (t1 - t2), (t1 - t2) / (t2 - t3), the voltages of up to 8
 *   neighbour nodes minus v0 (0 where absent), log10(a) / 5.
 * The guess is v0 + r (v0 - v1) + a * clamp(net, +-4): the network predicts
 * the linear extrapolation's error in units of the node's own recent
 * motion.  Nodes the network cannot matter for (guess_feature_row) take the
 * extrapolation without a network row; branch currents start from their
 * last accepted value.
 *
 * Run by the circuit's step workgroup (one wavefront) right after it asks
 * for a guess, in two sweeps so that the dependent loads of a row (its
 * neighbours) happen once per circuit rather than once per 64 nodes: the
 * first sweep reads each node's four accepted values, gives the nodes the
 * network cannot matter for their extrapolation and lists the others in
 * LDS (list, at least N ints, free once the step is accepted); one atomic
 * reserves the circuit's rows at the end of its half's row buffer; the
 * second sweep builds the listed rows.  The network kernel at the start of
 * the next pass turns them into the nodes' starting values.
 */
__device__ __forceinline__ void guess_rows(const TranDev &D, const CktCtl *K, int c, int *list)
{
    const int tid = threadIdx.x, nt = blockDim.x;
    const int lane = (int)__lane_id();
    const int s0 = (K->hpos + 3) & 3, s1 = (K->hpos + 2) & 3, s2 = (K->hpos + 1) & 3, s3 = K->hpos;
    const double t0 = K->ht[s0], t1 = K->ht[s1], t2 = K->ht[s2], t3 = K->ht[s3], h = K->delta;
    const int64_t stride = (int64_t)D.C * (D.N + 1);
    const double *h0p = D.hist + ((int64_t)s0 * D.C + c) * (D.N + 1);
    const int64_t xoff = ((int64_t)K->xsel * D.C + c) * (D.N + 1);
    double *xo = D.x + xoff;
    const int hs = c >= D.gsplit;   /* the half whose row buffer and counter this circuit uses */
    const int64_t r0 = hs ? (int64_t)D.gsplit * D.N : 0;
    const double nbz[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int n_on = 0;   /* uniform: rows listed so far */
    for (int b0 = 1; b0 <= D.N; b0 += nt) {
        const int u = b0 + tid;
        bool on = false;
        if (u <= D.N) {
            if (u > D.NN) {
                xo[u] = h0p[u];   /* branch currents start from their last accepted value */
            } else {
                double lin, act;
                on = guess_feature_row(h0p[u], h0p[u + ((s1 - s0) * stride)], h0p[u + ((s2 - s0) * stride)],
                                       h0p[u + ((s3 - s0) * stride)], t0, t1, t2, t3, h, nbz, 0u, D.reltol,
                                       D.vntol, nullptr, &lin, &act);
                if (!on) xo[u] = lin;
            }
        }
        const unsigned long long m = __ballot(on);
        if (on) list[n_on + __popcll(m & ((1ull << lane) - 1ull))] = u;
        n_on += __popcll(m);
    }
    if (n_on == 0) return;
    int first = 0;
    if (tid == 0) first = atomicAdd(D.gcount + 2 * hs, n_on);
    first = __shfl(first, 0);
    __syncthreads();   /* the list is complete */
    for (int k = tid; k < n_on; k += nt) {
        const int u = list[k];
        double nbv[8];
        unsigned nbmask = 0;
        for (int q = 0; q < 8; q++) {
            const int nb = D.nbr[u * 8 + q];
            nbv[q] = nb >= 0 ? h0p[nb] : 0.0;
            nbmask |= nb >= 0 ? 1u << q : 0u;
        }
        float f[16];
        double lin, act;
        guess_feature_row(h0p[u], h0p[u + ((s1 - s0) * stride)], h0p[u + ((s2 - s0) * stride)],
                          h0p[u + ((s3 - s0) * stride)], t0, t1, t2, t3, h, nbv, nbmask, D.reltol, D.vntol, f,
                          &lin, &act);
        const int64_t row = r0 + first + k;
        float4 *fr = reinterpret_cast<float4 *>(D.gF + row * 16);
#pragma unroll
        for (int q = 0; q < 4; q++) fr[q] = make_float4(f[4 * q], f[4 * q + 1], f[4 * q + 2], f[4 * q + 3]);
        D.gdest[row] = xoff + u;
        D.gbase[row] = make_double2(lin, act);
    }
}

/* ---------------- KLU's numeric pivoting ---------------- */

constexpr int PV_OK = 0, PV_SINGULAR = 1, PV_BUSY = 2;
/* a circuit waits for a free dense workspace at most this many polls; each
 * holder finishes its factorization without waiting on anything, so the
 * bound is only reached if the pool were never released */
constexpr int PV_POLLS = 1 << 22;
/* the failure status when it is */
constexpr int PV_NO_WORKSPACE = 108;

/* the largest of one value per lane (ties to the lowest key) with its
 * index, and a sum, over the one-wavefront workgroup; fixed exchange order */
__device__ __forceinline__ void wave_argmax(double &v, int &key, int &i)
{
#pragma unroll
    for (int s = 32; s > 0; s >>= 1) {
        const double ov = __shfl_xor(v, s);
        const int ok = __shfl_xor(key, s), oi = __shfl_xor(i, s);
        if (ov > v || (ov == v && ok < key)) {
            v = ov;
            key = ok;
            i = oi;
        }
    }
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int s = 32; s > 0; s >>= 1) v += __shfl_xor(v, s);
    return v;
}

/* klu_factor / klu_refactor for one circuit whose static pivots do not do,
 * by one wavefront on a dense workspace (the permuted matrix with its
 * right-hand side, columns in the static order).
 * full: KLU's threshold partial pivoting with diagonal preference (lpivot,
 * klu_kernel.c:370-530): in column k the diagonal candidate stays the pivot
 * when its scaled magnitude is at least pivrel times the largest candidate's,
 * else the largest is taken (ties to the lowest original row; KLU takes the
 * first in its column pattern, which lists the column's own entries by
 * ascending row, so the two can differ on some exact ties) and the
 * displaced candidate becomes the diagonal of the
 * column the chosen row was the candidate of (klu_kernel.c:930-945).  The
 * candidates start as the static rows (Symbolic->P); magnitudes are scaled
 * by the row maxima (klu_scale.c:19-159 with scale = 2, a zero row counting
 * 1).  The rows chosen go to pv_row.
 * !full: the rows in pv_row are refactored as they are (klu_refactor); a
 * zero pivot there restarts as a full factorization in the same call (the
 * same-iteration klu_factor of niiter.c:160-195), and *refailed is set.
 * lu and z hold the assembled slots and right-hand side; the LDS from sm on
 * (at least 36 N + 12 bytes, lu and z included) is scratch.  The solution of
 * permuted column k goes to z[k], as the static solve leaves it.  Returns
 * PV_OK, PV_SINGULAR (a zero or non-finite pivot, E_SINGULAR) or PV_BUSY. */
__device__ __forceinline__ int pivot_solve(const TranDev &D, CktCtl *K, int c, const double *lu, double *z,
                                           double *sm, bool full, bool *refailed)
{
    __shared__ int s_ws;
    const int tid = threadIdx.x, N = D.N, ld = D.pv_ld;
    const uint64_t below = (1ull << tid) - 1ull;
    if (tid == 0) {
        int got = -1;
        for (int poll = 0; got < 0 && poll < PV_POLLS; poll++) {
            for (int k = 0; k < D.pv_slots && got < 0; k++) {
                const int w = (c + k) % D.pv_slots;
                if (atomicCAS(D.pv_owner + w, 0, c + 1) == 0) got = w;
            }
            if (got < 0) __builtin_amdgcn_s_sleep(16);
        }
        s_ws = got;
    }
    __syncthreads();
    const int ws = s_ws;
    if (ws < 0) return PV_BUSY;
    const int64_t nld = (int64_t)N * ld;
    double *A0 = D.pv_pool + (int64_t)ws * D.pv_stride;   /* the assembled matrix */
    double *W = A0 + nld;                                   /* its elimination */
    double *Rs = D.pv_rs + (int64_t)c * N;
    int *prow_g = D.pv_row + (int64_t)c * N;

    for (int64_t e = tid; e < nld; e += STEP_THREADS) A0[e] = 0.0;
    __syncthreads();
    for (int a = tid; a < N; a += STEP_THREADS) {
        const int s0 = gload(D.row_start + a), s1 = gload(D.row_start + a + 1);
        double *row = A0 + (int64_t)a * ld;
        double m = 0.0;
        for (int s = s0; s < s1; s++) {
            const double v = lu[s];
            row[gload(D.slot_col + s)] = v;
            m = fmax(m, fabs(v));
        }
        row[N] = z[a];
        Rs[a] = m == 0.0 ? 1.0 : m;
    }
    __syncthreads();
    /* scratch over lu, z and what follows: multipliers (later the solution),
     * the pivot row's values, then the bookkeeping */
    double *lval = sm, *uval = sm + N, *x = sm;
    int *Pinv = reinterpret_cast<int *>(sm + 2 * N + 1);   /* row -> its column once pivotal, else
                                                              -1 - (column it is the candidate of) */
    int *Pd = Pinv + N;                                      /* column -> diagonal candidate row */
    int *rl = Pd + N, *cl = rl + N, *prw = cl + N + 1;

    int status;
    for (;;) {
        for (int64_t e = tid; e < nld; e += STEP_THREADS) W[e] = A0[e];
        for (int a = tid; a < N; a += STEP_THREADS) {
            Pinv[a] = full ? -1 - a : -1;
            Pd[a] = a;
            prw[a] = full ? a : prow_g[a];
        }
        if (tid == 0) {
            if (full) K->pvfull++;
            else K->pvrefac++;
        }
        __syncthreads();
        status = PV_OK;
        for (int k = 0; k < N; k++) {
            int p;
            if (full) {
                /* the candidates' original rows break ties */
                double best = -1.0;
                int bkey = INT_MAX, bi = N;
                for (int a = tid; a < N; a += STEP_THREADS)
                    if (Pinv[a] < 0) {
                        const double v = fabs(W[(int64_t)a * ld + k]) / Rs[a];
                        const int key = gload(D.perm_unknown + N + a);
                        if (v > best || (v == best && key < bkey)) {
                            best = v;
                            bkey = key;
                            bi = a;
                        }
                    }
                wave_argmax(best, bkey, bi);
                const int d = Pd[k];
                p = bi < N ? bi : d;
                if (Pinv[d] < 0 && fabs(W[(int64_t)d * ld + k]) / Rs[d] >= D.pivrel * best) p = d;
            } else {
                p = prw[k];
            }
            const double piv = W[(int64_t)p * ld + k];
            if (piv == 0.0 || !isfinite(piv)) {
                status = PV_SINGULAR;
                break;
            }
            __syncthreads();
            if (tid == 0) {
                if (full) {
                    const int d = Pd[k];
                    if (p != d && Pinv[d] < 0) {
                        const int kbar = -1 - Pinv[p];
                        Pd[kbar] = d;
                        Pinv[d] = -1 - kbar;
                    }
                    prw[k] = p;
                }
                Pinv[p] = k;
            }
            __syncthreads();
            /* rows still to eliminate with a nonzero in column k, and their
             * multipliers; the pivot row's nonzeros right of column k */
            int nr = 0, nc = 0;
            for (int a0 = 0; a0 < N; a0 += STEP_THREADS) {
                const int a = a0 + tid;
                const double xv = (a < N && Pinv[a] < 0) ? W[(int64_t)a * ld + k] : 0.0;
                const bool on = xv != 0.0;
                const uint64_t m = __ballot(on);
                if (on) {
                    const int q = nr + __popcll(m & below);
                    rl[q] = a;
                    lval[q] = xv / piv;
                }
                nr += __popcll(m);
            }
            const double *pr = W + (int64_t)p * ld;
            for (int j0 = k + 1; j0 <= N; j0 += STEP_THREADS) {
                const int j = j0 + tid;
                const double u = j <= N ? pr[j] : 0.0;
                const bool on = u != 0.0;
                const uint64_t m = __ballot(on);
                if (on) {
                    const int q = nc + __popcll(m & below);
                    cl[q] = j;
                    uval[q] = u;
                }
                nc += __popcll(m);
            }
            __syncthreads();
            const int tot = nr * nc;
            for (int e = tid; e < tot; e += STEP_THREADS) {
                const int r = e / nc, q = e - r * nc;
                double *w = W + (int64_t)rl[r] * ld + cl[q];
                *w = fma(-lval[r], uval[q], *w);
            }
            __syncthreads();
        }
        if (status == PV_OK || full) break;
        full = true;
        *refailed = true;
        __syncthreads();
    }
    if (status == PV_OK) {
        __syncthreads();
        for (int k = N - 1; k >= 0; k--) {
            const double *row = W + (int64_t)prw[k] * ld;
            double acc = 0.0;
            for (int j = k + 1 + tid; j < N; j += STEP_THREADS) acc = fma(row[j], x[j], acc);
            acc = wave_sum(acc);
            if (tid == 0) x[k] = (row[N] - acc) / row[k];
            __syncthreads();
        }
        if (full)
            for (int k = tid; k < N; k += STEP_THREADS) prow_g[k] = prw[k];
        __syncthreads();
        /* x lies in lu's part of the scratch (nslots >= N), clear of z */
        for (int k = tid; k < N; k += STEP_THREADS) z[k] = x[k];
    }
    __threadfence();
    __syncthreads();
    if (tid == 0) atomicExch(D.pv_owner + ws, 0);
    return status;
}

/* circuit c of the topology D, one workgroup */
__device__ __forceinline__ void tran_step_body(const TranDev &D, int c)
{
    extern __shared__ double sm[];
    __shared__ int s_flag[4];
    __shared__ int s_act;
    __shared__ int s_guess;
    __shared__ int s_pvbad;

    const int tid = threadIdx.x, nt = blockDim.x;
    CktCtl *K = D.ctl + c;
    if (K->phase >= PH_DONE) return;
    const int N = D.N;
    double *lu = sm;                 /* nslots + 1: lu[nslots] = 0 pads term groups */
    double *z = lu + D.nslots + 1;   /* N + 1: z[N] = 0 */
    double *capg = z + N + 1;
    double *capq = capg + D.NCAP;
    double *vsv = capq + D.NCAP;

    const int mode = K->mode, rb = K->rb, xs = K->xsel;
    /* OPtran steps its own time; the sources stay at CKTtime, 0 */
    const double tnow = K->optran ? 0.0 : K->time;
    const int64_t cnm = (int64_t)D.C * D.NM, cnc = (int64_t)D.C * D.NCAP;
    double *xo = D.x + ((int64_t)xs * D.C + c) * (N + 1);
    double *xn = D.x + ((int64_t)(xs ^ 1) * D.C + c) * (N + 1);
    if (tid < 4) s_flag[tid] = 0;
    if (tid == 0) s_pvbad = 0;
    /* NIiter factors afresh (SMPreorder, klu_factor) when NISHOULDREORDER
     * is set, on INITJCT and on the first INITTRAN iteration
     * (niiter.c:107-146), else refactors on the current pivot rows */
    const bool reorder = K->pvre || (mode & M_INITJCT) || ((mode & M_INITTRAN) && K->iterno == 0);
    const bool uic_op = (mode & M_TRANOP) && (mode & M_UIC);
    const bool can_pivot = D.pv_slots > 0;
    if (tid == 0) {
        K->passes++;   /* this circuit's Newton passes (its BSIM4 evaluations: passes x NM) */
        if (K->guess == 1) K->guess = 2;   /* this iteration starts from the learned guess */
    }

    /* capload.c:19-86 */
    for (int k = tid; k < D.NCAP; k += nt) {
        const int p = D.cap_nodes[k], q = D.cap_nodes[D.NCAP + k];
        const double Cv = D.cap_c[(int64_t)c * D.NCAP + k];
        const int64_t ci = (int64_t)c * D.NCAP + k;
        double geq = 0.0, ceq = 0.0;
        if (mode & (M_TRAN | M_AC | M_TRANOP)) {
            const int cond1 = ((mode & MODEDC_MASK) && (mode & M_INITJCT)) ||
                              ((mode & M_UIC) && (mode & M_INITTRAN));
            /* CAPinitCond: 0, or under UIC the .ic difference (capgetic.c) */
            const double vcap = cond1 ? D.cap_ic[ci] : xo[p] - xo[q];
            double *q0 = D.cstate + ((int64_t)lvl(rb, 0) * 2) * cnc + ci;
            double *q1 = D.cstate + ((int64_t)lvl(rb, 1) * 2) * cnc + ci;
            double *cc0 = D.cstate + ((int64_t)lvl(rb, 0) * 2 + 1) * cnc + ci;
            double *cc1 = D.cstate + ((int64_t)lvl(rb, 1) * 2 + 1) * cnc + ci;
            if (mode & (M_TRAN | M_AC)) {
                if (mode & M_INITPRED) *q0 = *q1;
                else {
                    *q0 = Cv * vcap;
                    if (mode & M_INITTRAN) *q1 = *q0;
                }
                /* NIintegrate, niinteg.c:18-80 */
                double cc;
                if (K->method == 1) {
                    if (K->order == 1) cc = K->ag[0] * *q0 + K->ag[1] * *q1;
                    else cc = -*cc1 * K->ag[1] + K->ag[0] * (*q0 - *q1);
                } else {
                    cc = 0;
                    for (int l = K->order; l >= 2; l--)
                        cc += K->ag[l] * D.cstate[((int64_t)lvl(rb, l) * 2) * cnc + ci];
                    cc += K->ag[1] * *q1;
                    cc += K->ag[0] * *q0;
                }
                *cc0 = cc;
                ceq = cc - K->ag[0] * *q0;
                geq = K->ag[0] * Cv;
                if (mode & M_INITTRAN) *cc1 = *cc0;
            } else {
                *q0 = Cv * vcap;
            }
        }
        capg[k] = geq;
        capq[k] = ceq;
    }
    for (int k = tid; k < D.NVS; k += nt)
    {
        /* vsrcload.c:453-457: the operating point scales sources by
         * CKTsrcFact */
        const double v = vsrc_value(D.vs_par + ((int64_t)c * D.NVS + k) * VSP, mode, tnow, K->step);
        vsv[k] = (mode & M_TRANOP) ? v * K->srcFact : v;
    }
    __syncthreads();

    /* dense: this iteration factors on pivot rows other than the static
     * ones (pivot_solve); full: afresh, with KLU's pivot search.  pvhold:
     * in the last pass the static pivots failed KLU's test or met a zero;
     * that iteration goes on here, its loads kept (the BSIM4 kernel skipped
     * the circuit) and the matrix assembled again from them */
    const bool hold = K->pvhold != 0;
    const bool dense = can_pivot && (hold || (K->pv && !reorder)), full = hold;
    if (tid == 0) {
        lu[D.nslots] = 0.0;
        z[N] = 0.0;
    }

    /* assembly into the permuted L+U slots and RHS rows.  Each leaf sums at
     * most a few groups of four contributions in element order; entries fed
     * by many elements (a supply node's diagonal) are split into leaves
     * writing partial sums that one combine step adds in order, so no lane
     * walks hundreds of contributions alone.  (CKTload's summation order is
     * not kept; the order here is fixed, so results repeat bit for bit.) */
    const bool icz = (mode & M_TRANOP) && !(mode & M_UIC);
    double *part = vsv + D.NVS;
    /* the level tables, read at every level, staged in LDS (visible after
     * the barrier that ends the assembly) */
    int *levf = reinterpret_cast<int *>(part + D.n_part);
    int *levw = levf + D.nlev_f + 1;
    int *levb = levw + D.nlev_fw + 1;
    for (int k = tid; k <= D.nlev_f; k += nt) levf[k] = gload(D.f_lev + k);
    for (int k = tid; k <= D.nlev_fw; k += nt) levw[k] = gload(D.fw_lev + k);
    for (int k = tid; k <= D.nlev_bw; k += nt) levb[k] = gload(D.bw_lev + k);
    const uint4 *AC = reinterpret_cast<const uint4 *>(D.a_code);
    const uint4 *RC = reinterpret_cast<const uint4 *>(D.r_code);
    /* the circuit's contribution values, four per group: consecutive leaves
     * read consecutive groups, so a wavefront's loads are contiguous */
    const double4 *AV = reinterpret_cast<const double4 *>(D.contrib + (int64_t)c * D.NA);
    const double4 *RV = reinterpret_cast<const double4 *>(D.rcontrib + (int64_t)c * D.NR);
    /* the sum of one leaf: its first group (codes from the record, values
     * already loaded), then its further groups in order */
    auto a_leaf_sum = [&](const Rec &lf, const double4 &m0, const uint4 &w1, const double4 &m1) {
        double acc = 0.0;
        acc += mat_value(lf.g.x, m0.x, capg, icz);
        acc += mat_value(lf.g.y, m0.y, capg, icz);
        acc += mat_value(lf.g.z, m0.z, capg, icz);
        acc += mat_value(lf.g.w, m0.w, capg, icz);
        if (lf.h.y < lf.h.z) {
            acc += mat_value(w1.x, m1.x, capg, icz);
            acc += mat_value(w1.y, m1.y, capg, icz);
            acc += mat_value(w1.z, m1.z, capg, icz);
            acc += mat_value(w1.w, m1.w, capg, icz);
        }
        for (int q = lf.h.y + 1; q < lf.h.z; q++) {
            const uint4 w = gload(AC + q);
            const double4 m = gload4(AV + q);
            acc += mat_value(w.x, m.x, capg, icz);
            acc += mat_value(w.y, m.y, capg, icz);
            acc += mat_value(w.z, m.z, capg, icz);
            acc += mat_value(w.w, m.w, capg, icz);
        }
        if (lf.h.x < D.nslots) lu[lf.h.x] = ((lf.h.w & 1) && icz) ? 1.0 : acc;
        else part[lf.h.x - D.nslots] = acc;
    };
    /* single-group leaves: no further groups, and no code that could load
     * them (a load anywhere in the body makes the compiler wait for all
     * outstanding loads after it, the prefetched leaves' included) */
    run_plain<PLAIN_UNROLL>(D.a_dest, AV, D.n_aplain, tid, nt, [&](int d, double4 m) {
        double acc = 0.0;
        acc += m.x;
        acc += m.y;
        acc += m.z;
        acc += m.w;
        if (d < D.nslots) lu[d] = acc;
        else part[d - D.nslots] = acc;
    });
    run_leaves<ASM_UNROLL, false>(D.a_leaf, AV, AC, D.n_aplain, D.n_aleaf1, tid, nt,
                                  [&](const Rec &lf, const double4 &m0, const uint4 &, const double4 &) {
        double acc = 0.0;
        acc += mat_value(lf.g.x, m0.x, capg, icz);
        acc += mat_value(lf.g.y, m0.y, capg, icz);
        acc += mat_value(lf.g.z, m0.z, capg, icz);
        acc += mat_value(lf.g.w, m0.w, capg, icz);
        if (lf.h.x < D.nslots) lu[lf.h.x] = ((lf.h.w & 1) && icz) ? 1.0 : acc;
        else part[lf.h.x - D.nslots] = acc;
    });
    run_leaves<ASM_UNROLL, true>(D.a_leaf, AV, AC, D.n_aleaf1, D.n_aleaf, tid, nt, a_leaf_sum);
    auto r_leaf_sum = [&](const Rec &lf, const double4 &m0, const uint4 &w1, const double4 &m1) {
        double acc = 0.0;
        acc += rhs_value(lf.g.x, m0.x, capq, vsv);
        acc += rhs_value(lf.g.y, m0.y, capq, vsv);
        acc += rhs_value(lf.g.z, m0.z, capq, vsv);
        acc += rhs_value(lf.g.w, m0.w, capq, vsv);
        if (lf.h.y < lf.h.z) {
            acc += rhs_value(w1.x, m1.x, capq, vsv);
            acc += rhs_value(w1.y, m1.y, capq, vsv);
            acc += rhs_value(w1.z, m1.z, capq, vsv);
            acc += rhs_value(w1.w, m1.w, capq, vsv);
        }
        for (int q = lf.h.y + 1; q < lf.h.z; q++) {
            const uint4 w = gload(RC + q);
            const double4 m = gload4(RV + q);
            acc += rhs_value(w.x, m.x, capq, vsv);
            acc += rhs_value(w.y, m.y, capq, vsv);
            acc += rhs_value(w.z, m.z, capq, vsv);
            acc += rhs_value(w.w, m.w, capq, vsv);
        }
        if (lf.h.x < N) z[lf.h.x] = acc;
        else part[D.n_apart + lf.h.x - N] = acc;
    };
    run_plain<PLAIN_UNROLL>(D.r_dest, RV, D.n_rplain, tid, nt, [&](int d, double4 m) {
        double acc = 0.0;
        acc += m.x;
        acc += m.y;
        acc += m.z;
        acc += m.w;
        if (d < N) z[d] = acc;
        else part[D.n_apart + d - N] = acc;
    });
    run_leaves<ASM_UNROLL, false>(D.r_leaf, RV, RC, D.n_rplain, D.n_rleaf1, tid, nt,
                                  [&](const Rec &lf, const double4 &m0, const uint4 &, const double4 &) {
        double acc = 0.0;
        acc += rhs_value(lf.g.x, m0.x, capq, vsv);
        acc += rhs_value(lf.g.y, m0.y, capq, vsv);
        acc += rhs_value(lf.g.z, m0.z, capq, vsv);
        acc += rhs_value(lf.g.w, m0.w, capq, vsv);
        if (lf.h.x < N) z[lf.h.x] = acc;
        else part[D.n_apart + lf.h.x - N] = acc;
    });
    run_leaves<ASM_UNROLL, true>(D.r_leaf, RV, RC, D.n_rleaf1, D.n_rleaf, tid, nt, r_leaf_sum);
    __syncthreads();
    for (int e = tid; e < D.n_acomb; e += nt) {
        const int4 cb = gload(D.a_comb + e);
        double acc = 0.0;
        for (int p = cb.y; p < cb.z; p++) acc += part[p];
        if ((cb.w & 1) && icz) acc = 1.0;
        lu[cb.x] = acc;
    }
    for (int e = tid; e < D.n_rcomb; e += nt) {
        const int4 cb = gload(D.r_comb + e);
        double acc = 0.0;
        for (int p = cb.y; p < cb.z; p++) acc += part[D.n_apart + p];
        z[cb.x] = acc;
    }
    if (icz && D.n_icrow) {
        __syncthreads();
        for (int e = tid; e < D.n_icrow; e += nt) {
            const int a = gload(D.ic_row + e);            /* permuted row of an .ic node */
            const int u = gload(D.perm_unknown + N + a);
            z[a] = gload(D.ic_val + (int64_t)c * (D.NN + 1) + u) * K->srcFact;   /* cktload.c:160 */
        }
    }
    const double dgmin = K->diagGmin;
    if (dgmin != 0.0) {
        /* LoadGmin_CSC (klusmp.c:1765-1777): CKTdiagGmin onto every diagonal
         * element the matrix has, after the load */
        __syncthreads();
        for (int u = tid + 1; u <= N; u += nt) {
            const int sl = gload(D.diag_slot + u);
            if (sl >= 0) lu[sl] += dgmin;
        }
    }
    __syncthreads();
    /* NIiter under UIC in the operating point (niiter.c:41-46): the load
     * above and nothing else, no solve, no iteration counted */
    if (!uic_op) {
        bool refailed = false;
        if (dense) {
            const int r = pivot_solve(D, K, c, lu, z, sm, full, &refailed);
            if (r == PV_BUSY) {
                if (tid == 0) fail(D, K, PV_NO_WORKSPACE);
                return;
            }
            if (r == PV_SINGULAR && tid == 0) s_flag[0] = 1;
            __syncthreads();
        } else {
            if (reorder && can_pivot) {
                /* KLU's row scale factors of the matrix being factored (klu_scale.c:
                 * 19-159, scale = 2): the largest magnitude in each row, 1 for a
                 * zero row */
                double *Rs = D.pv_rs + (int64_t)c * N;
                for (int a = tid; a < N; a += nt) {
                    const int s0 = gload(D.row_start + a), s1 = gload(D.row_start + a + 1);
                    double m = 0.0;
                    for (int s = s0; s < s1; s++) m = fmax(m, fabs(lu[s]));
                    Rs[a] = m == 0.0 ? 1.0 : m;
                }
            }

            /* LU factorisation, entry-wise Doolittle by dependency levels, then the
             * two substitutions; see run_levels for the load schedule */
            const uint4 *FT = reinterpret_cast<const uint4 *>(D.f_term);
            run_levels(levf, reinterpret_cast<const int4 *>(D.f_ent), D.nlev_f, tid, nt,
                       [&](const Prog &r) {
                           const int t = r.h.x & 0xffff, kind = (r.h.x >> 20) & 3;
                           if (kind == PROG_PARTIAL) {
                               part[t] = lu_row_terms(0.0, r, FT, lu, lu);
                               return;
                           }
                           const double acc = (kind == PROG_COMBINE) ? add_parts(lu[t], r, part)
                                                                     : lu_row_terms(lu[t], r, FT, lu, lu);
                           lu[t] = (r.h.y >= 0) ? acc / lu[r.h.y] : acc;
                       });
            /* forward substitution (unit L) */
            const uint4 *WT = reinterpret_cast<const uint4 *>(D.fw_term);
            run_levels(levw, reinterpret_cast<const int4 *>(D.fw_ent), D.nlev_fw, tid, nt,
                       [&](const Prog &r) {
                           const int a = r.h.x & 0xffff, kind = (r.h.x >> 20) & 3;
                           if (kind == PROG_PARTIAL) part[a] = lu_row_terms(0.0, r, WT, lu, z);
                           else if (kind == PROG_COMBINE) z[a] = add_parts(z[a], r, part);
                           else z[a] = lu_row_terms(z[a], r, WT, lu, z);
                       });
            /* back substitution */
            const uint4 *BT = reinterpret_cast<const uint4 *>(D.bw_term);
            run_levels(levb, reinterpret_cast<const int4 *>(D.bw_ent), D.nlev_bw, tid, nt,
                       [&](const Prog &r) {
                           const int a = r.h.x & 0xffff, kind = (r.h.x >> 20) & 3;
                           if (kind == PROG_PARTIAL) {
                               part[a] = lu_row_terms(0.0, r, BT, lu, z);
                               return;
                           }
                           const double acc = (kind == PROG_COMBINE) ? add_parts(z[a], r, part)
                                                                     : lu_row_terms(z[a], r, BT, lu, z);
                           if (r.h.y < 0) {   /* a forward row of the block-by-block solve */
                               z[a] = acc;
                               return;
                           }
                           const double piv = lu[r.h.y];
                           if (piv == 0.0 || !isfinite(piv)) s_flag[0] = 1;   /* E_SINGULAR */
                           z[a] = acc / piv;
                       });
            if (can_pivot && reorder) {
                /* would klu_factor keep the static rows?  Column k's diagonal stays
                 * the pivot when its scaled magnitude is at least pivrel times every
                 * other candidate's: with x(a, k) = L(a, k) U(k, k), when
                 * pivrel |L(a, k)| Rs[k] <= Rs[a] for each row a below it */
                __syncthreads();
                const double *Rs = D.pv_rs + (int64_t)c * N;
                int bad = 0;
                for (int a = tid; a < N; a += nt) {
                    const int s0 = gload(D.row_start + a), s1 = gload(D.row_start + a + 1);
                    const double ra = Rs[a];
                    for (int s = s0; s < s1; s++) {
                        const int k = gload(D.slot_col + s);
                        if (k >= a) break;
                        if (D.pivrel * fabs(lu[s]) * Rs[k] > ra) bad = 1;
                    }
                }
                if (bad) s_pvbad = 1;
            }
            if (can_pivot) {
                __syncthreads();
                if (s_flag[0] || s_pvbad) {
                    /* klu_factor would pivot elsewhere, or the refactor met a zero
                     * pivot: the iteration continues next pass with a dense
                     * factorization */
                    if (tid == 0) K->pvhold = 1;
                    return;
                }
            }
        }
        if (tid == 0 && can_pivot && !s_flag[0]) {
            /* a successful SMPreorder clears NISHOULDREORDER; a refactor that
             * met a zero pivot and was factored afresh in the same iteration
             * leaves it set (niiter.c:113-195) */
            if (reorder) K->pvre = 0;
            else if (full || refailed) K->pvre = 1;
            K->pv = dense ? 1 : 0;
        }
        if (tid == 0) K->pvhold = 0;
        /* new iterate; NIconvTest (niconv.c:21-90) and the device votes */
        /* unknown u is z[unk_pos[u]]; the loads of a lane's unknowns (and of
         * its device votes) are independent, so each chunk issues them all
         * before it waits */
        if (tid == 0) xn[0] = 0.0;
        int conv = 0;
        for (int u0 = tid + 1; u0 <= N; u0 += CONV_CHUNK * nt) {
            int pos[CONV_CHUNK];
            double old[CONV_CHUNK];
    #pragma unroll
            for (int k = 0; k < CONV_CHUNK; k++) {
                const int u = u0 + k * nt;
                if (u <= N) {
                    pos[k] = gload(D.unk_pos + u);
                    old[k] = xo[u];
                }
            }
    #pragma unroll
            for (int k = 0; k < CONV_CHUNK; k++) {
                const int u = u0 + k * nt;
                if (u > N) continue;
                const double nw = z[pos[k]];
                xn[u] = nw;
                if (isnan(nw)) { conv = 1; continue; }
                const double tol = D.reltol * fmax(fabs(old[k]), fabs(nw)) + (u <= D.NN ? D.vntol : D.abstol);
                if (fabs(nw - old[k]) > tol) conv = 1;
            }
        }
        int vote = 0;
        const int *nc = D.noncon + (int64_t)c * D.NM;
        for (int j0 = tid; j0 < D.NM; j0 += CONV_CHUNK * nt) {
            int v[CONV_CHUNK];
    #pragma unroll
            for (int k = 0; k < CONV_CHUNK; k++) v[k] = (j0 + k * nt < D.NM) ? gload(nc + j0 + k * nt) : 0;
    #pragma unroll
            for (int k = 0; k < CONV_CHUNK; k++) vote |= v[k];
        }
        if (vote) s_flag[1] = 1;
        if (conv) s_flag[2] = 1;
        __syncthreads();
    }   /* !uic_op */

    /* NIiter bookkeeping (niiter.c:95-365); thread 0 owns the scalars */
    if (tid == 0) {
        int m = K->mode;
        int rc = -1;   /* -1: keep iterating */
        const int iterno = K->iterno + (uic_op ? 0 : 1);
        K->iterno = iterno;
        if (!uic_op) K->iters++;
        if (uic_op) {
            rc = 0;
        } else if (s_flag[0]) {
            rc = 102;
            K->pvre = 1;   /* E_SINGULAR returns with NISHOULDREORDER set */
            K->pv = 0;
        } else if (iterno > 100) rc = 103;
        else {
            const int noncon = (s_flag[1] == 0 && iterno != 1) ? s_flag[2] : 1;
            if (m & M_INITFLOAT) {
                if (noncon == 0) rc = 0;
            } else if (m & M_INITJCT) {
                m = (m & ~INITF_MASK) | M_INITFIX;
                K->pvre = 1;   /* niiter.c:335-338 */
            } else if (m & M_INITFIX) {
                if (noncon == 0) m = (m & ~INITF_MASK) | M_INITFLOAT;
            } else {
                if ((m & M_INITTRAN) && iterno <= 1) K->pvre = 1;   /* niiter.c:343-347 */
                m = (m & ~INITF_MASK) | M_INITFLOAT;   /* INITSMSIG, INITTRAN, INITPRED */
            }
        }
        K->mode = m;
        if (rc == -1) K->xsel ^= 1;   /* SWAP(rhs, rhsOld) */
        s_act = ACT_NONE;
        if (rc >= 0) {
            int opa = OPA_OK;
            if (K->phase == PH_OP) opa = op_next(D, K, rc);
            if (K->phase == PH_OP && opa != OPA_OK) {
                if (opa == OPA_FAIL && D.opstep > 0.0) {
                    /* cktop.c:93-105: OPtran, last (optran.c:297-420) */
                    optran_start(D, K);
                    s_act = 10;   /* copy state0 -> state1, then accept */
                } else if (opa == OPA_FAIL) {
                    fail(D, K, K->status);
                    s_act = ACT_STOP;
                } else {
                    s_act = ACT_OPMEM + opa;   /* a stepping stage continues */
                }
            } else if (K->phase == PH_OP) {
                {
                    /* dctran.c:268-330: into the transient */
                    K->phase = PH_TRAN;
                    K->order = 1;
                    for (int i = 0; i < 7; i++) K->deltaOld[i] = K->maxStep;
                    K->saveDelta = K->finalTime / 50;
                    K->mode = (K->mode & M_UIC) | M_TRAN | M_INITTRAN;
                    K->ag[0] = K->ag[1] = 0;
                    s_act = 10;   /* copy state0 -> state1, then accept */
                }
            } else if (rc != 0 && K->guess == 2) {
                /* the Newton run from the learned guess failed: redo the same
                 * step from the plain predictor */
                K->guess = 0;
                K->fallbacks++;
                guess_outcome(K, true);
                K->mode = (K->mode & M_UIC) | M_TRAN | M_INITPRED;
                K->iterno = 0;
                s_act = ACT_STOP;
            } else {
                if (K->guess == 2) {
                    K->guess = 0;
                    K->guessed++;
                    guess_outcome(K, rc != 0 || iterno > 2);
                }
                K->mode = (K->mode & M_UIC) | M_TRAN | M_INITPRED;
                s_act = K->firsttime ? 20 : 21;   /* 20: copy state1 -> state2, state3 first */
                s_flag[3] = rc;
            }
        }
    }
    __syncthreads();
    int act = s_act;
    if (act == ACT_NONE || act == ACT_STOP) return;
    if (act >= ACT_OPMEM) {
        op_memory(D, c, K, act - ACT_OPMEM);
        return;
    }
    if (act == 10) {
        copy_level(D, c, K->rb, 0, 1);
        __syncthreads();
        act = ACT_ACCEPT;
    } else {
        if (act == 20) {
            copy_level(D, c, K->rb, 1, 2);
            copy_level(D, c, K->rb, 1, 3);
            __syncthreads();
        }
        if (tid == 0) {
            const int rc = s_flag[3];
            if (rc != 0) {
                K->time = K->time - K->delta;
                if (!K->optran) K->rejected++;
                K->delta = K->delta / 8;
                if (K->firsttime) K->mode = (K->mode & M_UIC) | M_TRAN | M_INITTRAN;
                K->order = 1;
                s_act = ACT_RETRY;
            } else if (K->firsttime) {
                K->firsttime = 0;
                s_act = ACT_ACCEPT;
            } else {
                s_act = ACT_LTE1;
            }
        }
        __syncthreads();
        act = s_act;
        if (act == ACT_LTE1) {
            const double t1 = circuit_terr(D, K, c);
            if (tid == 0) {
                const double nd = fmin(2 * K->delta, t1);

                if (nd > .9 * K->delta) {
                    s_act = (K->order == 1 && K->maxOrder > 1) ? 30 : ACT_ACCEPT;
                    if (s_act == 30) K->order = 2;
                    else K->delta = nd;
                } else {
                    K->time = K->time - K->delta;
                    if (!K->optran) K->rejected++;
                    K->delta = nd;
                    s_act = ACT_RETRY;
                }
            }
            __syncthreads();
            act = s_act;
            if (act == 30) {
                const double t2 = circuit_terr(D, K, c);
                if (tid == 0) {
                    const double nd = fmin(2 * K->delta, t2);
                    if (nd <= 1.05 * K->delta) K->order = 1;
                    K->delta = nd;
                }
                act = ACT_ACCEPT;
            }
        }
    }
    if (act == ACT_ACCEPT && K->optran) {
        /* OPtran's nextTime (optran.c:430-490): no output.  Its CKTaccept
         * runs VSRCaccept at CKTtime 0 into the transient's breakpoints,
         * which leaves them as the transient's own first accept at time 0
         * does; that accept runs below, so it is not repeated here. */
        if (tid == 0) {
            s_act = ACT_NONE;
            if (K->time > K->breaks[0]) clr_break(K);
            if (almost_equal_ulps(K->time, K->finalTime, 100)) {
                optran_end(K);
                s_act = 10;
            } else {
                resume_block(K);
                if (loop_top(K)) fail(D, K, 104);
            }
        }
        __syncthreads();
        if (s_act != 10) return;
        /* the operating point is found: into the transient, as after CKTop */
        copy_level(D, c, K->rb, 0, 1);
        __syncthreads();
    }
    if (act == ACT_ACCEPT) {
        /* nextTime: (dctran.c:352-455) */
        if (tid == 0) {
            s_act = ACT_ACCEPT;
            if (vsrc_accept(D, c, K)) {
                fail(D, K, 1);
                s_act = ACT_STOP;
            } else {
                if (K->time > K->breaks[0]) clr_break(K);
                K->accepted++;
            }
        }
        __syncthreads();
        if (s_act == ACT_STOP) return;
        /* CKTdump: rhsOld of the accepted point; under UIC not at time 0
         * (dctran.c:416-419) */
        const double *xold = D.x + ((int64_t)K->xsel * D.C + c) * (N + 1);
        const bool dump = !((K->mode & M_UIC) && K->time <= 0.0);
        const int64_t np = dump ? K->npts : D.cap_pts;
        if (np < D.cap_pts) {
            double *row = D.out_x + ((int64_t)c * D.cap_pts + np) * D.nprobe;
            for (int k = tid; k < D.nprobe; k += nt) row[k] = xold[D.probe[k]];
            if (tid == 0) D.out_t[(int64_t)c * D.cap_pts + np] = K->time;
        }
        if (K->learned) {
            double *hrow = D.hist + ((int64_t)K->hpos * D.C + c) * (N + 1);
            for (int u = tid; u <= N; u += nt) hrow[u] = xold[u];
        }
        __syncthreads();
        if (tid == 0) {
            if (dump) K->npts = np + 1;
            if (K->learned) {
                K->ht[K->hpos] = K->time;
                K->hpos = (K->hpos + 1) & 3;
                K->nhist++;
            }
            s_guess = 0;
            if (almost_equal_ulps(K->time, K->finalTime, 100)) {
                K->phase = PH_DONE;
                atomicAdd(D.done_count, 1);
            } else {
                resume_block(K);
                if (loop_top(K)) fail(D, K, 104);
                else s_guess = request_guess(K);
            }
        }
    } else if (act == ACT_RETRY) {
        if (tid == 0) {
            s_guess = 0;
            if (K->delta <= K->delmin && K->olddelta <= K->delmin) {
                fail(D, K, 106);
            } else {
                if (K->delta <= K->delmin) K->delta = K->delmin;
                if (loop_top(K)) fail(D, K, 104);
                else s_guess = request_guess(K);
            }
        }
    } else {
        return;
    }
    /* the next step starts from the learned guess: its feature rows (the
     * LU workspace in LDS is free by now) */
    __syncthreads();
    if (s_guess) guess_rows(D, K, c, reinterpret_cast<int *>(sm));
}

/* one topology, one workgroup per circuit, circuits c0.. */
__global__ void __launch_bounds__(STEP_THREADS)
k_tran_step(TranDev D, int c0)
{
    tran_step_body(D, c0 + blockIdx.x);
}

/* several topologies: blk[b] = (group, circuit) */
__global__ void __launch_bounds__(STEP_THREADS)
k_tran_step_groups(const TranDev *__restrict__ Ds, const int2 *__restrict__ blk)
{
    const int2 bg = blk[blockIdx.x];
    tran_step_body(Ds[bg.x], bg.y);
}

/* the same rows for caller-supplied windows: row i from t[4i..4i+3],
 * v[4i..4i+3], h[i], nbv[8i..8i+7], nbmask[i]; quiet rows get zeros and a
 * negative activity in act[i] */
__global__ void __launch_bounds__(256)
k_guess_feature_rows(int M, const double *t, const double *h, const double *v, const double *nbv,
                     const unsigned *nbmask, double reltol, double vntol, float *F, double *act)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
    float f[16];
    double lin, a;
    const bool on = guess_feature_row(v[4 * i], v[4 * i + 1], v[4 * i + 2], v[4 * i + 3], t[4 * i],
                                      t[4 * i + 1], t[4 * i + 2], t[4 * i + 3], h[i], nbv + 8 * i, nbmask[i],
                                      reltol, vntol, f, &lin, &a);
    for (int k = 0; k < 16; k++) F[(int64_t)i * 16 + k] = on ? f[k] : 0.0f;
    act[i] = on ? a : -a;
}

/* the network's output y[row] becomes the starting value of the row's node
 * (rows produced by k_guess_features; grid-stride) */
__global__ void __launch_bounds__(256)
k_guess_apply(double *x, const float *y, const int64_t *dest, const double2 *base, const int *count,
              unsigned long long *rows_total)
{
    const int n = *count;
    if (rows_total && blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(rows_total, (unsigned long long)n);
    for (int row = blockIdx.x * blockDim.x + threadIdx.x; row < n; row += gridDim.x * blockDim.x)
        x[dest[row]] = guess_value(base[row], y[row]);
}

int launch_guess_apply(const TranDev &D, const float *y, const int64_t *dest, const double2 *base,
                       const int *count, int maxrows, unsigned long long *rows_total, hipStream_t s)
{
    const int blocks = std::max(1, std::min((maxrows + 255) / 256, 1024));
    hipLaunchKernelGGL(k_guess_apply, dim3((unsigned)blocks), dim3(256), 0, s, D.x, y, dest, base, count,
                       rows_total);
    NGMI_HIP(hipGetLastError());
    return NGMI_OK;
}

extern "C" int ngmi_guess_features(int M, const double *t, const double *h, const double *v, const double *nbv,
                                   const unsigned *nbmask, double reltol, double vntol, float *F, double *act,
                                   void *stream)
{
    if (M < 0 || (M > 0 && (!t || !h || !v || !nbv || !nbmask || !F || !act))) {
        set_error("guess_features: bad arguments");
        return NGMI_E_PARMVAL;
    }
    if (M == 0) return NGMI_OK;
    hipLaunchKernelGGL(k_guess_feature_rows, dim3((unsigned)((M + 255) / 256)), dim3(256), 0,
                       (hipStream_t)stream, M, t, h, v, nbv, nbmask, reltol, vntol, F, act);
    NGMI_HIP(hipGetLastError());
    return NGMI_OK;
}

__global__ void k_zero_int(int *p) { *p = 0; }

int launch_zero_int(int *p, hipStream_t s)
{
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(1), 0, s, p);
    NGMI_HIP(hipGetLastError());
    return NGMI_OK;
}

int set_tran_lds_limit(size_t bytes)
{
    NGMI_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_tran_step),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    NGMI_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_tran_step_groups),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    return NGMI_OK;
}

/* one Newton pass over every circuit of every topology group; with one
 * group (tab0, D0 set) the descriptor goes by value and the pass covers
 * circuits [c_begin, c_end).  ev (optional) = 3
 * events recorded before the MOSFET kernel, between the two kernels and
 * after the step kernel. */
int launch_tran_pass(const B4Tab *tab0, const TranDev *D0, int c_begin, int c_end, const B4Tab *tabs,
                     const TranDev *Ds, const int2 *mos_blk, int n_mos_blk, const int2 *step_blk,
                     int n_step_blk, size_t lds_bytes, hipStream_t s, hipEvent_t *ev)
{
    /* mos_waves 0 (the default): one wave on a single 101-stage ring, 49.7
     * against 53.2 us per launch with two */
    const int waves = waves_for(TUNE_MOS_WAVES, D0 ? (int64_t)(c_end - c_begin) * D0->NM : (int64_t)n_mos_blk * 256);
    if (ev) NGMI_HIP(hipEventRecord(ev[0], s));
    if (n_mos_blk > 0) {
        const dim3 g((unsigned)n_mos_blk);
        if (D0) {
            const int64_t i0 = (int64_t)c_begin * D0->NM, i1 = (int64_t)c_end * D0->NM;
            const dim3 gh((unsigned)((i1 - i0 + 255) / 256));
            /* the default card's selectors compiled as constants when the
             * whole set shares them */
            const bool dflt = tab0->flags == B4FL_DEFAULT && tune_get(TUNE_SPECIALISE);
            if (i1 > i0) {
#define NGMI_MOS(W, F) hipLaunchKernelGGL((k_tran_mos<W, F>), gh, dim3(256), 0, s, *tab0, *D0, i0, i1)
                switch (waves) {
                case 1: if (dflt) NGMI_MOS(1, B4FL_DEFAULT); else NGMI_MOS(1, -1); break;
                case 3: if (dflt) NGMI_MOS(3, B4FL_DEFAULT); else NGMI_MOS(3, -1); break;
                case 4: if (dflt) NGMI_MOS(4, B4FL_DEFAULT); else NGMI_MOS(4, -1); break;
                default: if (dflt) NGMI_MOS(2, B4FL_DEFAULT); else NGMI_MOS(2, -1); break;
                }
#undef NGMI_MOS
            }
        } else {
            if (waves == 1) hipLaunchKernelGGL(k_tran_mos_groups<1>, g, dim3(256), 0, s, tabs, Ds, mos_blk);
            else hipLaunchKernelGGL(k_tran_mos_groups<2>, g, dim3(256), 0, s, tabs, Ds, mos_blk);
        }
        NGMI_HIP(hipGetLastError());
    }
    if (ev) NGMI_HIP(hipEventRecord(ev[1], s));
    if (D0)
        hipLaunchKernelGGL(k_tran_step, dim3((unsigned)(c_end - c_begin)), dim3(STEP_THREADS),
                           lds_bytes, s, *D0, c_begin);
    else
        hipLaunchKernelGGL(k_tran_step_groups, dim3((unsigned)n_step_blk), dim3(STEP_THREADS),
                           lds_bytes, s, Ds, step_blk);
    NGMI_HIP(hipGetLastError());
    if (ev) NGMI_HIP(hipEventRecord(ev[2], s));
    return NGMI_OK;
}

}  // namespace ngmi
