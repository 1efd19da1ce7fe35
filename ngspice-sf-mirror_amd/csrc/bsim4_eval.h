/*
 * BSIM4 4.8.3 instance evaluation for the engine: I-V, C-V, limiting,
 * convergence check, charge integration and the 16-entry Jacobian / 4-entry
 * RHS stamp of one MOSFET whose terminals are circuit nodes (no internal
 * nodes: rgateMod 0, rbodyMod 0, no S/D series resistance, no NQS).
 *
 * The sequence of floating-point operations follows
 * src/spicelib/devices/bsim4/b4ld.c:69-5433 region by region (line ranges
 * cited inline) so that per-instance results agree with the reference to the
 * last few ulps; the structure (accessors, stamp arrays, no device struct) is
 * this engine's.  The same source compiles as HIP device code and, for the
 * host-side self check of the library, as plain C++.
 *
 * Accessor contract (templates):
 *   P  prm(f)            -> double   instance table field f (bsim4_fields.h)
 *   S  s(lvl, k) / set(lvl, k, v)    state k at history level lvl (0 = current)
 *   C  c(k) / setc(k, v)             per-instance cache (NgmiB4Cache)
 */
#ifndef NGMI_BSIM4_EVAL_H
#define NGMI_BSIM4_EVAL_H

#include "bsim4_fields.h"

/* Every multiply and add is rounded on its own, as in the reference build:
 * contracting a*b + c*d into a fused multiply-add changes results whenever
 * the terms cancel (ag0 q0 + ag1 q1 right after INITTRAN is exactly 0 in
 * the reference and a rounding residual with FMA).  The remaining difference
 * to the reference is the last bit of exp/log/sqrt/pow. */
#pragma clang fp contract(off)

#ifdef __HIPCC__
#define NGMI_HD __host__ __device__ __forceinline__
#else
#include <cmath>
#define NGMI_HD inline
using std::exp; using std::log; using std::sqrt; using std::fabs;
#endif

namespace ngmi {

/* mode bits, same values as the reference's CKTmode (cktdefs.h:171-199) */
enum : int {
    M_TRAN = 0x1, M_AC = 0x2, M_DCOP = 0x10, M_TRANOP = 0x20, M_DCTRANCURVE = 0x40,
    M_INITFLOAT = 0x100, M_INITJCT = 0x200, M_INITFIX = 0x400, M_INITSMSIG = 0x800,
    M_INITTRAN = 0x1000, M_INITPRED = 0x2000, M_UIC = 0x10000,
    /* engine-internal, no reference counterpart: the first Newton iteration
     * of a step that starts from the learned guess.  The terminal voltages
     * come from the iterate as in INITFLOAT, and state0 first takes state1's
     * voltages as in INITPRED, so limiting is against the last accepted point */
    M_GUESSED = 0x40000000
};

struct B4Ckt {
    int mode;
    int order;          /* integration order */
    int method;         /* 1 trapezoidal, 2 Gear */
    double ag[7];       /* integration coefficients (NIcomCof) */
    double xfact;       /* delta / deltaOld[1], predictor extrapolation factor */
    double gmin, reltol, abstol;
};

namespace b4c {
constexpr double MAX_EXPL = 2.688117142e+43;
constexpr double MIN_EXPL = 3.720075976e-44;
constexpr double EXPL_THRESHOLD = 100.0;
constexpr double MAX_EXP = 5.834617425e14;
constexpr double MIN_EXP = 1.713908431e-15;
constexpr double EXP_THRESHOLD = 34.0;
constexpr double EPSSI = 1.03594e-10;
constexpr double Charge_q = 1.60219e-19;
constexpr double DELTA_1 = 0.02;
constexpr double DELTA_3 = 0.02;
constexpr double DELTA_4 = 0.02;
constexpr double CHARGE = 1.6021766208e-19;
constexpr double CONSTvt0 = 1.38064852e-23 * (27.0 + 273.15) / 1.6021766208e-19;
}

NGMI_HD double b4_max(double a, double b) { return a > b ? a : b; }
NGMI_HD double b4_min(double a, double b) { return a < b ? a : b; }

NGMI_HD void b4_dexp(double A, double &B, double &C)
{   /* b4ld.c:48-59 */
    using namespace b4c;
    if (A > EXP_THRESHOLD) { B = MAX_EXP * (1.0 + A - EXP_THRESHOLD); C = MAX_EXP; }
    else if (A < -EXP_THRESHOLD) { B = MIN_EXP; C = 0; }
    else { B = exp(A); C = B; }
}

/* devsup.c:21-40 */
NGMI_HD double dev_limvds(double vnew, double vold)
{
    if (vold >= 3.5) {
        if (vnew > vold) vnew = b4_min(vnew, (3 * vold) + 2);
        else if (vnew < 3.5) vnew = b4_max(vnew, 2);
    } else {
        if (vnew > vold) vnew = b4_min(vnew, 4);
        else vnew = b4_max(vnew, -.5);
    }
    return vnew;
}

/* devsup.c:50-84 */
NGMI_HD double dev_pnjlim(double vnew, double vold, double vt, double vcrit, int &icheck)
{
    double arg;
    if ((vnew > vcrit) && (fabs(vnew - vold) > (vt + vt))) {
        if (vold > 0) {
            arg = (vnew - vold) / vt;
            if (arg > 0) vnew = vold + vt * (2 + log(arg - 2));
            else vnew = vold - vt * (2 + log(2 - arg));
        } else vnew = vt * log(vnew / vt);
        icheck = 1;
    } else {
        if (vnew < 0) {
            arg = (vold > 0) ? -1 * vold - 1 : 2 * vold - 1;
            if (vnew < arg) { vnew = arg; icheck = 1; }
            else icheck = 0;
        } else icheck = 0;
    }
    return vnew;
}

/* devsup.c:93-151 */
NGMI_HD double dev_fetlim(double vnew, double vold, double vto)
{
    double vtsthi = fabs(2 * (vold - vto)) + 2;
    double vtstlo = fabs(vold - vto) + 1;
    double vtox = vto + 3.5;
    double delv = vnew - vold;
    if (vold >= vto) {
        if (vold >= vtox) {
            if (delv <= 0) {
                if (vnew >= vtox) { if (-delv > vtstlo) vnew = vold - vtstlo; }
                else vnew = b4_max(vnew, vto + 2);
            } else if (delv >= vtsthi) vnew = vold + vtsthi;
        } else {
            if (delv <= 0) vnew = b4_max(vnew, vto - .5);
            else vnew = b4_min(vnew, vto + 4);
        }
    } else {
        if (delv <= 0) { if (-delv > vtsthi) vnew = vold - vtsthi; }
        else {
            double vtemp = vto + .5;
            if (vnew <= vtemp) { if (delv > vtstlo) vnew = vold + vtstlo; }
            else vnew = vtemp;
        }
    }
    return vnew;
}

/* NIintegrate (niinteg.c:18-80) on charge state qcap; returns ceq */
template <class S>
NGMI_HD double b4_integrate(S &st, const B4Ckt &ck, int qcap)
{
    const int ccap = qcap + 1;
    double c;
    if (ck.method == 1) {
        if (ck.order == 1) c = ck.ag[0] * st.s(0, qcap) + ck.ag[1] * st.s(1, qcap);
        else c = -st.s(1, ccap) * ck.ag[1] + ck.ag[0] * (st.s(0, qcap) - st.s(1, qcap));
    } else {
        c = 0;
        for (int k = ck.order; k >= 2; k--) c += ck.ag[k] * st.s(k, qcap);
        c += ck.ag[1] * st.s(1, qcap);
        c += ck.ag[0] * st.s(0, qcap);
    }
    st.set(0, ccap, c);
    return c - ck.ag[0] * st.s(0, qcap);
}

/* junction diode DC model for one side (b4ld.c:700-860) */
NGMI_HD void b4_junction(int dioMod, double Isat, double vj, double Nvtm, double gmin,
                         double xjbv, double bv, double XExpBV, double vjmFwd, double IVjmFwd,
                         double vjmRev, double IVjmRev, double slpFwd, double slpRev,
                         double &g, double &c)
{
    using namespace b4c;
    double ev, T0, T1, T2, T3, dev;
    if (Isat <= 0.0) { g = gmin; c = g * vj; return; }
    if (dioMod == 0) {
        ev = exp(vj / Nvtm);
        T1 = xjbv * exp(-(bv + vj) / Nvtm);
        g = Isat * (ev + T1) / Nvtm + gmin;
        c = Isat * (ev + XExpBV - T1 - 1.0) + gmin * vj;
    } else if (dioMod == 1) {
        T2 = vj / Nvtm;
        if (T2 < -EXP_THRESHOLD) { g = gmin; c = Isat * (MIN_EXP - 1.0) + gmin * vj; }
        else if (vj <= vjmFwd) {
            ev = exp(T2);
            g = Isat * ev / Nvtm + gmin;
            c = Isat * (ev - 1.0) + gmin * vj;
        } else {
            T0 = IVjmFwd / Nvtm;
            g = T0 + gmin;
            c = IVjmFwd - Isat + T0 * (vj - vjmFwd) + gmin * vj;
        }
    } else {
        if (vj < vjmRev) {
            T0 = vj / Nvtm;
            if (T0 < -EXP_THRESHOLD) { ev = MIN_EXP; dev = 0.0; }
            else { ev = exp(T0); dev = ev / Nvtm; }
            T1 = ev - 1.0;
            T2 = IVjmRev + slpRev * (vj - vjmRev);
            g = dev * T2 + T1 * slpRev + gmin;
            c = T1 * T2 + gmin * vj;
        } else if (vj <= vjmFwd) {
            T0 = vj / Nvtm;
            if (T0 < -EXP_THRESHOLD) { ev = MIN_EXP; dev = 0.0; }
            else { ev = exp(T0); dev = ev / Nvtm; }
            T1 = (bv + vj) / Nvtm;
            if (T1 > EXP_THRESHOLD) { T2 = MIN_EXP; T3 = 0.0; }
            else { T2 = exp(-T1); T3 = -T2 / Nvtm; }
            g = Isat * (dev - xjbv * T3) + gmin;
            c = Isat * (ev + XExpBV - 1.0 - xjbv * T2) + gmin * vj;
        } else {
            g = slpFwd + gmin;
            c = IVjmFwd + slpFwd * (vj - vjmFwd) + gmin * vj;
        }
    }
}

/* trap-assisted tunnelling factor (b4ld.c:869-988) */
NGMI_HD void b4_tat(double vts, double vj, double Nvtmr, double &T, double &dT)
{
    double T0, T9, T10;
    if ((vts - vj) < (vts * 1e-3)) {
        T9 = 1.0e3;
        T0 = -vj / Nvtmr * T9;
        b4_dexp(T0, T, T10);
        dT = T10 / Nvtmr * T9;
    } else {
        T9 = 1.0 / (vts - vj);
        T0 = -vj / Nvtmr * vts * T9;
        double dT0 = vts / Nvtmr * (T9 + vj * T9 * T9);
        b4_dexp(T0, T, T10);
        dT = T10 * dT0;
    }
}

/* junction depletion charge (b4ld.c:3965-4063) */
NGMI_HD void b4_jcv(double vj, double czb, double czbsw, double czbswg, double MJ, double MJSW,
                    double MJSWG, double PhiB, double PhiBSW, double PhiBSWG, double &q, double &cap)
{
    double arg, sarg;
    if (vj == 0.0) { q = 0.0; cap = czb + czbsw + czbswg; }
    else if (vj < 0.0) {
        if (czb > 0.0) {
            arg = 1.0 - vj / PhiB;
            sarg = (MJ == 0.5) ? 1.0 / sqrt(arg) : exp(-MJ * log(arg));
            q = PhiB * czb * (1.0 - arg * sarg) / (1.0 - MJ);
            cap = czb * sarg;
        } else { q = 0.0; cap = 0.0; }
        if (czbsw > 0.0) {
            arg = 1.0 - vj / PhiBSW;
            sarg = (MJSW == 0.5) ? 1.0 / sqrt(arg) : exp(-MJSW * log(arg));
            q += PhiBSW * czbsw * (1.0 - arg * sarg) / (1.0 - MJSW);
            cap += czbsw * sarg;
        }
        if (czbswg > 0.0) {
            arg = 1.0 - vj / PhiBSWG;
            sarg = (MJSWG == 0.5) ? 1.0 / sqrt(arg) : exp(-MJSWG * log(arg));
            q += PhiBSWG * czbswg * (1.0 - arg * sarg) / (1.0 - MJSWG);
            cap += czbswg * sarg;
        }
    } else {
        double T0 = czb + czbsw + czbswg;
        double T1 = vj * (czb * MJ / PhiB + czbsw * MJSW / PhiBSW + czbswg * MJSWG / PhiBSWG);
        q = vj * (T0 + 0.5 * T1);
        cap = T0 + T1;
    }
}

/* poly-gate depletion (b4ld.c:5402-5433) */
NGMI_HD void b4_polydep(double phi, double ngate, double coxe, double Vgs, double &Vgs_eff,
                        double &dVgs_eff_dVg)
{
    using namespace b4c;
    if ((ngate > 1.0e18) && (ngate < 1.0e25) && (Vgs > phi)) {
        double T1 = 1.0e6 * CHARGE * EPSSI * ngate / (coxe * coxe);
        double T8 = Vgs - phi;
        double T4 = sqrt(1.0 + 2.0 * T8 / T1);
        double T2 = 2.0 * T8 / (T4 + 1.0);
        double T3 = 0.5 * T2 * T2 / T1;
        double T7 = 1.12 - T3 - 0.05;
        double T6 = sqrt(T7 * T7 + 0.224);
        double T5 = 1.12 - 0.5 * (T7 + T6);
        Vgs_eff = Vgs - T5;
        dVgs_eff_dVg = 1.0 - (0.5 - 0.5 / T4) * (1.0 + T7 / T6);
    } else {
        Vgs_eff = Vgs;
        dVgs_eff_dVg = 1.0;
    }
}

/*
 * One instance evaluation.  vnode = node voltages of (dp, gp, sp, bp) from
 * the previous Newton iterate (rhsOld).  Writes stamp[16], rhs[4], state
 * level 0 (and level 1 in INITTRAN), the cache, and returns the number of
 * non-convergence votes (0 or 1) this instance adds to CKTnoncon.
 */
/* FL >= 0: every instance has selector word FL, compiled as a constant so
 * the selector branches fold away; FL = -1 reads it per instance */
template <int FL = -1, class P, class S, class C>
NGMI_HD int b4_load(const P &prm, S &st, C &ca, const B4Ckt &ck, const double *vnode,
                    double *stamp, double *rhs)
{
    using namespace b4c;
#define PR(f) prm(B4F_##f)
#define CA(f) ca.c(B4C_##f)
#define SETC(f, v) ca.setc(B4C_##f, (v))
    const double type = PR(type);
    const int flags = (FL >= 0) ? FL : (int)PR(flags);
    const int mobMod = flags & B4FL_MOBMOD_MASK;
    const int dioMod = (flags >> B4FL_DIOMOD_SHIFT) & 3;
    const int igcMod = (flags >> B4FL_IGCMOD_SHIFT) & 3;
    const int igbMod = (flags & B4FL_IGBMOD) ? 1 : 0;
    const int tempMod = (flags >> B4FL_TEMPMOD_SHIFT) & 3;
    const int mode = ck.mode;
    const double gmin = ck.gmin;
    const double nf = PR(nf);

    double vds, vgs, vbs, vbd, vgd, vgb, qdef = 0.0;
    double Idtot = 0, cdhat = 0, Ibtot = 0, cbhat = 0, Igstot = 0, cgshat = 0;
    double Igdtot = 0, cgdhat = 0, Igbtot = 0, cgbhat = 0;
    int Check = 1;
    double T0, T1, T2, T3, T4, T5, T6, T7, T8, T9, T10, T11, T12, T13, T14;
    double dT0_dVg, dT0_dVd, dT0_dVb, dT1_dVg, dT1_dVd, dT1_dVb, dT2_dVg, dT2_dVd, dT2_dVb;
    double dT3_dVd, dT3_dVb, dT4_dVd, dT5_dVb;
    double dT6_dVg, dT6_dVd, dT6_dVb, dT7_dVg, dT7_dVd, dT7_dVb, dT8_dVg, dT8_dVd, dT8_dVb;
    double dT9_dVg, dT9_dVd, dT9_dVb, dT10_dVg, dT10_dVd, dT10_dVb;
    double tmp, tmp1, tmp2, tmp3, tmp4;

    /* ---- terminal voltages by mode (b4ld.c:261-407) */
    if (mode & M_INITSMSIG) {
        vds = st.s(0, B4X_vds); vgs = st.s(0, B4X_vgs); vbs = st.s(0, B4X_vbs);
    } else if (mode & M_INITTRAN) {
        vds = st.s(1, B4X_vds); vgs = st.s(1, B4X_vgs); vbs = st.s(1, B4X_vbs);
    } else if ((mode & M_INITJCT) && PR(off) == 0.0) {
        vds = type * PR(icVDS);
        vgs = type * PR(icVGS);
        vbs = type * PR(icVBS);
        if ((vds == 0.0) && (vgs == 0.0) && (vbs == 0.0) &&
            ((mode & (M_TRAN | M_AC | M_DCOP | M_DCTRANCURVE)) || (!(mode & M_UIC)))) {
            vds = 0.1;
            vgs = PR(tvth0) + 0.1;
            vbs = 0.0;
        }
    } else if ((mode & (M_INITJCT | M_INITFIX)) && PR(off) != 0.0) {
        vds = vgs = vbs = 0.0;
    } else {
        const double s0vds = st.s(0, B4X_vds), s0vgs = st.s(0, B4X_vgs), s0vbs = st.s(0, B4X_vbs);
        const double s0vbd = st.s(0, B4X_vbd);
        double delvbd, delvgd, delvds, delvgs, delvbs, vgdo;
        const int from1 = mode & (M_INITPRED | M_GUESSED);
        if (from1) {
            /* b4ld.c:322-372: state0 takes state1, and in INITPRED the
             * voltages extrapolate */
            const double s1vds = st.s(1, B4X_vds), s1vgs = st.s(1, B4X_vgs), s1vbs = st.s(1, B4X_vbs);
            const double s1vdbs = st.s(1, B4X_vdbs);
            st.set(0, B4X_vds, s1vds);
            st.set(0, B4X_vgs, s1vgs);
            st.set(0, B4X_vbs, s1vbs);
            st.set(0, B4X_vbd, s1vbs - s1vds);
            st.set(0, B4X_vdbs, s1vdbs);
            st.set(0, B4X_vdbd, s1vdbs - s1vds);
            if (mode & M_INITPRED) {
                const double xf = ck.xfact;
                vds = (1.0 + xf) * s1vds - (xf * st.s(2, B4X_vds));
                vgs = (1.0 + xf) * s1vgs - (xf * st.s(2, B4X_vgs));
                vbs = (1.0 + xf) * s1vbs - (xf * st.s(2, B4X_vbs));
            }
        }
        if (!(mode & M_INITPRED)) {
            vds = type * (vnode[0] - vnode[2]);
            vgs = type * (vnode[1] - vnode[2]);
            vbs = type * (vnode[3] - vnode[2]);
        }
        const double r0vds = from1 ? st.s(0, B4X_vds) : s0vds;
        const double r0vgs = from1 ? st.s(0, B4X_vgs) : s0vgs;
        const double r0vbs = from1 ? st.s(0, B4X_vbs) : s0vbs;
        const double r0vbd = from1 ? st.s(0, B4X_vbd) : s0vbd;
        vgdo = r0vgs - r0vds;
        vbd = vbs - vds;
        vgd = vgs - vds;
        delvbd = vbd - r0vbd;
        delvgd = vgd - vgdo;
        delvds = vds - r0vds;
        delvgs = vgs - r0vgs;
        delvbs = vbs - r0vbs;
        {   /* hat currents (b4ld.c:409-507), rbodyMod 0, rdsMod 0 */
            const double gbd = CA(gbd), gbs = CA(gbs), gm = CA(gm), gds = CA(gds), gmbs = CA(gmbs);
            const double gbbs = CA(gbbs), gbgs = CA(gbgs), gbds = CA(gbds);
            const double ggidld = CA(ggidld), ggidlg = CA(ggidlg), ggidlb = CA(ggidlb);
            const double ggisls = CA(ggisls), ggislg = CA(ggislg), ggislb = CA(ggislb);
            const double ggidls = -(ggidld + ggidlg + ggidlb);
            const double ggisld = -(ggisls + ggislg + ggislb);
            if (CA(mode) >= 0) {
                Idtot = CA(cd) + CA(csub) - CA(cbd) + CA(Igidl);
                cdhat = Idtot - gbd * delvbd + (gmbs + gbbs + ggidlb) * delvbs
                      + (gm + gbgs + ggidlg) * delvgs + (gds + gbds + ggidld) * delvds;
                Ibtot = CA(cbs) + CA(cbd) - CA(Igidl) - CA(Igisl) - CA(csub);
                cbhat = Ibtot + gbd * delvbd + gbs * delvbs - (gbbs + ggidlb) * delvbs
                      - (gbgs + ggidlg) * delvgs - (gbds + ggidld - ggisls) * delvds
                      - ggislg * delvgd - ggislb * delvbd;
                Igstot = CA(Igs) + CA(Igcs);
                cgshat = Igstot + (CA(gIgsg) + CA(gIgcsg)) * delvgs + CA(gIgcsd) * delvds + CA(gIgcsb) * delvbs;
                Igdtot = CA(Igd) + CA(Igcd);
                cgdhat = Igdtot + CA(gIgdg) * delvgd + CA(gIgcdg) * delvgs + CA(gIgcdd) * delvds + CA(gIgcdb) * delvbs;
                Igbtot = CA(Igb);
                cgbhat = CA(Igb) + CA(gIgbg) * delvgs + CA(gIgbd) * delvds + CA(gIgbb) * delvbs;
            } else {
                Idtot = CA(cd) + CA(cbd) - CA(Igidl);
                cdhat = Idtot + gbd * delvbd + gmbs * delvbd + gm * delvgd
                      - (gds + ggidls) * delvds - ggidlg * delvgs - ggidlb * delvbs;
                Ibtot = CA(cbs) + CA(cbd) - CA(Igidl) - CA(Igisl) - CA(csub);
                cbhat = Ibtot + gbs * delvbs + gbd * delvbd - (gbbs + ggislb) * delvbd
                      - (gbgs + ggislg) * delvgd + (gbds + ggisld - ggidls) * delvds
                      - ggidlg * delvgs - ggidlb * delvbs;
                Igstot = CA(Igs) + CA(Igcd);
                cgshat = Igstot + CA(gIgsg) * delvgs + CA(gIgcdg) * delvgd - CA(gIgcdd) * delvds + CA(gIgcdb) * delvbd;
                Igdtot = CA(Igd) + CA(Igcs);
                cgdhat = Igdtot + (CA(gIgdg) + CA(gIgcsg)) * delvgd - CA(gIgcsd) * delvds + CA(gIgcsb) * delvbd;
                Igbtot = CA(Igb);
                cgbhat = CA(Igb) + CA(gIgbg) * delvgd - CA(gIgbd) * delvds + CA(gIgbb) * delvbd;
            }
        }
        /* limiting (b4ld.c:605-686) */
        const double von = CA(von);
        if (r0vds >= 0.0) {
            vgs = dev_fetlim(vgs, r0vgs, von);
            vds = vgs - vgd;
            vds = dev_limvds(vds, r0vds);
            vgd = vgs - vds;
        } else {
            vgd = dev_fetlim(vgd, vgdo, von);
            vds = vgs - vgd;
            vds = -dev_limvds(-vds, -r0vds);
            vgs = vgd + vds;
        }
        if (vds >= 0.0) {
            vbs = dev_pnjlim(vbs, r0vbs, CONSTvt0, PR(vcrit), Check);
            vbd = vbs - vds;
        } else {
            vbd = dev_pnjlim(vbd, r0vbd, CONSTvt0, PR(vcrit), Check);
            vbs = vbd + vds;
        }
    }

    /* ---- DC currents (b4ld.c:688-1002) */
    vbd = vbs - vds;
    vgd = vgs - vds;
    vgb = vgs - vbs;
    const double vbs_jct = vbs, vbd_jct = vbd;
    double gbs, cbs, gbd, cbd;
    b4_junction(dioMod, PR(IsatS), vbs_jct, PR(Nvtms), gmin, PR(xjbvs), PR(bvs), PR(XExpBVS),
                PR(vjsmFwd), PR(IVjsmFwd), PR(vjsmRev), PR(IVjsmRev), PR(SslpFwd), PR(SslpRev), gbs, cbs);
    b4_junction(dioMod, PR(IsatD), vbd_jct, PR(Nvtmd), gmin, PR(xjbvd), PR(bvd), PR(XExpBVD),
                PR(vjdmFwd), PR(IVjdmFwd), PR(vjdmRev), PR(IVjdmRev), PR(DslpFwd), PR(DslpRev), gbd, cbd);
    {
        double Ta, Tb, Tc, Td, Te, Tf, d1, d2, d3, d4, d5, d6;
        b4_tat(PR(vtss), vbs_jct, PR(Nvtmrss), Ta, d1);
        b4_tat(PR(vtsd), vbd_jct, PR(Nvtmrsd), Tb, d2);
        b4_tat(PR(vtssws), vbs_jct, PR(Nvtmrssws), Tc, d3);
        b4_tat(PR(vtsswd), vbd_jct, PR(Nvtmrsswd), Td, d4);
        b4_tat(PR(vtsswgs), vbs_jct, PR(Nvtmrsswgs), Te, d5);
        b4_tat(PR(vtsswgd), vbd_jct, PR(Nvtmrsswgd), Tf, d6);
        gbs += PR(SjctTempRevSatCur) * d1 + PR(SswTempRevSatCur) * d3 + PR(SswgTempRevSatCur) * d5;
        cbs -= PR(SjctTempRevSatCur) * (Ta - 1.0) + PR(SswTempRevSatCur) * (Tc - 1.0)
             + PR(SswgTempRevSatCur) * (Te - 1.0);
        gbd += PR(DjctTempRevSatCur) * d2 + PR(DswTempRevSatCur) * d4 + PR(DswgTempRevSatCur) * d6;
        cbd -= PR(DjctTempRevSatCur) * (Tb - 1.0) + PR(DswTempRevSatCur) * (Td - 1.0)
             + PR(DswgTempRevSatCur) * (Tf - 1.0);
    }
    int imode;
    double Vds, Vgs, Vbs;
    if (vds >= 0.0) { imode = 1; Vds = vds; Vgs = vgs; Vbs = vbs; }
    else { imode = -1; Vds = -vds; Vgs = vgd; Vbs = vbd; }
    const double toxe = PR(toxe), epssub = EPSSI;
    const double coxe = PR(coxe), Vtm = PR(vtm), Vtm0 = PR(vtm0);

    /* ---- Vbseff, Vth, n (b4ld.c:1036-1245) */
    const double vbsc = PR(vbsc), phi = PR(phi), sqrtPhi = PR(sqrtPhi);
    double Vbseff, dVbseff_dVb;
    T0 = Vbs - vbsc - 0.001;
    T1 = sqrt(T0 * T0 - 0.004 * vbsc);
    if (T0 >= 0.0) { Vbseff = vbsc + 0.5 * (T0 + T1); dVbseff_dVb = 0.5 * (1.0 + T0 / T1); }
    else { T2 = -0.002 / (T1 - T0); Vbseff = vbsc * (1.0 + T2); dVbseff_dVb = T2 * vbsc / T1; }
    T9 = 0.95 * phi;
    T0 = T9 - Vbseff - 0.001;
    T1 = sqrt(T0 * T0 + 0.004 * T9);
    Vbseff = T9 - 0.5 * (T0 + T1);
    dVbseff_dVb *= 0.5 * (1.0 + T0 / T1);
    const double Phis = phi - Vbseff, dPhis_dVb = -1.0;
    const double sqrtPhis = sqrt(Phis), dsqrtPhis_dVb = -0.5 / sqrtPhis;
    const double Xdep0 = PR(Xdep0);
    const double Xdep = Xdep0 * sqrtPhis / sqrtPhi;
    const double dXdep_dVb = (Xdep0 / sqrtPhi) * dsqrtPhis_dVb;
    const double Leff = PR(leff), weffP = PR(weff), factor1 = PR(factor1);
    T3 = sqrt(Xdep);
    const double V0 = PR(vbi) - phi;
    double lt1, dlt1_dVb, ltw, dltw_dVb, Theta0, dTheta0_dVb;
    T0 = PR(dvt2) * Vbseff;
    if (T0 >= -0.5) { T1 = 1.0 + T0; T2 = PR(dvt2); }
    else { T4 = 1.0 / (3.0 + 8.0 * T0); T1 = (1.0 + 3.0 * T0) * T4; T2 = PR(dvt2) * T4 * T4; }
    lt1 = factor1 * T3 * T1;
    dlt1_dVb = factor1 * (0.5 / T3 * T1 * dXdep_dVb + T3 * T2);
    T0 = PR(dvt2w) * Vbseff;
    if (T0 >= -0.5) { T1 = 1.0 + T0; T2 = PR(dvt2w); }
    else { T4 = 1.0 / (3.0 + 8.0 * T0); T1 = (1.0 + 3.0 * T0) * T4; T2 = PR(dvt2w) * T4 * T4; }
    ltw = factor1 * T3 * T1;
    dltw_dVb = factor1 * (0.5 / T3 * T1 * dXdep_dVb + T3 * T2);
    T0 = PR(dvt1) * Leff / lt1;
    if (T0 < EXP_THRESHOLD) {
        T1 = exp(T0); T2 = T1 - 1.0; T3 = T2 * T2; T4 = T3 + 2.0 * T1 * MIN_EXP;
        Theta0 = T1 / T4;
        dT1_dVb = -T0 * T1 * dlt1_dVb / lt1;
        dTheta0_dVb = dT1_dVb * (T4 - 2.0 * T1 * (T2 + MIN_EXP)) / T4 / T4;
    } else { Theta0 = 1.0 / (MAX_EXP - 2.0); dTheta0_dVb = 0.0; }
    const double thetavth = PR(dvt0) * Theta0;
    const double Delt_vth = thetavth * V0;
    const double dDelt_vth_dVb = PR(dvt0) * dTheta0_dVb * V0;
    T0 = PR(dvt1w) * weffP * Leff / ltw;
    if (T0 < EXP_THRESHOLD) {
        T1 = exp(T0); T2 = T1 - 1.0; T3 = T2 * T2; T4 = T3 + 2.0 * T1 * MIN_EXP;
        T5 = T1 / T4;
        dT1_dVb = -T0 * T1 * dltw_dVb / ltw;
        dT5_dVb = dT1_dVb * (T4 - 2.0 * T1 * (T2 + MIN_EXP)) / T4 / T4;
    } else { T5 = 1.0 / (MAX_EXP - 2.0); dT5_dVb = 0.0; }
    T0 = PR(dvt0w) * T5;
    T2 = T0 * V0;
    dT2_dVb = PR(dvt0w) * dT5_dVb * V0;
    const double TempRatio = PR(TempRatio);
    T0 = sqrt(1.0 + PR(lpe0) / Leff);
    const double k1ox = PR(k1ox), k1 = PR(k1), k2ox = PR(k2ox);
    T1 = k1ox * (T0 - 1.0) * sqrtPhi + (PR(kt1) + PR(kt1l) / Leff + PR(kt2) * Vbseff) * TempRatio;
    const double Vth_NarrowW = toxe * phi / (weffP + PR(w0));
    T3 = PR(eta0) + PR(etab) * Vbseff;
    if (T3 < 1.0e-4) { T9 = 1.0 / (3.0 - 2.0e4 * T3); T3 = (2.0e-4 - T3) * T9; T4 = T9 * T9; }
    else T4 = 1.0;
    const double theta0vb0 = PR(theta0vb0);
    const double dDIBL_Sft_dVd = T3 * theta0vb0;
    const double DIBL_Sft = dDIBL_Sft_dVd * Vds;
    const double Lpe_Vb = sqrt(1.0 + PR(lpeb) / Leff);
    double Vth = PR(tvth0) + (k1ox * sqrtPhis - k1 * sqrtPhi) * Lpe_Vb
               - k2ox * Vbseff - Delt_vth - T2 + (PR(k3) + PR(k3b) * Vbseff) * Vth_NarrowW + T1 - DIBL_Sft;
    double dVth_dVb = Lpe_Vb * k1ox * dsqrtPhis_dVb - k2ox - dDelt_vth_dVb - dT2_dVb
                    + PR(k3b) * Vth_NarrowW - PR(etab) * Vds * theta0vb0 * T4 + PR(kt2) * TempRatio;
    double dVth_dVd = -dDIBL_Sft_dVd;
    double n, dn_dVb, dn_dVd;
    tmp1 = epssub / Xdep;
    tmp2 = PR(nfactor) * tmp1;
    tmp3 = PR(cdsc) + PR(cdscb) * Vbseff + PR(cdscd) * Vds;
    tmp4 = (tmp2 + tmp3 * Theta0 + PR(cit)) / coxe;
    if (tmp4 >= -0.5) {
        n = 1.0 + tmp4;
        dn_dVb = (-tmp2 / Xdep * dXdep_dVb + tmp3 * dTheta0_dVb + PR(cdscb) * Theta0) / coxe;
        dn_dVd = PR(cdscd) * Theta0 / coxe;
    } else {
        T0 = 1.0 / (3.0 + 8.0 * tmp4);
        n = (1.0 + 3.0 * tmp4) * T0;
        T0 *= T0;
        dn_dVb = (-tmp2 / Xdep * dXdep_dVb + tmp3 * dTheta0_dVb + PR(cdscb) * Theta0) / coxe * T0;
        dn_dVd = PR(cdscd) * Theta0 / coxe * T0;
    }
    if (PR(dvtp0) > 0.0) {
        T0 = -PR(dvtp1) * Vds;
        double dT2d;
        if (T0 < -EXP_THRESHOLD) { T2 = MIN_EXP; dT2d = 0.0; }
        else { T2 = exp(T0); dT2d = -PR(dvtp1) * T2; }
        T3 = Leff + PR(dvtp0) * (1.0 + T2);
        dT3_dVd = PR(dvtp0) * dT2d;
        if (tempMod < 2) { T4 = Vtm * log(Leff / T3); dT4_dVd = -Vtm * dT3_dVd / T3; }
        else { T4 = Vtm0 * log(Leff / T3); dT4_dVd = -Vtm0 * dT3_dVd / T3; }
        const double dDITS_Sft_dVd = dn_dVd * T4 + n * dT4_dVd;
        const double dDITS_Sft_dVb = T4 * dn_dVb;
        Vth -= n * T4;
        dVth_dVd -= dDITS_Sft_dVd;
        dVth_dVb -= dDITS_Sft_dVb;
    }
    if (!((PR(dvtp4) == 0.0) || (PR(dvtp2factor) == 0.0))) {
        T1 = 2.0 * PR(dvtp4) * Vds;
        b4_dexp(T1, T0, T10);
        const double DITS_Sft2 = PR(dvtp2factor) * (T0 - 1) / (T0 + 1);
        const double dDITS_Sft2_dVd = PR(dvtp2factor) * PR(dvtp4) * 4.0 * T10 / ((T0 + 1) * (T0 + 1));
        Vth -= DITS_Sft2;
        dVth_dVd -= dDITS_Sft2_dVd;
    }
    SETC(von, Vth);

    /* ---- poly depletion, Vgsteff (b4ld.c:1245-1332) */
    double vgs_eff, dvgs_eff_dvg, vgd_eff, dvgd_eff_dvg;
    T0 = PR(vfb) + phi;
    b4_polydep(T0, PR(ngate), coxe, vgs, vgs_eff, dvgs_eff_dvg);
    b4_polydep(T0, PR(ngate), coxe, vgd, vgd_eff, dvgd_eff_dvg);
    const double Vgs_eff = imode > 0 ? vgs_eff : vgd_eff;
    const double dVgs_eff_dVg = imode > 0 ? dvgs_eff_dvg : dvgd_eff_dvg;
    const double Vgst = Vgs_eff - Vth;
    const double mstar = PR(mstar), cdep0 = PR(cdep0);
    T0 = n * Vtm;
    T1 = mstar * Vgst;
    T2 = T1 / T0;
    if (T2 > EXP_THRESHOLD) {
        T10 = T1;
        dT10_dVg = mstar * dVgs_eff_dVg;
        dT10_dVd = -dVth_dVd * mstar;
        dT10_dVb = -dVth_dVb * mstar;
    } else if (T2 < -EXP_THRESHOLD) {
        T10 = Vtm * log(1.0 + MIN_EXP);
        dT10_dVg = 0.0;
        dT10_dVd = T10 * dn_dVd;
        dT10_dVb = T10 * dn_dVb;
        T10 *= n;
    } else {
        const double ExpVgst = exp(T2);
        T3 = Vtm * log(1.0 + ExpVgst);
        T10 = n * T3;
        dT10_dVg = mstar * ExpVgst / (1.0 + ExpVgst);
        dT10_dVb = T3 * dn_dVb - dT10_dVg * (dVth_dVb + Vgst * dn_dVb / n);
        dT10_dVd = T3 * dn_dVd - dT10_dVg * (dVth_dVd + Vgst * dn_dVd / n);
        dT10_dVg *= dVgs_eff_dVg;
    }
    T1 = PR(voffcbn) - (1.0 - mstar) * Vgst;
    T2 = T1 / T0;
    if (T2 < -EXP_THRESHOLD) {
        T3 = coxe * MIN_EXP / cdep0;
        T9 = mstar + T3 * n;
        dT9_dVg = 0.0; dT9_dVd = dn_dVd * T3; dT9_dVb = dn_dVb * T3;
    } else if (T2 > EXP_THRESHOLD) {
        T3 = coxe * MAX_EXP / cdep0;
        T9 = mstar + T3 * n;
        dT9_dVg = 0.0; dT9_dVd = dn_dVd * T3; dT9_dVb = dn_dVb * T3;
    } else {
        const double ExpVgst = exp(T2);
        T3 = coxe / cdep0;
        T4 = T3 * ExpVgst;
        T5 = T1 * T4 / T0;
        T9 = mstar + n * T4;
        dT9_dVg = T3 * (mstar - 1.0) * ExpVgst / Vtm;
        dT9_dVb = T4 * dn_dVb - dT9_dVg * dVth_dVb - T5 * dn_dVb;
        dT9_dVd = T4 * dn_dVd - dT9_dVg * dVth_dVd - T5 * dn_dVd;
        dT9_dVg *= dVgs_eff_dVg;
    }
    double Vgsteff = T10 / T9;
    T11 = T9 * T9;
    double dVgsteff_dVg = (T9 * dT10_dVg - T10 * dT9_dVg) / T11;
    double dVgsteff_dVd = (T9 * dT10_dVd - T10 * dT9_dVd) / T11;
    double dVgsteff_dVb = (T9 * dT10_dVb - T10 * dT9_dVb) / T11;

    /* ---- effective geometry, Rds, Abulk (b4ld.c:1334-1450) */
    T9 = sqrtPhis - sqrtPhi;
    double Weff = weffP - 2.0 * (PR(dwg) * Vgsteff + PR(dwb) * T9);
    double dWeff_dVg = -2.0 * PR(dwg);
    double dWeff_dVb = -2.0 * PR(dwb) * dsqrtPhis_dVb;
    if (Weff < 2.0e-8) {
        T0 = 1.0 / (6.0e-8 - 2.0 * Weff);
        Weff = 2.0e-8 * (4.0e-8 - Weff) * T0;
        T0 *= T0 * 4.0e-16;
        dWeff_dVg *= T0;
        dWeff_dVb *= T0;
    }
    T0 = 1.0 + PR(prwg) * Vgsteff;
    dT0_dVg = -PR(prwg) / T0 / T0;
    T1 = PR(prwb) * T9;
    dT1_dVb = PR(prwb) * dsqrtPhis_dVb;
    T2 = 1.0 / T0 + T1;
    T3 = T2 + sqrt(T2 * T2 + 0.01);
    double dT3_dVg = 1.0 + T2 / (T3 - T2);
    dT3_dVb = dT3_dVg * dT1_dVb;
    dT3_dVg *= dT0_dVg;
    T4 = PR(rds0) * 0.5;
    const double Rds = PR(rdswmin) + T3 * T4;
    const double dRds_dVg = T4 * dT3_dVg;
    const double dRds_dVb = T4 * dT3_dVb;

    T9 = 0.5 * k1ox * Lpe_Vb / sqrtPhis;
    T1 = T9 + k2ox - PR(k3b) * Vth_NarrowW;
    dT1_dVb = -T9 / sqrtPhis * dsqrtPhis_dVb;
    T9 = sqrt(PR(xj) * Xdep);
    tmp1 = Leff + 2.0 * T9;
    T5 = Leff / tmp1;
    tmp2 = PR(a0) * T5;
    tmp3 = weffP + PR(b1);
    tmp4 = PR(b0) / tmp3;
    T2 = tmp2 + tmp4;
    dT2_dVb = -T9 / tmp1 / Xdep * dXdep_dVb;
    T6 = T5 * T5;
    T7 = T5 * T6;
    double Abulk0 = 1.0 + T1 * T2;
    double dAbulk0_dVb = T1 * tmp2 * dT2_dVb + T2 * dT1_dVb;
    T8 = PR(ags) * PR(a0) * T7;
    double dAbulk_dVg = -T1 * T8;
    double Abulk = Abulk0 + dAbulk_dVg * Vgsteff;
    double dAbulk_dVb = dAbulk0_dVb - T8 * Vgsteff * (dT1_dVb + 3.0 * T1 * dT2_dVb);
    if (Abulk0 < 0.1) {
        T9 = 1.0 / (3.0 - 20.0 * Abulk0);
        Abulk0 = (0.2 - Abulk0) * T9;
        dAbulk0_dVb *= T9 * T9;
    }
    if (Abulk < 0.1) {
        T9 = 1.0 / (3.0 - 20.0 * Abulk);
        Abulk = (0.2 - Abulk) * T9;
        T10 = T9 * T9;
        dAbulk_dVb *= T10;
        dAbulk_dVg *= T10;
    }
    double dT0_dVbk;
    T2 = PR(keta) * Vbseff;
    if (T2 >= -0.9) { T0 = 1.0 / (1.0 + T2); dT0_dVbk = -PR(keta) * T0 * T0; }
    else { T1 = 1.0 / (0.8 + T2); T0 = (17.0 + 20.0 * T2) * T1; dT0_dVbk = -PR(keta) * T1 * T1; }
    dAbulk_dVg *= T0;
    dAbulk_dVb = dAbulk_dVb * T0 + Abulk * dT0_dVbk;
    double dAbulk0_Q_dVb = dAbulk0_dVb;
    dAbulk0_dVb = dAbulk0_dVb * T0 + Abulk0 * dT0_dVbk;
    Abulk *= T0;
    double Abulk0_Q = Abulk0;
    Abulk0 *= T0;
    if (PR(ketac) != PR(keta)) {
        T2 = PR(ketac) * Vbseff;
        if (T2 >= -0.9) { T0 = 1.0 / (1.0 + T2); dT0_dVbk = -PR(ketac) * T0 * T0; }
        else { T1 = 1.0 / (0.8 + T2); T0 = (17.0 + 20.0 * T2) * T1; dT0_dVbk = -PR(ketac) * T1 * T1; }
        dAbulk0_Q_dVb = dAbulk0_Q_dVb * T0 + Abulk0_Q * dT0_dVbk;
        Abulk0_Q *= T0;
    } else {
        dAbulk0_Q_dVb = dAbulk0_dVb;
        Abulk0_Q = Abulk0;
    }

    /* ---- mobility (b4ld.c:1450-1614) */
    double Denomi, dDenomi_dVg, dDenomi_dVd, dDenomi_dVb;
    {
        const double ua = PR(ua), ub = PR(ub), uc = PR(uc), ud = PR(ud);
        const double vtfbphi1 = PR(vtfbphi1);
        T14 = 0.0;
        if (mobMod == 0) {
            T0 = Vgsteff + Vth + Vth - T14;
            T2 = ua + uc * Vbseff;
            T3 = T0 / toxe;
            T12 = sqrt(Vth * Vth + 0.0001);
            T9 = 1.0 / (Vgsteff + 2 * T12);
            T10 = T9 * toxe;
            T8 = ud * T10 * T10 * Vth;
            T6 = T8 * Vth;
            T5 = T3 * (T2 + ub * T3) + T6;
            T7 = -2.0 * T6 * T9;
            T11 = T7 * Vth / T12;
            dDenomi_dVg = (T2 + 2.0 * ub * T3) / toxe;
            T13 = 2.0 * (dDenomi_dVg + T11 + T8);
            dDenomi_dVd = T13 * dVth_dVd;
            dDenomi_dVb = T13 * dVth_dVb + uc * T3;
            dDenomi_dVg += T7;
        } else if (mobMod == 1) {
            T0 = Vgsteff + Vth + Vth - T14;
            T2 = 1.0 + uc * Vbseff;
            T3 = T0 / toxe;
            T4 = T3 * (ua + ub * T3);
            T12 = sqrt(Vth * Vth + 0.0001);
            T9 = 1.0 / (Vgsteff + 2 * T12);
            T10 = T9 * toxe;
            T8 = ud * T10 * T10 * Vth;
            T6 = T8 * Vth;
            T5 = T4 * T2 + T6;
            T7 = -2.0 * T6 * T9;
            T11 = T7 * Vth / T12;
            dDenomi_dVg = (ua + 2.0 * ub * T3) * T2 / toxe;
            T13 = 2.0 * (dDenomi_dVg + T11 + T8);
            dDenomi_dVd = T13 * dVth_dVd;
            dDenomi_dVb = T13 * dVth_dVb + uc * T4;
            dDenomi_dVg += T7;
        } else if (mobMod == 2) {
            T0 = (Vgsteff + vtfbphi1) / toxe;
            T1 = exp(PR(eu) * log(T0));
            dT1_dVg = T1 * PR(eu) / T0 / toxe;
            T2 = ua + uc * Vbseff;
            T12 = sqrt(Vth * Vth + 0.0001);
            T9 = 1.0 / (Vgsteff + 2 * T12);
            T10 = T9 * toxe;
            T8 = ud * T10 * T10 * Vth;
            T6 = T8 * Vth;
            T5 = T1 * T2 + T6;
            T7 = -2.0 * T6 * T9;
            T11 = T7 * Vth / T12;
            dDenomi_dVg = T2 * dT1_dVg + T7;
            T13 = 2.0 * (T11 + T8);
            dDenomi_dVd = T13 * dVth_dVd;
            dDenomi_dVb = T13 * dVth_dVb + T1 * uc;
        } else if (mobMod == 4) {
            T0 = Vgsteff + vtfbphi1 - T14;
            T2 = ua + uc * Vbseff;
            T3 = T0 / toxe;
            T12 = sqrt(vtfbphi1 * vtfbphi1 + 0.0001);
            T9 = 1.0 / (Vgsteff + 2 * T12);
            T10 = T9 * toxe;
            T8 = ud * T10 * T10 * vtfbphi1;
            T6 = T8 * vtfbphi1;
            T5 = T3 * (T2 + ub * T3) + T6;
            T7 = -2.0 * T6 * T9;
            dDenomi_dVg = (T2 + 2.0 * ub * T3) / toxe;
            dDenomi_dVd = 0.0;
            dDenomi_dVb = uc * T3;
            dDenomi_dVg += T7;
        } else if (mobMod == 5) {
            T0 = Vgsteff + vtfbphi1 - T14;
            T2 = 1.0 + uc * Vbseff;
            T3 = T0 / toxe;
            T4 = T3 * (ua + ub * T3);
            T12 = sqrt(vtfbphi1 * vtfbphi1 + 0.0001);
            T9 = 1.0 / (Vgsteff + 2 * T12);
            T10 = T9 * toxe;
            T8 = ud * T10 * T10 * vtfbphi1;
            T6 = T8 * vtfbphi1;
            T5 = T4 * T2 + T6;
            T7 = -2.0 * T6 * T9;
            dDenomi_dVg = (ua + 2.0 * ub * T3) * T2 / toxe;
            dDenomi_dVd = 0.0;
            dDenomi_dVb = uc * T4;
            dDenomi_dVg += T7;
        } else if (mobMod == 6) {
            T0 = (Vgsteff + vtfbphi1) / toxe;
            T1 = exp(PR(eu) * log(T0));
            dT1_dVg = T1 * PR(eu) / T0 / toxe;
            T2 = ua + uc * Vbseff;
            T12 = sqrt(vtfbphi1 * vtfbphi1 + 0.0001);
            T9 = 1.0 / (Vgsteff + 2 * T12);
            T10 = T9 * toxe;
            T8 = ud * T10 * T10 * vtfbphi1;
            T6 = T8 * vtfbphi1;
            T5 = T1 * T2 + T6;
            T7 = -2.0 * T6 * T9;
            dDenomi_dVg = T2 * dT1_dVg + T7;
            dDenomi_dVd = 0;
            dDenomi_dVb = T1 * uc;
        } else {
            const double VgsteffVth = PR(VgsteffVth);
            T0 = (Vgsteff + vtfbphi1) * 1.0e-8 / toxe / 6.0;
            T1 = exp(PR(eu) * log(T0));
            dT1_dVg = T1 * PR(eu) * 1.0e-8 / T0 / toxe / 6.0;
            T2 = ua + uc * Vbseff;
            T10 = exp(PR(ucs) * log(0.5 + 0.5 * Vgsteff / VgsteffVth));
            T11 = ud / T10;
            const double dT11_dVg = -0.5 * PR(ucs) * T11 / (0.5 + 0.5 * Vgsteff / VgsteffVth) / VgsteffVth;
            dDenomi_dVg = T2 * dT1_dVg + dT11_dVg;
            dDenomi_dVd = 0.0;
            dDenomi_dVb = T1 * uc;
            T5 = T1 * T2 + T11;
        }
    }
    if (T5 >= -0.8) Denomi = 1.0 + T5;
    else {
        T9 = 1.0 / (7.0 + 10.0 * T5);
        Denomi = (0.6 + T5) * T9;
        T9 *= T9;
        dDenomi_dVg *= T9; dDenomi_dVd *= T9; dDenomi_dVb *= T9;
    }
    const double ueff = PR(u0temp) / Denomi;
    T9 = -ueff / Denomi;
    const double dueff_dVg = T9 * dDenomi_dVg;
    const double dueff_dVd = T9 * dDenomi_dVd;
    const double dueff_dVb = T9 * dDenomi_dVb;

    /* ---- Vdsat (b4ld.c:1614-1700) */
    const double vsattemp = PR(vsattemp);
    const double WVCox = Weff * vsattemp * coxe;
    const double WVCoxRds = WVCox * Rds;
    double Esat = 2.0 * vsattemp / ueff;
    double EsatL = Esat * Leff;
    T0 = -EsatL / ueff;
    double dEsatL_dVg = T0 * dueff_dVg;
    double dEsatL_dVd = T0 * dueff_dVd;
    double dEsatL_dVb = T0 * dueff_dVb;
    double Lambda, dLambda_dVg;
    const double a1 = PR(a1), a2 = PR(a2);
    if (a1 == 0.0) { Lambda = a2; dLambda_dVg = 0.0; }
    else if (a1 > 0.0) {
        T0 = 1.0 - a2;
        T1 = T0 - a1 * Vgsteff - 0.0001;
        T2 = sqrt(T1 * T1 + 0.0004 * T0);
        Lambda = a2 + T0 - 0.5 * (T1 + T2);
        dLambda_dVg = 0.5 * a1 * (1.0 + T1 / T2);
    } else {
        T1 = a2 + a1 * Vgsteff - 0.0001;
        T2 = sqrt(T1 * T1 + 0.0004 * a2);
        Lambda = 0.5 * (T1 + T2);
        dLambda_dVg = 0.5 * a1 * (1.0 + T1 / T2);
    }
    const double Vgst2Vtm = Vgsteff + 2.0 * Vtm;
    if (Rds > 0) { tmp2 = dRds_dVg / Rds + dWeff_dVg / Weff; tmp3 = dRds_dVb / Rds + dWeff_dVb / Weff; }
    else { tmp2 = dWeff_dVg / Weff; tmp3 = dWeff_dVb / Weff; }
    double Vdsat, dVdsat_dVg, dVdsat_dVd, dVdsat_dVb;
    if ((Rds == 0.0) && (Lambda == 1.0)) {
        T0 = 1.0 / (Abulk * EsatL + Vgst2Vtm);
        tmp1 = 0.0;
        T1 = T0 * T0;
        T2 = Vgst2Vtm * T0;
        T3 = EsatL * Vgst2Vtm;
        Vdsat = T3 * T0;
        dT0_dVg = -(Abulk * dEsatL_dVg + EsatL * dAbulk_dVg + 1.0) * T1;
        dT0_dVd = -(Abulk * dEsatL_dVd) * T1;
        dT0_dVb = -(Abulk * dEsatL_dVb + dAbulk_dVb * EsatL) * T1;
        dVdsat_dVg = T3 * dT0_dVg + T2 * dEsatL_dVg + EsatL * T0;
        dVdsat_dVd = T3 * dT0_dVd + T2 * dEsatL_dVd;
        dVdsat_dVb = T3 * dT0_dVb + T2 * dEsatL_dVb;
    } else {
        tmp1 = dLambda_dVg / (Lambda * Lambda);
        T9 = Abulk * WVCoxRds;
        T8 = Abulk * T9;
        T7 = Vgst2Vtm * T9;
        T6 = Vgst2Vtm * WVCoxRds;
        T0 = 2.0 * Abulk * (T9 - 1.0 + 1.0 / Lambda);
        dT0_dVg = 2.0 * (T8 * tmp2 - Abulk * tmp1 + (2.0 * T9 + 1.0 / Lambda - 1.0) * dAbulk_dVg);
        dT0_dVb = 2.0 * (T8 * (2.0 / Abulk * dAbulk_dVb + tmp3) + (1.0 / Lambda - 1.0) * dAbulk_dVb);
        dT0_dVd = 0.0;
        T1 = Vgst2Vtm * (2.0 / Lambda - 1.0) + Abulk * EsatL + 3.0 * T7;
        dT1_dVg = (2.0 / Lambda - 1.0) - 2.0 * Vgst2Vtm * tmp1 + Abulk * dEsatL_dVg
                + EsatL * dAbulk_dVg + 3.0 * (T9 + T7 * tmp2 + T6 * dAbulk_dVg);
        dT1_dVb = Abulk * dEsatL_dVb + EsatL * dAbulk_dVb + 3.0 * (T6 * dAbulk_dVb + T7 * tmp3);
        dT1_dVd = Abulk * dEsatL_dVd;
        T2 = Vgst2Vtm * (EsatL + 2.0 * T6);
        dT2_dVg = EsatL + Vgst2Vtm * dEsatL_dVg + T6 * (4.0 + 2.0 * Vgst2Vtm * tmp2);
        dT2_dVb = Vgst2Vtm * (dEsatL_dVb + 2.0 * T6 * tmp3);
        dT2_dVd = Vgst2Vtm * dEsatL_dVd;
        T3 = sqrt(T1 * T1 - 2.0 * T0 * T2);
        Vdsat = (T1 - T3) / T0;
        dVdsat_dVg = (dT1_dVg - (T1 * dT1_dVg - dT0_dVg * T2 - T0 * dT2_dVg) / T3 - Vdsat * dT0_dVg) / T0;
        dVdsat_dVb = (dT1_dVb - (T1 * dT1_dVb - dT0_dVb * T2 - T0 * dT2_dVb) / T3 - Vdsat * dT0_dVb) / T0;
        dVdsat_dVd = (dT1_dVd - (T1 * dT1_dVd - T0 * dT2_dVd) / T3) / T0;
    }

    /* ---- Vdseff (b4ld.c:1702-1757) */
    const double delta = PR(delta);
    double Vdseff, dVdseff_dVg, dVdseff_dVd, dVdseff_dVb;
    T1 = Vdsat - Vds - delta;
    dT1_dVg = dVdsat_dVg;
    dT1_dVd = dVdsat_dVd - 1.0;
    dT1_dVb = dVdsat_dVb;
    T2 = sqrt(T1 * T1 + 4.0 * delta * Vdsat);
    T0 = T1 / T2;
    T9 = 2.0 * delta;
    T3 = T9 / T2;
    dT2_dVg = T0 * dT1_dVg + T3 * dVdsat_dVg;
    dT2_dVd = T0 * dT1_dVd + T3 * dVdsat_dVd;
    dT2_dVb = T0 * dT1_dVb + T3 * dVdsat_dVb;
    if (T1 >= 0.0) {
        Vdseff = Vdsat - 0.5 * (T1 + T2);
        dVdseff_dVg = dVdsat_dVg - 0.5 * (dT1_dVg + dT2_dVg);
        dVdseff_dVd = dVdsat_dVd - 0.5 * (dT1_dVd + dT2_dVd);
        dVdseff_dVb = dVdsat_dVb - 0.5 * (dT1_dVb + dT2_dVb);
    } else {
        T4 = T9 / (T2 - T1);
        T5 = 1.0 - T4;
        T6 = Vdsat * T4 / (T2 - T1);
        Vdseff = Vdsat * T5;
        dVdseff_dVg = dVdsat_dVg * T5 + T6 * (dT2_dVg - dT1_dVg);
        dVdseff_dVd = dVdsat_dVd * T5 + T6 * (dT2_dVd - dT1_dVd);
        dVdseff_dVb = dVdsat_dVb * T5 + T6 * (dT2_dVb - dT1_dVb);
    }
    if (Vds == 0.0) { Vdseff = 0.0; dVdseff_dVg = 0.0; dVdseff_dVb = 0.0; }
    if (Vdseff > Vds) Vdseff = Vds;
    const double diffVds = Vds - Vdseff;
    const double litl = PR(litl);

    if (flags & B4FL_LAMBDA) {   /* velocity overshoot (b4ld.c:1759-1808) */
        double dT5_dVg, dT5_dVd, dT5b, dT6d, dT2g, dT2d, dT2b;
        T1 = Leff * ueff;
        T2 = PR(lambda) / T1;
        T3 = -T2 / T1 * Leff;
        dT2d = T3 * dueff_dVd;
        dT2g = T3 * dueff_dVg;
        dT2b = T3 * dueff_dVb;
        T5 = 1.0 / (Esat * litl);
        T4 = -T5 / EsatL;
        dT5_dVg = dEsatL_dVg * T4;
        dT5_dVd = dEsatL_dVd * T4;
        dT5b = dEsatL_dVb * T4;
        T6 = 1.0 + diffVds * T5;
        dT6_dVg = dT5_dVg * diffVds - dVdseff_dVg * T5;
        dT6d = dT5_dVd * diffVds + (1.0 - dVdseff_dVd) * T5;
        dT6_dVb = dT5b * diffVds - dVdseff_dVb * T5;
        T7 = 2.0 / (T6 * T6 + 1.0);
        T8 = 1.0 - T7;
        T9 = T6 * T7 * T7;
        dT8_dVg = T9 * dT6_dVg;
        dT8_dVd = T9 * dT6d;
        dT8_dVb = T9 * dT6_dVb;
        T10 = 1.0 + T2 * T8;
        dT10_dVg = dT2g * T8 + T2 * dT8_dVg;
        dT10_dVd = dT2d * T8 + T2 * dT8_dVd;
        dT10_dVb = dT2b * T8 + T2 * dT8_dVb;
        if (T10 == 1.0) dT10_dVg = dT10_dVd = dT10_dVb = 0.0;
        dEsatL_dVg *= T10; dEsatL_dVg += EsatL * dT10_dVg;
        dEsatL_dVd *= T10; dEsatL_dVd += EsatL * dT10_dVd;
        dEsatL_dVb *= T10; dEsatL_dVb += EsatL * dT10_dVb;
        EsatL *= T10;
        Esat = EsatL / Leff;
    }

    /* ---- Vasat (b4ld.c:1810-1838) */
    tmp4 = 1.0 - 0.5 * Abulk * Vdsat / Vgst2Vtm;
    T9 = WVCoxRds * Vgsteff;
    T8 = T9 / Vgst2Vtm;
    T0 = EsatL + Vdsat + 2.0 * T9 * tmp4;
    T7 = 2.0 * WVCoxRds * tmp4;
    dT0_dVg = dEsatL_dVg + dVdsat_dVg + T7 * (1.0 + tmp2 * Vgsteff)
            - T8 * (Abulk * dVdsat_dVg - Abulk * Vdsat / Vgst2Vtm + Vdsat * dAbulk_dVg);
    dT0_dVb = dEsatL_dVb + dVdsat_dVb + T7 * tmp3 * Vgsteff
            - T8 * (dAbulk_dVb * Vdsat + Abulk * dVdsat_dVb);
    dT0_dVd = dEsatL_dVd + dVdsat_dVd - T8 * Abulk * dVdsat_dVd;
    T9 = WVCoxRds * Abulk;
    T1 = 2.0 / Lambda - 1.0 + T9;
    dT1_dVg = -2.0 * tmp1 + WVCoxRds * (Abulk * tmp2 + dAbulk_dVg);
    dT1_dVb = dAbulk_dVb * WVCoxRds + T9 * tmp3;
    const double Vasat = T0 / T1;
    const double dVasat_dVg = (dT0_dVg - Vasat * dT1_dVg) / T1;
    const double dVasat_dVb = (dT0_dVb - Vasat * dT1_dVb) / T1;
    const double dVasat_dVd = dT0_dVd / T1;

    /* ---- Idl (b4ld.c:1840-1893) */
    tmp1 = PR(vtfbphi2);
    tmp2 = 2.0e8 * PR(toxp);
    dT0_dVg = 1.0 / tmp2;
    T0 = (Vgsteff + tmp1) * dT0_dVg;
    const double bdos = PR(bdos), ados = PR(ados), coxp = PR(coxp);
    tmp3 = exp(bdos * 0.7 * log(T0));
    T1 = 1.0 + tmp3;
    T2 = bdos * 0.7 * tmp3 / T0;
    double Tcen = ados * 1.9e-9 / T1;
    double dTcen_dVg = -Tcen * T2 * dT0_dVg / T1;
    double Coxeff = epssub * coxp / (epssub + coxp * Tcen);
    double dCoxeff_dVg = -Coxeff * Coxeff * dTcen_dVg / epssub;
    const double CoxeffWovL = Coxeff * Weff / Leff;
    const double beta = ueff * CoxeffWovL;
    T3 = ueff / Leff;
    const double dbeta_dVg = CoxeffWovL * dueff_dVg + T3 * (Weff * dCoxeff_dVg + Coxeff * dWeff_dVg);
    const double dbeta_dVd = CoxeffWovL * dueff_dVd;
    const double dbeta_dVb = CoxeffWovL * dueff_dVb + T3 * Coxeff * dWeff_dVb;
    const double AbovVgst2Vtm = Abulk / Vgst2Vtm;
    T0 = 1.0 - 0.5 * Vdseff * AbovVgst2Vtm;
    dT0_dVg = -0.5 * (Abulk * dVdseff_dVg - Abulk * Vdseff / Vgst2Vtm + Vdseff * dAbulk_dVg) / Vgst2Vtm;
    dT0_dVd = -0.5 * Abulk * dVdseff_dVd / Vgst2Vtm;
    dT0_dVb = -0.5 * (Abulk * dVdseff_dVb + dAbulk_dVb * Vdseff) / Vgst2Vtm;
    const double fgche1 = Vgsteff * T0;
    const double dfgche1_dVg = Vgsteff * dT0_dVg + T0;
    const double dfgche1_dVd = Vgsteff * dT0_dVd;
    const double dfgche1_dVb = Vgsteff * dT0_dVb;
    T9 = Vdseff / EsatL;
    const double fgche2 = 1.0 + T9;
    const double dfgche2_dVg = (dVdseff_dVg - T9 * dEsatL_dVg) / EsatL;
    const double dfgche2_dVd = (dVdseff_dVd - T9 * dEsatL_dVd) / EsatL;
    const double dfgche2_dVb = (dVdseff_dVb - T9 * dEsatL_dVb) / EsatL;
    const double gche = beta * fgche1 / fgche2;
    const double dgche_dVg = (beta * dfgche1_dVg + fgche1 * dbeta_dVg - gche * dfgche2_dVg) / fgche2;
    const double dgche_dVd = (beta * dfgche1_dVd + fgche1 * dbeta_dVd - gche * dfgche2_dVd) / fgche2;
    const double dgche_dVb = (beta * dfgche1_dVb + fgche1 * dbeta_dVb - gche * dfgche2_dVb) / fgche2;
    T0 = 1.0 + gche * Rds;
    const double Idl = gche / T0;
    T1 = (1.0 - Idl * Rds) / T0;
    T2 = Idl * Idl;
    const double dIdl_dVg = T1 * dgche_dVg - T2 * dRds_dVg;
    const double dIdl_dVd = T1 * dgche_dVd;
    const double dIdl_dVb = T1 * dgche_dVb - T2 * dRds_dVb;

    /* ---- output resistance (b4ld.c:1895-2044) */
    double FP, dFP_dVg;
    if (PR(fprout) <= 0.0) { FP = 1.0; dFP_dVg = 0.0; }
    else {
        T9 = PR(fprout) * sqrt(Leff) / Vgst2Vtm;
        FP = 1.0 / (1.0 + T9);
        dFP_dVg = FP * FP * T9 / Vgst2Vtm;
    }
    double PvagTerm, dPvagTerm_dVg, dPvagTerm_dVd, dPvagTerm_dVb;
    T8 = PR(pvag) / EsatL;
    T9 = T8 * Vgsteff;
    if (T9 > -0.9) {
        PvagTerm = 1.0 + T9;
        dPvagTerm_dVg = T8 * (1.0 - Vgsteff * dEsatL_dVg / EsatL);
        dPvagTerm_dVb = -T9 * dEsatL_dVb / EsatL;
        dPvagTerm_dVd = -T9 * dEsatL_dVd / EsatL;
    } else {
        T4 = 1.0 / (17.0 + 20.0 * T9);
        PvagTerm = (0.8 + T9) * T4;
        T4 *= T4;
        dPvagTerm_dVg = T8 * (1.0 - Vgsteff * dEsatL_dVg / EsatL) * T4;
        T9 *= T4 / EsatL;
        dPvagTerm_dVb = -T9 * dEsatL_dVb;
        dPvagTerm_dVd = -T9 * dEsatL_dVd;
    }
    double Cclm, dCclm_dVg, dCclm_dVd, dCclm_dVb, VACLM, dVACLM_dVg, dVACLM_dVd, dVACLM_dVb;
    if ((PR(pclm) > MIN_EXP) && (diffVds > 1.0e-10)) {
        T0 = 1.0 + Rds * Idl;
        dT0_dVg = dRds_dVg * Idl + Rds * dIdl_dVg;
        dT0_dVd = Rds * dIdl_dVd;
        dT0_dVb = dRds_dVb * Idl + Rds * dIdl_dVb;
        T2 = Vdsat / Esat;
        T1 = Leff + T2;
        dT1_dVg = (dVdsat_dVg - T2 * dEsatL_dVg / Leff) / Esat;
        dT1_dVd = (dVdsat_dVd - T2 * dEsatL_dVd / Leff) / Esat;
        dT1_dVb = (dVdsat_dVb - T2 * dEsatL_dVb / Leff) / Esat;
        Cclm = FP * PvagTerm * T0 * T1 / (PR(pclm) * litl);
        dCclm_dVg = Cclm * (dFP_dVg / FP + dPvagTerm_dVg / PvagTerm + dT0_dVg / T0 + dT1_dVg / T1);
        dCclm_dVb = Cclm * (dPvagTerm_dVb / PvagTerm + dT0_dVb / T0 + dT1_dVb / T1);
        dCclm_dVd = Cclm * (dPvagTerm_dVd / PvagTerm + dT0_dVd / T0 + dT1_dVd / T1);
        VACLM = Cclm * diffVds;
        dVACLM_dVg = dCclm_dVg * diffVds - dVdseff_dVg * Cclm;
        dVACLM_dVb = dCclm_dVb * diffVds - dVdseff_dVb * Cclm;
        dVACLM_dVd = dCclm_dVd * diffVds + (1.0 - dVdseff_dVd) * Cclm;
    } else {
        VACLM = Cclm = MAX_EXP;
        dVACLM_dVd = dVACLM_dVg = dVACLM_dVb = 0.0;
        dCclm_dVd = dCclm_dVg = dCclm_dVb = 0.0;
    }
    double VADIBL, dVADIBL_dVg, dVADIBL_dVd, dVADIBL_dVb;
    const double thetaRout = PR(thetaRout);
    if (thetaRout > MIN_EXP) {
        T8 = Abulk * Vdsat;
        T0 = Vgst2Vtm * T8;
        dT0_dVg = Vgst2Vtm * Abulk * dVdsat_dVg + T8 + Vgst2Vtm * Vdsat * dAbulk_dVg;
        dT0_dVb = Vgst2Vtm * (dAbulk_dVb * Vdsat + Abulk * dVdsat_dVb);
        dT0_dVd = Vgst2Vtm * Abulk * dVdsat_dVd;
        T1 = Vgst2Vtm + T8;
        dT1_dVg = 1.0 + Abulk * dVdsat_dVg + Vdsat * dAbulk_dVg;
        dT1_dVb = Abulk * dVdsat_dVb + dAbulk_dVb * Vdsat;
        dT1_dVd = Abulk * dVdsat_dVd;
        T9 = T1 * T1;
        T2 = thetaRout;
        VADIBL = (Vgst2Vtm - T0 / T1) / T2;
        dVADIBL_dVg = (1.0 - dT0_dVg / T1 + T0 * dT1_dVg / T9) / T2;
        dVADIBL_dVb = (-dT0_dVb / T1 + T0 * dT1_dVb / T9) / T2;
        dVADIBL_dVd = (-dT0_dVd / T1 + T0 * dT1_dVd / T9) / T2;
        T7 = PR(pdiblb) * Vbseff;
        if (T7 >= -0.9) {
            T3 = 1.0 / (1.0 + T7);
            VADIBL *= T3;
            dVADIBL_dVg *= T3;
            dVADIBL_dVb = (dVADIBL_dVb - VADIBL * PR(pdiblb)) * T3;
            dVADIBL_dVd *= T3;
        } else {
            T4 = 1.0 / (0.8 + T7);
            T3 = (17.0 + 20.0 * T7) * T4;
            dVADIBL_dVg *= T3;
            dVADIBL_dVb = dVADIBL_dVb * T3 - VADIBL * PR(pdiblb) * T4 * T4;
            dVADIBL_dVd *= T3;
            VADIBL *= T3;
        }
        dVADIBL_dVg = dVADIBL_dVg * PvagTerm + VADIBL * dPvagTerm_dVg;
        dVADIBL_dVb = dVADIBL_dVb * PvagTerm + VADIBL * dPvagTerm_dVb;
        dVADIBL_dVd = dVADIBL_dVd * PvagTerm + VADIBL * dPvagTerm_dVd;
        VADIBL *= PvagTerm;
    } else {
        VADIBL = MAX_EXP;
        dVADIBL_dVd = dVADIBL_dVg = dVADIBL_dVb = 0.0;
    }
    const double Va = Vasat + VACLM;
    const double dVa_dVg = dVasat_dVg + dVACLM_dVg;
    const double dVa_dVb = dVasat_dVb + dVACLM_dVb;
    const double dVa_dVd = dVasat_dVd + dVACLM_dVd;
    double VADITS, dVADITS_dVg, dVADITS_dVd;
    T0 = PR(pditsd) * Vds;
    if (T0 > EXP_THRESHOLD) { T1 = MAX_EXP; dT1_dVd = 0; }
    else { T1 = exp(T0); dT1_dVd = T1 * PR(pditsd); }
    if (PR(pdits) > MIN_EXP) {
        T2 = 1.0 + PR(pditsl) * Leff;
        VADITS = (1.0 + T2 * T1) / PR(pdits);
        dVADITS_dVg = VADITS * dFP_dVg;
        dVADITS_dVd = FP * T2 * dT1_dVd / PR(pdits);
        VADITS *= FP;
    } else {
        VADITS = MAX_EXP;
        dVADITS_dVg = dVADITS_dVd = 0;
    }
    double VASCBE, dVASCBE_dVg, dVASCBE_dVd, dVASCBE_dVb;
    const double pscbe1 = PR(pscbe1), pscbe2 = PR(pscbe2);
    if ((pscbe2 > 0.0) && (pscbe1 >= 0.0)) {
        if (diffVds > pscbe1 * litl / EXP_THRESHOLD) {
            T0 = pscbe1 * litl / diffVds;
            VASCBE = Leff * exp(T0) / pscbe2;
            T1 = T0 * VASCBE / diffVds;
            dVASCBE_dVg = T1 * dVdseff_dVg;
            dVASCBE_dVd = -T1 * (1.0 - dVdseff_dVd);
            dVASCBE_dVb = T1 * dVdseff_dVb;
        } else {
            VASCBE = MAX_EXP * Leff / pscbe2;
            dVASCBE_dVg = dVASCBE_dVd = dVASCBE_dVb = 0.0;
        }
    } else {
        VASCBE = MAX_EXP;
        dVASCBE_dVg = dVASCBE_dVd = dVASCBE_dVb = 0.0;
    }

    /* ---- Ids with DIBL, DITS, CLM; Isub; SCBE (b4ld.c:2046-2152) */
    T9 = diffVds / VADIBL;
    T0 = 1.0 + T9;
    double Idsa = Idl * T0;
    double dIdsa_dVg = T0 * dIdl_dVg - Idl * (dVdseff_dVg + T9 * dVADIBL_dVg) / VADIBL;
    double dIdsa_dVd = T0 * dIdl_dVd + Idl * (1.0 - dVdseff_dVd - T9 * dVADIBL_dVd) / VADIBL;
    double dIdsa_dVb = T0 * dIdl_dVb - Idl * (dVdseff_dVb + T9 * dVADIBL_dVb) / VADIBL;
    T9 = diffVds / VADITS;
    T0 = 1.0 + T9;
    dIdsa_dVg = T0 * dIdsa_dVg - Idsa * (dVdseff_dVg + T9 * dVADITS_dVg) / VADITS;
    dIdsa_dVd = T0 * dIdsa_dVd + Idsa * (1.0 - dVdseff_dVd - T9 * dVADITS_dVd) / VADITS;
    dIdsa_dVb = T0 * dIdsa_dVb - Idsa * dVdseff_dVb / VADITS;
    Idsa *= T0;
    T0 = log(Va / Vasat);
    dT0_dVg = dVa_dVg / Va - dVasat_dVg / Vasat;
    dT0_dVb = dVa_dVb / Va - dVasat_dVb / Vasat;
    dT0_dVd = dVa_dVd / Va - dVasat_dVd / Vasat;
    T1 = T0 / Cclm;
    T9 = 1.0 + T1;
    dT9_dVg = (dT0_dVg - T1 * dCclm_dVg) / Cclm;
    dT9_dVb = (dT0_dVb - T1 * dCclm_dVb) / Cclm;
    dT9_dVd = (dT0_dVd - T1 * dCclm_dVd) / Cclm;
    dIdsa_dVg = dIdsa_dVg * T9 + Idsa * dT9_dVg;
    dIdsa_dVb = dIdsa_dVb * T9 + Idsa * dT9_dVb;
    dIdsa_dVd = dIdsa_dVd * T9 + Idsa * dT9_dVd;
    Idsa *= T9;
    double Isub, Gbd, Gbg, Gbb;
    tmp = PR(alpha0) + PR(alpha1) * Leff;
    if ((tmp <= 0.0) || (PR(beta0) <= 0.0)) {
        Isub = Gbd = Gbb = Gbg = 0.0;
    } else {
        T2 = tmp / Leff;
        if (diffVds > PR(beta0) / EXP_THRESHOLD) {
            T0 = -PR(beta0) / diffVds;
            T1 = T2 * diffVds * exp(T0);
            T3 = T1 / diffVds * (T0 - 1.0);
            dT1_dVg = T3 * dVdseff_dVg;
            dT1_dVd = T3 * (dVdseff_dVd - 1.0);
            dT1_dVb = T3 * dVdseff_dVb;
        } else {
            T3 = T2 * MIN_EXP;
            T1 = T3 * diffVds;
            dT1_dVg = -T3 * dVdseff_dVg;
            dT1_dVd = T3 * (1.0 - dVdseff_dVd);
            dT1_dVb = -T3 * dVdseff_dVb;
        }
        T4 = Idsa * Vdseff;
        Isub = T1 * T4;
        Gbg = T1 * (dIdsa_dVg * Vdseff + Idsa * dVdseff_dVg) + T4 * dT1_dVg;
        Gbd = T1 * (dIdsa_dVd * Vdseff + Idsa * dVdseff_dVd) + T4 * dT1_dVd;
        Gbb = T1 * (dIdsa_dVb * Vdseff + Idsa * dVdseff_dVb) + T4 * dT1_dVb;
        Gbd += Gbg * dVgsteff_dVd;
        Gbb += Gbg * dVgsteff_dVb;
        Gbg *= dVgsteff_dVg;
        Gbb *= dVbseff_dVb;
    }
    double csub = Isub, gbbs = Gbb, gbgs = Gbg, gbds = Gbd;
    T9 = diffVds / VASCBE;
    T0 = 1.0 + T9;
    const double Ids = Idsa * T0;
    double Gm = T0 * dIdsa_dVg - Idsa * (dVdseff_dVg + T9 * dVASCBE_dVg) / VASCBE;
    double Gds = T0 * dIdsa_dVd + Idsa * (1.0 - dVdseff_dVd - T9 * dVASCBE_dVd) / VASCBE;
    double Gmb = T0 * dIdsa_dVb - Idsa * (dVdseff_dVb + T9 * dVASCBE_dVb) / VASCBE;
    tmp1 = Gds + Gm * dVgsteff_dVd;
    tmp2 = Gmb + Gm * dVgsteff_dVb;
    tmp3 = Gm;
    Gm = (Ids * dVdseff_dVg + Vdseff * tmp3) * dVgsteff_dVg;
    Gds = Ids * (dVdseff_dVd + dVdseff_dVg * dVgsteff_dVd) + Vdseff * tmp1;
    Gmb = (Ids * (dVdseff_dVb + dVdseff_dVg * dVgsteff_dVb) + Vdseff * tmp2) * dVbseff_dVb;
    double cdrain = Ids * Vdseff;
    if (flags & B4FL_VTL) {   /* b4ld.c:2154-2188 */
        T12 = 1.0 / Leff / CoxeffWovL;
        T11 = T12 / Vgsteff;
        T10 = -T11 / Vgsteff;
        const double vs = cdrain * T11;
        const double dvs_dVg = Gm * T11 + cdrain * T10 * dVgsteff_dVg;
        const double dvs_dVd = Gds * T11 + cdrain * T10 * dVgsteff_dVd;
        const double dvs_dVb = Gmb * T11 + cdrain * T10 * dVgsteff_dVb;
        double Fsevl, dFsevl_dVg, dFsevl_dVd, dFsevl_dVb;
        T0 = 2 * 3;
        T1 = vs / (PR(vtl) * PR(tfactor));
        if (T1 > 0.0) {
            T2 = 1.0 + exp(T0 * log(T1));
            T3 = (T2 - 1.0) * T0 / vs;
            Fsevl = 1.0 / exp(log(T2) / T0);
            T4 = -1.0 / T0 * Fsevl / T2;
            dFsevl_dVg = T4 * (T3 * dvs_dVg);
            dFsevl_dVd = T4 * (T3 * dvs_dVd);
            dFsevl_dVb = T4 * (T3 * dvs_dVb);
        } else {
            Fsevl = 1.0;
            dFsevl_dVg = dFsevl_dVd = dFsevl_dVb = 0.0;
        }
        Gm *= Fsevl; Gm += cdrain * dFsevl_dVg;
        Gmb *= Fsevl; Gmb += cdrain * dFsevl_dVb;
        Gds *= Fsevl; Gds += cdrain * dFsevl_dVd;
        cdrain *= Fsevl;
    }
    double gds = Gds, gm = Gm, gmbs = Gmb;

    /* ---- GIDL / GISL (b4ld.c:2308-2526), mtrlMod 0 */
    double Igidl, Ggidld, Ggidlg, Ggidlb, Igisl, Ggisls, Ggislg, Ggislb;
    T0 = 3.0 * toxe;
    if (!(flags & B4FL_GIDLMOD)) {
        T1 = (vds - vgs_eff - PR(egidl)) / T0;
        if ((PR(agidl) <= 0.0) || (PR(bgidl) <= 0.0) || (T1 <= 0.0) || (PR(cgidl) <= 0.0) || (vbd > 0.0))
            Igidl = Ggidld = Ggidlg = Ggidlb = 0.0;
        else {
            dT1_dVd = 1.0 / T0;
            dT1_dVg = -dvgs_eff_dvg * dT1_dVd;
            T2 = PR(bgidl) / T1;
            if (T2 < 100.0) {
                Igidl = PR(agidl) * PR(weffCJ) * T1 * exp(-T2);
                T3 = Igidl * (1.0 + T2) / T1;
                Ggidld = T3 * dT1_dVd;
                Ggidlg = T3 * dT1_dVg;
            } else {
                Igidl = PR(agidl) * PR(weffCJ) * 3.720075976e-44;
                Ggidld = Igidl * dT1_dVd;
                Ggidlg = Igidl * dT1_dVg;
                Igidl *= T1;
            }
            T4 = vbd * vbd;
            T5 = -vbd * T4;
            T6 = PR(cgidl) + T5;
            T7 = T5 / T6;
            T8 = 3.0 * PR(cgidl) * T4 / T6 / T6;
            Ggidld = Ggidld * T7 + Igidl * T8;
            Ggidlg = Ggidlg * T7;
            Ggidlb = -Igidl * T8;
            Igidl *= T7;
        }
        T1 = (-vds - vgd_eff - PR(egisl)) / T0;
        if ((PR(agisl) <= 0.0) || (PR(bgisl) <= 0.0) || (T1 <= 0.0) || (PR(cgisl) <= 0.0) || (vbs > 0.0))
            Igisl = Ggisls = Ggislg = Ggislb = 0.0;
        else {
            dT1_dVd = 1.0 / T0;
            dT1_dVg = -dvgd_eff_dvg * dT1_dVd;
            T2 = PR(bgisl) / T1;
            if (T2 < 100.0) {
                Igisl = PR(agisl) * PR(weffCJ) * T1 * exp(-T2);
                T3 = Igisl * (1.0 + T2) / T1;
                Ggisls = T3 * dT1_dVd;
                Ggislg = T3 * dT1_dVg;
            } else {
                Igisl = PR(agisl) * PR(weffCJ) * 3.720075976e-44;
                Ggisls = Igisl * dT1_dVd;
                Ggislg = Igisl * dT1_dVg;
                Igisl *= T1;
            }
            T4 = vbs * vbs;
            T5 = -vbs * T4;
            T6 = PR(cgisl) + T5;
            T7 = T5 / T6;
            T8 = 3.0 * PR(cgisl) * T4 / T6 / T6;
            Ggisls = Ggisls * T7 + Igisl * T8;
            Ggislg = Ggislg * T7;
            Ggislb = -Igisl * T8;
            Igisl *= T7;
        }
    } else {
        T1 = (-vds - PR(rgisl) * vgd_eff - PR(egisl)) / T0;
        if ((PR(agisl) <= 0.0) || (PR(bgisl) <= 0.0) || (T1 <= 0.0) || (PR(cgisl) < 0.0))
            Igisl = Ggisls = Ggislg = Ggislb = 0.0;
        else {
            dT1_dVd = 1 / T0;
            dT1_dVg = -PR(rgisl) * dT1_dVd * dvgd_eff_dvg;
            T2 = PR(bgisl) / T1;
            if (T2 < EXPL_THRESHOLD) {
                Igisl = PR(weffCJ) * PR(agisl) * T1 * exp(-T2);
                T3 = Igisl / T1 * (T2 + 1);
                Ggisls = T3 * dT1_dVd;
                Ggislg = T3 * dT1_dVg;
            } else {
                T3 = PR(weffCJ) * PR(agisl) * MIN_EXPL;
                Igisl = T3 * T1;
                Ggisls = T3 * dT1_dVd;
                Ggislg = T3 * dT1_dVg;
            }
            T4 = vbs - PR(fgisl);
            if (T4 > PR(gidlclamp)) T4 = PR(gidlclamp);
            T5 = (T4 == 0) ? EXPL_THRESHOLD : PR(kgisl) / T4;
            if (T5 < EXPL_THRESHOLD) { T6 = exp(T5); Ggislb = -Igisl * T6 * T5 / T4; }
            else { T6 = MAX_EXPL; Ggislb = 0.0; }
            Ggisls *= T6; Ggislg *= T6; Igisl *= T6;
        }
        T1 = (vds - PR(rgidl) * vgs_eff - PR(egidl)) / T0;
        if ((PR(agidl) <= 0.0) || (PR(bgidl) <= 0.0) || (T1 <= 0.0) || (PR(cgidl) < 0.0))
            Igidl = Ggidld = Ggidlg = Ggidlb = 0.0;
        else {
            dT1_dVd = 1 / T0;
            dT1_dVg = -PR(rgidl) * dT1_dVd * dvgs_eff_dvg;
            T2 = PR(bgidl) / T1;
            if (T2 < EXPL_THRESHOLD) {
                Igidl = PR(weffCJ) * PR(agidl) * T1 * exp(-T2);
                T3 = Igidl / T1 * (T2 + 1);
                Ggidld = T3 * dT1_dVd;
                Ggidlg = T3 * dT1_dVg;
            } else {
                T3 = PR(weffCJ) * PR(agidl) * MIN_EXPL;
                Igidl = T3 * T1;
                Ggidld = T3 * dT1_dVd;
                Ggidlg = T3 * dT1_dVg;
            }
            T4 = vbd - PR(fgidl);
            if (T4 > PR(gidlclamp)) T4 = PR(gidlclamp);
            T5 = (T4 == 0) ? EXPL_THRESHOLD : PR(kgidl) / T4;
            if (T5 < EXPL_THRESHOLD) { T6 = exp(T5); Ggidlb = -Igidl * T6 * T5 / T4; }
            else { T6 = MAX_EXPL; Ggidlb = 0.0; }
            Ggidld *= T6; Ggidlg *= T6; Igidl *= T6;
        }
    }

    /* ---- gate tunnelling (b4ld.c:2527-2978) */
    double Igcs = 0, gIgcsg = 0, gIgcsd = 0, gIgcsb = 0, Igcd = 0, gIgcdg = 0, gIgcdd = 0, gIgcdb = 0;
    double Igs = 0, gIgsg = 0, gIgss = 0, Igd = 0, gIgdg = 0, gIgdd = 0;
    double Igb = 0, gIgbg = 0, gIgbd = 0, gIgbb = 0;
    if (igcMod || igbMod) {
        double Vfb = PR(vfbzb), V3, Vfbeff, dVfbeff_dVg, dVfbeff_dVb;
        double Voxacc, dVoxacc_dVg, dVoxacc_dVb, Voxdepinv, dVoxdepinv_dVg, dVoxdepinv_dVd, dVoxdepinv_dVb;
        double Vaux = 0, dVaux_dVg = 0, dVaux_dVd = 0, dVaux_dVb = 0, VxNVt = 0, ExpVxNVt;
        V3 = Vfb - Vgs_eff + Vbseff - DELTA_3;
        if (Vfb <= 0.0) T0 = sqrt(V3 * V3 - 4.0 * DELTA_3 * Vfb);
        else T0 = sqrt(V3 * V3 + 4.0 * DELTA_3 * Vfb);
        T1 = 0.5 * (1.0 + V3 / T0);
        Vfbeff = Vfb - 0.5 * (V3 + T0);
        dVfbeff_dVg = T1 * dVgs_eff_dVg;
        dVfbeff_dVb = -T1;
        Voxacc = Vfb - Vfbeff;
        dVoxacc_dVg = -dVfbeff_dVg;
        dVoxacc_dVb = -dVfbeff_dVb;
        if (Voxacc < 0.0) Voxacc = dVoxacc_dVg = dVoxacc_dVb = 0.0;
        T0 = 0.5 * k1ox;
        T3 = Vgs_eff - Vfbeff - Vbseff - Vgsteff;
        if (k1ox == 0.0)
            Voxdepinv = dVoxdepinv_dVg = dVoxdepinv_dVd = dVoxdepinv_dVb = 0.0;
        else if (T3 < 0.0) {
            Voxdepinv = -T3;
            dVoxdepinv_dVg = -dVgs_eff_dVg + dVfbeff_dVg + dVgsteff_dVg;
            dVoxdepinv_dVd = dVgsteff_dVd;
            dVoxdepinv_dVb = dVfbeff_dVb + 1.0 + dVgsteff_dVb;
        } else {
            T1 = sqrt(T0 * T0 + T3);
            T2 = T0 / T1;
            Voxdepinv = k1ox * (T1 - T0);
            dVoxdepinv_dVg = T2 * (dVgs_eff_dVg - dVfbeff_dVg - dVgsteff_dVg);
            dVoxdepinv_dVd = -T2 * dVgsteff_dVd;
            dVoxdepinv_dVb = -T2 * (dVfbeff_dVb + 1.0 + dVgsteff_dVb);
        }
        Voxdepinv += Vgsteff;
        dVoxdepinv_dVg += dVgsteff_dVg;
        dVoxdepinv_dVd += dVgsteff_dVd;
        dVoxdepinv_dVb += dVgsteff_dVb;
        const double tmpv = (tempMod < 2) ? Vtm : Vtm0;
        if (igcMod) {
            double dT2_dVgi, dT2_dVdi, dT2_dVbi;
            T0 = tmpv * PR(nigc);
            if (igcMod == 1) {
                VxNVt = (Vgs_eff - PR(tvth0)) / T0;
                if (VxNVt > EXP_THRESHOLD) {
                    Vaux = Vgs_eff - PR(tvth0);
                    dVaux_dVg = dVgs_eff_dVg; dVaux_dVd = 0.0; dVaux_dVb = 0.0;
                }
            } else {
                VxNVt = (Vgs_eff - Vth) / T0;
                if (VxNVt > EXP_THRESHOLD) {
                    Vaux = Vgs_eff - Vth;
                    dVaux_dVg = dVgs_eff_dVg; dVaux_dVd = -dVth_dVd; dVaux_dVb = -dVth_dVb;
                }
            }
            if (VxNVt < -EXP_THRESHOLD) {
                Vaux = T0 * log(1.0 + MIN_EXP);
                dVaux_dVg = dVaux_dVd = dVaux_dVb = 0.0;
            } else if ((VxNVt >= -EXP_THRESHOLD) && (VxNVt <= EXP_THRESHOLD)) {
                ExpVxNVt = exp(VxNVt);
                Vaux = T0 * log(1.0 + ExpVxNVt);
                dVaux_dVg = ExpVxNVt / (1.0 + ExpVxNVt);
                if (igcMod == 1) { dVaux_dVd = 0.0; dVaux_dVb = 0.0; }
                else { dVaux_dVd = -dVaux_dVg * dVth_dVd; dVaux_dVb = -dVaux_dVg * dVth_dVb; }
                dVaux_dVg *= dVgs_eff_dVg;
            }
            T2 = Vgs * Vaux;
            dT2_dVgi = Vaux + Vgs * dVaux_dVg;
            dT2_dVdi = Vgs * dVaux_dVd;
            dT2_dVbi = Vgs * dVaux_dVb;
            T11 = PR(Aechvb);
            T12 = PR(Bechvb);
            T3 = PR(aigc) * PR(cigc) - PR(bigc);
            T4 = PR(bigc) * PR(cigc);
            T5 = T12 * (PR(aigc) + T3 * Voxdepinv - T4 * Voxdepinv * Voxdepinv);
            if (T5 > EXP_THRESHOLD) { T6 = MAX_EXP; dT6_dVg = dT6_dVd = dT6_dVb = 0.0; }
            else if (T5 < -EXP_THRESHOLD) { T6 = MIN_EXP; dT6_dVg = dT6_dVd = dT6_dVb = 0.0; }
            else {
                T6 = exp(T5);
                dT6_dVg = T6 * T12 * (T3 - 2.0 * T4 * Voxdepinv);
                dT6_dVd = dT6_dVg * dVoxdepinv_dVd;
                dT6_dVb = dT6_dVg * dVoxdepinv_dVb;
                dT6_dVg *= dVoxdepinv_dVg;
            }
            const double Igc = T11 * T2 * T6;
            const double dIgc_dVg = T11 * (T2 * dT6_dVg + T6 * dT2_dVgi);
            const double dIgc_dVd = T11 * (T2 * dT6_dVd + T6 * dT2_dVdi);
            const double dIgc_dVb = T11 * (T2 * dT6_dVb + T6 * dT2_dVbi);
            double Pigcd, dPigcd_dVg, dPigcd_dVd, dPigcd_dVb;
            if (flags & B4FL_PIGCD_GIVEN) {
                Pigcd = PR(pigcd);
                dPigcd_dVg = dPigcd_dVd = dPigcd_dVb = 0.0;
            } else {
                T11 = -PR(Bechvb);
                T12 = Vgsteff + 1.0e-20;
                T13 = T11 / T12 / T12;
                T14 = -T13 / T12;
                Pigcd = T13 * (1.0 - 0.5 * Vdseff / T12);
                dPigcd_dVg = T14 * (2.0 + 0.5 * (dVdseff_dVg - 3.0 * Vdseff / T12));
                dPigcd_dVd = 0.5 * T14 * dVdseff_dVd;
                dPigcd_dVb = 0.5 * T14 * dVdseff_dVb;
            }
            T7 = -Pigcd * Vdseff;
            dT7_dVg = -Vdseff * dPigcd_dVg - Pigcd * dVdseff_dVg;
            dT7_dVd = -Vdseff * dPigcd_dVd - Pigcd * dVdseff_dVd + dT7_dVg * dVgsteff_dVd;
            dT7_dVb = -Vdseff * dPigcd_dVb - Pigcd * dVdseff_dVb + dT7_dVg * dVgsteff_dVb;
            dT7_dVg *= dVgsteff_dVg;
            T8 = T7 * T7 + 2.0e-4;
            dT8_dVg = 2.0 * T7;
            dT8_dVd = dT8_dVg * dT7_dVd;
            dT8_dVb = dT8_dVg * dT7_dVb;
            dT8_dVg *= dT7_dVg;
            if (T7 > EXP_THRESHOLD) { T9 = MAX_EXP; dT9_dVg = dT9_dVd = dT9_dVb = 0.0; }
            else if (T7 < -EXP_THRESHOLD) { T9 = MIN_EXP; dT9_dVg = dT9_dVd = dT9_dVb = 0.0; }
            else { T9 = exp(T7); dT9_dVg = T9 * dT7_dVg; dT9_dVd = T9 * dT7_dVd; dT9_dVb = T9 * dT7_dVb; }
            T1 = T9 - 1.0 + 1.0e-4;
            T10 = (T1 - T7) / T8;
            dT10_dVg = (dT9_dVg - dT7_dVg - T10 * dT8_dVg) / T8;
            dT10_dVd = (dT9_dVd - dT7_dVd - T10 * dT8_dVd) / T8;
            dT10_dVb = (dT9_dVb - dT7_dVb - T10 * dT8_dVb) / T8;
            Igcs = Igc * T10;
            gIgcsg = dIgc_dVg * T10 + Igc * dT10_dVg;
            gIgcsd = dIgc_dVd * T10 + Igc * dT10_dVd;
            gIgcsb = (dIgc_dVb * T10 + Igc * dT10_dVb) * dVbseff_dVb;
            T1 = T9 - 1.0 - 1.0e-4;
            T10 = (T7 * T9 - T1) / T8;
            dT10_dVg = (dT7_dVg * T9 + (T7 - 1.0) * dT9_dVg - T10 * dT8_dVg) / T8;
            dT10_dVd = (dT7_dVd * T9 + (T7 - 1.0) * dT9_dVd - T10 * dT8_dVd) / T8;
            dT10_dVb = (dT7_dVb * T9 + (T7 - 1.0) * dT9_dVb - T10 * dT8_dVb) / T8;
            Igcd = Igc * T10;
            gIgcdg = dIgc_dVg * T10 + Igc * dT10_dVg;
            gIgcdd = dIgc_dVd * T10 + Igc * dT10_dVd;
            gIgcdb = (dIgc_dVb * T10 + Igc * dT10_dVb) * dVbseff_dVb;

            double vgse, dvgse, vgde, dvgde;
            T0 = vgs - (PR(vfbsd) + PR(vfbsdoff));
            vgse = sqrt(T0 * T0 + 1.0e-4);
            dvgse = T0 / vgse;
            T2 = vgs * vgse;
            double dT2g = vgs * dvgse + vgse;
            T11 = PR(AechvbEdgeS);
            T12 = PR(BechvbEdge);
            T3 = PR(aigs) * PR(cigs) - PR(bigs);
            T4 = PR(bigs) * PR(cigs);
            T5 = T12 * (PR(aigs) + T3 * vgse - T4 * vgse * vgse);
            if (T5 > EXP_THRESHOLD) { T6 = MAX_EXP; dT6_dVg = 0.0; }
            else if (T5 < -EXP_THRESHOLD) { T6 = MIN_EXP; dT6_dVg = 0.0; }
            else { T6 = exp(T5); dT6_dVg = T6 * T12 * (T3 - 2.0 * T4 * vgse) * dvgse; }
            Igs = T11 * T2 * T6;
            gIgsg = T11 * (T2 * dT6_dVg + T6 * dT2g);
            gIgss = -gIgsg;
            T0 = vgd - (PR(vfbsd) + PR(vfbsdoff));
            vgde = sqrt(T0 * T0 + 1.0e-4);
            dvgde = T0 / vgde;
            T2 = vgd * vgde;
            dT2g = vgd * dvgde + vgde;
            T11 = PR(AechvbEdgeD);
            T3 = PR(aigd) * PR(cigd) - PR(bigd);
            T4 = PR(bigd) * PR(cigd);
            T5 = T12 * (PR(aigd) + T3 * vgde - T4 * vgde * vgde);
            if (T5 > EXP_THRESHOLD) { T6 = MAX_EXP; dT6_dVg = 0.0; }
            else if (T5 < -EXP_THRESHOLD) { T6 = MIN_EXP; dT6_dVg = 0.0; }
            else { T6 = exp(T5); dT6_dVg = T6 * T12 * (T3 - 2.0 * T4 * vgde) * dvgde; }
            Igd = T11 * T2 * T6;
            gIgdg = T11 * (T2 * dT6_dVg + T6 * dT2g);
            gIgdd = -gIgdg;
        }
        if (igbMod) {
            double dT2g, dT2d = 0.0, dT2b, dT6g, dT6d = 0.0, dT6b;
            T0 = tmpv * PR(nigbacc);
            T1 = -Vgs_eff + Vbseff + Vfb;
            VxNVt = T1 / T0;
            if (VxNVt > EXP_THRESHOLD) { Vaux = T1; dVaux_dVg = -dVgs_eff_dVg; dVaux_dVb = 1.0; }
            else if (VxNVt < -EXP_THRESHOLD) { Vaux = T0 * log(1.0 + MIN_EXP); dVaux_dVg = dVaux_dVb = 0.0; }
            else {
                ExpVxNVt = exp(VxNVt);
                Vaux = T0 * log(1.0 + ExpVxNVt);
                dVaux_dVb = ExpVxNVt / (1.0 + ExpVxNVt);
                dVaux_dVg = -dVaux_dVb * dVgs_eff_dVg;
            }
            T2 = (Vgs - Vbs) * Vaux;
            dT2g = Vaux + (Vgs - Vbs) * dVaux_dVg;
            dT2b = -Vaux + (Vgs - Vbs) * dVaux_dVb;
            T11 = PR(Tigb);
            T12 = -7.45669e11 * toxe;
            T3 = PR(aigbacc) * PR(cigbacc) - PR(bigbacc);
            T4 = PR(bigbacc) * PR(cigbacc);
            T5 = T12 * (PR(aigbacc) + T3 * Voxacc - T4 * Voxacc * Voxacc);
            if (T5 > EXP_THRESHOLD) { T6 = MAX_EXP; dT6g = dT6b = 0.0; }
            else if (T5 < -EXP_THRESHOLD) { T6 = MIN_EXP; dT6g = dT6b = 0.0; }
            else {
                T6 = exp(T5);
                dT6g = T6 * T12 * (T3 - 2.0 * T4 * Voxacc);
                dT6b = dT6g * dVoxacc_dVb;
                dT6g *= dVoxacc_dVg;
            }
            const double Igbacc = T11 * T2 * T6;
            const double dIgbacc_dVg = T11 * (T2 * dT6g + T6 * dT2g);
            const double dIgbacc_dVb = T11 * (T2 * dT6b + T6 * dT2b);
            T0 = tmpv * PR(nigbinv);
            T1 = Voxdepinv - PR(eigbinv);
            VxNVt = T1 / T0;
            if (VxNVt > EXP_THRESHOLD) {
                Vaux = T1; dVaux_dVg = dVoxdepinv_dVg; dVaux_dVd = dVoxdepinv_dVd; dVaux_dVb = dVoxdepinv_dVb;
            } else if (VxNVt < -EXP_THRESHOLD) {
                Vaux = T0 * log(1.0 + MIN_EXP); dVaux_dVg = dVaux_dVd = dVaux_dVb = 0.0;
            } else {
                ExpVxNVt = exp(VxNVt);
                Vaux = T0 * log(1.0 + ExpVxNVt);
                dVaux_dVg = ExpVxNVt / (1.0 + ExpVxNVt);
                dVaux_dVd = dVaux_dVg * dVoxdepinv_dVd;
                dVaux_dVb = dVaux_dVg * dVoxdepinv_dVb;
                dVaux_dVg *= dVoxdepinv_dVg;
            }
            T2 = (Vgs - Vbs) * Vaux;
            dT2g = Vaux + (Vgs - Vbs) * dVaux_dVg;
            dT2d = (Vgs - Vbs) * dVaux_dVd;
            dT2b = -Vaux + (Vgs - Vbs) * dVaux_dVb;
            T11 *= 0.75610;
            T12 *= 1.31724;
            T3 = PR(aigbinv) * PR(cigbinv) - PR(bigbinv);
            T4 = PR(bigbinv) * PR(cigbinv);
            T5 = T12 * (PR(aigbinv) + T3 * Voxdepinv - T4 * Voxdepinv * Voxdepinv);
            if (T5 > EXP_THRESHOLD) { T6 = MAX_EXP; dT6g = dT6d = dT6b = 0.0; }
            else if (T5 < -EXP_THRESHOLD) { T6 = MIN_EXP; dT6g = dT6d = dT6b = 0.0; }
            else {
                T6 = exp(T5);
                dT6g = T6 * T12 * (T3 - 2.0 * T4 * Voxdepinv);
                dT6d = dT6g * dVoxdepinv_dVd;
                dT6b = dT6g * dVoxdepinv_dVb;
                dT6g *= dVoxdepinv_dVg;
            }
            const double Igbinv = T11 * T2 * T6;
            const double dIgbinv_dVg = T11 * (T2 * dT6g + T6 * dT2g);
            const double dIgbinv_dVd = T11 * (T2 * dT6d + T6 * dT2d);
            const double dIgbinv_dVb = T11 * (T2 * dT6b + T6 * dT2b);
            Igb = Igbinv + Igbacc;
            gIgbg = dIgbinv_dVg + dIgbacc_dVg;
            gIgbd = dIgbinv_dVd;
            gIgbb = (dIgbinv_dVb + dIgbacc_dVb) * dVbseff_dVb;
        }
    }
    if (nf != 1.0) {   /* b4ld.c:2929-2978 */
        cdrain *= nf; gds *= nf; gm *= nf; gmbs *= nf;
        gbbs *= nf; gbgs *= nf; gbds *= nf; csub *= nf;
        Igidl *= nf; Ggidld *= nf; Ggidlg *= nf; Ggidlb *= nf;
        Igisl *= nf; Ggisls *= nf; Ggislg *= nf; Ggislb *= nf;
        Igcs *= nf; gIgcsg *= nf; gIgcsd *= nf; gIgcsb *= nf;
        Igcd *= nf; gIgcdg *= nf; gIgcdd *= nf; gIgcdb *= nf;
        Igs *= nf; gIgsg *= nf; gIgss *= nf;
        Igd *= nf; gIgdg *= nf; gIgdd *= nf;
        Igb *= nf; gIgbg *= nf; gIgbd *= nf; gIgbb *= nf;
    }
    const double ggidls = -(Ggidld + Ggidlg + Ggidlb);
    const double ggisld = -(Ggisls + Ggislg + Ggislb);
    const double gIgbs = -(gIgbg + gIgbd + gIgbb);
    const double gIgcss = -(gIgcsg + gIgcsd + gIgcsb);
    const double gIgcds = -(gIgcdg + gIgcdd + gIgcdb);
    (void)ggidls; (void)ggisld;

    /* ---- C-V, capMod 2 (b4ld.c:3012-3914) */
    const bool ChargeComputationNeeded =
        ((mode & (M_DCTRANCURVE | M_AC | M_TRAN | M_INITSMSIG)) ||
         ((mode & M_TRANOP) && (mode & M_UIC)));
    double qgate = 0.0, qbulk = 0.0, qdrn = 0.0, qsrc = 0.0;
    double cggb = 0, cgsb = 0, cgdb = 0, cdgb = 0, cdsb = 0, cddb = 0, cbgb = 0, cbsb = 0, cbdb = 0;
    if (!((PR(xpart) < 0) || !ChargeComputationNeeded)) {
        const double Vbs_cv_phi = phi - Phis;
        double VbseffCV, dVbseffCV_dVb;
        if (Vbseff < 0.0) { VbseffCV = Vbseff; dVbseffCV_dVb = 1.0; }
        else { VbseffCV = Vbs_cv_phi; dVbseffCV_dVb = -dPhis_dVb; }
        const double weffCV = PR(weffCV), leffCV = PR(leffCV);
        const double CoxWL = coxe * weffCV * leffCV * nf;
        if (!(flags & B4FL_CVCHARGE)) {
            const double noff = n * PR(noff);
            const double dnoff_dVd = PR(noff) * dn_dVd;
            const double dnoff_dVb = PR(noff) * dn_dVb;
            T0 = Vtm * noff;
            const double voffcv = PR(voffcv);
            const double VgstNVt = (Vgst - voffcv) / T0;
            if (VgstNVt > EXP_THRESHOLD) {
                Vgsteff = Vgst - voffcv;
                dVgsteff_dVg = dVgs_eff_dVg;
                dVgsteff_dVd = -dVth_dVd;
                dVgsteff_dVb = -dVth_dVb;
            } else if (VgstNVt < -EXP_THRESHOLD) {
                Vgsteff = T0 * log(1.0 + MIN_EXP);
                dVgsteff_dVg = 0.0;
                dVgsteff_dVd = Vgsteff / noff;
                dVgsteff_dVb = dVgsteff_dVd * dnoff_dVb;
                dVgsteff_dVd *= dnoff_dVd;
            } else {
                const double ExpVgst = exp(VgstNVt);
                Vgsteff = T0 * log(1.0 + ExpVgst);
                dVgsteff_dVg = ExpVgst / (1.0 + ExpVgst);
                dVgsteff_dVd = -dVgsteff_dVg * (dVth_dVd + (Vgst - voffcv) / noff * dnoff_dVd)
                             + Vgsteff / noff * dnoff_dVd;
                dVgsteff_dVb = -dVgsteff_dVg * (dVth_dVb + (Vgst - voffcv) / noff * dnoff_dVb)
                             + Vgsteff / noff * dnoff_dVb;
                dVgsteff_dVg *= dVgs_eff_dVg;
            }
        } else {
            const double mstarcv = PR(mstarcv);
            T0 = n * Vtm;
            T1 = mstarcv * Vgst;
            T2 = T1 / T0;
            if (T2 > EXP_THRESHOLD) {
                T10 = T1;
                dT10_dVg = mstarcv * dVgs_eff_dVg;
                dT10_dVd = -dVth_dVd * mstarcv;
                dT10_dVb = -dVth_dVb * mstarcv;
            } else if (T2 < -EXP_THRESHOLD) {
                T10 = Vtm * log(1.0 + MIN_EXP);
                dT10_dVg = 0.0;
                dT10_dVd = T10 * dn_dVd;
                dT10_dVb = T10 * dn_dVb;
                T10 *= n;
            } else {
                const double ExpVgst = exp(T2);
                T3 = Vtm * log(1.0 + ExpVgst);
                T10 = n * T3;
                dT10_dVg = mstarcv * ExpVgst / (1.0 + ExpVgst);
                dT10_dVb = T3 * dn_dVb - dT10_dVg * (dVth_dVb + Vgst * dn_dVb / n);
                dT10_dVd = T3 * dn_dVd - dT10_dVg * (dVth_dVd + Vgst * dn_dVd / n);
                dT10_dVg *= dVgs_eff_dVg;
            }
            T1 = PR(voffcbncv) - (1.0 - mstarcv) * Vgst;
            T2 = T1 / T0;
            if (T2 < -EXP_THRESHOLD) {
                T3 = coxe * MIN_EXP / cdep0;
                T9 = mstarcv + T3 * n;
                dT9_dVg = 0.0; dT9_dVd = dn_dVd * T3; dT9_dVb = dn_dVb * T3;
            } else if (T2 > EXP_THRESHOLD) {
                T3 = coxe * MAX_EXP / cdep0;
                T9 = mstarcv + T3 * n;
                dT9_dVg = 0.0; dT9_dVd = dn_dVd * T3; dT9_dVb = dn_dVb * T3;
            } else {
                const double ExpVgst = exp(T2);
                T3 = coxe / cdep0;
                T4 = T3 * ExpVgst;
                T5 = T1 * T4 / T0;
                T9 = mstarcv + n * T4;
                dT9_dVg = T3 * (mstarcv - 1.0) * ExpVgst / Vtm;
                dT9_dVb = T4 * dn_dVb - dT9_dVg * dVth_dVb - T5 * dn_dVb;
                dT9_dVd = T4 * dn_dVd - dT9_dVg * dVth_dVd - T5 * dn_dVd;
                dT9_dVg *= dVgs_eff_dVg;
            }
            Vgsteff = T10 / T9;
            T11 = T9 * T9;
            dVgsteff_dVg = (T9 * dT10_dVg - T10 * dT9_dVg) / T11;
            dVgsteff_dVd = (T9 * dT10_dVd - T10 * dT9_dVd) / T11;
            dVgsteff_dVb = (T9 * dT10_dVb - T10 * dT9_dVb) / T11;
        }
        const double vfbzb = PR(vfbzb), toxp = PR(toxp), ldeb = PR(ldeb), acde = PR(acde);
        double V3, V4, Vfbeff, dVfbeff_dVg, dVfbeff_dVb, dTcen_dVd, dTcen_dVb, dCoxeff_dVd, dCoxeff_dVb;
        V3 = vfbzb - Vgs_eff + VbseffCV - DELTA_3;
        if (vfbzb <= 0.0) T0 = sqrt(V3 * V3 - 4.0 * DELTA_3 * vfbzb);
        else T0 = sqrt(V3 * V3 + 4.0 * DELTA_3 * vfbzb);
        T1 = 0.5 * (1.0 + V3 / T0);
        Vfbeff = vfbzb - 0.5 * (V3 + T0);
        dVfbeff_dVg = T1 * dVgs_eff_dVg;
        dVfbeff_dVb = -T1 * dVbseffCV_dVb;
        const double Cox = coxp;
        double Tox = 1.0e8 * toxp;
        T0 = (Vgs_eff - VbseffCV - vfbzb) / Tox;
        dT0_dVg = dVgs_eff_dVg / Tox;
        dT0_dVb = -dVbseffCV_dVb / Tox;
        tmp = T0 * acde;
        if ((-EXP_THRESHOLD < tmp) && (tmp < EXP_THRESHOLD)) {
            Tcen = ldeb * exp(tmp);
            dTcen_dVg = acde * Tcen;
            dTcen_dVb = dTcen_dVg * dT0_dVb;
            dTcen_dVg *= dT0_dVg;
        } else if (tmp <= -EXP_THRESHOLD) {
            Tcen = ldeb * MIN_EXP;
            dTcen_dVg = dTcen_dVb = 0.0;
        } else {
            Tcen = ldeb * MAX_EXP;
            dTcen_dVg = dTcen_dVb = 0.0;
        }
        const double LINK = 1.0e-3 * toxp;
        V3 = ldeb - Tcen - LINK;
        V4 = sqrt(V3 * V3 + 4.0 * LINK * ldeb);
        Tcen = ldeb - 0.5 * (V3 + V4);
        T1 = 0.5 * (1.0 + V3 / V4);
        dTcen_dVg *= T1;
        dTcen_dVb *= T1;
        double Ccen = epssub / Tcen;
        T2 = Cox / (Cox + Ccen);
        Coxeff = T2 * Ccen;
        T3 = -Ccen / Tcen;
        dCoxeff_dVg = T2 * T2 * T3;
        dCoxeff_dVb = dCoxeff_dVg * dTcen_dVb;
        dCoxeff_dVg *= dTcen_dVg;
        double CoxWLcen = CoxWL * Coxeff / coxe;
        const double Qac0 = CoxWLcen * (Vfbeff - vfbzb);
        double QovCox = Qac0 / Coxeff;
        const double dQac0_dVg = CoxWLcen * dVfbeff_dVg + QovCox * dCoxeff_dVg;
        const double dQac0_dVb = CoxWLcen * dVfbeff_dVb + QovCox * dCoxeff_dVb;
        T0 = 0.5 * k1ox;
        T3 = Vgs_eff - Vfbeff - VbseffCV - Vgsteff;
        if (k1ox == 0.0) { T1 = 0.0; T2 = 0.0; }
        else if (T3 < 0.0) { T1 = T0 + T3 / k1ox; T2 = CoxWLcen; }
        else { T1 = sqrt(T0 * T0 + T3); T2 = CoxWLcen * T0 / T1; }
        const double Qsub0 = CoxWLcen * k1ox * (T1 - T0);
        QovCox = Qsub0 / Coxeff;
        const double dQsub0_dVg = T2 * (dVgs_eff_dVg - dVfbeff_dVg - dVgsteff_dVg) + QovCox * dCoxeff_dVg;
        const double dQsub0_dVd = -T2 * dVgsteff_dVd;
        const double dQsub0_dVb = -T2 * (dVfbeff_dVb + dVbseffCV_dVb + dVgsteff_dVb) + QovCox * dCoxeff_dVb;
        double DenomiC;
        if (k1ox <= 0.0) { DenomiC = 0.25 * PR(moin) * Vtm; T0 = 0.5 * sqrtPhi; }
        else { DenomiC = PR(moin) * Vtm * k1ox * k1ox; T0 = k1ox * sqrtPhi; }
        T1 = 2.0 * T0 + Vgsteff;
        const double DeltaPhi = Vtm * log(1.0 + T1 * Vgsteff / DenomiC);
        const double dDeltaPhi_dVg = 2.0 * Vtm * (T1 - T0) / (DenomiC + T1 * Vgsteff);
        T0 = Vgsteff - DeltaPhi - 0.001;
        dT0_dVg = 1.0 - dDeltaPhi_dVg;
        T1 = sqrt(T0 * T0 + Vgsteff * 0.004);
        const double VgDP = 0.5 * (T0 + T1);
        const double dVgDP_dVg = 0.5 * (dT0_dVg + (T0 * dT0_dVg + 0.002) / T1);
        Tox += Tox;
        T0 = (Vgsteff + PR(vtfbphi2)) / Tox;
        tmp = exp(bdos * 0.7 * log(T0));
        T1 = 1.0 + tmp;
        T2 = bdos * 0.7 * tmp / (T0 * Tox);
        Tcen = ados * 1.9e-9 / T1;
        dTcen_dVg = -Tcen * T2 / T1;
        dTcen_dVd = dTcen_dVg * dVgsteff_dVd;
        dTcen_dVb = dTcen_dVg * dVgsteff_dVb;
        dTcen_dVg *= dVgsteff_dVg;
        Ccen = epssub / Tcen;
        T0 = Cox / (Cox + Ccen);
        Coxeff = T0 * Ccen;
        T1 = -Ccen / Tcen;
        dCoxeff_dVg = T0 * T0 * T1;
        dCoxeff_dVd = dCoxeff_dVg * dTcen_dVd;
        dCoxeff_dVb = dCoxeff_dVg * dTcen_dVb;
        dCoxeff_dVg *= dTcen_dVg;
        CoxWLcen = CoxWL * Coxeff / coxe;
        const double AbulkCV = Abulk0_Q * PR(abulkCVfactor);
        const double dAbulkCV_dVb = PR(abulkCVfactor) * dAbulk0_Q_dVb;
        const double VdsatCV = VgDP / AbulkCV;
        double VdseffCV, dVdseffCV_dVg, dVdseffCV_dVd, dVdseffCV_dVb;
        T0 = VdsatCV - Vds - DELTA_4;
        dT0_dVg = dVgDP_dVg / AbulkCV;
        dT0_dVb = -VdsatCV * dAbulkCV_dVb / AbulkCV;
        T1 = sqrt(T0 * T0 + 4.0 * DELTA_4 * VdsatCV);
        dT1_dVg = (T0 + DELTA_4 + DELTA_4) / T1;
        dT1_dVd = -T0 / T1;
        dT1_dVb = dT1_dVg * dT0_dVb;
        dT1_dVg *= dT0_dVg;
        if (T0 >= 0.0) {
            VdseffCV = VdsatCV - 0.5 * (T0 + T1);
            dVdseffCV_dVg = 0.5 * (dT0_dVg - dT1_dVg);
            dVdseffCV_dVd = 0.5 * (1.0 - dT1_dVd);
            dVdseffCV_dVb = 0.5 * (dT0_dVb - dT1_dVb);
        } else {
            T3 = (DELTA_4 + DELTA_4) / (T1 - T0);
            T4 = 1.0 - T3;
            T5 = VdsatCV * T3 / (T1 - T0);
            VdseffCV = VdsatCV * T4;
            dVdseffCV_dVg = dT0_dVg * T4 + T5 * (dT1_dVg - dT0_dVg);
            dVdseffCV_dVd = T5 * (dT1_dVd + 1.0);
            dVdseffCV_dVb = dT0_dVb * (T4 - T5) + T5 * dT1_dVb;
        }
        if (Vds == 0.0) { VdseffCV = 0.0; dVdseffCV_dVg = 0.0; dVdseffCV_dVb = 0.0; }
        T0 = AbulkCV * VdseffCV;
        T1 = VgDP;
        T2 = 12.0 * (T1 - 0.5 * T0 + 1.0e-20);
        T3 = T0 / T2;
        T4 = 1.0 - 12.0 * T3 * T3;
        T5 = AbulkCV * (6.0 * T0 * (4.0 * T1 - T0) / (T2 * T2) - 0.5);
        T6 = T5 * VdseffCV / AbulkCV;
        qgate = CoxWLcen * (T1 - T0 * (0.5 - T3));
        QovCox = qgate / Coxeff;
        double Cgg1 = CoxWLcen * (T4 * dVgDP_dVg + T5 * dVdseffCV_dVg);
        const double Cgd1 = CoxWLcen * T5 * dVdseffCV_dVd + Cgg1 * dVgsteff_dVd + QovCox * dCoxeff_dVd;
        const double Cgb1 = CoxWLcen * (T5 * dVdseffCV_dVb + T6 * dAbulkCV_dVb) + Cgg1 * dVgsteff_dVb + QovCox * dCoxeff_dVb;
        Cgg1 = Cgg1 * dVgsteff_dVg + QovCox * dCoxeff_dVg;
        T7 = 1.0 - AbulkCV;
        T8 = T2 * T2;
        T9 = 12.0 * T7 * T0 * T0 / (T8 * AbulkCV);
        T10 = T9 * dVgDP_dVg;
        T11 = -T7 * T5 / AbulkCV;
        T12 = -(T9 * T1 / AbulkCV + VdseffCV * (0.5 - T0 / T2));
        qbulk = CoxWLcen * T7 * (0.5 * VdseffCV - T0 * VdseffCV / T2);
        QovCox = qbulk / Coxeff;
        double Cbg1 = CoxWLcen * (T10 + T11 * dVdseffCV_dVg);
        const double Cbd1 = CoxWLcen * T11 * dVdseffCV_dVd + Cbg1 * dVgsteff_dVd + QovCox * dCoxeff_dVd;
        const double Cbb1 = CoxWLcen * (T11 * dVdseffCV_dVb + T12 * dAbulkCV_dVb) + Cbg1 * dVgsteff_dVb + QovCox * dCoxeff_dVb;
        Cbg1 = Cbg1 * dVgsteff_dVg + QovCox * dCoxeff_dVg;
        double Csg, Csd, Csb;
        const double xpart = PR(xpart);
        if (xpart > 0.5) {
            qsrc = -CoxWLcen * (T1 / 2.0 + T0 / 4.0 - 0.5 * T0 * T0 / T2);
            QovCox = qsrc / Coxeff;
            T2 += T2;
            T3 = T2 * T2;
            T7 = -(0.25 - 12.0 * T0 * (4.0 * T1 - T0) / T3);
            T4 = -(0.5 + 24.0 * T0 * T0 / T3) * dVgDP_dVg;
            T5 = T7 * AbulkCV;
            T6 = T7 * VdseffCV;
            Csg = CoxWLcen * (T4 + T5 * dVdseffCV_dVg);
            Csd = CoxWLcen * T5 * dVdseffCV_dVd + Csg * dVgsteff_dVd + QovCox * dCoxeff_dVd;
            Csb = CoxWLcen * (T5 * dVdseffCV_dVb + T6 * dAbulkCV_dVb) + Csg * dVgsteff_dVb + QovCox * dCoxeff_dVb;
            Csg = Csg * dVgsteff_dVg + QovCox * dCoxeff_dVg;
        } else if (xpart < 0.5) {
            T2 = T2 / 12.0;
            T3 = 0.5 * CoxWLcen / (T2 * T2);
            T4 = T1 * (2.0 * T0 * T0 / 3.0 + T1 * (T1 - 4.0 * T0 / 3.0)) - 2.0 * T0 * T0 * T0 / 15.0;
            qsrc = -T3 * T4;
            QovCox = qsrc / Coxeff;
            T8 = 4.0 / 3.0 * T1 * (T1 - T0) + 0.4 * T0 * T0;
            T5 = -2.0 * qsrc / T2 - T3 * (T1 * (3.0 * T1 - 8.0 * T0 / 3.0) + 2.0 * T0 * T0 / 3.0);
            T6 = AbulkCV * (qsrc / T2 + T3 * T8);
            T7 = T6 * VdseffCV / AbulkCV;
            Csg = T5 * dVgDP_dVg + T6 * dVdseffCV_dVg;
            Csd = Csg * dVgsteff_dVd + T6 * dVdseffCV_dVd + QovCox * dCoxeff_dVd;
            Csb = Csg * dVgsteff_dVb + T6 * dVdseffCV_dVb + T7 * dAbulkCV_dVb + QovCox * dCoxeff_dVb;
            Csg = Csg * dVgsteff_dVg + QovCox * dCoxeff_dVg;
        } else {
            qsrc = -0.5 * qgate;
            Csg = -0.5 * Cgg1;
            Csd = -0.5 * Cgd1;
            Csb = -0.5 * Cgb1;
        }
        qgate += Qac0 + Qsub0 - qbulk;
        qbulk -= (Qac0 + Qsub0);
        qdrn = -(qgate + qbulk + qsrc);
        const double Cbg = Cbg1 - dQac0_dVg - dQsub0_dVg;
        const double Cbd = Cbd1 - dQsub0_dVd;
        double Cbb = Cbb1 - dQac0_dVb - dQsub0_dVb;
        const double Cgg = Cgg1 - Cbg;
        const double Cgd = Cgd1 - Cbd;
        double Cgb = Cgb1 - Cbb;
        Cgb *= dVbseff_dVb;
        Cbb *= dVbseff_dVb;
        Csb *= dVbseff_dVb;
        cggb = Cgg;
        cgsb = -(Cgg + Cgd + Cgb);
        cgdb = Cgd;
        cdgb = -(Cgg + Cbg + Csg);
        cdsb = (Cgg + Cgd + Cgb + Cbg + Cbd + Cbb + Csg + Csd + Csb);
        cddb = -(Cgd + Cbd + Csd);
        cbgb = Cbg;
        cbsb = -(Cbg + Cbd + Cbb);
        cbdb = Cbd;
    }

    /* ---- junction C-V (b4ld.c:3944-4063) */
    double capbs = 0.0, capbd = 0.0;
    if (ChargeComputationNeeded) {
        double qbs, qbd;
        b4_jcv(vbs_jct, PR(czbs), PR(czbssw), PR(czbsswg), PR(MJS), PR(MJSWS), PR(MJSWGS),
               PR(PhiBS), PR(PhiBSWS), PR(PhiBSWGS), qbs, capbs);
        b4_jcv(vbd_jct, PR(czbd), PR(czbdsw), PR(czbdswg), PR(MJD), PR(MJSWD), PR(MJSWGD),
               PR(PhiBD), PR(PhiBSWD), PR(PhiBSWGD), qbd, capbd);
        st.set(0, B4X_qbs, qbs);
        st.set(0, B4X_qbd, qbd);
    }

    /* ---- convergence check (b4ld.c:4066-4114) */
    int noncon = 0;
    if ((PR(off) == 0.0) || (!(mode & M_INITFIX))) {
        if (Check == 1) noncon = 1;
        else {
            double Idtot2 = (imode >= 0) ? cdrain + csub + Igidl - cbd : cdrain + cbd - Igidl;
            double tol0 = ck.reltol * b4_max(fabs(cdhat), fabs(Idtot2)) + ck.abstol;
            double tol3 = ck.reltol * b4_max(fabs(cgshat), fabs(Igstot)) + ck.abstol;
            double tol4 = ck.reltol * b4_max(fabs(cgdhat), fabs(Igdtot)) + ck.abstol;
            double tol5 = ck.reltol * b4_max(fabs(cgbhat), fabs(Igbtot)) + ck.abstol;
            /* the series-resistance terms (tol1/tol2) are 0 against abstol here */
            if (fabs(cdhat - Idtot2) >= tol0 || 0.0 >= ck.abstol) noncon = 1;
            else if ((fabs(cgshat - Igstot) >= tol3) || (fabs(cgdhat - Igdtot) >= tol4)
                     || (fabs(cgbhat - Igbtot) >= tol5)) noncon = 1;
            else {
                double Ibtot2 = cbs + cbd - Igidl - Igisl - csub;
                double tol6 = ck.reltol * b4_max(fabs(cbhat), fabs(Ibtot2)) + ck.abstol;
                if (fabs(cbhat - Ibtot2) > tol6) noncon = 1;
            }
            (void)Idtot; (void)Ibtot;
        }
    }
    st.set(0, B4X_vds, vds); st.set(0, B4X_vgs, vgs); st.set(0, B4X_vbs, vbs);
    st.set(0, B4X_vbd, vbd); st.set(0, B4X_vges, vgs); st.set(0, B4X_vgms, vgs);
    st.set(0, B4X_vdbs, vbs); st.set(0, B4X_vdbd, vbd); st.set(0, B4X_vsbs, vbs);
    st.set(0, B4X_vses, 0.0); st.set(0, B4X_vdes, vds); st.set(0, B4X_qdef, qdef);

    /* cache for the next iteration */
    SETC(cd, cdrain); SETC(csub, csub); SETC(cbd, cbd); SETC(cbs, cbs); SETC(gbd, gbd);
    SETC(gbs, gbs); SETC(gm, gm); SETC(gds, gds); SETC(gmbs, gmbs); SETC(gbbs, gbbs);
    SETC(gbgs, gbgs); SETC(gbds, gbds); SETC(Igidl, Igidl); SETC(Igisl, Igisl);
    SETC(ggidld, Ggidld); SETC(ggidlg, Ggidlg); SETC(ggidlb, Ggidlb); SETC(ggisls, Ggisls);
    SETC(ggislg, Ggislg); SETC(ggislb, Ggislb); SETC(Igs, Igs); SETC(Igd, Igd); SETC(Igcs, Igcs);
    SETC(Igcd, Igcd); SETC(Igb, Igb); SETC(gIgsg, gIgsg); SETC(gIgdg, gIgdg); SETC(gIgcsg, gIgcsg);
    SETC(gIgcsd, gIgcsd); SETC(gIgcsb, gIgcsb); SETC(gIgcdg, gIgcdg); SETC(gIgcdd, gIgcdd);
    SETC(gIgcdb, gIgcdb); SETC(gIgbg, gIgbg); SETC(gIgbd, gIgbd); SETC(gIgbb, gIgbb);
    SETC(mode, (double)imode);

    /* ---- capacitive Jacobian and charge integration (b4ld.c:4132-4743) */
    double gcggb = 0, gcgdb = 0, gcgsb = 0, gcgbb = 0, gcdgb = 0, gcddb = 0, gcdsb = 0, gcdbb = 0;
    double gcsgb = 0, gcsdb = 0, gcssb = 0, gcsbb = 0, gcbgb = 0, gcbdb = 0, gcbsb = 0, gcbbb = 0;
    double ceqqg = 0, ceqqb = 0, ceqqd = 0;
    if (ChargeComputationNeeded) {
        double cgdo, qgdo, cgso, qgso;
        T0 = vgd + DELTA_1;
        T1 = sqrt(T0 * T0 + 4.0 * DELTA_1);
        T2 = 0.5 * (T0 - T1);
        T3 = PR(weffCV) * PR(cgdl);
        T4 = sqrt(1.0 - 4.0 * T2 / PR(ckappad));
        cgdo = PR(cgdoP) + T3 - T3 * (1.0 - 1.0 / T4) * (0.5 - 0.5 * T0 / T1);
        qgdo = (PR(cgdoP) + T3) * vgd - T3 * (T2 + 0.5 * PR(ckappad) * (T4 - 1.0));
        T0 = vgs + DELTA_1;
        T1 = sqrt(T0 * T0 + 4.0 * DELTA_1);
        T2 = 0.5 * (T0 - T1);
        T3 = PR(weffCV) * PR(cgsl);
        T4 = sqrt(1.0 - 4.0 * T2 / PR(ckappas));
        cgso = PR(cgsoP) + T3 - T3 * (1.0 - 1.0 / T4) * (0.5 - 0.5 * T0 / T1);
        qgso = (PR(cgsoP) + T3) * vgs - T3 * (T2 + 0.5 * PR(ckappas) * (T4 - 1.0));
        if (nf != 1.0) { cgdo *= nf; cgso *= nf; qgdo *= nf; qgso *= nf; }
        const double ag0 = ck.ag[0], cgbo = PR(cgboP);
        double qgb;
        if (imode > 0) {
            qdrn -= qgdo;
            gcggb = (cggb + cgdo + cgso + cgbo) * ag0;
            gcgdb = (cgdb - cgdo) * ag0;
            gcgsb = (cgsb - cgso) * ag0;
            gcgbb = -(gcggb + gcgdb + gcgsb);
            gcdgb = (cdgb - cgdo) * ag0;
            gcsgb = -(cggb + cbgb + cdgb + cgso) * ag0;
            gcbgb = (cbgb - cgbo) * ag0;
            qgb = cgbo * vgb;
            qgate += qgdo + qgso + qgb;
            qbulk -= qgb;
            qsrc = -(qgate + qbulk + qdrn);
            gcddb = (cddb + capbd + cgdo) * ag0;
            gcdsb = cdsb * ag0;
            gcsdb = -(cgdb + cbdb + cddb) * ag0;
            gcssb = (capbs + cgso - (cgsb + cbsb + cdsb)) * ag0;
            gcdbb = -(gcdgb + gcddb + gcdsb);
            gcsbb = -(gcsgb + gcsdb + gcssb);
            gcbdb = (cbdb - capbd) * ag0;
            gcbsb = (cbsb - capbs) * ag0;
            gcbbb = -(gcbdb + gcbgb + gcbsb);
        } else {
            qsrc = qdrn - qgso;
            gcggb = (cggb + cgdo + cgso + cgbo) * ag0;
            gcgdb = (cgsb - cgdo) * ag0;
            gcgsb = (cgdb - cgso) * ag0;
            gcgbb = -(gcggb + gcgdb + gcgsb);
            gcdgb = -(cggb + cbgb + cdgb + cgdo) * ag0;
            gcsgb = (cdgb - cgso) * ag0;
            gcbgb = (cbgb - cgbo) * ag0;
            qgb = cgbo * vgb;
            qgate += qgdo + qgso + qgb;
            qbulk -= qgb;
            qdrn = -(qgate + qbulk + qsrc);
            gcddb = (capbd + cgdo - (cgsb + cbsb + cdsb)) * ag0;
            gcdsb = -(cgdb + cbdb + cddb) * ag0;
            gcsdb = cdsb * ag0;
            gcssb = (cddb + capbs + cgso) * ag0;
            gcdbb = -(gcdgb + gcddb + gcdsb);
            gcsbb = -(gcsgb + gcsdb + gcssb);
            gcbdb = (cbsb - capbd) * ag0;
            gcbsb = (cbdb - capbs) * ag0;
            gcbbb = -(gcbgb + gcbdb + gcbsb);
        }
        st.set(0, B4X_qg, qgate);
        st.set(0, B4X_qd, qdrn - st.s(0, B4X_qbd));
        st.set(0, B4X_qs, qsrc - st.s(0, B4X_qbs));
        st.set(0, B4X_qb, qbulk + st.s(0, B4X_qbd) + st.s(0, B4X_qbs));
        if (!(mode & M_INITSMSIG) && !(mode & M_DCTRANCURVE)) {
            if (mode & M_INITTRAN) {
                st.set(1, B4X_qb, st.s(0, B4X_qb));
                st.set(1, B4X_qg, st.s(0, B4X_qg));
                st.set(1, B4X_qd, st.s(0, B4X_qd));
            }
            b4_integrate(st, ck, B4X_qb);
            b4_integrate(st, ck, B4X_qg);
            b4_integrate(st, ck, B4X_qd);
            ceqqg = st.s(0, B4X_cqg) - gcggb * vgb + gcgdb * vbd + gcgsb * vbs;
            ceqqd = st.s(0, B4X_cqd) - gcdgb * vgb + gcddb * vbd + gcdsb * vbs;
            ceqqb = st.s(0, B4X_cqb) - gcbgb * vgb + gcbdb * vbd + gcbsb * vbs;
            if (mode & M_INITTRAN) {
                st.set(1, B4X_cqb, st.s(0, B4X_cqb));
                st.set(1, B4X_cqg, st.s(0, B4X_cqg));
                st.set(1, B4X_cqd, st.s(0, B4X_cqd));
            }
        } else {
            gcggb = gcgdb = gcgsb = gcgbb = gcdgb = gcddb = gcdsb = gcdbb = 0.0;
            gcsgb = gcsdb = gcssb = gcsbb = gcbgb = gcbdb = gcbsb = gcbbb = 0.0;
        }
    }

    /* ---- equivalent current sources (b4ld.c:4750-4982) */
    double Gmr, Gmbsr, FwdSum, RevSum, ceqdrn, ceqbd, ceqbs;
    double gbbdp, gbbsp, gbdpg, gbdpdp, gbdpb, gbdpsp, gbspg, gbspdp, gbspb, gbspsp;
    double gIstotg = 0, gIstotd = 0, gIstots = 0, gIstotb = 0, Istoteq = 0;
    double gIdtotg = 0, gIdtotd = 0, gIdtots = 0, gIdtotb = 0, Idtoteq = 0;
    double gIbtotg = 0, gIbtotd = 0, gIbtots = 0, gIbtotb = 0, Ibtoteq = 0;
    if (imode >= 0) {
        Gmr = gm; Gmbsr = gmbs;
        FwdSum = Gmr + Gmbsr; RevSum = 0.0;
        ceqdrn = type * (cdrain - gds * vds - Gmr * vgs - Gmbsr * vbs);
        ceqbd = type * (csub + Igidl - (gbds + Ggidld) * vds - (gbgs + Ggidlg) * vgs - (gbbs + Ggidlb) * vbs);
        ceqbs = type * (Igisl + Ggisls * vds - Ggislg * vgd - Ggislb * vbd);
        gbbdp = -(gbds);
        gbbsp = gbds + gbgs + gbbs;
        gbdpg = gbgs; gbdpdp = gbds; gbdpb = gbbs;
        gbdpsp = -(gbdpg + gbdpdp + gbdpb);
        gbspg = 0.0; gbspdp = 0.0; gbspb = 0.0; gbspsp = 0.0;
        if (igcMod) {
            gIstotg = gIgsg + gIgcsg; gIstotd = gIgcsd; gIstots = gIgss + gIgcss; gIstotb = gIgcsb;
            Istoteq = type * (Igs + Igcs - gIstotg * vgs - gIgcsd * vds - gIgcsb * vbs);
            gIdtotg = gIgdg + gIgcdg; gIdtotd = gIgdd + gIgcdd; gIdtots = gIgcds; gIdtotb = gIgcdb;
            Idtoteq = type * (Igd + Igcd - gIgdg * vgd - gIgcdg * vgs - gIgcdd * vds - gIgcdb * vbs);
        }
        if (igbMod) {
            gIbtotg = gIgbg; gIbtotd = gIgbd; gIbtots = gIgbs; gIbtotb = gIgbb;
            Ibtoteq = type * (Igb - gIgbg * vgs - gIgbd * vds - gIgbb * vbs);
        }
    } else {
        Gmr = -gm; Gmbsr = -gmbs;
        FwdSum = 0.0; RevSum = -(Gmr + Gmbsr);
        ceqdrn = -type * (cdrain + gds * vds + Gmr * vgd + Gmbsr * vbd);
        ceqbs = type * (csub + Igisl + (gbds + Ggisls) * vds - (gbgs + Ggislg) * vgd - (gbbs + Ggislb) * vbd);
        ceqbd = type * (Igidl - Ggidld * vds - Ggidlg * vgs - Ggidlb * vbs);
        gbbsp = -(gbds);
        gbbdp = gbds + gbgs + gbbs;
        gbdpg = 0.0; gbdpsp = 0.0; gbdpb = 0.0; gbdpdp = 0.0;
        gbspg = gbgs; gbspsp = gbds; gbspb = gbbs;
        gbspdp = -(gbspg + gbspsp + gbspb);
        if (igcMod) {
            gIstotg = gIgsg + gIgcdg; gIstotd = gIgcds; gIstots = gIgss + gIgcdd; gIstotb = gIgcdb;
            Istoteq = type * (Igs + Igcd - gIgsg * vgs - gIgcdg * vgd + gIgcdd * vds - gIgcdb * vbd);
            gIdtotg = gIgdg + gIgcsg; gIdtotd = gIgdd + gIgcss; gIdtots = gIgcsd; gIdtotb = gIgcsb;
            Idtoteq = type * (Igd + Igcs - (gIgdg + gIgcsg) * vgd + gIgcsd * vds - gIgcsb * vbd);
        }
        if (igbMod) {
            gIbtotg = gIgbg; gIbtotd = gIgbs; gIbtots = gIgbd; gIbtotb = gIgbb;
            Ibtoteq = type * (Igb - gIgbg * vgd + gIgbd * vds - gIgbb * vbd);
        }
    }
    double gIgtotg = 0, gIgtotd = 0, gIgtots = 0, gIgtotb = 0, Igtoteq = 0;
    if (igcMod || igbMod) {
        gIgtotg = gIstotg + gIdtotg + gIbtotg;
        gIgtotd = gIstotd + gIdtotd + gIbtotd;
        gIgtots = gIstots + gIdtots + gIbtots;
        gIgtotb = gIstotb + gIdtotb + gIbtotb;
        Igtoteq = Istoteq + Idtoteq + Ibtoteq;
    }
    double ceqjs, ceqjd;
    if (type > 0) {
        ceqjs = (cbs - gbs * vbs_jct);
        ceqjd = (cbd - gbd * vbd_jct);
    } else {
        ceqjs = -(cbs - gbs * vbs_jct);
        ceqjd = -(cbd - gbd * vbd_jct);
        ceqqg = -ceqqg;
        ceqqd = -ceqqd;
        ceqqb = -ceqqb;
    }
    const double mult_i = PR(mult_i), mult_q = PR(mult_q);

    /* ---- RHS (b4ld.c:4992-5022 OpenMP form, applied as 5446-5472) */
    rhs[B4R_D] = (mult_i * (ceqjd - ceqbd - ceqdrn + Idtoteq) - mult_q * ceqqd);
    rhs[B4R_G] = -(mult_q * ceqqg - mult_i * (-Igtoteq));
    rhs[B4R_B] = (mult_i * (ceqbd + ceqbs - ceqjd - ceqjs + Ibtoteq) - mult_q * ceqqb);
    rhs[B4R_S] = (mult_i * (ceqdrn - ceqbs + ceqjs + Istoteq) + mult_q * (ceqqg + ceqqb + ceqqd));

    /* ---- Jacobian (b4ld.c:5060-5235, gidl/gisl 5561-5578) */
    const double gjbd = gbd, gjbs = gbs;
    double *J = stamp;
    J[B4S_GG] = mult_q * gcggb + mult_i * gIgtotg;
    J[B4S_GD] = mult_q * gcgdb + mult_i * gIgtotd;
    J[B4S_GS] = mult_q * gcgsb + mult_i * gIgtots;
    J[B4S_GB] = mult_q * gcgbb + mult_i * gIgtotb;
    J[B4S_DD] = mult_i * (gds + gbd + RevSum + gbdpdp - gIdtotd) + mult_q * gcddb;
    J[B4S_DG] = mult_i * (Gmr + gbdpg - gIdtotg) + mult_q * gcdgb;
    J[B4S_DS] = -(mult_i * (gds + gIdtots + FwdSum - gbdpsp) - mult_q * gcdsb);
    J[B4S_DB] = -(mult_i * (gjbd - Gmbsr - gbdpb + gIdtotb) - mult_q * gcdbb);
    J[B4S_SD] = -(mult_i * (gds + RevSum - gbspdp + gIstotd) - mult_q * gcsdb);
    J[B4S_SG] = mult_q * gcsgb + mult_i * (gbspg - Gmr - gIstotg);
    J[B4S_SS] = mult_i * (gds + gbs - gIstots + FwdSum + gbspsp) + mult_q * gcssb;
    J[B4S_SB] = -(mult_i * (gjbs + Gmbsr - gbspb + gIstotb) - mult_q * gcsbb);
    J[B4S_BD] = mult_q * gcbdb - mult_i * (gjbd - gbbdp + gIbtotd);
    J[B4S_BG] = mult_q * gcbgb - mult_i * (gbgs + gIbtotg);
    J[B4S_BS] = mult_q * gcbsb - mult_i * (gjbs - gbbsp + gIbtots);
    J[B4S_BB] = mult_i * (gjbd + gjbs - gbbs - gIbtotb) + mult_q * gcbbb;
    J[B4S_DD] += mult_i * Ggidld;
    J[B4S_DG] += mult_i * Ggidlg;
    J[B4S_DS] -= mult_i * (Ggidlg + Ggidld + Ggidlb);
    J[B4S_DB] += mult_i * Ggidlb;
    J[B4S_BD] -= mult_i * Ggidld;
    J[B4S_BG] -= mult_i * Ggidlg;
    J[B4S_BS] += mult_i * (Ggidlg + Ggidld + Ggidlb);
    J[B4S_BB] -= mult_i * Ggidlb;
    J[B4S_SD] -= mult_i * (Ggisls + Ggislg + Ggislb);
    J[B4S_SG] += mult_i * Ggislg;
    J[B4S_SS] += mult_i * Ggisls;
    J[B4S_SB] += mult_i * Ggislb;
    J[B4S_BD] += mult_i * (Ggislg + Ggisls + Ggislb);
    J[B4S_BG] -= mult_i * Ggislg;
    J[B4S_BS] -= mult_i * Ggisls;
    J[B4S_BB] -= mult_i * Ggislb;
    (void)Tcen; (void)dTcen_dVg; (void)dCoxeff_dVg; (void)beta;
    return noncon;
#undef PR
#undef CA
#undef SETC
}

}  // namespace ngmi

#endif
