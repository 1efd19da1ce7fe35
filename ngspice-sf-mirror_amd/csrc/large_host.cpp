/*
 * Transient of one circuit too large for the per-circuit wavefront programs
 * (configuration 3's power-grid mesh: ~1.1M unknowns): the reference's
 * DCtran (dctran.c:66-971) and NIiter (niiter.c:28-367) driven from the
 * host, with every vector operation on the GPU:
 *
 *   CKTload (cktload.c:31-180)    BSIM4 evaluation of every MOSFET
 *                                 (k_bsim4_eval), CAPload + NIintegrate and
 *                                 VSRCload as element kernels, then the
 *                                 matrix and right-hand side assembled by
 *                                 signed gathers in a fixed element order;
 *   SMPluFac / SMPsolve           the whole-GPU multifrontal LU
 *                                 (ngmi_splu_factor / ngmi_splu_solve);
 *   NIconvTest (niconv.c:21-90)   one grid kernel and a flag;
 *   CKTtrunc / CKTterr            one grid min-reduction over every charge.
 *
 * The host keeps only scalars: the NIiter mode machine, DCtran's step
 * control and the breakpoint table (CKTsetBreak / CKTclrBreak, with one
 * pending PULSE corner per source, vsrcacct.c:20-160).  Each Newton
 * iteration costs one device-to-host read (the convergence flag); the LU
 * refactor dominates (about 11 ms at 1.1M unknowns).
 *
 * The operating point is CKTop's direct Newton (INITJCT -> INITFIX ->
 * INITFLOAT); the gmin / source stepping fallbacks of the batched engine are
 * not repeated here, a failing operating point returns its error.  .ic
 * values and Gear orders above 2 are not supported (NGMI_E_UNSUPP).
 */
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>
#include "ngmi_internal.h"
#include "bsim4_eval.h"
#include "circuit_types.h"
#include "large_types.h"

namespace ngmi {
const B4Tab *set_table(const ngmi_bsim4_set *s);
int launch_bsim4_eval(const B4Tab &tab, const B4Ckt &ck, const double *vterm, double *const state[3],
                      double *cache, double *stamp, double *rhs, int *noncon, hipStream_t stream);
int lg_launch_vterm(const LgDev &D, const double *x, hipStream_t s);
int lg_launch_sources(const LgDev &D, const LgStep &S, const double *x, hipStream_t s);
int lg_launch_gather(int64_t nrows, const int64_t *ptr, const int64_t *ent, const double *src,
                     const double *base, double *out, hipStream_t s);
int lg_launch_conv(const LgDev &D, const double *xn, const double *xo, const int *votes, int *flag,
                   hipStream_t s);
int lg_launch_terr(const LgDev &D, const LgStep &S, unsigned long long *out, hipStream_t s);
int lg_launch_pick(int n, const int *idx, const double *x, double *out, hipStream_t s);
int lg_launch_kcl(int64_t n, int64_t nnodes, const int64_t *Ap, const int *Ai, const double *Ax, const double *x,
                  const double *b, double *y, double *sabs, unsigned long long *out, hipStream_t s);
}  // namespace ngmi

using namespace ngmi;

namespace {

/* equality.c:22-60 */
bool almost_equal_ulps(double A, double B, int maxUlps)
{
    if (A == B) return true;
    int64_t a, b;
    std::memcpy(&a, &A, 8);
    std::memcpy(&b, &B, 8);
    if (a < 0) a = INT64_MIN - a;
    if (b < 0) b = INT64_MIN - b;
    return std::llabs(a - b) <= maxUlps;
}

struct Src {
    int p, n, br;
    bool pulse;
    int npulse;
    double dc, c[7];
    double break_time;
};

}  // namespace

struct ngmi_large {
    int64_t n = 0, NN = 0, NM = 0, NCAP = 0, NRES = 0, NVS = 0;
    ngmi_bsim4_set *set = nullptr;
    std::vector<Src> vs;
    /* options */
    double reltol, abstol, vntol, chgtol, trtol, gmin;
    int method, maxord;
    /* pattern and plans (host copies kept for the info call) */
    std::vector<int64_t> Ap;
    std::vector<int> Ai;
    ngmi_splu *lu = nullptr;
    /* device */
    LgDev D{};
    std::vector<void *> allocs;
    double *x[2] = {nullptr, nullptr};   /* x[xnew], x[1 - xnew]: CKTrhs, CKTrhsOld */
    int xnew = 0;
    double *Ax = nullptr, *ax_base = nullptr;
    int64_t *ax_ptr = nullptr, *ax_ent = nullptr, *rhs_ptr = nullptr, *rhs_ent = nullptr;
    double *cs[LG_NLEV] = {}, *ms[LG_NLEV] = {};
    double *cache = nullptr;
    int *votes = nullptr, *flag = nullptr, *h_flag = nullptr;
    unsigned long long *terr = nullptr, *h_terr = nullptr;
    LgPulse *d_pulse = nullptr;
    /* probes and the recorded waveform */
    std::vector<int> probe;
    int *d_probe = nullptr;
    double *d_pick = nullptr, *h_pick = nullptr;
    int64_t max_points = 0;
    std::vector<double> wt, wx;
    /* run state */
    hipStream_t s = nullptr;
    LgStep S{};
    int64_t accepted = 0, rejected = 0, iters = 0;
    double lmax = 0.0;   /* the largest |l| multiplier of the run's factorizations */
    double tstep_run = 0.0, finalTime = 0.0, minBreak = 0.0, time = 0.0;
    std::vector<double> breaks;
    double t_load = 0, t_factor = 0, t_solve = 0, t_other = 0;

    ~ngmi_large()
    {
        if (set) ngmi_bsim4_set_free(set);
        ngmi_splu_free(lu);
        for (void *p : allocs) (void)hipFree(p);
        if (h_flag) (void)hipHostFree(h_flag);
        if (h_terr) (void)hipHostFree(h_terr);
        if (h_pick) (void)hipHostFree(h_pick);
    }

    template <class T>
    int dalloc(T **p, size_t count)
    {
        void *q = nullptr;
        NGMI_HIP(hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T)));
        allocs.push_back(q);
        *p = (T *)q;
        return NGMI_OK;
    }
    template <class T>
    int upload(T **p, const std::vector<T> &v)
    {
        int rc = dalloc(p, v.size());
        if (rc) return rc;
        if (!v.empty()) NGMI_HIP(hipMemcpy(*p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
        return NGMI_OK;
    }

    /* cktsetbk.c:19-100 */
    int set_break(double t)
    {
        if (almost_equal_ulps(t, time, 3)) return NGMI_OK;
        if (time > t) return NGMI_E_PANIC;
        for (size_t i = 0; i < breaks.size(); i++) {
            if (breaks[i] > t) {
                if (breaks[i] - t <= minBreak) {
                    breaks[i] = t;
                    return NGMI_OK;
                }
                if (i > 0 && t - breaks[i - 1] <= minBreak) return NGMI_OK;
                breaks.insert(breaks.begin() + (int64_t)i, t);
                return NGMI_OK;
            }
        }
        if (!breaks.empty() && t - breaks.back() <= minBreak) return NGMI_OK;
        breaks.push_back(t);
        return NGMI_OK;
    }
    /* cktclrbk.c:15-40 */
    void clr_break()
    {
        if (breaks.size() > 2) {
            breaks.erase(breaks.begin());
        } else {
            breaks[0] = breaks[1];
            breaks[1] = finalTime;
        }
    }
    /* vsrcload.c:96-125 defaults */
    LgPulse pulse_of(const Src &v) const
    {
        LgPulse P{};
        P.pulse = v.pulse ? 1 : 0;
        P.dc = v.dc;
        if (!v.pulse) return P;
        P.v1 = v.c[0];
        P.v2 = v.c[1];
        P.td = v.npulse > 2 ? v.c[2] : 0.0;
        P.tr = v.npulse > 3 && v.c[3] > 0.0 ? v.c[3] : tstep_run;
        P.tf = v.npulse > 4 && v.c[4] > 0.0 ? v.c[4] : tstep_run;
        P.pw = v.npulse > 5 && v.c[5] >= 0.0 ? v.c[5] : 0.0;
        P.per = v.npulse > 6 && v.c[6] > 0.0 ? v.c[6] : P.tr + P.tf + P.pw;
        return P;
    }
    /* vsrcacct.c:20-160 for PULSE: the next corner of each source */
    int vsrc_accept(int mode)
    {
        if (!(mode & (NGMI_MODETRAN | NGMI_MODETRANOP))) return NGMI_OK;
        for (Src &v : vs) {
            if (!v.pulse) continue;
            const LgPulse P = pulse_of(v);
            double t = time - P.td;
            if (time >= v.break_time) {
                if (t >= P.per) t -= P.per * std::floor(t / P.per);
                const double atime = t + minBreak;
                double wait;
                if (atime < 0.0) wait = -t;
                else if (atime < P.tr) wait = P.tr - t;
                else if (atime < P.tr + P.pw) wait = P.tr + P.pw - t;
                else if (atime < P.tr + P.pw + P.tf) wait = P.tr + P.pw + P.tf - t;
                else wait = P.per - t;
                v.break_time = time + wait;
                if (set_break(v.break_time)) return NGMI_E_PANIC;
                v.break_time -= minBreak;
            }
        }
        return NGMI_OK;
    }
    /* nicomcof.c:17-130 for trapezoidal and Gear up to order 2 */
    void com_cof()
    {
        for (int k = 0; k < 7; k++) S.ag[k] = 0.0;
        if (S.method == 1) {
            if (S.order == 1) {
                S.ag[0] = 1 / S.delta;
                S.ag[1] = -1 / S.delta;
            } else {
                const double xmu = 0.5;
                S.ag[0] = 1.0 / S.delta / (1.0 - xmu);
                S.ag[1] = xmu / (1.0 - xmu);
            }
            return;
        }
        double mat[3][3], arg = 0, arg1;
        S.ag[1] = -1 / S.delta;
        const int o = S.order;
        for (int i = 0; i <= o; i++) mat[0][i] = 1;
        for (int i = 1; i <= o; i++) mat[i][0] = 0;
        for (int i = 1; i <= o; i++) {
            arg += S.deltaOld[i - 1];
            arg1 = 1;
            for (int j = 1; j <= o; j++) {
                arg1 *= arg / S.delta;
                mat[j][i] = arg1;
            }
        }
        for (int i = 1; i <= o; i++)
            for (int j = i + 1; j <= o; j++) {
                mat[j][i] /= mat[i][i];
                for (int k = i + 1; k <= o; k++) mat[j][k] -= mat[j][i] * mat[i][k];
            }
        for (int i = 1; i <= o; i++)
            for (int j = i + 1; j <= o; j++) S.ag[j] = S.ag[j] - mat[j][i] * S.ag[i];
        S.ag[o] /= mat[o][o];
        for (int i = o - 1; i >= 0; i--) {
            for (int j = i + 1; j <= o; j++) S.ag[i] = S.ag[i] - mat[i][j] * S.ag[j];
            S.ag[i] /= mat[i][i];
        }
    }

    /* CKTload into Ax and CKTrhs (= x[xnew][1..n]) from CKTrhsOld */
    int load()
    {
        const double *xold = x[1 - xnew];
        int rc;
        for (int l = 0; l < LG_NLEV; l++) {
            S.cs[l] = cs[l];
            S.ms[l] = ms[l];
        }
        if (NM) {
            if ((rc = lg_launch_vterm(D, xold, s))) return rc;
            B4Ckt ck;
            ck.mode = S.mode;
            ck.order = S.order;
            ck.method = S.method;
            for (int k = 0; k < 7; k++) ck.ag[k] = S.ag[k];
            ck.xfact = S.deltaOld[1] != 0.0 ? S.delta / S.deltaOld[1] : 0.0;
            ck.gmin = gmin;
            ck.reltol = reltol;
            ck.abstol = abstol;
            double *st[3] = {ms[0], ms[1], ms[2]};
            if ((rc = launch_bsim4_eval(*set_table(set), ck, D.vterm, st, cache, D.gsrc + NCAP, D.rsrc + NCAP, votes,
                                        s)))
                return rc;
        }
        if ((rc = lg_launch_sources(D, S, xold, s))) return rc;
        if ((rc = lg_launch_gather(Ap[n], ax_ptr, ax_ent, D.gsrc, ax_base, Ax, s))) return rc;
        return lg_launch_gather(n, rhs_ptr, rhs_ent, D.rsrc, nullptr, x[xnew] + 1, s);
    }

    /* niiter.c:28-367 without nodesets or damping */
    int niiter(int maxIter)
    {
        using clk = std::chrono::steady_clock;
        if (maxIter < 100) maxIter = 100;
        int iterno = 0;
        for (;;) {
            auto t0 = clk::now();
            int rc = load();
            if (rc) return rc;
            NGMI_HIP(hipStreamSynchronize(s));
            auto t1 = clk::now();
            iterno++;
            rc = ngmi_splu_factor(lu, Ax, s);   /* synchronizes */
            auto t2 = clk::now();
            iters++;
            if (rc) return rc;
            double lm = 0.0;
            if (ngmi_splu_multiplier(lu, &lm) == NGMI_OK) lmax = std::max(lmax, lm);
            if ((rc = ngmi_splu_solve(lu, x[xnew] + 1, s))) return rc;
            NGMI_HIP(hipStreamSynchronize(s));
            auto t3 = clk::now();
            int noncon = 1;
            if (iterno > maxIter) return NGMI_E_ITERLIM;
            if (iterno != 1) {
                NGMI_HIP(hipMemsetAsync(flag, 0, sizeof(int), s));
                if ((rc = lg_launch_conv(D, x[xnew], x[1 - xnew], votes, flag, s))) return rc;
                NGMI_HIP(hipMemcpyAsync(h_flag, flag, sizeof(int), hipMemcpyDeviceToHost, s));
                NGMI_HIP(hipStreamSynchronize(s));
                noncon = *h_flag;
            }
            auto t4 = clk::now();
            t_load += std::chrono::duration<double, std::milli>(t1 - t0).count();
            t_factor += std::chrono::duration<double, std::milli>(t2 - t1).count();
            t_solve += std::chrono::duration<double, std::milli>(t3 - t2).count();
            t_other += std::chrono::duration<double, std::milli>(t4 - t3).count();
            const int m = S.mode;
            const int keep = m & ~0x3f00;
            if (m & NGMI_MODEINITFLOAT) {
                if (noncon == 0) return NGMI_OK;
            } else if (m & NGMI_MODEINITJCT) {
                S.mode = keep | NGMI_MODEINITFIX;
            } else if (m & NGMI_MODEINITFIX) {
                if (noncon == 0) S.mode = keep | NGMI_MODEINITFLOAT;
            } else if (m & (NGMI_MODEINITSMSIG | NGMI_MODEINITTRAN | NGMI_MODEINITPRED)) {
                S.mode = keep | NGMI_MODEINITFLOAT;
            } else {
                return NGMI_E_PANIC;
            }
            xnew = 1 - xnew;   /* CKTrhs <-> CKTrhsOld */
        }
    }

    /* ckttrunc.c:20-56 */
    int trunc(double *timeStep)
    {
        for (int l = 0; l < LG_NLEV; l++) {
            S.cs[l] = cs[l];
            S.ms[l] = ms[l];
        }
        const double big = 1e308;
        unsigned long long init;
        std::memcpy(&init, &big, 8);
        NGMI_HIP(hipMemcpyAsync(terr, &init, 8, hipMemcpyHostToDevice, s));
        int rc = lg_launch_terr(D, S, terr, s);
        if (rc) return rc;
        NGMI_HIP(hipMemcpyAsync(h_terr, terr, 8, hipMemcpyDeviceToHost, s));
        NGMI_HIP(hipStreamSynchronize(s));
        double timetemp;
        std::memcpy(&timetemp, h_terr, 8);
        *timeStep = std::fmin(2 * *timeStep, timetemp);
        return NGMI_OK;
    }

    void rotate_states()
    {
        double *tc = cs[maxord + 1], *tm = ms[maxord + 1];
        for (int i = maxord; i >= 0; i--) {
            cs[i + 1] = cs[i];
            ms[i + 1] = ms[i];
        }
        cs[0] = tc;
        ms[0] = tm;
    }
    int copy_level(int from, int to)
    {
        if (NCAP) NGMI_HIP(hipMemcpyAsync(cs[to], cs[from], (size_t)2 * NCAP * 8, hipMemcpyDeviceToDevice, s));
        if (NM) NGMI_HIP(hipMemcpyAsync(ms[to], ms[from], (size_t)B4X__COUNT * NM * 8, hipMemcpyDeviceToDevice, s));
        return NGMI_OK;
    }

    /* CKTdump (cktdump.c:33-44): the probed unknowns of CKTrhsOld */
    int dump()
    {
        if ((int64_t)wt.size() >= max_points) return NGMI_OK;
        const int np = (int)probe.size();
        int rc = lg_launch_pick(np, d_probe, x[1 - xnew], d_pick, s);
        if (rc) return rc;
        NGMI_HIP(hipMemcpyAsync(h_pick, d_pick, (size_t)std::max(np, 1) * 8, hipMemcpyDeviceToHost, s));
        NGMI_HIP(hipStreamSynchronize(s));
        wt.push_back(time);
        wx.insert(wx.end(), h_pick, h_pick + np);
        return NGMI_OK;
    }

    /* traninit.c + dctran.c (XSPICE build, no event instances) */
    int tran(double tstep, double tstop, double tmax)
    {
        tstep_run = tstep;
        finalTime = tstop;
        double maxStep = tmax;
        if (maxStep == 0) maxStep = tstep < tstop / 50.0 ? tstep : tstop / 50.0;
        const double delmin = 1e-11 * maxStep;
        double delta = std::fmin(tstop / 100, tstep) / 10;
        breaks.assign({0.0, tstop});
        minBreak = 10.0 * delmin;
        time = 0;
        accepted = rejected = iters = 0;
        lmax = 0.0;
        t_load = t_factor = t_solve = t_other = 0;
        wt.clear();
        wx.clear();
        std::vector<LgPulse> pulses(NVS);
        for (int64_t k = 0; k < NVS; k++) {
            vs[k].break_time = -1.0;
            pulses[k] = pulse_of(vs[k]);
        }
        if (NVS) NGMI_HIP(hipMemcpyAsync(d_pulse, pulses.data(), NVS * sizeof(LgPulse), hipMemcpyHostToDevice, s));
        /* zero iterates, states and the kept device values */
        for (int i = 0; i < 2; i++) NGMI_HIP(hipMemsetAsync(x[i], 0, (size_t)(n + 1) * 8, s));
        for (int l = 0; l < LG_NLEV; l++) {
            if (NCAP) NGMI_HIP(hipMemsetAsync(cs[l], 0, (size_t)2 * NCAP * 8, s));
            if (NM) NGMI_HIP(hipMemsetAsync(ms[l], 0, (size_t)B4X__COUNT * NM * 8, s));
        }
        if (NM) NGMI_HIP(hipMemsetAsync(cache, 0, (size_t)B4C__COUNT * NM * 8, s));
        xnew = 0;
        S = LgStep{};
        S.method = method;
        S.order = 1;
        S.srcFact = 1.0;
        S.time = 0;
        S.delta = 0;
        S.mode = NGMI_MODETRANOP | NGMI_MODEINITJCT;
        int rc = niiter(100);
        if (rc) {
            set_error("large transient: operating point failed (%d)", rc);
            return rc;
        }
        S.order = 1;
        for (int i = 0; i < 7; i++) S.deltaOld[i] = maxStep;
        S.delta = delta;
        double saveDelta = tstop / 50;
        S.mode = NGMI_MODETRAN | NGMI_MODEINITTRAN;
        S.ag[0] = S.ag[1] = 0;
        if ((rc = copy_level(0, 1))) return rc;
        bool firsttime = true;
        for (;;) {
            /* nextTime */
            if (vsrc_accept(S.mode)) {
                set_error("large transient: breakpoint table");
                return NGMI_E_PANIC;
            }
            if (time > breaks[0]) clr_break();
            accepted++;
            if ((rc = dump())) return rc;
            if (almost_equal_ulps(time, finalTime, 100)) return NGMI_OK;
            /* resume */
            delta = std::fmin(delta, maxStep);
            if (almost_equal_ulps(time, breaks[0], 100) || breaks[0] - time <= delmin) {
                S.order = 1;
                delta = std::fmin(delta, .1 * std::fmin(saveDelta, breaks[1] - breaks[0]));
                if (firsttime) delta /= 10;
            }
            while ((breaks[0] <= time + minBreak || almost_equal_ulps(breaks[0], time, 100)) &&
                   breaks[0] < finalTime)
                clr_break();
            if (time + delta > breaks[0]) {
                saveDelta = delta;
                delta = breaks[0] - time;
            } else if (time + 1.9 * delta > breaks[0]) {
                saveDelta = delta;
                delta = (breaks[0] - time) / 2.;
            }
            for (int i = 5; i >= 0; i--) S.deltaOld[i + 1] = S.deltaOld[i];
            S.deltaOld[0] = delta;
            rotate_states();
            for (;;) {
                bool accept = false;
                const double olddelta = delta;
                time += delta;
                S.time = time;
                S.delta = delta;
                S.deltaOld[0] = delta;
                com_cof();
                const int conv = niiter(10);
                if (conv && conv != NGMI_E_ITERLIM && conv != NGMI_E_SINGULAR) return conv;
                S.mode = NGMI_MODETRAN | NGMI_MODEINITPRED;
                if (firsttime) {
                    if ((rc = copy_level(1, 2))) return rc;
                    if ((rc = copy_level(1, 3))) return rc;
                }
                if (conv != 0) {
                    time -= delta;
                    rejected++;
                    delta = delta / 8;
                    if (firsttime) S.mode = NGMI_MODETRAN | NGMI_MODEINITTRAN;
                    S.order = 1;
                } else if (firsttime) {
                    firsttime = false;
                    accept = true;
                } else {
                    double newdelta = delta;
                    if ((rc = trunc(&newdelta))) return rc;
                    if (newdelta > .9 * delta) {
                        if (S.order == 1 && maxord > 1) {
                            newdelta = delta;
                            S.order = 2;
                            if ((rc = trunc(&newdelta))) return rc;
                            if (newdelta <= 1.05 * delta) S.order = 1;
                        }
                        delta = newdelta;
                        accept = true;
                    } else {
                        time -= delta;
                        rejected++;
                        delta = newdelta;
                    }
                }
                if (accept) break;
                if (delta <= delmin) {
                    if (olddelta > delmin) {
                        delta = delmin;
                    } else {
                        set_error("large transient: timestep too small at t = %g", time);
                        return NGMI_E_TIMESTEP;
                    }
                }
            }
        }
    }
};

extern "C" {

ngmi_large *ngmi_large_new(const ngmi_circuit *c, const int *probe, int nprobe, int64_t max_points)
{
    if (!c) {
        set_error("large_new: no circuit");
        return nullptr;
    }
    for (char g : c->ic_given)
        if (g) {
            set_error("large_new: .ic values are not supported for a large circuit");
            return nullptr;
        }
    if (c->uic) {
        set_error("large_new: .tran uic is not supported for a large circuit");
        return nullptr;
    }
    if (c->method == 2 && c->maxord > 2) {
        set_error("large_new: Gear order %d above 2 is not supported for a large circuit", c->maxord);
        return nullptr;
    }
    ngmi_large *L = new ngmi_large;
    auto fail = [&](const char *what) -> ngmi_large * {
        if (what) set_error("large_new: %s", what);
        delete L;
        return nullptr;
    };
    L->NN = c->nnodes;
    L->NM = (int64_t)c->mos.size();
    L->NCAP = (int64_t)c->caps.size();
    L->NRES = (int64_t)c->res.size();
    L->NVS = (int64_t)c->vs.size();
    L->n = L->NN + L->NVS;
    const int64_t n = L->n, NN = L->NN, NM = L->NM, NCAP = L->NCAP, NRES = L->NRES, NVS = L->NVS;
    L->reltol = c->reltol;
    L->abstol = c->abstol;
    L->vntol = c->vntol;
    L->chgtol = c->chgtol;
    L->trtol = c->trtol;
    L->gmin = c->gmin;
    L->method = c->method;
    L->maxord = std::min(c->maxord, 2);
    for (int64_t k = 0; k < NVS; k++) {
        const auto &v = c->vs[k];
        Src sv{};
        sv.p = v.p;
        sv.n = v.n;
        sv.br = (int)(NN + 1 + k);
        sv.pulse = v.npulse > 0;
        sv.npulse = v.npulse;
        sv.dc = v.dc;
        for (int q = 0; q < 7; q++) sv.c[q] = v.pulse[q];
        L->vs.push_back(sv);
    }

    /* ---- elements in CKTload order: BSIM4, CAP, RES, VSRC, instances newest
     * first; each is (row, col, signed source index or a constant) */
    struct El {
        int r, c;
        int64_t src;     /* gsrc index, ~index for a negated one; INT64_MIN: constant */
        double v;
    };
    std::vector<El> el;
    el.reserve((size_t)(16 * NM + 4 * (NCAP + NRES + NVS)));
    for (int64_t j = NM - 1; j >= 0; j--)
        for (int a = 0; a < 4; a++)
            for (int b = 0; b < 4; b++) {
                const int r = c->mos[j].node[a], cc = c->mos[j].node[b];
                if (r && cc) el.push_back({r, cc, NCAP + (int64_t)(a * 4 + b) * NM + j, 0.0});
            }
    for (int64_t k = NCAP - 1; k >= 0; k--) {
        const int p = (int)c->caps[k][0], q = (int)c->caps[k][1];
        if (p) el.push_back({p, p, k, 0.0});
        if (q) el.push_back({q, q, k, 0.0});
        if (p && q) {
            el.push_back({p, q, ~k, 0.0});
            el.push_back({q, p, ~k, 0.0});
        }
    }
    for (int64_t k = NRES - 1; k >= 0; k--) {
        const int p = (int)c->res[k][0], q = (int)c->res[k][1];
        const double g = 1.0 / c->res[k][2];
        if (p) el.push_back({p, p, INT64_MIN, g});
        if (q) el.push_back({q, q, INT64_MIN, g});
        if (p && q) {
            el.push_back({p, q, INT64_MIN, -g});
            el.push_back({q, p, INT64_MIN, -g});
        }
    }
    for (int64_t k = NVS - 1; k >= 0; k--) {
        const Src &v = L->vs[k];
        if (v.p) el.push_back({v.p, v.br, INT64_MIN, 1.0});
        if (v.n) el.push_back({v.n, v.br, INT64_MIN, -1.0});
        if (v.p) el.push_back({v.br, v.p, INT64_MIN, 1.0});
        if (v.n) el.push_back({v.br, v.n, INT64_MIN, -1.0});
    }
    /* CSC of the pattern (0-based), element -> position */
    std::vector<int> er(el.size()), ec(el.size());
    for (size_t e = 0; e < el.size(); e++) {
        er[e] = el[e].r - 1;
        ec[e] = el[e].c - 1;
    }
    L->Ap.resize(n + 1);
    L->Ai.resize(el.size());
    std::vector<int64_t> e2z(el.size());
    int64_t nnz = 0;
    if (ngmi_csc_from_elements((int)n, (int64_t)el.size(), er.data(), ec.data(), L->Ap.data(), L->Ai.data(),
                               e2z.data(), &nnz) != NGMI_OK)
        return fail(nullptr);
    L->Ai.resize(nnz);
    /* matrix plan: per position, the varying sources in element order, the
     * constants summed into a base value in element order */
    std::vector<int64_t> ptr(nnz + 1, 0), ent;
    std::vector<double> base(nnz, 0.0);
    for (size_t e = 0; e < el.size(); e++)
        if (el[e].src != INT64_MIN) ptr[e2z[e] + 1]++;
    for (int64_t z = 0; z < nnz; z++) ptr[z + 1] += ptr[z];
    ent.resize(ptr[nnz]);
    {
        std::vector<int64_t> fill(ptr.begin(), ptr.end() - 1);
        for (size_t e = 0; e < el.size(); e++) {
            if (el[e].src == INT64_MIN) base[e2z[e]] += el[e].v;
            else ent[fill[e2z[e]]++] = el[e].src;
        }
    }
    /* right-hand side plan: rsrc = [ceq | MOSFET rhs | source values] */
    std::vector<std::vector<int64_t>> rrow(n);
    for (int64_t j = NM - 1; j >= 0; j--)
        for (int a = 0; a < 4; a++) {
            const int u = c->mos[j].node[a];
            if (u) rrow[u - 1].push_back(NCAP + (int64_t)a * NM + j);
        }
    for (int64_t k = NCAP - 1; k >= 0; k--) {
        const int p = (int)c->caps[k][0], q = (int)c->caps[k][1];
        if (p) rrow[p - 1].push_back(~k);
        if (q) rrow[q - 1].push_back(k);
    }
    for (int64_t k = NVS - 1; k >= 0; k--) rrow[L->vs[k].br - 1].push_back(NCAP + 4 * NM + k);
    std::vector<int64_t> rptr(n + 1, 0), rent;
    for (int64_t u = 0; u < n; u++) rptr[u + 1] = rptr[u] + (int64_t)rrow[u].size();
    rent.reserve(rptr[n]);
    for (int64_t u = 0; u < n; u++) rent.insert(rent.end(), rrow[u].begin(), rrow[u].end());
    rrow.clear();
    rrow.shrink_to_fit();

    /* ---- symbolic analysis of the LU */
    if (ngmi_splu_analyze((int)n, L->Ap.data(), L->Ai.data(), &L->lu) != NGMI_OK) return fail(nullptr);

    /* ---- BSIM4 instances */
    if (NM) {
        L->set = ngmi_bsim4_set_new();
        for (int64_t j = 0; j < NM; j++) {
            const ngmi_circuit::Mos &mo = c->mos[j];
            std::vector<const char *> names;
            for (const auto &s : mo.names) names.push_back(s.c_str());
            ngmi_bsim4_model *mp = mo.model;
            const int zero = 0;
            if (ngmi_bsim4_set_add(L->set, &mp, 1, &zero, 1, names.data(), (int)names.size(), mo.values.data()) < 0)
                return fail(nullptr);
        }
        if (ngmi_bsim4_set_finalize(L->set) != NGMI_OK) return fail(nullptr);
    }

    /* ---- device arrays */
    std::vector<int> mos_nodes((size_t)4 * NM);
    for (int64_t j = 0; j < NM; j++)
        for (int a = 0; a < 4; a++) mos_nodes[(size_t)a * NM + j] = c->mos[j].node[a];
    std::vector<int> cap_p(NCAP), cap_n(NCAP);
    std::vector<double> cap_c(NCAP);
    for (int64_t k = 0; k < NCAP; k++) {
        cap_p[k] = (int)c->caps[k][0];
        cap_n[k] = (int)c->caps[k][1];
        cap_c[k] = c->caps[k][2];
    }
    int *d_nodes = nullptr, *d_cp = nullptr, *d_cn = nullptr;
    double *d_cc = nullptr;
    if (L->upload(&d_nodes, mos_nodes) || L->upload(&d_cp, cap_p) || L->upload(&d_cn, cap_n) ||
        L->upload(&d_cc, cap_c) || L->upload(&L->ax_ptr, ptr) || L->upload(&L->ax_ent, ent) ||
        L->upload(&L->ax_base, base) || L->upload(&L->rhs_ptr, rptr) || L->upload(&L->rhs_ent, rent) ||
        L->dalloc(&L->Ax, nnz) || L->dalloc(&L->x[0], n + 1) || L->dalloc(&L->x[1], n + 1) ||
        L->dalloc(&L->D.vterm, 4 * NM) || L->dalloc(&L->D.gsrc, NCAP + 16 * NM) ||
        L->dalloc(&L->D.rsrc, NCAP + 4 * NM + NVS) || L->dalloc(&L->cache, (size_t)B4C__COUNT * NM) ||
        L->dalloc(&L->votes, NM) || L->dalloc(&L->flag, 1) || L->dalloc(&L->terr, 1) ||
        L->dalloc(&L->d_pulse, NVS))
        return fail(nullptr);
    for (int l = 0; l < LG_NLEV; l++)
        if (L->dalloc(&L->cs[l], 2 * NCAP) || L->dalloc(&L->ms[l], (size_t)B4X__COUNT * NM)) return fail(nullptr);
    if (hipHostMalloc((void **)&L->h_flag, sizeof(int), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void **)&L->h_terr, 8, hipHostMallocDefault) != hipSuccess)
        return fail("pinned host allocation failed");
    /* probes: 1-based unknowns, all when none given */
    if (probe && nprobe > 0) L->probe.assign(probe, probe + nprobe);
    else
        for (int64_t u = 1; u <= n; u++) L->probe.push_back((int)u);
    for (int u : L->probe)
        if (u < 1 || u > n) return fail("probe outside the unknowns");
    if (L->upload(&L->d_probe, L->probe) || L->dalloc(&L->d_pick, L->probe.size()) ||
        hipHostMalloc((void **)&L->h_pick, std::max<size_t>(1, L->probe.size()) * 8, hipHostMallocDefault) !=
            hipSuccess)
        return fail(nullptr);
    L->max_points = max_points > 0 ? max_points : 1;
    LgDev &D = L->D;
    D.n = n;
    D.nnodes = NN;
    D.nm = NM;
    D.ncap = NCAP;
    D.nvs = NVS;
    D.mos_nodes = d_nodes;
    D.cap_p = d_cp;
    D.cap_n = d_cn;
    D.cap_c = d_cc;
    D.pulse = L->d_pulse;
    D.reltol = L->reltol;
    D.abstol = L->abstol;
    D.vntol = L->vntol;
    D.chgtol = L->chgtol;
    D.trtol = L->trtol;
    return L;
}

void ngmi_large_free(ngmi_large *L) { delete L; }

int ngmi_large_info(const ngmi_large *L, int64_t *n_unknowns, int64_t *nnz, int64_t *lu_nonzeros, double *flops)
{
    if (!L) return NGMI_E_PANIC;
    if (n_unknowns) *n_unknowns = L->n;
    if (nnz) *nnz = L->Ap[L->n];
    return ngmi_splu_info(L->lu, nullptr, nullptr, nullptr, nullptr, lu_nonzeros, flops);
}

int ngmi_large_tran(ngmi_large *L, double tstep, double tstop, double tmax, void *stream)
{
    if (!L || !(tstep > 0) || !(tstop > 0)) {
        set_error("large_tran: bad arguments");
        return NGMI_E_PARMVAL;
    }
    L->s = (hipStream_t)stream;
    return L->tran(tstep, tstop, tmax);
}

int ngmi_large_stats(const ngmi_large *L, int64_t *npts, int64_t *accepted, int64_t *rejected, int64_t *iterations)
{
    if (!L) return NGMI_E_PANIC;
    if (npts) *npts = (int64_t)L->wt.size();
    if (accepted) *accepted = L->accepted;
    if (rejected) *rejected = L->rejected;
    if (iterations) *iterations = L->iters;
    return NGMI_OK;
}

int ngmi_large_multiplier(const ngmi_large *L, double *lmax)
{
    if (!L || !lmax) return NGMI_E_PANIC;
    *lmax = L->lmax;
    return NGMI_OK;
}

int ngmi_large_waveform(const ngmi_large *L, double *t, double *x, int64_t maxpts, int64_t *npts)
{
    if (!L) return NGMI_E_PANIC;
    const int64_t m = std::min<int64_t>(maxpts, (int64_t)L->wt.size());
    const size_t np = L->probe.size();
    if (t) std::copy(L->wt.begin(), L->wt.begin() + m, t);
    if (x) std::copy(L->wx.begin(), L->wx.begin() + (size_t)m * np, x);
    if (npts) *npts = m;
    return NGMI_OK;
}

int ngmi_large_kcl(ngmi_large *L, double *scaled, double *per_kind)
{
    if (!L || L->wt.empty()) {
        set_error("large_kcl: no finished run");
        return NGMI_E_PANIC;
    }
    /* the circuit equations at the last accepted solution (CKTrhsOld): one
     * more load in MODETRAN | INITFLOAT gives A(x) and b(x); F = A x - b */
    const int64_t n = L->n, nnz = L->Ap[n];
    int64_t *dAp = nullptr;
    int *dAi = nullptr;
    double *y = nullptr, *sabs = nullptr, *b = nullptr;
    unsigned long long *out = nullptr;
    int rc = 0;
    if ((rc = L->upload(&dAp, L->Ap)) || (rc = L->upload(&dAi, L->Ai)) || (rc = L->dalloc(&y, n)) ||
        (rc = L->dalloc(&sabs, n)) || (rc = L->dalloc(&b, n)) || (rc = L->dalloc(&out, 3)))
        return rc;
    L->S.mode = NGMI_MODETRAN | NGMI_MODEINITFLOAT;
    if ((rc = L->load())) return rc;
    NGMI_HIP(hipMemcpyAsync(b, L->x[L->xnew] + 1, (size_t)n * 8, hipMemcpyDeviceToDevice, L->s));
    if ((rc = lg_launch_kcl(n, L->NN, dAp, dAi, L->Ax, L->x[1 - L->xnew], b, y, sabs, out, L->s))) return rc;
    unsigned long long h[3];
    NGMI_HIP(hipMemcpyAsync(h, out, sizeof h, hipMemcpyDeviceToHost, L->s));
    NGMI_HIP(hipStreamSynchronize(L->s));
    (void)nnz;
    double d[3];
    std::memcpy(d, h, sizeof d);
    if (scaled) *scaled = d[0];
    if (per_kind) {
        per_kind[0] = d[1];
        per_kind[1] = d[2];
    }
    return NGMI_OK;
}

int ngmi_large_times(const ngmi_large *L, double *load_ms, double *factor_ms, double *solve_ms, double *other_ms)
{
    if (!L) return NGMI_E_PANIC;
    if (load_ms) *load_ms = L->t_load;
    if (factor_ms) *factor_ms = L->t_factor;
    if (solve_ms) *solve_ms = L->t_solve;
    if (other_ms) *other_ms = L->t_other;
    return NGMI_OK;
}

}  // extern "C"
