/*
 * Kernels of the one-large-circuit transient (large_host.cpp): the vector
 * work of CKTload, NIconvTest and CKTterr for a circuit too large for one
 * wavefront (the power-grid mesh: ~1.1M unknowns, ~1M capacitors, ~2M
 * resistors, tens of thousands of MOSFETs), as grid-wide kernels.  The
 * matrix and right-hand side are assembled by signed gathers in a fixed
 * order, so every run gives the same bits; reductions (convergence flag,
 * LTE minimum) are order-independent (or / min).
 */
#include <cmath>
#include "ngmi_internal.h"
#include "large_types.h"

namespace ngmi {

constexpr int LG_WG = 256;

/* terminal voltages of every MOSFET from the current iterate */
__global__ void __launch_bounds__(LG_WG)
k_lg_vterm(int64_t nm, const int *__restrict__ nodes, const double *__restrict__ x, double *__restrict__ vterm)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nm) return;
#pragma unroll
    for (int k = 0; k < 4; k++) vterm[k * nm + i] = x[nodes[k * nm + i]];
}

/* CAPload (capload.c:19-86) with NIintegrate (niinteg.c:18-80) for every
 * capacitor: geq into gsrc[i], ceq into rsrc[i] */
__global__ void __launch_bounds__(LG_WG)
k_lg_caps(LgDev D, LgStep S, const double *__restrict__ x)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= D.ncap) return;
    const int64_t nc = D.ncap;
    double *q0 = S.cs[0], *q1 = S.cs[1];
    const double C = D.cap_c[i];
    const int mode = S.mode;
    const bool cond1 = ((mode & (NGMI_MODEDCOP | NGMI_MODETRANOP | NGMI_MODEDCTRANCURVE)) &&
                        (mode & NGMI_MODEINITJCT)) ||
                       ((mode & NGMI_MODEUIC) && (mode & NGMI_MODEINITTRAN));
    const double vcap = cond1 ? 0.0 : x[D.cap_p[i]] - x[D.cap_n[i]];
    double geq = 0.0, ceq = 0.0;
    if (mode & (NGMI_MODETRAN | NGMI_MODEAC)) {
        if (mode & NGMI_MODEINITPRED) {
            q0[i] = q1[i];
        } else {
            q0[i] = C * vcap;
            if (mode & NGMI_MODEINITTRAN) q1[i] = q0[i];
        }
        /* NIintegrate: charge at [i], current at [nc + i] of each level */
        double ccap;
        if (S.method == 1) {
            if (S.order == 1) ccap = S.ag[0] * q0[i] + S.ag[1] * q1[i];
            else ccap = -q1[nc + i] * S.ag[1] + S.ag[0] * (q0[i] - q1[i]);
        } else {
            ccap = 0.0;
            for (int k = 0; k <= S.order; k++) ccap += S.ag[k] * S.cs[k][i];
        }
        q0[nc + i] = ccap;
        if (mode & NGMI_MODEINITTRAN) q1[nc + i] = q0[nc + i];
        geq = S.ag[0] * C;
        ceq = ccap - S.ag[0] * q0[i];
    } else if (mode & NGMI_MODETRANOP) {
        q0[i] = C * vcap;
    }
    D.gsrc[i] = geq;
    D.rsrc[i] = ceq;
}

/* VSRCload values (vsrcload.c:78-170, DC and PULSE; CKTsrcFact in TRANOP)
 * into the right-hand-side sources after the capacitor and MOSFET ones */
__global__ void __launch_bounds__(LG_WG)
k_lg_vsrc(LgDev D, LgStep S)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= D.nvs) return;
    const LgPulse P = D.pulse[k];
    double value;
    if (!P.pulse) {
        value = P.dc;
    } else {
        double time = (S.mode & (NGMI_MODEDCOP | NGMI_MODETRANOP | NGMI_MODEDCTRANCURVE)) ? 0.0 : S.time;
        time -= P.td;
        if (time > P.per) {
            const double basetime = P.per * floor(time / P.per);
            time -= basetime;
        }
        if (time <= 0 || time >= P.tr + P.pw + P.tf) value = P.v1;
        else if (time >= P.tr && time <= P.tr + P.pw) value = P.v2;
        else if (time > 0 && time < P.tr) value = P.v1 + (P.v2 - P.v1) * (time) / P.tr;
        else value = P.v2 + (P.v1 - P.v2) * (time - (P.tr + P.pw)) / P.tf;
    }
    if (S.mode & NGMI_MODETRANOP) value *= S.srcFact;
    D.rsrc[D.ncap + 4 * D.nm + k] = value;
}

/* out[r] = sum over the plan entries of row r, in plan order, of
 * +-src[index] (entry e >= 0 adds src[e], e < 0 subtracts src[~e]), then
 * + base[r]; rows outside [0, nrows) untouched */
__global__ void __launch_bounds__(LG_WG)
k_lg_gather(int64_t nrows, const int64_t *__restrict__ ptr, const int64_t *__restrict__ ent,
            const double *__restrict__ src, const double *__restrict__ base, double *__restrict__ out)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrows) return;
    double acc = 0.0;
    for (int64_t p = ptr[r], e = ptr[r + 1]; p < e; p++) {
        const int64_t k = ent[p];
        acc += k >= 0 ? src[k] : -src[~k];
    }
    out[r] = base ? acc + base[r] : acc;
}

/* NIconvTest (niconv.c:21-90) over every unknown plus the device votes;
 * flag[0] |= 1 on any failure */
__global__ void __launch_bounds__(LG_WG)
k_lg_conv(LgDev D, const double *__restrict__ xn, const double *__restrict__ xo, const int *__restrict__ votes,
          int *flag)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool bad = false;
    if (i < D.n) {
        const int64_t u = i + 1;
        const double nw = xn[u], old = xo[u];
        const double tol = D.reltol * fmax(fabs(old), fabs(nw)) + (u <= D.nnodes ? D.vntol : D.abstol);
        bad = nw != nw || fabs(nw - old) > tol;
    }
    if (i < D.nm && votes[i]) bad = true;
    if (__any(bad) && (threadIdx.x & 63) == 0) atomicOr(flag, 1);
}

/* CKTterr (cktterr.c:13-78) for one charge state at trapezoidal or Gear
 * order 1..2: the step its truncation error allows */
__device__ __forceinline__ double lg_terr(double q0, double q1, double q2, double q3, double c0, double c1,
                                          const LgDev &D, const LgStep &S)
{
    const double volttol = D.abstol + D.reltol * fmax(fabs(c0), fabs(c1));
    double chargetol = fmax(fabs(q0), fabs(q1));
    chargetol = D.reltol * fmax(chargetol, D.chgtol) / S.delta;
    const double tol = fmax(volttol, chargetol);
    double diff0, factor;
    if (S.order == 1) {
        const double a = (q0 - q1) / S.deltaOld[0];
        const double b = (q1 - q2) / S.deltaOld[1];
        diff0 = (a - b) / (S.deltaOld[1] + S.deltaOld[0]);
        factor = .5;
    } else {
        double a = (q0 - q1) / S.deltaOld[0], b = (q1 - q2) / S.deltaOld[1];
        const double c = (q2 - q3) / S.deltaOld[2];
        const double t0 = S.deltaOld[1] + S.deltaOld[0], t1 = S.deltaOld[2] + S.deltaOld[1];
        a = (a - b) / t0;
        b = (b - c) / t1;
        diff0 = (a - b) / (t1 + S.deltaOld[0]);
        factor = S.method == 2 ? .2222222222 : .08333333333;
    }
    double del = D.trtol * tol / fmax(D.abstol, factor * fabs(diff0));
    if (S.order == 2) del = sqrt(del);
    return del;
}

/* CKTtrunc (ckttrunc.c:20-56): CAPtrunc over the capacitors and BSIM4trunc
 * (qb, qg, qd; b4trunc.c:32-71) over the MOSFETs; the minimum goes to
 * *out (non-negative doubles order as their bit patterns) */
__global__ void __launch_bounds__(LG_WG)
k_lg_terr(LgDev D, LgStep S, unsigned long long *out)
{
    __shared__ double red[LG_WG / 64];
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double m = 1e308;
    const int64_t nc = D.ncap, nm = D.nm;
    if (i < nc) {
        const double q3 = S.order > 1 ? S.cs[3][i] : 0.0;
        m = lg_terr(S.cs[0][i], S.cs[1][i], S.cs[2][i], q3, S.cs[0][nc + i], S.cs[1][nc + i], D, S);
    } else if (i < nc + 3 * nm) {
        const int64_t j = (i - nc) / 3, r = (i - nc) - 3 * j;
        const int which = r == 0 ? B4X_qb : (r == 1 ? B4X_qg : B4X_qd);
        const int64_t o = which * nm + j;
        const double q3 = S.order > 1 ? S.ms[3][o] : 0.0;
        m = lg_terr(S.ms[0][o], S.ms[1][o], S.ms[2][o], q3, S.ms[0][o + nm], S.ms[1][o + nm], D, S);
    }
    for (int s = 32; s > 0; s >>= 1) m = fmin(m, __shfl_xor(m, s));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < LG_WG / 64; w++) m = fmin(m, red[w]);
        if (!(m >= 0.0)) m = 0.0;
        atomicMin(out, (unsigned long long)__double_as_longlong(m));
    }
}

/* out[k] = x[idx[k]] */
__global__ void __launch_bounds__(LG_WG)
k_lg_pick(int n, const int *__restrict__ idx, const double *__restrict__ x, double *__restrict__ out)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) out[k] = x[idx[k]];
}

/* KCL check, step 1: y[r] += A[r][c] x[c] and s[r] += |A[r][c] x[c]| for
 * column c of the CSC matrix (a diagnostic; summation order is free) */
__global__ void __launch_bounds__(LG_WG)
k_lg_spmv(int64_t n, const int64_t *__restrict__ Ap, const int *__restrict__ Ai, const double *__restrict__ Ax,
          const double *__restrict__ x, double *y, double *sabs)
{
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const double xc = x[c + 1];
    for (int64_t p = Ap[c]; p < Ap[c + 1]; p++) {
        const double v = Ax[p] * xc;
        atomicAdd(&y[Ai[p]], v);
        atomicAdd(&sabs[Ai[p]], fabs(v));
    }
}

/* step 2: F = A x - b per row; out[0] = max |F| / (sum |A x| + |b|),
 * out[1] / out[2] = max |F| over node / branch rows (bit patterns of
 * non-negative doubles) */
__global__ void __launch_bounds__(LG_WG)
k_lg_resid(int64_t n, int64_t nnodes, const double *__restrict__ y, const double *__restrict__ sabs,
           const double *__restrict__ b, unsigned long long *out)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const double F = fabs(y[r] - b[r]);
    const double den = sabs[r] + fabs(b[r]);
    const double rel = den > 0 ? F / den : 0.0;
    atomicMax(&out[0], (unsigned long long)__double_as_longlong(rel));
    atomicMax(&out[r < nnodes ? 1 : 2], (unsigned long long)__double_as_longlong(F));
}

static inline unsigned blocks_for(int64_t n) { return (unsigned)((n + LG_WG - 1) / LG_WG); }

int lg_launch_vterm(const LgDev &D, const double *x, hipStream_t s)
{
    if (D.nm == 0) return NGMI_OK;
    hipLaunchKernelGGL(k_lg_vterm, dim3(blocks_for(D.nm)), dim3(LG_WG), 0, s, D.nm, D.mos_nodes, x, D.vterm);
    NGMI_HIP(hipGetLastError());
    return NGMI_OK;
}

int lg_launch_sources(const LgDev &D, const LgStep &S, const double *x, hipStream_t s)
{
    if (D.ncap) hipLaunchKernelGGL(k_lg_caps, dim3(blocks_for(D.ncap)), dim3(LG_WG), 0, s, D, S, x);
    if (D.nvs) hipLaunchKernelGGL(k_lg_vsrc, dim3(blocks_for(D.nvs)), dim3(LG_WG), 0, s, D, S);
    NGMI_HIP(hipGetLastError());
    return NGMI_OK;
}

int lg_launch_gather(int64_t nrows, const int64_t *ptr, const int64_t *ent, const double *src, const double *base,
                     double *out, hipStream_t s)
{
    if (nrows == 0) return NGMI_OK;
    hipLaunchKernelGGL(k_lg_gather, dim3(blocks_for(nrows)), dim3(LG_WG), 0, s, nrows, ptr, ent, src, base, out);
    NGMI_HIP(hipGetLastError());
    return NGMI_OK;
}

int lg_launch_conv(const LgDev &D, const double *xn, const double *xo, const int *votes, int *flag, hipStream_t s)
{
    const int64_t n = std::max(D.n, D.nm);
    hipLaunchKernelGGL(k_lg_conv, dim3(blocks_for(n)), dim3(LG_WG), 0, s, D, xn, xo, votes, flag);
    NGMI_HIP(hipGetLastError());
    return NGMI_OK;
}

int lg_launch_terr(const LgDev &D, const LgStep &S, unsigned long long *out, hipStream_t s)
{
    const int64_t n = D.ncap + 3 * D.nm;
    if (n == 0) return NGMI_OK;
    hipLaunchKernelGGL(k_lg_terr, dim3(blocks_for(n)), dim3(LG_WG), 0, s, D, S, out);
    NGMI_HIP(hipGetLastError());
    return NGMI_OK;
}

int lg_launch_pick(int n, const int *idx, const double *x, double *out, hipStream_t s)
{
    if (n == 0) return NGMI_OK;
    hipLaunchKernelGGL(k_lg_pick, dim3(blocks_for(n)), dim3(LG_WG), 0, s, n, idx, x, out);
    NGMI_HIP(hipGetLastError());
    return NGMI_OK;
}

int lg_launch_kcl(int64_t n, int64_t nnodes, const int64_t *Ap, const int *Ai, const double *Ax, const double *x,
                  const double *b, double *y, double *sabs, unsigned long long *out, hipStream_t s)
{
    if (n == 0) return NGMI_OK;
    NGMI_HIP(hipMemsetAsync(y, 0, (size_t)n * 8, s));
    NGMI_HIP(hipMemsetAsync(sabs, 0, (size_t)n * 8, s));
    NGMI_HIP(hipMemsetAsync(out, 0, 3 * 8, s));
    hipLaunchKernelGGL(k_lg_spmv, dim3(blocks_for(n)), dim3(LG_WG), 0, s, n, Ap, Ai, Ax, x, y, sabs);
    hipLaunchKernelGGL(k_lg_resid, dim3(blocks_for(n)), dim3(LG_WG), 0, s, n, nnodes, y, sabs, b, out);
    NGMI_HIP(hipGetLastError());
    return NGMI_OK;
}

}  // namespace ngmi
