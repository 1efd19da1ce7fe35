/*
 * Whole-GPU multifrontal LU of one large sparse system: the replacement of
 * klu_refactor (klu_refactor.c:18-482) and klu_solve (klu_solve.c:14-400)
 * for systems far beyond one wave's LDS program, such as the 1024 x 1024
 * power-grid mesh (~1.1M unknowns).  The symbolic phase (splu_host.cpp)
 * fixes the pivot sequence and the frontal tree; this file holds the numeric
 * kernels and the C ABI.
 *
 * Refactor, per call:
 *   k_sp_zero     clears every front (and the zero-pivot flag and the
 *                 largest-multiplier word);
 *   k_sp_scatter  stores each CSC value at its front position;
 *   then by tree height, leaves first:
 *   k_sp_extadd   adds each child's Schur complement into its parent, one
 *                 child slot per launch so no two writes meet;
 *   k_sp_panel    per 32-column panel: the diagonal block factored in LDS
 *                 (Doolittle, no pivoting), then L21 rows and U12 columns by
 *                 triangular solves, one row or column per thread;
 *   k_sp_update   the trailing block minus L21 * U12, 32 x 32 tiles of 2 x 2
 *                 per thread with both operands staged in LDS.
 * Solve: rows permuted (k_sp_perm), forward substitution leaves first with
 * per-front update vectors passed to the parent (k_sp_fwd), backward from
 * the root down (k_sp_bwd), each blocked by 32 with one wave solving the
 * triangle.  Every write has one writer and a fixed order, so the results
 * are the same run to run.
 */
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "ngmi_internal.h"
#include "splu_types.h"

using namespace ngmi;

namespace {

constexpr int WG = 256;
constexpr int NB = SPLU_NB;
constexpr int TL = SPLU_TILE;

/* a wave-uniform lane's double, through two scalar readlanes */
__device__ __forceinline__ double rdlane(double v, int lane)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), lane);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

__global__ void __launch_bounds__(WG) k_sp_zero(double *__restrict__ F, int64_t n, int *__restrict__ flag)
{
    const int64_t stride = (int64_t)gridDim.x * WG;
    for (int64_t i = (int64_t)blockIdx.x * WG + threadIdx.x; i < n; i += stride) F[i] = 0.0;
    if (blockIdx.x == 0 && threadIdx.x < 4) flag[threadIdx.x] = 0;
}

__global__ void __launch_bounds__(WG) k_sp_scatter(const int64_t *__restrict__ pos, const double *__restrict__ Ax,
                                                   double *__restrict__ F, int64_t nnz)
{
    const int64_t e = (int64_t)blockIdx.x * WG + threadIdx.x;
    if (e < nnz) F[pos[e]] = Ax[e];
}

struct SpDev {
    const int *ns, *nb, *parent, *idx, *rel, *kid_ptr, *kids, *lev, *ea, *rowsrc;
    const int64_t *off, *idx_off, *upd_off;
};

/* child slot list ea[0..count): F_parent[rel b, rel a] += F_child[ns+a, ns+b] */
__global__ void __launch_bounds__(WG) k_sp_extadd(SpDev d, const int *__restrict__ list, double *__restrict__ F)
{
    const int c = list[blockIdx.x];
    const int p = d.parent[c];
    const int nsc = d.ns[c], nbc = d.nb[c], nfc = nsc + nbc;
    const int nfp = d.ns[p] + d.nb[p];
    const double *Fc = F + d.off[c];
    double *Fp = F + d.off[p];
    const int *rl = d.rel + d.idx_off[c] + nsc;
    const int64_t tot = (int64_t)nbc * nbc;
    for (int64_t e = (int64_t)blockIdx.y * WG + threadIdx.x; e < tot; e += (int64_t)gridDim.y * WG) {
        const int a = (int)(e % nbc), b = (int)(e / nbc);
        Fp[(int64_t)rl[b] * nfp + rl[a]] += Fc[(int64_t)(nsc + b) * nfc + nsc + a];
    }
}

/* Panel p of every front in lev[0..gridDim.x): factor the w x w diagonal
 * block, write it to dfac (the update launch copies it back, after every
 * block of this launch has read the unfactored one), then L21 rows
 * (x U11 = row) and U12 columns (L11 y = column).  The largest |l| of the
 * panel's multipliers goes into flag[2..3] (a double's bits, atomic max):
 * a pivot sequence fixed in advance has no threshold test, and KLU's
 * partial pivoting (tol 1e-3, klu_kernel.c:370-530) would have exchanged
 * any pivot whose column holds a multiplier above 1/tol. */
__global__ void __launch_bounds__(WG) k_sp_panel(SpDev d, const int *__restrict__ lev, int p, double *__restrict__ F,
                                                 double *__restrict__ dfac, int *__restrict__ flag)
{
    const int f = lev[blockIdx.x];
    const int nsf = d.ns[f], nf = nsf + d.nb[f];
    const int kb = p * NB;
    if (kb >= nsf) return;
    const int w = min(NB, nsf - kb), t0 = kb + w, m = nf - t0;
    if (blockIdx.y > 0 && (int64_t)blockIdx.y * WG >= 2 * (int64_t)m) return;
    double *Fr = F + d.off[f];
    __shared__ double D[NB][NB + 1];
    __shared__ double dinv[NB];
    const int tid = threadIdx.x;
    double lmax = 0.0;
    if (tid < 64) {
        /* one wave: lane j holds column j; step k scales lane k's column by
         * the pivot's reciprocal and every later column subtracts the
         * multipliers, read from lane k, times its own row-k entry */
        const int j = tid;
        double c[NB];
#pragma unroll
        for (int i = 0; i < NB; i++)
            c[i] = (j < w && i < w) ? Fr[(int64_t)(kb + j) * nf + kb + i] : (i == j ? 1.0 : 0.0);
        bool bad = false;
        double diag = 1.0;
#pragma unroll
        for (int k = 0; k < NB; k++) {
            const double piv = rdlane(c[k], k);
            if (j == k) diag = piv;
            bad |= !(fabs(piv) > 0.0) || !isfinite(piv);
            const double inv = 1.0 / piv;
#pragma unroll
            for (int i = 0; i < NB; i++)
                if (i > k && j == k) c[i] *= inv;
            const double ukj = c[k];
#pragma unroll
            for (int i = 0; i < NB; i++) {
                if (i > k) {
                    const double l = rdlane(c[i], k);
                    if (j > k) c[i] -= l * ukj;
                }
            }
        }
        if (j < NB) {
#pragma unroll
            for (int i = 0; i < NB; i++) D[i][j] = c[i];
            dinv[j] = 1.0 / diag;
        }
        if (blockIdx.y == 0 && j < w) {
#pragma unroll
            for (int i = 0; i < NB; i++)
                if (i > j && i < w) lmax = fmax(lmax, fabs(c[i]));
        }
        if (blockIdx.y == 0 && j == 0 && bad) flag[0] = 1;
        if (blockIdx.y == 0 && j < w) {
            double *df = dfac + (int64_t)blockIdx.x * NB * NB + (int64_t)j * w;
#pragma unroll
            for (int i = 0; i < NB; i++)
                if (i < w) df[i] = c[i];
        }
    }
    __syncthreads();
    const int t = blockIdx.y * WG + tid;
    if (t < m) {
        const int i = t0 + t;
        double x[NB];
#pragma unroll
        for (int k = 0; k < NB; k++) x[k] = k < w ? Fr[(int64_t)(kb + k) * nf + i] : 0.0;
        /* column sweep.  Each step reads its row of U from LDS in one
         * batch (the memory clobber keeps the batch in the step) and pins
         * its results, so neither the 528 operands nor the products pile
         * up in registers across steps. */
#pragma unroll
        for (int k = 0; k < NB; k++) {
            asm volatile("" ::: "memory");
            double u[NB];
#pragma unroll
            for (int q = k + 1; q < NB; q++) u[q] = D[k][q];
            const double xk = x[k] * dinv[k];
            x[k] = xk;
#pragma unroll
            for (int q = k + 1; q < NB; q++) x[q] -= xk * u[q];
#pragma unroll
            for (int q = k + 1; q < NB; q++) asm volatile("" : "+v"(x[q]));
        }
#pragma unroll
        for (int k = 0; k < NB; k++)
            if (k < w) {
                Fr[(int64_t)(kb + k) * nf + i] = x[k];
                lmax = fmax(lmax, fabs(x[k]));
            }
    } else if (t < 2 * m) {
        const int j = t0 + (t - m);
        double *col = Fr + (int64_t)j * nf + kb;
        double y[NB];
#pragma unroll
        for (int k = 0; k < NB; k++) y[k] = k < w ? col[k] : 0.0;
#pragma unroll
        for (int k = 0; k < NB - 1; k++) {
            asm volatile("" ::: "memory");
            double l[NB];
#pragma unroll
            for (int q = k + 1; q < NB; q++) l[q] = D[q][k];
            const double yk = y[k];
#pragma unroll
            for (int q = k + 1; q < NB; q++) y[q] -= l[q] * yk;
#pragma unroll
            for (int q = k + 1; q < NB; q++) asm volatile("" : "+v"(y[q]));
        }
#pragma unroll
        for (int k = 0; k < NB; k++)
            if (k < w) col[k] = y[k];
    }
    /* every lane of the block reaches here: the wave's largest multiplier */
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) lmax = fmax(lmax, __shfl_xor(lmax, o));
    if ((tid & 63) == 0 && lmax > 0.0)
        atomicMax(reinterpret_cast<unsigned long long *>(flag + 2), (unsigned long long)__double_as_longlong(lmax));
}

/* F[t0:, t0:] -= L21 * U12 for panel p, TL x TL tile blockIdx.y; block 0
 * also puts the factored diagonal block back from dfac. */
__global__ void __launch_bounds__(WG) k_sp_update(SpDev d, const int *__restrict__ lev, int p, double *__restrict__ F,
                                                  const double *__restrict__ dfac)
{
    const int f = lev[blockIdx.x];
    const int nsf = d.ns[f], nf = nsf + d.nb[f];
    const int kb = p * NB;
    if (kb >= nsf) return;
    const int w = min(NB, nsf - kb), t0 = kb + w, m = nf - t0;
    double *Fr = F + d.off[f];
    const int tid = threadIdx.x;
    if (blockIdx.y == 0) {
        const double *df = dfac + (int64_t)blockIdx.x * NB * NB;
        for (int e = tid; e < w * w; e += WG) Fr[(int64_t)(kb + e / w) * nf + kb + e % w] = df[e];
    }
    const int T = (m + TL - 1) / TL;
    if ((int)blockIdx.y >= T * T) return;
    const int r0 = t0 + (blockIdx.y % T) * TL, c0 = t0 + (blockIdx.y / T) * TL;
    __shared__ double As[NB][TL + 1];
    __shared__ double Bs[NB][TL + 1];
    for (int e = tid; e < NB * TL; e += WG) {
        const int r = e % TL, k = e / TL;
        As[k][r] = (k < w && r0 + r < nf) ? Fr[(int64_t)(kb + k) * nf + r0 + r] : 0.0;
    }
    for (int e = tid; e < NB * TL; e += WG) {
        const int k = e % NB, c = e / NB;
        Bs[k][c] = (k < w && c0 + c < nf) ? Fr[(int64_t)(c0 + c) * nf + kb + k] : 0.0;
    }
    __syncthreads();
    const int tx = tid % 16, ty = tid / 16;
    constexpr int RB = TL / 16;   /* accumulators per thread and dimension */
    double acc[RB][RB] = {};
#pragma unroll 8
    for (int k = 0; k < NB; k++) {
        double a[RB], b[RB];
#pragma unroll
        for (int q = 0; q < RB; q++) a[q] = As[k][tx + 16 * q], b[q] = Bs[k][ty + 16 * q];
#pragma unroll
        for (int q = 0; q < RB; q++)
#pragma unroll
            for (int s = 0; s < RB; s++) acc[q][s] = fma(a[q], b[s], acc[q][s]);
    }
#pragma unroll
    for (int s = 0; s < RB; s++) {
        const int c = c0 + ty + 16 * s;
        if (c >= nf) continue;
#pragma unroll
        for (int q = 0; q < RB; q++) {
            const int r = r0 + tx + 16 * q;
            if (r < nf) Fr[(int64_t)c * nf + r] -= acc[q][s];
        }
    }
}

__global__ void __launch_bounds__(WG) k_sp_perm(const int *__restrict__ rowsrc, const double *__restrict__ b,
                                                double *__restrict__ out, int n)
{
    const int r = blockIdx.x * WG + threadIdx.x;
    if (r < n) out[r] = b[rowsrc[r]];
}

constexpr int SWG = 512;   /* triangle kernels: more rows in flight per CU */

/* forward substitution, pivot block, for the fronts of one height: r = P b
 * on the pivots plus the children's update vectors, then L11 by blocks of
 * 32 (one wave solves the triangle, all threads apply it to the later
 * pivot rows).  The boundary part of r goes to upd; k_sp_fwd_off then
 * subtracts L21 y from it. */
__global__ void __launch_bounds__(SWG) k_sp_fwd(SpDev d, const int *__restrict__ lev, const double *__restrict__ F,
                                                const double *__restrict__ bp, double *__restrict__ upd,
                                                double *__restrict__ y)
{
    extern __shared__ double r[];
    __shared__ double cy[NB];
    const int f = lev[blockIdx.x];
    const int nsf = d.ns[f], nbf = d.nb[f], nf = nsf + nbf;
    const double *Fr = F + d.off[f];
    const int *id = d.idx + d.idx_off[f];
    const int tid = threadIdx.x;
    for (int k = tid; k < nf; k += SWG) r[k] = k < nsf ? bp[id[k]] : 0.0;
    __syncthreads();
    for (int q = d.kid_ptr[f]; q < d.kid_ptr[f + 1]; q++) {
        const int c = d.kids[q];
        const int *rl = d.rel + d.idx_off[c] + d.ns[c];
        const double *u = upd + d.upd_off[c];
        for (int k = tid; k < d.nb[c]; k += SWG) r[rl[k]] += u[k];
        __syncthreads();
    }
    for (int kb = 0; kb < nsf; kb += NB) {
        const int w = min(NB, nsf - kb);
        if (tid < 64) {
            const int l = tid;
            double lv[NB];
#pragma unroll
            for (int k = 0; k < NB; k++)
                lv[k] = Fr[(int64_t)(kb + min(k, w - 1)) * nf + kb + min(l, w - 1)];
            double v = l < w ? r[kb + l] : 0.0;
#pragma unroll
            for (int k = 0; k < NB; k++) {
                const double yk = rdlane(v, k);
                if (k < w && l > k && l < w) v -= lv[k] * yk;
            }
            if (l < w) r[kb + l] = v;
            if (l < NB) cy[l] = l < w ? v : 0.0;
        }
        __syncthreads();
        for (int i = kb + w + tid; i < nsf; i += SWG) {
            double s = 0.0;
#pragma unroll
            for (int k = 0; k < NB; k++) s += Fr[(int64_t)(kb + min(k, w - 1)) * nf + i] * cy[k];
            r[i] -= s;
        }
        __syncthreads();
    }
    for (int k = tid; k < nsf; k += SWG) y[id[k]] = r[k];
    double *uo = upd + d.upd_off[f];
    for (int k = tid; k < nbf; k += SWG) uo[k] = r[nsf + k];
}

constexpr int OFF_ROWS = 64;   /* rows per workgroup of the off-diagonal products */

/* upd[f] -= L21 y_f: 64 boundary rows per workgroup (blockIdx.y), the four
 * waves taking interleaved 64-column slices of each 256-column chunk of y
 * (staged in LDS); the four partial sums are added in wave order, so the
 * result does not depend on timing */
__global__ void __launch_bounds__(WG) k_sp_fwd_off(SpDev d, const int *__restrict__ lev, const double *__restrict__ F,
                                                   const double *__restrict__ y, double *__restrict__ upd)
{
    __shared__ double yv[WG];
    __shared__ double part[WG / 64][OFF_ROWS];
    const int f = lev[blockIdx.x];
    const int nsf = d.ns[f], nbf = d.nb[f], nf = nsf + nbf;
    if ((int)blockIdx.y * OFF_ROWS >= nbf) return;
    const double *Fr = F + d.off[f];
    const int *id = d.idx + d.idx_off[f];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int k = blockIdx.y * OFF_ROWS + lane;
    const int kr = min(k, nbf - 1);
    double s = 0.0;
    for (int j0 = 0; j0 < nsf; j0 += WG) {
        __syncthreads();
        yv[tid] = j0 + tid < nsf ? y[id[j0 + tid]] : 0.0;
        __syncthreads();
        const int js = j0 + wv * 64;
        const int jn = min(64, nsf - js);
        const double *col = Fr + (int64_t)js * nf + nsf + kr;
#pragma unroll 8
        for (int j = 0; j < jn; j++) s += col[(int64_t)j * nf] * yv[wv * 64 + j];
    }
    part[wv][lane] = s;
    __syncthreads();
    if (wv == 0 && k < nbf) {
        double t = part[0][lane];
#pragma unroll
        for (int q = 1; q < WG / 64; q++) t += part[q][lane];
        upd[d.upd_off[f] + k] -= t;
    }
}

/* y_f -= U12 x_boundary for the fronts of one height (ancestors solved):
 * 64 pivot rows per workgroup, the boundary's x staged in LDS and split
 * over the four waves as above */
__global__ void __launch_bounds__(WG) k_sp_bwd_off(SpDev d, const int *__restrict__ lev, const double *__restrict__ F,
                                                   double *__restrict__ y, const double *__restrict__ x)
{
    __shared__ double xv[WG];
    __shared__ double part[WG / 64][OFF_ROWS];
    const int f = lev[blockIdx.x];
    const int nsf = d.ns[f], nbf = d.nb[f], nf = nsf + nbf;
    if ((int)blockIdx.y * OFF_ROWS >= nsf || nbf == 0) return;
    const double *Fr = F + d.off[f];
    const int *id = d.idx + d.idx_off[f];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int i = blockIdx.y * OFF_ROWS + lane;
    const int ir = min(i, nsf - 1);
    double s = 0.0;
    for (int j0 = 0; j0 < nbf; j0 += WG) {
        __syncthreads();
        xv[tid] = j0 + tid < nbf ? x[id[nsf + j0 + tid]] : 0.0;
        __syncthreads();
        const int js = j0 + wv * 64;
        const int jn = min(64, nbf - js);
        const double *col = Fr + (int64_t)(nsf + js) * nf + ir;
#pragma unroll 8
        for (int j = 0; j < jn; j++) s += col[(int64_t)j * nf] * xv[wv * 64 + j];
    }
    part[wv][lane] = s;
    __syncthreads();
    if (wv == 0 && i < nsf) {
        double t = part[0][lane];
#pragma unroll
        for (int q = 1; q < WG / 64; q++) t += part[q][lane];
        y[id[i]] -= t;
    }
}

/* backward substitution, pivot block: U11 x = y by blocks of 32 from the
 * last, one wave per triangle */
__global__ void __launch_bounds__(SWG) k_sp_bwd(SpDev d, const int *__restrict__ lev, const double *__restrict__ F,
                                                const double *__restrict__ y, double *__restrict__ x)
{
    extern __shared__ double r[];
    __shared__ double cx[NB];
    const int f = lev[blockIdx.x];
    const int nsf = d.ns[f], nf = nsf + d.nb[f];
    const double *Fr = F + d.off[f];
    const int *id = d.idx + d.idx_off[f];
    const int tid = threadIdx.x;
    for (int k = tid; k < nsf; k += SWG) r[k] = y[id[k]];
    __syncthreads();
    for (int kb = ((nsf - 1) / NB) * NB; kb >= 0; kb -= NB) {
        const int w = min(NB, nsf - kb);
        if (tid < 64) {
            const int l = tid;
            double uv[NB];
#pragma unroll
            for (int k = 0; k < NB; k++)
                uv[k] = Fr[(int64_t)(kb + min(k, w - 1)) * nf + kb + min(l, w - 1)];
            double v = l < w ? r[kb + l] : 0.0;
#pragma unroll
            for (int k = NB - 1; k >= 0; k--) {
                if (k < w) {
                    if (l == k) v /= uv[k];
                    const double xk = rdlane(v, k);
                    if (l < k) v -= uv[k] * xk;
                }
            }
            if (l < w) r[kb + l] = v;
            if (l < NB) cx[l] = l < w ? v : 0.0;
        }
        __syncthreads();
        for (int i = tid; i < kb; i += SWG) {
            double s = 0.0;
#pragma unroll
            for (int k = 0; k < NB; k++) s += Fr[(int64_t)(kb + min(k, w - 1)) * nf + i] * cx[k];
            r[i] -= s;
        }
        __syncthreads();
    }
    for (int k = tid; k < nsf; k += SWG) x[id[k]] = r[k];
}

enum { L_EXTADD, L_PANEL, L_UPDATE };
struct Step {
    int kind, h, p, slot;
    unsigned gx, gy;
};

template <class T>
int upload(T **dst, const std::vector<T> &v)
{
    const size_t bytes = std::max<size_t>(1, v.size()) * sizeof(T);
    NGMI_HIP(hipMalloc((void **)dst, bytes));
    if (!v.empty()) NGMI_HIP(hipMemcpy(*dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return NGMI_OK;
}

}  // namespace

struct ngmi_splu {
    SpluSym S;
    bool ready = false;      /* device buffers allocated and plan built */
    bool factored = false;   /* the last ngmi_splu_factor succeeded */
    double lmax = 0.0;       /* its largest |l| multiplier */
    SpDev d{};
    int *d_ns = nullptr, *d_nb = nullptr, *d_parent = nullptr, *d_idx = nullptr, *d_rel = nullptr;
    int *d_kid_ptr = nullptr, *d_kids = nullptr, *d_lev = nullptr, *d_ea = nullptr, *d_rowsrc = nullptr;
    int *d_flag = nullptr;
    int64_t *d_off = nullptr, *d_idx_off = nullptr, *d_upd_off = nullptr, *d_ent = nullptr;
    double *F = nullptr, *dfac = nullptr, *upd = nullptr, *vb = nullptr, *vy = nullptr, *vx = nullptr;
    std::vector<Step> plan;
    std::vector<int> lev_lds;   /* dynamic LDS bytes of the solve launches by height */
    std::vector<int> lev_ns, lev_nb;   /* largest pivot and boundary counts by height */
};

/* frees every device buffer and nulls its pointer, so a later
 * splu_device call starts from nothing */
static void splu_release(ngmi_splu *s)
{
    void **bufs[] = {(void **)&s->d_ns, (void **)&s->d_nb, (void **)&s->d_parent, (void **)&s->d_idx,
                     (void **)&s->d_rel, (void **)&s->d_kid_ptr, (void **)&s->d_kids, (void **)&s->d_lev,
                     (void **)&s->d_ea, (void **)&s->d_rowsrc, (void **)&s->d_flag, (void **)&s->d_off,
                     (void **)&s->d_idx_off, (void **)&s->d_upd_off, (void **)&s->d_ent, (void **)&s->F,
                     (void **)&s->dfac, (void **)&s->upd, (void **)&s->vb, (void **)&s->vy, (void **)&s->vx};
    for (void **b : bufs) {
        if (*b) (void)hipFree(*b);
        *b = nullptr;
    }
    s->ready = false;
    s->factored = false;
}

static int splu_device_alloc(ngmi_splu *s);

static int splu_device(ngmi_splu *s)
{
    if (s->ready) return NGMI_OK;
    const int rc = splu_device_alloc(s);
    if (rc != NGMI_OK) splu_release(s);
    return rc;
}

static int splu_device_alloc(ngmi_splu *s)
{
    const SpluSym &S = s->S;
    int rc;
#define UP(dst, v) \
    if ((rc = upload(&dst, v)) != NGMI_OK) return rc
    UP(s->d_ns, S.ns);
    UP(s->d_nb, S.nb);
    UP(s->d_parent, S.parent);
    UP(s->d_idx, S.idx);
    UP(s->d_rel, S.rel);
    UP(s->d_kid_ptr, S.kid_ptr);
    UP(s->d_kids, S.kids);
    UP(s->d_lev, S.lev);
    UP(s->d_ea, S.ea);
    UP(s->d_rowsrc, S.rowsrc);
    UP(s->d_off, S.off);
    UP(s->d_idx_off, S.idx_off);
    UP(s->d_upd_off, S.upd_off);
    UP(s->d_ent, S.ent_pos);
#undef UP
    NGMI_HIP(hipMalloc((void **)&s->d_flag, 4 * sizeof(int)));   /* zero-pivot flag, pad, largest |l| */
    NGMI_HIP(hipMalloc((void **)&s->F, std::max<int64_t>(1, S.store) * sizeof(double)));
    NGMI_HIP(hipMalloc((void **)&s->dfac, (size_t)std::max(1, S.max_lev) * NB * NB * sizeof(double)));
    NGMI_HIP(hipMalloc((void **)&s->upd, std::max<int64_t>(1, S.upd_off[S.nf]) * sizeof(double)));
    NGMI_HIP(hipMalloc((void **)&s->vb, (size_t)S.n * sizeof(double)));
    NGMI_HIP(hipMalloc((void **)&s->vy, (size_t)S.n * sizeof(double)));
    NGMI_HIP(hipMalloc((void **)&s->vx, (size_t)S.n * sizeof(double)));
    s->d = SpDev{s->d_ns, s->d_nb, s->d_parent, s->d_idx, s->d_rel, s->d_kid_ptr, s->d_kids,
                 s->d_lev, s->d_ea, s->d_rowsrc, s->d_off, s->d_idx_off, s->d_upd_off};

    /* launch plan: for each height, the extend-add slots, then the panels */
    s->plan.clear();
    s->lev_lds.assign(S.H + 1, 0);
    s->lev_ns.assign(S.H + 1, 0);
    s->lev_nb.assign(S.H + 1, 0);
    for (int h = 0; h <= S.H; h++) {
        for (int sl = 0; sl < S.max_slots; sl++) {
            const int b = S.ea_ptr[(size_t)h * S.max_slots + sl], e = S.ea_ptr[(size_t)h * S.max_slots + sl + 1];
            if (e == b) continue;
            int64_t most = 1;
            for (int q = b; q < e; q++) most = std::max<int64_t>(most, (int64_t)S.nb[S.ea[q]] * S.nb[S.ea[q]]);
            s->plan.push_back({L_EXTADD, h, 0, sl, (unsigned)(e - b),
                               (unsigned)std::min<int64_t>(1024, (most + WG - 1) / WG)});
        }
        int maxns = 0, maxnf = 0;
        for (int q = S.lev_ptr[h]; q < S.lev_ptr[h + 1]; q++) {
            maxns = std::max(maxns, S.ns[S.lev[q]]);
            maxnf = std::max(maxnf, S.ns[S.lev[q]] + S.nb[S.lev[q]]);
        }
        s->lev_lds[h] = maxnf * (int)sizeof(double);
        s->lev_ns[h] = maxns;
        s->lev_nb[h] = 0;
        for (int q = S.lev_ptr[h]; q < S.lev_ptr[h + 1]; q++) s->lev_nb[h] = std::max(s->lev_nb[h], S.nb[S.lev[q]]);
        const unsigned nl = (unsigned)(S.lev_ptr[h + 1] - S.lev_ptr[h]);
        for (int p = 0; p * NB < maxns; p++) {
            int64_t chunks = 1, tiles = 1;
            for (int q = S.lev_ptr[h]; q < S.lev_ptr[h + 1]; q++) {
                const int f = S.lev[q];
                if (p * NB >= S.ns[f]) continue;
                const int t0 = p * NB + std::min(NB, S.ns[f] - p * NB);
                const int64_t m = (int64_t)S.ns[f] + S.nb[f] - t0;
                chunks = std::max<int64_t>(chunks, (2 * m + WG - 1) / WG);
                const int64_t T = (m + TL - 1) / TL;
                tiles = std::max<int64_t>(tiles, T * T);
            }
            if (chunks > 65535 || tiles > 65535) {
                set_error("splu: front too large for the launch grid");
                return NGMI_E_UNSUPP;
            }
            s->plan.push_back({L_PANEL, h, p, 0, nl, (unsigned)chunks});
            s->plan.push_back({L_UPDATE, h, p, 0, nl, (unsigned)tiles});
        }
    }
    s->ready = true;
    return NGMI_OK;
}

extern "C" {

int ngmi_splu_analyze(int n, const int64_t *Ap, const int *Ai, ngmi_splu **out)
{
    if (!out) return NGMI_E_PANIC;
    *out = nullptr;
    ngmi_splu *s = new ngmi_splu();
    const int rc = splu_analyze(n, Ap, Ai, s->S);
    if (rc != NGMI_OK) {
        delete s;
        return rc;
    }
    *out = s;
    return NGMI_OK;
}

void ngmi_splu_free(ngmi_splu *s)
{
    if (!s) return;
    splu_release(s);
    delete s;
}

int ngmi_splu_info(const ngmi_splu *s, int64_t *fronts, int64_t *heights, int64_t *max_front,
                   int64_t *front_doubles, int64_t *lu_nonzeros, double *flops)
{
    if (!s) return NGMI_E_PANIC;
    if (fronts) *fronts = s->S.nf;
    if (heights) *heights = s->S.H + 1;
    if (max_front) *max_front = s->S.max_front;
    if (front_doubles) *front_doubles = s->S.store;
    if (lu_nonzeros) *lu_nonzeros = s->S.lu_nnz;
    if (flops) *flops = s->S.flops;
    return NGMI_OK;
}

int ngmi_splu_front(const ngmi_splu *s, int f, int *ns, int *nb, int *parent, int *idx)
{
    if (!s || f < 0 || f >= s->S.nf) {
        set_error("splu_front: no front %d", f);
        return NGMI_E_PARMVAL;
    }
    const SpluSym &S = s->S;
    if (ns) *ns = S.ns[f];
    if (nb) *nb = S.nb[f];
    if (parent) *parent = S.parent[f];
    if (idx) std::copy(S.idx.begin() + S.idx_off[f], S.idx.begin() + S.idx_off[f + 1], idx);
    return NGMI_OK;
}

int ngmi_splu_rowsrc(const ngmi_splu *s, int *rowsrc)
{
    if (!s || !rowsrc) return NGMI_E_PANIC;
    std::copy(s->S.rowsrc.begin(), s->S.rowsrc.end(), rowsrc);
    return NGMI_OK;
}

int ngmi_splu_factor(ngmi_splu *s, const double *Ax, void *stream)
{
    if (!s || !Ax) return NGMI_E_PANIC;
    int rc = splu_device(s);
    if (rc != NGMI_OK) return rc;
    s->factored = false;
    const SpluSym &S = s->S;
    hipStream_t st = (hipStream_t)stream;
    const unsigned zb = (unsigned)std::min<int64_t>(8192, (S.store + WG - 1) / WG + 1);
    hipLaunchKernelGGL(k_sp_zero, dim3(zb), dim3(WG), 0, st, s->F, S.store, s->d_flag);
    if (S.nnz > 0)
        hipLaunchKernelGGL(k_sp_scatter, dim3((unsigned)((S.nnz + WG - 1) / WG)), dim3(WG), 0, st, s->d_ent, Ax,
                           s->F, S.nnz);
    for (const Step &p : s->plan) {
        const int *lev = s->d_lev + S.lev_ptr[p.h];
        if (p.kind == L_EXTADD)
            hipLaunchKernelGGL(k_sp_extadd, dim3(p.gx, p.gy), dim3(WG), 0, st, s->d,
                               s->d_ea + S.ea_ptr[(size_t)p.h * S.max_slots + p.slot], s->F);
        else if (p.kind == L_PANEL)
            hipLaunchKernelGGL(k_sp_panel, dim3(p.gx, p.gy), dim3(WG), 0, st, s->d, lev, p.p, s->F, s->dfac,
                               s->d_flag);
        else
            hipLaunchKernelGGL(k_sp_update, dim3(p.gx, p.gy), dim3(WG), 0, st, s->d, lev, p.p, s->F, s->dfac);
    }
    NGMI_HIP(hipGetLastError());
    int fl[4] = {0, 0, 0, 0};
    NGMI_HIP(hipMemcpyAsync(fl, s->d_flag, sizeof(fl), hipMemcpyDeviceToHost, st));
    NGMI_HIP(hipStreamSynchronize(st));
    std::memcpy(&s->lmax, fl + 2, sizeof(double));
    if (fl[0]) {
        set_error("splu_factor: zero or non-finite pivot");
        return NGMI_E_SINGULAR;
    }
    s->factored = true;
    return NGMI_OK;
}

int ngmi_splu_multiplier(const ngmi_splu *s, double *lmax)
{
    if (!s || !lmax) return NGMI_E_PANIC;
    if (!s->factored) {
        set_error("splu_multiplier: no successful factorization");
        return NGMI_E_SINGULAR;
    }
    *lmax = s->lmax;
    return NGMI_OK;
}

int ngmi_splu_solve(ngmi_splu *s, double *b, void *stream)
{
    if (!s || !b) return NGMI_E_PANIC;
    if (!s->ready || !s->factored) {
        set_error("splu_solve: no successful factorization");
        return NGMI_E_SINGULAR;
    }
    const SpluSym &S = s->S;
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(k_sp_perm, dim3((S.n + WG - 1) / WG), dim3(WG), 0, st, s->d_rowsrc, b, s->vb, S.n);
    for (int h = 0; h <= S.H; h++) {
        const unsigned nl = (unsigned)(S.lev_ptr[h + 1] - S.lev_ptr[h]);
        const int *lev = s->d_lev + S.lev_ptr[h];
        hipLaunchKernelGGL(k_sp_fwd, dim3(nl), dim3(SWG), s->lev_lds[h], st, s->d, lev, s->F, s->vb, s->upd, s->vy);
        if (s->lev_nb[h] > 0)
            hipLaunchKernelGGL(k_sp_fwd_off, dim3(nl, (s->lev_nb[h] + OFF_ROWS - 1) / OFF_ROWS), dim3(WG), 0, st, s->d, lev,
                               s->F, s->vy, s->upd);
    }
    for (int h = S.H; h >= 0; h--) {
        const unsigned nl = (unsigned)(S.lev_ptr[h + 1] - S.lev_ptr[h]);
        const int *lev = s->d_lev + S.lev_ptr[h];
        if (s->lev_nb[h] > 0)
            hipLaunchKernelGGL(k_sp_bwd_off, dim3(nl, (s->lev_ns[h] + OFF_ROWS - 1) / OFF_ROWS), dim3(WG), 0, st, s->d, lev,
                               s->F, s->vy, s->vx);
        hipLaunchKernelGGL(k_sp_bwd, dim3(nl), dim3(SWG), s->lev_lds[h], st, s->d, lev, s->F, s->vy, s->vx);
    }
    NGMI_HIP(hipGetLastError());
    NGMI_HIP(hipMemcpyAsync(b, s->vx, (size_t)S.n * sizeof(double), hipMemcpyDeviceToDevice, st));
    return NGMI_OK;
}

}  // extern "C"
