/*
 * Newton initial-guess network: a 16 -> 256 -> 256 -> 256 -> 1 MLP with SiLU,
 * BF16 weights and activations on the matrix cores (v_mfma_f32_32x32x16
 * bf16, FP32 accumulate), FP32 master weights and Adam for training.
 *
 * One kernel computes C[i][j] = sum_k A(i,k) B(j,k) (+ bias[j]) for strided
 * BF16 operands, so the forward pass (X W^T), the input gradient (dY W) and
 * the weight gradient (dY^T X) are the same kernel with different strides.
 * Each wavefront owns a 32 x 32 tile of C; lane l feeds rows/cols l % 32 and
 * the k octet (l / 32) * 8 .. +7 of every 16-deep step, and receives
 * C rows (v / 4) * 8 + (l / 32) * 4 + v % 4, column l % 32, v = 0..15.
 *
 * Inference inside the transient uses k_guess_net instead: one kernel from
 * the feature rows to the nodes' starting values, with the weights held in
 * registers and the activations in LDS.  It issues the same matrix-core
 * steps in the same order and the same element-wise code, so its outputs
 * equal the layer-by-layer path's bit for bit.
 */
#include <algorithm>
#include "ngmi_internal.h"
#include "mlp_types.h"
#include "guess_types.h"

namespace ngmi {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ f32x16 mfma16(bf16x8 a, bf16x8 b, f32x16 acc)
{
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
}

/* epilogue modes */
enum { EPI_NONE = 0, EPI_BIAS = 1, EPI_BIAS_SILU = 2 };

/*
 * grid: (N/32, tiles) wavefronts of 64 lanes; M, N multiples of 32, K a
 * multiple of 16.  Row tiles are visited grid-stride, so with Mdev (a device
 * row count below M) the launch can be sized to the machine and the work
 * follows the rows actually present.  Outputs: C (fp32, ldc) and, when Cb
 * is set, a BF16 copy of the activated values; Z (fp32) receives the
 * pre-activation.  CONTIG: as1 == bs1 == 1, so each lane's k octet is one
 * 16-byte load per operand.
 */
template <bool CONTIG>
__global__ void __launch_bounds__(64)
k_gemm_bf16(const unsigned short *__restrict__ A, int64_t as0, int64_t as1,
            const unsigned short *__restrict__ B, int64_t bs0, int64_t bs1,
            const float *__restrict__ bias, float *__restrict__ Z, float *__restrict__ C,
            unsigned short *__restrict__ Cb, int ldc, int M, int N, int K, int epi,
            const int *__restrict__ Mdev)
{
    const int l = threadIdx.x;
    const int j0 = blockIdx.x * 32;
    if (Mdev) M = min(M, *Mdev);
    const int r = l & 31, kq = (l >> 5) * 8;
    for (int i0 = blockIdx.y * 32; i0 < M; i0 += gridDim.y * 32) {
        /* with a bias the accumulator starts from it (the column is the
         * lane's, the same for all 16 values) */
        const float acc0 = (epi != EPI_NONE && j0 + r < N) ? bias[j0 + r] : 0.0f;
        f32x16 acc;
#pragma unroll
        for (int v = 0; v < 16; v++) acc[v] = acc0;
        const unsigned short *pa = A + (int64_t)(i0 + r) * as0;
        const unsigned short *pb = B + (int64_t)(j0 + r) * bs0;
        for (int k0 = 0; k0 < K; k0 += 16) {
            bf16x8 a, b;
            if (CONTIG) {
                a = *reinterpret_cast<const bf16x8 *>(pa + k0 + kq);
                b = *reinterpret_cast<const bf16x8 *>(pb + k0 + kq);
            } else {
#pragma unroll
                for (int q = 0; q < 8; q++) {
                    a[q] = __builtin_bit_cast(__bf16, pa[(int64_t)(k0 + kq + q) * as1]);
                    b[q] = __builtin_bit_cast(__bf16, pb[(int64_t)(k0 + kq + q) * bs1]);
                }
            }
            acc = mfma16(a, b, acc);
        }
#pragma unroll
        for (int v = 0; v < 16; v++) {
            const int i = i0 + (v >> 2) * 8 + (l >> 5) * 4 + (v & 3);
            const int j = j0 + r;
            if (i >= M || j >= N) continue;
            const float z = acc[v];
            if (Z) Z[(int64_t)i * ldc + j] = z;
            /* with EPI_BIAS_SILU, Cb gets hidden_out(z) */
            const float y = (epi == EPI_BIAS_SILU) ? silu(z) : z;
            if (C) C[(int64_t)i * ldc + j] = y;
            if (Cb) Cb[(int64_t)i * ldc + j] = f2bf(y);
        }
    }
}

__global__ void k_f32_to_bf16(const float *__restrict__ x, unsigned short *__restrict__ y, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = f2bf(x[i]);
}

/* dZ = dY * silu'(Z), written as BF16 (next GEMM operand) and FP32 */
__global__ void k_silu_back(const float *__restrict__ dY, const float *__restrict__ Z,
                            float *__restrict__ dZ, unsigned short *__restrict__ dZb, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float z = Z[i];
    const float s = 1.0f / (1.0f + __expf(-z));
    const float g = dY[i] * (s * (1.0f + z * (1.0f - s)));
    dZ[i] = g;
    dZb[i] = f2bf(g);
}

/* two-stage column sums, deterministic: part[p][j] = sum over the rows of
 * chunk p of wt[i] * X[i][j] (X fp32 or BF16, wt optional), block = 64
 * columns x 4 row lanes; then out[j] = sum_p part[p][j] in order */
__global__ void __launch_bounds__(256)
k_colsum_part(const float *__restrict__ Xf, const unsigned short *__restrict__ Xb,
              const float *__restrict__ wt, float *__restrict__ part, int M, int N, int chunk)
{
    __shared__ float red[4][64];
    const int cx = threadIdx.x & 63, ry = threadIdx.x >> 6;
    const int j = blockIdx.x * 64 + cx;
    const int i0 = blockIdx.y * chunk, i1 = min(M, i0 + chunk);
    float s = 0.0f;
    if (j < N) {
        for (int i = i0 + ry; i < i1; i += 4) {
            const float x = Xf ? Xf[(int64_t)i * N + j] : bf2f(Xb[(int64_t)i * N + j]);
            s += wt ? wt[i] * x : x;
        }
    }
    red[ry][cx] = s;
    __syncthreads();
    if (ry == 0 && j < N)
        part[(int64_t)blockIdx.y * N + j] = ((red[0][cx] + red[1][cx]) + red[2][cx]) + red[3][cx];
}

__global__ void k_colsum_final(const float *__restrict__ part, float *__restrict__ out, int P, int N)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= N) return;
    float s = 0.0f;
    for (int p = 0; p < P; p++) s += part[(int64_t)p * N + j];
    out[j] = s;
}

constexpr int COLSUM_PARTS = 64;

/* out[j] = sum_i wt[i] X[i][j] for an M x N matrix; part holds
 * COLSUM_PARTS x N floats */
int launch_colsum(const float *Xf, const unsigned short *Xb, const float *wt, float *part, float *out,
                  int M, int N, hipStream_t s)
{
    const int P = std::max(1, std::min(COLSUM_PARTS, (M + 127) / 128));
    const int chunk = (M + P - 1) / P;
    hipLaunchKernelGGL(k_colsum_part, dim3((unsigned)((N + 63) / 64), (unsigned)P), dim3(256), 0, s,
                       Xf, Xb, wt, part, M, N, chunk);
    NGMI_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_colsum_final, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, part, out, P,
                       N);
    NGMI_HIP(hipGetLastError());
    return NGMI_OK;
}

/* Adam on FP32 master weights, refreshing the BF16 copy */
__global__ void k_adam(float *__restrict__ w, unsigned short *__restrict__ wb,
                       const float *__restrict__ g, float *__restrict__ m, float *__restrict__ v,
                       int64_t n, float lr, float b1, float b2, float eps, float c1, float c2)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[i] = wi;
    if (wb) wb[i] = f2bf(wi);
}

/* output layer (256 -> 1) forward: y[i] = b + the in-order sum of the
 * blocks' dot products over MLP_HEAD_BLOCK hidden units, each accumulated
 * in order with FMA (k_guess_net sums the same way); and the MSE gradient
 * dy = 2 (y - t) / M.  K is a multiple of MLP_HEAD_BLOCK. */
__global__ void k_head(const unsigned short *__restrict__ H, const unsigned short *__restrict__ w,
                       const float *__restrict__ b, const float *__restrict__ target,
                       float *__restrict__ y, float *__restrict__ dy, int M, int K,
                       const int *__restrict__ Mdev)
{
    if (Mdev) M = min(M, *Mdev);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < M; i += gridDim.x * blockDim.x) {
        float s = b[0];
        for (int k0 = 0; k0 < K; k0 += MLP_HEAD_BLOCK) {
            float p = 0.0f;
            for (int k = k0; k < k0 + MLP_HEAD_BLOCK; k++) p = fmaf(bf2f(H[(int64_t)i * K + k]), bf2f(w[k]), p);
            s += p;
        }
        y[i] = s;
        if (target) dy[i] = 2.0f * (s - target[i]) / (float)M;
    }
}

/* output layer backward, input gradient: dH[i][k] = dy[i] w[k] (fp32) */
__global__ void k_head_back(const unsigned short *__restrict__ w, const float *__restrict__ dy,
                            float *__restrict__ dH, int M, int K)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)M * K) return;
    const int i = (int)(e / K), k = (int)(e - (int64_t)i * K);
    dH[e] = dy[i] * bf2f(w[k]);
}

/* ---------------- fused inference for the transient ---------------- */

constexpr int GN_THREADS = 512;                 /* 8 waves: wave w owns hidden units 32w .. 32w + 31 */
constexpr int GN_XLD = MLP_IN_PAD + 8;          /* LDS row strides in elements, padded by 16 bytes */
constexpr int GN_HLD = MLP_HID + 8;
constexpr int GN_HEAD_BLOCKS = MLP_HID / MLP_HEAD_BLOCK;
static_assert(GN_THREADS == 32 * MLP_IN, "the input tile is packed one element per thread");
static_assert(GN_THREADS / 64 * 32 == MLP_HID, "one wave per 32 hidden units");

/* one hidden layer for a tile of 32 rows: A (32 x 16 STEPS, BF16, LDS) times
 * this wave's register-resident weight octets, written to out (LDS) as
 * BF16 activations of columns col0 .. col0 + 31.  The matrix-core steps and
 * the epilogue are those of k_gemm_bf16 (the accumulator starts at the
 * bias). */
template <int STEPS>
__device__ __forceinline__ void gn_layer(const unsigned short *A, int lda, const bf16x8 (&w)[STEPS], float bias,
                                         unsigned short *out, int col0, int l)
{
    const unsigned short *pa = A + (l & 31) * lda + (l >> 5) * 8;
    f32x16 acc;
#pragma unroll
    for (int v = 0; v < 16; v++) acc[v] = bias;
#pragma unroll
    for (int s = 0; s < STEPS; s++) acc = mfma16(*reinterpret_cast<const bf16x8 *>(pa + 16 * s), w[s], acc);
#pragma unroll
    for (int v = 0; v < 16; v++) {
        const int i = (v >> 2) * 8 + (l >> 5) * 4 + (v & 3);
        out[i * GN_HLD + col0 + (l & 31)] = hidden_out(acc[v]);
    }
}

/*
 * Feature rows -> network -> starting values, one launch.  count[0] holds
 * the number of rows (at most maxRows), count[1] is a finish counter: the
 * last workgroup to finish clears both, so the next pass's features kernel
 * starts from zero without a separate launch.  Without count the rows are
 * maxRows and no counter is touched.  Each workgroup loads the
 * BF16 weights once into registers (8 + 64 + 64 VGPRs a lane) and walks
 * tiles of 32 rows grid-stride; per tile the packed input and the hidden
 * activations stay in LDS.  Row r's node value x[dest[r]] becomes
 * guess_value(base[r], y) (and yout[r] = y when yout is set); rows_total,
 * when set, accumulates the rows processed.
 */
__global__ void __launch_bounds__(GN_THREADS)
k_guess_net(const float *__restrict__ X, const unsigned short *__restrict__ wb, const float *__restrict__ wf,
            int *count, int maxRows, const int64_t *__restrict__ dest, const double2 *__restrict__ base,
            double *__restrict__ x, float *__restrict__ yout, unsigned long long *rows_total)
{
    __shared__ __align__(16) unsigned short xs[32 * GN_XLD];
    __shared__ __align__(16) unsigned short ha[32 * GN_HLD];
    __shared__ __align__(16) unsigned short hb[32 * GN_HLD];
    __shared__ float w4s[MLP_HID];
    __shared__ float ps[32][GN_HEAD_BLOCKS + 1];
    const int tid = threadIdx.x, l = tid & 63, col0 = (tid >> 6) * 32;
    const int M = count ? min(count[0], maxRows) : maxRows;
    if ((int64_t)blockIdx.x * 32 < M) {
        const int j = col0 + (l & 31), kh = (l >> 5) * 8;
        bf16x8 w1[MLP_IN_PAD / 16], w2[MLP_HID / 16], w3[MLP_HID / 16];
        const unsigned short *p1 = wb + MLP_OFF_W1 + (int64_t)j * MLP_IN_PAD + kh;
        const unsigned short *p2 = wb + MLP_OFF_W2 + (int64_t)j * MLP_HID + kh;
        const unsigned short *p3 = wb + MLP_OFF_W3 + (int64_t)j * MLP_HID + kh;
#pragma unroll
        for (int s = 0; s < MLP_IN_PAD / 16; s++) w1[s] = *reinterpret_cast<const bf16x8 *>(p1 + 16 * s);
#pragma unroll
        for (int s = 0; s < MLP_HID / 16; s++) {
            w2[s] = *reinterpret_cast<const bf16x8 *>(p2 + 16 * s);
            w3[s] = *reinterpret_cast<const bf16x8 *>(p3 + 16 * s);
        }
        const float b1 = wf[MLP_OFF_B1 + j], b2 = wf[MLP_OFF_B2 + j], b3 = wf[MLP_OFF_B3 + j];
        const float b4 = wf[MLP_OFF_B4];
        for (int e = tid; e < MLP_HID; e += GN_THREADS) w4s[e] = bf2f(wb[MLP_OFF_W4 + e]);
        for (int e = tid; e < 32 * GN_XLD; e += GN_THREADS) xs[e] = 0;   /* the zero columns stay */
        __syncthreads();
        /* this thread's input element (row tid / 16, feature tid % 16) of
         * the tile, loaded one tile ahead so its latency hides behind the
         * previous tile's layers */
        const int rr = tid >> 4, cc = tid & 15;
        const int64_t step = (int64_t)gridDim.x * 32;
        int64_t i0 = (int64_t)blockIdx.x * 32;
        float xin = i0 + rr < M ? X[(i0 + rr) * MLP_IN + cc] : 0.0f;
        for (; i0 < M; i0 += step) {
            xs[rr * GN_XLD + cc] = f2bf(xin);
            xin = i0 + step + rr < M ? X[(i0 + step + rr) * MLP_IN + cc] : 0.0f;
            /* the apply's operands, loaded now for use after the layers */
            const bool mine = tid < 32 && i0 + tid < M;
            const int64_t d_row = (mine && x) ? dest[i0 + tid] : 0;
            const double2 b_row = (mine && x) ? base[i0 + tid] : make_double2(0.0, 0.0);
            __syncthreads();
            gn_layer(xs, GN_XLD, w1, b1, ha, col0, l);
            __syncthreads();
            gn_layer(ha, GN_HLD, w2, b2, hb, col0, l);
            __syncthreads();
            gn_layer(hb, GN_HLD, w3, b3, ha, col0, l);
            __syncthreads();
            if (tid < 32 * GN_HEAD_BLOCKS) {
                const int rr = tid / GN_HEAD_BLOCKS, cb = tid % GN_HEAD_BLOCKS;
                const uint4 *h = reinterpret_cast<const uint4 *>(ha + rr * GN_HLD + cb * MLP_HEAD_BLOCK);
                const float *w4 = w4s + cb * MLP_HEAD_BLOCK;
                float p = 0.0f;
#pragma unroll
                for (int q = 0; q < MLP_HEAD_BLOCK / 8; q++) {
                    const uint4 u = h[q];
                    const unsigned int wd[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
                    for (int e = 0; e < 4; e++) {
                        p = fmaf(bf2f((unsigned short)(wd[e] & 0xffffu)), w4[8 * q + 2 * e], p);
                        p = fmaf(bf2f((unsigned short)(wd[e] >> 16)), w4[8 * q + 2 * e + 1], p);
                    }
                }
                ps[rr][cb] = p;
            }
            __syncthreads();
            if (mine) {
                float y = b4;
#pragma unroll
                for (int cb = 0; cb < GN_HEAD_BLOCKS; cb++) y += ps[tid][cb];
                if (yout) yout[i0 + tid] = y;
                if (x) x[d_row] = guess_value(b_row, y);
            }
        }
    }
    if (!count) return;
    __syncthreads();   /* every wave has read count[0] */
    if (tid == 0) {
        __threadfence();
        if (atomicAdd(reinterpret_cast<unsigned int *>(count + 1), 1u) == gridDim.x - 1) {
            if (rows_total) atomicAdd(rows_total, (unsigned long long)M);
            count[0] = 0;
            count[1] = 0;
        }
    }
}

/* The same network with the weights streamed from L2 per tile instead of
 * held in registers: 4 waves, wave w owning hidden units 64w .. 64w + 63 as
 * two 32-column groups, each group's weights for one layer loaded just
 * before its matrix-core steps.  About 100 VGPRs a lane, so a workgroup
 * fits where one BSIM4 workgroup of the other stream's half finished
 * instead of needing a whole compute unit.  Every output column takes the
 * same matrix-core steps in the same order as in k_guess_net, so both give
 * the same values bit for bit. */
constexpr int GL_THREADS = 256;
static_assert(GL_THREADS / 64 * 64 == MLP_HID, "one wave per 64 hidden units");
static_assert(GL_THREADS == 32 * GN_HEAD_BLOCKS, "one head partial per thread");

template <int STEPS>
__device__ __forceinline__ void gl_layer(const unsigned short *A, int lda, const unsigned short *__restrict__ W,
                                         int ldw, const float *__restrict__ bias, unsigned short *out, int col0,
                                         int l)
{
#pragma unroll
    for (int g = 0; g < 2; g++) {
        const int j = col0 + 32 * g + (l & 31);
        const unsigned short *pw = W + (int64_t)j * ldw + (l >> 5) * 8;
        bf16x8 w[STEPS];
#pragma unroll
        for (int q = 0; q < STEPS; q++) w[q] = *reinterpret_cast<const bf16x8 *>(pw + 16 * q);
        gn_layer(A, lda, w, bias[j], out, col0 + 32 * g, l);
    }
}

__global__ void __launch_bounds__(GL_THREADS) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_guess_net_lite(const float *__restrict__ X, const unsigned short *__restrict__ wb, const float *__restrict__ wf,
                 int *count, int maxRows, const int64_t *__restrict__ dest, const double2 *__restrict__ base,
                 double *__restrict__ x, float *__restrict__ yout, unsigned long long *rows_total)
{
    __shared__ __align__(16) unsigned short xs[32 * GN_XLD];
    __shared__ __align__(16) unsigned short ha[32 * GN_HLD];
    __shared__ __align__(16) unsigned short hb[32 * GN_HLD];
    __shared__ float w4s[MLP_HID];
    __shared__ float ps[32][GN_HEAD_BLOCKS + 1];
    const int tid = threadIdx.x, l = tid & 63, col0 = (tid >> 6) * 64;
    const int M = count ? min(count[0], maxRows) : maxRows;
    if ((int64_t)blockIdx.x * 32 < M) {
        const float b4 = wf[MLP_OFF_B4];
        for (int e = tid; e < MLP_HID; e += GL_THREADS) w4s[e] = bf2f(wb[MLP_OFF_W4 + e]);
        for (int e = tid; e < 32 * GN_XLD; e += GL_THREADS) xs[e] = 0;   /* the zero columns stay */
        __syncthreads();
        /* this thread's two input elements (rows rr and rr + 16, feature
         * cc), loaded one tile ahead */
        const int rr = tid >> 4, cc = tid & 15;
        const int64_t step = (int64_t)gridDim.x * 32;
        int64_t i0 = (int64_t)blockIdx.x * 32;
        float xa = i0 + rr < M ? X[(i0 + rr) * MLP_IN + cc] : 0.0f;
        float xb = i0 + rr + 16 < M ? X[(i0 + rr + 16) * MLP_IN + cc] : 0.0f;
        for (; i0 < M; i0 += step) {
            xs[rr * GN_XLD + cc] = f2bf(xa);
            xs[(rr + 16) * GN_XLD + cc] = f2bf(xb);
            xa = i0 + step + rr < M ? X[(i0 + step + rr) * MLP_IN + cc] : 0.0f;
            xb = i0 + step + rr + 16 < M ? X[(i0 + step + rr + 16) * MLP_IN + cc] : 0.0f;
            const bool mine = tid < 32 && i0 + tid < M;
            const int64_t d_row = (mine && x) ? dest[i0 + tid] : 0;
            const double2 b_row = (mine && x) ? base[i0 + tid] : make_double2(0.0, 0.0);
            /* the weight pointers made opaque per tile, so the compiler does
             * not hoist every layer's weights out of the loop into registers */
            const unsigned short *wt = wb;
            const float *ft = wf;
            asm volatile("" : "+s"(wt), "+s"(ft));
            __syncthreads();
            gl_layer<MLP_IN_PAD / 16>(xs, GN_XLD, wt + MLP_OFF_W1, MLP_IN_PAD, ft + MLP_OFF_B1, ha, col0, l);
            __syncthreads();
            gl_layer<MLP_HID / 16>(ha, GN_HLD, wt + MLP_OFF_W2, MLP_HID, ft + MLP_OFF_B2, hb, col0, l);
            __syncthreads();
            gl_layer<MLP_HID / 16>(hb, GN_HLD, wt + MLP_OFF_W3, MLP_HID, ft + MLP_OFF_B3, ha, col0, l);
            __syncthreads();
            {
                const int hr = tid / GN_HEAD_BLOCKS, cb = tid % GN_HEAD_BLOCKS;
                const uint4 *h = reinterpret_cast<const uint4 *>(ha + hr * GN_HLD + cb * MLP_HEAD_BLOCK);
                const float *w4 = w4s + cb * MLP_HEAD_BLOCK;
                float p = 0.0f;
#pragma unroll
                for (int q = 0; q < MLP_HEAD_BLOCK / 8; q++) {
                    const uint4 u = h[q];
                    const unsigned int wd[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
                    for (int e = 0; e < 4; e++) {
                        p = fmaf(bf2f((unsigned short)(wd[e] & 0xffffu)), w4[8 * q + 2 * e], p);
                        p = fmaf(bf2f((unsigned short)(wd[e] >> 16)), w4[8 * q + 2 * e + 1], p);
                    }
                }
                ps[hr][cb] = p;
            }
            __syncthreads();
            if (mine) {
                float y = b4;
#pragma unroll
                for (int cb = 0; cb < GN_HEAD_BLOCKS; cb++) y += ps[tid][cb];
                if (yout) yout[i0 + tid] = y;
                if (x) x[d_row] = guess_value(b_row, y);
            }
        }
    }
    if (!count) return;
    __syncthreads();   /* every wave has read count[0] */
    if (tid == 0) {
        __threadfence();
        if (atomicAdd(reinterpret_cast<unsigned int *>(count + 1), 1u) == gridDim.x - 1) {
            if (rows_total) atomicAdd(rows_total, (unsigned long long)M);
            count[0] = 0;
            count[1] = 0;
        }
    }
}

int launch_guess_net(const unsigned short *wb, const float *wf, const float *X, int *count, int maxRows,
                     const int64_t *dest, const double2 *base, double *x, float *yout,
                     unsigned long long *rows_total, hipStream_t s)
{
    static int cus = 0;
    if (!cus) {
        int dev = 0;
        NGMI_HIP(hipGetDevice(&dev));
        NGMI_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        if (cus <= 0) cus = 256;
    }
    /* guess_grid workgroups (each holds a whole compute unit's registers),
     * 0 = one per compute unit: 64 leave the other compute units to the
     * BSIM4 kernel of the other stream's half (measured 0.933x plain
     * against 0.923x with all 256) */
    const int tiles = std::max(1, (maxRows + 31) / 32);
    const int want = tune_get(TUNE_GUESS_GRID);
    if (tune_get(TUNE_GUESS_FUSED) == 2) {
        /* streamed weights: guess_grid small workgroups (0 = one per
         * compute unit) */
        const int grid = std::min(tiles, want > 0 ? want : cus);
        hipLaunchKernelGGL(k_guess_net_lite, dim3((unsigned)grid), dim3(GL_THREADS), 0, s, X, wb, wf, count,
                           maxRows, dest, base, x, yout, rows_total);
        NGMI_HIP(hipGetLastError());
        return NGMI_OK;
    }
    const int grid = std::min(tiles, want > 0 ? want : cus);
    hipLaunchKernelGGL(k_guess_net, dim3((unsigned)grid), dim3(GN_THREADS), 0, s, X, wb, wf, count, maxRows,
                       dest, base, x, yout, rows_total);
    NGMI_HIP(hipGetLastError());
    return NGMI_OK;
}

int launch_gemm_bf16(const unsigned short *A, int64_t as0, int64_t as1, const unsigned short *B,
                     int64_t bs0, int64_t bs1, const float *bias, float *Z, float *C,
                     unsigned short *Cb, int ldc, int M, int N, int K, int epi, hipStream_t s,
                     const int *Mdev)
{
    if (M % 32 || N % 32 || K % 16) {
        set_error("gemm shape %d x %d x %d not a multiple of 32 x 32 x 16", M, N, K);
        return NGMI_E_PARMVAL;
    }
    /* with a device row count the grid is sized to the machine (8 XCDs x
     * 32 CUs x a few waves per tile column) and strides over the tiles */
    const int tiles = Mdev ? std::min(M / 32, 2048) : M / 32;
    if (as1 == 1 && bs1 == 1)
        hipLaunchKernelGGL(k_gemm_bf16<true>, dim3(N / 32, tiles), dim3(64), 0, s, A, as0, as1, B, bs0,
                           bs1, bias, Z, C, Cb, ldc, M, N, K, epi, Mdev);
    else
        hipLaunchKernelGGL(k_gemm_bf16<false>, dim3(N / 32, tiles), dim3(64), 0, s, A, as0, as1, B,
                           bs0, bs1, bias, Z, C, Cb, ldc, M, N, K, epi, Mdev);
    NGMI_HIP(hipGetLastError());
    return NGMI_OK;
}

}  // namespace ngmi
