/*
 * Newton initial-guess network: a 16 -> 256 -> 256 -> 256 -> 1 MLP with SiLU,
 * BF16 weights and activations on the matrix cores (v_mfma_f32_32x32x8
 * bf16, FP32 accumulate), FP32 master weights and Adam for training.
 *
 * One kernel computes C[i][j] = sum_k A(i,k) B(j,k) (+ bias[j]) for strided
 * BF16 operands, so the forward pass (X W^T), the input gradient (dY W) and
 * the weight gradient (dY^T X) are the same kernel with different strides.
 * Each wavefront owns a 32 x 32 tile of C; lane l feeds rows/cols l % 32 and
 * the k quad (l / 32) * 4 .. +3 of every 8-deep step, and receives
 * C rows (v / 4) * 8 + (l / 32) * 4 + v % 4, column l % 32, v = 0..15.
 */
#include <algorithm>
#include "ngmi_internal.h"
#include "mlp_types.h"

namespace ngmi {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ unsigned short f2bf(float f)
{   /* round to nearest even */
    unsigned int u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short)(u >> 16 | ((u & 0xffff) ? 0x40 : 0));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float((unsigned int)h << 16); }

__device__ __forceinline__ float silu(float z) { return z / (1.0f + __expf(-z)); }

/* epilogue modes */
enum { EPI_NONE = 0, EPI_BIAS = 1, EPI_BIAS_SILU = 2 };

/*
 * grid: (N/32, tiles) wavefronts of 64 lanes; M, N multiples of 32, K a
 * multiple of 8.  Row tiles are visited grid-stride, so with Mdev (a device
 * row count below M) the launch can be sized to the machine and the work
 * follows the rows actually present.  Outputs: C (fp32, ldc) and, when Cb
 * is set, a BF16 copy of the activated values; Z (fp32) receives the
 * pre-activation.  CONTIG: as1 == bs1 == 1, so each lane's k quad is one
 * 8-byte load per operand.
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
    const int r = l & 31, kq = (l >> 5) * 4;
    for (int i0 = blockIdx.y * 32; i0 < M; i0 += gridDim.y * 32) {
        f32x16 acc;
#pragma unroll
        for (int v = 0; v < 16; v++) acc[v] = 0.0f;
        const unsigned short *pa = A + (int64_t)(i0 + r) * as0;
        const unsigned short *pb = B + (int64_t)(j0 + r) * bs0;
        for (int k0 = 0; k0 < K; k0 += 8) {
            bf16x4 a, b;
            if (CONTIG) {
                a = *reinterpret_cast<const bf16x4 *>(pa + k0 + kq);
                b = *reinterpret_cast<const bf16x4 *>(pb + k0 + kq);
            } else {
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    a[q] = (short)pa[(int64_t)(k0 + kq + q) * as1];
                    b[q] = (short)pb[(int64_t)(k0 + kq + q) * bs1];
                }
            }
            acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
        }
#pragma unroll
        for (int v = 0; v < 16; v++) {
            const int i = i0 + (v >> 2) * 8 + (l >> 5) * 4 + (v & 3);
            const int j = j0 + r;
            if (i >= M || j >= N) continue;
            float z = acc[v];
            if (epi != EPI_NONE) z += bias[j];
            if (Z) Z[(int64_t)i * ldc + j] = z;
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

/* output layer (256 -> 1) forward: y[i] = sum_k h[i][k] w[k] + b, and the
 * MSE gradient dy = 2 (y - t) / M */
__global__ void k_head(const unsigned short *__restrict__ H, const unsigned short *__restrict__ w,
                       const float *__restrict__ b, const float *__restrict__ target,
                       float *__restrict__ y, float *__restrict__ dy, int M, int K,
                       const int *__restrict__ Mdev)
{
    if (Mdev) M = min(M, *Mdev);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < M; i += gridDim.x * blockDim.x) {
        float s = b[0];
        for (int k = 0; k < K; k++) s += bf2f(H[(int64_t)i * K + k]) * bf2f(w[k]);
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

int launch_gemm_bf16(const unsigned short *A, int64_t as0, int64_t as1, const unsigned short *B,
                     int64_t bs0, int64_t bs1, const float *bias, float *Z, float *C,
                     unsigned short *Cb, int ldc, int M, int N, int K, int epi, hipStream_t s,
                     const int *Mdev)
{
    if (M % 32 || N % 32 || K % 8) {
        set_error("gemm shape %d x %d x %d not a multiple of 32 x 32 x 8", M, N, K);
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
