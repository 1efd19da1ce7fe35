/*
 * Host side of the Newton initial-guess network: parameter storage (FP32
 * master weights + BF16 copies on the device), forward pass, and one
 * training step (MSE loss, backward through the three SiLU layers, Adam).
 * The input layer is 32 wide on the device: 16 features and 16 zero
 * columns, so every matrix-core tile is full.
 */
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <random>
#include <vector>
#include "ngmi_internal.h"
#include "mlp_types.h"

namespace ngmi {
int launch_gemm_bf16(const unsigned short *A, int64_t as0, int64_t as1, const unsigned short *B,
                     int64_t bs0, int64_t bs1, const float *bias, float *Z, float *C,
                     unsigned short *Cb, int ldc, int M, int N, int K, int epi, hipStream_t s,
                     const int *Mdev = nullptr);
__global__ void k_f32_to_bf16(const float *x, unsigned short *y, int64_t n);
__global__ void k_silu_back(const float *dY, const float *Z, float *dZ, unsigned short *dZb,
                            int64_t n);
int launch_colsum(const float *Xf, const unsigned short *Xb, const float *wt, float *part, float *out,
                  int M, int N, hipStream_t s);
__global__ void k_adam(float *w, unsigned short *wb, const float *g, float *m, float *v, int64_t n,
                       float lr, float b1, float b2, float eps, float c1, float c2);
__global__ void k_head(const unsigned short *H, const unsigned short *w, const float *b,
                       const float *target, float *y, float *dy, int M, int K, const int *Mdev);
__global__ void k_head_back(const unsigned short *w, const float *dy, float *dH, int M, int K);
int launch_guess_net(const unsigned short *wb, const float *wf, const float *X, int *count, int maxRows,
                     const int64_t *dest, const double2 *base, double *x, float *yout,
                     unsigned long long *rows_total, hipStream_t s);
}  // namespace ngmi

using namespace ngmi;

namespace {
constexpr int IN_PAD = MLP_IN_PAD;
constexpr int H = MLP_HID;
constexpr int64_t OFF_W1 = MLP_OFF_W1, OFF_B1 = MLP_OFF_B1, OFF_W2 = MLP_OFF_W2, OFF_B2 = MLP_OFF_B2,
                  OFF_W3 = MLP_OFF_W3, OFF_B3 = MLP_OFF_B3, OFF_W4 = MLP_OFF_W4, OFF_B4 = MLP_OFF_B4,
                  NPARAM = MLP_NPARAM;
}  // namespace

struct ngmi_mlp {
    float *w = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;   /* NPARAM each */
    unsigned short *wb = nullptr;                                   /* BF16 copy */
    int cap = 0;                                                    /* rows allocated */
    unsigned short *xb = nullptr, *h1 = nullptr, *h2 = nullptr, *h3 = nullptr, *dzb = nullptr;
    float *z1 = nullptr, *z2 = nullptr, *z3 = nullptr, *dh = nullptr, *dz = nullptr;
    float *y = nullptr, *dy = nullptr;
    float *part = nullptr;                                          /* column-sum partials */
    int64_t step = 0;
    /* inference-only buffers (BF16 activations, ping-pong) */
    int icap = 0;
    unsigned short *ixb = nullptr, *ia = nullptr, *ib = nullptr;
    ~ngmi_mlp()
    {
        void *ps[] = {w, g, m, v, wb, xb, h1, h2, h3, dzb, z1, z2, z3, dh, dz, y, dy, part, ixb, ia, ib};
        for (void *p : ps) (void)hipFree(p);
    }
    int ensure_infer(int rows)
    {
        if (rows <= icap) return NGMI_OK;
        void *ps[] = {ixb, ia, ib};
        for (void *p : ps) (void)hipFree(p);
        NGMI_HIP(hipMalloc(&ixb, (size_t)rows * 32 * 2));
        NGMI_HIP(hipMalloc(&ia, (size_t)rows * MLP_HID * 2));
        NGMI_HIP(hipMalloc(&ib, (size_t)rows * MLP_HID * 2));
        icap = rows;
        return NGMI_OK;
    }
    int ensure(int rows)
    {
        if (rows <= cap) return NGMI_OK;
        void *ps[] = {xb, h1, h2, h3, dzb, z1, z2, z3, dh, dz, y, dy};
        for (void *p : ps) (void)hipFree(p);
        const size_t r = (size_t)rows;
        NGMI_HIP(hipMalloc(&xb, r * IN_PAD * 2));
        NGMI_HIP(hipMalloc(&h1, r * H * 2));
        NGMI_HIP(hipMalloc(&h2, r * H * 2));
        NGMI_HIP(hipMalloc(&h3, r * H * 2));
        NGMI_HIP(hipMalloc(&dzb, r * H * 2));
        NGMI_HIP(hipMalloc(&z1, r * H * 4));
        NGMI_HIP(hipMalloc(&z2, r * H * 4));
        NGMI_HIP(hipMalloc(&z3, r * H * 4));
        NGMI_HIP(hipMalloc(&dh, r * H * 4));
        NGMI_HIP(hipMalloc(&dz, r * H * 4));
        NGMI_HIP(hipMalloc(&y, r * 4));
        NGMI_HIP(hipMalloc(&dy, r * 4));
        if (!part) NGMI_HIP(hipMalloc(&part, (size_t)64 * H * 4));
        cap = rows;
        return NGMI_OK;
    }
};

namespace {

int blocks(int64_t n) { return (int)((n + 255) / 256); }

/* X (device fp32, M x 16, row-major) -> padded BF16 input rows */
__global__ void k_pack_input(const float *__restrict__ X, unsigned short *__restrict__ xb, int M,
                             int Mpad, const int *__restrict__ Mdev = nullptr)
{
    if (Mdev) {
        M = min(M, *Mdev);
        Mpad = min(Mpad, (M + 31) / 32 * 32);
    }
    const int64_t n = (int64_t)Mpad * IN_PAD;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int r = (int)(i / IN_PAD), c = (int)(i % IN_PAD);
        xb[i] = f2bf((r < M && c < MLP_IN) ? X[(int64_t)r * MLP_IN + c] : 0.0f);
    }
}

int forward(ngmi_mlp *n, const float *X, int M, int Mpad, hipStream_t s)
{
    int rc;
    hipLaunchKernelGGL(k_pack_input, dim3(blocks((int64_t)Mpad * IN_PAD)), dim3(256), 0, s, X, n->xb,
                       M, Mpad);
    NGMI_HIP(hipGetLastError());
    const unsigned short *wb = n->wb;
    /* Z = X W^T + b: A(i,k) = X[i*K + k], B(j,k) = W[j*K + k] */
    if ((rc = launch_gemm_bf16(n->xb, IN_PAD, 1, wb + OFF_W1, IN_PAD, 1, n->w + OFF_B1, n->z1,
                               nullptr, n->h1, H, Mpad, H, IN_PAD, 2, s)))
        return rc;
    if ((rc = launch_gemm_bf16(n->h1, H, 1, wb + OFF_W2, H, 1, n->w + OFF_B2, n->z2, nullptr, n->h2,
                               H, Mpad, H, H, 2, s)))
        return rc;
    if ((rc = launch_gemm_bf16(n->h2, H, 1, wb + OFF_W3, H, 1, n->w + OFF_B3, n->z3, nullptr, n->h3,
                               H, Mpad, H, H, 2, s)))
        return rc;
    return NGMI_OK;
}

}  // namespace

namespace {
/* inference over at most maxRows rows (the first *dcount when dcount is
 * set) with caller-provided work buffers: xb (Mpad x 32 BF16 input rows),
 * ia and ib (Mpad x 256 BF16 activations, ping-pong) */
int forward_infer(const ngmi_mlp *n, const float *X, float *y, int maxRows, int Mpad, unsigned short *xb,
                  unsigned short *ia, unsigned short *ib, const int *dcount, hipStream_t s)
{
    int rc;
    hipLaunchKernelGGL(k_pack_input, dim3(std::min(blocks((int64_t)Mpad * IN_PAD), 2048)), dim3(256), 0, s,
                       X, xb, maxRows, Mpad, dcount);
    NGMI_HIP(hipGetLastError());
    const unsigned short *wb = n->wb;
    if ((rc = launch_gemm_bf16(xb, IN_PAD, 1, wb + OFF_W1, IN_PAD, 1, n->w + OFF_B1, nullptr,
                               nullptr, ia, H, Mpad, H, IN_PAD, 2, s, dcount)))
        return rc;
    if ((rc = launch_gemm_bf16(ia, H, 1, wb + OFF_W2, H, 1, n->w + OFF_B2, nullptr, nullptr, ib,
                               H, Mpad, H, H, 2, s, dcount)))
        return rc;
    if ((rc = launch_gemm_bf16(ib, H, 1, wb + OFF_W3, H, 1, n->w + OFF_B3, nullptr, nullptr, ia,
                               H, Mpad, H, H, 2, s, dcount)))
        return rc;
    hipLaunchKernelGGL(k_head, dim3(std::min(blocks(maxRows), 1024)), dim3(256), 0, s, ia, wb + OFF_W4,
                       n->w + OFF_B4, nullptr, y, nullptr, maxRows, H, dcount);
    NGMI_HIP(hipGetLastError());
    return NGMI_OK;
}

/* Work memory of ngmi_mlp_forward_noalloc: NGMI_MLP_NOALLOC_MAX_ROWS rows of
 * packed BF16 input (32 wide) and two BF16 activation buffers (256 wide),
 * 17 MB, in device memory of every GPU the library runs on until the
 * process ends.  Host code addresses it only through the pointer
 * hipGetSymbolAddress returns for the current device. */
constexpr int64_t NOALLOC_XB = (int64_t)NGMI_MLP_NOALLOC_MAX_ROWS * IN_PAD;
constexpr int64_t NOALLOC_ACT = (int64_t)NGMI_MLP_NOALLOC_MAX_ROWS * H;
}  // namespace

static __device__ __align__(256) unsigned short scratch[NOALLOC_XB + 2 * NOALLOC_ACT];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex scratch_mu;
unsigned short *scratch_ptr[MAX_DEVICES] = {};

/* the device address of this device's scratch, looked up once per device */
int scratch_for_current_device(unsigned short **out)
{
    int dev = 0;
    NGMI_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= MAX_DEVICES) {
        set_error("mlp_forward_noalloc: device %d beyond the %d the library tracks", dev, MAX_DEVICES);
        return NGMI_E_DEVICE;
    }
    std::lock_guard<std::mutex> lk(scratch_mu);
    if (!scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error("mlp_forward_noalloc: hipGetSymbolAddress failed: %s", hipGetErrorString(e));
            return NGMI_E_DEVICE;
        }
        scratch_ptr[dev] = (unsigned short *)p;
    }
    *out = scratch_ptr[dev];
    return NGMI_OK;
}
}  // namespace

namespace ngmi {
/* forward over the first *dcount of at most maxRows rows (device-side count) */
int mlp_forward_rows(ngmi_mlp *n, const float *X, float *y, int maxRows, const int *dcount,
                     hipStream_t s)
{
    const int Mpad = (maxRows + 31) / 32 * 32;
    int rc = n->ensure_infer(Mpad);
    if (rc) return rc;
    return forward_infer(n, X, y, maxRows, Mpad, n->ixb, n->ia, n->ib, dcount, s);
}

/* the fused inference of the transient's guess pass: rows from count[0]
 * (cleared, with the finish counter count[1], by the kernel) */
int mlp_guess_rows(const ngmi_mlp *n, const float *X, int *count, int maxRows, const int64_t *dest,
                   const double2 *base, double *x, float *yout, unsigned long long *rows_total, hipStream_t s)
{
    return launch_guess_net(n->wb, n->w, X, count, maxRows, dest, base, x, yout, rows_total, s);
}
}  // namespace ngmi

extern "C" {

ngmi_mlp *ngmi_mlp_new(unsigned long long seed)
{
    ngmi_mlp *n = new ngmi_mlp;
    std::vector<float> w(NPARAM, 0.0f);
    std::mt19937_64 g(seed);
    auto init = [&](int64_t off, int fan_out, int fan_in, int used_in) {
        const double lim = std::sqrt(6.0 / (fan_in + fan_out));
        std::uniform_real_distribution<double> u(-lim, lim);
        for (int o = 0; o < fan_out; o++)
            for (int i = 0; i < fan_in; i++)
                w[off + (int64_t)o * fan_in + i] = i < used_in ? (float)u(g) : 0.0f;
    };
    init(OFF_W1, H, IN_PAD, MLP_IN);
    init(OFF_W2, H, H, H);
    init(OFF_W3, H, H, H);
    {
        std::uniform_real_distribution<double> u(-1e-3, 1e-3);
        for (int k = 0; k < H; k++) w[OFF_W4 + k] = (float)u(g);
    }
    if (hipMalloc(&n->w, NPARAM * 4) != hipSuccess || hipMalloc(&n->g, NPARAM * 4) != hipSuccess ||
        hipMalloc(&n->m, NPARAM * 4) != hipSuccess || hipMalloc(&n->v, NPARAM * 4) != hipSuccess ||
        hipMalloc(&n->wb, NPARAM * 2) != hipSuccess ||
        hipMemcpy(n->w, w.data(), NPARAM * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(n->m, 0, NPARAM * 4) != hipSuccess || hipMemset(n->v, 0, NPARAM * 4) != hipSuccess ||
        hipMemset(n->g, 0, NPARAM * 4) != hipSuccess) {
        set_error("ngmi_mlp_new: device allocation failed");
        delete n;
        return nullptr;
    }
    hipLaunchKernelGGL(k_f32_to_bf16, dim3(blocks(NPARAM)), dim3(256), 0, 0, n->w, n->wb,
                       (int64_t)NPARAM);
    if (hipDeviceSynchronize() != hipSuccess) {
        delete n;
        return nullptr;
    }
    return n;
}

void ngmi_mlp_free(ngmi_mlp *n) { delete n; }

int64_t ngmi_mlp_num_params(void) { return NPARAM; }

int ngmi_mlp_get_params(const ngmi_mlp *n, float *host, int grads)
{
    if (!n || !host) return NGMI_E_PANIC;
    NGMI_HIP(hipMemcpy(host, grads ? n->g : n->w, NPARAM * 4, hipMemcpyDeviceToHost));
    return NGMI_OK;
}

int ngmi_mlp_set_params(ngmi_mlp *n, const float *host)
{
    if (!n || !host) return NGMI_E_PANIC;
    NGMI_HIP(hipMemcpy(n->w, host, NPARAM * 4, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_f32_to_bf16, dim3(blocks(NPARAM)), dim3(256), 0, 0, n->w, n->wb,
                       (int64_t)NPARAM);
    NGMI_HIP(hipGetLastError());
    NGMI_HIP(hipDeviceSynchronize());
    return NGMI_OK;
}

/* y[M] = network(X[M x 16]); device pointers */
int ngmi_mlp_forward(ngmi_mlp *n, const float *X, float *y, int M, void *stream)
{
    if (!n || M < 0) return NGMI_E_PANIC;
    if (M == 0) return NGMI_OK;
    const hipStream_t s = (hipStream_t)stream;
    if (tune_get(TUNE_GUESS_FUSED))
        return launch_guess_net(n->wb, n->w, X, nullptr, M, nullptr, nullptr, nullptr, y, nullptr, s);
    const int Mpad = (M + 31) / 32 * 32;
    int rc = n->ensure(Mpad);
    if (rc) return rc;
    if ((rc = forward(n, X, M, Mpad, s))) return rc;
    hipLaunchKernelGGL(k_head, dim3(blocks(M)), dim3(256), 0, s, n->h3, n->wb + OFF_W4,
                       n->w + OFF_B4, nullptr, y, nullptr, M, H, nullptr);
    NGMI_HIP(hipGetLastError());
    return NGMI_OK;
}

/* ngmi_mlp_forward without allocating: the work buffers are the library's
 * fixed per-device scratch, so calls that overlap on one GPU race */
int ngmi_mlp_forward_noalloc(const ngmi_mlp *n, const float *X, float *y, int M, void *stream)
{
    if (!n || M < 0 || (M > 0 && (!X || !y))) {
        set_error("mlp_forward_noalloc: bad arguments");
        return NGMI_E_PANIC;
    }
    if (M > NGMI_MLP_NOALLOC_MAX_ROWS) {
        set_error("mlp_forward_noalloc: %d rows exceed the scratch maximum of %d", M, NGMI_MLP_NOALLOC_MAX_ROWS);
        return NGMI_E_PARMVAL;
    }
    if (M == 0) return NGMI_OK;
    unsigned short *base = nullptr;
    int rc = scratch_for_current_device(&base);
    if (rc) return rc;
    const int Mpad = (M + 31) / 32 * 32;
    return forward_infer(n, X, y, M, Mpad, base, base + NOALLOC_XB, base + NOALLOC_XB + NOALLOC_ACT, nullptr,
                         (hipStream_t)stream);
}

/* gradients of the mean squared error over M samples into the gradient
 * buffer; with lr > 0 also one Adam step (beta 0.9 / 0.999, eps 1e-8) */
int ngmi_mlp_train_step(ngmi_mlp *n, const float *X, const float *target, int M, float lr,
                        float *loss_host, void *stream)
{
    if (!n || M <= 0) return NGMI_E_PANIC;
    const hipStream_t s = (hipStream_t)stream;
    const int Mpad = (M + 31) / 32 * 32;
    int rc = n->ensure(Mpad);
    if (rc) return rc;
    if ((rc = forward(n, X, M, Mpad, s))) return rc;
    float *g = n->g;
    const unsigned short *wb = n->wb;
    hipLaunchKernelGGL(k_head, dim3(blocks(M)), dim3(256), 0, s, n->h3, wb + OFF_W4, n->w + OFF_B4,
                       target, n->y, n->dy, M, H, nullptr);
    NGMI_HIP(hipGetLastError());
    /* padded rows carry no gradient */
    if (Mpad > M) NGMI_HIP(hipMemsetAsync(n->dy + M, 0, (size_t)(Mpad - M) * 4, s));
    if ((rc = launch_colsum(nullptr, n->h3, n->dy, n->part, g + OFF_W4, Mpad, H, s))) return rc;
    if ((rc = launch_colsum(n->dy, nullptr, nullptr, n->part, g + OFF_B4, Mpad, 1, s))) return rc;
    hipLaunchKernelGGL(k_head_back, dim3(blocks((int64_t)Mpad * H)), dim3(256), 0, s, wb + OFF_W4, n->dy,
                       n->dh, Mpad, H);
    NGMI_HIP(hipGetLastError());
    struct L { const unsigned short *hin; int kin; float *z; int64_t ow, ob; };
    const L layers[3] = {{n->xb, IN_PAD, n->z1, OFF_W1, OFF_B1},
                         {n->h1, H, n->z2, OFF_W2, OFF_B2},
                         {n->h2, H, n->z3, OFF_W3, OFF_B3}};
    for (int li = 2; li >= 0; li--) {
        const L &Ly = layers[li];
        const int64_t nz = (int64_t)Mpad * H;
        hipLaunchKernelGGL(k_silu_back, dim3(blocks(nz)), dim3(256), 0, s, n->dh, Ly.z, n->dz, n->dzb,
                           nz);
        NGMI_HIP(hipGetLastError());
        /* dW[o][i] = sum_m dZ[m][o] Hin[m][i] */
        if ((rc = launch_gemm_bf16(n->dzb, 1, H, Ly.hin, 1, Ly.kin, nullptr, nullptr, g + Ly.ow,
                                   nullptr, Ly.kin, H, Ly.kin, Mpad, 0, s)))
            return rc;
        if ((rc = launch_colsum(n->dz, nullptr, nullptr, n->part, g + Ly.ob, Mpad, H, s))) return rc;
        if (li > 0) {
            /* dHin[m][i] = sum_o dZ[m][o] W[o][i] */
            if ((rc = launch_gemm_bf16(n->dzb, H, 1, wb + Ly.ow, 1, Ly.kin, nullptr, nullptr, n->dh,
                                       nullptr, Ly.kin, Mpad, Ly.kin, H, 0, s)))
                return rc;
        }
    }
    if (loss_host) {
        std::vector<float> yh(M), th(M);
        NGMI_HIP(hipMemcpyAsync(yh.data(), n->y, (size_t)M * 4, hipMemcpyDeviceToHost, s));
        NGMI_HIP(hipMemcpyAsync(th.data(), target, (size_t)M * 4, hipMemcpyDeviceToHost, s));
        NGMI_HIP(hipStreamSynchronize(s));
        double acc = 0;
        for (int i = 0; i < M; i++) acc += (double)(yh[i] - th[i]) * (yh[i] - th[i]);
        *loss_host = (float)(acc / M);
    }
    if (lr > 0) {
        n->step++;
        const float b1 = 0.9f, b2 = 0.999f;
        const float c1 = 1.0f - std::pow(b1, (float)n->step), c2 = 1.0f - std::pow(b2, (float)n->step);
        hipLaunchKernelGGL(k_adam, dim3(blocks(NPARAM)), dim3(256), 0, s, n->w, n->wb, g, n->m, n->v,
                           (int64_t)NPARAM, lr, b1, b2, 1e-8f, c1, c2);
        NGMI_HIP(hipGetLastError());
    }
    return NGMI_OK;
}

}  // extern "C"
