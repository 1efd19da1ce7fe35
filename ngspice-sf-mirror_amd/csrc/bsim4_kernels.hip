/*
 * BSIM4 evaluation kernels: one thread per MOSFET instance, FP64.
 *
 * Layout: every per-instance quantity is a row of an n-strided array
 * (row r of instance i at [r * n + i]), so lane i of a wavefront touches
 * consecutive doubles of each row and every load/store is coalesced.
 * Model-constant table fields are read from a per-model column (a broadcast
 * within the wavefront), per-instance fields from the instance column.
 */
#include "ngmi_internal.h"
#include "bsim4_eval.h"

namespace ngmi {

struct DevPrm {
    const B4Tab &t;
    int64_t i;
    int mi;
    __device__ __forceinline__ double operator()(int f) const
    {
        const int c = t.code[f];
        return c >= 0 ? t.itab[(int64_t)c * t.n + i] : t.mtab[(int64_t)mi * t.mcols + (~c)];
    }
};

/* states: level l row k of instance i at lv[l][k * n + i] */
struct DevState {
    double *lv[3];
    int64_t n, i;
    __device__ __forceinline__ double s(int l, int k) const { return lv[l][k * n + i]; }
    __device__ __forceinline__ void set(int l, int k, double v) { lv[l][k * n + i] = v; }
};

struct DevCache {
    double *c_;
    int64_t n, i;
    __device__ __forceinline__ double c(int k) const { return c_[k * n + i]; }
    __device__ __forceinline__ void setc(int k, double v) { c_[k * n + i] = v; }
};

struct EvalIO {
    const double *vterm;
    double *st0, *st1, *st2;
    double *cache, *stamp, *rhs;
    int *noncon;
};

/* WAVES = wavefronts per SIMD the register allocation aims for: 1 gives the
 * whole evaluation its registers, 2..4 trade spills for latency hiding
 * (chosen per workload, see ngmi_tune).  FL >= 0: the selector word every
 * instance shares, compiled as a constant. */
template <int WAVES, int FL>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WAVES, WAVES)))
k_bsim4_eval(B4Tab tab, B4Ckt ck, EvalIO io)
{
    const int64_t n = tab.n;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int mi = tab.model_of[i];
    DevPrm p{tab, i, mi};
    DevState st{{io.st0, io.st1, io.st2}, n, i};
    DevCache ca{io.cache, n, i};
    double v[4];
#pragma unroll
    for (int k = 0; k < 4; k++) v[k] = io.vterm[k * n + i];
    double J[B4S__COUNT], R[B4R__COUNT];
    const int nc = b4_load<FL>(p, st, ca, ck, v, J, R);
#pragma unroll
    for (int k = 0; k < B4S__COUNT; k++) io.stamp[k * n + i] = J[k];
#pragma unroll
    for (int k = 0; k < B4R__COUNT; k++) io.rhs[k * n + i] = R[k];
    io.noncon[i] = nc;
}

int launch_bsim4_eval(const B4Tab &tab, const B4Ckt &ck, const double *vterm,
                      double *const state[3], double *cache, double *stamp, double *rhs,
                      int *noncon, hipStream_t stream)
{
    if (tab.n == 0) return NGMI_OK;
    EvalIO io{vterm, state[0], state[1], state[2], cache, stamp, rhs, noncon};
    const int bs = 256;
    const int64_t nb = (tab.n + bs - 1) / bs;
    const bool dflt = tab.flags == B4FL_DEFAULT && tune_get(TUNE_SPECIALISE);
#define NGMI_EVAL(W, F) \
    hipLaunchKernelGGL((k_bsim4_eval<W, F>), dim3((unsigned)nb), dim3(bs), 0, stream, tab, ck, io)
    switch (waves_for(TUNE_EVAL_WAVES, tab.n)) {
    case 1: if (dflt) NGMI_EVAL(1, B4FL_DEFAULT); else NGMI_EVAL(1, -1); break;
    case 3: if (dflt) NGMI_EVAL(3, B4FL_DEFAULT); else NGMI_EVAL(3, -1); break;
    case 4: if (dflt) NGMI_EVAL(4, B4FL_DEFAULT); else NGMI_EVAL(4, -1); break;
    default: if (dflt) NGMI_EVAL(2, B4FL_DEFAULT); else NGMI_EVAL(2, -1); break;
    }
#undef NGMI_EVAL
    NGMI_HIP(hipGetLastError());
    return NGMI_OK;
}

}  // namespace ngmi
