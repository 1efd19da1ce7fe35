/*
 * Error text, device memory, timing and device queries of the C ABI.
 */
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include "ngmi_internal.h"

namespace ngmi {

static thread_local char g_err[512] = "";

static int g_tune[TUNE__COUNT] = {0, 0, 2, 1, 1, 1, 2, 64};

int tune_get(int knob) { return (knob >= 0 && knob < TUNE__COUNT) ? g_tune[knob] : 0; }

int waves_for(int knob, int64_t instances)
{
    const int w = tune_get(knob);
    if (w > 0) return w;
    static int lanes[64] = {0};   /* per device: compute units x 4 SIMDs x 64 lanes */
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 2;
    if (!lanes[dev]) {
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            return 2;
        lanes[dev] = cus * 4 * 64;
    }
    return instances <= lanes[dev] ? 1 : 2;
}

void set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

}  // namespace ngmi

using namespace ngmi;

extern "C" {

const char *ngmi_last_error(void) { return g_err; }

static int knob_index(const char *knob)
{
    if (!knob) return -1;
    if (!std::strcmp(knob, "eval_waves")) return TUNE_EVAL_WAVES;
    if (!std::strcmp(knob, "mos_waves")) return TUNE_MOS_WAVES;
    if (!std::strcmp(knob, "streams")) return TUNE_STREAMS;
    if (!std::strcmp(knob, "specialise")) return TUNE_SPECIALISE;
    if (!std::strcmp(knob, "order_slack")) return TUNE_ORDER_SLACK;
    if (!std::strcmp(knob, "graph")) return TUNE_GRAPH;
    if (!std::strcmp(knob, "guess_fused")) return TUNE_GUESS_FUSED;
    if (!std::strcmp(knob, "guess_grid")) return TUNE_GUESS_GRID;
    return -1;
}

int ngmi_tune_get(const char *knob)
{
    const int k = knob_index(knob);
    if (k < 0) {
        set_error("unknown tuning knob '%s'", knob ? knob : "(null)");
        return -NGMI_E_BADPARM;
    }
    return tune_get(k);
}

int ngmi_tune(const char *knob, int value)
{
    if (!knob) return NGMI_E_PANIC;
    const int k = knob_index(knob);
    if (k < 0) {
        set_error("unknown tuning knob '%s'", knob);
        return NGMI_E_BADPARM;
    }
    const bool flag = k == TUNE_SPECIALISE || k == TUNE_GRAPH || k == TUNE_GUESS_FUSED;
    const int vmin = (flag || k == TUNE_ORDER_SLACK || k == TUNE_GUESS_GRID || k == TUNE_MOS_WAVES || k == TUNE_EVAL_WAVES) ? 0 : 1;
    const int vmax = (k == TUNE_GUESS_FUSED) ? 2 : flag ? 1 : (k == TUNE_GUESS_GRID) ? 1024 : 4;
    if (value < vmin || value > vmax) {
        set_error("knob '%s' takes %d..%d", knob, vmin, vmax);
        return NGMI_E_PARMVAL;
    }
    g_tune[k] = value;
    return NGMI_OK;
}

void *ngmi_dev_alloc(size_t bytes)
{
    void *p = nullptr;
    if (bytes == 0) bytes = 8;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
        set_error("hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
        return nullptr;
    }
    return p;
}

int ngmi_dev_free(void *p)
{
    if (p) NGMI_HIP(hipFree(p));
    return NGMI_OK;
}

int ngmi_copy_h2d(void *dst, const void *src, size_t bytes)
{
    NGMI_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return NGMI_OK;
}

int ngmi_copy_d2h(void *dst, const void *src, size_t bytes)
{
    NGMI_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return NGMI_OK;
}

int ngmi_dev_memset(void *dst, int byte, size_t bytes)
{
    NGMI_HIP(hipMemset(dst, byte, bytes));
    return NGMI_OK;
}

int ngmi_sync(void)
{
    NGMI_HIP(hipDeviceSynchronize());
    return NGMI_OK;
}

int ngmi_device_info(char *name, int name_len, int *cus, int *clock_khz)
{
    int dev = 0;
    NGMI_HIP(hipGetDevice(&dev));
    hipDeviceProp_t p;
    NGMI_HIP(hipGetDeviceProperties(&p, dev));
    if (name && name_len > 0) {
        std::snprintf(name, (size_t)name_len, "%s (%s)", p.name, p.gcnArchName);
    }
    if (cus) *cus = p.multiProcessorCount;
    if (clock_khz) *clock_khz = p.clockRate;
    return NGMI_OK;
}

struct ngmi_timer {
    hipEvent_t a, b;
};

ngmi_timer *ngmi_timer_new(void)
{
    ngmi_timer *t = new ngmi_timer;
    if (hipEventCreate(&t->a) != hipSuccess || hipEventCreate(&t->b) != hipSuccess) {
        set_error("hipEventCreate failed");
        delete t;
        return nullptr;
    }
    return t;
}

void ngmi_timer_free(ngmi_timer *t)
{
    if (!t) return;
    (void)hipEventDestroy(t->a);
    (void)hipEventDestroy(t->b);
    delete t;
}

int ngmi_timer_start(ngmi_timer *t, void *stream)
{
    NGMI_HIP(hipEventRecord(t->a, (hipStream_t)stream));
    return NGMI_OK;
}

int ngmi_timer_stop(ngmi_timer *t, void *stream)
{
    NGMI_HIP(hipEventRecord(t->b, (hipStream_t)stream));
    return NGMI_OK;
}

double ngmi_timer_ms(ngmi_timer *t)
{
    float ms = -1.0f;
    if (hipEventSynchronize(t->b) != hipSuccess) return -1.0;
    if (hipEventElapsedTime(&ms, t->a, t->b) != hipSuccess) return -1.0;
    return (double)ms;
}

}  // extern "C"
