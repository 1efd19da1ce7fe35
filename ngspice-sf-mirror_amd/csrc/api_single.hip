/*
 * Per-circuit drop-in cores: what a reference-shaped DEVload / SMP binding
 * calls for one circuit whose matrix, right-hand side and state vectors live
 * in host memory (the CKTcircuit of a single ngspice run).
 *
 *   ngmi_devload_*  BSIM4load (b4ld.c:69-5399) for the instances of one
 *                   circuit: gathers CKTrhsOld and each instance's GENstate
 *                   slices (b4set.c:2318) into the engine's instance-strided
 *                   layout on the device, evaluates, scatters the states
 *                   back, and adds the 16 stamps and 4 right-hand-side values
 *                   of every instance through its bound matrix positions
 *                   (BSIM4bindCSC, b4bindCSC.c:12-111) in instance order.
 *   ngmi_smp_*      SMPpreOrder / SMPreorder / SMPluFac / SMPsolve
 *                   (klusmp.c:573-1285) for that circuit's CSC matrix on the
 *                   GPU, through the whole-GPU sparse LU (ngmi_splu_*):
 *                   LoadGmin_CSC (klusmp.c:1765-1777) on the device, 1-based
 *                   right-hand side solved in place.
 */
#include <algorithm>
#include <cstring>
#include <vector>
#include "ngmi_internal.h"
#include "bsim4_eval.h"

namespace ngmi {
const B4Tab *set_table(const ngmi_bsim4_set *s);
int launch_bsim4_eval(const B4Tab &tab, const B4Ckt &ck, const double *vterm, double *const state[3],
                      double *cache, double *stamp, double *rhs, int *noncon, hipStream_t stream);
}  // namespace ngmi

using namespace ngmi;

namespace {

/* instance i: terminal voltages from the circuit's rhsOld (node 0 reads
 * rhsOld[0], the ground slot) and its 29 states of each level from the
 * circuit's state vectors at its own offset */
__global__ void __launch_bounds__(256)
k_dl_gather(int n, const int *__restrict__ nodes, const int *__restrict__ soff,
            const double *__restrict__ rhsold, const double *__restrict__ cs0,
            const double *__restrict__ cs1, const double *__restrict__ cs2, double *__restrict__ vterm,
            double *__restrict__ st0, double *__restrict__ st1, double *__restrict__ st2)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int k = 0; k < 4; k++) vterm[(int64_t)k * n + i] = rhsold[nodes[(int64_t)k * n + i]];
    const int64_t o = soff[i];
    for (int k = 0; k < B4X__COUNT; k++) {
        st0[(int64_t)k * n + i] = cs0[o + k];
        st1[(int64_t)k * n + i] = cs1[o + k];
        st2[(int64_t)k * n + i] = cs2[o + k];
    }
}

/* the levels the load writes (state0; state1 at INITTRAN) back into the
 * circuit's vectors */
__global__ void __launch_bounds__(256)
k_dl_scatter(int n, const int *__restrict__ soff, const double *__restrict__ st0,
             const double *__restrict__ st1, double *__restrict__ cs0, double *__restrict__ cs1)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t o = soff[i];
    for (int k = 0; k < B4X__COUNT; k++) {
        cs0[o + k] = st0[(int64_t)k * n + i];
        cs1[o + k] = st1[(int64_t)k * n + i];
    }
}

/* LoadGmin_CSC: gmin onto every diagonal entry the pattern has */
__global__ void __launch_bounds__(256)
k_smp_gmin(double *__restrict__ Ax, const int64_t *__restrict__ diag, int nd, double gmin)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nd) Ax[diag[i]] += gmin;
}

template <class T>
int dev_copy(T **dst, const std::vector<T> &v)
{
    NGMI_HIP(hipMalloc((void **)dst, std::max<size_t>(1, v.size()) * sizeof(T)));
    if (!v.empty()) NGMI_HIP(hipMemcpy(*dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return NGMI_OK;
}

}  // namespace

struct ngmi_devload {
    const ngmi_bsim4_set *set = nullptr;
    int n = 0, nunk = 0, nstates = 0;
    std::vector<int> nodes;            /* [4][n] */
    std::vector<int64_t> ax_pos;       /* [n][16] */
    hipStream_t stream = nullptr;
    int *d_nodes = nullptr, *d_soff = nullptr, *d_nc = nullptr;
    double *d_rhsold = nullptr, *d_cs[3] = {}, *d_vterm = nullptr, *d_st[3] = {}, *d_cache = nullptr;
    double *d_stamp = nullptr, *d_rhs = nullptr;
    double *h_stamp = nullptr, *h_rhs = nullptr;   /* pinned staging */
    int *h_nc = nullptr;
    ~ngmi_devload()
    {
        void *d[] = {d_nodes, d_soff, d_nc, d_rhsold, d_cs[0], d_cs[1], d_cs[2], d_vterm, d_st[0], d_st[1],
                     d_st[2], d_cache, d_stamp, d_rhs};
        for (void *p : d)
            if (p) (void)hipFree(p);
        void *h[] = {h_stamp, h_rhs, h_nc};
        for (void *p : h)
            if (p) (void)hipHostFree(p);
        if (stream) (void)hipStreamDestroy(stream);
    }
    int init(const int *state_off)
    {
        std::vector<int> so(state_off, state_off + n);
        int rc;
        if ((rc = dev_copy(&d_nodes, nodes))) return rc;
        if ((rc = dev_copy(&d_soff, so))) return rc;
        const size_t nn = (size_t)std::max(n, 1);
        NGMI_HIP(hipMalloc((void **)&d_nc, nn * sizeof(int)));
        NGMI_HIP(hipMalloc((void **)&d_rhsold, (size_t)(nunk + 1) * sizeof(double)));
        for (int l = 0; l < 3; l++) {
            NGMI_HIP(hipMalloc((void **)&d_cs[l], (size_t)std::max(nstates, 1) * sizeof(double)));
            NGMI_HIP(hipMalloc((void **)&d_st[l], nn * B4X__COUNT * sizeof(double)));
        }
        NGMI_HIP(hipMalloc((void **)&d_vterm, nn * 4 * sizeof(double)));
        NGMI_HIP(hipMalloc((void **)&d_cache, nn * B4C__COUNT * sizeof(double)));
        NGMI_HIP(hipMemset(d_cache, 0, nn * B4C__COUNT * sizeof(double)));
        NGMI_HIP(hipMalloc((void **)&d_stamp, nn * B4S__COUNT * sizeof(double)));
        NGMI_HIP(hipMalloc((void **)&d_rhs, nn * B4R__COUNT * sizeof(double)));
        NGMI_HIP(hipHostMalloc((void **)&h_stamp, nn * B4S__COUNT * sizeof(double), hipHostMallocDefault));
        NGMI_HIP(hipHostMalloc((void **)&h_rhs, nn * B4R__COUNT * sizeof(double), hipHostMallocDefault));
        NGMI_HIP(hipHostMalloc((void **)&h_nc, nn * sizeof(int), hipHostMallocDefault));
        NGMI_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        return NGMI_OK;
    }
};

struct ngmi_smp {
    ngmi_splu *lu = nullptr;
    int n = 0;
    int64_t nnz = 0;
    std::vector<int64_t> diag;
    int64_t *d_diag = nullptr;
    double *d_Ax = nullptr, *d_b = nullptr;
    hipStream_t stream = nullptr;
    ~ngmi_smp()
    {
        ngmi_splu_free(lu);
        void *d[] = {d_diag, d_Ax, d_b};
        for (void *p : d)
            if (p) (void)hipFree(p);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

extern "C" {

ngmi_devload *ngmi_devload_new(const ngmi_bsim4_set *set, int n_unknowns, int num_states, const int *nodes,
                               const int *state_off, const int64_t *ax_pos)
{
    const B4Tab *tab = set_table(set);
    if (!tab) {
        set_error("devload_new: instance set not finalized");
        return nullptr;
    }
    const int n = (int)tab->n;
    if (n_unknowns < 0 || num_states < 0 || (n > 0 && (!nodes || !state_off || !ax_pos))) {
        set_error("devload_new: bad arguments");
        return nullptr;
    }
    for (int i = 0; i < n; i++) {
        for (int k = 0; k < 4; k++)
            if (nodes[4 * i + k] < 0 || nodes[4 * i + k] > n_unknowns) {
                set_error("devload_new: instance %d terminal %d node %d outside 0..%d", i, k, nodes[4 * i + k],
                          n_unknowns);
                return nullptr;
            }
        if (state_off[i] < 0 || (int64_t)state_off[i] + B4X__COUNT > num_states) {
            set_error("devload_new: instance %d states at %d do not fit %d states", i, state_off[i], num_states);
            return nullptr;
        }
    }
    ngmi_devload *d = new ngmi_devload;
    d->set = set;
    d->n = n;
    d->nunk = n_unknowns;
    d->nstates = num_states;
    d->nodes.resize((size_t)4 * n);
    for (int i = 0; i < n; i++)
        for (int k = 0; k < 4; k++) d->nodes[(size_t)k * n + i] = nodes[4 * i + k];
    d->ax_pos.assign(ax_pos, ax_pos + (size_t)16 * n);
    if (d->init(state_off) != NGMI_OK) {
        delete d;
        return nullptr;
    }
    return d;
}

void ngmi_devload_free(ngmi_devload *d) { delete d; }

int ngmi_devload_run(ngmi_devload *d, const ngmi_ckt_scalars *c, const double *rhsOld, double *const states[3],
                     double *Ax, double *rhs, int *noncon)
{
    if (!d || !c || !rhsOld || !states || !states[0] || !states[1] || !states[2] || !Ax || !rhs) {
        set_error("devload_run: missing arguments");
        return NGMI_E_PANIC;
    }
    if (c->order > 2 && c->method != 1) {
        set_error("devload_run keeps three state levels: Gear order %d needs more", c->order);
        return NGMI_E_UNSUPP;
    }
    const int n = d->n;
    if (n == 0) return NGMI_OK;
    hipStream_t s = d->stream;
    NGMI_HIP(hipMemcpyAsync(d->d_rhsold, rhsOld, (size_t)(d->nunk + 1) * sizeof(double), hipMemcpyHostToDevice, s));
    for (int l = 0; l < 3; l++)
        NGMI_HIP(hipMemcpyAsync(d->d_cs[l], states[l], (size_t)d->nstates * sizeof(double), hipMemcpyHostToDevice,
                                s));
    const unsigned nb = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(k_dl_gather, dim3(nb), dim3(256), 0, s, n, d->d_nodes, d->d_soff, d->d_rhsold, d->d_cs[0],
                       d->d_cs[1], d->d_cs[2], d->d_vterm, d->d_st[0], d->d_st[1], d->d_st[2]);
    NGMI_HIP(hipGetLastError());
    B4Ckt ck;
    ck.mode = c->mode;
    ck.order = c->order;
    ck.method = c->method;
    for (int k = 0; k < 7; k++) ck.ag[k] = c->ag[k];
    ck.xfact = c->xfact;
    ck.gmin = c->gmin;
    ck.reltol = c->reltol;
    ck.abstol = c->abstol;
    int rc = launch_bsim4_eval(*set_table(d->set), ck, d->d_vterm, d->d_st, d->d_cache, d->d_stamp, d->d_rhs,
                               d->d_nc, s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_dl_scatter, dim3(nb), dim3(256), 0, s, n, d->d_soff, d->d_st[0], d->d_st[1], d->d_cs[0],
                       d->d_cs[1]);
    NGMI_HIP(hipGetLastError());
    NGMI_HIP(hipMemcpyAsync(states[0], d->d_cs[0], (size_t)d->nstates * sizeof(double), hipMemcpyDeviceToHost, s));
    NGMI_HIP(hipMemcpyAsync(states[1], d->d_cs[1], (size_t)d->nstates * sizeof(double), hipMemcpyDeviceToHost, s));
    NGMI_HIP(hipMemcpyAsync(d->h_stamp, d->d_stamp, (size_t)n * B4S__COUNT * sizeof(double), hipMemcpyDeviceToHost,
                            s));
    NGMI_HIP(hipMemcpyAsync(d->h_rhs, d->d_rhs, (size_t)n * B4R__COUNT * sizeof(double), hipMemcpyDeviceToHost, s));
    NGMI_HIP(hipMemcpyAsync(d->h_nc, d->d_nc, (size_t)n * sizeof(int), hipMemcpyDeviceToHost, s));
    NGMI_HIP(hipStreamSynchronize(s));
    /* the stamp phase: instance by instance in the set's order, each entry
     * added through its bound position, as the reference's load adds
     * through its matrix pointers */
    int nc = 0;
    for (int i = 0; i < n; i++) {
        const int64_t *pos = &d->ax_pos[(size_t)16 * i];
        for (int k = 0; k < B4S__COUNT; k++)
            if (pos[k] >= 0) Ax[pos[k]] += d->h_stamp[(size_t)k * n + i];
        for (int k = 0; k < B4R__COUNT; k++) {
            const int node = d->nodes[(size_t)k * n + i];
            if (node > 0) rhs[node] += d->h_rhs[(size_t)k * n + i];
        }
        nc += d->h_nc[i];
    }
    if (noncon) *noncon += nc;
    return NGMI_OK;
}

int ngmi_smp_new(int n, const int64_t *Ap, const int *Ai, ngmi_smp **out)
{
    if (!out) return NGMI_E_PANIC;
    *out = nullptr;
    ngmi_splu *lu = nullptr;
    int rc = ngmi_splu_analyze(n, Ap, Ai, &lu);
    if (rc) return rc;
    ngmi_smp *m = new ngmi_smp;
    m->lu = lu;
    m->n = n;
    m->nnz = Ap[n];
    for (int c = 0; c < n; c++)
        for (int64_t e = Ap[c]; e < Ap[c + 1]; e++)
            if (Ai[e] == c) m->diag.push_back(e);
    auto fail = [&](void) {
        delete m;
        return NGMI_E_DEVICE;
    };
    if (dev_copy(&m->d_diag, m->diag) || hipMalloc((void **)&m->d_Ax, (size_t)std::max<int64_t>(1, m->nnz) * 8) ||
        hipMalloc((void **)&m->d_b, (size_t)n * 8) ||
        hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess)
        return fail();
    *out = m;
    return NGMI_OK;
}

void ngmi_smp_free(ngmi_smp *m) { delete m; }

int ngmi_smp_factor(ngmi_smp *m, const double *Ax, double gmin)
{
    if (!m || !Ax) return NGMI_E_PANIC;
    hipStream_t s = m->stream;
    NGMI_HIP(hipMemcpyAsync(m->d_Ax, Ax, (size_t)m->nnz * sizeof(double), hipMemcpyHostToDevice, s));
    if (gmin != 0.0 && !m->diag.empty()) {
        const int nd = (int)m->diag.size();
        hipLaunchKernelGGL(k_smp_gmin, dim3((unsigned)((nd + 255) / 256)), dim3(256), 0, s, m->d_Ax, m->d_diag, nd,
                           gmin);
        NGMI_HIP(hipGetLastError());
    }
    return ngmi_splu_factor(m->lu, m->d_Ax, s);
}

int ngmi_smp_solve(ngmi_smp *m, double *rhs)
{
    if (!m || !rhs) return NGMI_E_PANIC;
    hipStream_t s = m->stream;
    NGMI_HIP(hipMemcpyAsync(m->d_b, rhs + 1, (size_t)m->n * sizeof(double), hipMemcpyHostToDevice, s));
    int rc = ngmi_splu_solve(m->lu, m->d_b, s);
    if (rc) return rc;
    NGMI_HIP(hipMemcpyAsync(rhs + 1, m->d_b, (size_t)m->n * sizeof(double), hipMemcpyDeviceToHost, s));
    NGMI_HIP(hipStreamSynchronize(s));
    rhs[0] = 0.0;
    return NGMI_OK;
}

}  // extern "C"
