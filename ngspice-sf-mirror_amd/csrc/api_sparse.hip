/*
 * Sparse pattern construction and deterministic assembly.
 *
 * ngmi_csc_from_elements restates what SMPmakeElt (klusmp.c:417-440) and
 * SMPconvertCOOtoCSC (klusmp.c:137-323) produce: the distinct (row, col)
 * pairs of all element requests as a CSC pattern with ascending rows per
 * column, and for every request the CSC position it binds to (the reference
 * binds a double* into KLU's Ax; the engine keeps the position).
 *
 * ngmi_gather_* sums element contributions into Ax in a fixed order (one
 * thread per nonzero, contributions in ascending element order), so results
 * are bitwise reproducible run to run; no atomics.
 */
#include <algorithm>
#include <numeric>
#include <vector>
#include "ngmi_internal.h"

using namespace ngmi;

struct ngmi_gather {
    int64_t nnz = 0, ne = 0;
    int64_t *d_ptr = nullptr;
    int64_t *d_src = nullptr;
};

namespace ngmi {

__global__ void __launch_bounds__(256)
k_gather(const int64_t *__restrict__ ptr, const int64_t *__restrict__ srcIdx,
         const double *__restrict__ src, double *__restrict__ Ax, int64_t nnz)
{
    const int64_t z = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (z >= nnz) return;
    double acc = 0.0;
    const int64_t b = ptr[z], e = ptr[z + 1];
    for (int64_t p = b; p < e; p++) acc += src[srcIdx[p]];
    Ax[z] = acc;
}

}  // namespace ngmi

extern "C" {

int ngmi_csc_from_elements(int n, int64_t ne, const int *row, const int *col, int64_t *Ap,
                           int *Ai, int64_t *elt2nz, int64_t *nnz)
{
    if (n < 0 || ne < 0 || (ne > 0 && (!row || !col || !Ai || !elt2nz)) || !Ap || !nnz)
        return NGMI_E_PANIC;
    for (int64_t e = 0; e < ne; e++) {
        if (row[e] < 0 || row[e] >= n || col[e] < 0 || col[e] >= n) {
            set_error("element %lld: (%d, %d) outside a %d x %d matrix", (long long)e, row[e],
                      col[e], n, n);
            return NGMI_E_PARMVAL;
        }
    }
    /* counting sort by column, then by row within each column (stable on e) */
    std::vector<int64_t> ccount((size_t)n + 1, 0);
    for (int64_t e = 0; e < ne; e++) ccount[(size_t)col[e] + 1]++;
    for (int c = 0; c < n; c++) ccount[(size_t)c + 1] += ccount[c];
    std::vector<int64_t> byCol((size_t)ne);
    {
        std::vector<int64_t> pos(ccount.begin(), ccount.end() - 1);
        for (int64_t e = 0; e < ne; e++) byCol[(size_t)pos[col[e]]++] = e;
    }
    int64_t z = 0;
    Ap[0] = 0;
    for (int c = 0; c < n; c++) {
        const int64_t b = ccount[c], en = ccount[(size_t)c + 1];
        std::stable_sort(byCol.begin() + b, byCol.begin() + en,
                         [&](int64_t x, int64_t y) { return row[x] < row[y]; });
        int last = -1;
        for (int64_t p = b; p < en; p++) {
            const int64_t e = byCol[(size_t)p];
            if (row[e] != last) {
                Ai[z++] = row[e];
                last = row[e];
            }
            elt2nz[e] = z - 1;
        }
        Ap[c + 1] = z;
    }
    *nnz = z;
    return NGMI_OK;
}

ngmi_gather *ngmi_gather_new(int64_t nnz, int64_t ne, const int64_t *elt2nz)
{
    if (nnz < 0 || ne < 0 || (ne > 0 && !elt2nz)) {
        set_error("ngmi_gather_new: bad arguments");
        return nullptr;
    }
    std::vector<int64_t> ptr((size_t)nnz + 1, 0), src((size_t)std::max<int64_t>(ne, 1));
    for (int64_t e = 0; e < ne; e++) {
        if (elt2nz[e] < 0 || elt2nz[e] >= nnz) {
            set_error("ngmi_gather_new: element %lld maps to %lld outside [0, %lld)",
                      (long long)e, (long long)elt2nz[e], (long long)nnz);
            return nullptr;
        }
        ptr[(size_t)elt2nz[e] + 1]++;
    }
    for (int64_t z = 0; z < nnz; z++) ptr[(size_t)z + 1] += ptr[(size_t)z];
    {
        std::vector<int64_t> pos(ptr.begin(), ptr.end() - 1);
        for (int64_t e = 0; e < ne; e++) src[(size_t)pos[(size_t)elt2nz[e]]++] = e;
    }
    ngmi_gather *g = new ngmi_gather;
    g->nnz = nnz;
    g->ne = ne;
    if (hipMalloc(&g->d_ptr, ptr.size() * sizeof(int64_t)) != hipSuccess ||
        hipMalloc(&g->d_src, src.size() * sizeof(int64_t)) != hipSuccess ||
        hipMemcpy(g->d_ptr, ptr.data(), ptr.size() * sizeof(int64_t), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(g->d_src, src.data(), src.size() * sizeof(int64_t), hipMemcpyHostToDevice) != hipSuccess) {
        set_error("ngmi_gather_new: device allocation or copy failed");
        ngmi_gather_free(g);
        return nullptr;
    }
    return g;
}

void ngmi_gather_free(ngmi_gather *g)
{
    if (!g) return;
    (void)hipFree(g->d_ptr);
    (void)hipFree(g->d_src);
    delete g;
}

int ngmi_gather_run(const ngmi_gather *g, const double *src, double *Ax, void *stream)
{
    if (!g) return NGMI_E_PANIC;
    if (g->nnz == 0) return NGMI_OK;
    const int bs = 256;
    const int64_t nb = (g->nnz + bs - 1) / bs;
    hipLaunchKernelGGL(k_gather, dim3((unsigned)nb), dim3(bs), 0, (hipStream_t)stream, g->d_ptr,
                       g->d_src, src, Ax, g->nnz);
    NGMI_HIP(hipGetLastError());
    return NGMI_OK;
}

}  // extern "C"
