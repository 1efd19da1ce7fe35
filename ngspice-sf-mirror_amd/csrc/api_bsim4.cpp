/*
 * C ABI of the BSIM4 device: model cards, instance sets and the
 * per-instance evaluation entry point.
 *
 * A set resolves every instance on the host (b4temp.c size / temperature
 * dependence), then splits the table columns: a field whose value is the
 * same for every instance of a model card becomes a per-model column, every
 * other field a per-instance column.  Only the per-instance columns cost
 * HBM traffic per evaluation.
 */
#include <algorithm>
#include <cstring>
#include <thread>
#include <unordered_map>
#include <vector>
#include "ngmi_internal.h"
#include "bsim4_setup.h"
#include "bsim4_eval.h"

namespace ngmi {
int launch_bsim4_eval(const B4Tab &tab, const B4Ckt &ck, const double *vterm,
                      double *const state[3], double *cache, double *stamp, double *rhs,
                      int *noncon, hipStream_t stream);

static const char *const kFieldNames[] = {
#define NGMI_X(f) #f,
    NGMI_B4_FIELDS(NGMI_X)
#undef NGMI_X
};

int field_index(const char *name)
{
    for (int f = 0; f < B4F__COUNT; f++)
        if (std::strcmp(kFieldNames[f], name) == 0) return f;
    return -1;
}

}  // namespace ngmi

using namespace ngmi;

struct ngmi_bsim4_model {
    B4Model m;
    bool ready = false;
};

struct ngmi_bsim4_set {
    std::vector<ngmi_bsim4_model *> models;   /* distinct cards, in first-use order */
    std::unordered_map<const ngmi_bsim4_model *, int> model_pos;
    std::vector<int> model_of;
    std::vector<B4Inst> inst;
    std::vector<double> rows;                 /* n x B4F__COUNT, after finalize */
    std::vector<int> code;
    int icols = 0, mcols = 0;
    double *d_itab = nullptr, *d_mtab = nullptr;
    int *d_model_of = nullptr, *d_code = nullptr;
    size_t dev_bytes = 0;
    B4Tab tab{};
    bool final_ = false;

    void release()
    {
        (void)hipFree(d_itab);
        (void)hipFree(d_mtab);
        (void)hipFree(d_model_of);
        (void)hipFree(d_code);
        d_itab = d_mtab = nullptr;
        d_model_of = d_code = nullptr;
    }
};

namespace ngmi {
const B4Tab *set_table(const ngmi_bsim4_set *s) { return s && s->final_ ? &s->tab : nullptr; }
}

extern "C" {

ngmi_bsim4_model *ngmi_bsim4_model_new(void) { return new ngmi_bsim4_model; }

void ngmi_bsim4_model_free(ngmi_bsim4_model *m) { delete m; }

int ngmi_bsim4_model_param(ngmi_bsim4_model *m, const char *name, double value)
{
    if (!m || !name) return NGMI_E_PANIC;
    if (m->m.set(name, value) != 0) {
        set_error("unknown BSIM4 model parameter '%s'", name);
        return NGMI_E_BADPARM;
    }
    m->ready = false;
    return NGMI_OK;
}

int ngmi_bsim4_model_setup(ngmi_bsim4_model *m, double tnom, double temp)
{
    if (!m) return NGMI_E_PANIC;
    if (m->m.finish(tnom, temp) != 0) {
        const char *why = m->m.check();
        set_error("BSIM4 model card not supported by the engine: %s", why ? why : "setup failed");
        return NGMI_E_UNSUPP;
    }
    m->ready = true;
    return NGMI_OK;
}

double ngmi_bsim4_model_get(const ngmi_bsim4_model *m, const char *what)
{
    if (!m || !what) return 0.0;
    if (!std::strcmp(what, "type")) return m->m.type;
    if (!std::strcmp(what, "coxe")) return m->m.coxe;
    if (!std::strcmp(what, "vtm")) return m->m.vtm;
    if (!std::strcmp(what, "vcrit")) return m->m.vcrit;
    return 0.0;
}

ngmi_bsim4_set *ngmi_bsim4_set_new(void) { return new ngmi_bsim4_set; }

void ngmi_bsim4_set_free(ngmi_bsim4_set *s)
{
    if (!s) return;
    s->release();
    delete s;
}

int ngmi_bsim4_set_add(ngmi_bsim4_set *s, ngmi_bsim4_model *const *models, int nmodels,
                       const int *model_idx, int n, const char *const *names, int nparams,
                       const double *params)
{
    if (!s || !models || n < 0 || (n > 0 && !model_idx)) return NGMI_E_PANIC;
    if (s->final_) {
        set_error("instance set already finalized");
        return NGMI_E_PANIC;
    }
    for (int k = 0; k < nmodels; k++) {
        if (!models[k] || !models[k]->ready) {
            set_error("model %d not set up (call ngmi_bsim4_model_setup first)", k);
            return NGMI_E_NOMOD;
        }
    }
    /* map caller model indices to set model indices */
    std::vector<int> local(nmodels, -1);
    const int first = (int)s->inst.size();
    for (int k = 0; k < n; k++) {
        const int mk = model_idx[k];
        if (mk < 0 || mk >= nmodels) {
            set_error("instance %d: model index %d out of range", k, mk);
            return NGMI_E_NOMOD;
        }
        if (local[mk] < 0) {
            auto it = s->model_pos.find(models[mk]);
            if (it == s->model_pos.end()) {
                s->models.push_back(models[mk]);
                local[mk] = (int)s->models.size() - 1;
                s->model_pos[models[mk]] = local[mk];
            } else {
                local[mk] = it->second;
            }
        }
        B4Inst in;
        in.model = &models[mk]->m;
        for (int j = 0; j < nparams; j++) {
            if (in.set(names[j], params[(size_t)j * n + k]) != 0) {
                set_error("unknown BSIM4 instance parameter '%s'", names[j]);
                s->inst.resize(first);
                s->model_of.resize(first);
                return -NGMI_E_BADPARM;
            }
        }
        in.defaults();
        if (const char *why = in.check()) {
            set_error("instance %d not supported by the engine: %s", first + k, why);
            s->inst.resize(first);
            s->model_of.resize(first);
            return -NGMI_E_UNSUPP;
        }
        s->inst.push_back(in);
        s->model_of.push_back(local[mk]);
    }
    return first;
}

int ngmi_bsim4_set_finalize(ngmi_bsim4_set *s)
{
    if (!s) return NGMI_E_PANIC;
    const size_t n = s->inst.size();
    const int F = B4F__COUNT;
    const int nm = (int)s->models.size();
    s->rows.assign(n * F, 0.0);
    /* b4check clamps act on the first instance of each card: resolve those
     * first and in order, the rest in parallel */
    std::vector<char> done(n, 0);
    std::vector<char> seen(nm, 0);
    for (size_t k = 0; k < n; k++) {
        const int m = s->model_of[k];
        if (seen[m]) continue;
        seen[m] = 1;
        if (s->inst[k].resolve(&s->rows[k * F]) != 0) {
            set_error("instance %zu: effective channel length or width <= 0", k);
            return NGMI_E_PARMVAL;
        }
        done[k] = 1;
    }
    unsigned nt = std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
    if (n < 4096) nt = 1;
    std::vector<int> bad(nt, -1);
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; t++) {
        th.emplace_back([&, t]() {
            for (size_t k = t; k < n; k += nt) {
                if (done[k]) continue;
                if (s->inst[k].resolve(&s->rows[k * F]) != 0 && bad[t] < 0) bad[t] = (int)k;
            }
        });
    }
    for (auto &x : th) x.join();
    for (unsigned t = 0; t < nt; t++) {
        if (bad[t] >= 0) {
            set_error("instance %d: effective channel length or width <= 0", bad[t]);
            return NGMI_E_PARMVAL;
        }
    }
    /* column split */
    std::vector<char> perinst(F, 0);
    std::vector<size_t> firstOf(nm, (size_t)-1);
    for (size_t k = 0; k < n; k++)
        if (firstOf[s->model_of[k]] == (size_t)-1) firstOf[s->model_of[k]] = k;
    for (int f = 0; f < F; f++) {
        for (size_t k = 0; k < n && !perinst[f]; k++) {
            const double a = s->rows[k * F + f];
            const double b = s->rows[firstOf[s->model_of[k]] * F + f];
            if (std::memcmp(&a, &b, sizeof(double)) != 0) perinst[f] = 1;
        }
    }
    s->code.assign(F, 0);
    s->icols = s->mcols = 0;
    for (int f = 0; f < F; f++) s->code[f] = perinst[f] ? s->icols++ : ~(s->mcols++);
    std::vector<double> itab((size_t)s->icols * n), mtab((size_t)s->mcols * std::max(nm, 1));
    for (int f = 0; f < F; f++) {
        if (s->code[f] >= 0) {
            double *dst = &itab[(size_t)s->code[f] * n];
            for (size_t k = 0; k < n; k++) dst[k] = s->rows[k * F + f];
        } else {
            /* model-major: one card's per-model values are contiguous, so
             * the few cards a workgroup reads stay in its L1 */
            for (int m = 0; m < nm; m++)
                mtab[(size_t)m * s->mcols + (~s->code[f])] = s->rows[firstOf[m] * F + f];
        }
    }
    /* one selector word for the whole set lets the kernels compile the
     * selectors as constants */
    s->tab.flags = n ? (int)s->rows[B4F_flags] : -1;
    for (size_t k = 1; k < n && s->tab.flags >= 0; k++)
        if ((int)s->rows[k * F + B4F_flags] != s->tab.flags) s->tab.flags = -1;
    s->release();
    const size_t bi = std::max<size_t>(itab.size(), 1) * sizeof(double);
    const size_t bm = std::max<size_t>(mtab.size(), 1) * sizeof(double);
    const size_t bo = std::max<size_t>(n, 1) * sizeof(int);
    NGMI_HIP(hipMalloc(&s->d_itab, bi));
    NGMI_HIP(hipMalloc(&s->d_mtab, bm));
    NGMI_HIP(hipMalloc(&s->d_model_of, bo));
    NGMI_HIP(hipMalloc(&s->d_code, F * sizeof(int)));
    if (!itab.empty()) NGMI_HIP(hipMemcpy(s->d_itab, itab.data(), itab.size() * 8, hipMemcpyHostToDevice));
    if (!mtab.empty()) NGMI_HIP(hipMemcpy(s->d_mtab, mtab.data(), mtab.size() * 8, hipMemcpyHostToDevice));
    if (n) NGMI_HIP(hipMemcpy(s->d_model_of, s->model_of.data(), n * sizeof(int), hipMemcpyHostToDevice));
    NGMI_HIP(hipMemcpy(s->d_code, s->code.data(), F * sizeof(int), hipMemcpyHostToDevice));
    s->dev_bytes = bi + bm + bo + F * sizeof(int);
    s->tab.itab = s->d_itab;
    s->tab.mtab = s->d_mtab;
    s->tab.model_of = s->d_model_of;
    s->tab.n = (int64_t)n;
    s->tab.nm = nm;
    s->tab.mcols = s->mcols;
    s->tab.code = s->d_code;
    s->final_ = true;
    return NGMI_OK;
}

int ngmi_bsim4_set_size(const ngmi_bsim4_set *s) { return s ? (int)s->inst.size() : 0; }

double ngmi_bsim4_set_field(const ngmi_bsim4_set *s, int k, const char *field)
{
    if (!s || !s->final_ || k < 0 || (size_t)k >= s->inst.size()) return 0.0;
    const int f = field_index(field);
    if (f < 0) return 0.0;
    return s->rows[(size_t)k * B4F__COUNT + f];
}

int ngmi_bsim4_set_layout(const ngmi_bsim4_set *s, int *inst_cols, int *model_cols,
                          size_t *device_bytes)
{
    if (!s || !s->final_) return NGMI_E_PANIC;
    if (inst_cols) *inst_cols = s->icols;
    if (model_cols) *model_cols = s->mcols;
    if (device_bytes) *device_bytes = s->dev_bytes;
    return NGMI_OK;
}

int ngmi_bsim4_eval(const ngmi_bsim4_set *s, const ngmi_ckt_scalars *c, const double *vterm,
                    double *const state[3], double *cache, double *stamp, double *rhs,
                    int *noncon, void *stream)
{
    if (!s || !s->final_ || !c || !state) {
        set_error("ngmi_bsim4_eval: set not finalized or missing arguments");
        return NGMI_E_PANIC;
    }
    if (c->order > 2 && c->method != 1) {
        set_error("ngmi_bsim4_eval keeps three state levels: Gear order %d needs more", c->order);
        return NGMI_E_UNSUPP;
    }
    B4Ckt ck;
    ck.mode = c->mode;
    ck.order = c->order;
    ck.method = c->method;
    for (int k = 0; k < 7; k++) ck.ag[k] = c->ag[k];
    ck.xfact = c->xfact;
    ck.gmin = c->gmin;
    ck.reltol = c->reltol;
    ck.abstol = c->abstol;
    return launch_bsim4_eval(s->tab, ck, vterm, state, cache, stamp, rhs, noncon,
                             (hipStream_t)stream);
}

}  // extern "C"
