/*
 * Test harness (tests only, never linked into the engine): the engine's host
 * BSIM4 setup (bsim4_setup.cpp) and its evaluation template (bsim4_eval.h)
 * compiled for the CPU with g++, so that tests/test_host_bsim4.py can hold
 * them against the oracle without a GPU.  With contraction off on both sides
 * the two are expected to agree bit for bit.
 */
#include <cstring>
#include <vector>
#include "../../ngspice-sf-mirror_amd/csrc/bsim4_setup.h"
#include "../../ngspice-sf-mirror_amd/csrc/bsim4_eval.h"

using namespace ngmi;

namespace {
struct HInst {
    B4Inst in;
    std::vector<double> row;
};
struct Prm {
    const double *r;
    double operator()(int f) const { return r[f]; }
};
struct St {
    double **p;
    double s(int l, int k) const { return p[l][k]; }
    void set(int l, int k, double v) { p[l][k] = v; }
};
struct Ca {
    double *c_;
    double c(int k) const { return c_[k]; }
    void setc(int k, double v) { c_[k] = v; }
};
}  // namespace

extern "C" {

void *hc_model_new() { return new B4Model(); }
void hc_model_free(void *m) { delete (B4Model *)m; }
int hc_model_param(void *m, const char *n, double v) { return ((B4Model *)m)->set(n, v); }
int hc_model_finish(void *m, double tn, double t) { return ((B4Model *)m)->finish(tn, t); }

void *hc_inst_new(void *m)
{
    HInst *h = new HInst();
    h->in.model = (B4Model *)m;
    return h;
}
void hc_inst_free(void *h) { delete (HInst *)h; }
int hc_inst_param(void *h, const char *n, double v) { return ((HInst *)h)->in.set(n, v); }

int hc_inst_finish(void *hp)
{
    HInst *h = (HInst *)hp;
    h->in.defaults();
    if (h->in.check()) return -2;
    h->row.assign(B4F__COUNT, 0.0);
    return h->in.resolve(h->row.data());
}

int hc_load(void *hp, int mode, double gmin, double reltol, double abstol, const double *ag,
            int order, int method, double xfact, const double *vnode, double *s0, double *s1,
            double *s2, double *s3, double *cache, double *stamp, double *rhs)
{
    HInst *h = (HInst *)hp;
    B4Ckt ck;
    ck.mode = mode;
    ck.order = order;
    ck.method = method;
    for (int k = 0; k < 7; k++) ck.ag[k] = ag[k];
    ck.xfact = xfact;
    ck.gmin = gmin;
    ck.reltol = reltol;
    ck.abstol = abstol;
    double *ss[4] = {s0, s1, s2, s3};
    Prm p{h->row.data()};
    St st{ss};
    Ca ca{cache};
    return b4_load(p, st, ca, ck, vnode, stamp, rhs);
}

}  // extern "C"
