/*
 * Autotuner of the batched transient's launch knobs (no counterpart in the
 * reference): for a given set of circuits it times short runs of
 * ngmi_batch_tran (a fixed number of Newton passes from t = 0, the same
 * work for every setting) under candidate values of each knob, one knob at
 * a time from the current settings (coordinate descent), keeps the fastest,
 * and appends the choice to a text cache keyed by the batch's shape so a
 * later run applies it without timing (ngmi_autotune_load).
 *
 * Knobs tried, in this order: order_slack (pivot order, rebuilds the
 * batch), mos_waves (BSIM4 kernel occupancy), streams (two-stream overlap,
 * single-topology batches only), graph (HIP-graph replay of pass chunks).
 */
#include <chrono>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>
#include "ngmi_internal.h"

namespace {

struct Knob {
    const char *name;
    int tune_id;
    std::vector<int> values;
    bool rebuild;   /* takes effect when the batch is built */
};

std::string signature(ngmi_circuit *const *ckts, int C)
{
    /* shape of the batch: circuits, topology groups, unknowns and MOSFETs
     * summed over circuits */
    ngmi_batch *b = ngmi_batch_new(ckts, C, nullptr, 0, 4);
    if (!b) return std::string();
    int groups = 0;
    ngmi_batch_groups(b, &groups);
    long long unk = 0;
    for (int c = 0; c < C; c++) {
        int g = 0, n = 0, np = 0, ns = 0;
        ngmi_batch_circuit_info(b, c, &g, &n, &np, &ns);
        unk += n;
    }
    ngmi_batch_free(b);
    char buf[160];
    std::snprintf(buf, sizeof buf, "circuits=%d groups=%d unknowns=%lld", C, groups, unk);
    return buf;
}

}  // namespace

extern "C" {

int ngmi_autotune_batch(ngmi_circuit *const *ckts, int C, double tstep, double tstop, int64_t trial_passes,
                        const char *cache_path, char *report, int report_len)
{
    if (!ckts || C <= 0 || !(tstep > 0) || !(tstop > 0) || trial_passes <= 0) {
        ngmi::set_error("autotune: bad arguments");
        return NGMI_E_PARMVAL;
    }
    using clk = std::chrono::steady_clock;
    std::vector<Knob> knobs = {{"order_slack", ngmi::TUNE_ORDER_SLACK, {0, 1, 2}, true},
                               {"mos_waves", ngmi::TUNE_MOS_WAVES, {1, 2, 3, 4}, false},
                               {"streams", ngmi::TUNE_STREAMS, {1, 2, 3, 4}, false},
                               {"graph", ngmi::TUNE_GRAPH, {0, 1}, false}};
    std::ostringstream log;
    ngmi_batch *b = nullptr;
    auto build = [&]() -> int {
        if (b) ngmi_batch_free(b);
        b = ngmi_batch_new(ckts, C, nullptr, 0, 4);
        return b ? NGMI_OK : NGMI_E_PANIC;
    };
    /* passes per second of one trial; a run that ends early (every circuit
     * finished) counts the passes it made */
    auto trial = [&](double *rate) -> int {
        int64_t passes = 0;
        (void)hipDeviceSynchronize();
        const auto t0 = clk::now();
        const int rc = ngmi_batch_tran(b, tstep, tstop, 0.0, trial_passes, nullptr, &passes);
        (void)hipDeviceSynchronize();
        const double s = std::chrono::duration<double>(clk::now() - t0).count();
        if (rc != NGMI_OK && rc != NGMI_E_ITERLIM) return rc;
        *rate = s > 0 ? passes / s : 0.0;
        return NGMI_OK;
    };
    int rc = build();
    if (rc) return rc;
    int groups = 1;
    ngmi_batch_groups(b, &groups);
    double rate = 0.0;
    if ((rc = trial(&rate)) != NGMI_OK) {   /* warm-up: code objects, lazy allocations */
        ngmi_batch_free(b);
        return rc;
    }
    double best = 0.0;
    if ((rc = trial(&best)) != NGMI_OK) {
        ngmi_batch_free(b);
        return rc;
    }
    log << "start " << best << " passes/s;";
    for (const Knob &k : knobs) {
        if (!std::strcmp(k.name, "streams") && groups != 1) continue;
        const int keep = ngmi::tune_get(k.tune_id);
        int choice = keep;
        for (int v : k.values) {
            if (v == keep) continue;
            ngmi_tune(k.name, v);
            if (k.rebuild && (rc = build()) != NGMI_OK) break;
            if ((rc = trial(&rate)) != NGMI_OK) break;
            log << " " << k.name << "=" << v << ": " << rate << ";";
            if (rate > best * 1.02) {   /* a change must win by more than the timing noise */
                best = rate;
                choice = v;
            }
        }
        ngmi_tune(k.name, choice);
        if (rc != NGMI_OK) break;
        if (k.rebuild && (rc = build()) != NGMI_OK) break;
    }
    if (b) ngmi_batch_free(b);
    if (rc != NGMI_OK) return rc;
    std::ostringstream choice;
    for (const Knob &k : knobs) choice << " " << k.name << "=" << ngmi::tune_get(k.tune_id);
    log << " chosen" << choice.str() << " (" << best << " passes/s)";
    if (cache_path && *cache_path) {
        const std::string sig = signature(ckts, C);
        std::ofstream out(cache_path, std::ios::app);
        if (!out) {
            ngmi::set_error("autotune: cannot write %s", cache_path);
            return NGMI_E_PANIC;
        }
        out << sig << " |" << choice.str() << " | passes_per_s=" << best << "\n";
    }
    if (report && report_len > 0) std::snprintf(report, (size_t)report_len, "%s", log.str().c_str());
    return NGMI_OK;
}

int ngmi_autotune_load(const char *cache_path, ngmi_circuit *const *ckts, int C)
{
    if (!cache_path || !ckts || C <= 0) return NGMI_E_PARMVAL;
    std::ifstream in(cache_path);
    if (!in) {
        ngmi::set_error("autotune: no cache %s", cache_path);
        return NGMI_E_NOTFOUND;
    }
    const std::string sig = signature(ckts, C);
    std::string line, last;
    while (std::getline(in, line))
        if (line.rfind(sig + " |", 0) == 0) last = line;   /* the newest entry wins */
    if (last.empty()) {
        ngmi::set_error("autotune: no entry for %s", sig.c_str());
        return NGMI_E_NOTFOUND;
    }
    const size_t a = last.find('|') + 1, z = last.find('|', a);
    std::istringstream kv(last.substr(a, z - a));
    std::string item;
    while (kv >> item) {
        const size_t eq = item.find('=');
        if (eq == std::string::npos) continue;
        const int rc = ngmi_tune(item.substr(0, eq).c_str(), std::stoi(item.substr(eq + 1)));
        if (rc) return rc;
    }
    return NGMI_OK;
}

}  // extern "C"
