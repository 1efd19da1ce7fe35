/*
 * Raw file output of a batch circuit's waveform in the layout ngspice's
 * batch mode writes (-r file): fileInit / fileInit_pass2 / fileStartPoint /
 * fileAddRealValue / fileEnd, frontend/outitf.c:881-1110, with the
 * variable naming of guess_type (outitf.c:928-990): "name#branch" vectors
 * are currents written i(name), every other node vector a voltage written
 * v(name), the scale "time".  ASCII values carry 15 digits after the point
 * (DOUBLE_PRECISION, outitf.c:89); binary rows are native doubles.
 */
#include <cstdio>
#include <cstring>
#include <ctime>
#include <string>
#include <vector>

#include "ngmi_internal.h"

using namespace ngmi;

extern "C" int ngmi_batch_write_raw(const ngmi_batch *b, int c, const char *path, const char *title,
                                    const char *const *names, int binary)
{
    if (!b || !path) return NGMI_E_PANIC;
    int nprobe = 0;
    int rc = ngmi_batch_circuit_info(b, c, nullptr, nullptr, &nprobe, nullptr);
    if (rc != NGMI_OK) return rc;
    int status = 0;
    int64_t npts = 0;
    rc = ngmi_batch_stats(b, c, &status, &npts, nullptr, nullptr, nullptr);
    if (rc != NGMI_OK) return rc;
    std::vector<double> t((size_t)npts), x((size_t)npts * nprobe);
    int64_t got = 0;
    if (npts > 0 && (rc = ngmi_batch_waveform(b, c, t.data(), x.data(), npts, &got)) != NGMI_OK) return rc;
    FILE *fp = std::fopen(path, binary ? "wb" : "w");
    if (!fp) {
        set_error("cannot open %s for writing", path);
        return NGMI_E_PARMVAL;
    }
    /* datestring() (misc_time.c:59-76): asctime without its newline, with
     * the year moved one column right */
    const std::time_t now = std::time(nullptr);
    std::tm tmv;
    localtime_r(&now, &tmv);
    char ab[64];
    asctime_r(&tmv, ab);
    std::string date = std::string(ab, 20) + std::string(ab + 19);
    if (!date.empty() && date.back() == '\n') date.pop_back();
    std::fprintf(fp, "Title: %s\n", title ? title : "");
    std::fprintf(fp, "Date: %s\n", date.c_str());
    std::fprintf(fp, "Command: ngspice_mi355x, batched transient engine\n");
    std::fprintf(fp, "Plotname: Transient Analysis\n");
    std::fprintf(fp, "Flags: real\n");
    std::fprintf(fp, "No. Variables: %d\n", nprobe + 1);
    /* the point count is written over eight reserved characters */
    char cnt[32];
    std::snprintf(cnt, sizeof cnt, "%lld", (long long)got);
    std::string field(cnt);
    if (field.size() < 8) field += std::string(8 - field.size(), ' ');
    std::fprintf(fp, "No. Points: %s\n", field.c_str());
    std::fprintf(fp, "Variables:\n");
    std::fprintf(fp, "\t0\ttime\ttime\n");
    for (int k = 0; k < nprobe; k++) {
        std::string nm = names && names[k] ? names[k] : std::to_string(k + 1);
        const size_t br = nm.find("#branch");
        if (br != std::string::npos)
            std::fprintf(fp, "\t%d\ti(%s)\tcurrent\n", k + 1, nm.substr(0, br).c_str());
        else
            std::fprintf(fp, "\t%d\tv(%s)\tvoltage\n", k + 1, nm.c_str());
    }
    std::fprintf(fp, "%s:\n", binary ? "Binary" : "Values");
    std::vector<double> row((size_t)nprobe + 1);
    for (int64_t p = 0; p < got; p++) {
        if (binary) {
            row[0] = t[p];
            for (int k = 0; k < nprobe; k++) row[k + 1] = x[(size_t)p * nprobe + k];
            std::fwrite(row.data(), sizeof(double), row.size(), fp);
        } else {
            std::fprintf(fp, "%lld\t", (long long)p);
            std::fprintf(fp, "\t%.*e\n", 15, t[p]);
            for (int k = 0; k < nprobe; k++) std::fprintf(fp, "\t%.*e\n", 15, x[(size_t)p * nprobe + k]);
        }
    }
    const bool ok = std::fclose(fp) == 0;
    if (!ok) {
        set_error("write error on %s", path);
        return NGMI_E_PANIC;
    }
    return NGMI_OK;
}
