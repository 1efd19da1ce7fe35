/*
 * The embedding interface of ngspice's shared library (sharedspice.h:
 * 384-486; behaviour of sharedspice.c) over the engine: a deck arrives by
 * ngSpice_Circ, "source" or "circbyline", is parsed by the engine's front end,
 * and "run" / "bg_run" / "tran" run its transient as a one-circuit batch on
 * the GPU.  Every accepted point goes to SendData as it arrives (the
 * engine's live data stream), the vectors of the finished run are served by
 * ngGet_Vec_Info.  Output to the caller goes through SendChar as
 * "stdout ..." / "stderr ..." lines, as the reference's does.
 */
#include <atomic>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <fstream>
#include <mutex>
#include <sstream>
#include <string>
#include <thread>
#include <vector>
#include "ngmi_internal.h"
#include "../../include/ngspice_mi355x_shared.h"

namespace {

struct Vec {
    std::string name;      /* lower case, as ngspice stores vector names */
    std::vector<double> data;
    vector_info info{};
};

struct Shared {
    bool init = false;
    SendChar *sendchar = nullptr;
    SendStat *sendstat = nullptr;
    ControlledExit *exitfn = nullptr;
    SendData *senddata = nullptr;
    SendInitData *sendinit = nullptr;
    BGThreadRunning *bgrun = nullptr;
    void *user = nullptr;

    std::mutex mu;                       /* deck, results, thread */
    std::vector<std::string> lines;      /* circbyline collection */
    std::string deck;                    /* the current circuit's text */
    bool have_circuit = false;

    std::vector<Vec> vecs;               /* the last run: time, then unknowns */
    std::string plot = "const";
    std::vector<char *> plot_list, vec_list;
    std::thread worker;
    std::atomic<bool> running{false};

    /* per-run SendData buffers */
    std::vector<std::string> names;
    std::vector<vecvalues> vals;
    std::vector<pvecvalues> valp;
    int64_t next_index = 0;
};

Shared &G()
{
    static Shared g;
    return g;
}

void say(const char *stream, const std::string &text)
{
    Shared &g = G();
    std::string line = std::string(stream) + " " + text;
    if (g.sendchar) g.sendchar(const_cast<char *>(line.c_str()), 0, g.user);
    else std::fprintf(std::strcmp(stream, "stderr") == 0 ? stderr : stdout, "%s\n", text.c_str());
}

std::string lower(std::string s)
{
    for (char &c : s) c = (char)std::tolower((unsigned char)c);
    return s;
}

std::string trim(const std::string &s)
{
    size_t a = 0, b = s.size();
    while (a < b && std::isspace((unsigned char)s[a])) a++;
    while (b > a && std::isspace((unsigned char)s[b - 1])) b--;
    return s.substr(a, b - a);
}

/* SendData for every new point of the running circuit */
int on_points(int /*circuit*/, int64_t first, int64_t count, const double *t, const double *x, int nprobe, void *)
{
    Shared &g = G();
    for (int64_t p = 0; p < count; p++) {
        g.vals[0].creal = t[p];
        for (int k = 0; k < nprobe; k++) g.vals[1 + k].creal = x[p * nprobe + k];
        vecvaluesall all{(int)g.vals.size(), (int)(first + p), g.valp.data()};
        if (g.senddata) g.senddata(&all, (int)g.vals.size(), 0, g.user);
    }
    g.next_index = first + count;
    return 0;
}

/* one transient of the current deck; tstep <= 0 takes the deck's .tran */
int run_tran(double tstep, double tstop, double tstart, double tmax)
{
    Shared &g = G();
    std::string deck;
    {
        std::lock_guard<std::mutex> lk(g.mu);
        if (!g.have_circuit) {
            say("stderr", "Error: there aren't any circuits loaded.");
            return 1;
        }
        deck = g.deck;
    }
    ngmi_netlist *nl = ngmi_netlist_parse(deck.c_str(), 0);
    if (!nl) {
        say("stderr", std::string("Error: ") + ngmi_last_error());
        return 1;
    }
    int uic = 0;
    if (tstep <= 0.0) {
        if (ngmi_netlist_tran(nl, &tstep, &tstop, &tstart, &tmax, &uic) != NGMI_OK || tstep <= 0.0) {
            say("stderr", "Error: the deck has no .tran analysis");
            ngmi_netlist_free(nl);
            return 1;
        }
    }
    int nnodes = 0, nunk = 0, nel = 0;
    ngmi_netlist_size(nl, &nnodes, &nunk, &nel);
    ngmi_circuit *ckt = ngmi_netlist_circuit(nl);
    if (!ckt) {
        say("stderr", std::string("Error: ") + ngmi_last_error());
        ngmi_netlist_free(nl);
        return 1;
    }
    const int64_t maxpts = (int64_t)std::min(4194304.0, std::max(4096.0, 2.0 * tstop / tstep + 1024.0));
    ngmi_batch *b = ngmi_batch_new(&ckt, 1, nullptr, 0, maxpts);
    if (!b) {
        say("stderr", std::string("Error: ") + ngmi_last_error());
        ngmi_circuit_free(ckt);
        ngmi_netlist_free(nl);
        return 1;
    }
    /* vectors: the scale, then every unknown under ngspice's names */
    g.names.assign(1, "time");
    for (int k = 1; k <= nunk; k++) g.names.push_back(lower(ngmi_netlist_name(nl, k)));
    g.vals.assign(g.names.size(), vecvalues{});
    g.valp.resize(g.names.size());
    for (size_t k = 0; k < g.names.size(); k++) {
        g.vals[k].name = const_cast<char *>(g.names[k].c_str());
        g.vals[k].is_scale = k == 0;
        g.vals[k].is_complex = false;
        g.valp[k] = &g.vals[k];
    }
    if (g.sendinit) {
        std::vector<vecinfo> vi(g.names.size());
        std::vector<pvecinfo> vip(g.names.size());
        for (size_t k = 0; k < g.names.size(); k++) {
            vi[k] = vecinfo{(int)k, const_cast<char *>(g.names[k].c_str()), true, nullptr, nullptr};
            vip[k] = &vi[k];
        }
        char name[] = "tran1", type[] = "transient";
        std::string title = ngmi_netlist_title(nl) ? ngmi_netlist_title(nl) : "";
        std::time_t now = std::time(nullptr);
        std::string date = std::ctime(&now);
        if (!date.empty() && date.back() == '\n') date.pop_back();
        vecinfoall va{name, const_cast<char *>(title.c_str()), const_cast<char *>(date.c_str()), type,
                      (int)g.names.size(), vip.data()};
        g.sendinit(&va, 0, g.user);
    }
    g.next_index = 0;
    ngmi_batch_set_data_callback(b, on_points, nullptr, 0, nullptr);
    int64_t passes = 0;
    const int rc = ngmi_batch_tran(b, tstep, tstop, tmax, (int64_t)1 << 40, nullptr, &passes);
    int status = 0;
    int64_t npts = 0, acc = 0, rej = 0, its = 0;
    ngmi_batch_stats(b, 0, &status, &npts, &acc, &rej, &its);
    std::vector<double> t((size_t)std::max<int64_t>(npts, 1)), x((size_t)std::max<int64_t>(npts, 1) * nunk);
    int64_t got = 0;
    if (npts > 0) ngmi_batch_waveform(b, 0, t.data(), x.data(), npts, &got);
    ngmi_batch_free(b);
    ngmi_circuit_free(ckt);
    ngmi_netlist_free(nl);
    if (rc != NGMI_OK || status != 0) {
        say("stderr", std::string("Error: transient analysis failed: ") + ngmi_last_error());
        return 1;
    }
    /* results: points at or after tstart, as ngspice saves them */
    std::lock_guard<std::mutex> lk(g.mu);
    g.vecs.clear();
    g.vecs.resize(g.names.size());
    for (size_t k = 0; k < g.names.size(); k++) g.vecs[k].name = g.names[k];
    for (int64_t p = 0; p < got; p++) {
        if (t[p] < tstart) continue;
        g.vecs[0].data.push_back(t[p]);
        for (int k = 0; k < nunk; k++) g.vecs[1 + k].data.push_back(x[(size_t)p * nunk + k]);
    }
    for (Vec &v : g.vecs) {
        v.info.v_name = const_cast<char *>(v.name.c_str());
        v.info.v_type = 0;
        v.info.v_flags = 1;   /* VF_REAL */
        v.info.v_realdata = v.data.data();
        v.info.v_compdata = nullptr;
        v.info.v_length = (int)v.data.size();
    }
    g.plot = "tran1";
    char buf[160];
    std::snprintf(buf, sizeof buf, "Transient: %lld accepted, %lld rejected timepoints, %lld Newton iterations",
                  (long long)acc, (long long)rej, (long long)its);
    say("stdout", buf);
    if (g.sendstat) {
        char st[] = "--ready--";
        g.sendstat(st, 0, g.user);
    }
    return 0;
}

void join_worker()
{
    Shared &g = G();
    if (g.worker.joinable()) g.worker.join();
}

int set_deck(const std::vector<std::string> &lines)
{
    Shared &g = G();
    std::string text;
    for (const std::string &l : lines) text += l + "\n";
    ngmi_netlist *nl = ngmi_netlist_parse(text.c_str(), 0);
    if (!nl) {
        say("stderr", std::string("Error: ") + ngmi_last_error());
        return 1;
    }
    const char *title = ngmi_netlist_title(nl);
    say("stdout", std::string("Circuit: ") + (title ? title : ""));
    ngmi_netlist_free(nl);
    std::lock_guard<std::mutex> lk(g.mu);
    g.deck = text;
    g.have_circuit = true;
    return 0;
}

std::vector<std::string> split_words(const std::string &s)
{
    std::istringstream in(s);
    std::vector<std::string> w;
    std::string x;
    while (in >> x) w.push_back(x);
    return w;
}

/* a SPICE number with an optional scale suffix (f p n u m k meg g t) */
bool number(const std::string &s, double &out)
{
    const char *p = s.c_str();
    char *end = nullptr;
    double v = std::strtod(p, &end);
    if (end == p) return false;
    std::string suf = lower(end);
    double m = 1.0;
    if (suf.rfind("meg", 0) == 0) m = 1e6;
    else if (!suf.empty()) {
        switch (suf[0]) {
        case 'f': m = 1e-15; break;
        case 'p': m = 1e-12; break;
        case 'n': m = 1e-9; break;
        case 'u': m = 1e-6; break;
        case 'm': m = 1e-3; break;
        case 'k': m = 1e3; break;
        case 'g': m = 1e9; break;
        case 't': m = 1e12; break;
        case 's': break;
        default: return false;
        }
    }
    out = v * m;
    return true;
}

}  // namespace

extern "C" {

int ngSpice_Init(SendChar *printfcn, SendStat *statfcn, ControlledExit *ngexit, SendData *sdata,
                 SendInitData *sinitdata, BGThreadRunning *bgtrun, void *userData)
{
    Shared &g = G();
    join_worker();
    g.sendchar = printfcn;
    g.sendstat = statfcn;
    g.exitfn = ngexit;
    g.senddata = sdata;
    g.sendinit = sinitdata;
    g.bgrun = bgtrun;
    g.user = userData;
    g.init = true;
    return 0;
}

int ngSpice_Circ(char **circarray)
{
    Shared &g = G();
    if (!g.init || !circarray) return 1;
    std::vector<std::string> lines;
    for (char **p = circarray; *p; p++) lines.push_back(*p);
    return set_deck(lines);
}

int ngSpice_Command(char *command)
{
    Shared &g = G();
    if (!g.init) return 1;
    if (!command) {
        std::lock_guard<std::mutex> lk(g.mu);
        g.lines.clear();
        return 0;
    }
    const std::string cmd = trim(command);
    if (cmd.empty()) return 1;
    std::vector<std::string> w = split_words(cmd);
    const std::string verb = lower(w[0]);
    if (verb == "circbyline") {
        const std::string line = trim(cmd.substr(w[0].size()));
        std::vector<std::string> lines;
        {
            std::lock_guard<std::mutex> lk(g.mu);
            g.lines.push_back(line);
            if (lower(line).rfind(".end", 0) != 0 || lower(line).rfind(".ends", 0) == 0) return 0;
            lines.swap(g.lines);
        }
        return set_deck(lines);
    }
    if (verb == "source") {
        if (w.size() < 2) {
            say("stderr", "Error: source: no file name");
            return 1;
        }
        std::ifstream in(trim(cmd.substr(w[0].size())));
        if (!in) {
            say("stderr", "Error: source: cannot open " + w[1]);
            return 1;
        }
        std::vector<std::string> lines;
        std::string l;
        while (std::getline(in, l)) lines.push_back(l);
        return set_deck(lines);
    }
    if (verb == "run" || verb == "tran") {
        if (g.running.load()) {
            say("stderr", "Error: a background run is active");
            return 1;
        }
        join_worker();
        double v[4] = {0.0, 0.0, 0.0, 0.0};
        if (verb == "tran") {
            if (w.size() < 3) {
                say("stderr", "Error: tran: tstep and tstop are required");
                return 1;
            }
            for (size_t k = 1; k < w.size() && k <= 4; k++)
                if (!number(w[k], v[k - 1])) {
                    say("stderr", "Error: tran: bad number " + w[k]);
                    return 1;
                }
        }
        return run_tran(v[0], v[1], v[2], v[3]);
    }
    if (verb == "bg_run") {
        if (g.running.load()) {
            say("stderr", "Error: a background run is active");
            return 1;
        }
        join_worker();
        g.running = true;
        g.worker = std::thread([] {
            Shared &gg = G();
            if (gg.bgrun) gg.bgrun(false, 0, gg.user);   /* false: the thread is running */
            run_tran(0.0, 0.0, 0.0, 0.0);
            gg.running = false;
            if (gg.bgrun) gg.bgrun(true, 0, gg.user);    /* true: no thread running */
        });
        return 0;
    }
    if (verb == "remcirc" || verb == "reset") {
        join_worker();
        std::lock_guard<std::mutex> lk(g.mu);
        g.have_circuit = false;
        g.deck.clear();
        return 0;
    }
    if (verb == "destroy") {
        join_worker();
        std::lock_guard<std::mutex> lk(g.mu);
        g.vecs.clear();
        g.plot = "const";
        return 0;
    }
    if (verb == "echo") {
        say("stdout", trim(cmd.substr(w[0].size())));
        return 0;
    }
    if (verb == "quit" || verb == "exit") {
        join_worker();
        if (g.exitfn) g.exitfn(0, false, true, 0, g.user);
        return 0;
    }
    say("stderr", "Error: " + verb + ": command not available in this library");
    return 1;
}

pvector_info ngGet_Vec_Info(char *vecname)
{
    Shared &g = G();
    if (!vecname) return nullptr;
    std::string n = lower(trim(vecname));
    std::string alt;
    if (n.size() > 3 && (n.rfind("v(", 0) == 0) && n.back() == ')') n = n.substr(2, n.size() - 3);
    else if (n.size() > 3 && (n.rfind("i(", 0) == 0) && n.back() == ')') n = n.substr(2, n.size() - 3) + "#branch";
    if (n.find('.') != std::string::npos && n.rfind("tran1.", 0) == 0) n = n.substr(6);
    std::lock_guard<std::mutex> lk(g.mu);
    for (Vec &v : g.vecs)
        if (v.name == n) return &v.info;
    return nullptr;
}

char *ngSpice_CurPlot(void) { return const_cast<char *>(G().plot.c_str()); }

char **ngSpice_AllPlots(void)
{
    Shared &g = G();
    std::lock_guard<std::mutex> lk(g.mu);
    static char cst[] = "const";
    static char tr[] = "tran1";
    g.plot_list.clear();
    if (!g.vecs.empty()) g.plot_list.push_back(tr);
    g.plot_list.push_back(cst);
    g.plot_list.push_back(nullptr);
    return g.plot_list.data();
}

char **ngSpice_AllVecs(char *plotname)
{
    Shared &g = G();
    std::lock_guard<std::mutex> lk(g.mu);
    g.vec_list.clear();
    if (plotname && lower(plotname) == "tran1")
        for (Vec &v : g.vecs) g.vec_list.push_back(const_cast<char *>(v.name.c_str()));
    g.vec_list.push_back(nullptr);
    return g.vec_list.data();
}

NG_BOOL ngSpice_running(void) { return G().running.load(); }

}  // extern "C"
