/*
 * SPICE deck front end for the batched transient engine: the subset of
 * ngspice's input language a BSIM4 transient deck uses, turned into an
 * ngmi_circuit with the MNA numbering ngspice gives the same deck.
 *
 * Cards: title (first line), '*' comments, '+' continuations, inline
 * comments after ';' or " $ ", M / C / R / V / X elements, .subckt/.ends,
 * .model <name> nmos|pmos level=54 (or 14), .param, .tran, .options,
 * .temp, .ic v(node)=value, .global, .end.  Names are case-insensitive
 * (lower-cased, as ngspice does); "gnd" is ground, like "0".
 *
 * Numbering (what a raw file or a KLU pivot sequence refers to):
 *   - nodes are numbered 1, 2, ... in order of first appearance as the
 *     flattened deck's element cards are parsed, terminals in card order
 *     (INPpas2 -> INPtermInsert -> CKTnewNode, inppas2.c, inpsymt.c,
 *     cktnewn.c);
 *   - subcircuits are flattened in place (inp_subcktexpand,
 *     frontend/subckt.c); a node local to instance Xname is "name.node"
 *     (translate_node_name, subckt.c:1107-1127), nested instances prefix
 *     their path the same way;
 *   - voltage-source branch unknowns follow the node unknowns in VSRCsetup
 *     order (vsrcset.c:41-44), which walks the model's instance list; new
 *     instances are put at its head (cktcrte.c:62-64), so branches come in
 *     reverse deck order.
 *
 * Values: SPICE numbers with scale suffixes (t g meg k mil m u n p f a,
 * trailing unit letters ignored), parameters, and expressions in braces or
 * single quotes with + - * / ^ ** and the usual functions.  The Monte
 * Carlo functions agauss / gauss / aunif / unif draw from a generator
 * seeded by the caller (not ngspice's, whose sequence is not reproduced).
 */
#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <set>
#include <string>
#include <vector>

#include "ngmi_internal.h"

using namespace ngmi;

namespace {

/* ---------------- numbers and expressions ---------------- */

/* SPICE number at s[pos..]: digits, fraction, exponent, scale suffix,
 * then any letters (units); false if no number starts there */
bool spice_number(const std::string &s, size_t &pos, double &out)
{
    size_t p = pos;
    const size_t n = s.size();
    size_t q = p;
    if (q < n && (s[q] == '+' || s[q] == '-')) q++;
    const size_t d0 = q;
    while (q < n && std::isdigit((unsigned char)s[q])) q++;
    if (q < n && s[q] == '.') {
        q++;
        while (q < n && std::isdigit((unsigned char)s[q])) q++;
    }
    if (q == d0 || (q == d0 + 1 && s[d0] == '.')) return false;
    if (q < n && (s[q] == 'e' || s[q] == 'E')) {
        size_t r = q + 1;
        if (r < n && (s[r] == '+' || s[r] == '-')) r++;
        if (r < n && std::isdigit((unsigned char)s[r])) {
            while (r < n && std::isdigit((unsigned char)s[r])) r++;
            q = r;
        }
    }
    double v = std::strtod(s.substr(p, q - p).c_str(), nullptr);
    auto pre = [&](const char *t) { return s.compare(q, std::strlen(t), t) == 0; };
    if (pre("meg")) { v *= 1e6; q += 3; }
    else if (pre("mil")) { v *= 25.4e-6; q += 3; }
    else if (q < n) {
        switch (s[q]) {
        case 't': v *= 1e12; q++; break;
        case 'g': v *= 1e9; q++; break;
        case 'k': v *= 1e3; q++; break;
        case 'm': v *= 1e-3; q++; break;
        case 'u': v *= 1e-6; q++; break;
        case 'n': v *= 1e-9; q++; break;
        case 'p': v *= 1e-12; q++; break;
        case 'f': v *= 1e-15; q++; break;
        case 'a': v *= 1e-18; q++; break;
        default: break;
        }
    }
    while (q < n && std::isalpha((unsigned char)s[q])) q++;
    pos = q;
    out = v;
    return true;
}

/* seeded normal / uniform draws for the Monte Carlo functions */
struct Draws {
    uint64_t state;
    uint64_t next()
    {   /* splitmix64 */
        uint64_t z = (state += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    }
    double uniform() { return ((next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
    double normal()
    {   /* Box-Muller, one value per call */
        const double u1 = uniform(), u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

struct Scope {
    const Scope *up = nullptr;
    std::map<std::string, double> v;
    bool find(const std::string &k, double &out) const
    {
        for (const Scope *s = this; s; s = s->up) {
            auto it = s->v.find(k);
            if (it != s->v.end()) {
                out = it->second;
                return true;
            }
        }
        return false;
    }
};

struct Expr {
    const std::string &s;
    size_t p = 0;
    const Scope &sc;
    Draws &rng;
    std::string err;

    Expr(const std::string &src, const Scope &scope, Draws &r) : s(src), sc(scope), rng(r) {}
    void ws() { while (p < s.size() && std::isspace((unsigned char)s[p])) p++; }
    bool eat(char c)
    {
        ws();
        if (p < s.size() && s[p] == c) { p++; return true; }
        return false;
    }
    double fail(const std::string &m)
    {
        if (err.empty()) err = m;
        return 0.0;
    }
    double call(const std::string &f, const std::vector<double> &a)
    {
        auto need = [&](size_t k) { return a.size() == k ? true : (fail(f + ": wrong argument count"), false); };
        if (f == "sqrt" && need(1)) return std::sqrt(a[0]);
        if (f == "exp" && need(1)) return std::exp(a[0]);
        if ((f == "log" || f == "ln") && need(1)) return std::log(a[0]);
        if (f == "log10" && need(1)) return std::log10(a[0]);
        if (f == "abs" && need(1)) return std::fabs(a[0]);
        if (f == "sin" && need(1)) return std::sin(a[0]);
        if (f == "cos" && need(1)) return std::cos(a[0]);
        if (f == "tan" && need(1)) return std::tan(a[0]);
        if (f == "atan" && need(1)) return std::atan(a[0]);
        if (f == "floor" && need(1)) return std::floor(a[0]);
        if (f == "ceil" && need(1)) return std::ceil(a[0]);
        if (f == "int" && need(1)) return std::trunc(a[0]);
        if (f == "min" && need(2)) return std::fmin(a[0], a[1]);
        if (f == "max" && need(2)) return std::fmax(a[0], a[1]);
        if ((f == "pow" || f == "pwr") && need(2)) return std::pow(a[0], a[1]);
        /* Monte Carlo: nominal, variation, sigma (ngspice's agauss/gauss
         * meaning: the variation is at `sigma` standard deviations) */
        if (f == "agauss" && need(3)) return a[0] + a[1] / a[2] * rng.normal();
        if (f == "gauss" && need(3)) return a[0] * (1.0 + a[1] / a[2] * rng.normal());
        if (f == "aunif" && need(2)) return a[0] + a[1] * (2.0 * rng.uniform() - 1.0);
        if (f == "unif" && need(2)) return a[0] * (1.0 + a[1] * (2.0 * rng.uniform() - 1.0));
        if (err.empty()) fail("unknown function " + f);
        return 0.0;
    }
    double atom()
    {
        ws();
        if (p >= s.size()) return fail("expression ends early");
        if (s[p] == '(' || s[p] == '{') {
            const char close = s[p] == '(' ? ')' : '}';
            p++;
            const double v = sum();
            if (!eat(close)) return fail(std::string("missing ") + close);
            return v;
        }
        if (s[p] == '-') { p++; return -power(); }
        if (s[p] == '+') { p++; return power(); }
        double v;
        if (std::isdigit((unsigned char)s[p]) || s[p] == '.') {
            size_t q = p;
            if (!spice_number(s, q, v)) return fail("bad number");
            p = q;
            return v;
        }
        if (std::isalpha((unsigned char)s[p]) || s[p] == '_') {
            size_t q = p;
            while (q < s.size() && (std::isalnum((unsigned char)s[q]) || s[q] == '_' || s[q] == '.')) q++;
            const std::string id = s.substr(p, q - p);
            p = q;
            if (eat('(')) {
                std::vector<double> a;
                if (!eat(')')) {
                    do a.push_back(sum()); while (eat(','));
                    if (!eat(')')) return fail("missing ) after arguments of " + id);
                }
                return call(id, a);
            }
            if (id == "pi") return 3.141592653589793;
            if (!sc.find(id, v)) return fail("unknown parameter " + id);
            return v;
        }
        return fail(std::string("unexpected '") + s[p] + "'");
    }
    double power()
    {
        const double b = atom();
        ws();
        if (p < s.size() && s[p] == '^') { p++; return std::pow(b, power()); }
        if (p + 1 < s.size() && s[p] == '*' && s[p + 1] == '*') { p += 2; return std::pow(b, power()); }
        return b;
    }
    double product()
    {
        double v = power();
        for (;;) {
            ws();
            if (p < s.size() && s[p] == '*' && !(p + 1 < s.size() && s[p + 1] == '*')) { p++; v *= power(); }
            else if (p < s.size() && s[p] == '/') { p++; v /= power(); }
            else return v;
        }
    }
    double sum()
    {
        double v = product();
        for (;;) {
            ws();
            if (p < s.size() && s[p] == '+') { p++; v += product(); }
            else if (p < s.size() && s[p] == '-') { p++; v -= product(); }
            else return v;
        }
    }
};

/* a value token: number, parameter name, or {expr} / 'expr' */
bool eval_value(const std::string &tok, const Scope &sc, Draws &rng, double &out, std::string &err)
{
    std::string t = tok;
    if (t.size() >= 2 && ((t.front() == '{' && t.back() == '}') || (t.front() == '\'' && t.back() == '\'')))
        t = t.substr(1, t.size() - 2);
    size_t q = 0;
    double v;
    if (spice_number(t, q, v) && q == t.size()) {
        out = v;
        return true;
    }
    Expr e(t, sc, rng);
    v = e.sum();
    e.ws();
    if (e.err.empty() && e.p != t.size()) e.err = "trailing text in '" + tok + "'";
    if (!e.err.empty()) {
        err = e.err;
        return false;
    }
    out = v;
    return true;
}

/* ---------------- cards ---------------- */

struct Card {
    int line;
    std::vector<std::string> tok;
    std::string raw;   /* the logical line, lower-cased */
};

bool ident_char(char ch) { return std::isalnum((unsigned char)ch) || ch == '_' || ch == '.'; }

/* .param assignments from the raw text after the keyword: name = expr,
 * where expr runs (parentheses and braces balanced) up to the next
 * "name =" at top level or the end of the card */
bool param_pairs(const std::string &raw, size_t pos,
                 std::vector<std::pair<std::string, std::string>> &out, std::string &err)
{
    const size_t n = raw.size();
    auto skip = [&](size_t q) { while (q < n && std::isspace((unsigned char)raw[q])) q++; return q; };
    auto assign_at = [&](size_t q) {   /* "ident =" starts at q? */
        if (q >= n || !(std::isalpha((unsigned char)raw[q]) || raw[q] == '_')) return false;
        while (q < n && ident_char(raw[q])) q++;
        q = skip(q);
        return q < n && raw[q] == '=' && !(q + 1 < n && raw[q + 1] == '=');
    };
    size_t q = skip(pos);
    while (q < n) {
        if (!assign_at(q)) {
            err = "expected name=value";
            return false;
        }
        size_t a = q;
        while (q < n && ident_char(raw[q])) q++;
        const std::string name = raw.substr(a, q - a);
        q = skip(q) + 1;   /* past '=' */
        q = skip(q);
        const size_t e0 = q;
        int depth = 0;
        while (q < n) {
            const char ch = raw[q];
            if (ch == '(' || ch == '{') depth++;
            else if (ch == ')' || ch == '}') depth--;
            else if (ch == '\'') {
                const size_t z = raw.find('\'', q + 1);
                q = z == std::string::npos ? n : z + 1;
                continue;
            } else if (depth == 0 && std::isspace((unsigned char)ch)) {
                const size_t nx = skip(q);
                if (nx >= n || assign_at(nx)) break;
            }
            q++;
        }
        std::string ex = raw.substr(e0, q - e0);
        while (!ex.empty() && std::isspace((unsigned char)ex.back())) ex.pop_back();
        if (ex.empty()) {
            err = "missing value after " + name + "=";
            return false;
        }
        out.push_back({name, ex});
        q = skip(q);
    }
    return true;
}

/* tokens: whitespace, ',' separate; '(' ')' '=' are tokens of their own;
 * {...} and '...' stay whole */
std::vector<std::string> tokenize(const std::string &s)
{
    std::vector<std::string> out;
    std::string cur;
    auto flush = [&]() {
        if (!cur.empty()) out.push_back(cur);
        cur.clear();
    };
    for (size_t i = 0; i < s.size(); i++) {
        const char ch = s[i];
        if (ch == '{' || ch == '\'') {
            const char close = ch == '{' ? '}' : '\'';
            flush();
            size_t j = s.find(close, i + 1);
            if (j == std::string::npos) j = s.size() - 1;
            out.push_back(s.substr(i, j - i + 1));
            i = j;
        } else if (std::isspace((unsigned char)ch) || ch == ',') {
            flush();
        } else if (ch == '(' || ch == ')' || ch == '=') {
            flush();
            out.push_back(std::string(1, ch));
        } else {
            cur += ch;
        }
    }
    flush();
    return out;
}

std::string lower(std::string s)
{
    for (char &c : s) c = (char)std::tolower((unsigned char)c);
    return s;
}

struct Subckt {
    std::vector<std::string> ports;
    std::vector<std::pair<std::string, std::string>> params;   /* defaults, unevaluated */
    std::vector<Card> body;
    std::map<std::string, std::shared_ptr<Subckt>> subs;     /* local definitions */
    std::vector<Card> models;                                /* local .model cards */
};

/* one flattened element */
struct Elt {
    char kind;                  /* m c r v */
    std::string name;
    std::vector<std::string> nodes;
    std::string model;          /* m: resolved model key */
    double value = 0.0;         /* c r: value; v: dc */
    bool dc_given = false;
    std::vector<double> pulse;  /* v */
    std::vector<std::pair<std::string, double>> params;   /* m: instance parameters */
};

}  // namespace

struct ngmi_netlist {
    std::string title;
    std::vector<Elt> elts;
    std::vector<std::string> names;        /* unknown k -> name; [0] = "0" */
    std::map<std::string, int> index;      /* name -> unknown */
    int nnodes = 0;
    std::map<std::string, ngmi_bsim4_model *> models;
    double tstep = 0, tstop = 0, tstart = 0, tmax = 0;
    int uic = 0, have_tran = 0;
    double reltol = 1e-3, abstol = 1e-12, vntol = 1e-6, chgtol = 1e-14, trtol = 7.0, gmin = 1e-12;
    int method = 1, maxord = 2;
    int noopiter = 0, gminsteps = 1, srcsteps = 1, itl2 = 50;
    double gminfactor = 10.0, gshunt = 0.0, pivrel = 1e-3;
    double temp = 300.15, tnom = 300.15;
    std::vector<std::pair<int, double>> ic;
    ~ngmi_netlist()
    {
        for (auto &m : models) ngmi_bsim4_model_free(m.second);
    }
};

namespace {

struct Parser {
    ngmi_netlist *nl;
    Draws rng;
    std::string err;
    std::set<std::string> globals;
    Scope top;
    /* model cards by scoped key, unevaluated (evaluated once, at top scope
     * or at the scope of the subcircuit instance that owns them) */
    std::map<std::string, std::shared_ptr<Subckt>> subs;
    std::vector<Card> top_models;
    std::vector<Card> top_body;

    bool fail(int line, const std::string &m)
    {
        if (err.empty()) err = "line " + std::to_string(line) + ": " + m;
        return false;
    }

    /* logical lines: continuation joined, comments dropped, lower-cased */
    std::vector<Card> read(const char *text)
    {
        std::vector<std::pair<int, std::string>> lines;
        int ln = 0;
        const char *s = text;
        while (*s) {
            const char *e = std::strchr(s, '\n');
            std::string l = e ? std::string(s, e - s) : std::string(s);
            s = e ? e + 1 : s + l.size();
            ln++;
            if (!l.empty() && l.back() == '\r') l.pop_back();
            if (ln == 1) {
                nl->title = l;
                continue;
            }
            size_t a = l.find_first_not_of(" \t");
            if (a == std::string::npos) continue;
            if (l[a] == '*') continue;
            for (size_t k = a; k < l.size(); k++) {   /* inline comments */
                if (l[k] == ';' || (l[k] == '$' && k > 0 && std::isspace((unsigned char)l[k - 1]))) {
                    l.resize(k);
                    break;
                }
            }
            l = lower(l.substr(a));
            if (l.empty()) continue;
            if (l[0] == '+') {
                if (!lines.empty()) lines.back().second += " " + l.substr(1);
                continue;
            }
            lines.push_back({ln, l});
        }
        std::vector<Card> cards;
        for (auto &x : lines) cards.push_back({x.first, tokenize(x.second), x.second});
        return cards;
    }

    /* key = value pairs from tokens[i..] */
    bool pairs(const Card &c, size_t i, std::vector<std::pair<std::string, std::string>> &out,
               bool allow_flags = false)
    {
        while (i < c.tok.size()) {
            if (c.tok[i] == "(" || c.tok[i] == ")") { i++; continue; }
            if (i + 1 < c.tok.size() && c.tok[i + 1] == "=") {
                if (i + 2 >= c.tok.size()) return fail(c.line, "missing value after " + c.tok[i] + "=");
                out.push_back({c.tok[i], c.tok[i + 2]});
                i += 3;
            } else if (allow_flags) {
                out.push_back({c.tok[i], ""});
                i++;
            } else {
                return fail(c.line, "expected name=value, got '" + c.tok[i] + "'");
            }
        }
        return true;
    }

    bool params(const Card &c, Scope &sc)
    {
        std::vector<std::pair<std::string, std::string>> pr;
        std::string e;
        const size_t at = c.raw.find(".param") + 6;
        if (!param_pairs(c.raw, at, pr, e)) return fail(c.line, e);
        for (auto &kv : pr) {
            double v;
            if (!value(c, "{" + kv.second + "}", sc, v)) return false;
            sc.v[kv.first] = v;
        }
        return true;
    }

    bool value(const Card &c, const std::string &tok, const Scope &sc, double &v)
    {
        std::string e;
        if (!eval_value(tok, sc, rng, v, e)) return fail(c.line, e);
        return true;
    }

    /* split the deck into top-level cards, subcircuit definitions and models */
    bool structure(const std::vector<Card> &cards, size_t &i, Subckt *into, bool toplevel)
    {
        for (; i < cards.size(); i++) {
            const Card &c = cards[i];
            const std::string &k = c.tok[0];
            if (k == ".end") {
                if (!toplevel) return fail(c.line, ".end inside .subckt");
                i = cards.size();
                return true;
            }
            if (k == ".ends") {
                if (toplevel) return fail(c.line, ".ends without .subckt");
                return true;
            }
            if (k == ".subckt") {
                if (c.tok.size() < 2) return fail(c.line, ".subckt needs a name");
                auto sub = std::make_shared<Subckt>();
                size_t t = 2;
                for (; t < c.tok.size(); t++) {
                    if (c.tok[t] == "params:") { t++; break; }
                    if (t + 1 < c.tok.size() && c.tok[t + 1] == "=") break;
                    sub->ports.push_back(c.tok[t]);
                }
                std::vector<std::pair<std::string, std::string>> pr;
                if (!pairs(c, t, pr)) return false;
                sub->params = pr;
                i++;
                if (!structure(cards, i, sub.get(), false)) return false;
                if (i >= cards.size()) return fail(c.line, ".subckt " + c.tok[1] + " has no .ends");
                (into ? into->subs : subs)[c.tok[1]] = sub;
                continue;
            }
            if (k == ".model") {
                (into ? into->models : top_models).push_back(c);
                continue;
            }
            if (into) {
                if (k[0] == '.' && k != ".param")
                    return fail(c.line, k + " inside .subckt is not supported");
                into->body.push_back(c);
            } else {
                top_body.push_back(c);
            }
        }
        return true;
    }

    /* .model card -> engine model (evaluated in scope sc) */
    bool make_model(const Card &c, const std::string &key, const Scope &sc)
    {
        if (c.tok.size() < 3) return fail(c.line, ".model needs a name and a type");
        const std::string &type = c.tok[2];
        if (type != "nmos" && type != "pmos")
            return fail(c.line, "model type '" + type + "' is not supported (nmos/pmos BSIM4 only)");
        std::vector<std::pair<std::string, std::string>> pr;
        if (!pairs(c, 3, pr)) return false;
        ngmi_bsim4_model *m = ngmi_bsim4_model_new();
        nl->models[key] = m;
        ngmi_bsim4_model_param(m, type.c_str(), 1.0);
        bool level_ok = false;
        for (auto &kv : pr) {
            if (kv.first == "version") continue;   /* the engine is BSIM4 4.8.3 */
            double v;
            if (!value(c, kv.second, sc, v)) return false;
            if (kv.first == "level") {
                if (v != 54 && v != 14) return fail(c.line, "only BSIM4 (level=54 or 14) is supported");
                level_ok = true;
                continue;
            }
            if (ngmi_bsim4_model_param(m, kv.first.c_str(), v) != NGMI_OK)
                return fail(c.line, "unknown BSIM4 model parameter '" + kv.first + "'");
        }
        if (!level_ok) return fail(c.line, "model " + c.tok[1] + " needs level=54");
        if (ngmi_bsim4_model_setup(m, nl->tnom, nl->temp) != NGMI_OK)
            return fail(c.line, std::string("model setup: ") + ngmi_last_error());
        return true;
    }

    struct Frame {
        std::string path;                          /* "" at top, "a." inside Xa */
        std::map<std::string, std::string> port;   /* formal -> actual node */
        const Subckt *def = nullptr;
        const Frame *up = nullptr;
        Scope *scope = nullptr;
        std::map<std::string, std::string> local_models;   /* card name -> model key */
    };

    std::string node(const Frame &f, const std::string &n)
    {
        if (n == "0" || n == "gnd") return "0";
        auto it = f.port.find(n);
        if (it != f.port.end()) return it->second;
        if (globals.count(n)) return n;
        return f.path + n;
    }

    std::string model_key(const Frame &f, const std::string &name)
    {
        for (const Frame *g = &f; g; g = g->up) {
            auto it = g->local_models.find(name);
            if (it != g->local_models.end()) return it->second;
        }
        return name;
    }

    const Subckt *find_sub(const Frame &f, const std::string &name)
    {
        for (const Frame *g = &f; g; g = g->up) {
            if (!g->def) continue;
            auto it = g->def->subs.find(name);
            if (it != g->def->subs.end()) return it->second.get();
        }
        auto it = subs.find(name);
        return it == subs.end() ? nullptr : it->second.get();
    }

    bool element(const Card &c, Frame &f, int depth)
    {
        const std::string &name = c.tok[0];
        const char kind = name[0];
        Scope &sc = *f.scope;
        if (kind == 'x') {
            /* nodes..., subckt name, then name=value overrides */
            size_t t = 1;
            while (t < c.tok.size() && !(t + 1 < c.tok.size() && c.tok[t + 1] == "=")) t++;
            if (t < 2) return fail(c.line, "subcircuit instance needs a subcircuit name");
            const std::string subname = c.tok[t - 1];
            const Subckt *sub = find_sub(f, subname);
            if (!sub) return fail(c.line, "unknown subcircuit " + subname);
            if (t - 2 != sub->ports.size())
                return fail(c.line, "subcircuit " + subname + " has " + std::to_string(sub->ports.size()) +
                                        " ports, instance gives " + std::to_string(t - 2));
            if (depth > 64) return fail(c.line, "subcircuits nested too deep");
            std::vector<std::pair<std::string, std::string>> over;
            if (!pairs(c, t, over)) return false;
            Scope inner;
            inner.up = &top;   /* a subcircuit sees global parameters and its own */
            for (auto &kv : sub->params) {
                double v;
                if (!value(c, kv.second, inner, v)) return false;
                inner.v[kv.first] = v;
            }
            for (auto &kv : over) {
                double v;
                if (!value(c, kv.second, sc, v)) return false;   /* caller's scope */
                inner.v[kv.first] = v;
            }
            Frame g;
            g.path = f.path + name.substr(1) + ".";
            g.def = sub;
            g.up = &f;
            g.scope = &inner;
            for (size_t k = 0; k < sub->ports.size(); k++) g.port[sub->ports[k]] = node(f, c.tok[1 + k]);
            for (const Card &m : sub->models) {
                if (m.tok.size() < 2) return fail(m.line, ".model needs a name");
                const std::string key = g.path + m.tok[1];
                g.local_models[m.tok[1]] = key;
                if (!make_model(m, key, inner)) return false;
            }
            for (const Card &b : sub->body)
                if (!card(b, g, depth + 1)) return false;
            return true;
        }
        Elt e;
        e.kind = kind;
        e.name = f.path.empty() ? name : std::string(1, kind) + "." + f.path + name;
        if (kind == 'm') {
            if (c.tok.size() < 6) return fail(c.line, "MOSFET needs d g s b and a model");
            for (int k = 1; k <= 4; k++) e.nodes.push_back(node(f, c.tok[k]));
            e.model = model_key(f, c.tok[5]);
            if (!nl->models.count(e.model)) return fail(c.line, "unknown model " + c.tok[5]);
            std::vector<std::pair<std::string, std::string>> pr;
            if (!pairs(c, 6, pr)) return false;
            for (auto &kv : pr) {
                double v;
                if (!value(c, kv.second, sc, v)) return false;
                e.params.push_back({kv.first == "mult" ? "m" : kv.first, v});
            }
        } else if (kind == 'c' || kind == 'r') {
            if (c.tok.size() < 4) return fail(c.line, "element needs two nodes and a value");
            e.nodes = {node(f, c.tok[1]), node(f, c.tok[2])};
            size_t t = 3;
            if (c.tok[t] == (kind == 'c' ? "c" : "r") && t + 2 < c.tok.size() && c.tok[t + 1] == "=") t += 2;
            if (!value(c, c.tok[t], sc, e.value)) return false;
            for (size_t k = t + 1; k < c.tok.size(); k++)
                return fail(c.line, "element parameter '" + c.tok[k] + "' is not supported");
        } else if (kind == 'v') {
            if (c.tok.size() < 3) return fail(c.line, "voltage source needs two nodes");
            e.nodes = {node(f, c.tok[1]), node(f, c.tok[2])};
            size_t t = 3;
            while (t < c.tok.size()) {
                const std::string &k = c.tok[t];
                if (k == "dc") {
                    if (t + 1 >= c.tok.size()) return fail(c.line, "dc needs a value");
                    if (!value(c, c.tok[t + 1], sc, e.value)) return false;
                    e.dc_given = true;
                    t += 2;
                } else if (k == "pulse") {
                    t++;
                    const bool paren = t < c.tok.size() && c.tok[t] == "(";
                    if (paren) t++;
                    while (t < c.tok.size() && c.tok[t] != ")") {
                        double v;
                        if (e.pulse.size() >= 7) return fail(c.line, "pulse takes at most 7 values");
                        if (!value(c, c.tok[t], sc, v)) return false;
                        e.pulse.push_back(v);
                        t++;
                        if (!paren && e.pulse.size() == 7) break;
                    }
                    if (paren) {
                        if (t >= c.tok.size()) return fail(c.line, "pulse( without )");
                        t++;
                    }
                    if (e.pulse.size() < 2) return fail(c.line, "pulse needs v1 and v2");
                } else if (k == "ac" || k == "sin" || k == "pwl" || k == "exp" || k == "sffm" || k == "am") {
                    return fail(c.line, "source function '" + k + "' is not supported");
                } else if (!e.dc_given) {
                    if (!value(c, k, sc, e.value)) return false;
                    e.dc_given = true;
                    t++;
                } else {
                    return fail(c.line, "unexpected '" + k + "'");
                }
            }
        } else {
            return fail(c.line, std::string("element type '") + kind + "' is not supported");
        }
        nl->elts.push_back(std::move(e));
        return true;
    }

    bool card(const Card &c, Frame &f, int depth)
    {
        const std::string &k = c.tok[0];
        if (k[0] != '.') return element(c, f, depth);
        if (k == ".param") return params(c, *f.scope);
        return fail(c.line, k + " inside .subckt is not supported");
    }

    bool control(const Card &c)
    {
        const std::string &k = c.tok[0];
        if (k == ".tran") {
            std::vector<double> v;
            for (size_t t = 1; t < c.tok.size(); t++) {
                if (c.tok[t] == "uic") { nl->uic = 1; continue; }
                double x;
                if (!value(c, c.tok[t], top, x)) return false;
                v.push_back(x);
            }
            if (v.size() < 2) return fail(c.line, ".tran needs tstep and tstop");
            nl->tstep = v[0];
            nl->tstop = v[1];
            nl->tstart = v.size() > 2 ? v[2] : 0.0;
            nl->tmax = v.size() > 3 ? v[3] : 0.0;
            nl->have_tran = 1;
            return true;
        }
        if (k == ".options" || k == ".option" || k == ".opt") {
            std::vector<std::pair<std::string, std::string>> pr;
            if (!pairs(c, 1, pr, true)) return false;
            for (auto &kv : pr) {
                const std::string &o = kv.first;
                if (kv.second.empty()) {   /* flags */
                    if (o == "noopiter") nl->noopiter = 1;
                    else if (o == "klu" || o == "acct" || o == "noacct" || o == "nopage" || o == "savecurrents") {}
                    else return fail(c.line, "option '" + o + "' is not supported");
                    continue;
                }
                if (o == "method") {
                    if (kv.second == "trap" || kv.second == "trapezoidal") nl->method = 1;
                    else if (kv.second == "gear") nl->method = 2;
                    else return fail(c.line, "method '" + kv.second + "' is not supported");
                    continue;
                }
                double v;
                if (!value(c, kv.second, top, v)) return false;
                if (o == "reltol") nl->reltol = v;
                else if (o == "abstol") nl->abstol = v;
                else if (o == "vntol") nl->vntol = v;
                else if (o == "chgtol") nl->chgtol = v;
                else if (o == "trtol") nl->trtol = v;
                else if (o == "gmin") nl->gmin = v;
                else if (o == "maxord") nl->maxord = (int)v;
                else if (o == "gminsteps") nl->gminsteps = (int)v;
                else if (o == "srcsteps") nl->srcsteps = (int)v;
                else if (o == "itl2") nl->itl2 = (int)v;
                else if (o == "gminfactor") nl->gminfactor = v;
                else if (o == "gshunt") nl->gshunt = v;
                else if (o == "pivrel") nl->pivrel = v;
                else if (o == "temp") nl->temp = v + 273.15;
                else if (o == "tnom") nl->tnom = v + 273.15;
                else if (o == "seed" || o == "itl1" || o == "itl4") {}
                else return fail(c.line, "option '" + o + "' is not supported");
            }
            return true;
        }
        if (k == ".temp") {
            double v;
            if (c.tok.size() < 2 || !value(c, c.tok[1], top, v)) return fail(c.line, ".temp needs a value");
            nl->temp = v + 273.15;
            return true;
        }
        if (k == ".global") {
            for (size_t t = 1; t < c.tok.size(); t++) globals.insert(c.tok[t]);
            return true;
        }
        if (k == ".ic" || k == ".param") return true;   /* handled in their own passes */
        return fail(c.line, "card " + k + " is not supported");
    }

    bool run(const char *text)
    {
        const std::vector<Card> cards = read(text);
        size_t i = 0;
        if (!structure(cards, i, nullptr, true)) return false;
        /* options, temperature, globals and top-level parameters first
         * (ngspice reads .options and .param before it builds devices) */
        for (const Card &c : top_body)
            if (c.tok[0] == ".param" && !params(c, top)) return false;
        for (const Card &c : top_body)
            if (c.tok[0][0] == '.' && c.tok[0] != ".ic" && !control(c)) return false;
        Frame f;
        f.scope = &top;
        for (const Card &m : top_models) {
            if (m.tok.size() < 2) return fail(m.line, ".model needs a name");
            if (!make_model(m, m.tok[1], top)) return false;
        }
        for (const Card &c : top_body)
            if (c.tok[0][0] != '.' && !element(c, f, 0)) return false;
        /* numbering: nodes by first appearance, then source branches in
         * reverse deck order */
        nl->names.push_back("0");
        nl->index["0"] = 0;
        for (const Elt &e : nl->elts)
            for (const std::string &n : e.nodes)
                if (!nl->index.count(n)) {
                    nl->index[n] = (int)nl->names.size();
                    nl->names.push_back(n);
                }
        nl->nnodes = (int)nl->names.size() - 1;
        for (size_t k = nl->elts.size(); k-- > 0;) {
            const Elt &e = nl->elts[k];
            if (e.kind != 'v') continue;
            const std::string bn = e.name + "#branch";
            nl->index[bn] = (int)nl->names.size();
            nl->names.push_back(bn);
        }
        /* .ic v(node)=value */
        for (const Card &c : top_body) {
            if (c.tok[0] != ".ic") continue;
            size_t t = 1;
            while (t < c.tok.size()) {
                if (t + 5 >= c.tok.size() || c.tok[t] != "v" || c.tok[t + 1] != "(" || c.tok[t + 3] != ")" || c.tok[t + 4] != "=")
                    return fail(c.line, ".ic expects v(node)=value");
                const std::string n = c.tok[t + 2] == "gnd" ? "0" : c.tok[t + 2];
                auto it = nl->index.find(n);
                if (it == nl->index.end() || it->second == 0 || it->second > nl->nnodes)
                    return fail(c.line, ".ic on unknown node " + c.tok[t + 2]);
                double v;
                if (!value(c, c.tok[t + 5], top, v)) return false;
                nl->ic.push_back({it->second, v});
                t += 6;
            }
        }
        return true;
    }
};

}  // namespace

extern "C" {

ngmi_netlist *ngmi_netlist_parse(const char *text, uint64_t seed)
{
    if (!text) {
        set_error("ngmi_netlist_parse: no text");
        return nullptr;
    }
    std::unique_ptr<ngmi_netlist> nl(new ngmi_netlist);
    Parser P;
    P.nl = nl.get();
    P.rng.state = seed;
    if (!P.run(text)) {
        set_error("%s", P.err.c_str());
        return nullptr;
    }
    return nl.release();
}

void ngmi_netlist_free(ngmi_netlist *nl) { delete nl; }

const char *ngmi_netlist_title(const ngmi_netlist *nl) { return nl ? nl->title.c_str() : ""; }

int ngmi_netlist_size(const ngmi_netlist *nl, int *n_nodes, int *n_unknowns, int *n_elements)
{
    if (!nl) return NGMI_E_PANIC;
    if (n_nodes) *n_nodes = nl->nnodes;
    if (n_unknowns) *n_unknowns = (int)nl->names.size() - 1;
    if (n_elements) *n_elements = (int)nl->elts.size();
    return NGMI_OK;
}

const char *ngmi_netlist_name(const ngmi_netlist *nl, int k)
{
    if (!nl || k < 0 || k >= (int)nl->names.size()) return nullptr;
    return nl->names[k].c_str();
}

int ngmi_netlist_find(const ngmi_netlist *nl, const char *name)
{
    if (!nl || !name) return -1;
    auto it = nl->index.find(lower(name));
    return it == nl->index.end() ? -1 : it->second;
}

int ngmi_netlist_tran(const ngmi_netlist *nl, double *tstep, double *tstop, double *tstart,
                      double *tmax, int *uic)
{
    if (!nl) return NGMI_E_PANIC;
    if (!nl->have_tran) {
        set_error("the deck has no .tran card");
        return NGMI_E_PARMVAL;
    }
    if (tstep) *tstep = nl->tstep;
    if (tstop) *tstop = nl->tstop;
    if (tstart) *tstart = nl->tstart;
    if (tmax) *tmax = nl->tmax;
    if (uic) *uic = nl->uic;
    return NGMI_OK;
}

int ngmi_netlist_element(const ngmi_netlist *nl, int k, char *kind, int *nodes, double *value)
{
    if (!nl || k < 0 || k >= (int)nl->elts.size()) return -1;
    const Elt &e = nl->elts[k];
    if (kind) *kind = e.kind;
    if (nodes)
        for (size_t q = 0; q < e.nodes.size(); q++) nodes[q] = nl->index.at(e.nodes[q]);
    if (value) *value = e.value;
    return (int)e.nodes.size();
}

const char *ngmi_netlist_element_name(const ngmi_netlist *nl, int k)
{
    if (!nl || k < 0 || k >= (int)nl->elts.size()) return nullptr;
    return nl->elts[k].name.c_str();
}

ngmi_circuit *ngmi_netlist_circuit(const ngmi_netlist *nl)
{
    if (!nl) return nullptr;
    ngmi_circuit *c = ngmi_circuit_new(nl->nnodes);
    if (!c) return nullptr;
    auto bail = [&](const std::string &what) -> ngmi_circuit * {
        set_error("%s: %s", what.c_str(), ngmi_last_error());
        ngmi_circuit_free(c);
        return nullptr;
    };
    if (ngmi_circuit_set_options(c, nl->reltol, nl->abstol, nl->vntol, nl->chgtol, nl->trtol, nl->gmin,
                                 nl->method, nl->maxord) != NGMI_OK)
        return bail("options");
    if (ngmi_circuit_set_op_options(c, nl->noopiter, nl->gminsteps, nl->srcsteps, nl->gminfactor,
                                    nl->gshunt, nl->itl2) != NGMI_OK)
        return bail("operating-point options");
    if (ngmi_circuit_set_pivrel(c, nl->pivrel) != NGMI_OK) return bail("pivrel");
    if (ngmi_circuit_set_uic(c, nl->uic) != NGMI_OK) return bail("uic");
    auto num = [&](const std::string &n) { return nl->index.at(n); };
    for (const Elt &e : nl->elts) {
        int rc = NGMI_OK;
        if (e.kind == 'm') {
            std::vector<const char *> pn;
            std::vector<double> pv;
            for (auto &kv : e.params) {
                pn.push_back(kv.first.c_str());
                pv.push_back(kv.second);
            }
            rc = ngmi_circuit_add_bsim4(c, nl->models.at(e.model), num(e.nodes[0]), num(e.nodes[1]),
                                        num(e.nodes[2]), num(e.nodes[3]), (int)pn.size(), pn.data(),
                                        pv.data());
        } else if (e.kind == 'c') {
            rc = ngmi_circuit_add_cap(c, num(e.nodes[0]), num(e.nodes[1]), e.value);
        } else if (e.kind == 'r') {
            rc = ngmi_circuit_add_res(c, num(e.nodes[0]), num(e.nodes[1]), e.value);
        }
        if (rc != NGMI_OK) return bail(e.name);
    }
    /* sources in branch order: reverse deck order */
    for (size_t k = nl->elts.size(); k-- > 0;) {
        const Elt &e = nl->elts[k];
        if (e.kind != 'v') continue;
        double pulse[7];
        for (size_t q = 0; q < e.pulse.size(); q++) pulse[q] = e.pulse[q];
        if (ngmi_circuit_add_vsrc(c, num(e.nodes[0]), num(e.nodes[1]), e.value,
                                  e.pulse.empty() ? nullptr : pulse, (int)e.pulse.size()) != NGMI_OK)
            return bail(e.name);
    }
    for (auto &x : nl->ic)
        if (ngmi_circuit_set_ic(c, x.first, x.second) != NGMI_OK) return bail(".ic");
    return c;
}

}  // extern "C"
