// jst_host.cpp — host driver of the gfx950 JST + SA solver and its C ABI (include/jst_gpu.h).
//
// Mirrors the jst-rs crate surface: Config (config.rs), Mesh (mesh.rs), Solver (solver.rs).
// Everything numeric that the host computes (derived constants, unpack for field export and
// for the NonPhysical record) uses the reference's operation order; this file is compiled
// with -ffp-contract=off.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <charconv>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/jst_gpu.h"
#include "jst_dev.h"
#include "jst_launch.h"

#include "jst_common.h"

namespace jsth {

thread_local std::string g_err = "ok";

jst_status fail(jst_status st, const std::string &msg)
{
    g_err = msg;
    return st;
}

// ------------------------------------------------------------------------------------
// Minimal JSON reader (the subset serde_json accepts for config.json).
// ------------------------------------------------------------------------------------
struct JVal {
    enum Kind { NUL, BOOL, NUM, STR, ARR, OBJ } kind = NUL;
    bool b = false;
    double num = 0.0;
    bool integral = false; // literal without '.', exponent or sign
    std::string str;
    std::vector<JVal> arr;
    std::vector<std::pair<std::string, JVal>> obj;
    const JVal *get(const char *key) const
    {
        if (kind != OBJ) return nullptr;
        for (auto &kv : obj)
            if (kv.first == key) return &kv.second;
        return nullptr;
    }
};

struct JParser {
    const char *p, *e;
    std::string err;
    void ws()
    {
        while (p < e && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')) ++p;
    }
    bool value(JVal &v)
    {
        ws();
        if (p >= e) return set_err("EOF while parsing a value");
        char c = *p;
        if (c == '{') return object(v);
        if (c == '[') return array(v);
        if (c == '"') {
            v.kind = JVal::STR;
            return string(v.str);
        }
        if (!strncmp(p, "true", 4)) { p += 4; v.kind = JVal::BOOL; v.b = true; return true; }
        if (!strncmp(p, "false", 5)) { p += 5; v.kind = JVal::BOOL; v.b = false; return true; }
        if (!strncmp(p, "null", 4)) { p += 4; v.kind = JVal::NUL; return true; }
        return number(v);
    }
    bool set_err(const char *m)
    {
        if (err.empty()) err = m;
        return false;
    }
    bool string(std::string &out)
    {
        ++p;
        while (p < e && *p != '"') {
            if (*p == '\\') {
                ++p;
                if (p >= e) return set_err("EOF while parsing a string");
                switch (*p) {
                case 'n': out += '\n'; break;
                case 't': out += '\t'; break;
                case 'r': out += '\r'; break;
                case 'b': out += '\b'; break;
                case 'f': out += '\f'; break;
                case 'u': p += 4; out += '?'; break;
                default: out += *p;
                }
                ++p;
            } else {
                out += *p++;
            }
        }
        if (p >= e) return set_err("EOF while parsing a string");
        ++p;
        return true;
    }
    bool number(JVal &v)
    {
        const char *s = p;
        bool integral = true;
        if (p < e && *p == '-') { ++p; integral = false; }
        if (p >= e || !isdigit((unsigned char)*p)) return set_err("expected value");
        while (p < e && isdigit((unsigned char)*p)) ++p;
        if (p < e && *p == '.') {
            integral = false;
            ++p;
            if (p >= e || !isdigit((unsigned char)*p)) return set_err("invalid number");
            while (p < e && isdigit((unsigned char)*p)) ++p;
        }
        if (p < e && (*p == 'e' || *p == 'E')) {
            integral = false;
            ++p;
            if (p < e && (*p == '+' || *p == '-')) ++p;
            if (p >= e || !isdigit((unsigned char)*p)) return set_err("invalid number");
            while (p < e && isdigit((unsigned char)*p)) ++p;
        }
        std::string tok(s, p);
        v.kind = JVal::NUM;
        v.num = strtod(tok.c_str(), nullptr);
        v.integral = integral;
        return true;
    }
    bool array(JVal &v)
    {
        v.kind = JVal::ARR;
        ++p;
        ws();
        if (p < e && *p == ']') { ++p; return true; }
        for (;;) {
            JVal x;
            if (!value(x)) return false;
            v.arr.push_back(std::move(x));
            ws();
            if (p < e && *p == ',') { ++p; continue; }
            if (p < e && *p == ']') { ++p; return true; }
            return set_err("expected `,` or `]`");
        }
    }
    bool object(JVal &v)
    {
        v.kind = JVal::OBJ;
        ++p;
        ws();
        if (p < e && *p == '}') { ++p; return true; }
        for (;;) {
            ws();
            if (p >= e || *p != '"') return set_err("key must be a string");
            std::string key;
            if (!string(key)) return false;
            ws();
            if (p >= e || *p != ':') return set_err("expected `:`");
            ++p;
            JVal x;
            if (!value(x)) return false;
            v.obj.emplace_back(std::move(key), std::move(x));
            ws();
            if (p < e && *p == ',') { ++p; continue; }
            if (p < e && *p == '}') { ++p; return true; }
            return set_err("expected `,` or `}`");
        }
    }
};

// ------------------------------------------------------------------------------------
// Config (jst-rs config.rs)
// ------------------------------------------------------------------------------------


double powi_c(double x, int n)
{
    double res = 0.0, sq = x;
    bool have = false;
    while (n) {
        if (n & 1) {
            if (have) res = res * sq;
            else { res = sq; have = true; }
        }
        sq = sq * sq;
        n >>= 1;
    }
    return have ? res : 1.0;
}

// config.rs:153-156
double sutherland(double mu0, double t, double t_ref, double t_suth)
{
    double r = t / t_ref;
    return mu0 * (r * std::sqrt(r)) * (t_ref + t_suth) / (t + t_suth);
}

// config.rs:173-206 (Config::finish), plus the Reynolds / nut_ratio extensions.
Derived finish(const Config &c)
{
    Derived d{};
    double cv = c.r_gas / (c.gamma - 1.0);
    double c_inf = std::sqrt(c.gamma * c.r_gas * c.t_inf);
    double aoa = c.aoa * (M_PI / 180.0);
    double rho_inf = c.p_inf / (c.r_gas * c.t_inf);
    double mu0 = c.mu0;
    double u_inf = c_inf * c.mach * std::cos(aoa);
    double v_inf = c_inf * c.mach * std::sin(aoa);
    if (c.reynolds > 0.0) {
        double mu_ref = sutherland(mu0, c.t_inf, c.t_ref, c.t_suth);
        double re_now = rho_inf * std::hypot(u_inf, v_inf) * c.ref_length / mu_ref;
        mu0 = mu0 * (re_now / c.reynolds);
    }
    double mu_inf = sutherland(mu0, c.t_inf, c.t_ref, c.t_suth);
    d.mu0_eff = mu0;
    d.cv = cv;
    d.cp = c.gamma * cv;
    d.cw1 = c.cb1 / (c.kappa * c.kappa) + (1.0 + c.cb2) * c.sigma;
    d.cv1_cubed = powi_c(c.cv1, 3);
    d.c_inf = c_inf;
    d.u_inf = u_inf;
    d.v_inf = v_inf;
    d.rho_inf = rho_inf;
    d.mu_inf = mu_inf;
    d.nut_inf = c.nut_ratio * mu_inf / rho_inf;
    d.supersonic = std::hypot(u_inf, v_inf) >= c_inf;
    d.suth_num = mu0 * (c.t_ref + c.t_suth);
    d.inv_t_ref = 1.0 / c.t_ref;
    d.inv_pr = 1.0 / c.pr;
    d.inv_prt = 1.0 / c.prt;
    d.inv_kappa2 = 1.0 / (c.kappa * c.kappa);
    d.cw3_6 = powi_c(c.cw3, 6);
    d.one_plus_cw3_6 = 1.0 + powi_c(c.cw3, 6);
    return d;
}

// config.rs:210-213 (Config::mu)
double config_mu(const Config &c, const Derived &d, double t)
{
    double r = t * d.inv_t_ref;
    return r * std::sqrt(r) * d.suth_num / (t + c.t_suth);
}

jst_status parse_config(const char *text, Config &c)
{
    if (!text) return fail(JST_ERR_ARG, "config text is NULL");
    JParser jp{text, text + strlen(text), {}};
    JVal root;
    if (!jp.value(root)) return fail(JST_ERR_CONFIG, jp.err);
    jp.ws();
    if (jp.p != jp.e) return fail(JST_ERR_CONFIG, "trailing characters");
    if (root.kind != JVal::OBJ) return fail(JST_ERR_CONFIG, "invalid type: expected struct Config");
    std::string err;
    auto sec = [&](const char *name) -> const JVal * {
        const JVal *s = root.get(name);
        if (!s && err.empty()) err = std::string("missing field `") + name + "`";
        else if (s && s->kind != JVal::OBJ && err.empty())
            err = std::string("invalid type for `") + name + "`: expected a map";
        return (s && s->kind == JVal::OBJ) ? s : nullptr;
    };
    auto num = [&](const JVal *s, const char *key, double &out) {
        if (!s) return;
        const JVal *v = s->get(key);
        if (!v) {
            if (err.empty()) err = std::string("missing field `") + key + "`";
            return;
        }
        if (v->kind != JVal::NUM) {
            if (err.empty()) err = std::string("invalid type for `") + key + "`: expected f64";
            return;
        }
        out = v->num;
    };
    auto uint = [&](const JVal *s, const char *key, long &out, bool required) {
        if (!s) return;
        const JVal *v = s->get(key);
        if (!v) {
            if (required && err.empty()) err = std::string("missing field `") + key + "`";
            return;
        }
        if (v->kind != JVal::NUM || !v->integral) {
            if (err.empty()) err = std::string("invalid type for `") + key + "`: expected usize";
            return;
        }
        out = (long)v->num;
    };
    auto optnum = [&](const JVal *s, const char *key, double &out) {
        if (!s) return;
        const JVal *v = s->get(key);
        if (v && v->kind == JVal::NUM) out = v->num;
        else if (v && err.empty()) err = std::string("invalid type for `") + key + "`: expected f64";
    };
    const JVal *ph = sec("physics");
    num(ph, "gamma", c.gamma);
    num(ph, "R", c.r_gas);
    num(ph, "T0", c.t_ref);
    num(ph, "Ts", c.t_suth);
    num(ph, "mu0", c.mu0);
    const JVal *sim = sec("simulation");
    num(sim, "AOA", c.aoa);
    num(sim, "Ma", c.mach);
    num(sim, "CFL", c.cfl);
    long halo = 0;
    uint(sim, "IM", halo, true);
    c.halo = (int)halo;
    num(sim, "T", c.t_inf);
    num(sim, "P", c.p_inf);
    optnum(sim, "Re", c.reynolds);
    optnum(sim, "ref_length", c.ref_length);
    if (sim) {
        const JVal *v = sim->get("local_time_stepping");
        if (v && v->kind == JVal::BOOL) c.lts = v->b;
        else if (v && err.empty()) err = "invalid type for `local_time_stepping`: expected bool";
        const JVal *w = sim->get("viscous_time_step");
        if (w && w->kind == JVal::BOOL) c.visc_dt = w->b;
        else if (w && err.empty()) err = "invalid type for `viscous_time_step`: expected bool";
    }
    const JVal *sol = root.get("solver");
    if (sol && sol->kind == JVal::OBJ) {
        uint(sol, "iteration", c.iteration, false);
        optnum(sol, "targetres", c.targetres);
        optnum(sol, "residual_drop", c.residual_drop);
    } else if (sol && sol->kind != JVal::NUL && err.empty()) {
        err = "invalid type for `solver`: expected a map";
    }
    const JVal *sa = sec("spalart_allmaras");
    num(sa, "Cv1", c.cv1);
    num(sa, "Pr", c.pr);
    num(sa, "Prt", c.prt);
    num(sa, "sigma", c.sigma);
    num(sa, "Cb1", c.cb1);
    num(sa, "Cb2", c.cb2);
    num(sa, "Ct3", c.ct3);
    num(sa, "Ct4", c.ct4);
    num(sa, "Cw2", c.cw2);
    num(sa, "Cw3", c.cw3);
    num(sa, "fv3", c.fv3);
    num(sa, "kappa", c.kappa);
    num(sa, "rmax", c.rmax);
    optnum(sa, "nut_ratio", c.nut_ratio);
    if (sa) {
        const JVal *v = sa->get("negative");
        if (v && v->kind == JVal::BOOL) c.sa_neg = v->b;
        else if (v && err.empty()) err = "invalid type for `negative`: expected bool";
    }
    const JVal *dis = sec("dissipation");
    num(dis, "k2", c.k2);
    num(dis, "k4", c.k4);
    if (!err.empty()) return fail(JST_ERR_CONFIG, err);
    return JST_OK;
}

// ------------------------------------------------------------------------------------
// Mesh (jst-rs mesh.rs:109-189)
// ------------------------------------------------------------------------------------

bool parse_usize(const std::string &s, long &out)
{
    size_t k = 0;
    long v = 0;
    if (!s.empty() && s[0] == '+') k = 1;
    if (k >= s.size()) return false;
    for (; k < s.size(); ++k) {
        if (s[k] < '0' || s[k] > '9') return false;
        if (v > 1000000000000000L) return false;
        v = v * 10 + (s[k] - '0');
    }
    out = v;
    return true;
}

bool parse_f64(const std::string &s, double &out)
{
    if (s.empty()) return false;
    for (char ch : s)
        if (ch == 'x' || ch == 'X' || ch == '(' || ch == ')' || ch == 'p' || ch == 'P') return false;
    char *end = nullptr;
    out = strtod(s.c_str(), &end);
    return end == s.c_str() + s.size();
}

std::vector<std::string> split_ws(const std::string &line)
{
    std::vector<std::string> out;
    std::istringstream is(line);
    std::string t;
    while (is >> t) out.push_back(t);
    return out;
}

// Whitespace of the mesh text format (the ASCII set of Rust's str::trim / split_whitespace).
inline bool mesh_ws(char c)
{
    return c == ' ' || c == '\t' || c == '\r' || c == '\n' || c == '\v' || c == '\f';
}

// parse_f64 on the token [a, b) without copying it: std::from_chars (correctly rounded, as
// strtod) is the fast path; what it declines — a leading '+', values out of the double range
// (Rust parses those to +-inf or 0) — goes through parse_f64, so every token gets exactly
// the value and verdict parse_f64 gives it.
bool parse_f64_view(const char *a, const char *b, double &out)
{
    if (a == b) return false;
    for (const char *q = a; q < b; ++q)
        if (*q == 'x' || *q == 'X' || *q == '(' || *q == ')' || *q == 'p' || *q == 'P') return false;
    const auto r = std::from_chars(a, b, out);
    if (r.ec == std::errc() && r.ptr == b) return true;
    return parse_f64(std::string(a, b), out);
}

jst_status parse_mesh(const char *text, Mesh &m)
{
    if (!text) return fail(JST_ERR_ARG, "mesh text is NULL");
    const char *const end = text + strlen(text);
    if (text == end) return fail(JST_ERR_EMPTY, "mesh file contains no points");
    // lines split at '\n' and trimmed in place (a mesh of 513 x 257 points is 5 MB of text)
    long ni = 0, nj = 0;
    std::vector<double> xy;
    size_t k = 0; // line index, 0 = header
    for (const char *s = text;;) {
        const char *nl = (const char *)memchr(s, '\n', (size_t)(end - s));
        const char *a = s, *b = nl ? nl : end;
        while (a < b && mesh_ws(*a)) ++a;
        while (b > a && mesh_ws(b[-1])) --b;
        if (k == 0) {
            const std::string header(a, b);
            auto ht = split_ws(header);
            if (ht.size() < 2 || !parse_usize(ht[0], ni) || !parse_usize(ht[1], nj)) {
                char buf[512];
                snprintf(buf, sizeof buf, "bad header \"%s\": expected `i_total j_total`", header.c_str());
                return fail(JST_ERR_BAD_HEADER, buf);
            }
            xy.reserve((size_t)std::max(0L, ni * nj) * 2);
        } else if (a < b) {
            const char *e0 = a;
            while (e0 < b && !mesh_ws(*e0)) ++e0;
            const char *a1 = e0;
            while (a1 < b && mesh_ws(*a1)) ++a1;
            const char *e1 = a1;
            while (e1 < b && !mesh_ws(*e1)) ++e1;
            double x, y;
            if (a1 == b || !parse_f64_view(a, e0, x) || !parse_f64_view(a1, e1, y)) {
                char buf[512];
                snprintf(buf, sizeof buf, "line %zu: cannot parse point \"%s\"", k + 1,
                         std::string(a, b).c_str());
                return fail(JST_ERR_BAD_POINT, buf);
            }
            xy.push_back(x);
            xy.push_back(y);
        }
        ++k;
        if (!nl) break;
        s = nl + 1;
        if (s >= end) break;
    }
    long got = (long)(xy.size() / 2);
    if (got == 0) return fail(JST_ERR_EMPTY, "mesh file contains no points");
    if (got != ni * nj) {
        char buf[256];
        snprintf(buf, sizeof buf, "expected %ld points from header, found %ld", ni * nj, got);
        return fail(JST_ERR_COUNT_MISMATCH, buf);
    }
    if (nj >= 2) {
        bool closed = true;
        for (long r = 0; r < ni && closed; ++r) {
            double ax = xy[2 * (r * nj)], ay = xy[2 * (r * nj) + 1];
            double bx = xy[2 * (r * nj + nj - 1)], by = xy[2 * (r * nj + nj - 1) + 1];
            closed = std::fabs(ax - bx) <= 1e-10 && std::fabs(ay - by) <= 1e-10;
        }
        if (closed) {
            long nn = nj - 1;
            std::vector<double> t;
            t.reserve((size_t)(ni * nn * 2));
            for (long r = 0; r < ni; ++r)
                for (long q = 0; q < nn; ++q) {
                    t.push_back(xy[2 * (r * nj + q)]);
                    t.push_back(xy[2 * (r * nj + q) + 1]);
                }
            xy.swap(t);
            nj = nn;
        }
    }
    if (ni < 4 || nj < 8) {
        char buf[256];
        snprintf(buf, sizeof buf,
                 "mesh %ldx%ld too small: the JST 4th-order stencil needs at least 4 rings and 8 "
                 "circumferential points",
                 ni, nj);
        return fail(JST_ERR_TOO_SMALL, buf);
    }
    m.n_rings = (int)ni;
    m.n_theta = (int)nj;
    m.xy.swap(xy);
    return JST_OK;
}

bool read_file(const char *path, std::string &out)
{
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    std::ostringstream ss;
    ss << f.rdbuf();
    out = ss.str();
    return true;
}

// Rust `{:.Ne}` (LowerExp with precision): mantissa as printf, exponent without '+' or
// leading zeros ("1.00000000e0", "-3.50000000e-5").
// prec < 0 gives Rust's `{:e}` (shortest digits that round-trip).
std::string rust_exp(double v, int prec)
{
    if (std::isnan(v)) return "NaN";
    if (std::isinf(v)) return v > 0 ? "inf" : "-inf";
    char buf[64];
    if (prec < 0) {
        for (prec = 0; prec < 17; ++prec) {
            snprintf(buf, sizeof buf, "%.*e", prec, v);
            if (strtod(buf, nullptr) == v) break;
        }
    }
    snprintf(buf, sizeof buf, "%.*e", prec, v);
    std::string s(buf);
    size_t e = s.find('e');
    std::string mant = s.substr(0, e);
    int ex = atoi(s.c_str() + e + 1);
    return mant + "e" + std::to_string(ex);
}

} // namespace jsth

using namespace jsth;

// ------------------------------------------------------------------------------------
// Context
// ------------------------------------------------------------------------------------
#define HIPCHK(x)                                                                            \
    do {                                                                                     \
        hipError_t e_ = (x);                                                                 \
        if (e_ != hipSuccess)                                                                \
            return fail(JST_ERR_HIP, std::string("HIP error: ") + hipGetErrorString(e_) +    \
                                         " at " #x);                                         \
    } while (0)

struct jst_ctx {
    Config cfg{};
    Derived d{};
    Mesh mesh;
    jst_precision prec = JST_FP64;
    int ni = 0, nj = 0, pitch = 0;
    size_t plane = 0, fplane = 0; // ni*pitch, (ni+1)*pitch
    hipStream_t stream = nullptr;
    std::vector<void *> allocs;
    jst::GeomPtrs g{};
    double *ubuf[3] = {nullptr, nullptr, nullptr};
    double *ratio = nullptr, *localdt = nullptr, *partial = nullptr, *res_hist = nullptr;
    double *dbg = nullptr, *scratch = nullptr;
    CaseCtl *ctl = nullptr;
    CaseDev *d_case = nullptr;
    TileDesc *d_tiles = nullptr;
    CaseDev hcase{};
    int ntiles = 0;
    int res_cap = 4096;
    CaseCtl *h_ctl = nullptr; // pinned mirror
    hipEvent_t ev[2] = {nullptr, nullptr};
    jst_batch *mix = nullptr; // FP32-mixed: a one-case batch
    int mix_id = -1;

    ~jst_ctx()
    {
        if (mix) jst_batch_destroy(mix);
        if (stream) (void)hipStreamSynchronize(stream);
        for (void *p : allocs) (void)hipFree(p);
        if (h_ctl) (void)hipHostFree(h_ctl);
        for (auto e : ev)
            if (e) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
    }

    template <class T> jst_status dalloc(T *&p, size_t n)
    {
        void *q = nullptr;
        HIPCHK(hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T)));
        HIPCHK(hipMemsetAsync(q, 0, std::max<size_t>(n, 1) * sizeof(T), stream));
        allocs.push_back(q);
        p = (T *)q;
        return JST_OK;
    }
};

namespace jsth {
PhysConst make_const(const Config &c, const Derived &d)
{
    PhysConst k{};
    k.gamma = c.gamma;
    k.r_gas = c.r_gas;
    k.cfl = c.cfl;
    k.k2 = c.k2;
    k.k4 = c.k4;
    k.t_inf = c.t_inf;
    k.p_inf = c.p_inf;
    k.rho_inf = d.rho_inf;
    k.u_inf = d.u_inf;
    k.v_inf = d.v_inf;
    k.c_inf = d.c_inf;
    k.nut_inf = d.nut_inf;
    k.suth_num = d.suth_num;
    k.inv_t_ref = d.inv_t_ref;
    k.t_suth = c.t_suth;
    k.inv_pr = d.inv_pr;
    k.inv_prt = d.inv_prt;
    k.cp = d.cp;
    k.sigma = c.sigma;
    k.cv1_cubed = d.cv1_cubed;
    k.cb1 = c.cb1;
    k.cb2 = c.cb2;
    k.ct3 = c.ct3;
    k.ct4 = c.ct4;
    k.cw1 = d.cw1;
    k.cw2 = c.cw2;
    k.cw3_6 = d.cw3_6;
    k.one_plus_cw3_6 = d.one_plus_cw3_6;
    k.fv3 = c.fv3;
    k.inv_kappa2 = d.inv_kappa2;
    k.rmax = c.rmax;
    k.targetres = c.targetres;
    k.residual_drop = c.residual_drop;
    k.supersonic = d.supersonic ? 1 : 0;
    k.lts = c.lts ? 1 : 0;
    k.visc_dt = c.visc_dt ? 1 : 0;
    k.sa_neg = c.sa_neg ? 1 : 0;
    return k;
}

jst_status wall_distribution(const double *d_terms, const jst::GeomPtrs &g, int nj,
                             const Config &c, const Derived &d, hipStream_t s, double *x,
                             double *y, double *cp, double *cf)
{
    const size_t n = (size_t)nj;
    std::vector<double> t(8 * n);
    HIPCHK(hipMemcpyAsync(t.data(), d_terms, 4 * n * sizeof(double), hipMemcpyDeviceToHost, s));
    const double *rows[4] = {g.tnx, g.tny, g.tmx, g.tmy}; // tau faces of row 0: first nj entries
    for (int q = 0; q < 4; ++q)
        HIPCHK(hipMemcpyAsync(t.data() + (4 + q) * n, rows[q], n * sizeof(double),
                              hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    const double q = 0.5 * d.rho_inf * (d.u_inf * d.u_inf + d.v_inf * d.v_inf);
    for (size_t j = 0; j < n; ++j) {
        const double fx = t[j], fy = t[n + j], vx = t[2 * n + j], vy = t[3 * n + j];
        const double nx = t[4 * n + j], ny = t[5 * n + j];
        const double n2 = nx * nx + ny * ny;
        const double pw = (fx * nx + fy * ny) / n2;
        const double tau = (vx * (-ny) + vy * nx) / n2;
        if (x) x[j] = t[6 * n + j];
        if (y) y[j] = t[7 * n + j];
        if (cp) cp[j] = (pw - c.p_inf) / q;
        if (cf) cf[j] = tau / q;
    }
    return JST_OK;
}
} // namespace jsth

namespace {

// Tile list in an XCD-aware order: workgroup b is dispatched to XCD b % 8, so consecutive
// row-major tiles (which share halo rows in L2) are dealt to the same XCD.
std::vector<TileDesc> make_tiles(int case_id, int ni, int nj, int ti, int tj)
{
    const int nti = (ni + ti - 1) / ti, ntj = (nj + tj - 1) / tj, n = nti * ntj;
    std::vector<TileDesc> nat(n);
    for (int a = 0; a < nti; ++a)
        for (int b = 0; b < ntj; ++b) nat[a * ntj + b] = TileDesc{case_id, a * ti, b * tj, a * ntj + b};
    const int per = (n + 7) / 8;
    std::vector<TileDesc> out;
    out.reserve(n);
    for (int q = 0; q < per; ++q)
        for (int x = 0; x < 8; ++x) {
            int t = x * per + q;
            if (t < n) out.push_back(nat[t]);
        }
    return out;
}

// jst-rs solver.rs:151-185 unpack of one cell on the host (for exports and error records).
struct HPrim {
    double rho, vx, vy, e, p, t, h, c, nut;
};
HPrim host_unpack(const Config &c, const double u[5])
{
    HPrim q;
    double rho = u[0];
    double inv = 1.0 / rho;
    q.rho = rho;
    q.vx = u[1] * inv;
    q.vy = u[2] * inv;
    q.e = u[3] * inv;
    q.p = (c.gamma - 1.0) * (u[3] - rho * (q.vx * q.vx + q.vy * q.vy) * 0.5);
    q.t = q.p / (c.r_gas * rho);
    q.nut = u[4] * inv;
    q.h = q.e + q.p * inv;
    q.c = std::sqrt(c.r_gas * c.gamma * q.t);
    return q;
}

jst_status ctl_fetch(jst_ctx *x)
{
    HIPCHK(hipMemcpyAsync(x->h_ctl, x->ctl, sizeof(CaseCtl), hipMemcpyDeviceToHost, x->stream));
    HIPCHK(hipStreamSynchronize(x->stream));
    return JST_OK;
}

template <class T> jst_status ctl_set(jst_ctx *x, T CaseCtl::*field, T value)
{
    // small host->device write of one control field (never inside a captured graph)
    static_assert(sizeof(T) <= 8, "scalar field");
    x->h_ctl->*field = value;
    size_t off = (size_t)((char *)&(x->h_ctl->*field) - (char *)x->h_ctl);
    HIPCHK(hipMemcpyAsync((char *)x->ctl + off, &(x->h_ctl->*field), sizeof(T),
                          hipMemcpyHostToDevice, x->stream));
    HIPCHK(hipStreamSynchronize(x->stream));
    return JST_OK;
}

static const double RK_ALPHA[5] = {0.25, 1.0 / 6.0, 0.375, 0.5, 1.0}; // config.rs:217

// Steps are launched directly (five stages, then the two step-end launches).  Replaying them
// from a HIP graph was measured as no faster (a small case's step is bound by the per-kernel
// turnaround of about 10 us on the device either way, and slightly slower in graph form).
jst_status launch_steps(jst_ctx *x, int n)
{
    for (int s = 0; s < n; ++s) {
        for (int k = 0; k < 5; ++k)
            HIPCHK(jst::launch_stage_fp64(x->d_case, x->d_tiles, x->ntiles, k, RK_ALPHA[k], false,
                                          x->stream));
        HIPCHK(jst::launch_post_fp64(x->d_case, x->d_tiles, x->ntiles, 1, x->stream));
    }
    return JST_OK;
}

jst_status refresh_dt(jst_ctx *x)
{
    HIPCHK(jst::launch_post_fp64(x->d_case, x->d_tiles, x->ntiles, 0, x->stream));
    return JST_OK;
}

jst_status download(jst_ctx *x, const double *dsrc, size_t n, std::vector<double> &out)
{
    out.resize(n);
    HIPCHK(hipMemcpyAsync(out.data(), dsrc, n * sizeof(double), hipMemcpyDeviceToHost, x->stream));
    HIPCHK(hipStreamSynchronize(x->stream));
    return JST_OK;
}

// Decode the first non-physical cell (stage sequence, row-major index) and recompute
// {rho, p} from the state the failing stage wrote, as Solver::unpack reports them.
jst_status report_nonphysical(jst_ctx *x, jst_nonphysical *err)
{
    unsigned long long key = x->h_ctl->err_key;
    long long seq = (long long)(key >> 32);
    long long cell = (long long)(key & 0xFFFFFFFFull);
    int stage = (int)(seq % 5);
    int rot = x->h_ctl->rot;
    int b1 = (rot + 1) % 3, b2 = (rot + 2) % 3;
    const double *ubad = x->ubuf[(stage % 2 == 0) ? b1 : b2];
    long i = (long)(cell / x->nj), j = (long)(cell % x->nj);
    double u[5];
    for (int m = 0; m < 5; ++m) {
        HIPCHK(hipMemcpyAsync(&u[m], ubad + m * x->plane + (size_t)i * x->pitch + j, sizeof(double),
                              hipMemcpyDeviceToHost, x->stream));
    }
    HIPCHK(hipStreamSynchronize(x->stream));
    double rho = u[0], p = NAN;
    if (!(std::isnan(rho) || rho <= 1e-15)) {
        double inv = 1.0 / rho;
        double vx = u[1] * inv, vy = u[2] * inv;
        p = (x->cfg.gamma - 1.0) * (u[3] - rho * (vx * vx + vy * vy) * 0.5);
    }
    if (err) *err = jst_nonphysical{i, j, rho, p};
    char buf[256];
    snprintf(buf, sizeof buf, "non-physical state at cell (%ld, %ld): rho = %s, p = %s", i, j,
             rust_exp(rho, -1).c_str(), rust_exp(p, -1).c_str());
    return fail(JST_ERR_NONPHYSICAL, buf);
}

} // namespace

// ------------------------------------------------------------------------------------
// C ABI
// ------------------------------------------------------------------------------------
namespace jsth {

jst_status check_injected_state(const double *u_planes, int ni, int nj, double gamma,
                                jst_nonphysical *err)
{
    const size_t ncell = (size_t)ni * nj;
    for (int i = 0; i < ni; ++i)
        for (int j = 0; j < nj; ++j) {
            double u[5];
            for (int m = 0; m < 5; ++m) u[m] = u_planes[m * ncell + (size_t)i * nj + j];
            const double rho = u[0];
            bool bad = std::isnan(rho) || rho <= 1e-15;
            double p = NAN;
            if (!bad) {
                const double inv = 1.0 / rho, vx = u[1] * inv, vy = u[2] * inv;
                p = (gamma - 1.0) * (u[3] - rho * (vx * vx + vy * vy) * 0.5);
                bad = std::isnan(p) || p <= 1e-15;
            }
            if (bad) {
                if (err) *err = jst_nonphysical{i, j, rho, p};
                char buf[256];
                snprintf(buf, sizeof buf, "non-physical state at cell (%d, %d): rho = %s, p = %s", i,
                         j, rust_exp(rho, -1).c_str(), rust_exp(p, -1).c_str());
                return fail(JST_ERR_NONPHYSICAL, buf);
            }
        }
    return JST_OK;
}

} // namespace jsth

extern "C" {

const char *jst_last_error(void) { return g_err.c_str(); }

jst_status jst_config_check(const char *config_json, double *out)
{
    Config c;
    jst_status st = parse_config(config_json, c);
    if (st != JST_OK) return st;
    Derived d = finish(c);
    if (out) {
        double v[20] = {d.cv, d.cp, d.cw1, d.cv1_cubed, d.c_inf, d.u_inf, d.v_inf, d.rho_inf,
                        d.mu_inf, d.nut_inf, d.supersonic ? 1.0 : 0.0, d.suth_num, d.inv_t_ref,
                        d.inv_pr, d.inv_prt, d.inv_kappa2, d.cw3_6, d.one_plus_cw3_6, d.mu0_eff,
                        (double)c.halo};
        memcpy(out, v, sizeof v);
    }
    return JST_OK;
}

jst_status jst_mesh_parse(const char *mesh_text, int *n_rings, int *n_theta, double *xy,
                          size_t xy_len)
{
    Mesh m;
    jst_status st = parse_mesh(mesh_text, m);
    if (st != JST_OK) return st;
    if (n_rings) *n_rings = m.n_rings;
    if (n_theta) *n_theta = m.n_theta;
    if (xy) {
        if (xy_len < m.xy.size()) return fail(JST_ERR_ARG, "xy buffer too small");
        memcpy(xy, m.xy.data(), m.xy.size() * sizeof(double));
    }
    return JST_OK;
}

jst_status jst_create(const char *config_json, const char *mesh_text, jst_precision prec,
                      jst_ctx **out)
{
    if (!out) return fail(JST_ERR_ARG, "out is NULL");
    *out = nullptr;
    std::unique_ptr<jst_ctx> x(new jst_ctx());
    jst_status st = parse_config(config_json, x->cfg);
    if (st != JST_OK) return st;
    st = parse_mesh(mesh_text, x->mesh);
    if (st != JST_OK) return st;
    // The stencils reach three cells past a boundary (the pressure sensors of a face's four
    // cells), so any IM >= 3 gives the same results, as in the crate, whose extra ghost layers
    // are filled but never read (config.rs:33-35).  IM < 3 would index past the crate's halo.
    if (x->cfg.halo < JST_H) {
        char buf[128];
        snprintf(buf, sizeof buf, "unsupported halo IM=%d: the stencils need IM >= 3", x->cfg.halo);
        return fail(JST_ERR_CONFIG, buf);
    }
    x->prec = prec;
    if (prec == JST_MIXED) {
        size_t bytes = 0;
        jst_status sm = jst_case_footprint(config_json, mesh_text, prec, &bytes);
        if (sm != JST_OK) return sm;
        if ((sm = jst_batch_create(JST_MIXED, bytes + (4u << 20), 1, &x->mix)) != JST_OK) return sm;
        // one case alone takes its own row-segment length, chosen from its size: the
        // shortest whose work items all fit in the 2048 resident waves.  Measured: 256x128
        // cells 67 / 77 / 91 / 120 us per step with 1 / 2 / 4 / 8 rows (1 is chosen);
        // 1024x512 cells 324 / 208 / 185 / 147 / 172 us with 1 / 2 / 4 / 5 / 8 rows (5).
        if ((sm = jst_batch_set_tiling(x->mix, MIX_VARIANT_PER_CASE, 0)) != JST_OK) return sm;
        if ((sm = jst_batch_add(x->mix, config_json, mesh_text, 0, &x->mix_id)) != JST_OK) return sm;
        x->d = finish(x->cfg);
        x->ni = x->mesh.n_rings - 1;
        x->nj = x->mesh.n_theta;
        *out = x.release();
        return JST_OK;
    }
    if (prec != JST_FP64) return fail(JST_ERR_ARG, "unknown precision");
    x->d = finish(x->cfg);
    x->ni = x->mesh.n_rings - 1;
    x->nj = x->mesh.n_theta;
    x->pitch = (x->nj + 15) / 16 * 16;
    x->plane = (size_t)x->ni * x->pitch;
    x->fplane = (size_t)(x->ni + 1) * x->pitch;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return fail(JST_ERR_HIP, "no HIP device available");
    HIPCHK(hipStreamCreateWithFlags(&x->stream, hipStreamNonBlocking));
    HIPCHK(hipEventCreate(&x->ev[0]));
    HIPCHK(hipEventCreate(&x->ev[1]));
    HIPCHK(hipHostMalloc((void **)&x->h_ctl, sizeof(CaseCtl), hipHostMallocDefault));
    // allocations
    jst_ctx &c = *x;
#define ALLOC(p, n)                                                                          \
    do {                                                                                     \
        jst_status s_ = c.dalloc(p, n);                                                      \
        if (s_ != JST_OK) return s_;                                                         \
    } while (0)
    for (int b = 0; b < 3; ++b) ALLOC(c.ubuf[b], 5 * c.plane);
    ALLOC(c.g.vol, c.plane);
    ALLOC(c.g.cx, c.plane);
    ALLOC(c.g.cy, c.plane);
    ALLOC(c.g.wd, c.plane);
    ALLOC(c.g.inv_vol, c.plane);
    ALLOC(c.g.inv_wd2, c.plane);
    ALLOC(c.g.tnx, c.fplane);
    ALLOC(c.g.tny, c.fplane);
    ALLOC(c.g.tmx, c.fplane);
    ALLOC(c.g.tmy, c.fplane);
    ALLOC(c.g.nnx, c.plane);
    ALLOC(c.g.nny, c.plane);
    ALLOC(c.g.nmx, c.plane);
    ALLOC(c.g.nmy, c.plane);
    ALLOC(c.ratio, 2 * c.plane);   // two planes: this step's (step & 1) and the next one's
    ALLOC(c.localdt, 2 * c.plane);
    ALLOC(c.res_hist, (size_t)c.res_cap);
    ALLOC(c.scratch, 4 * (size_t)c.nj + 2); // wall-force terms [4][nj] + {Fx, Fy}
    auto tiles = make_tiles(0, c.ni, c.nj, FP64_TI, FP64_TJ);
    c.ntiles = (int)tiles.size();
    ALLOC(c.partial, 2 * (size_t)c.ntiles); // per-tile residual sums | per-tile min dt
    ALLOC(c.ctl, 1);
    ALLOC(c.d_case, 1);
    ALLOC(c.d_tiles, tiles.size());
    double *d_xy = nullptr;
    ALLOC(d_xy, c.mesh.xy.size());
#undef ALLOC
    HIPCHK(hipMemcpyAsync(c.d_tiles, tiles.data(), tiles.size() * sizeof(TileDesc),
                          hipMemcpyHostToDevice, c.stream));
    HIPCHK(hipMemcpyAsync(d_xy, c.mesh.xy.data(), c.mesh.xy.size() * sizeof(double),
                          hipMemcpyHostToDevice, c.stream));
    HIPCHK(jst::launch_geometry(d_xy, c.mesh.n_theta, c.ni, c.nj, c.pitch, c.g, c.stream));

    // Cells::initialize (state.rs:174-190): freestream primitives, packed (state.rs:142-156)
    {
        const Config &k = c.cfg;
        double rho = k.p_inf / (k.r_gas * k.t_inf);
        double vx = c.d.u_inf, vy = c.d.v_inf, p = k.p_inf, nut = c.d.nut_inf;
        double e = p / (rho * (k.gamma - 1.0)) + 0.5 * (vx * vx + vy * vy);
        double uinit[5] = {rho, rho * vx, rho * vy, rho * e, rho * nut};
        std::vector<double> h(5 * c.plane, 0.0);
        for (int m = 0; m < 5; ++m)
            for (int i = 0; i < c.ni; ++i)
                for (int j = 0; j < c.nj; ++j) h[m * c.plane + (size_t)i * c.pitch + j] = uinit[m];
        HIPCHK(hipMemcpyAsync(c.ubuf[0], h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice,
                              c.stream));
        HIPCHK(hipStreamSynchronize(c.stream));
    }
    // device case descriptor
    CaseDev &hc = c.hcase;
    hc.ni = c.ni;
    hc.nj = c.nj;
    hc.pitch = c.pitch;
    hc.ntiles = c.ntiles;
    for (int b = 0; b < 3; ++b) hc.ubuf[b] = c.ubuf[b];
    hc.tnx = c.g.tnx;
    hc.tny = c.g.tny;
    hc.tmx = c.g.tmx;
    hc.tmy = c.g.tmy;
    hc.nnx = c.g.nnx;
    hc.nny = c.g.nny;
    hc.vol = c.g.vol;
    hc.inv_vol = c.g.inv_vol;
    hc.inv_wd2 = c.g.inv_wd2;
    hc.cx = c.g.cx;
    hc.cy = c.g.cy;
    hc.ratio = c.ratio;
    hc.localdt = c.localdt;
    hc.partial = c.partial;
    hc.res_hist = c.res_hist;
    hc.res_cap = c.res_cap;
    hc.ctl = c.ctl;
    hc.dbg = nullptr;
    hc.k = make_const(c.cfg, c.d);
    HIPCHK(hipMemcpyAsync(c.d_case, &hc, sizeof(CaseDev), hipMemcpyHostToDevice, c.stream));
    CaseCtl init{};
    init.dt = 0.0;
    init.dt_acc = JST_INF_BITS;
    init.totaltime = 0.0;
    init.step = 0;
    init.stop_step = 0;
    init.res_first = 0.0;
    init.residual = NAN;
    init.err_key = JST_ERR_NONE;
    init.status = JST_RUN_ACTIVE;
    init.rot = 0;
    init.blocks_done = 0;
    init.timing = 0;
    *c.h_ctl = init;
    HIPCHK(hipMemcpyAsync(c.ctl, c.h_ctl, sizeof(CaseCtl), hipMemcpyHostToDevice, c.stream));
    jst_status s2 = refresh_dt(&c);
    if (s2 != JST_OK) return s2;
    HIPCHK(hipStreamSynchronize(c.stream));
    s2 = ctl_fetch(&c);
    if (s2 != JST_OK) return s2;
    *out = x.release();
    return JST_OK;
}

jst_status jst_create_from_paths(const char *config_path, const char *mesh_path,
                                 jst_precision prec, jst_ctx **out)
{
    std::string ct, mt;
    if (!config_path || !read_file(config_path, ct))
        return fail(JST_ERR_IO, std::string("I/O error: cannot read ") + (config_path ? config_path : "(null)"));
    if (!mesh_path || !read_file(mesh_path, mt))
        return fail(JST_ERR_IO, std::string("I/O error: cannot read ") + (mesh_path ? mesh_path : "(null)"));
    return jst_create(ct.c_str(), mt.c_str(), prec, out);
}

void jst_destroy(jst_ctx *ctx) { delete ctx; }

jst_status jst_dims(const jst_ctx *ctx, int *ni, int *nj)
{
    if (!ctx) return fail(JST_ERR_ARG, "ctx is NULL");
    if (ni) *ni = ctx->ni;
    if (nj) *nj = ctx->nj;
    return JST_OK;
}

jst_status jst_progress(const jst_ctx *ctx, long *step, double *totaltime)
{
    if (!ctx) return fail(JST_ERR_ARG, "ctx is NULL");
    if (ctx->mix) {
        jst_case_info ci;
        jst_status st = jst_batch_info(ctx->mix, ctx->mix_id, &ci);
        if (st != JST_OK) return st;
        if (step) *step = ci.step;
        if (totaltime) *totaltime = ci.totaltime;
        return JST_OK;
    }
    if (step) *step = (long)ctx->h_ctl->step;
    if (totaltime) *totaltime = ctx->h_ctl->totaltime;
    return JST_OK;
}

jst_status jst_run(jst_ctx *x, long max_steps, jst_step_cb cb, void *user, jst_report *out,
                   jst_nonphysical *err)
{
    if (!x) return fail(JST_ERR_ARG, "ctx is NULL");
    const long limit = max_steps > 0 ? max_steps : x->cfg.iteration;
    if (x->mix) return batch_run_case(x->mix, x->mix_id, limit, cb, user, out, false);
    jst_status st = ctl_fetch(x);
    if (st != JST_OK) return st;
    if (x->h_ctl->err_key != JST_ERR_NONE) return report_nonphysical(x, err);
    const long long start = x->h_ctl->step;
    if ((st = ctl_set(x, &CaseCtl::stop_step, (long long)(start + limit))) != JST_OK) return st;
    if ((st = ctl_set(x, &CaseCtl::status, (int)JST_RUN_ACTIVE)) != JST_OK) return st;
    long done = 0;
    int batch = 8;
    std::vector<double> hist;
    while (done < limit) {
        int n = (int)std::min<long>(batch, limit - done);
        if ((st = launch_steps(x, n)) != JST_OK) return st;
        if ((st = ctl_fetch(x)) != JST_OK) return st;
        long nd = (long)(x->h_ctl->step - start) - done;
        if (nd > 0) {
            if ((st = download(x, x->res_hist, (size_t)x->res_cap, hist)) != JST_OK) return st;
            for (long q = 0; q < nd; ++q) {
                long long sidx = start + done + q; // 0-based absolute step index
                if (cb) cb(done + q + 1, hist[(size_t)(sidx % x->res_cap)], user);
            }
            done += nd;
        }
        if (x->h_ctl->err_key != JST_ERR_NONE) {
            if (out) *out = jst_report{done, x->h_ctl->residual, x->h_ctl->totaltime, 0};
            return report_nonphysical(x, err);
        }
        if (x->h_ctl->status != JST_RUN_ACTIVE) break;
        batch = std::min(batch * 2, std::min(512, x->res_cap));
    }
    if (out) {
        out->steps = done;
        out->residual = done > 0 ? x->h_ctl->residual : NAN;
        out->totaltime = x->h_ctl->totaltime;
        out->converged = x->h_ctl->status == JST_RUN_CONVERGED ? 1 : 0;
    }
    return JST_OK;
}

jst_status jst_advance(jst_ctx *x, int nsteps, double *res_out, jst_nonphysical *err)
{
    if (!x) return fail(JST_ERR_ARG, "ctx is NULL");
    if (nsteps <= 0) return JST_OK;
    struct Acc {
        double *out;
        long n;
    } acc{res_out, 0};
    jst_report rep{};
    auto acb = [](long, double r, void *u) {
        Acc *a = (Acc *)u;
        if (a->out) a->out[a->n] = r;
        a->n++;
    };
    if (x->mix) return batch_run_case(x->mix, x->mix_id, nsteps, acb, &acc, &rep, true);
    // advance never stops on convergence: the reference's advance() has no stop criterion
    jst_status st = ctl_set(x, &CaseCtl::timing, 1);
    if (st != JST_OK) return st;
    st = jst_run(x, nsteps, acb, &acc, &rep, err);
    jst_status st2 = ctl_set(x, &CaseCtl::timing, 0);
    return st != JST_OK ? st : st2;
}

jst_status jst_residual(const jst_ctx *x, double *out)
{
    if (!x || !out) return fail(JST_ERR_ARG, "NULL argument");
    if (x->mix) {
        jst_case_info ci;
        jst_status st = jst_batch_info(x->mix, x->mix_id, &ci);
        if (st != JST_OK) return st;
        *out = ci.step > 0 ? ci.residual : NAN;
        return JST_OK;
    }
    *out = x->h_ctl->step > 0 ? x->h_ctl->residual : NAN;
    return JST_OK;
}

jst_status jst_residual_terms(jst_ctx *x, double *dt_out)
{
    if (!x) return fail(JST_ERR_ARG, "ctx is NULL");
    if (x->mix) return jst_batch_residual_terms(x->mix, x->mix_id, dt_out);
    if (!x->dbg) {
        jst_status st = x->dalloc(x->dbg, (size_t)DBG_NPLANES * x->plane);
        if (st != JST_OK) return st;
        x->hcase.dbg = x->dbg;
        HIPCHK(hipMemcpyAsync(x->d_case, &x->hcase, sizeof(CaseDev), hipMemcpyHostToDevice, x->stream));
    }
    jst_status st = refresh_dt(x);
    if (st != JST_OK) return st;
    HIPCHK(jst::launch_stage_fp64(x->d_case, x->d_tiles, x->ntiles, 0, RK_ALPHA[0], true, x->stream));
    st = ctl_fetch(x);
    if (st != JST_OK) return st;
    if (dt_out) *dt_out = x->h_ctl->dt;
    return JST_OK;
}

jst_status jst_get_field(jst_ctx *x, const char *name, double *dst, size_t n)
{
    if (!x || !name || !dst) return fail(JST_ERR_ARG, "NULL argument");
    if (x->mix) return jst_batch_get_field(x->mix, x->mix_id, name, dst, n);
    const int ni = x->ni, nj = x->nj, pitch = x->pitch;
    const size_t ncell = (size_t)ni * nj;
    std::string s(name);
    auto gather = [&](const std::vector<double> &h, size_t base, int rows) {
        for (int i = 0; i < rows; ++i)
            for (int j = 0; j < nj; ++j) dst[(size_t)i * nj + j] = h[base + (size_t)i * pitch + j];
    };
    jst_status st;
    std::vector<double> h;
    struct G {
        const char *n;
        const double *p;
        bool face;
    } geo[] = {{"vol", x->g.vol, false},         {"cx", x->g.cx, false},
               {"cy", x->g.cy, false},           {"wall_dist", x->g.wd, false},
               {"inv_vol", x->g.inv_vol, false}, {"inv_wd2", x->g.inv_wd2, false},
               {"tau_nx", x->g.tnx, true},       {"tau_ny", x->g.tny, true},
               {"tau_mx", x->g.tmx, true},       {"tau_my", x->g.tmy, true},
               {"n_nx", x->g.nnx, false},        {"n_ny", x->g.nny, false},
               {"n_mx", x->g.nmx, false},        {"n_my", x->g.nmy, false}};
    if (s == "localdt" || s == "ratio") {
        // the frozen pair of the current step: plane (step & 1)
        if (n < ncell) return fail(JST_ERR_ARG, "destination too small");
        if ((st = ctl_fetch(x)) != JST_OK) return st;
        const double *base = (s == "ratio" ? x->ratio : x->localdt) + (size_t)(x->h_ctl->step & 1) * x->plane;
        if ((st = download(x, base, x->plane, h)) != JST_OK) return st;
        gather(h, 0, ni);
        return JST_OK;
    }
    for (auto &gd : geo) {
        if (s == gd.n) {
            int rows = gd.face ? ni + 1 : ni;
            if (n < (size_t)rows * nj) return fail(JST_ERR_ARG, "destination too small");
            if ((st = download(x, gd.p, (size_t)rows * pitch, h)) != JST_OK) return st;
            gather(h, 0, rows);
            return JST_OK;
        }
    }
    if (n < ncell) return fail(JST_ERR_ARG, "destination too small");
    auto term_plane = [&](int p) -> jst_status {
        if (!x->dbg) return fail(JST_ERR_ARG, "residual terms not computed: call jst_residual_terms");
        jst_status q = download(x, x->dbg + (size_t)p * x->plane, x->plane, h);
        if (q != JST_OK) return q;
        gather(h, 0, ni);
        return JST_OK;
    };
    if (s.size() == 3 && (s.rfind("Fc", 0) == 0 || s.rfind("Fv", 0) == 0 || s.rfind("Fd", 0) == 0) &&
        s[2] >= '0' && s[2] <= '4') {
        int base = s[1] == 'c' ? DBG_FC : (s[1] == 'v' ? DBG_FV : DBG_FD);
        return term_plane(base + (s[2] - '0'));
    }
    if (s.size() == 5 && s.rfind("grad", 0) == 0 && s[4] >= '0' && s[4] <= '7')
        return term_plane(DBG_GRAD + (s[4] - '0'));
    if (s == "S") return term_plane(DBG_S);
    if (s == "chi") return term_plane(DBG_CHI);
    if (s == "fv1") return term_plane(DBG_FV1);
    // state: current conserved planes
    if ((st = ctl_fetch(x)) != JST_OK) return st;
    const double *ucur = x->ubuf[x->h_ctl->rot];
    if ((st = download(x, ucur, 5 * x->plane, h)) != JST_OK) return st;
    if (s.size() == 2 && s[0] == 'U' && s[1] >= '0' && s[1] <= '4') {
        gather(h, (size_t)(s[1] - '0') * x->plane, ni);
        return JST_OK;
    }
    static const char *prims[] = {"rho", "p", "T", "u", "v", "E", "H", "c", "nut"};
    int which = -1;
    for (int q = 0; q < 9; ++q)
        if (s == prims[q]) which = q;
    if (which < 0) return fail(JST_ERR_ARG, "unknown field " + s);
    for (int i = 0; i < ni; ++i)
        for (int j = 0; j < nj; ++j) {
            double u[5];
            for (int m = 0; m < 5; ++m) u[m] = h[m * x->plane + (size_t)i * pitch + j];
            HPrim q = host_unpack(x->cfg, u);
            double v[9] = {q.rho, q.p, q.t, q.vx, q.vy, q.e, q.h, q.c, q.nut};
            dst[(size_t)i * nj + j] = v[which];
        }
    return JST_OK;
}

jst_status jst_set_state(jst_ctx *x, const double *u_planes, jst_nonphysical *err)
{
    if (!x || !u_planes) return fail(JST_ERR_ARG, "NULL argument");
    if (x->mix) return jst_batch_set_state(x->mix, x->mix_id, u_planes, err);
    // the unpack check first: a rejected state leaves the context as it was
    jst_status st = check_injected_state(u_planes, x->ni, x->nj, x->cfg.gamma, err);
    if (st != JST_OK) return st;
    if ((st = ctl_fetch(x)) != JST_OK) return st;
    const int ni = x->ni, nj = x->nj, pitch = x->pitch;
    const size_t ncell = (size_t)ni * nj;
    std::vector<double> h(5 * x->plane, 0.0);
    for (int m = 0; m < 5; ++m)
        for (int i = 0; i < ni; ++i)
            for (int j = 0; j < nj; ++j)
                h[m * x->plane + (size_t)i * pitch + j] = u_planes[m * ncell + (size_t)i * nj + j];
    HIPCHK(hipMemcpyAsync(x->ubuf[x->h_ctl->rot], h.data(), h.size() * sizeof(double),
                          hipMemcpyHostToDevice, x->stream));
    if ((st = ctl_set(x, &CaseCtl::err_key, (unsigned long long)JST_ERR_NONE)) != JST_OK) return st;
    if ((st = ctl_set(x, &CaseCtl::status, (int)JST_RUN_ACTIVE)) != JST_OK) return st;
    if ((st = refresh_dt(x)) != JST_OK) return st;
    return ctl_fetch(x);
}

jst_status jst_forces(jst_ctx *x, double *fx, double *fy, double *cl, double *cd)
{
    if (!x) return fail(JST_ERR_ARG, "ctx is NULL");
    if (x->mix) {
        double f4[4];
        jst_status st = batch_forces(x->mix, x->mix_id, f4);
        if (st != JST_OK) return st;
        if (fx) *fx = f4[0];
        if (fy) *fy = f4[1];
        if (cl) *cl = f4[2];
        if (cd) *cd = f4[3];
        return JST_OK;
    }
    double *dout = x->scratch + 4 * (size_t)x->nj;
    HIPCHK(jst::launch_forces_fp64(x->d_case, 0, x->scratch, dout, x->stream));
    double f[2];
    HIPCHK(hipMemcpyAsync(f, dout, sizeof f, hipMemcpyDeviceToHost, x->stream));
    HIPCHK(hipStreamSynchronize(x->stream));
    const double aoa = x->cfg.aoa * (M_PI / 180.0);
    const double ca = std::cos(aoa), sa = std::sin(aoa);
    const double q = 0.5 * x->d.rho_inf * (x->d.u_inf * x->d.u_inf + x->d.v_inf * x->d.v_inf);
    const double ql = q * x->cfg.ref_length;
    if (fx) *fx = f[0];
    if (fy) *fy = f[1];
    if (cl) *cl = (-f[0] * sa + f[1] * ca) / ql;
    if (cd) *cd = (f[0] * ca + f[1] * sa) / ql;
    return JST_OK;
}

jst_status jst_wall(jst_ctx *x, double *wx, double *wy, double *cp, double *cf)
{
    if (!x) return fail(JST_ERR_ARG, "ctx is NULL");
    if (x->mix) return batch_wall(x->mix, x->mix_id, wx, wy, cp, cf);
    double *dout = x->scratch + 4 * (size_t)x->nj;
    HIPCHK(jst::launch_forces_fp64(x->d_case, 0, x->scratch, dout, x->stream));
    return wall_distribution(x->scratch, x->g, x->nj, x->cfg, x->d, x->stream, wx, wy, cp, cf);
}

jst_status jst_write_wall_csv(jst_ctx *x, const char *path)
{
    if (!x || !path) return fail(JST_ERR_ARG, "NULL argument");
    const size_t n = (size_t)x->nj;
    std::vector<double> wx(n), wy(n), cp(n), cf(n);
    jst_status st = jst_wall(x, wx.data(), wy.data(), cp.data(), cf.data());
    if (st != JST_OK) return st;
    FILE *f = fopen(path, "w");
    if (!f) return fail(JST_ERR_IO, std::string("I/O error: cannot create ") + path);
    fprintf(f, "j,x,y,cp,cf\n");
    for (size_t j = 0; j < n; ++j)
        fprintf(f, "%zu,%s,%s,%s,%s\n", j + 1, rust_exp(wx[j], 8).c_str(), rust_exp(wy[j], 8).c_str(),
                rust_exp(cp[j], 8).c_str(), rust_exp(cf[j], 8).c_str());
    fclose(f);
    return JST_OK;
}

jst_status jst_write_csv(jst_ctx *x, const char *path)
{
    if (!x || !path) return fail(JST_ERR_ARG, "NULL argument");
    const size_t n = (size_t)x->ni * x->nj;
    std::vector<double> cx(n), cy(n), rho(n), p(n), t(n), u(n), v(n), nut(n);
    const char *names[] = {"cx", "cy", "rho", "p", "T", "u", "v", "nut"};
    double *bufs[] = {cx.data(), cy.data(), rho.data(), p.data(), t.data(), u.data(), v.data(), nut.data()};
    for (int q = 0; q < 8; ++q) {
        jst_status st = jst_get_field(x, names[q], bufs[q], n);
        if (st != JST_OK) return st;
    }
    FILE *f = fopen(path, "w");
    if (!f) return fail(JST_ERR_IO, std::string("I/O error: cannot create ") + path);
    fprintf(f, "i,j,x,y,rho,p,T,u,v,miubl\n");
    for (int i = 0; i < x->ni; ++i)
        for (int j = 0; j < x->nj; ++j) {
            size_t k = (size_t)i * x->nj + j;
            fprintf(f, "%d,%d,%s,%s,%s,%s,%s,%s,%s,%s\n", i + 1, j + 1, rust_exp(cx[k], 8).c_str(),
                    rust_exp(cy[k], 8).c_str(), rust_exp(rho[k], 8).c_str(), rust_exp(p[k], 8).c_str(),
                    rust_exp(t[k], 8).c_str(), rust_exp(u[k], 8).c_str(), rust_exp(v[k], 8).c_str(),
                    rust_exp(nut[k], 8).c_str());
        }
    fclose(f);
    return JST_OK;
}

jst_status jst_time_steps(jst_ctx *x, int nsteps, double *total_ms, double *stage_ms)
{
    if (!x || nsteps <= 0) return fail(JST_ERR_ARG, "bad arguments");
    if (x->mix) return jst_batch_time_steps(x->mix, nsteps, total_ms, stage_ms);
    jst_status st;
    if ((st = ctl_set(x, &CaseCtl::timing, 1)) != JST_OK) return st;
    if ((st = ctl_set(x, &CaseCtl::status, (int)JST_RUN_ACTIVE)) != JST_OK) return st;
    HIPCHK(hipEventRecord(x->ev[0], x->stream));
    if ((st = launch_steps(x, nsteps)) != JST_OK) return st;
    HIPCHK(hipEventRecord(x->ev[1], x->stream));
    HIPCHK(hipEventSynchronize(x->ev[1]));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, x->ev[0], x->ev[1]));
    if (total_ms) *total_ms = ms;
    if (stage_ms) {
        std::vector<hipEvent_t> evs(10 * (size_t)nsteps);
        for (auto &e : evs) HIPCHK(hipEventCreate(&e));
        for (int s = 0; s < nsteps; ++s) {
            for (int k = 0; k < 5; ++k) {
                HIPCHK(hipEventRecord(evs[10 * s + 2 * k], x->stream));
                HIPCHK(jst::launch_stage_fp64(x->d_case, x->d_tiles, x->ntiles, k, RK_ALPHA[k], false,
                                              x->stream));
                HIPCHK(hipEventRecord(evs[10 * s + 2 * k + 1], x->stream));
            }
            HIPCHK(jst::launch_post_fp64(x->d_case, x->d_tiles, x->ntiles, 1, x->stream));
        }
        HIPCHK(hipStreamSynchronize(x->stream));
        double sum = 0.0;
        for (int q = 0; q < 5 * nsteps; ++q) {
            float e = 0.f;
            HIPCHK(hipEventElapsedTime(&e, evs[2 * q], evs[2 * q + 1]));
            sum += e;
        }
        for (auto &e : evs) (void)hipEventDestroy(e);
        *stage_ms = sum / (5.0 * nsteps);
    }
    if ((st = ctl_set(x, &CaseCtl::timing, 0)) != JST_OK) return st;
    st = ctl_fetch(x);
    if (st != JST_OK) return st;
    if (x->h_ctl->err_key != JST_ERR_NONE) return report_nonphysical(x, nullptr);
    return JST_OK;
}


jst_status jst_time_begin(jst_ctx *x, int nsteps)
{
    if (!x || nsteps <= 0) return fail(JST_ERR_ARG, "bad arguments");
    if (x->mix) return jst_batch_time_begin(x->mix, nsteps);
    jst_status st;
    if ((st = ctl_set(x, &CaseCtl::timing, 1)) != JST_OK) return st;
    if ((st = ctl_set(x, &CaseCtl::status, (int)JST_RUN_ACTIVE)) != JST_OK) return st;
    HIPCHK(hipStreamSynchronize(x->stream));
    return JST_OK;
}

jst_status jst_time_run(jst_ctx *x, int nsteps)
{
    if (!x || nsteps <= 0) return fail(JST_ERR_ARG, "bad arguments");
    if (x->mix) return jst_batch_step(x->mix, nsteps, 1);
    jst_status st = launch_steps(x, nsteps);
    if (st != JST_OK) return st;
    HIPCHK(hipStreamSynchronize(x->stream));
    return JST_OK;
}

jst_status jst_time_end(jst_ctx *x)
{
    if (!x) return fail(JST_ERR_ARG, "bad arguments");
    if (x->mix) return jst_batch_time_end(x->mix);
    jst_status st;
    if ((st = ctl_set(x, &CaseCtl::timing, 0)) != JST_OK) return st;
    if ((st = ctl_fetch(x)) != JST_OK) return st;
    if (x->h_ctl->err_key != JST_ERR_NONE) return report_nonphysical(x, nullptr);
    return JST_OK;
}

} // extern "C"
