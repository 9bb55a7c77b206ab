// bfm_io.cpp -- frames, lineage and JSON configuration of the 3-D driver
// (ABM-3d/ABM-3d-cpp/include/utils.hpp:68-311, src/abmConstAdhesion.cpp:31-457,
// include/simulation.hpp:1444-1572, 2023-2079), host code only.
//
// * Frames: "# key = value" header lines, then one tab-separated row per cell
//   with 10 significant digits (utils.hpp:146-198); the id and group columns
//   as integers, the other required columns as doubles, extra columns only
//   where the caller lists them.
// * Lineage: "child<TAB>parent" per line (utils.hpp:207-258).
// * Configuration: the JSON keys of abmConstAdhesion.cpp, read with the
//   value kinds Boost.JSON would give them (an integer literal is not a
//   double, and the reverse), required keys and optional keys as there.
// * bfm_run: the driver main() -- founder cell or initial population plus
//   lineage, surface energy from the target equilibrium distance, the JKR
//   table of runSimulation, the growth loop on the device with a frame at
//   every dt_write crossing, the final frame and the lineage file.
#include <algorithm>
#include <array>
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <sstream>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include "../../include/bfm.h"
#include "bfm_device.hpp"
#include "bfm_internal.hpp"

namespace {

struct IoError {
    int code;
    std::string msg;
};

[[noreturn]] void io_fail(int code, const std::string& msg) { throw IoError{code, msg}; }

thread_local std::string g_io_error;

template <class F>
int guarded(F&& f)
{
    try {
        f();
        return BFM_OK;
    } catch (const IoError& e) {
        g_io_error = e.msg;
        return e.code;
    } catch (const std::exception& e) {
        g_io_error = e.what();
        return BFM_ERR_ARG;
    }
}

// ---- number formatting ----------------------------------------------------
// `ostream << setprecision(p) << x` formats as printf's %.{p}g.
std::string fmt_g(double x, int precision = 10)
{
    char buf[64];
    std::snprintf(buf, sizeof(buf), "%.*g", precision, x);
    return buf;
}

// ---- frames (utils.hpp:95-198) --------------------------------------------
constexpr int kColId = 0, kColGroup = 20, kNcolsRequired = 21;

void write_frame(const std::string& path, const double* cells, int n, int ld, int ncols,
                 const std::map<std::string, std::string>& params, const std::unordered_map<int, int>& extra)
{
    std::string out;
    out.reserve((size_t)n * ncols * 14 + 4096);
    for (const auto& kv : params) {
        out += "# ";
        out += kv.first;
        out += " = ";
        out += kv.second;
        out += '\n';
    }
    char buf[64];
    for (int i = 0; i < n; ++i) {
        bool any = false;
        for (int j = 0; j < ncols; ++j) {
            const double v = cells[(size_t)j * ld + i];
            int kind;   // 0: int, 1: double, -1: skipped
            if (j == kColId || j == kColGroup) kind = 0;
            else if (j > kColId && j < kColGroup) kind = 1;
            else {
                auto it = extra.find(j);
                kind = it == extra.end() ? -1 : (it->second == 0 ? 0 : 1);
            }
            if (kind < 0) continue;
            if (kind == 0) std::snprintf(buf, sizeof(buf), "%d", (int)v);
            else std::snprintf(buf, sizeof(buf), "%.10g", v);
            out += buf;
            out += '\t';
            any = true;
        }
        if (any) out.pop_back();   // the last tab
        out += '\n';
    }
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) io_fail(BFM_ERR_ARG, "cannot open " + path + " for writing: " + std::strerror(errno));
    const size_t w = std::fwrite(out.data(), 1, out.size(), f);
    const int rc = std::fclose(f);
    if (w != out.size() || rc != 0) io_fail(BFM_ERR_ARG, "write to " + path + " failed");
}

struct Frame {
    std::vector<std::vector<double>> rows;
    std::vector<std::pair<std::string, std::string>> params;   // in file order, later duplicates win on lookup
    int ncols = kNcolsRequired;
};

Frame read_frame(const std::string& path)
{
    std::ifstream in(path);
    if (!in) io_fail(BFM_ERR_ARG, "cannot open " + path);
    Frame fr;
    std::string line;
    while (std::getline(in, line)) {
        if (!line.empty() && line.back() == '\r') line.pop_back();
        if (line.empty()) continue;
        if (line[0] == '#') {
            const size_t eq = line.find(" = ");
            if (eq == std::string::npos || eq < 2) io_fail(BFM_ERR_ARG, "malformed parameter line in " + path + ": " + line);
            fr.params.emplace_back(line.substr(2, eq - 2), line.substr(eq + 3));
            continue;
        }
        std::vector<double> row;
        size_t pos = 0;
        while (pos <= line.size()) {
            size_t tab = line.find('\t', pos);
            if (tab == std::string::npos) tab = line.size();
            const std::string tok = line.substr(pos, tab - pos);
            char* end = nullptr;
            errno = 0;
            const double v = std::strtod(tok.c_str(), &end);
            if (tok.empty() || end == tok.c_str())
                io_fail(BFM_ERR_ARG, "bad number '" + tok + "' in " + path + " row " + std::to_string(fr.rows.size()));
            row.push_back(v);
            pos = tab + 1;
        }
        fr.ncols = std::max<int>(fr.ncols, (int)row.size());
        fr.rows.push_back(std::move(row));
    }
    return fr;
}

void write_lineage_file(const std::string& path, const int* parents, int n)
{
    std::string out;
    char buf[48];
    for (int i = 0; i < n; ++i) {
        std::snprintf(buf, sizeof(buf), "%d\t%d\n", i, parents[i]);
        out += buf;
    }
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) io_fail(BFM_ERR_ARG, "cannot open " + path + " for writing: " + std::strerror(errno));
    const size_t w = std::fwrite(out.data(), 1, out.size(), f);
    const int rc = std::fclose(f);
    if (w != out.size() || rc != 0) io_fail(BFM_ERR_ARG, "write to " + path + " failed");
}

// readLineage (utils.hpp:207-233) and the vector the driver builds from it
// (abmConstAdhesion.cpp:386-393): parents[child] = parent, one entry per line.
std::vector<int> read_lineage_file(const std::string& path)
{
    std::ifstream in(path);
    if (!in) io_fail(BFM_ERR_ARG, "cannot open " + path);
    std::map<int, int> lin;
    std::string line;
    while (std::getline(in, line)) {
        if (line.empty() || line == "\r") continue;
        char* end = nullptr;
        const long child = std::strtol(line.c_str(), &end, 10);
        if (end == line.c_str() || *end != '\t') io_fail(BFM_ERR_ARG, "malformed lineage line: " + line);
        const char* p2 = end + 1;
        const long parent = std::strtol(p2, &end, 10);
        if (end == p2) io_fail(BFM_ERR_ARG, "malformed lineage line: " + line);
        lin[(int)child] = (int)parent;
    }
    std::vector<int> parents(lin.size(), -1);
    for (const auto& kv : lin) {
        if (kv.first < 0 || kv.first >= (int)parents.size())
            io_fail(BFM_ERR_ARG, "lineage ids must be 0 .. (lines - 1); found " + std::to_string(kv.first));
        parents[kv.first] = kv.second;
    }
    return parents;
}

// ---- anisotropic JKR tables (simulation.hpp:550-1008) ----------------------
// Fields of a tab-separated line, as successive getline(ss, token, '\t')
// calls see them (a missing field reads as the empty string).
std::vector<std::string> tab_fields(const std::string& line)
{
    std::vector<std::string> f;
    size_t pos = 0;
    while (true) {
        const size_t tab = line.find('\t', pos);
        f.push_back(line.substr(pos, tab == std::string::npos ? std::string::npos : tab - pos));
        if (tab == std::string::npos) break;
        pos = tab + 1;
    }
    return f;
}

// std::stod: leading white space, the longest number prefix; no number or a
// value out of range is an error.
double stod_like(const std::string& tok, const std::string& path, int64_t lineno)
{
    char* end = nullptr;
    errno = 0;
    const double v = std::strtod(tok.c_str(), &end);
    if (end == tok.c_str() || errno == ERANGE)
        io_fail(BFM_ERR_ARG, "bad number '" + tok + "' in " + path + " line " + std::to_string(lineno));
    return v;
}

double field(const std::vector<std::string>& f, size_t k, const std::string& path, int64_t lineno)
{
    return stod_like(k < f.size() ? f[k] : std::string(), path, lineno);
}

struct AnisoTables {
    std::vector<double> theta, half_l, coords, curv, rx, phi, delta, gamma, fr;
    double fixed_gamma = 0;
    bfm_aniso_jkr_tables view;
};

// parseCurvatureRadiiTable (simulation.hpp:565-660): lines (theta, half_l,
// coord, Rx, Ry) enumerated theta-major. A new theta value starts the next
// theta index (half_l and coord indices back to 0), else a new half_l value
// the next half_l index, else the line is the next coordinate; the grids are
// the theta values, the half_l values under the first theta and the
// coordinates under the first (theta, half_l). The reference keeps (Rx, Ry)
// per index triple in a hash map; only Rx enters the forces.
void parse_curvature_table(const std::string& path, AnisoTables& T)
{
    std::ifstream in(path);
    if (!in) io_fail(BFM_ERR_ARG, "cannot open " + path);
    std::string line;
    if (!std::getline(in, line)) io_fail(BFM_ERR_ARG, path + " is empty");
    int64_t lineno = 1;
    auto f = tab_fields(line);
    double th = field(f, 0, path, lineno), hl = field(f, 1, path, lineno), co = field(f, 2, path, lineno);
    double rx = field(f, 3, path, lineno);
    field(f, 4, path, lineno);
    T.theta.assign(1, th);
    T.half_l.assign(1, hl);
    T.coords.assign(1, co);
    std::vector<std::pair<std::array<int, 3>, double>> ent;
    ent.push_back({{0, 0, 0}, rx});
    int ti = 0, hi = 0, ci = 0;
    while (std::getline(in, line)) {
        ++lineno;
        f = tab_fields(line);
        const double th2 = field(f, 0, path, lineno), hl2 = field(f, 1, path, lineno);
        const double co2 = field(f, 2, path, lineno);
        rx = field(f, 3, path, lineno);
        field(f, 4, path, lineno);
        if (th2 != th) {
            T.theta.push_back(th2);
            ++ti;
            hi = ci = 0;
        } else if (hl2 != hl) {
            if (ti == 0) T.half_l.push_back(hl2);
            ++hi;
            ci = 0;
        } else {
            if (ti == 0 && hi == 0) T.coords.push_back(co2);
            ++ci;
        }
        ent.push_back({{ti, hi, ci}, rx});
        th = th2;
        hl = hl2;
        co = co2;
    }
    const int nt = (int)T.theta.size(), nh = (int)T.half_l.size(), nc = (int)T.coords.size();
    T.curv.assign((size_t)nt * nh * nc, 0.0);
    // index triples outside the grids (ragged files) cannot be looked up
    for (const auto& e : ent)
        if (e.first[0] < nt && e.first[1] < nh && e.first[2] < nc)
            T.curv[((size_t)e.first[0] * nh + e.first[1]) * nc + e.first[2]] = e.second;
}

// parseReducedJKRForceTable (simulation.hpp:662-836, first line: the fixed
// gamma) and parseJKRForceTable (:838-1008, no header, a gamma column). The
// Rx grid is the first column's values in order of first appearance, keyed
// by the text of the field (a second-column text the first column never
// holds maps to index 0, as the reference's map lookup inserts it). Lines are
// enumerated (Rx1, Rx2, phi, delta[, gamma]); a new Rx1 or Rx2 value resets
// the inner indices, a new phi value starts the next phi index, and so on;
// the inner grids come from the first block.
void parse_force_table(const std::string& path, bool variable_gamma, AnisoTables& T)
{
    std::ifstream in(path);
    if (!in) io_fail(BFM_ERR_ARG, "cannot open " + path);
    std::string line;
    int64_t lineno = 0;
    if (!variable_gamma) {
        if (!std::getline(in, line)) io_fail(BFM_ERR_ARG, path + " is empty");
        T.fixed_gamma = stod_like(line, path, ++lineno);
    }
    const int64_t first_row = lineno + 1;
    std::unordered_map<std::string, int> rx_idx;
    while (std::getline(in, line)) {
        ++lineno;
        const std::string tok = tab_fields(line)[0];
        if (rx_idx.find(tok) == rx_idx.end()) {
            T.rx.push_back(stod_like(tok, path, lineno));
            rx_idx[tok] = (int)T.rx.size() - 1;
        }
    }
    in.clear();
    in.seekg(0);
    lineno = 0;
    if (!variable_gamma) { std::getline(in, line); ++lineno; }
    if (!std::getline(in, line)) io_fail(BFM_ERR_ARG, path + " has no force rows");
    lineno = first_row;
    auto rx_index = [&](const std::string& tok) {
        auto it = rx_idx.find(tok);
        if (it == rx_idx.end()) { rx_idx[tok] = 0; return 0; }
        return it->second;
    };
    const int nv = variable_gamma ? 7 : 6;
    auto f = tab_fields(line);
    std::vector<double> v(nv);
    for (int k = 0; k < nv; ++k) v[k] = field(f, k, path, lineno);
    int i1 = rx_index(f[0]), i2 = rx_index(f.size() > 1 ? f[1] : std::string());
    T.phi.assign(1, v[2]);
    T.delta.assign(1, v[3]);
    if (variable_gamma) T.gamma.assign(1, v[4]);
    std::vector<std::pair<std::array<int, 5>, std::array<double, 2>>> ent;
    ent.push_back({{i1, i2, 0, 0, 0}, {v[nv - 2], v[nv - 1]}});
    double r1 = v[0], r2 = v[1], ph = v[2], de = v[3], ga = variable_gamma ? v[4] : 0.0;
    int pi = 0, di = 0, gi = 0;
    while (std::getline(in, line)) {
        ++lineno;
        f = tab_fields(line);
        for (int k = 0; k < nv; ++k) v[k] = field(f, k, path, lineno);
        i1 = rx_index(f[0]);
        i2 = rx_index(f[1]);
        if (v[0] != r1 || v[1] != r2) {
            pi = di = gi = 0;
        } else if (v[2] != ph) {
            if (i1 == 0 && i2 == 0) T.phi.push_back(v[2]);
            ++pi;
            di = gi = 0;
        } else if (!variable_gamma || v[3] != de) {
            if (i1 == 0 && i2 == 0 && pi == 0) T.delta.push_back(v[3]);
            ++di;
            gi = 0;
        } else {
            if (i1 == 0 && i2 == 0 && pi == 0 && di == 0) T.gamma.push_back(v[4]);
            ++gi;
        }
        ent.push_back({{i1, i2, pi, di, gi}, {v[nv - 2], v[nv - 1]}});
        r1 = v[0];
        r2 = v[1];
        ph = v[2];
        de = v[3];
        if (variable_gamma) ga = v[4];
    }
    (void)ga;
    const int nr = (int)T.rx.size(), np = (int)T.phi.size(), nd = (int)T.delta.size();
    const int ng = variable_gamma ? (int)T.gamma.size() : 1;
    T.fr.assign((size_t)nr * nr * np * nd * ng * 2, 0.0);
    for (const auto& e : ent) {
        const auto& k = e.first;
        if (k[0] >= nr || k[1] >= nr || k[2] >= np || k[3] >= nd || k[4] >= ng) continue;
        const size_t q = (((((size_t)k[0] * nr + k[1]) * np + k[2]) * nd + k[3]) * ng + k[4]) * 2;
        T.fr[q] = e.second[0];
        T.fr[q + 1] = e.second[1];
    }
}

void aniso_view(AnisoTables& T, bool variable_gamma)
{
    bfm_aniso_jkr_tables& v = T.view;
    v.theta = T.theta.data();   v.n_theta = (int)T.theta.size();
    v.half_l = T.half_l.data(); v.n_half_l = (int)T.half_l.size();
    v.coords = T.coords.data(); v.n_coords = (int)T.coords.size();
    v.curvature_rx = T.curv.data();
    v.rx = T.rx.data();         v.n_rx = (int)T.rx.size();
    v.phi = T.phi.data();       v.n_phi = (int)T.phi.size();
    v.delta = T.delta.data();   v.n_delta = (int)T.delta.size();
    v.gamma = variable_gamma ? T.gamma.data() : nullptr;
    v.n_gamma = variable_gamma ? (int)T.gamma.size() : 0;
    v.force_radius = T.fr.data();
}

std::mutex g_aniso_mu;
std::map<const bfm_aniso_jkr_tables*, std::unique_ptr<AnisoTables>> g_aniso;

// ---- JSON (the subset Boost.JSON's parser accepts without extensions) ------
struct JVal {
    enum Kind { Null, Bool, Int, Uint, Dbl, Str, Arr, Obj } kind = Null;
    bool b = false;
    int64_t i = 0;
    uint64_t u = 0;
    double d = 0;
    std::string s;
    std::vector<JVal> a;
    std::vector<std::pair<std::string, JVal>> o;
};

struct JParser {
    const std::string& t;
    size_t p = 0;
    explicit JParser(const std::string& text) : t(text) {}
    [[noreturn]] void bad(const char* what)
    {
        io_fail(BFM_ERR_ARG, std::string("config is not valid JSON (") + what + " at offset " + std::to_string(p) + ")");
    }
    void ws()
    {
        while (p < t.size() && (t[p] == ' ' || t[p] == '\t' || t[p] == '\n' || t[p] == '\r')) ++p;
    }
    bool lit(const char* w)
    {
        const size_t n = std::strlen(w);
        if (t.compare(p, n, w) == 0) { p += n; return true; }
        return false;
    }
    static void put_utf8(std::string& s, uint32_t c)
    {
        if (c < 0x80) s += (char)c;
        else if (c < 0x800) { s += (char)(0xC0 | (c >> 6)); s += (char)(0x80 | (c & 0x3F)); }
        else if (c < 0x10000) {
            s += (char)(0xE0 | (c >> 12)); s += (char)(0x80 | ((c >> 6) & 0x3F)); s += (char)(0x80 | (c & 0x3F));
        } else {
            s += (char)(0xF0 | (c >> 18)); s += (char)(0x80 | ((c >> 12) & 0x3F));
            s += (char)(0x80 | ((c >> 6) & 0x3F)); s += (char)(0x80 | (c & 0x3F));
        }
    }
    uint32_t hex4()
    {
        if (p + 4 > t.size()) bad("short \\u escape");
        uint32_t v = 0;
        for (int k = 0; k < 4; ++k) {
            const char ch = t[p++];
            v <<= 4;
            if (ch >= '0' && ch <= '9') v |= ch - '0';
            else if (ch >= 'a' && ch <= 'f') v |= ch - 'a' + 10;
            else if (ch >= 'A' && ch <= 'F') v |= ch - 'A' + 10;
            else bad("bad \\u escape");
        }
        return v;
    }
    std::string str()
    {
        if (t[p] != '"') bad("expected string");
        ++p;
        std::string s;
        for (;;) {
            if (p >= t.size()) bad("unterminated string");
            const char ch = t[p++];
            if (ch == '"') break;
            if ((unsigned char)ch < 0x20) bad("control character in string");
            if (ch != '\\') { s += ch; continue; }
            if (p >= t.size()) bad("unterminated escape");
            const char e = t[p++];
            switch (e) {
            case '"': s += '"'; break;
            case '\\': s += '\\'; break;
            case '/': s += '/'; break;
            case 'b': s += '\b'; break;
            case 'f': s += '\f'; break;
            case 'n': s += '\n'; break;
            case 'r': s += '\r'; break;
            case 't': s += '\t'; break;
            case 'u': {
                uint32_t c = hex4();
                if (c >= 0xD800 && c < 0xDC00 && t.compare(p, 2, "\\u") == 0) {
                    p += 2;
                    const uint32_t lo = hex4();
                    if (lo < 0xDC00 || lo > 0xDFFF) bad("bad surrogate pair");
                    c = 0x10000 + ((c - 0xD800) << 10) + (lo - 0xDC00);
                }
                put_utf8(s, c);
                break;
            }
            default: bad("bad escape");
            }
        }
        return s;
    }
    JVal num()
    {
        const size_t s0 = p;
        if (t[p] == '-') ++p;
        if (p >= t.size() || !std::isdigit((unsigned char)t[p])) bad("bad number");
        if (t[p] == '0') ++p;
        else while (p < t.size() && std::isdigit((unsigned char)t[p])) ++p;
        bool is_int = true;
        if (p < t.size() && t[p] == '.') {
            is_int = false;
            ++p;
            if (p >= t.size() || !std::isdigit((unsigned char)t[p])) bad("bad fraction");
            while (p < t.size() && std::isdigit((unsigned char)t[p])) ++p;
        }
        if (p < t.size() && (t[p] == 'e' || t[p] == 'E')) {
            is_int = false;
            ++p;
            if (p < t.size() && (t[p] == '+' || t[p] == '-')) ++p;
            if (p >= t.size() || !std::isdigit((unsigned char)t[p])) bad("bad exponent");
            while (p < t.size() && std::isdigit((unsigned char)t[p])) ++p;
        }
        const std::string tok = t.substr(s0, p - s0);
        JVal v;
        if (is_int) {
            // int64 where it fits, else uint64, else double (Boost.JSON's kinds)
            errno = 0;
            const long long ll = std::strtoll(tok.c_str(), nullptr, 10);
            if (errno == 0) { v.kind = JVal::Int; v.i = ll; return v; }
            if (tok[0] != '-') {
                errno = 0;
                const unsigned long long ull = std::strtoull(tok.c_str(), nullptr, 10);
                if (errno == 0) { v.kind = JVal::Uint; v.u = ull; return v; }
            }
        }
        v.kind = JVal::Dbl;
        v.d = std::strtod(tok.c_str(), nullptr);
        return v;
    }
    JVal value(int depth = 0)
    {
        if (depth > 64) bad("nesting too deep");
        ws();
        if (p >= t.size()) bad("unexpected end");
        JVal v;
        const char ch = t[p];
        if (ch == '{') {
            ++p;
            v.kind = JVal::Obj;
            ws();
            if (p < t.size() && t[p] == '}') { ++p; return v; }
            for (;;) {
                ws();
                if (p >= t.size()) bad("unexpected end in object");
                std::string k = str();
                ws();
                if (p >= t.size() || t[p] != ':') bad("expected ':'");
                ++p;
                JVal x = value(depth + 1);
                // a repeated key replaces the earlier value
                bool found = false;
                for (auto& kv : v.o)
                    if (kv.first == k) { kv.second = std::move(x); found = true; break; }
                if (!found) v.o.emplace_back(std::move(k), std::move(x));
                ws();
                if (p < t.size() && t[p] == ',') { ++p; continue; }
                if (p < t.size() && t[p] == '}') { ++p; return v; }
                bad("expected ',' or '}'");
            }
        }
        if (ch == '[') {
            ++p;
            v.kind = JVal::Arr;
            ws();
            if (p < t.size() && t[p] == ']') { ++p; return v; }
            for (;;) {
                v.a.push_back(value(depth + 1));
                ws();
                if (p < t.size() && t[p] == ',') { ++p; continue; }
                if (p < t.size() && t[p] == ']') { ++p; return v; }
                bad("expected ',' or ']'");
            }
        }
        if (ch == '"') { v.kind = JVal::Str; v.s = str(); return v; }
        if (lit("true")) { v.kind = JVal::Bool; v.b = true; return v; }
        if (lit("false")) { v.kind = JVal::Bool; v.b = false; return v; }
        if (lit("null")) return v;
        if (ch == '-' || std::isdigit((unsigned char)ch)) return num();
        bad("unexpected character");
    }
};

JVal parse_json_text(const std::string& text)
{
    JParser jp(text);
    JVal v = jp.value();
    jp.ws();
    if (jp.p != text.size()) jp.bad("trailing characters");
    return v;
}

// Config lookups with the reference's value kinds: as_double() needs a JSON
// number with a fraction or exponent, as_int64() an integer literal.
struct Cfg {
    const JVal& obj;
    const JVal* find(const char* key) const
    {
        for (const auto& kv : obj.o)
            if (kv.first == key) return &kv.second;
        return nullptr;
    }
    bool has(const char* key) const
    {
        const JVal* v = find(key);
        return v && v->kind != JVal::Null;
    }
    double dbl(const char* key) const
    {
        const JVal* v = find(key);
        if (!v || v->kind == JVal::Null) io_fail(BFM_ERR_ARG, std::string("config: missing key \"") + key + "\"");
        if (v->kind != JVal::Dbl)
            io_fail(BFM_ERR_ARG, std::string("config: key \"") + key + "\" must be a floating-point number (e.g. 1.0)");
        return v->d;
    }
    int64_t i64(const char* key) const
    {
        const JVal* v = find(key);
        if (!v || v->kind == JVal::Null) io_fail(BFM_ERR_ARG, std::string("config: missing key \"") + key + "\"");
        if (v->kind == JVal::Int) return v->i;
        if (v->kind == JVal::Uint && v->u <= (uint64_t)INT64_MAX) return (int64_t)v->u;
        io_fail(BFM_ERR_ARG, std::string("config: key \"") + key + "\" must be an integer");
    }
    std::string string(const char* key) const
    {
        const JVal* v = find(key);
        if (!v || v->kind == JVal::Null) io_fail(BFM_ERR_ARG, std::string("config: missing key \"") + key + "\"");
        if (v->kind != JVal::Str) io_fail(BFM_ERR_ARG, std::string("config: key \"") + key + "\" must be a string");
        return v->s;
    }
    // optional keys: the reference keeps its default when the key is absent or
    // of another kind (the failed conversion is caught)
    double opt_dbl(const char* key, double dflt) const
    {
        const JVal* v = find(key);
        return (v && v->kind == JVal::Dbl) ? v->d : dflt;
    }
    int64_t opt_i64(const char* key, int64_t dflt) const
    {
        const JVal* v = find(key);
        if (v && v->kind == JVal::Int) return v->i;
        return dflt;
    }
    std::string opt_string(const char* key) const
    {
        const JVal* v = find(key);
        return (v && v->kind == JVal::Str) ? v->s : std::string();
    }
};

void copy_name(char* dst, size_t cap, const std::string& s, const char* what)
{
    if (s.size() + 1 > cap) io_fail(BFM_ERR_ARG, std::string("config: ") + what + " is longer than " + std::to_string(cap - 1) + " bytes");
    std::memcpy(dst, s.c_str(), s.size() + 1);
}

// abmConstAdhesion.cpp:31-333
void load_config(const std::string& path, bfm_run_config* o)
{
    std::ifstream in(path);
    if (!in) io_fail(BFM_ERR_ARG, "cannot open " + path);
    // the reference feeds the file line by line without the newlines
    std::string text, line;
    while (std::getline(in, line)) text += line;
    const JVal root = parse_json_text(text);
    if (root.kind != JVal::Obj) io_fail(BFM_ERR_ARG, "config: the top-level JSON value must be an object");
    const Cfg c{root};
    std::memset(o, 0, sizeof(*o));
    o->R = c.dbl("R");
    o->Rcell = c.dbl("Rcell");
    o->L0 = c.dbl("L0");
    o->Ldiv = 2 * o->L0 + 2 * o->R;
    o->E0 = c.dbl("E0");
    o->Ecell = c.dbl("Ecell");
    o->sigma0 = c.dbl("sigma0");
    o->eta_ambient = c.dbl("eta_ambient");
    o->eta_surface = c.dbl("eta_surface");
    o->max_stepsize = c.dbl("max_stepsize");
    o->min_stepsize = c.dbl("min_stepsize");
    o->dt_write = c.dbl("dt_write");
    o->iter_update_stepsize = (int)c.i64("iter_update_stepsize");
    o->iter_update_neighbors = (int)c.i64("iter_update_neighbors");
    o->neighbor_threshold = 0.5 * (2 * o->R + o->Ldiv);
    o->max_iter = (int)c.i64("max_iter");
    o->n_cells = (int)c.i64("n_cells");
    o->max_time = c.dbl("max_time");
    o->growth_mean = c.dbl("growth_mean");
    o->growth_std = c.dbl("growth_std");
    o->daughter_length_std = c.dbl("daughter_length_std");
    o->daughter_angle_xy_bound = c.dbl("daughter_angle_xy_bound");
    o->daughter_angle_z_bound = c.dbl("daughter_angle_z_bound");
    o->nz_threshold = c.dbl("nz_threshold");
    o->max_rxy_noise = c.dbl("max_rxy_noise");
    o->max_rz_noise = c.dbl("max_rz_noise");
    o->max_nxy_noise = c.dbl("max_nxy_noise");
    o->max_nz_noise = c.dbl("max_nz_noise");
    o->max_error_allowed = c.dbl("max_error_allowed");
    o->basal_only = c.i64("basal_only") != 0;
    o->basal_min_overlap = o->basal_only ? c.dbl("basal_min_overlap") : 0.0;
    const int64_t am = c.i64("adhesion_mode");
    if (am < 0 || am > 2) io_fail(BFM_ERR_ARG, "Invalid cell-cell adhesion mode specified");
    o->adhesion_mode = (int)am;
    const int64_t fm = c.i64("cell_cell_friction_mode");
    if (fm < 0 || fm > 1) io_fail(BFM_ERR_ARG, "Invalid cell-cell friction mode specified");
    o->friction_mode = (int)fm;
    o->jkr_imag_tol = 1e-8;
    o->jkr_aberth_tol = 1e-8;
    o->n_mesh_overlap = 100;
    o->n_mesh_gamma = 100;
    if (o->adhesion_mode != 0) {
        o->adhesion_eqdist = c.dbl("adhesion_eqdist");
        o->jkr_imag_tol = c.opt_dbl("adhesion_jkr_imag_tol", 1e-8);
        o->jkr_aberth_tol = c.opt_dbl("adhesion_jkr_aberth_tol", 1e-8);
        o->n_mesh_overlap = (int)c.opt_i64("adhesion_n_mesh_overlap", 100);
        o->n_mesh_gamma = (int)c.opt_i64("adhesion_n_mesh_gamma", 100);
    }
    // anisotropic tables (abmConstAdhesion.cpp:133-138, 242-249)
    o->precompute_jkr_forces = 0;
    o->curvature_filename[0] = o->jkr_forces_filename[0] = '\0';
    o->calibrate_endpoint_radii = 1;
    o->min_aspect_ratio = 0.01;   // not read from the file (abmConstAdhesion.cpp:146)
    o->ellipsoid_project_tol = 1e-8;
    o->brent_tol = 1e-8;
    o->init_bracket_dx = 1e-3;
    o->n_mesh_theta = o->n_mesh_half_l = o->n_mesh_centerline_coords = o->n_mesh_curvature_radii = 100;
    o->ellipsoid_project_max_iter = 100;
    o->brent_max_iter = 1000;
    o->n_tries_bracket = 5;
    if (o->adhesion_mode == 2) {
        o->precompute_jkr_forces = (int)c.i64("adhesion_precompute_jkr_forces");
        o->calibrate_endpoint_radii = (double)c.opt_i64("adhesion_calibrate_endpoint_radii", 1);
        o->n_mesh_theta = (int)c.opt_i64("adhesion_n_mesh_theta", 100);
        o->n_mesh_half_l = (int)c.opt_i64("adhesion_n_mesh_half_l", 100);
        o->n_mesh_centerline_coords = (int)c.opt_i64("adhesion_n_mesh_centerline_coords", 100);
        o->n_mesh_curvature_radii = (int)c.opt_i64("adhesion_n_mesh_curvature_radii", 100);
        o->ellipsoid_project_tol = c.opt_dbl("adhesion_ellipsoid_project_tol", 1e-8);
        o->ellipsoid_project_max_iter = (int)c.opt_i64("adhesion_ellipsoid_project_max_iter", 100);
        o->brent_tol = c.opt_dbl("adhesion_brent_tol", 1e-8);
        o->brent_max_iter = (int)c.opt_i64("adhesion_brent_max_iter", 1000);
        o->init_bracket_dx = c.opt_dbl("adhesion_init_bracket_dx", 1e-3);
        o->n_tries_bracket = (int)c.opt_i64("adhesion_n_tries_bracket", 5);
        if (o->precompute_jkr_forces == 0) {
            copy_name(o->curvature_filename, sizeof(o->curvature_filename), c.string("adhesion_curvature_filename"),
                      "adhesion_curvature_filename");
            copy_name(o->jkr_forces_filename, sizeof(o->jkr_forces_filename),
                      c.string("adhesion_jkr_forces_filename"), "adhesion_jkr_forces_filename");
        }
    }
    o->eta_cell_cell = o->friction_mode == 1 ? c.dbl("eta_cell_cell") : 0.0;
    o->cell_surface_coulomb_coeff = c.opt_dbl("cell_surface_coulomb_coeff", 1.0);
    o->cell_cell_coulomb_coeff = o->friction_mode == 1 ? c.opt_dbl("cell_cell_coulomb_coeff", 1.0) : 1.0;
    o->integration_mode = 1;   // Heun-Euler
    if (const JVal* v = c.find("integration_mode")) {
        if (v->kind == JVal::Int) {
            if (v->i < 0 || v->i > 4) io_fail(BFM_ERR_ARG, "Invalid integration mode specified");
            o->integration_mode = (int)v->i;
        }
    }
    o->M0 = o->integration_mode == 0 ? c.dbl("M0") : 0.0;
    o->no_surface = (int)(c.opt_i64("no_surface", 0) != 0);
    copy_name(o->init_filename, sizeof(o->init_filename), c.opt_string("init_filename"), "init_filename");
    copy_name(o->lineage_filename, sizeof(o->lineage_filename), c.opt_string("lineage_filename"), "lineage_filename");
}

// The parameter dictionary runSimulation writes into every frame
// (simulation.hpp:1444-1549) for the modes this driver runs: one group, no
// switching, isotropic JKR or none, no friction, no pole tracking.
std::map<std::string, std::string> frame_params(const bfm_run_config& o, int rng_seed, double max_gamma_unused)
{
    (void)max_gamma_unused;
    std::map<std::string, std::string> p;
    p["n_cells"] = std::to_string(o.n_cells);
    p["R"] = fmt_g(o.R);
    p["Rcell"] = fmt_g(o.Rcell);
    p["L0"] = fmt_g(o.L0);
    p["Ldiv"] = fmt_g(o.Ldiv);
    p["E0"] = fmt_g(o.E0);
    p["Ecell"] = fmt_g(o.Ecell);
    p["integration_mode"] = std::to_string(o.integration_mode);
    p["max_stepsize"] = fmt_g(o.max_stepsize);
    p["min_stepsize"] = fmt_g(o.min_stepsize);
    p["dt_write"] = fmt_g(o.dt_write);
    p["iter_update_neighbors"] = std::to_string(o.iter_update_neighbors);
    p["iter_update_stepsize"] = std::to_string(o.iter_update_stepsize);
    p["max_error_allowed"] = fmt_g(o.max_error_allowed);
    p["max_tries_update_stepsize"] = std::to_string(3);
    p["neighbor_threshold"] = fmt_g(o.neighbor_threshold);
    p["nz_threshold"] = fmt_g(o.nz_threshold);
    p["random_seed"] = std::to_string(rng_seed);
    p["n_groups"] = "1";
    p["growth_mean1"] = fmt_g(o.growth_mean);
    p["growth_std1"] = fmt_g(o.growth_std);
    p["switch_mode"] = "0";
    p["switch_timescale"] = fmt_g(0.0);
    p["daughter_length_std"] = fmt_g(o.daughter_length_std);
    p["daughter_angle_xy_bound"] = fmt_g(o.daughter_angle_xy_bound);
    p["daughter_angle_z_bound"] = fmt_g(o.daughter_angle_z_bound);
    p["max_rxy_noise"] = fmt_g(o.max_rxy_noise);
    p["max_rz_noise"] = fmt_g(o.max_rz_noise);
    p["max_nxy_noise"] = fmt_g(o.max_nxy_noise);
    p["max_nz_noise"] = fmt_g(o.max_nz_noise);
    p["basal_only"] = o.basal_only ? "1" : "0";
    p["basal_min_overlap"] = fmt_g(o.basal_min_overlap);
    p["adhesion_mode"] = std::to_string(o.adhesion_mode);
    if (o.adhesion_mode != 0) {
        // the adhesion_params map of the driver (abmConstAdhesion.cpp:95-131)
        p["adhesion_eqdist"] = fmt_g(o.adhesion_eqdist);
        p["adhesion_jkr_imag_tol"] = fmt_g(o.jkr_imag_tol);
        p["adhesion_jkr_aberth_tol"] = fmt_g(o.jkr_aberth_tol);
        p["adhesion_n_mesh_overlap"] = fmt_g((double)o.n_mesh_overlap);
        p["adhesion_n_mesh_gamma"] = fmt_g((double)o.n_mesh_gamma);
    }
    if (o.adhesion_mode == 2) {
        // the anisotropic entries of adhesion_params (abmConstAdhesion.cpp:136-240)
        p["adhesion_precompute_jkr_forces"] = fmt_g((double)o.precompute_jkr_forces);
        p["adhesion_calibrate_endpoint_radii"] = fmt_g(o.calibrate_endpoint_radii);
        p["adhesion_n_mesh_theta"] = fmt_g((double)o.n_mesh_theta);
        p["adhesion_n_mesh_half_l"] = fmt_g((double)o.n_mesh_half_l);
        p["adhesion_n_mesh_centerline_coords"] = fmt_g((double)o.n_mesh_centerline_coords);
        p["adhesion_n_mesh_curvature_radii"] = fmt_g((double)o.n_mesh_curvature_radii);
        p["adhesion_min_aspect_ratio"] = fmt_g(o.min_aspect_ratio);
        p["adhesion_ellipsoid_project_tol"] = fmt_g(o.ellipsoid_project_tol);
        p["adhesion_ellipsoid_project_max_iter"] = fmt_g((double)o.ellipsoid_project_max_iter);
        p["adhesion_brent_tol"] = fmt_g(o.brent_tol);
        p["adhesion_brent_max_iter"] = fmt_g((double)o.brent_max_iter);
        p["adhesion_init_bracket_dx"] = fmt_g(o.init_bracket_dx);
        p["adhesion_n_tries_bracket"] = fmt_g((double)o.n_tries_bracket);
    }
    p["cell_cell_friction_mode"] = std::to_string(o.friction_mode);
    p["cell_cell_coulomb_coeff"] = fmt_g(o.cell_cell_coulomb_coeff);
    p["cell_surface_coulomb_coeff"] = fmt_g(o.cell_surface_coulomb_coeff);
    p["track_poles"] = "0";
    p["no_surface"] = o.no_surface ? "1" : "0";
    p["n_cells_start_switch"] = "0";
    return p;
}

std::string cstr(const char* s) { return s ? std::string(s) : std::string(); }

}  // namespace

extern "C" {

const char* bfm_io_last_error(void) { return g_io_error.c_str(); }

int bfm_write_cells(const char* path, const double* cells, int n, int ld, int ncols, const char* const* keys,
                    const char* const* values, int n_params, const int* extra_cols, const int* extra_types,
                    int n_extra)
{
    return guarded([&] {
        if (!path || (n > 0 && !cells) || n < 0 || ncols < kNcolsRequired || ld < n || n_params < 0 || n_extra < 0 ||
            (n_params > 0 && (!keys || !values)) || (n_extra > 0 && (!extra_cols || !extra_types)))
            io_fail(BFM_ERR_ARG, "bfm_write_cells: need path, n >= 0, ld >= n, ncols >= 21 and the listed arrays");
        std::map<std::string, std::string> params;
        for (int k = 0; k < n_params; ++k) {
            if (!keys[k] || !values[k]) io_fail(BFM_ERR_ARG, "bfm_write_cells: NULL key or value");
            params[keys[k]] = values[k];
        }
        std::unordered_map<int, int> extra;
        for (int k = 0; k < n_extra; ++k) extra[extra_cols[k]] = extra_types[k];
        write_frame(path, cells, n, ld, ncols, params, extra);
    });
}

int bfm_read_cells(const char* path, double* cells, int cap, int ld, int max_cols, int* n_out, int* ncols_out,
                   char* params_out, int params_cap, int* params_len_out)
{
    return guarded([&] {
        if (!path || !n_out || !ncols_out) io_fail(BFM_ERR_ARG, "bfm_read_cells: need path, n_out and ncols_out");
        Frame fr = read_frame(path);
        *n_out = (int)fr.rows.size();
        *ncols_out = fr.ncols;
        std::string ptxt;
        for (const auto& kv : fr.params) ptxt += kv.first + " = " + kv.second + "\n";
        if (params_len_out) *params_len_out = (int)ptxt.size();
        if (params_out && params_cap > 0) {
            if ((int)ptxt.size() + 1 > params_cap) io_fail(BFM_ERR_CAPACITY, "bfm_read_cells: params buffer too small");
            std::memcpy(params_out, ptxt.c_str(), ptxt.size() + 1);
        }
        if (!cells) return;
        if (*n_out > cap || *n_out > ld || fr.ncols > max_cols)
            io_fail(BFM_ERR_CAPACITY, "bfm_read_cells: the frame has " + std::to_string(*n_out) + " rows x " +
                                          std::to_string(fr.ncols) + " columns, more than the buffer");
        for (int j = 0; j < fr.ncols; ++j)
            for (int i = 0; i < *n_out; ++i) {
                const auto& row = fr.rows[i];
                cells[(size_t)j * ld + i] = j < (int)row.size() ? row[j] : 0.0;
            }
    });
}

int bfm_write_lineage(const char* path, const int* parents, int n)
{
    return guarded([&] {
        if (!path || n < 0 || (n > 0 && !parents)) io_fail(BFM_ERR_ARG, "bfm_write_lineage: need path and parents");
        write_lineage_file(path, parents, n);
    });
}

int bfm_read_lineage(const char* path, int* parents, int cap, int* n_out)
{
    return guarded([&] {
        if (!path || !n_out) io_fail(BFM_ERR_ARG, "bfm_read_lineage: need path and n_out");
        const std::vector<int> v = read_lineage_file(path);
        *n_out = (int)v.size();
        if (!parents) return;
        if ((int)v.size() > cap) io_fail(BFM_ERR_CAPACITY, "bfm_read_lineage: more entries than cap");
        std::copy(v.begin(), v.end(), parents);
    });
}

int bfm_load_config(const char* json_path, bfm_run_config* out)
{
    return guarded([&] {
        if (!json_path || !out) io_fail(BFM_ERR_ARG, "bfm_load_config: need a path and an output struct");
        load_config(json_path, out);
    });
}

int bfm_read_aniso_jkr_tables(const char* curvature_path, const char* forces_path, int variable_gamma,
                              bfm_aniso_jkr_tables** out, double* fixed_gamma_out)
{
    return guarded([&] {
        if (!curvature_path || !forces_path || !out) io_fail(BFM_ERR_ARG, "bfm_read_aniso_jkr_tables: NULL argument");
        auto T = std::make_unique<AnisoTables>();
        parse_curvature_table(curvature_path, *T);
        parse_force_table(forces_path, variable_gamma != 0, *T);
        aniso_view(*T, variable_gamma != 0);
        if (fixed_gamma_out) *fixed_gamma_out = T->fixed_gamma;
        *out = &T->view;
        std::lock_guard<std::mutex> lk(g_aniso_mu);
        g_aniso[&T->view] = std::move(T);
    });
}

void bfm_free_aniso_jkr_tables(bfm_aniso_jkr_tables* tables)
{
    std::lock_guard<std::mutex> lk(g_aniso_mu);
    g_aniso.erase(tables);
}

int bfm_run(const char* json_path, const char* outprefix, int rng_seed, int device, int jkr_mode, int max_cells,
            int verbose, bfm_sim_stats* stats_out)
{
    bfm_ctx* ctx = nullptr;
    const int rc = guarded([&] {
        if (!json_path || !outprefix) io_fail(BFM_ERR_ARG, "bfm_run: need the config path and the output prefix");
        bfm_run_config o;
        load_config(json_path, &o);
        // runSimulation reads the anisotropic tables from the two files
        // (simulation.hpp:1405-1441); computing them at start-up is not built
        if (o.adhesion_mode == 2 && o.precompute_jkr_forces != 0)
            io_fail(BFM_ERR_UNSUPPORTED,
                    "anisotropic JKR tables computed at start-up (adhesion_precompute_jkr_forces != 0) are not supported");
        if (o.integration_mode == 0 && !(o.M0 > 0))
            io_fail(BFM_ERR_ARG, "Initial cell mass must be defined when using velocity Verlet");
        if (o.basal_only) io_fail(BFM_ERR_UNSUPPORTED, "basal_only is not supported");
        if (o.max_rxy_noise > 0 || o.max_rz_noise > 0 || o.max_nxy_noise > 0 || o.max_nz_noise > 0)
            io_fail(BFM_ERR_UNSUPPORTED, "noise in the growth loop is not supported");
        if (o.adhesion_mode != 2 && jkr_mode != BFM_JKR_EXACT && jkr_mode != BFM_JKR_TABLE_1D)
            io_fail(BFM_ERR_ARG, "bfm_run: jkr_mode must be BFM_JKR_EXACT or BFM_JKR_TABLE_1D");
        const bool adhesion = o.adhesion_mode != 0, friction = o.friction_mode != 0;
        // extra columns: gamma (21 with adhesion), then the friction
        // coefficient (simulation.hpp:1176-1188)
        const int ncols = kNcolsRequired + (adhesion ? 1 : 0) + (friction ? 1 : 0);
        const int col_gamma = adhesion ? 21 : -1;
        const int col_eta = friction ? (adhesion ? 22 : 21) : -1;

        // initial population (abmConstAdhesion.cpp:333-436)
        std::vector<double> cells;
        std::vector<int> parents;
        int n0 = 0;
        if (o.init_filename[0]) {
            Frame fr = read_frame(o.init_filename);
            if (fr.ncols != ncols)
                io_fail(BFM_ERR_ARG, "File specifying initial population contains incorrect number of columns");
            if (!o.lineage_filename[0])
                io_fail(BFM_ERR_ARG,
                        "File specifying initial population was specified, but without accompanying lineage file");
            n0 = (int)fr.rows.size();
            if (n0 < 1) io_fail(BFM_ERR_ARG, "the initial population is empty");
            cells.assign((size_t)n0 * ncols, 0.0);
            for (int i = 0; i < n0; ++i)
                for (int j = 0; j < (int)fr.rows[i].size(); ++j) cells[(size_t)j * n0 + i] = fr.rows[i][j];
            const std::vector<int> lin = read_lineage_file(o.lineage_filename);
            for (int i = 0; i < n0; ++i) {
                const int id = (int)cells[(size_t)kColId * n0 + i];
                if (id < 0 || id >= (int)lin.size())
                    io_fail(BFM_ERR_ARG, "Input lineage file does not specify parent of cell in initial population");
            }
            parents = lin;
        } else {
            n0 = 1;
            cells.assign(ncols, 0.0);
            const double rz = o.R - std::pow(o.sigma0 * std::sqrt(o.R) / (4 * o.E0), 2. / 3.);
            const double row[21] = {0, 0, 0, rz, 1, 0, 0, 0, 0, 0, 0, 0, 0, o.L0, o.L0 / 2, 0, o.growth_mean,
                                    o.eta_ambient, o.eta_surface, o.sigma0, 1};
            for (int j = 0; j < 21; ++j) cells[j] = row[j];
            if (friction) cells[col_eta] = o.eta_cell_cell;
            parents.push_back(-1);
        }

        // surface energy density from the target distance; group 1 -> max_gamma
        // (simulation.hpp:1354-1374), and the 1-D table of runSimulation (:1376-1405)
        double max_gamma = 0;
        std::vector<double> overlaps, radii;
        std::unique_ptr<AnisoTables> aniso;
        if (o.adhesion_mode == 2) {
            // fixed gamma (no switching): the reduced force table
            aniso = std::make_unique<AnisoTables>();
            parse_curvature_table(o.curvature_filename, *aniso);
            parse_force_table(o.jkr_forces_filename, false, *aniso);
            aniso_view(*aniso, false);
        }
        if (adhesion) {
            double resid = 0;
            const int grc = bfm_jkr_optimal_surface_energy_density(o.R, o.Rcell, o.E0, o.adhesion_eqdist, 100.0, 1e-6,
                                                                   1e-8, 1e-8, 1e-8, 1000, 1000, &max_gamma, &resid);
            if (grc != BFM_OK) io_fail(grc, "surface energy density for adhesion_eqdist could not be found");
            for (int i = 0; i < n0; ++i)
                cells[(size_t)col_gamma * n0 + i] = cells[(size_t)kColGroup * n0 + i] == 1 ? max_gamma : 0.0;
        }
        if (o.adhesion_mode == 1) {
            const int nm = o.n_mesh_overlap;
            if (nm < 2) io_fail(BFM_ERR_ARG, "adhesion_n_mesh_overlap must be >= 2");
            overlaps.resize(nm);
            radii.resize(nm);
            // LinSpaced(n, 0, hi) as the host table builder (bfm/jkr.py) forms it:
            // the low half upward from 0, the high half downward from hi
            const double hi = 2 * (o.R - o.Rcell), step = hi / (nm - 1);
            const int flip = (nm - 1) / 2;
            for (int k = 0; k < nm; ++k) {
                overlaps[k] = k < flip ? 0.0 + k * step : hi - (nm - 1 - k) * step;
                radii[k] = bfm::jkr_radius_newton(overlaps[k], o.R / 2, o.E0, max_gamma, 0.0, 400, nullptr);
            }
        }

        // the device run
        int cap = max_cells;
        if (cap <= 0) {
            if (o.n_cells <= 0) io_fail(BFM_ERR_ARG, "bfm_run: max_cells is required when n_cells is not given");
            cap = 2 * std::max(o.n_cells, n0) + 16;
        }
        if (cap < n0) io_fail(BFM_ERR_ARG, "bfm_run: max_cells is below the initial population");
        int rc2 = bfm_create(&ctx, device, cap, std::max<int64_t>(16 * (int64_t)cap, 1024));
        if (rc2 != BFM_OK) io_fail(rc2, "bfm_create failed");
        bfm_params prm;
        std::memset(&prm, 0, sizeof(prm));
        prm.R = o.R;
        prm.Rcell = o.Rcell;
        prm.E0 = o.E0;
        prm.repulsion_prefactors[0] = (4. / 3.) * o.E0 * std::sqrt(o.R / 2);
        prm.repulsion_prefactors[1] = (4. / 3.) * o.Ecell * std::sqrt(o.Rcell / 2);
        prm.repulsion_prefactors[2] = (4. / 3.) * o.E0 * std::sqrt(o.R / 2) * std::pow(2 * (o.R - o.Rcell), 1.5);
        prm.nz_threshold = o.nz_threshold;
        prm.cell_cell_coulomb_coeff = o.cell_cell_coulomb_coeff;
        prm.cell_surface_coulomb_coeff = o.cell_surface_coulomb_coeff;
        prm.no_surface = o.no_surface;
        prm.adhesion_mode = o.adhesion_mode == 2 ? BFM_ADHESION_JKR_ANISOTROPIC
                            : adhesion          ? BFM_ADHESION_JKR_ISOTROPIC
                                                : BFM_ADHESION_NONE;
        prm.jkr_mode = o.adhesion_mode == 2 ? BFM_JKR_EXACT : jkr_mode;
        prm.colidx_gamma = col_gamma;
        prm.friction_mode = friction ? 1 : 0;
        prm.ncols = ncols;
        prm.colidx_eta_cell_cell = col_eta;
        if ((rc2 = bfm_set_params(ctx, &prm)) != BFM_OK) io_fail(rc2, bfm_last_error(ctx));
        if (aniso && (rc2 = bfm_set_aniso_jkr_tables(ctx, &aniso->view)) != BFM_OK) io_fail(rc2, bfm_last_error(ctx));
        if (o.adhesion_mode == 1 && jkr_mode == BFM_JKR_TABLE_1D &&
            (rc2 = bfm_set_jkr_table(ctx, overlaps.data(), (int)overlaps.size(), nullptr, 0, radii.data())) != BFM_OK)
            io_fail(rc2, bfm_last_error(ctx));

        bfm_sim_config sc;
        std::memset(&sc, 0, sizeof(sc));
        sc.R = o.R;
        sc.Rcell = o.Rcell;
        sc.Ldiv = o.Ldiv;
        sc.max_stepsize = o.max_stepsize;
        sc.min_stepsize = o.min_stepsize;
        sc.max_error_allowed = o.max_error_allowed;
        sc.min_error = 1e-30;
        sc.neighbor_threshold = o.neighbor_threshold;
        sc.growth_mean[0] = sc.growth_mean[1] = o.growth_mean;
        sc.growth_std[0] = sc.growth_std[1] = o.growth_std;
        sc.daughter_length_std = o.daughter_length_std;
        sc.daughter_angle_xy_bound = o.daughter_angle_xy_bound;
        sc.daughter_angle_z_bound = o.daughter_angle_z_bound;
        sc.max_time = o.max_time;
        sc.iter_update_neighbors = o.iter_update_neighbors;
        sc.iter_update_stepsize = o.iter_update_stepsize;
        sc.max_tries_update_stepsize = 3;
        sc.n_cells = o.n_cells;
        sc.max_iter = o.max_iter;
        sc.tableau = o.integration_mode - 1;   // VELOCITY_VERLET = 0 -> -1, HEUN_EULER = 1 ... DORMAND_PRINCE = 4
        if (o.integration_mode == 0)
            sc.rho = o.M0 / (bfm::kPi * ((4. / 3.) * std::pow(o.R, 3) + o.R * o.R * o.L0));
        sc.seed = (uint64_t)(uint32_t)rng_seed;
        sc.report_seconds = verbose ? 30.0 : 0.0;

        std::map<std::string, std::string> params = frame_params(o, rng_seed, max_gamma);
        std::unordered_map<int, int> extra;
        if (adhesion) extra[col_gamma] = 1;
        if (friction) extra[col_eta] = 1;
        const std::string prefix = outprefix;
        params["t_curr"] = fmt_g(0.0);
        write_frame(prefix + "_init.txt", cells.data(), n0, n0, ncols, params, extra);

        bfm::FrameHook hook;
        hook.dt_write = o.dt_write;
        hook.fn = [&](const double* hc, int n, int ld, int64_t iter, double t, double dt, double max_err,
                      double wall_elapsed, const std::vector<int>& par) {
            if (verbose) {
                std::printf("Iteration %lld: %d cells, time = %g, time elapsed = %g sec, max error = %g, dt = %g\n",
                            (long long)iter, n, t, wall_elapsed, max_err, dt);
                std::fflush(stdout);
            }
            params["t_curr"] = fmt_g(t);
            write_frame(prefix + "_iter" + std::to_string(iter) + ".txt", hc, n, ld, ncols, params, extra);
            write_lineage_file(prefix + "_lineage.txt", par.data(), (int)par.size());
        };
        std::vector<double> out((size_t)cap * ncols);
        std::vector<int> par_out(parents.size() + 2 * (size_t)cap + 8);
        int n_out = 0, n_par = 0;
        bfm_sim_stats st;
        std::memset(&st, 0, sizeof(st));
        rc2 = bfm::simulate_with_frames(ctx, &sc, cells.data(), n0, ncols, parents.data(), (int)parents.size(),
                                        out.data(), cap, &n_out, par_out.data(), &n_par, &st, &hook);
        if (rc2 != BFM_OK) io_fail(rc2, bfm_last_error(ctx));
        params["t_curr"] = fmt_g(st.t_end);
        write_frame(prefix + "_final.txt", out.data(), n_out, cap, ncols, params, extra);
        write_lineage_file(prefix + "_lineage.txt", par_out.data(), n_par);
        if (stats_out) *stats_out = st;
    });
    if (ctx) bfm_destroy(ctx);
    return rc;
}

}  // extern "C"
