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
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <functional>
#include <map>
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

int bfm_run(const char* json_path, const char* outprefix, int rng_seed, int device, int jkr_mode, int max_cells,
            int verbose, bfm_sim_stats* stats_out)
{
    bfm_ctx* ctx = nullptr;
    const int rc = guarded([&] {
        if (!json_path || !outprefix) io_fail(BFM_ERR_ARG, "bfm_run: need the config path and the output prefix");
        bfm_run_config o;
        load_config(json_path, &o);
        if (o.adhesion_mode == 2) io_fail(BFM_ERR_UNSUPPORTED, "anisotropic JKR adhesion is not supported");
        if (o.integration_mode == 0 && !(o.M0 > 0))
            io_fail(BFM_ERR_ARG, "Initial cell mass must be defined when using velocity Verlet");
        if (o.basal_only) io_fail(BFM_ERR_UNSUPPORTED, "basal_only is not supported");
        if (o.max_rxy_noise > 0 || o.max_rz_noise > 0 || o.max_nxy_noise > 0 || o.max_nz_noise > 0)
            io_fail(BFM_ERR_UNSUPPORTED, "noise in the growth loop is not supported");
        if (jkr_mode != BFM_JKR_EXACT && jkr_mode != BFM_JKR_TABLE_1D)
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
        if (adhesion) {
            double resid = 0;
            const int grc = bfm_jkr_optimal_surface_energy_density(o.R, o.Rcell, o.E0, o.adhesion_eqdist, 100.0, 1e-6,
                                                                   1e-8, 1e-8, 1e-8, 1000, 1000, &max_gamma, &resid);
            if (grc != BFM_OK) io_fail(grc, "surface energy density for adhesion_eqdist could not be found");
            for (int i = 0; i < n0; ++i)
                cells[(size_t)col_gamma * n0 + i] = cells[(size_t)kColGroup * n0 + i] == 1 ? max_gamma : 0.0;
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
        prm.adhesion_mode = adhesion ? BFM_ADHESION_JKR_ISOTROPIC : BFM_ADHESION_NONE;
        prm.jkr_mode = jkr_mode;
        prm.colidx_gamma = col_gamma;
        prm.friction_mode = friction ? 1 : 0;
        prm.ncols = ncols;
        prm.colidx_eta_cell_cell = col_eta;
        if ((rc2 = bfm_set_params(ctx, &prm)) != BFM_OK) io_fail(rc2, bfm_last_error(ctx));
        if (adhesion && jkr_mode == BFM_JKR_TABLE_1D &&
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
