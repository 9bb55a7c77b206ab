// workloads.cpp -- seeded synthetic inputs of the benchmark configurations
// (host C++, libbfm_workloads.so). Not part of the mechanical-step ABI.
//
// bfm_make_dome: N spherocylinders in a hemispherical dome at a target
// packing fraction by seeded random sequential addition. Centers uniform in
// the dome (z >= 0), orientations uniform on the upper unit hemisphere,
// half-lengths uniform in [hl_lo, hl_hi]. A candidate is rejected if its core
// pierces the surface (lowest endpoint z < Rcell) or if its centerline comes
// within 2 Rcell of an accepted cell's centerline (EPS overlaps are allowed).
#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <random>
#include <vector>

namespace {

const double kPi = 3.14159265358979323846;

struct Rng {
    std::mt19937_64 g;
    explicit Rng(uint64_t seed) : g(seed) {}
    double u() { return (double)(g() >> 11) * (1.0 / 9007199254740992.0); }
};

struct Seg {
    double r[3], n[3], hl;
};

double clamp(double x, double lo, double hi) { return x < lo ? lo : (x > hi ? hi : x); }

// Squared distance between two centerline segments (clamped closed form).
double seg_dist2(const Seg& a, const Seg& b)
{
    double r12[3] = {b.r[0] - a.r[0], b.r[1] - a.r[1], b.r[2] - a.r[2]};
    double bb = a.n[0] * b.n[0] + a.n[1] * b.n[1] + a.n[2] * b.n[2];
    double e = a.n[0] * r12[0] + a.n[1] * r12[1] + a.n[2] * r12[2];
    double f = b.n[0] * r12[0] + b.n[1] * r12[1] + b.n[2] * r12[2];
    double det = 1 - bb * bb;
    double s, t;
    if (det > 1e-12) {
        s = clamp((e - bb * f) / det, -a.hl, a.hl);
    } else {
        s = 0;
    }
    t = bb * s - f;
    if (t < -b.hl || t > b.hl) {
        t = clamp(t, -b.hl, b.hl);
        s = clamp(e + bb * t, -a.hl, a.hl);
    }
    double d2 = 0;
    for (int k = 0; k < 3; ++k) {
        double dk = r12[k] + t * b.n[k] - s * a.n[k];
        d2 += dk * dk;
    }
    return d2;
}

}  // namespace

extern "C" {

// Fills cells (n x ncols column-major, leading dimension n). Returns the
// number of accepted cells (n on success) and the dome radius in *dome_out.
int bfm_make_dome(int n, int ncols, double R, double Rcell, double phi, double hl_lo, double hl_hi,
                  double gamma, double eta0, double eta1, double sigma0, double growth, uint64_t seed,
                  double* cells, double* dome_out)
{
    const double mean_hl = 0.5 * (hl_lo + hl_hi);
    const double mean_vol = kPi * R * R * (2 * mean_hl) + 4.0 / 3.0 * kPi * R * R * R;
    const double dome_vol = n * mean_vol / phi;
    const double Rd = std::cbrt(3 * dome_vol / (2 * kPi));
    if (dome_out) *dome_out = Rd;
    Rng rng(seed);
    const double excl = 2 * Rcell;
    const double edge = excl + 2 * hl_hi;
    const int dims = (int)std::ceil(2 * Rd / edge) + 1;
    const int dimz = (int)std::ceil(Rd / edge) + 1;
    std::vector<std::vector<int>> grid((size_t)dims * dims * dimz);
    std::vector<Seg> acc;
    acc.reserve(n);
    auto bin = [&](double x, double y, double z, int& bx, int& by, int& bz) {
        bx = std::min(dims - 1, std::max(0, (int)((x + Rd) / edge)));
        by = std::min(dims - 1, std::max(0, (int)((y + Rd) / edge)));
        bz = std::min(dimz - 1, std::max(0, (int)(z / edge)));
    };
    int64_t tries = 0, max_tries = (int64_t)n * 400 + 100000;
    while ((int)acc.size() < n && tries < max_tries) {
        ++tries;
        Seg c;
        double x = (2 * rng.u() - 1) * Rd, y = (2 * rng.u() - 1) * Rd, z = rng.u() * Rd;
        double nzv = rng.u(), ph = 2 * kPi * rng.u();
        double hl = hl_lo + (hl_hi - hl_lo) * rng.u();
        if (x * x + y * y + z * z > Rd * Rd) continue;
        double sxy = std::sqrt(std::max(0.0, 1 - nzv * nzv));
        c.r[0] = x; c.r[1] = y; c.r[2] = z;
        c.n[0] = sxy * std::cos(ph); c.n[1] = sxy * std::sin(ph); c.n[2] = nzv;
        c.hl = hl;
        if (z - hl * nzv < Rcell) continue;
        int bx, by, bz;
        bin(x, y, z, bx, by, bz);
        bool ok = true;
        for (int zz = std::max(0, bz - 1); ok && zz <= std::min(dimz - 1, bz + 1); ++zz)
            for (int yy = std::max(0, by - 1); ok && yy <= std::min(dims - 1, by + 1); ++yy)
                for (int xx = std::max(0, bx - 1); ok && xx <= std::min(dims - 1, bx + 1); ++xx)
                    for (int q : grid[((size_t)zz * dims + yy) * dims + xx])
                        if (seg_dist2(c, acc[q]) < excl * excl) { ok = false; break; }
        if (!ok) continue;
        grid[((size_t)bz * dims + by) * dims + bx].push_back((int)acc.size());
        acc.push_back(c);
    }
    const int m = (int)acc.size();
    for (int i = 0; i < m; ++i) {
        const Seg& c = acc[i];
        double row[32] = {0};
        row[0] = i;
        for (int k = 0; k < 3; ++k) { row[1 + k] = c.r[k]; row[4 + k] = c.n[k]; }
        row[13] = 2 * c.hl; row[14] = c.hl; row[15] = 0; row[16] = growth;
        row[17] = eta0; row[18] = eta1; row[19] = sigma0; row[20] = 1;
        if (ncols > 21) row[21] = gamma;
        for (int col = 0; col < ncols && col < 32; ++col) cells[(size_t)col * n + i] = row[col];
    }
    return m;
}

}  // extern "C"
