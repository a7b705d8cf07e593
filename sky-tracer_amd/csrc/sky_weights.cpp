// sky_weights.cpp -- quadrature weights of sparse band combos (host side of
// include/sky_scan.h), following crates/sky-wavelength-optimizer/src/main.rs:
//   band widths                 main.rs:714-733
//   Voronoi widths              main.rs:748-784
//   CMF least squares           main.rs:786-897 (rows), 899-955 (moments)
//   non-negative fixed sum      main.rs:957-1000 (subset enumeration)
//   KKT system per subset       main.rs:1002-1034
//   Gauss-Jordan elimination    main.rs:1036-1077
//   fit error                   main.rs:1079-1094
// All in FP64 with the reference's operation order, so the weights match its
// printed weights digit for digit.
#include <algorithm>
#include <cmath>
#include <string>
#include <thread>
#include <vector>

#include "../../include/sky_scan.h"
#include "sky_internal.hpp"

namespace {

constexpr int kMaxP = SKY_SCAN_MAX_PRIMARIES;
constexpr int kMaxRows = 9;
constexpr int kKkt = kMaxP + 1;

struct Grid {
    const float *wl;
    uint32_t n;
    std::vector<double> width;
    double lo, hi;  // outer edges of the first and last bands
};

void widths(const float *wl, uint32_t n, double *w)
{
    for (uint32_t i = 0; i < n; ++i) {
        // differences and midpoints in f32, as the reference forms them
        double lo = i == 0 ? (double)wl[i] - 0.5 * (double)(wl[i + 1] - wl[i]) : 0.5 * (double)(wl[i - 1] + wl[i]);
        double hi = i + 1 == n ? (double)wl[i] + 0.5 * (double)(wl[i] - wl[i - 1]) : 0.5 * (double)(wl[i] + wl[i + 1]);
        w[i] = hi - lo;
    }
}

void voronoi(const Grid &g, const uint8_t *combo, uint32_t count, double *out)
{
    for (int k = 0; k < kMaxP; ++k) out[k] = 0.0;
    if (count == 1) {
        out[0] = g.hi - g.lo;
        return;
    }
    double split[kMaxP];
    for (uint32_t k = 0; k + 1 < count; ++k) split[k] = 0.5 * ((double)g.wl[combo[k]] + (double)g.wl[combo[k + 1]]);
    out[0] = split[0] - g.lo;
    for (uint32_t k = 1; k + 1 < count; ++k) out[k] = split[k] - split[k - 1];
    out[count - 1] = g.hi - split[count - 2];
}

// rows of one wavelength: basis (1, m, m^2) x (x, y, z), m in [-1, 1] over the grid
void cmf_rows(const Grid &g, int row_count, double lambda, const double *cmf, double *rows)
{
    const double center = 0.5 * (g.lo + g.hi);
    const double half = 0.5 * (g.hi - g.lo);
    const double m = (lambda - center) / std::max(half, 1.0e-6);
    const double basis[3] = {1.0, m, m * m};
    for (int b = 0; b < row_count / 3; ++b)
        for (int c = 0; c < 3; ++c) rows[b * 3 + c] = basis[b] * cmf[c];
}

bool solve(double a[kKkt][kKkt], double *rhs, int dim)
{
    for (int p = 0; p < dim; ++p) {
        int best = p;
        double best_v = std::fabs(a[p][p]);
        for (int r = p + 1; r < dim; ++r)
            if (std::fabs(a[r][p]) > best_v) {
                best_v = std::fabs(a[r][p]);
                best = r;
            }
        if (best_v <= 1.0e-12) return false;
        if (best != p) {
            for (int c = 0; c < kKkt; ++c) std::swap(a[p][c], a[best][c]);
            std::swap(rhs[p], rhs[best]);
        }
        const double pv = a[p][p];
        for (int c = p; c < dim; ++c) a[p][c] /= pv;
        rhs[p] /= pv;
        for (int r = 0; r < dim; ++r) {
            if (r == p) continue;
            const double f = a[r][p];
            if (std::fabs(f) <= 1.0e-18) continue;
            for (int c = p; c < dim; ++c) a[r][c] -= f * a[p][c];
            rhs[r] -= f * rhs[p];
        }
    }
    return true;
}

bool cmf_fit(const Grid &g, const double *cmf, const uint8_t *combo, uint32_t count, bool moments, double *out)
{
    const int rows = moments ? 9 : 3;
    const double total = g.hi - g.lo;
    double col[kMaxRows][kMaxP] = {};
    for (uint32_t k = 0; k < count; ++k) {
        double r[kMaxRows];
        cmf_rows(g, rows, (double)g.wl[combo[k]], cmf + 3 * combo[k], r);
        for (int i = 0; i < rows; ++i) col[i][k] = r[i];
    }
    double target[kMaxRows] = {};
    for (uint32_t b = 0; b < g.n; ++b) {
        double r[kMaxRows];
        cmf_rows(g, rows, (double)g.wl[b], cmf + 3 * b, r);
        for (int i = 0; i < rows; ++i) target[i] += g.width[b] * r[i];
    }
    bool found = false;
    double best_err = INFINITY;
    for (uint32_t mask = 1; mask < (1u << count); ++mask) {
        int sub[kMaxP], ns = 0;
        for (uint32_t k = 0; k < count; ++k)
            if (mask & (1u << k)) sub[ns++] = (int)k;
        double a[kKkt][kKkt] = {}, rhs[kKkt] = {};
        for (int i = 0; i < ns; ++i) {
            for (int j = 0; j < ns; ++j) {
                double s = 0.0;
                for (int r = 0; r < rows; ++r) s += col[r][sub[i]] * col[r][sub[j]];
                a[i][j] = s;
            }
            a[i][ns] = 1.0;
            a[ns][i] = 1.0;
            double s = 0.0;
            for (int r = 0; r < rows; ++r) s += col[r][sub[i]] * target[r];
            rhs[i] = s;
        }
        rhs[ns] = total;
        if (!solve(a, rhs, ns + 1)) continue;
        double w[kMaxP] = {};
        for (int i = 0; i < ns; ++i) w[sub[i]] = rhs[i];
        bool negative = false;
        for (uint32_t k = 0; k < count; ++k) negative = negative || w[k] < -1.0e-7 * total;
        if (negative) continue;
        for (uint32_t k = 0; k < count; ++k)
            if (w[k] < 0.0) w[k] = 0.0;
        double err = 0.0;
        for (int r = 0; r < rows; ++r) {
            double v = 0.0;
            for (int k = 0; k < kMaxP; ++k) v += col[r][k] * w[k];
            const double d = v - target[r];
            err += d * d;
        }
        if (err < best_err) {
            best_err = err;
            for (int k = 0; k < kMaxP; ++k) out[k] = w[k];
            found = true;
        }
    }
    return found;
}

}  // namespace

extern "C" {

sky_status sky_band_widths(const float *wavelengths_nm, uint32_t n_bands, double *widths_nm)
{
    if (!wavelengths_nm || !widths_nm) return sky::set_error(SKY_ERR_INVALID_CONFIG, "null argument");
    if (n_bands < 2) return sky::set_error(SKY_ERR_INVALID_CONFIG, "band widths need at least two bands");
    widths(wavelengths_nm, n_bands, widths_nm);
    return SKY_OK;
}

sky_status sky_quadrature_weights(const float *wavelengths_nm, uint32_t n_bands, const double *cmf_xyz,
                                  const uint8_t *combos, uint32_t count, uint64_t n_combos, uint32_t strategy,
                                  uint32_t threads, double *weights)
{
    if (!wavelengths_nm || !combos || !weights) return sky::set_error(SKY_ERR_INVALID_CONFIG, "null argument");
    if (n_bands < 2 || n_bands > SKY_SCAN_MAX_BANDS)
        return sky::set_error(SKY_ERR_INVALID_CONFIG, "band count must be in 2..=64");
    if (count < 1 || count > SKY_SCAN_MAX_PRIMARIES)
        return sky::set_error(SKY_ERR_INVALID_CONFIG, "combo size must be in 1..=8");
    if (strategy > SKY_WEIGHTS_CMF_MOMENTS) return sky::set_error(SKY_ERR_INVALID_CONFIG, "unknown weight strategy");
    if (strategy != SKY_WEIGHTS_VORONOI && !cmf_xyz)
        return sky::set_error(SKY_ERR_INVALID_CONFIG, "CMF fit weights need the colour-matching functions");
    for (uint64_t i = 0; i < n_combos * count; ++i)
        if (combos[i] >= n_bands) return sky::set_error(SKY_ERR_INVALID_CONFIG, "combo band index out of range");
    Grid g{wavelengths_nm, n_bands, std::vector<double>(n_bands), 0.0, 0.0};
    widths(wavelengths_nm, n_bands, g.width.data());
    g.lo = (double)wavelengths_nm[0] - 0.5 * g.width[0];
    g.hi = (double)wavelengths_nm[n_bands - 1] + 0.5 * g.width[n_bands - 1];
    auto work = [&](uint64_t first, uint64_t last) {
        for (uint64_t c = first; c < last; ++c) {
            const uint8_t *combo = combos + c * count;
            double w[kMaxP];
            if (strategy == SKY_WEIGHTS_VORONOI ||
                !cmf_fit(g, cmf_xyz, combo, count, strategy == SKY_WEIGHTS_CMF_MOMENTS, w))
                voronoi(g, combo, count, w);
            for (uint32_t k = 0; k < count; ++k) weights[c * count + k] = w[k];
        }
    };
    uint32_t nt = threads ? threads : std::max(1u, std::thread::hardware_concurrency());
    nt = (uint32_t)std::min<uint64_t>(nt, std::max<uint64_t>(1, n_combos / 256));
    if (nt <= 1) {
        work(0, n_combos);
        return SKY_OK;
    }
    std::vector<std::thread> pool;
    const uint64_t per = (n_combos + nt - 1) / nt;
    for (uint32_t t = 0; t < nt; ++t) {
        const uint64_t a = t * per, b = std::min<uint64_t>(n_combos, a + per);
        if (a < b) pool.emplace_back(work, a, b);
    }
    for (auto &th : pool) th.join();
    return SKY_OK;
}

}  // extern "C"
