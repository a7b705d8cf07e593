// sky_scene.cpp -- host-side scene build for the MI355X integrator.
//
// Restates the scene assembly of crates/sky-core/src/data.rs:15-51 and the
// GPU packing of crates/sky-spectral-path-tracer/src/integrator.rs:319-418 for
// the analytic atmosphere: per-wavelength coefficients, the Planck sun, the
// CIE weights, the per-shell majorant grid (medium.rs:63-83) and the constant
// block the kernels read.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>

#include "sky_internal.hpp"

namespace sky {

namespace {

constexpr double kPiD = 3.14159265358979323846;

// One lobe of the Wyman-Sloan-Shirley multi-lobe fit of the CIE 1931 2-degree
// colour-matching functions (JCGT 2(2), 2013).  Replaces the table lerp of
// spectrum.rs:196-219.
double wss_lobe(double l, double mu, double s_lo, double s_hi)
{
    double t = (l - mu) / (l < mu ? s_lo : s_hi);
    return std::exp(-0.5 * t * t);
}

void cmf_xyz(double l, double out[3])
{
    out[0] = 1.056 * wss_lobe(l, 599.8, 37.9, 31.0) + 0.362 * wss_lobe(l, 442.0, 16.0, 26.7) -
             0.065 * wss_lobe(l, 501.1, 20.4, 26.2);
    out[1] = 0.821 * wss_lobe(l, 568.8, 46.9, 40.5) + 0.286 * wss_lobe(l, 530.9, 16.3, 31.1);
    out[2] = 1.217 * wss_lobe(l, 437.0, 11.8, 36.0) + 0.681 * wss_lobe(l, 459.0, 26.0, 13.8);
}

// Planck B_lambda(T), W m^-2 sr^-1 nm^-1 (replaces E_band / Omega_sun of
// data.rs:30-33).
double planck(double lambda_nm, double temperature_k)
{
    const double h = 6.62607015e-34, c = 299792458.0, k = 1.380649e-23;
    double l = lambda_nm * 1e-9;
    return 2.0 * h * c * c / (l * l * l * l * l) / std::expm1(h * c / (l * k * temperature_k)) * 1e-9;
}

double default_weight(const sky_wavelengths *wl)
{
    if (wl->count < 2) return 1.0;
    return (wl->center_nm[wl->count - 1] - wl->center_nm[0]) / (wl->count - 1);
}

}  // namespace

// km^-1 coefficients at the profiles' reference heights: Rayleigh
// beta (550/l)^4, ozone Gaussian band, aerosol Angstrom law (replaces the
// cross-section tables of data.rs:26-29 and medium.rs:104-106).
static void coeffs_d(const sky_atmosphere *a, double l, double c[4])
{
    double r = 550.0 / l;
    double dl = (l - a->ozone_lambda0_nm) / a->ozone_sigma_nm;
    double ang = std::pow(l / 550.0, -a->angstrom);
    c[0] = a->rayleigh_sca_550_m_inv * 1000.0 * r * r * r * r;
    c[1] = a->ozone_peak_m_inv * 1000.0 * std::exp(-0.5 * dl * dl);
    c[2] = a->aerosol_sca_550_m_inv * 1000.0 * ang;
    c[3] = a->aerosol_abs_550_m_inv * 1000.0 * ang;
}

// bracket + lerp of one profile column at altitude x (medium.rs:108-155):
// clamped to the end points, t = (x - z_lo) / (z_hi - z_lo).
static double profile_at(const double *rows, uint32_t n, uint32_t stride, uint32_t col, double x)
{
    uint32_t hi = 0;
    while (hi < n && rows[(size_t)hi * stride] < x) ++hi;  // partition_point(z < x)
    if (hi == 0) return rows[col];
    if (hi >= n) return rows[(size_t)(n - 1) * stride + col];
    const double *lo = rows + (size_t)(hi - 1) * stride, *up = rows + (size_t)hi * stride;
    double t = up[0] > lo[0] ? (x - lo[0]) / (up[0] - lo[0]) : 0.0;
    t = std::fmin(std::fmax(t, 0.0), 1.0);
    return lo[col] * (1.0 - t) + up[col] * t;
}

void table_coeffs(const sky_medium_tables *t, uint32_t b, double lambda_nm, double h, double out[4])
{
    const double air = profile_at(t->atmosphere, t->atmosphere_points, 3, 1, h);
    const double o3 = profile_at(t->atmosphere, t->atmosphere_points, 3, 2, h);
    const double r = 550.0 / lambda_nm;
    const double xs_r = t->rayleigh_cross_section_m2 ? t->rayleigh_cross_section_m2[b] : 5.8e-31 * r * r * r * r;
    out[0] = std::fmax(air * xs_r * 1.0e9, 0.0);
    out[2] = std::fmax(o3 * t->ozone_cross_section_cm2[b] * 1.0e5, 0.0);
    out[1] = out[3] = 0.0;
    for (int sp = 0; sp < SKY_AEROSOL_SPECIES; ++sp) {
        const double m = std::fmax(profile_at(t->aerosol, t->aerosol_points, 5, 1 + sp, h), 0.0);
        out[1] += m * t->aerosol_optics[((size_t)b * SKY_AEROSOL_SPECIES + sp) * 2];
        out[3] += m * t->aerosol_optics[((size_t)b * SKY_AEROSOL_SPECIES + sp) * 2 + 1];
    }
}

// Merged breakpoint grid: 0, top and every profile altitude in between.
static std::vector<double> table_breakpoints(const sky_medium_tables *t, double top)
{
    std::vector<double> z{0.0, top};
    for (uint32_t i = 0; i < t->atmosphere_points; ++i) z.push_back(t->atmosphere[3 * (size_t)i]);
    for (uint32_t i = 0; i < t->aerosol_points; ++i) z.push_back(t->aerosol[5 * (size_t)i]);
    std::sort(z.begin(), z.end());
    std::vector<double> out;
    for (double v : z)
        if (v >= 0.0 && v <= top && (out.empty() || v > out.back() + 1e-9)) out.push_back(v);
    if (out.back() < top) out.push_back(top);
    return out;
}

// Interval of h among the FP32 breakpoints: the number of z[1..K-2] <= h,
// which is what tab_interval's index-and-walk search returns.
static uint32_t interval_by_count(const std::vector<float> &z, float h)
{
    return (uint32_t)(std::upper_bound(z.begin() + 1, z.end() - 1, h) - (z.begin() + 1));
}

// The kernel's one-load lookup (tab_medium) on the host.
static uint32_t interval_by_cell(const MediumTable &m, float h, float &z0, float &inv)
{
    const uint32_t n = (uint32_t)(m.cell.size() / 8);
    uint32_t j = (uint32_t)(h * m.cell_inv_dh);
    j = j < n ? j : n - 1;
    const float *c = &m.cell[(size_t)j * 8];
    const bool ua = h >= c[3], ub = h >= c[5];
    uint32_t k0;
    std::memcpy(&k0, c, 4);
    z0 = ub ? c[5] : (ua ? c[3] : c[1]);
    inv = ub ? c[6] : (ua ? c[4] : c[2]);
    return k0 + (ua ? 1u : 0u) + (ub ? 1u : 0u);
}

// Records for the kernel's one-load interval lookup.  An altitude h in
// [0, top] goes to cell j = trunc(fl(h * fl(n / top))), clamped to n - 1.  The
// two roundings put h * n / top within 2^-23 n (1.01) < 0.008 cells of the
// computed product for n <= 65536, so every h sent to cell j lies strictly
// within 0.01 cells of [j, j + 1) (all of [0, top] for the clamped last
// cell).  Breakpoints at or below that range's start are counted into k0; the
// ones inside it (at most two, else the next finer grid is tried) become
// za <= zb, unused ones +inf.  The records are then checked against the
// counting search at every breakpoint and cell boundary and their FP32
// neighbours; with no grid of up to 65536 cells passing, the table keeps the
// index-and-walk search alone.
static void build_cells(MediumTable &m, double top)
{
    const std::vector<float> &z = m.z;
    const size_t K = z.size();
    m.cell.clear();
    if (K < 2) return;
    for (uint32_t n = 1024; n <= 65536; n *= 2) {
        const double dh = top / n;
        std::vector<float> rec((size_t)n * 8, 0.0f);
        bool fits = true;
        uint32_t k0 = 0;  // #{z[1..K-2] <= lo}, non-decreasing in j
        for (uint32_t j = 0; j < n && fits; ++j) {
            const double lo = j == 0 ? -INFINITY : (j - 0.01) * dh;
            const double hi = j + 1 == n ? INFINITY : (j + 1.01) * dh;
            while (k0 + 2 < K && (double)z[k0 + 1] <= lo) ++k0;
            uint32_t ns = 0;
            while (k0 + ns + 2 < K && (double)z[k0 + ns + 1] <= hi) ++ns;
            if (ns > 2) {
                fits = false;
                break;
            }
            float *c = &rec[(size_t)j * 8];
            std::memcpy(c, &k0, 4);
            c[1] = z[k0];
            c[2] = m.inv[k0];
            c[3] = ns > 0 ? z[k0 + 1] : INFINITY;
            c[4] = ns > 0 ? m.inv[k0 + 1] : 0.0f;
            c[5] = ns > 1 ? z[k0 + 2] : INFINITY;
            c[6] = ns > 1 ? m.inv[k0 + 2] : 0.0f;
        }
        if (!fits) continue;
        m.cell.swap(rec);
        m.cell_inv_dh = (float)(n / top);
        const float topf = (float)top;
        std::vector<float> probe{0.0f, topf};
        for (size_t i = 0; i < K; ++i) probe.push_back(z[i]);
        for (uint32_t j = 0; j <= n; ++j) probe.push_back((float)(j * dh));
        bool ok = true;
        for (float p : probe) {
            float up = p, dn = p;
            for (int s = 0; s <= 4 && ok; ++s) {
                for (float q : {up, dn}) {
                    if (!(q >= 0.0f && q <= topf)) continue;
                    float z0, inv;
                    const uint32_t k = interval_by_cell(m, q, z0, inv);
                    const uint32_t kr = interval_by_count(z, q);
                    if (k != kr || z0 != z[kr] || inv != m.inv[kr]) ok = false;
                }
                up = std::nextafter(up, INFINITY);
                dn = std::nextafter(dn, -INFINITY);
            }
            if (!ok) break;
        }
        if (!ok) m.cell.clear();
        return;
    }
}

static void build_medium_table(const sky_atmosphere *a, const sky_wavelengths *wl, double top, MediumTable &m)
{
    const sky_medium_tables *t = a->tables;
    const std::vector<double> z = table_breakpoints(t, top);
    const size_t K = z.size();
    m.z.resize(K);
    m.inv.assign(K, 0.0f);
    for (size_t k = 0; k < K; ++k) {
        m.z[k] = (float)z[k];
        if (k + 1 < K) m.inv[k] = (float)(1.0 / (z[k + 1] - z[k]));
    }
    m.sig.resize((size_t)wl->count * K * 4);
    for (uint32_t b = 0; b < wl->count; ++b)
        for (size_t k = 0; k < K; ++k) {
            double c[4];
            table_coeffs(t, b, wl->center_nm[b], z[k], c);
            for (int i = 0; i < 4; ++i) m.sig[((size_t)b * K + k) * 4 + i] = (float)c[i];
        }
    const uint32_t nidx = 1024;
    m.idx.resize(nidx);
    m.inv_dh = (float)(nidx / top);
    size_t k = 0;
    for (uint32_t j = 0; j < nidx; ++j) {
        const double x = j * (top / nidx);
        while (k + 2 < K && z[k + 1] <= x) ++k;
        m.idx[j] = (uint16_t)k;
    }
    build_cells(m, top);
}

void MediumTable::pack(std::vector<uint8_t> &blob) const
{
    blob.assign(bytes(), 0);
    std::memcpy(blob.data(), sig.data(), sig.size() * 4);
    std::memcpy(blob.data() + off_z(), z.data(), z.size() * 4);
    std::memcpy(blob.data() + off_inv(), inv.data(), inv.size() * 4);
    std::memcpy(blob.data() + off_idx(), idx.data(), idx.size() * 2);
    if (!cell.empty()) std::memcpy(blob.data() + off_cell(), cell.data(), cell.size() * 4);
}

void MediumTable::attach(DevScene &s, const void *dev) const
{
    if (empty()) {
        s.tab = nullptr;
        return;
    }
    const uint8_t *p = (const uint8_t *)dev;
    s.tab = (const float4 *)p;
    s.tab_z = (const float *)(p + off_z());
    s.tab_inv = (const float *)(p + off_inv());
    s.tab_idx = (const uint16_t *)(p + off_idx());
    s.tab_k = (uint32_t)z.size();
    s.tab_nidx = (uint32_t)idx.size();
    s.tab_inv_dh = inv_dh;
    s.tab_cell = cell.empty() ? nullptr : (const float4 *)(p + off_cell());
    s.tab_ncell = (uint32_t)(cell.size() / 8);
    s.tab_cell_inv_dh = cell_inv_dh;
}

bool validate_tables(const sky_atmosphere *a, const sky_wavelengths *wl, std::string &err)
{
    const sky_medium_tables *t = a->tables;
    if (!t->atmosphere || !t->aerosol || !t->aerosol_optics || !t->ozone_cross_section_cm2 ||
        t->atmosphere_points == 0 || t->aerosol_points == 0 || t->atmosphere_points > 4096 ||
        t->aerosol_points > 4096) {
        err = "medium tables need 1..4096 profile points, optics and ozone cross sections";
        return false;
    }
    auto check_profile = [&](const double *rows, uint32_t n, uint32_t stride) {
        for (uint32_t i = 0; i < n; ++i)
            for (uint32_t c = 0; c < stride; ++c) {
                const double v = rows[(size_t)i * stride + c];
                if (!std::isfinite(v) || (c > 0 && v < 0.0)) return false;
                if (c == 0 && i > 0 && v < rows[(size_t)(i - 1) * stride]) return false;
            }
        return true;
    };
    if (!check_profile(t->atmosphere, t->atmosphere_points, 3) || !check_profile(t->aerosol, t->aerosol_points, 5)) {
        err = "profiles must be finite, non-negative and in ascending altitude";
        return false;
    }
    for (uint32_t b = 0; b < wl->count; ++b) {
        bool ok = std::isfinite(t->ozone_cross_section_cm2[b]) && t->ozone_cross_section_cm2[b] >= 0.0;
        if (t->rayleigh_cross_section_m2)
            ok = ok && std::isfinite(t->rayleigh_cross_section_m2[b]) && t->rayleigh_cross_section_m2[b] >= 0.0;
        if (t->solar_irradiance_w_m2)
            ok = ok && std::isfinite(t->solar_irradiance_w_m2[b]) && t->solar_irradiance_w_m2[b] >= 0.0;
        for (int i = 0; i < 2 * SKY_AEROSOL_SPECIES; ++i) {
            const double v = t->aerosol_optics[(size_t)b * 2 * SKY_AEROSOL_SPECIES + i];
            ok = ok && std::isfinite(v) && v >= 0.0;
        }
        if (!ok) {
            err = "cross sections, optics and irradiance must be finite and non-negative";
            return false;
        }
    }
    return true;
}

void fill_wavelength_rows(const sky_atmosphere *a, const sky_wavelengths *wl,
                          std::vector<DevLambda> &rows)
{
    rows.resize(wl->count);
    double dw = default_weight(wl);
    for (uint32_t i = 0; i < wl->count; ++i) {
        double l = wl->center_nm[i];
        double w = wl->weight_nm ? wl->weight_nm[i] : dw;
        double cmf[3], c[4];
        cmf_xyz(l, cmf);
        DevLambda &d = rows[i];
        if (a->tables) {  // sea-level coefficients, for inspection; the kernels read the table
            double t4[4];
            table_coeffs(a->tables, i, l, 0.0, t4);
            c[0] = t4[0];
            c[1] = t4[2];
            c[2] = t4[1];
            c[3] = t4[3];
        } else {
            coeffs_d(a, l, c);
        }
        d.br = (float)c[0];
        d.bo = (float)c[1];
        d.bas = (float)c[2];
        d.baa = (float)c[3];
        if (a->tables && a->tables->solar_irradiance_w_m2) {
            // band irradiance over the sun's solid angle 2 pi (1 - cos r) (data.rs:30-33)
            const double hs = std::sin(0.5 * a->sun_angular_radius_rad);
            d.lsun = (float)(a->tables->solar_irradiance_w_m2[i] / (4.0 * kPiD * hs * hs));
        } else {
            d.lsun = (float)planck(l, a->sun_temperature_k);
        }
        d.wx = (float)(cmf[0] * w);
        d.wy = (float)(cmf[1] * w);
        d.wz = (float)(cmf[2] * w);
    }
}

// margin * max over shell [k dz, (k+1) dz] of sigma_t, floored at 1e-8
// (medium.rs:77, sky_trace.wgsl:260).  Exact for these profiles: the
// exponentials peak at the shell's lower edge, the tent at the point of the
// shell nearest its centre.
// Lower boundary of shell k (k = 0..L, h_L = top): the reference's uniform
// shells (medium.rs:63-83) for growth 1, else thicknesses growing by the
// factor g from the ground up: h_k = top (g^k - 1) / (g^L - 1).
double shell_height(double top, uint32_t layers, double growth, uint32_t k)
{
    if (k == 0) return 0.0;
    if (k >= layers) return top;
    if (!(growth > 1.0)) return k * (top / layers);
    return top * (std::pow(growth, (double)k) - 1.0) / (std::pow(growth, (double)layers) - 1.0);
}

// Majorant grid (medium.rs:63-83): margin x the exact maximum of sigma_t
// over each shell (both exponentials peak at the shell's lower edge, the
// ozone tent at the point nearest its centre), floored at 1e-8.
void fill_majorants(const sky_planet *planet, const sky_atmosphere *a, const sky_wavelengths *wl,
                    uint32_t layers, double margin, double growth, std::vector<float> &out)
{
    double top = planet->atmosphere_radius_km - planet->ground_radius_km;
    out.resize((size_t)wl->count * layers);
    if (a->tables) {
        // compute_majorant_grid (medium.rs:63-102): the maximum over the shell's
        // ends, its middle and every profile altitude inside, which is exact
        // for coefficients linear between profile altitudes.
        const std::vector<double> z = table_breakpoints(a->tables, top);
        for (uint32_t b = 0; b < wl->count; ++b)
            for (uint32_t k = 0; k < layers; ++k) {
                const double lo = shell_height(top, layers, growth, k), hi = shell_height(top, layers, growth, k + 1);
                std::vector<double> probes{lo, hi, 0.5 * (lo + hi)};
                for (double v : z)
                    if (v > lo && v < hi) probes.push_back(v);
                double m = 0.0;
                for (double h : probes) {
                    double c[4];
                    table_coeffs(a->tables, b, wl->center_nm[b], h, c);
                    m = std::fmax(m, c[0] + c[1] + c[2] + c[3]);
                }
                out[(size_t)b * layers + k] = (float)std::fmax(m * margin, 1e-8);
            }
        return;
    }
    for (size_t b = 0; b < wl->count; ++b) {
        double c[4];
        coeffs_d(a, wl->center_nm[b], c);
        for (uint32_t k = 0; k < layers; ++k) {
            double lo = shell_height(top, layers, growth, k), hi = shell_height(top, layers, growth, k + 1);
            double d = 0.0;
            if (a->ozone_center_km < lo) d = lo - a->ozone_center_km;
            if (a->ozone_center_km > hi) d = a->ozone_center_km - hi;
            double tent = std::fmax(1.0 - d / a->ozone_half_width_km, 0.0);
            double m = c[0] * std::exp(-lo / a->rayleigh_h_km) +
                       (c[2] + c[3]) * std::exp(-lo / a->aerosol_h_km) + c[1] * tent;
            m *= margin;
            out[b * layers + k] = (float)std::fmax(m, 1e-8);
        }
    }
}

static bool finite_pos(double v) { return std::isfinite(v) && v > 0.0; }

sky_status build_scene(const sky_planet *planet, const sky_atmosphere *a, const sky_wavelengths *wl,
                       const sky_render_config *cfg, HostScene &out, std::string &err)
{
    // validate_config, integrator.rs:130-163 (bands generalised to 1..64)
    if (!planet || !a || !wl || !cfg) {
        err = "null scene or config pointer";
        return SKY_ERR_INVALID_CONFIG;
    }
    if (wl->count == 0 || wl->count > (uint32_t)kMaxLambda || !wl->center_nm) {
        err = "expected 1..64 wavelengths, found " + std::to_string(wl->count);
        return SKY_ERR_INVALID_CONFIG;
    }
    for (uint32_t i = 0; i < wl->count; ++i) {
        if (!finite_pos(wl->center_nm[i]) || (wl->weight_nm && !std::isfinite(wl->weight_nm[i]))) {
            err = "wavelengths and weights must be finite and positive";
            return SKY_ERR_INVALID_CONFIG;
        }
    }
    if (cfg->width == 0 || cfg->height == 0 || cfg->spp == 0) {
        err = "width, height, and spp must be greater than zero";
        return SKY_ERR_INVALID_CONFIG;
    }
    uint64_t cells = (uint64_t)cfg->width * cfg->height * wl->count;
    if (cells > 0xFFFFFFFFull / 2) {
        err = "film is too large for the GPU shader index type";
        return SKY_ERR_INVALID_CONFIG;
    }
    if ((uint64_t)cfg->sample_offset + cfg->spp > 0xFFFFFFFFull) {
        err = "sample range overflows the 32-bit sample key";
        return SKY_ERR_INVALID_CONFIG;
    }
    if (cfg->majorant_layers == 0 || cfg->majorant_layers > (uint32_t)kMaxLayers) {
        err = "majorant grid must contain 1..256 layers";
        return SKY_ERR_INVALID_CONFIG;
    }
    if (!finite_pos(planet->ground_radius_km) ||
        !(planet->atmosphere_radius_km > planet->ground_radius_km) ||
        !std::isfinite(planet->atmosphere_radius_km)) {
        err = "planet radii must satisfy 0 < ground < atmosphere";
        return SKY_ERR_INVALID_CONFIG;
    }
    if (!finite_pos(a->rayleigh_h_km) || !finite_pos(a->aerosol_h_km) ||
        !finite_pos(a->ozone_half_width_km) || !finite_pos(a->ozone_sigma_nm) ||
        !finite_pos(a->sun_temperature_k) || !finite_pos(a->sun_angular_radius_rad) ||
        !(a->sun_angular_radius_rad < 0.5 * kPiD)) {
        err = "scale heights, widths, sun temperature and radius must be positive";
        return SKY_ERR_INVALID_CONFIG;
    }
    if (!(a->aerosol_g > -1.0 && a->aerosol_g < 1.0)) {
        err = "aerosol asymmetry g must lie in (-1, 1)";
        return SKY_ERR_INVALID_CONFIG;
    }
    if (!(a->ground_albedo >= 0.0 && a->ground_albedo <= 1.0)) {
        err = "ground albedo must lie in [0, 1]";
        return SKY_ERR_INVALID_CONFIG;
    }
    if (!(a->rayleigh_sca_550_m_inv >= 0.0) || !(a->ozone_peak_m_inv >= 0.0) ||
        !(a->aerosol_sca_550_m_inv >= 0.0) || !(a->aerosol_abs_550_m_inv >= 0.0) ||
        !std::isfinite(a->angstrom) || !std::isfinite(a->ozone_center_km)) {
        err = "medium coefficients must be finite and non-negative";
        return SKY_ERR_INVALID_CONFIG;
    }
    if (!(cfg->majorant_margin >= 1.0) || !std::isfinite(cfg->majorant_margin)) {
        err = "majorant margin must be >= 1";
        return SKY_ERR_INVALID_CONFIG;
    }
    if (cfg->projection > 1u) {
        err = "unknown output projection";
        return SKY_ERR_INVALID_CONFIG;
    }
    const double growth = cfg->shell_growth == 0.0 ? 1.0 : cfg->shell_growth;
    if (!(growth >= 1.0 && growth <= 16.0) || cfg->majorant_layers * std::log2(growth) > 1000.0) {
        err = "shell growth must lie in [1, 16] (0 means 1)";
        return SKY_ERR_INVALID_CONFIG;
    }

    if (a->tables) {
        if (!validate_tables(a, wl, err)) return SKY_ERR_INVALID_CONFIG;
        build_medium_table(a, wl, planet->atmosphere_radius_km - planet->ground_radius_km, out.tab);
    } else {
        out.tab = MediumTable{};
    }

    out.nlam = wl->count;
    fill_wavelength_rows(a, wl, out.lam);
    if (cfg->majorant_table) {
        size_t n = (size_t)wl->count * cfg->majorant_layers;
        out.maj.assign(cfg->majorant_table, cfg->majorant_table + n);
        for (float m : out.maj) {
            if (!(m > 0.0f) || !std::isfinite(m)) {
                err = "majorant table entries must be finite and positive";
                return SKY_ERR_INVALID_CONFIG;
            }
        }
        // Delta tracking samples the free-flight law only where the majorant
        // bounds sigma_t; below it the accept ratio would be clamped to 1 and
        // the render silently biased.  Each entry must reach its shell's
        // exact maximum extinction (the reference's 1e-8 floor included),
        // up to the FP32 rounding of a table made with margin 1.
        std::vector<float> exact;
        fill_majorants(planet, a, wl, cfg->majorant_layers, 1.0, growth, exact);
        for (size_t i = 0; i < n; ++i) {
            if (out.maj[i] < exact[i] * (1.0f - 1e-6f)) {
                err = "majorant table entry " + std::to_string(i) + " (band " +
                      std::to_string(i / cfg->majorant_layers) + ", shell " +
                      std::to_string(i % cfg->majorant_layers) + ") is below the shell's maximum extinction " +
                      std::to_string(exact[i]) + " km^-1";
                return SKY_ERR_INVALID_CONFIG;
            }
        }
    } else {
        fill_majorants(planet, a, wl, cfg->majorant_layers, cfg->majorant_margin, growth, out.maj);
    }

    DevScene &s = out.dev;
    std::memset(&s, 0, sizeof(s));
    s.R = planet->ground_radius_km;
    s.top = planet->atmosphere_radius_km - planet->ground_radius_km;
    s.L = cfg->majorant_layers;
    s.dz = s.top / s.L;
    out.shell_w.resize(s.L + 1);
    for (uint32_t k = 0; k <= s.L; ++k) {
        double h = shell_height(s.top, s.L, growth, k);
        out.shell_w[k] = (float)(h * (2.0 * s.R + h));
    }
    s.shell_tolf = (float)(2e-9 * s.R);  // 1e-9 km of height
    if (growth > 1.0) {
        s.shell_c1 = (float)((std::pow(growth, (double)s.L) - 1.0) / s.top);
        s.shell_c2 = (float)(1.0 / std::log2(growth));
    } else {
        s.shell_c1 = 0.0f;
        s.shell_c2 = (float)(s.L / s.top);
    }
    s.delta_topf = (float)((planet->atmosphere_radius_km - s.R) * (planet->atmosphere_radius_km + s.R));
    const double h_obs = std::fmax(cfg->observer_altitude_km, 0.0);
    s.obs_y = s.R + h_obs;
    s.obs_yf = (float)s.obs_y;
    s.obs_wf = (float)(h_obs * (2.0 * s.R + h_obs));
    s.Rf = (float)s.R;
    s.R2f = (float)(s.R * s.R);
    s.topf = (float)s.top;
    s.eps = kRayEpsilonKm;
    // Sun::from_degrees, atmosphere.rs:20-29
    double e = cfg->sun_elevation_deg * kPiD / 180.0, az = cfg->sun_azimuth_deg * kPiD / 180.0;
    double sd[3] = {std::cos(e) * std::sin(az), std::sin(e), std::cos(e) * std::cos(az)};
    double sn = std::sqrt(sd[0] * sd[0] + sd[1] * sd[1] + sd[2] * sd[2]);
    for (int i = 0; i < 3; ++i) s.sun[i] = (float)(sd[i] / sn);
    double hsin = std::sin(0.5 * a->sun_angular_radius_rad);
    double omc = 2.0 * hsin * hsin;
    s.omc = (float)omc;
    s.cone_pdf = (float)(1.0 / (2.0 * kPiD * omc));
    s.cone_omega = (float)(2.0 * kPiD * omc);
    s.inv_hr = (float)(1.0 / a->rayleigh_h_km);
    s.inv_ha = (float)(1.0 / a->aerosol_h_km);
    s.l2_hr = (float)(1.4426950408889634 / a->rayleigh_h_km);
    s.l2_ha = (float)(1.4426950408889634 / a->aerosol_h_km);
    s.hc = (float)a->ozone_center_km;
    s.inv_hw = (float)(1.0 / a->ozone_half_width_km);
    s.g = (float)a->aerosol_g;
    s.inv_2g = std::fabs(s.g) < 1e-3f ? 0.0f : 1.0f / (2.0f * s.g);
    s.albedo = (float)a->ground_albedo;
    // Sky-view LUT frame, sky_trace.wgsl:632-681
    {
        double vh = s.obs_y, R = s.R;
        double v_horizon = std::sqrt(std::fmax(vh * vh - R * R, 0.0));
        double cos_beta = std::fmin(std::fmax(v_horizon / std::fmax(vh, 1e-6), 0.0), 1.0);
        double beta = std::acos(cos_beta);
        s.svl_zha = (float)(kPiD - beta);
        s.svl_beta = (float)beta;
        double sh[3] = {sd[0] / sn, 0.0, sd[2] / sn};
        double shl = std::sqrt(sh[0] * sh[0] + sh[2] * sh[2]);
        double fw[3] = {1.0, 0.0, 0.0};
        if (shl > 1e-5) {
            fw[0] = sh[0] / shl;
            fw[2] = sh[2] / shl;
        }
        double side[3] = {fw[2], 0.0, -fw[0]};
        double sl = std::sqrt(side[0] * side[0] + side[2] * side[2]);
        for (int i = 0; i < 3; ++i) {
            s.svl_fw[i] = (float)fw[i];
            s.svl_side[i] = (float)(side[i] / sl);
        }
    }
    s.width = cfg->width;
    s.height = cfg->height;
    s.inv_w = (float)(1.0 / cfg->width);
    s.inv_h = (float)(1.0 / cfg->height);
    s.spp = cfg->spp;
    s.sample_offset = cfg->sample_offset;
    s.seed_lo = (uint32_t)cfg->seed;
    s.seed_hi = (uint32_t)(cfg->seed >> 32);
    s.nlam = wl->count;
    s.inv_nlam = (float)(1.0 / wl->count);
    s.dls = cfg->direct_light_samples ? cfg->direct_light_samples : 1u;  // integrator.rs:325
    s.inv_dls = 1.0f / (float)s.dls;
    s.projection = cfg->projection;
    s.max_depth = cfg->max_depth;
    s.rr_start = cfg->rr_start_depth;
    s.watchdog_limit = cfg->watchdog_limit;
    s.max_segments = cfg->max_segments;
    s.flags = cfg->flags;
    // shell crossings a lane takes per tracking step (the general kernels and
    // the ratio-tracked specialised ones): 2 up to 4 shells, 4 up to 16, then
    // 8 (on 1024x512x64 renders, the paths unchanged: ratio tracking in 4
    // shells growing 2.5x 385 -> 370 ms per 128-spp launch; 8 uniform shells
    // 310 -> 249 ms and 9 shells 310 -> 277 ms per 64-spp launch, 16 shells
    // 235 -> 157 ms and 32 shells 529 -> 287 ms per 32-spp launch, 64 shells
    // 338 -> 163 ms per 16-spp launch)
    s.walk_cap = s.L <= 4u ? 2u : (s.L <= 16u ? 4u : 8u);
    return SKY_OK;
}

}  // namespace sky
