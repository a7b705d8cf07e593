/*
 * sky_oracle.c -- CPU oracle (TEST INFRASTRUCTURE ONLY; see sky_oracle.h).
 *
 * Precision-independent parts: RNG, colorimetry, Planck source, scene build,
 * reference-literal geometry used to pin the shell walk, and the
 * deterministic optical-depth quadrature.  The path tracer itself is in
 * sky_oracle_trace.inc, compiled once in float and once in double.
 */
#include "sky_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ORC_PI 3.14159265358979323846
#define ORC_TAU 6.28318530717958647692
#define ORC_MAX_LAYERS 256
#define ORC_MAX_LAMBDA 64
/* sky_trace.wgsl:8 -- ray offset used by the GPU reference (5 m). */
#define ORC_RAY_EPSILON_KM 0.005

/* ------------------------------------------------------------------ RNG */

/* sky_trace.wgsl:156-168 (pcg4d, Jarzynski & Olano 2020). */
void orc_pcg4d(const uint32_t in[4], uint32_t out[4])
{
    uint32_t x = in[0] * 1664525u + 1013904223u;
    uint32_t y = in[1] * 1664525u + 1013904223u;
    uint32_t z = in[2] * 1664525u + 1013904223u;
    uint32_t w = in[3] * 1664525u + 1013904223u;
    x += y * w;
    y += z * x;
    z += x * y;
    w += y * z;
    x ^= x >> 16;
    y ^= y >> 16;
    z ^= z >> 16;
    w ^= w >> 16;
    x += y * w;
    y += z * x;
    z += x * y;
    w += y * z;
    out[0] = x;
    out[1] = y;
    out[2] = z;
    out[3] = w;
}

/* sky_trace.wgsl:174-184: key (pixel ^ seed_lo, sample ^ seed_hi, band, dim),
 * xor-fold of the four hash words. */
uint32_t orc_rng_u32(uint32_t pixel, uint32_t sample, uint32_t band, uint32_t dim, uint64_t seed)
{
    uint32_t in[4] = {pixel ^ (uint32_t)seed, sample ^ (uint32_t)(seed >> 32), band, dim};
    uint32_t v[4];
    orc_pcg4d(in, v);
    return v[0] ^ v[1] ^ v[2] ^ v[3];
}

/* sky_trace.wgsl:186-189: top 24 bits scaled by 2^-24. */
float orc_rng_f32(uint32_t pixel, uint32_t sample, uint32_t band, uint32_t dim, uint64_t seed)
{
    return (float)(orc_rng_u32(pixel, sample, band, dim, seed) >> 8) * 5.9604644775390625e-8f;
}

/* ------------------------------------------------------- colorimetry */

/* Piecewise Gaussian lobe of the CIE 1931 2-degree fit of Wyman, Sloan and
 * Shirley, "Simple Analytic Approximations to the CIE XYZ Color Matching
 * Functions", JCGT 2(2), 2013 (multi-lobe fit, their eq. 2).  Replaces the
 * table lerp of spectrum.rs:196-219 (the table itself is the test fixture). */
static double wss_lobe(double lambda, double mu, double s_lo, double s_hi)
{
    double s = lambda < mu ? s_lo : s_hi;
    double t = (lambda - mu) / s;
    return exp(-0.5 * t * t);
}

void orc_cmf_xyz(double l, double xyz[3])
{
    xyz[0] = 1.056 * wss_lobe(l, 599.8, 37.9, 31.0) + 0.362 * wss_lobe(l, 442.0, 16.0, 26.7) -
             0.065 * wss_lobe(l, 501.1, 20.4, 26.2);
    xyz[1] = 0.821 * wss_lobe(l, 568.8, 46.9, 40.5) + 0.286 * wss_lobe(l, 530.9, 16.3, 31.1);
    xyz[2] = 1.217 * wss_lobe(l, 437.0, 11.8, 36.0) + 0.681 * wss_lobe(l, 459.0, 26.0, 13.8);
}

/* Planck spectral radiance B_lambda(T) in W m^-2 sr^-1 nm^-1.  Replaces the
 * tabulated band irradiance / solid angle of data.rs:30-33. */
double orc_planck_radiance(double lambda_nm, double temperature_k)
{
    const double h = 6.62607015e-34, c = 299792458.0, k = 1.380649e-23;
    double l = lambda_nm * 1e-9;
    double x = h * c / (l * k * temperature_k);
    return 2.0 * h * c * c / (l * l * l * l * l) / expm1(x) * 1e-9;
}

/* -------------------------------------------------------- scene build */

void orc_default_params(orc_params *p)
{
    memset(p, 0, sizeof(*p));
    p->ground_radius_km = 6360.0;
    p->atmosphere_radius_km = 6460.0;
    p->rayleigh_sca_550_m_inv = 1.3558e-5;
    p->rayleigh_h_km = 8.0;
    p->ozone_peak_m_inv = 1.9e-6;
    p->ozone_center_km = 25.0;
    p->ozone_half_width_km = 15.0;
    p->ozone_lambda0_nm = 600.0;
    p->ozone_sigma_nm = 80.0;
    p->aerosol_sca_550_m_inv = 3.996e-6;
    p->aerosol_abs_550_m_inv = 4.4e-6;
    p->aerosol_h_km = 1.2;
    p->aerosol_g = 0.8;
    p->angstrom = 1.3;
    p->ground_albedo = 0.3;
    p->sun_temperature_k = 5778.0;
    p->sun_angular_radius_rad = 0.2675 * ORC_PI / 180.0;
    p->width = 128;
    p->height = 64;
    p->spp = 64;
    p->seed = 1;
    p->sun_elevation_deg = 10.0;
    p->observer_altitude_km = 0.2;     /* config.rs:47 */
    p->direct_light_samples = 1;       /* config.rs:48 */
    p->projection = ORC_PROJ_PANORAMA; /* config.rs:49 */
    p->max_depth = 8;
    p->rr_start_depth = 4;      /* sky_trace.wgsl:692 */
    p->watchdog_limit = 1024;   /* integrator.rs:14 */
    p->max_segments = 100000;   /* sky_trace.wgsl:412 */
    p->majorant_layers = 64;    /* data.rs:48 */
    p->majorant_margin = 1.05;  /* medium.rs:77 */
    p->shell_growth = 1.0;      /* uniform shells, as the reference */
}

/* Per-wavelength coefficients of the analytic medium (data.rs:26-33 and
 * medium.rs:104-106 restated): Rayleigh beta*(550/l)^4, ozone Gaussian band,
 * aerosol Angstrom law; all converted from m^-1 to km^-1. */
/* bracket + lerp of one profile column (medium.rs:108-155): partition_point
 * of z < x, clamped to the end points, a (1 - t) + b t. */
static double profile_lerp(const double *rows, uint32_t n, uint32_t stride, uint32_t col, double x)
{
    uint32_t hi = 0;
    while (hi < n && rows[hi * stride] < x) ++hi;
    if (hi == 0) return rows[col];
    if (hi >= n) return rows[(n - 1) * stride + col];
    const double *lo = rows + (hi - 1) * stride, *up = rows + hi * stride;
    double t = up[0] > lo[0] ? (x - lo[0]) / (up[0] - lo[0]) : 0.0;
    if (t < 0.0) t = 0.0;
    if (t > 1.0) t = 1.0;
    return lo[col] * (1.0 - t) + up[col] * t;
}

static double rayleigh_xs(const orc_params *p, uint32_t b)
{
    const orc_tables *t = p->tables;
    if (t->rayleigh_cross_section_m2) return t->rayleigh_cross_section_m2[b];
    double r = 550.0 / p->lambda_nm[b];
    return 5.8e-31 * r * r * r * r; /* medium.rs:104-106 */
}

/* coefficients_at (medium.rs:28-57) for band b at altitude h (0 <= h <= top). */
static void table_coefficients(const orc_params *p, uint32_t b, double h, double *sr, double *sas, double *so,
                               double *saa)
{
    const orc_tables *t = p->tables;
    double air = profile_lerp(t->atmosphere, t->atmosphere_points, 3, 1, h);
    double o3 = profile_lerp(t->atmosphere, t->atmosphere_points, 3, 2, h);
    *sr = air * rayleigh_xs(p, b) * 1.0e9;
    *so = o3 * t->ozone_cross_section_cm2[b] * 1.0e5;
    if (*sr < 0.0) *sr = 0.0;
    if (*so < 0.0) *so = 0.0;
    *sas = *saa = 0.0;
    for (int s = 0; s < ORC_SPECIES; ++s) {
        double m = profile_lerp(t->aerosol, t->aerosol_points, 5, 1 + s, h);
        if (m < 0.0) m = 0.0;
        *sas += m * t->aerosol_optics[(b * ORC_SPECIES + s) * 2];
        *saa += m * t->aerosol_optics[(b * ORC_SPECIES + s) * 2 + 1];
    }
}

int orc_lambda_table(const orc_params *p, orc_lambda *out)
{
    if (p->n_lambda == 0 || p->n_lambda > ORC_MAX_LAMBDA || !p->lambda_nm) return 1;
    for (uint32_t i = 0; i < p->n_lambda; ++i) {
        double l = p->lambda_nm[i];
        double r = 550.0 / l;
        double dl = (l - p->ozone_lambda0_nm) / p->ozone_sigma_nm;
        double ang = pow(l / 550.0, -p->angstrom);
        out[i].beta_r = p->rayleigh_sca_550_m_inv * 1000.0 * r * r * r * r;
        out[i].beta_o = p->ozone_peak_m_inv * 1000.0 * exp(-0.5 * dl * dl);
        out[i].beta_as = p->aerosol_sca_550_m_inv * 1000.0 * ang;
        out[i].beta_aa = p->aerosol_abs_550_m_inv * 1000.0 * ang;
        out[i].sun_radiance = orc_planck_radiance(l, p->sun_temperature_k);
        if (p->tables) {
            /* sea-level coefficients (for inspection; the tracer reads the profiles) */
            table_coefficients(p, i, 0.0, &out[i].beta_r, &out[i].beta_as, &out[i].beta_o, &out[i].beta_aa);
            if (p->tables->solar_irradiance_w_m2) {
                /* solar_irradiance / solid angle 2 pi (1 - cos r) (data.rs:30-33, atmosphere.rs:26) */
                double hs = sin(0.5 * p->sun_angular_radius_rad);
                out[i].sun_radiance = p->tables->solar_irradiance_w_m2[i] / (4.0 * ORC_PI * hs * hs);
            }
        }
        double cmf[3];
        orc_cmf_xyz(l, cmf);
        double w = p->weight_nm ? p->weight_nm[i] : 1.0;
        out[i].wx = cmf[0] * w;
        out[i].wy = cmf[1] * w;
        out[i].wz = cmf[2] * w;
    }
    return 0;
}

/* atmosphere.rs:20-29 (Sun::from_degrees). */
void orc_sun_direction(const orc_params *p, double dir[3])
{
    double e = p->sun_elevation_deg * ORC_PI / 180.0;
    double a = p->sun_azimuth_deg * ORC_PI / 180.0;
    double v[3] = {cos(e) * sin(a), sin(e), cos(e) * cos(a)};
    double n = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    for (int i = 0; i < 3; ++i) dir[i] = v[i] / n;
}

/* medium.rs:28-55 for the analytic profiles; zero outside [0, top]
 * (medium.rs:30-34, sky_trace.wgsl:358-360). */
void orc_coefficients_band(const orc_params *p, uint32_t band, double h, double *sr, double *sas, double *so,
                           double *saa)
{
    double top = p->atmosphere_radius_km - p->ground_radius_km;
    if (h < 0.0 || h > top) {
        *sr = *sas = *so = *saa = 0.0;
        return;
    }
    if (p->tables) {
        table_coefficients(p, band, h, sr, sas, so, saa);
        return;
    }
    orc_lambda lam[ORC_MAX_LAMBDA];
    orc_lambda_table(p, lam);
    orc_coefficients(p, &lam[band], h, sr, sas, so, saa);
}

void orc_coefficients(const orc_params *p, const orc_lambda *lam, double h, double *sr, double *sas,
                      double *so, double *saa)
{
    double top = p->atmosphere_radius_km - p->ground_radius_km;
    if (h < 0.0 || h > top) {
        *sr = *sas = *so = *saa = 0.0;
        return;
    }
    double rr = exp(-h / p->rayleigh_h_km);
    double ra = exp(-h / p->aerosol_h_km);
    double ro = 1.0 - fabs(h - p->ozone_center_km) / p->ozone_half_width_km;
    if (ro < 0.0) ro = 0.0;
    *sr = lam->beta_r * rr;
    *sas = lam->beta_as * ra;
    *so = lam->beta_o * ro;
    *saa = lam->beta_aa * ra;
}

/* Shell boundaries: h_k = top k / L (uniform), else
 * h_k = top (g^k - 1) / (g^L - 1) for growth g > 1. */
double orc_shell_height(const orc_params *p, int k)
{
    double top = p->atmosphere_radius_km - p->ground_radius_km;
    int L = (int)p->majorant_layers;
    double g = p->shell_growth;
    if (k <= 0) return 0.0;
    if (k >= L) return top;
    if (!(g > 1.0)) return k * (top / L);
    return top * (pow(g, k) - 1.0) / (pow(g, L) - 1.0);
}

int orc_shell_of(const orc_params *p, double h)
{
    int L = (int)p->majorant_layers;
    int k = 0;
    while (k < L - 1 && orc_shell_height(p, k + 1) <= h) ++k;
    return k;
}

/* Majorant grid, medium.rs:63-83: margin * max extinction over each shell,
 * floored at 1e-8 (medium.rs:77, sky_trace.wgsl:260).  For these profiles the
 * maximum is exact: both exponentials peak at the shell's lower edge and the
 * tent peaks at the point of the shell closest to its centre. */
int orc_majorant_grid(const orc_params *p, double *out)
{
    orc_lambda lam[ORC_MAX_LAMBDA];
    if (orc_lambda_table(p, lam)) return 1;
    uint32_t L = p->majorant_layers;
    if (L == 0 || L > ORC_MAX_LAYERS) return 1;
    if (p->tables) {
        /* compute_majorant_grid (medium.rs:63-102): max over the shell's ends,
         * middle and the profile altitudes inside it, times the margin */
        const orc_tables *t = p->tables;
        for (uint32_t b = 0; b < p->n_lambda; ++b)
            for (uint32_t k = 0; k < L; ++k) {
                double lo = orc_shell_height(p, (int)k), hi = orc_shell_height(p, (int)k + 1);
                double probes[3] = {lo, hi, 0.5 * (lo + hi)}, m = 0.0, c[4];
                for (int i = 0; i < 3; ++i) {
                    orc_coefficients_band(p, b, probes[i], &c[0], &c[1], &c[2], &c[3]);
                    double e = c[0] + c[1] + c[2] + c[3];
                    if (e > m) m = e;
                }
                for (int pass = 0; pass < 2; ++pass) {
                    uint32_t n = pass ? t->aerosol_points : t->atmosphere_points;
                    const double *rows = pass ? t->aerosol : t->atmosphere;
                    uint32_t stride = pass ? 5 : 3;
                    for (uint32_t i = 0; i < n; ++i) {
                        double z = rows[i * stride];
                        if (!(z > lo && z < hi)) continue;
                        orc_coefficients_band(p, b, z, &c[0], &c[1], &c[2], &c[3]);
                        double e = c[0] + c[1] + c[2] + c[3];
                        if (e > m) m = e;
                    }
                }
                m *= p->majorant_margin;
                out[b * L + k] = m > 1e-8 ? m : 1e-8;
            }
        return 0;
    }
    for (uint32_t b = 0; b < p->n_lambda; ++b) {
        for (uint32_t k = 0; k < L; ++k) {
            double lo = orc_shell_height(p, (int)k), hi = orc_shell_height(p, (int)k + 1);
            double d = 0.0;
            if (p->ozone_center_km < lo) d = lo - p->ozone_center_km;
            if (p->ozone_center_km > hi) d = p->ozone_center_km - hi;
            double tent = 1.0 - d / p->ozone_half_width_km;
            if (tent < 0.0) tent = 0.0;
            double m = lam[b].beta_r * exp(-lo / p->rayleigh_h_km) +
                       (lam[b].beta_as + lam[b].beta_aa) * exp(-lo / p->aerosol_h_km) +
                       lam[b].beta_o * tent;
            m *= p->majorant_margin;
            out[b * L + k] = m > 1e-8 ? m : 1e-8;
        }
    }
    return 0;
}

/* ---------------------------------------------- reference-literal geometry */

static double dot3(const double a[3], const double b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

/* sky_trace.wgsl:211-221 in double. */
static int ref_intersect_sphere(const double o[3], const double d[3], double radius, double *t0,
                                double *t1)
{
    double b = dot3(o, d);
    double r = sqrt(dot3(o, o));
    double c = (r - radius) * (r + radius);
    double disc = b * b - c;
    if (disc < 0.0) return 0;
    double s = sqrt(disc);
    *t0 = -b - s;
    *t1 = -b + s;
    return 1;
}

/* sky_trace.wgsl:223-250 in double: kind 0 ground, 1 atmosphere exit. */
int orc_next_boundary_reference(const orc_params *p, const double o[3], const double d[3],
                                double *t_km, int *kind)
{
    const double eps = ORC_RAY_EPSILON_KM;
    double a0, a1;
    if (!ref_intersect_sphere(o, d, p->atmosphere_radius_km, &a0, &a1)) return 0;
    double tg = -1.0, g0, g1;
    if (ref_intersect_sphere(o, d, p->ground_radius_km, &g0, &g1)) {
        if (g0 > eps) tg = g0;
        else if (g1 > eps) tg = g1;
    }
    double t_exit = a1 > eps ? a1 : eps;
    if (tg > eps && tg < t_exit) {
        *t_km = tg;
        *kind = 0;
    } else {
        *t_km = t_exit;
        *kind = 1;
    }
    return 1;
}

static void ray_at(const double o[3], const double d[3], double t, double out[3])
{
    for (int i = 0; i < 3; ++i) out[i] = o[i] + d[i] * t;
}

/* sky_trace.wgsl:252-257. */
static int ref_layer_for_altitude(const orc_params *p, double h)
{
    double top = p->atmosphere_radius_km - p->ground_radius_km;
    double c = h < 0.0 ? 0.0 : (h > top - 1e-6 ? top - 1e-6 : h);
    int layer = (int)floor(c / top * p->majorant_layers);
    return layer < (int)p->majorant_layers - 1 ? layer : (int)p->majorant_layers - 1;
}

/* sky_trace.wgsl:263-282. */
static double ref_boundary_candidate(const orc_params *p, const double o[3], const double d[3],
                                     double t, double cur, double hb)
{
    double top = p->atmosphere_radius_km - p->ground_radius_km;
    if (hb <= 0.0 || hb >= top) return cur;
    double t0, t1;
    if (!ref_intersect_sphere(o, d, p->ground_radius_km + hb, &t0, &t1)) return cur;
    if (t0 > t + ORC_RAY_EPSILON_KM && t0 < cur) cur = t0;
    if (t1 > t + ORC_RAY_EPSILON_KM && t1 < cur) cur = t1;
    return cur;
}

/* Segment sequence of the reference tracking loops (sky_trace.wgsl:284-294
 * and the outer loop of 405-439) in double: the end of each majorant
 * segment and the layer whose majorant it uses. */
int orc_segments_reference(const orc_params *p, const double o[3], const double d[3], double t_max,
                           double *ends, int *layers, int cap)
{
    double top = p->atmosphere_radius_km - p->ground_radius_km;
    double dz = top / p->majorant_layers;
    double t = 0.0;
    int n = 0;
    while (t < t_max && n < cap) {
        double probe_t = t + ORC_RAY_EPSILON_KM < t_max ? t + ORC_RAY_EPSILON_KM : t_max;
        double pos[3];
        ray_at(o, d, probe_t, pos);
        int layer = ref_layer_for_altitude(p, sqrt(dot3(pos, pos)) - p->ground_radius_km);
        double seg = t_max;
        seg = ref_boundary_candidate(p, o, d, t, seg, layer * dz);
        seg = ref_boundary_candidate(p, o, d, t, seg, (layer + 1) * dz);
        double mt = t + ORC_RAY_EPSILON_KM < seg ? t + ORC_RAY_EPSILON_KM : seg;
        ray_at(o, d, mt, pos);
        layers[n] = ref_layer_for_altitude(p, sqrt(dot3(pos, pos)) - p->ground_radius_km);
        ends[n] = seg;
        ++n;
        t = seg;
    }
    return n;
}

/* Shell walk used by the oracle tracers and the GPU kernel: same segments as
 * sky_trace.wgsl:284-294 (consecutive crossings of the shells k*dz), found
 * from the ray's impact parameter instead of two sphere intersections per
 * segment.  With b = o.d and q = |o|^2 - b^2 - R^2, the ray reaches radius
 * R + h_k at s = t + b = +-sqrt(D_k - q), D_k = h_k (2R + h_k). */
typedef struct walk_ray {
    double b, q;
    int layer, outward;
} walk_ray;

static void walk_init(const orc_params *p, const double o[3], const double d[3], walk_ray *w)
{
    double R = p->ground_radius_km;
    w->b = dot3(o, d);
    w->q = dot3(o, o) - w->b * w->b - R * R;
    double h0 = sqrt(dot3(o, o)) - R;
    int k = orc_shell_of(p, h0);
    w->outward = w->b >= 0.0;
    if (!w->outward && k > 0 && h0 - orc_shell_height(p, k) <= 1e-9) k -= 1;
    w->layer = k;
}

static double walk_next(const orc_params *p, const walk_ray *w, int k, int out, int *kn, int *on)
{
    double R = p->ground_radius_km;
    int L = (int)p->majorant_layers;
    if (!out) {
        double hk = orc_shell_height(p, k);
        double disc = hk * (2.0 * R + hk) - w->q;
        if (disc > 0.0) {
            *on = 0;
            if (k == 0) { /* ground: the boundary ends the ray */
                *kn = 0;
                return INFINITY;
            }
            *kn = k - 1;
            return -sqrt(disc) - w->b;
        }
    }
    *kn = k + 1;
    *on = 1;
    if (k + 1 >= L) {
        *kn = k;
        return INFINITY;
    }
    double hk = orc_shell_height(p, k + 1);
    double disc = hk * (2.0 * R + hk) - w->q;
    return sqrt(disc > 0.0 ? disc : 0.0) - w->b;
}

int orc_segments_walk(const orc_params *p, const double o[3], const double d[3], double t_max,
                      double *ends, int *layers, int cap)
{
    walk_ray w;
    walk_init(p, o, d, &w);
    int k = w.layer, out = w.outward, n = 0;
    double t = 0.0;
    while (t < t_max && n < cap) {
        int kn, on;
        double tc = walk_next(p, &w, k, out, &kn, &on);
        double seg = tc < t_max ? tc : t_max;
        if (seg < t) seg = t;
        ends[n] = seg;
        layers[n] = k;
        ++n;
        t = seg;
        k = kn;
        out = on;
    }
    return n;
}

/* --------------------------------------------- optical-depth quadrature */

/* 16-point Gauss-Legendre rule on [-1, 1] (positive half), generated by
 * Newton iteration on P_16 with the three-term recurrence. */
static double gl16_x[8], gl16_w[8];

static void gl16_init(void)
{
    static int done = 0;
    if (done) return;
    const int n = 16;
    for (int i = 0; i < n / 2; ++i) {
        double x = cos(ORC_PI * (i + 0.75) / (n + 0.5)), dp = 0.0;
        for (int it = 0; it < 100; ++it) {
            double p0 = 1.0, p1 = x;
            for (int k = 2; k <= n; ++k) {
                double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            dp = n * (x * p1 - p0) / (x * x - 1.0);
            double dx = p1 / dp;
            x -= dx;
            if (fabs(dx) < 1e-16) break;
        }
        gl16_x[n / 2 - 1 - i] = x;
        gl16_w[n / 2 - 1 - i] = 2.0 / ((1.0 - x * x) * dp * dp);
    }
    done = 1;
}

/* Band of the current optical-depth integral (tabulated medium only). */
static _Thread_local uint32_t od_band;

static double sigma_t_at(const orc_params *p, const orc_lambda *lam, const double o[3],
                         const double d[3], double t)
{
    double pos[3], sr, sas, so, saa;
    ray_at(o, d, t, pos);
    double h = sqrt(dot3(pos, pos)) - p->ground_radius_km;
    if (p->tables)
        orc_coefficients_band(p, od_band, h, &sr, &sas, &so, &saa);
    else
        orc_coefficients(p, lam, h, &sr, &sas, &so, &saa);
    return sr + sas + so + saa;
}

static double gl_piece(const orc_params *p, const orc_lambda *lam, const double o[3],
                       const double d[3], double a, double b)
{
    double c = 0.5 * (a + b), hw = 0.5 * (b - a), s = 0.0;
    for (int i = 0; i < 8; ++i)
        s += gl16_w[i] * (sigma_t_at(p, lam, o, d, c - hw * gl16_x[i]) +
                          sigma_t_at(p, lam, o, d, c + hw * gl16_x[i]));
    return s * hw;
}

static int cmp_double(const void *a, const void *b)
{
    double x = *(const double *)a, y = *(const double *)b;
    return (x > y) - (x < y);
}

/* Deterministic optical depth along [0, t_max]: the integrand is split at
 * every crossing of the majorant shells, of the ozone tent's kinks and of
 * the perigee, then each piece is integrated by 16-point Gauss-Legendre
 * (split further so that no piece exceeds 8 km). */
double orc_optical_depth_band(const orc_params *p, uint32_t band, const double o[3], const double d[3],
                              double t_max)
{
    orc_lambda lam[ORC_MAX_LAMBDA];
    if (band >= p->n_lambda || orc_lambda_table(p, lam)) return NAN;
    od_band = band;
    return orc_optical_depth(p, &lam[band], o, d, t_max);
}

/* orc_optical_depth_band over n rays (origins and directions [n][3], lengths
 * [n]) in one call, for quadratures that need many transmittances. */
int orc_optical_depth_batch(const orc_params *p, uint32_t band, uint32_t n, const double *o, const double *d,
                            const double *t_max, double *out)
{
    orc_lambda lam[ORC_MAX_LAMBDA];
    if (band >= p->n_lambda || orc_lambda_table(p, lam)) return 1;
    gl16_init();
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        od_band = band;
        out[i] = orc_optical_depth(p, &lam[band], o + 3 * i, d + 3 * i, t_max[i]);
    }
    return 0;
}

double orc_optical_depth(const orc_params *p, const orc_lambda *lam, const double o[3],
                         const double d[3], double t_max)
{
    gl16_init();
    double R = p->ground_radius_km;
    double b = dot3(o, d), q = dot3(o, o) - b * b - R * R;
    double cuts[2048];
    int n = 0;
    cuts[n++] = 0.0;
    cuts[n++] = t_max;
    if (-b > 0.0 && -b < t_max) cuts[n++] = -b;
    /* kinks of the integrand: the ozone tent, or every profile altitude */
    double kinks[512];
    int nk = 0;
    if (p->tables) {
        const orc_tables *t = p->tables;
        for (uint32_t i = 0; i < t->atmosphere_points && nk < 256; ++i) kinks[nk++] = t->atmosphere[3 * i];
        for (uint32_t i = 0; i < t->aerosol_points && nk < 512; ++i) kinks[nk++] = t->aerosol[5 * i];
    } else {
        kinks[nk++] = p->ozone_center_km - p->ozone_half_width_km;
        kinks[nk++] = p->ozone_center_km;
        kinks[nk++] = p->ozone_center_km + p->ozone_half_width_km;
    }
    double top = p->atmosphere_radius_km - R;
    int L = (int)p->majorant_layers;
    for (int k = 0; k <= L + nk && n < 2000; ++k) {
        double hk = k <= L ? k * top / L : kinks[k - L - 1];
        double disc = hk * (2.0 * R + hk) - q;
        if (disc <= 0.0) continue;
        double s = sqrt(disc);
        double t0 = -s - b, t1 = s - b;
        if (t0 > 0.0 && t0 < t_max) cuts[n++] = t0;
        if (t1 > 0.0 && t1 < t_max) cuts[n++] = t1;
    }
    qsort(cuts, n, sizeof(double), cmp_double);
    double tau = 0.0;
    for (int i = 0; i + 1 < n; ++i) {
        double a = cuts[i], e = cuts[i + 1];
        if (e <= a) continue;
        int m = (int)ceil((e - a) / 8.0);
        for (int j = 0; j < m; ++j)
            tau += gl_piece(p, lam, o, d, a + (e - a) * j / m, a + (e - a) * (j + 1) / m);
    }
    return tau;
}

/* ------------------------------------------------------------- camera */

/* sky_trace.wgsl:614-623. */
static void panorama_dir_d(double u, double v, double dir[3])
{
    double az = (u - 0.5) * ORC_TAU, el = (0.5 - v) * ORC_PI;
    double ce = cos(el);
    double x[3] = {ce * sin(az), sin(el), ce * cos(az)};
    double n = sqrt(dot3(x, x));
    for (int i = 0; i < 3; ++i) dir[i] = n > 0 ? x[i] / n : (i == 1);
}

void orc_camera_dir(const orc_params *p, double u, double v, double dir[3])
{
    if (p->projection != ORC_PROJ_SKY_VIEW_LUT) {
        panorama_dir_d(u, v, dir);
        return;
    }
    /* sky_trace.wgsl:625-681 */
    double uu = p->width <= 1 ? 0.5 : (u - 0.5 / p->width) * (p->width / (p->width - 1.0));
    double vv = p->height <= 1 ? 0.5 : (v - 0.5 / p->height) * (p->height / (p->height - 1.0));
    uu = uu < 0 ? 0 : (uu > 1 ? 1 : uu);
    vv = vv < 0 ? 0 : (vv > 1 ? 1 : vv);
    double R = p->ground_radius_km;
    double vh = R + (p->observer_altitude_km > 0 ? p->observer_altitude_km : 0);
    double v_horizon = sqrt(fmax(vh * vh - R * R, 0.0));
    double cos_beta = fmin(fmax(v_horizon / fmax(vh, 1e-6), 0.0), 1.0);
    double beta = acos(cos_beta);
    double zha = ORC_PI - beta;
    double vzc;
    if (vv < 0.75) {
        double c = vv / 0.75;
        c = 1.0 - c;
        c = c * c;
        c = 1.0 - c;
        vzc = cos(zha * c);
    } else {
        double c = (vv - 0.75) / 0.25;
        c = c * c;
        vzc = cos(zha + beta * c);
    }
    double c = uu * uu;
    double lvc = -(c * 2.0 - 1.0);
    double sun[3];
    orc_sun_direction(p, sun);
    double sh[3] = {sun[0], 0.0, sun[2]};
    double shl = sqrt(dot3(sh, sh));
    double fw[3] = {1.0, 0.0, 0.0};
    if (shl > 1e-5)
        for (int i = 0; i < 3; ++i) fw[i] = sh[i] / shl;
    /* side = normalize(cross(up, forward)), up = (0,1,0) */
    double side[3] = {fw[2], 0.0, -fw[0]};
    double sl = sqrt(dot3(side, side));
    for (int i = 0; i < 3; ++i) side[i] /= sl;
    double sv = sqrt(fmax(1.0 - vzc * vzc, 0.0));
    double ss = sqrt(fmax(1.0 - lvc * lvc, 0.0));
    double x[3];
    for (int i = 0; i < 3; ++i) x[i] = sv * (fw[i] * lvc + side[i] * ss);
    x[1] += vzc;
    double n = sqrt(dot3(x, x));
    for (int i = 0; i < 3; ++i) dir[i] = n > 0 ? x[i] / n : (i == 1);
}

/* ------------------------------------------------ validation helper */

/* integrator.rs:130-163 restated for the generalised wavelength set. */
static int validate(const orc_params *p)
{
    if (p->width == 0 || p->height == 0 || p->spp == 0) return 1;
    if ((uint64_t)p->width * p->height * (p->n_lambda ? p->n_lambda : 1) > 0xFFFFFFFFull) return 1;
    if (p->n_lambda == 0 || p->n_lambda > ORC_MAX_LAMBDA || !p->lambda_nm) return 1;
    if (p->majorant_layers == 0 || p->majorant_layers > ORC_MAX_LAYERS) return 1;
    if (!(p->shell_growth >= 1.0 && p->shell_growth <= 16.0)) return 1;
    if (p->majorant_layers * log2(p->shell_growth) > 1000.0) return 1;
    if (!(p->rayleigh_h_km > 0) || !(p->aerosol_h_km > 0) || !(p->ozone_half_width_km > 0)) return 1;
    if (!(p->aerosol_g > -1 && p->aerosol_g < 1)) return 1;
    if (!(p->atmosphere_radius_km > p->ground_radius_km) || !(p->ground_radius_km > 0)) return 1;
    for (uint32_t i = 0; i < p->n_lambda; ++i)
        if (!(p->lambda_nm[i] > 0)) return 1;
    if (p->tables) {
        const orc_tables *t = p->tables;
        if (!t->atmosphere || !t->aerosol || !t->aerosol_optics || !t->ozone_cross_section_cm2 ||
            t->atmosphere_points == 0 || t->aerosol_points == 0)
            return 1;
    }
    return 0;
}

/* ------------------------------------------ precision-templated tracer */

typedef struct orc_rng {
    uint32_t x, y, z, w;
} orc_rng;

static inline uint32_t rng_next_u32(orc_rng *r, orc_counters *c)
{
    uint32_t in[4] = {r->x, r->y, r->z, r->w}, v[4];
    r->w += 1u;
    orc_pcg4d(in, v);
    c->rng_draws++;
    return v[0] ^ v[1] ^ v[2] ^ v[3];
}

static void add_counters(orc_counters *dst, const orc_counters *src)
{
    uint64_t *d = (uint64_t *)dst;
    const uint64_t *s = (const uint64_t *)src;
    for (size_t i = 0; i < sizeof(orc_counters) / sizeof(uint64_t); ++i) {
#pragma omp atomic
        d[i] += s[i];
    }
}

#define REAL float
#define FN(name) name##_f32
#define RF(x) x##f
#define ORC_GRAZE64 1
#define RFMA(a, b, c) fmaf(a, b, c) /* the GPU's fused multiply-adds of the geometry */
#include "sky_oracle_trace.inc"
#undef REAL
#undef FN
#undef RF
#undef ORC_GRAZE64
#undef RFMA

#define REAL double
#define FN(name) name##_f64
#define RF(x) x
#define ORC_GRAZE64 0
#define RFMA(a, b, c) ((a) * (b) + (c))
#include "sky_oracle_trace.inc"
#undef REAL
#undef FN
#undef RF

int orc_ray_setup_f32(const orc_params *p, const double o[3], const double d[3], double e, double out[7])
{
    scene_f32 S;
    if (scene_init_f32(p, &S)) return -1;
    orc_counters c;
    memset(&c, 0, sizeof(c));
    float of[3] = {(float)o[0], (float)o[1], (float)o[2]};
    float df[3] = {(float)d[0], (float)d[1], (float)d[2]};
    float w = (float)(o[0] * o[0] + o[1] * o[1] + o[2] * o[2] - S.R * S.R);
    ray_f32 r;
    memset(&r, 0, sizeof(r));
    ray_setup_f32(&S, of, w, df, (float)e, &r, &c);
    double v[7] = {r.hit, r.kind, r.t_max, r.layer, r.bf, r.qf, r.wo};
    memcpy(out, v, sizeof(v));
    scene_free_f32(&S);
    return c.graze64 ? 1 : 0;
}

/* Statistical tracking estimator on one ray (tests of E[T] = exp(-tau)). */
double orc_transmittance_mc_f32(const orc_params *p, uint32_t band, const double o[3],
                                const double d[3], uint32_t n, int ratio)
{
    scene_f32 S;
    if (scene_init_f32(p, &S) || band >= p->n_lambda) return -1.0;
    orc_counters c;
    memset(&c, 0, sizeof(c));
    float df[3] = {(float)d[0], (float)d[1], (float)d[2]};
    float of[3] = {(float)o[0], (float)o[1], (float)o[2]};
    float w = (float)(o[0] * o[0] + o[1] * o[1] + o[2] * o[2] - S.R * S.R);
    ray_f32 r;
    ray_setup_f32(&S, of, w, df, 0.0f, &r, &c);
    double sum = 0.0;
    int wd = 0;
    for (uint32_t i = 0; i < n; ++i) {
        orc_rng rng = {i ^ (uint32_t)p->seed, 0x2545F491u, band, 0};
        float T = 0.0f, th, co[4];
        track_f32(&S, &r, (int)band, &rng, ratio ? TRACK_RATIO : TRACK_DELTA, &th, co, &T, &c, &wd);
        sum += T;
    }
    scene_free_f32(&S);
    return sum / n;
}

/* ---------------------------------------- exported sampling primitives */
/* Thin FP64 entry points onto the tracer's own functions, for the sampling
 * known-answer tests (crates/sky-core/src/sampling.rs:234-274). */

double orc_sample_rayleigh_mu(double xi) { return sample_rayleigh_mu_f64(xi); }
double orc_sample_hg_mu(double g, double xi) { return sample_hg_mu_f64(g, xi); }
double orc_phase_rayleigh(double mu) { return phase_rayleigh_f64(mu); }
double orc_phase_hg(double g, double mu) { return phase_hg_f64(g, mu); }

void orc_frame_dir(const double axis[3], double mu, double st, double phi, double out[3])
{
    frame_dir_f64(axis, mu, st, phi, out);
}

/* sample_uniform_cone around the parameters' sun; returns the pdf. */
double orc_sun_cone_dir(const orc_params *p, double xi0, double xi1, double out[3])
{
    scene_f64 S;
    if (scene_init_f64(p, &S)) return -1.0;
    orc_rng dummy = {0, 0, 0, 0};
    (void)dummy;
    double xi = xi0 > 0.99999994 ? 0.99999994 : xi0;
    double x = xi * S.omc;
    frame_dir_f64(S.sun, 1.0 - x, sqrt(x * (2.0 - x)), ORC_TAU * xi1, out);
    double pdf = S.cone_pdf;
    int inside = in_sun_cone_f64(&S, out);
    scene_free_f64(&S);
    return inside ? pdf : -pdf;
}

/* sample_cosine_hemisphere (sky_trace.wgsl:548-560); returns the pdf. */
double orc_cosine_dir(const double n[3], double u, double xi, double out[3])
{
    double ct = sqrt(fmax(1.0 - u, 0.0));
    frame_dir_f64(n, ct, sqrt(u), ORC_TAU * xi, out);
    double nn[3];
    normalize_or_y_f64(n, nn);
    double c = nn[0] * out[0] + nn[1] * out[1] + nn[2] * out[2];
    return (c > 0 ? c : 0.0) / ORC_PI;
}
