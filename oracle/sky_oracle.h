/*
 * sky_oracle.h -- CPU oracle for the spectral path-tracing integrator.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * estimator (crates/sky-spectral-path-tracer/src/sky_trace.wgsl and the
 * sky-core helpers it mirrors) over the analytic atmosphere used by the
 * benchmark configurations.  It exists so that tests, the smoke check and the
 * CPU-baseline leg of bench.py can check the HIP product against it.  Nothing
 * in the product links, loads or calls this code.
 *
 * Two precisions are built from one source (sky_oracle_trace.inc):
 *   *_f32 : per-step arithmetic in float, per-ray set-up (ray origin, o.d,
 *           impact parameter, shell discriminants) in double -- the same
 *           precision split as the GPU kernel, so paths pair with it.
 *   *_f64 : everything in double -- the accuracy oracle.
 *
 * Every function cites the reference file:line it restates.  Where the
 * configurations change the reference (analytic medium, HG aerosol phase,
 * Planck sun, depth cap, phase importance sampling, ratio tracking) the
 * deviation is named at the function.
 */
#ifndef SKY_ORACLE_H
#define SKY_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_FLAG_PHASE_IS 1u       /* sample Rayleigh/HG lobes instead of isotropic */
#define ORC_FLAG_RATIO_TRACKING 2u /* ratio tracking for shadow-ray transmittance */
#define ORC_FLAG_CARRIED_TAU 4u    /* one exponential draw carried across shells  */

#define ORC_PROJ_PANORAMA 0u
#define ORC_PROJ_SKY_VIEW_LUT 1u

/* Event counters for the algorithmic-flop model (all totals over a render). */
typedef struct orc_counters {
    uint64_t paths;            /* path x lambda samples started            */
    uint64_t boundary_queries; /* next_boundary calls                       */
    uint64_t segments;         /* majorant segments entered (shell walk)    */
    uint64_t free_flights;     /* free-flight draws (incl. overshoots)      */
    uint64_t density_evals;    /* tentative collisions (sigma_t evaluated)  */
    uint64_t scatter_vertices; /* real collisions that scattered            */
    uint64_t ground_vertices;  /* ground hits processed                     */
    uint64_t nee_samples;      /* sun cone samples drawn                    */
    uint64_t shadow_rays;      /* shadow rays traced                        */
    uint64_t rng_draws;        /* pcg4d evaluations                         */
    uint64_t watchdog;         /* paths that tripped a watchdog             */
    uint64_t graze64;          /* FP32 ray set-ups formed in FP64 (ground-grazing rays) */
} orc_counters;

/* The reference's tabulated SceneData (data.rs:53-152): profiles in
 * ascending altitude, per-wavelength optics.  NULL in orc_params: the
 * analytic profiles. */
#define ORC_SPECIES 4 /* inso, waso, soot, suso (atmosphere.rs:6-8) */
typedef struct orc_tables {
    uint32_t atmosphere_points, aerosol_points;
    const double *atmosphere;      /* [n][3]: altitude_km, air_cm3, ozone_cm3 */
    const double *aerosol;         /* [n][5]: altitude_km, mass_g_m3 x 4      */
    const double *aerosol_optics;  /* [n_lambda][4][2]: scattering, absorption per g/m^3 */
    const double *ozone_cross_section_cm2;   /* [n_lambda] */
    const double *rayleigh_cross_section_m2; /* [n_lambda] or NULL: 5.8e-31 (550/l)^4 */
    const double *solar_irradiance_w_m2;     /* [n_lambda] or NULL: Planck */
} orc_tables;

typedef struct orc_params {
    /* Planet (geometry.rs:12-17; the configurations use a 6460 km top). */
    double ground_radius_km;
    double atmosphere_radius_km;
    /* Analytic atmosphere: beta in m^-1 at 550 nm (sea level), heights km. */
    double rayleigh_sca_550_m_inv;
    double rayleigh_h_km;
    double ozone_peak_m_inv;
    double ozone_center_km;
    double ozone_half_width_km;
    double ozone_lambda0_nm;
    double ozone_sigma_nm;
    double aerosol_sca_550_m_inv;
    double aerosol_abs_550_m_inv;
    double aerosol_h_km;
    double aerosol_g;
    double angstrom;
    double ground_albedo;
    double sun_temperature_k;
    double sun_angular_radius_rad;
    /* Render configuration (config.rs:2-12) plus the extensions. */
    uint32_t width, height, spp, sample_offset;
    uint64_t seed;
    double sun_elevation_deg, sun_azimuth_deg, observer_altitude_km;
    uint32_t direct_light_samples;
    uint32_t projection;
    uint32_t max_depth;      /* 0 = uncapped (reference), else vertex cap   */
    uint32_t rr_start_depth; /* reference: 4 (sky_trace.wgsl:692)           */
    uint32_t watchdog_limit; /* reference: 1024 (integrator.rs:14)          */
    uint32_t max_segments;   /* reference: 100000 (sky_trace.wgsl:412)      */
    uint32_t majorant_layers;/* reference: 64 (data.rs:48)                  */
    uint32_t flags;
    double majorant_margin;  /* reference: 1.05 (medium.rs:77)              */
    uint32_t n_lambda;
    const double *lambda_nm;
    const double *weight_nm;
    /* Majorant shell thicknesses grow by this factor from the ground up
     * (1: the reference's uniform shells, medium.rs:63-83).  Any shells give
     * the same delta-tracking law; thin shells near the ground, where the
     * density falls fastest, give tighter majorants for the same count. */
    double shell_growth;
    /* Tabulated medium (NULL: analytic).  Aerosol phase stays HG (aerosol_g):
     * the reference's Mie table is absent from its checkout. */
    const orc_tables *tables;
} orc_params;

/* Per-wavelength constants derived from the parameters. */
typedef struct orc_lambda {
    double beta_r, beta_o, beta_as, beta_aa; /* km^-1 at the reference height */
    double sun_radiance;                     /* W m^-2 sr^-1 nm^-1            */
    double wx, wy, wz;                       /* cmf(lambda) * weight_nm       */
} orc_lambda;

/* Height of the lower boundary of shell k (k = 0..L; h_L = top). */
double orc_shell_height(const orc_params *p, int k);
/* Shell containing altitude h: the largest k < L with h_k <= h. */
int orc_shell_of(const orc_params *p, double h);

/* RNG: pcg4d counter hash (sky_trace.wgsl:156-189). */
void orc_pcg4d(const uint32_t in[4], uint32_t out[4]);
uint32_t orc_rng_u32(uint32_t pixel, uint32_t sample, uint32_t band, uint32_t dim, uint64_t seed);
float orc_rng_f32(uint32_t pixel, uint32_t sample, uint32_t band, uint32_t dim, uint64_t seed);

/* Colorimetry and source. */
void orc_cmf_xyz(double lambda_nm, double xyz[3]);
double orc_planck_radiance(double lambda_nm, double temperature_k);

/* Scene build (data.rs:15-51 restated for the analytic medium). */
void orc_default_params(orc_params *p);
int orc_lambda_table(const orc_params *p, orc_lambda *out);
void orc_sun_direction(const orc_params *p, double dir[3]);
int orc_majorant_grid(const orc_params *p, double *out /* n_lambda*layers */);
void orc_coefficients(const orc_params *p, const orc_lambda *lam, double altitude_km,
                      double *sigma_r, double *sigma_as, double *sigma_o, double *sigma_aa);
/* Same for band index `band`, either medium (tabulated: medium.rs:28-57). */
void orc_coefficients_band(const orc_params *p, uint32_t band, double altitude_km, double *sigma_r,
                           double *sigma_as, double *sigma_o, double *sigma_aa);

/* Geometry. */
int orc_next_boundary_reference(const orc_params *p, const double o[3], const double d[3],
                                double *t_km, int *kind);
int orc_segments_reference(const orc_params *p, const double o[3], const double d[3], double t_max,
                           double *ends, int *layers, int cap);
int orc_segments_walk(const orc_params *p, const double o[3], const double d[3], double t_max,
                      double *ends, int *layers, int cap);
/* Same for band index `band` of either medium. */
double orc_optical_depth_band(const orc_params *p, uint32_t band, const double o[3], const double d[3],
                              double t_max);
double orc_optical_depth(const orc_params *p, const orc_lambda *lam, const double o[3],
                         const double d[3], double t_max);
/* orc_optical_depth_band for n rays: o, d [n][3], t_max [n] -> out [n]; 0 on success. */
int orc_optical_depth_batch(const orc_params *p, uint32_t band, uint32_t n, const double *o, const double *d,
                            const double *t_max, double *out);

/* Camera (sky_trace.wgsl:610-689). */
void orc_camera_dir(const orc_params *p, double u, double v, double dir[3]);

/* Estimators.  film / film_m2 are [pixel*n_lambda + band] (may be NULL),
 * xyz is [pixel*3 + c] (may be NULL).  Returns 0 on success, 6 if a
 * watchdog tripped, 1 on invalid parameters. */
int orc_render_f32(const orc_params *p, double *film, double *film_m2, double *xyz,
                   orc_counters *counters, int threads);
int orc_render_f64(const orc_params *p, double *film, double *film_m2, double *xyz,
                   orc_counters *counters, int threads);
/* Renders only the listed pixels (pixel indices into the full image). */
int orc_render_pixels_f32(const orc_params *p, const uint32_t *pixels, uint32_t n_pixels,
                          double *film, double *film_m2, orc_counters *counters, int threads);
int orc_render_pixels_f64(const orc_params *p, const uint32_t *pixels, uint32_t n_pixels,
                          double *film, double *film_m2, orc_counters *counters, int threads);

/* One path (pixel, sample, band): radiance of trace_path. */
double orc_trace_path_f32(const orc_params *p, uint32_t x, uint32_t y, uint32_t sample,
                          uint32_t band, orc_counters *counters);
double orc_trace_path_f64(const orc_params *p, uint32_t x, uint32_t y, uint32_t sample,
                          uint32_t band, orc_counters *counters);

/* Tracking estimators on one ray, for statistical known-answer tests.
 * Returns the mean over n keys of the escape indicator (delta tracking,
 * sky_trace.wgsl:441-475) or of the ratio-tracking weight. */
/* The FP32 tracer's ray set-up from the vertex o (rounded to FP32, its
 * r^2 - R^2 formed in double) along d with the origin moved by e:
 * out = {hit, kind, t_max, layer, b, q, w_o}.  Returns 1 where the set-up
 * took the FP64 grazing branch, 0 otherwise, -1 on a bad configuration. */
int orc_ray_setup_f32(const orc_params *p, const double o[3], const double d[3], double e, double out[7]);

double orc_transmittance_mc_f32(const orc_params *p, uint32_t band, const double o[3],
                                const double d[3], uint32_t n, int ratio);

/* Sampling primitives (FP64) for the sampling known-answer tests. */
double orc_sample_rayleigh_mu(double xi);
double orc_sample_hg_mu(double g, double xi);
double orc_phase_rayleigh(double mu);
double orc_phase_hg(double g, double mu);
void orc_frame_dir(const double axis[3], double mu, double st, double phi, double out[3]);
double orc_sun_cone_dir(const orc_params *p, double xi0, double xi1, double out[3]);
double orc_cosine_dir(const double n[3], double u, double xi, double out[3]);

#ifdef __cplusplus
}
#endif
#endif
