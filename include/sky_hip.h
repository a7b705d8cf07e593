/*
 * sky_hip.h -- C ABI of the MI355X spectral sky path tracer (libsky_hip.so).
 *
 * Drop-in boundary for the reference's integrator entry point
 *   sky_spectral_path_tracer::integrator::render(&SceneData, &RenderConfig)
 *       -> Result<Film, RenderError>
 * (crates/sky-spectral-path-tracer/src/integrator.rs:125-128, re-exported at
 * lib.rs:7).  Plain C types, caller-owned buffers, no torch types.
 *
 * Threading: a context is single-threaded (not reentrant); distinct contexts
 * may be used concurrently from different threads.  The library never keeps
 * caller pointers after a call returns.
 */
#ifndef SKY_HIP_H
#define SKY_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SKY_HIP_ABI_VERSION 4

/* RenderError (integrator.rs:19-26) as status codes. */
typedef enum sky_status {
    SKY_OK = 0,
    SKY_ERR_INVALID_CONFIG = 1, /* RenderError::InvalidConfig                   */
    SKY_ERR_NO_DEVICE = 2,      /* RenderError::NoAdapter                       */
    SKY_ERR_DEVICE_INIT = 3,    /* RenderError::RequestDevice                   */
    SKY_ERR_COPY = 4,           /* RenderError::BufferMap                       */
    SKY_ERR_SYNC = 5,           /* RenderError::Poll                            */
    SKY_ERR_WATCHDOG = 6,       /* RenderError::Watchdog                        */
    SKY_ERR_OOM = 7             /* device allocation failed (no reference twin) */
} sky_status;

/* Estimator flags (sky_render_config.flags). */
#define SKY_FLAG_PHASE_IS 1u       /* Rayleigh/HG lobe sampling instead of isotropic (sky_trace.wgsl:735-738) */
#define SKY_FLAG_RATIO_TRACKING 2u /* ratio tracking for shadow rays instead of delta (sky_trace.wgsl:441-475) */
#define SKY_FLAG_CARRIED_TAU 4u    /* carry the free-flight optical depth across majorant shells */
#define SKY_FLAG_SIMPLE_KERNEL 256u /* one-thread-per-(pixel, lambda) kernel instead of the wave-pool kernel */

/* OutputProjection (config.rs:14-36). */
#define SKY_PROJECTION_PANORAMA 0u
#define SKY_PROJECTION_SKY_VIEW_LUT 1u

/* Opaque: device, stream, cached device buffers. */
typedef struct sky_context sky_context;

/* Planet (crates/sky-core/src/geometry.rs:6-17). */
typedef struct sky_planet {
    double ground_radius_km;
    double atmosphere_radius_km;
} sky_planet;

/* Tabulated medium: the reference's SceneData profiles and optics
 * (crates/sky-core/src/data.rs:53-152, atmosphere.rs:40-58, medium.rs:28-57):
 *   sigma_R = air_cm3(h) * rayleigh_cross_section_m2 * 1e9       (km^-1)
 *   sigma_O = ozone_cm3(h) * ozone_cross_section_cm2 * 1e5
 *   sigma_As = sum_s mass_s(h) * scattering_s,  sigma_Aa likewise
 * with every profile interpolated linearly in altitude, clamped to its end
 * points (medium.rs:108-155).  Profiles are given in ascending altitude.  The
 * aerosol phase is Henyey-Greenstein with sky_atmosphere.aerosol_g (the
 * reference's tabulated Mie phase, data/mie_phase.csv, is absent from its
 * checkout). */
#define SKY_AEROSOL_SPECIES 4 /* inso, waso, soot, suso (atmosphere.rs:6-8) */
typedef struct sky_medium_tables {
    uint32_t atmosphere_points;    /* rows of atmosphere_profile.csv (>= 1) */
    uint32_t aerosol_points;       /* rows of aerosol_profile.csv (>= 1)    */
    const double *atmosphere;      /* [atmosphere_points][3]: altitude_km, air_cm3, ozone_cm3 */
    const double *aerosol;         /* [aerosol_points][5]: altitude_km, mass_g_m3 of the 4 species */
    const double *aerosol_optics;  /* [count][4][2]: scattering, absorption (km^-1 per g/m^3) */
    const double *ozone_cross_section_cm2;   /* [count] */
    const double *rayleigh_cross_section_m2; /* [count]; NULL: 5.8e-31 (550/lambda)^4 (medium.rs:104-106) */
    const double *solar_irradiance_w_m2;     /* [count] band irradiance, radiance = E / sun solid angle
                                                (data.rs:30-33); NULL: Planck at sun_temperature_k */
} sky_medium_tables;

/* Analytic replacement of the tabulated SceneData medium
 * (crates/sky-core/src/atmosphere.rs:121-133, medium.rs:28-55):
 *   sigma_R = rayleigh_sca_550 (550/l)^4 exp(-h/H_R)
 *   sigma_O = ozone_peak exp(-(l-l0)^2/(2 s^2)) max(0, 1-|h-h_c|/h_w)
 *   sigma_As = aerosol_sca_550 (l/550)^-angstrom exp(-h/H_A), sigma_Aa likewise
 * HG aerosol phase (g), Lambertian ground, Planck sun of the given radius. */
typedef struct sky_atmosphere {
    double rayleigh_sca_550_m_inv; /* 1.3558e-5 */
    double rayleigh_h_km;          /* 8         */
    double ozone_peak_m_inv;       /* 1.9e-6    */
    double ozone_center_km;        /* 25        */
    double ozone_half_width_km;    /* 15        */
    double ozone_lambda0_nm;       /* 600       */
    double ozone_sigma_nm;         /* 80        */
    double aerosol_sca_550_m_inv;  /* 3.996e-6  */
    double aerosol_abs_550_m_inv;  /* 4.4e-6    */
    double aerosol_h_km;           /* 1.2       */
    double aerosol_g;              /* 0.8       */
    double angstrom;               /* 1.3       */
    double ground_albedo;          /* 0.3       */
    double sun_temperature_k;      /* 5778      */
    double sun_angular_radius_rad; /* 0.2675 deg */
    /* NULL: the analytic profiles above.  Otherwise the tabulated medium
     * (count = the wavelength set's count); only aerosol_g, ground_albedo,
     * sun_temperature_k and sun_angular_radius_rad are then used. */
    const sky_medium_tables *tables;
} sky_atmosphere;

/* Wavelength set (replaces the fixed 41 bands, spectrum.rs:5).  weight_nm
 * multiplies cmf(lambda) in the in-kernel XYZ sum (NULL: uniform spacing). */
typedef struct sky_wavelengths {
    uint32_t count; /* 1..64 */
    const double *center_nm;
    const double *weight_nm;
} sky_wavelengths;

/* RenderConfig (config.rs:2-12) plus the extensions. */
typedef struct sky_render_config {
    uint32_t width, height, spp;
    uint32_t sample_offset; /* first sample index: resume / shard point */
    uint64_t seed;          /* RNG key: seed_lo ^ pixel, seed_hi ^ sample */
    double sun_elevation_deg, sun_azimuth_deg, observer_altitude_km;
    uint32_t direct_light_samples; /* >= 1 */
    uint32_t projection;           /* SKY_PROJECTION_*                 */
    uint32_t max_depth;            /* 0 = uncapped (reference)         */
    uint32_t rr_start_depth;       /* 4 in the reference               */
    uint32_t watchdog_limit;       /* 1024 in the reference            */
    uint32_t max_segments;         /* 100000 in the reference          */
    uint32_t majorant_layers;      /* 64 in the reference (data.rs:48) */
    uint32_t flags;                /* SKY_FLAG_*                       */
    double majorant_margin;        /* 1.05 in the reference            */
    /* Optional host table [count * majorant_layers] (km^-1) replacing the
     * computed majorant grid, e.g. to hold it fixed while parameters move
     * (finite differences with common random numbers, the adjoint).  Every
     * entry must bound its shell's extinction (sky_majorant_grid with
     * margin 1 is the smallest valid table); a smaller one is rejected with
     * SKY_ERR_INVALID_CONFIG, since delta tracking would be biased. */
    const float *majorant_table;
    /* Majorant shell thickness growth from the ground up: 1 (or 0) = the
     * reference's uniform shells; g > 1 gives h_k = top (g^k-1)/(g^L-1),
     * thin shells where the density changes fastest.  Any shells sample the
     * same free-flight law; this only moves cost between shell crossings and
     * null collisions. */
    double shell_growth;
} sky_render_config;

typedef struct sky_render_stats {
    double kernel_ms;      /* device time of the integrator launch(es)  */
    double total_ms;       /* host wall time of the call                */
    uint64_t path_samples; /* width * height * spp * count              */
    uint32_t watchdog;     /* nonzero if a watchdog tripped             */
    uint32_t workgroups;   /* launch shape actually used                */
    uint32_t block;
    uint32_t items;
    /* sky_render_adjoint only: the objective sum dl_dxyz * XYZ over the
     * replayed paths, summed in FP64 (the render's own XYZ image is FP32;
     * this is the same estimator, for finite differences finer than FP32
     * image rounding resolves) */
    double adjoint_value;
} sky_render_stats;

/* Launch tuning of the wave-pool integrator (autotuner output); zero fields
 * mean "use the default". */
typedef struct sky_tuning {
    uint32_t waves_per_cu;    /* persistent waves per compute unit (24)        */
    uint32_t pixels_per_item; /* pixels per wave work item (the fewest giving
                                 16384 paths per item, at most 256 / count)   */
    uint32_t event_threshold; /* lanes waiting before events are handled (8)   */
    uint32_t track_steps;     /* tracking micro-steps between handler passes (2) */
    uint32_t occupancy;       /* forward renders: 0 the compiler's register budget; 4, 5, 6, 8 at
                                 least that many waves per SIMD (the adjoint's replay kernels are
                                 always built for 4) */
} sky_tuning;

/* ---- context ---- */
sky_status sky_context_create(int device, sky_context **out);
void sky_context_destroy(sky_context *ctx);
sky_status sky_context_set_tuning(sky_context *ctx, const sky_tuning *tuning);
sky_status sky_context_get_tuning(const sky_context *ctx, sky_tuning *tuning);

/* ---- defaults (config.rs:38-52, BASELINE analytic atmosphere) ---- */
void sky_default_planet(sky_planet *planet);
void sky_default_atmosphere(sky_atmosphere *atm);
void sky_default_config(sky_render_config *cfg);

/* ---- scene build (data.rs:15-51; integrator.rs:319-418 PackedScene) ----
 * Per-wavelength table, row = [beta_r, beta_o, beta_as, beta_aa (km^-1),
 * sun_radiance (W m^-2 sr^-1 nm^-1), wx, wy, wz]; n_out >= 8 * count. */
sky_status sky_wavelength_table(const sky_atmosphere *atm, const sky_wavelengths *wl, float *out,
                                uint32_t n_out);
/* Majorant grid (medium.rs:63-83): [lambda * layers + layer], km^-1, over
 * the shells of sky_render_config.shell_growth. */
sky_status sky_majorant_grid(const sky_planet *planet, const sky_atmosphere *atm,
                             const sky_wavelengths *wl, uint32_t layers, double margin,
                             double shell_growth, float *out, uint32_t n_out);

/* validate_config (integrator.rs:130-163) for the generalised inputs:
 * SKY_OK or SKY_ERR_INVALID_CONFIG with the reason in sky_last_error(). */
sky_status sky_validate_config(const sky_planet *planet, const sky_atmosphere *atm,
                               const sky_wavelengths *wl, const sky_render_config *cfg);

/* ---- the integrator: render() (integrator.rs:125-128) ----
 * Host buffers, any may be NULL:
 *   film    [(y*W + x) * count + band]  mean radiance, clamped >= 0 (integrator.rs:307-310)
 *   film_m2 same layout, mean of squared path radiance (for standard errors)
 *   xyz     [(y*W + x) * 3 + c]  sum_band film * cmf * weight_nm
 * Synchronous, like pollster::block_on (integrator.rs:127). */
sky_status sky_render(sky_context *ctx, const sky_planet *planet, const sky_atmosphere *atm,
                      const sky_wavelengths *wl, const sky_render_config *cfg, float *film,
                      float *film_m2, float *xyz, sky_render_stats *stats);

/* Same, with device pointers (e.g. torch tensors); ordered on `stream`
 * (hipStream_t, NULL = the context's stream) and synchronous on return. */
sky_status sky_render_device(sky_context *ctx, const sky_planet *planet, const sky_atmosphere *atm,
                             const sky_wavelengths *wl, const sky_render_config *cfg, float *d_film,
                             float *d_film_m2, float *d_xyz, void *stream, sky_render_stats *stats);

/* ---- allocation-free render ----
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory: code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc.  Same inputs, outputs and results as sky_render_device
 * (outputs are the caller's device buffers), but the working buffers -- the
 * packed scene tables, work counters, job record and, when the render's
 * samples are split into chunks, the chunk partial sums -- live in a static
 * device array of SKY_NOALLOC_SCRATCH_BYTES that the library holds on every
 * GPU it runs on until the process ends; the call makes no device
 * allocation.  If the usual sample-chunk partials do not fit, the render
 * runs one chunk per pixel group; a render whose tables alone, or whose
 * samples per pixel (beyond 2^24 / (pixels per item x count)) need more,
 * fails with SKY_ERR_INVALID_CONFIG.  Every call on one GPU shares that
 * GPU's array: calls that overlap on one GPU (other threads, other
 * contexts, other streams) race and must be serialised by the caller.
 * Synchronous on return, ordered on `stream` (NULL: the context's). */
#define SKY_NOALLOC_SCRATCH_BYTES (4u << 20)
sky_status sky_render_device_noalloc(sky_context *ctx, const sky_planet *planet, const sky_atmosphere *atm,
                                     const sky_wavelengths *wl, const sky_render_config *cfg, float *d_film,
                                     float *d_film_m2, float *d_xyz, void *stream, sky_render_stats *stats);

/* ---- adjoint: gradient of the rendered XYZ image ----
 * Returns grad[k] = d/dtheta_k sum_{pixel,c} dl_dxyz[pixel*3+c] * XYZ[pixel*3+c]
 * for the ten parameters, in this order: rayleigh_sca_550_m_inv,
 * rayleigh_h_km, ozone_peak_m_inv, ozone_center_km, aerosol_sca_550_m_inv,
 * aerosol_abs_550_m_inv, aerosol_h_km, aerosol_g, angstrom, ground_albedo.
 * The paths of sky_render with the same configuration are replayed (same
 * RNG keys) with the majorant held fixed; each path carries the
 * log-derivative (score) of its sampling decisions and contribution weights.
 * No counterpart in the reference, which has no derivatives. */
#define SKY_NUM_PARAMS 10
sky_status sky_render_adjoint(sky_context *ctx, const sky_planet *planet, const sky_atmosphere *atm,
                              const sky_wavelengths *wl, const sky_render_config *cfg,
                              const float *dl_dxyz, double *grad, sky_render_stats *stats);
/* Same with dl_dxyz in device memory (ordered on `stream`); grad is host. */
sky_status sky_render_adjoint_device(sky_context *ctx, const sky_planet *planet,
                                     const sky_atmosphere *atm, const sky_wavelengths *wl,
                                     const sky_render_config *cfg, const float *d_dl_dxyz,
                                     double *grad, void *stream, sky_render_stats *stats);

/* ---- ragged batch: many renders in one persistent launch ----
 * The reference renders one image per process and device
 * (scripts/bake_skyview_elevations.ps1:13-35 calls the CLI per elevation);
 * this runs a whole batch of independent renders -- each with its own
 * scene, resolution, wavelength set and spp -- as one wave-pool launch,
 * captured in a HIP graph and replayed when the batch layout repeats.
 * Outputs are host buffers per item (either may be NULL). */
typedef struct sky_batch_item {
    const sky_planet *planet;
    const sky_atmosphere *atm;
    const sky_wavelengths *wl;
    const sky_render_config *cfg;
    float *xyz;  /* [H*W*3]     */
    float *film; /* [H*W*count] */
} sky_batch_item;

typedef struct sky_batch_stats {
    double kernel_ms;      /* device time of the launch                      */
    double total_ms;       /* host wall time of the call                     */
    uint64_t path_samples; /* sum over renders of W * H * spp * count        */
    uint32_t renders, items, watchdog, graph_reused;
    double latency_p50_ms, latency_p99_ms, latency_max_ms; /* completion after launch */
} sky_batch_stats;

#define SKY_BATCH_NO_GRAPH 1u /* launch eagerly instead of through the cached graph */

/* finish_ms (may be NULL): per item, completion time after the launch began. */
sky_status sky_render_batch(sky_context *ctx, const sky_batch_item *items, uint32_t n, uint32_t flags,
                            double *finish_ms, sky_batch_stats *stats);

/* ---- deterministic optical depth (Gauss-Legendre per shell) ----
 * rays: n x 7 doubles (origin xyz km, direction xyz, t_max km);
 * tau: n x count doubles. */
sky_status sky_optical_depth(sky_context *ctx, const sky_planet *planet, const sky_atmosphere *atm,
                             const sky_wavelengths *wl, const double *rays, uint32_t n,
                             double *tau);

/* Thread-local message of the last error (RenderError's String). */
const char *sky_last_error(void);
const char *sky_version(void);

#ifdef __cplusplus
}
#endif
#endif
