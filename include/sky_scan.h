/*
 * sky_scan.h -- C ABI of the sparse-wavelength combo scan (libsky_hip.so).
 *
 * Drop-in for the evaluation loop of the reference's wavelength optimizer
 * (crates/sky-wavelength-optimizer/src/main.rs): given spectral sky-view LUT
 * samples on a band grid and a set of band combos with quadrature weights,
 * score every combo by the Oklab distance between the combo's XYZ quadrature
 * and the full-grid reference colour.
 *
 *   evaluate_combo      main.rs:414-488  -> sky_scan_combos (screening: mean,
 *                                           RMSE) and sky_refine_combos (exact
 *                                           p50/p90/p95/p99/max as well)
 *   quadrature_weights  main.rs:735-746  -> sky_quadrature_weights (host)
 *   band_widths_from_centers main.rs:714-733 -> sky_band_widths (host)
 *   cie_1931_2deg       spectrum.rs:196-219 -> sky_cie_1931_2deg (host)
 *
 * Per sample s and combo (b_0..b_{n-1}, coefficients c_k):
 *   xyz = sum_k band[s][b_k] * c_k,   c_k = weight_k / width_{b_k} * cmf(b_k) / scale
 *   error = |oklab(xyz) - reference_oklab[s]|
 * (main.rs:443-458); the caller folds the reference's 1/reference_scale
 * (main.rs:337-341) into c_k.  Arithmetic is FP32 per sample with FP64
 * accumulation; the reference computes in FP64.
 *
 * Errors are reported through sky_last_error() (sky_hip.h).  A dataset is
 * single-threaded like a context.
 */
#ifndef SKY_SCAN_H
#define SKY_SCAN_H

#include <stdint.h>

#include "sky_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

#define SKY_SCAN_MAX_PRIMARIES 8 /* MAX_PRIMARY_COUNT, main.rs:38 */
#define SKY_SCAN_MAX_BANDS 64

/* WeightStrategy (main.rs:44-53). */
#define SKY_WEIGHTS_VORONOI 0u
#define SKY_WEIGHTS_CMF_MASS 1u
#define SKY_WEIGHTS_CMF_MOMENTS 2u

/* Device-resident samples: band values and the reference Oklab colour. */
typedef struct sky_scan_dataset sky_scan_dataset;

/* band_values: [n_samples][n_bands] host floats (the per-band EXR values of
 * every LUT pixel, main.rs:301-316); reference_oklab: [n_samples][3] host
 * floats (main.rs:337-341).  n_bands <= SKY_SCAN_MAX_BANDS, n_samples < 2^32. */
sky_status sky_scan_dataset_create(int device, const float *band_values, uint32_t n_bands, uint64_t n_samples,
                                   const float *reference_oklab, sky_scan_dataset **out);
void sky_scan_dataset_destroy(sky_scan_dataset *ds);

typedef struct sky_scan_stats {
    double kernel_ms; /* device time of the scan kernels */
    double total_ms;  /* host wall time of the call, copies included */
    uint64_t evaluations; /* combos x samples */
} sky_scan_stats;

/* Screening pass (main.rs:414-488 with exact_percentiles = false).
 * combos: [n_combos][count] band indices (ascending, distinct, < n_bands);
 * coef:   [n_combos][count][3] c_k as above;
 * samples: indices into the dataset, or NULL for all samples;
 * sum, sum_sq: [n_combos] sum of errors and of squared errors (host). */
sky_status sky_scan_combos(sky_scan_dataset *ds, const uint8_t *combos, const float *coef, uint32_t count,
                           uint64_t n_combos, const uint32_t *samples, uint64_t n_samples, double *sum,
                           double *sum_sq, sky_scan_stats *stats);

/* Same with device pointers (e.g. torch tensors) for combos, coef, samples and
 * the sums, ordered on `stream` (hipStream_t, NULL = the dataset's stream)
 * and synchronous on return.  Out-of-range band or sample indices are caught
 * on the device (clamped, then reported as SKY_ERR_INVALID_CONFIG). */
sky_status sky_scan_combos_device(sky_scan_dataset *ds, const uint8_t *d_combos, const float *d_coef, uint32_t count,
                                  uint64_t n_combos, const uint32_t *d_samples, uint64_t n_samples, double *d_sum,
                                  double *d_sum_sq, void *stream, sky_scan_stats *stats);

/* ComboResult statistics (main.rs:97-110) of one combo. */
typedef struct sky_combo_stats {
    double mean, rmse, p50, p90, p95, p99, max;
    uint64_t sample_count;
} sky_combo_stats;

/* Refinement pass (exact_percentiles = true): mean, RMSE, the percentiles
 * of the finite errors by index round((n-1) q) of the sorted list
 * (main.rs:1107-1117) and the maximum, per combo.  n_combos <= 65536. */
sky_status sky_refine_combos(sky_scan_dataset *ds, const uint8_t *combos, const float *coef, uint32_t count,
                             uint64_t n_combos, const uint32_t *samples, uint64_t n_samples,
                             sky_combo_stats *out, sky_scan_stats *stats);

/* Host: Voronoi widths of a band grid (main.rs:714-733); n_bands >= 2. */
sky_status sky_band_widths(const float *wavelengths_nm, uint32_t n_bands, double *widths_nm);

/* Host: quadrature weights [n_combos][count] for each combo (main.rs:735-1070):
 * Voronoi widths, or the non-negative fixed-sum least-squares fit of the
 * CMF response (cmf_xyz: [n_bands][3], the colour-matching functions at the
 * band centres), falling back to Voronoi when no subset solves.  Runs on
 * up to `threads` host threads (0: all). */
sky_status sky_quadrature_weights(const float *wavelengths_nm, uint32_t n_bands, const double *cmf_xyz,
                                  const uint8_t *combos, uint32_t count, uint64_t n_combos, uint32_t strategy,
                                  uint32_t threads, double *weights);

/* Host: CIE 1931 2-degree colour-matching functions at n wavelengths,
 * xyz: [n][3] (cie_1931_2deg, spectrum.rs:196-219: clamped outside
 * 360-830 nm, linear between table points). */
sky_status sky_cie_1931_2deg(const float *lambda_nm, uint32_t n, double *xyz);

#ifdef __cplusplus
}
#endif

#endif
