/*
 * scan_oracle.h -- CPU oracle for the sparse-wavelength combo scan.
 *
 * TEST INFRASTRUCTURE ONLY: an FP64 restatement of the evaluation side of the
 * reference's wavelength optimizer (crates/sky-wavelength-optimizer/src/
 * main.rs) for tests and bench.py's CPU baseline.  The product never links,
 * loads or calls it.
 */
#ifndef SCAN_ORACLE_H
#define SCAN_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_SCAN_MAX_BANDS 64
#define ORC_WEIGHTS_VORONOI 0u
#define ORC_WEIGHTS_CMF_MASS 1u
#define ORC_WEIGHTS_CMF_MOMENTS 2u

void orc_band_widths(const float *wl, uint32_t n, double *w);
/* weights: [8]; returns 0 on success */
int orc_quadrature_weights(const float *wl, uint32_t n, const double *cie, const uint8_t *combo, uint32_t count,
                           uint32_t strategy, double *w);
void orc_xyz_to_oklab(const double *xyz, double *lab);
double orc_percentile(const double *values, uint64_t n, double q);
/* reference Oklab [n_samples][3]; returns the p99-of-Y scale */
double orc_reference_oklab(const float *values, uint32_t n_bands, uint64_t n_samples, const double *cie,
                           double *ref_oklab);
/* out[8]: mean, rmse, p50, p90, p95, p99, max, score */
void orc_evaluate_combo(const float *values, uint32_t n_bands, uint64_t n_samples, const double *ref_oklab,
                        double scale, const float *wl, const double *cie, const uint8_t *combo, uint32_t count,
                        const double *weights, const uint32_t *idx, uint64_t n_idx, int exact, double *out);
void orc_evaluate_combos(const float *values, uint32_t n_bands, uint64_t n_samples, const double *ref_oklab,
                         double scale, const float *wl, const double *cie, const uint8_t *combos, uint32_t count,
                         uint64_t n_combos, const double *weights, const uint32_t *idx, uint64_t n_idx, int exact,
                         int threads, double *out);

#ifdef __cplusplus
}
#endif

#endif
