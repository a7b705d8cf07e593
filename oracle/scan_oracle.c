/*
 * scan_oracle.c -- CPU oracle for the sparse-wavelength combo scan.
 *
 * TEST INFRASTRUCTURE ONLY.  Plain-C FP64 restatement of the evaluation side
 * of the reference's wavelength optimizer
 * (crates/sky-wavelength-optimizer/src/main.rs) and of the CIE table lookup
 * it uses, for tests and the CPU-baseline leg of bench.py.  Nothing in the
 * product links, loads or calls this code.
 */
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#include "scan_oracle.h"

#define MAXP 8
#define MAXROWS 9
#define KKT (MAXP + 1)

/* main.rs:714-733 (differences and midpoints formed in f32, then widened) */
void orc_band_widths(const float *wl, uint32_t n, double *w)
{
    for (uint32_t i = 0; i < n; ++i) {
        double lo, hi;
        if (i == 0) {
            float d = wl[1] - wl[0];
            lo = (double)wl[0] - 0.5 * (double)d;
        } else {
            float m = wl[i - 1] + wl[i];
            lo = 0.5 * (double)m;
        }
        if (i + 1 == n) {
            float d = wl[i] - wl[i - 1];
            hi = (double)wl[i] + 0.5 * (double)d;
        } else {
            float m = wl[i] + wl[i + 1];
            hi = 0.5 * (double)m;
        }
        w[i] = hi - lo;
    }
}

typedef struct {
    const float *wl;
    uint32_t n;
    double width[ORC_SCAN_MAX_BANDS];
    double lo, hi;
} grid_t;

static void grid_init(grid_t *g, const float *wl, uint32_t n)
{
    g->wl = wl;
    g->n = n;
    orc_band_widths(wl, n, g->width);
    g->lo = (double)wl[0] - 0.5 * g->width[0];                 /* main.rs:754-755 */
    g->hi = (double)wl[n - 1] + 0.5 * g->width[n - 1];
}

/* main.rs:748-784 */
static void voronoi(const grid_t *g, const uint8_t *c, uint32_t count, double *w)
{
    double centers[MAXP], splits[MAXP];
    for (int k = 0; k < MAXP; ++k) w[k] = 0.0;
    for (uint32_t k = 0; k < count; ++k) centers[k] = (double)g->wl[c[k]];
    if (count == 1) {
        w[0] = g->hi - g->lo;
        return;
    }
    for (uint32_t k = 0; k + 1 < count; ++k) splits[k] = 0.5 * (centers[k] + centers[k + 1]);
    w[0] = splits[0] - g->lo;
    for (uint32_t k = 1; k + 1 < count; ++k) w[k] = splits[k] - splits[k - 1];
    w[count - 1] = g->hi - splits[count - 2];
}

/* main.rs:921-955: basis (1, moment, moment^2) times the CMF */
static void target_rows(const grid_t *g, int rows, double lambda, const double *cie, double *out)
{
    double center = 0.5 * (g->lo + g->hi);
    double half = 0.5 * (g->hi - g->lo);
    double m = (lambda - center) / (half > 1.0e-6 ? half : 1.0e-6);
    double basis[3] = {1.0, m, m * m};
    for (int b = 0; b < rows / 3; ++b)
        for (int c = 0; c < 3; ++c) out[b * 3 + c] = basis[b] * cie[c];
}

/* main.rs:1036-1077: Gauss-Jordan with partial pivoting */
static int solve_linear(double a[KKT][KKT], double *rhs, int dim)
{
    for (int p = 0; p < dim; ++p) {
        int best = p;
        double bv = fabs(a[p][p]);
        for (int r = p + 1; r < dim; ++r) {
            double v = fabs(a[r][p]);
            if (v > bv) {
                bv = v;
                best = r;
            }
        }
        if (bv <= 1.0e-12) return 0;
        if (best != p) {
            for (int c = 0; c < KKT; ++c) {
                double t = a[p][c];
                a[p][c] = a[best][c];
                a[best][c] = t;
            }
            double t = rhs[p];
            rhs[p] = rhs[best];
            rhs[best] = t;
        }
        double pv = a[p][p];
        for (int c = p; c < dim; ++c) a[p][c] /= pv;
        rhs[p] /= pv;
        for (int r = 0; r < dim; ++r) {
            if (r == p) continue;
            double f = a[r][p];
            if (fabs(f) <= 1.0e-18) continue;
            for (int c = p; c < dim; ++c) a[r][c] -= f * a[p][c];
            rhs[r] -= f * rhs[p];
        }
    }
    return 1;
}

/* main.rs:786-897, 957-1034, 1079-1094 */
static int cmf_fit(const grid_t *g, const double *cie, const uint8_t *c, uint32_t count, int moments, double *w)
{
    int rows = moments ? 9 : 3;
    double total = g->hi - g->lo;
    double col[MAXROWS][MAXP];
    memset(col, 0, sizeof col);
    for (uint32_t k = 0; k < count; ++k) {
        double r[MAXROWS];
        target_rows(g, rows, (double)g->wl[c[k]], cie + 3 * c[k], r);
        for (int i = 0; i < rows; ++i) col[i][k] = r[i];
    }
    double target[MAXROWS] = {0};
    for (uint32_t b = 0; b < g->n; ++b) {
        double r[MAXROWS];
        target_rows(g, rows, (double)g->wl[b], cie + 3 * b, r);
        for (int i = 0; i < rows; ++i) target[i] += g->width[b] * r[i];
    }
    int found = 0;
    double best = INFINITY;
    for (uint32_t mask = 1; mask < (1u << count); ++mask) {
        int sub[MAXP], ns = 0;
        for (uint32_t k = 0; k < count; ++k)
            if (mask & (1u << k)) sub[ns++] = (int)k;
        double a[KKT][KKT], rhs[KKT];
        memset(a, 0, sizeof a);
        memset(rhs, 0, sizeof rhs);
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
        if (!solve_linear(a, rhs, ns + 1)) continue;
        double x[MAXP] = {0};
        for (int i = 0; i < ns; ++i) x[sub[i]] = rhs[i];
        int neg = 0;
        for (uint32_t k = 0; k < count; ++k)
            if (x[k] < -1.0e-7 * total) neg = 1;
        if (neg) continue;
        for (uint32_t k = 0; k < count; ++k)
            if (x[k] < 0.0) x[k] = 0.0;
        double err = 0.0;
        for (int r = 0; r < rows; ++r) {
            double v = 0.0;
            for (int k = 0; k < MAXP; ++k) v += col[r][k] * x[k];
            double d = v - target[r];
            err += d * d;
        }
        if (err < best) {
            best = err;
            memcpy(w, x, sizeof x);
            found = 1;
        }
    }
    return found;
}

int orc_quadrature_weights(const float *wl, uint32_t n, const double *cie, const uint8_t *combo, uint32_t count,
                           uint32_t strategy, double *w)
{
    if (n < 2 || n > ORC_SCAN_MAX_BANDS || count < 1 || count > MAXP) return 1;
    grid_t g;
    grid_init(&g, wl, n);
    if (strategy == ORC_WEIGHTS_VORONOI || !cmf_fit(&g, cie, combo, count, strategy == ORC_WEIGHTS_CMF_MOMENTS, w))
        voronoi(&g, combo, count, w);
    return 0;
}

/* main.rs:1088-1098 */
void orc_xyz_to_oklab(const double *xyz, double *lab)
{
    double l = cbrt(0.8189330101 * xyz[0] + 0.3618667424 * xyz[1] - 0.1288597137 * xyz[2]);
    double m = cbrt(0.0329845436 * xyz[0] + 0.9293118715 * xyz[1] + 0.0361456387 * xyz[2]);
    double s = cbrt(0.0482003018 * xyz[0] + 0.2643662691 * xyz[1] + 0.633851707 * xyz[2]);
    lab[0] = 0.2104542553 * l + 0.793617785 * m - 0.0040720468 * s;
    lab[1] = 1.9779984951 * l - 2.428592205 * m + 0.4505937099 * s;
    lab[2] = 0.0259040371 * l + 0.7827717662 * m - 0.808675766 * s;
}

static int cmp_double(const void *a, const void *b)
{
    double x = *(const double *)a, y = *(const double *)b;
    return x < y ? -1 : x > y ? 1 : 0;
}

/* main.rs:1107-1117: finite values sorted, index round((n-1) q) */
double orc_percentile(const double *values, uint64_t n, double q)
{
    double *v = (double *)malloc((n ? n : 1) * sizeof(double));
    uint64_t m = 0;
    for (uint64_t i = 0; i < n; ++i)
        if (isfinite(values[i])) v[m++] = values[i];
    double r = NAN;
    if (m) {
        qsort(v, m, sizeof(double), cmp_double);
        if (q < 0.0) q = 0.0;
        if (q > 1.0) q = 1.0;
        r = v[(uint64_t)round((double)(m - 1) * q)];
    }
    free(v);
    return r;
}

/* main.rs:276-341: reference XYZ per sample, p99 Y scale, reference Oklab */
double orc_reference_oklab(const float *values, uint32_t n_bands, uint64_t n_samples, const double *cie,
                           double *ref_oklab)
{
    double *xyz = (double *)calloc(3 * (n_samples ? n_samples : 1), sizeof(double));
    double *y = (double *)calloc(n_samples ? n_samples : 1, sizeof(double));
    for (uint64_t s = 0; s < n_samples; ++s) {
        double a[3] = {0.0, 0.0, 0.0};
        for (uint32_t b = 0; b < n_bands; ++b) {
            double v = (double)values[s * n_bands + b];
            for (int c = 0; c < 3; ++c) a[c] += v * cie[3 * b + c];
        }
        memcpy(xyz + 3 * s, a, sizeof a);
        y[s] = a[1];
    }
    double scale = orc_percentile(y, n_samples, 0.99);
    if (!(scale > 1.0e-12)) scale = 1.0e-12; /* percentile(..).max(1e-12) */
    for (uint64_t s = 0; s < n_samples; ++s) {
        double v[3] = {xyz[3 * s] / scale, xyz[3 * s + 1] / scale, xyz[3 * s + 2] / scale};
        orc_xyz_to_oklab(v, ref_oklab + 3 * s);
    }
    free(xyz);
    free(y);
    return scale;
}

/* main.rs:414-488 (evaluate_combo) */
void orc_evaluate_combo(const float *values, uint32_t n_bands, uint64_t n_samples, const double *ref_oklab,
                        double scale, const float *wl, const double *cie, const uint8_t *combo, uint32_t count,
                        const double *weights, const uint32_t *idx, uint64_t n_idx, int exact, double *out)
{
    double width[ORC_SCAN_MAX_BANDS];
    orc_band_widths(wl, n_bands, width);
    const uint64_t n = idx ? n_idx : n_samples;
    double *errors = exact ? (double *)malloc((n ? n : 1) * sizeof(double)) : NULL;
    double sum = 0.0, sq = 0.0, mx = -INFINITY;
    for (uint64_t i = 0; i < n; ++i) {
        const uint64_t s = idx ? idx[i] : i;
        double xyz[3] = {0.0, 0.0, 0.0};
        for (uint32_t k = 0; k < count; ++k) {
            double v = (double)values[s * n_bands + combo[k]] * weights[k] / width[combo[k]];
            for (int c = 0; c < 3; ++c) xyz[c] += v * cie[3 * combo[k] + c];
        }
        double lab[3], x[3] = {xyz[0] / scale, xyz[1] / scale, xyz[2] / scale};
        orc_xyz_to_oklab(x, lab);
        const double *r = ref_oklab + 3 * s;
        double dl = lab[0] - r[0], da = lab[1] - r[1], db = lab[2] - r[2];
        double e = sqrt(dl * dl + da * da + db * db);
        sum += e;
        sq += e * e;
        if (errors) {
            errors[i] = e;
            mx = fmax(mx, e); /* f64::max ignores NaN */
        }
    }
    out[0] = sum / (double)n;
    out[1] = sqrt(sq / (double)n);
    if (exact) {
        out[2] = orc_percentile(errors, n, 0.50);
        out[3] = orc_percentile(errors, n, 0.90);
        out[4] = orc_percentile(errors, n, 0.95);
        out[5] = orc_percentile(errors, n, 0.99);
        out[6] = mx;
        out[7] = out[1] + 0.25 * out[4];
        free(errors);
    } else {
        for (int k = 2; k < 7; ++k) out[k] = NAN;
        out[7] = out[1];
    }
}

/* evaluate_combos over many combos, screening form, on host threads */
void orc_evaluate_combos(const float *values, uint32_t n_bands, uint64_t n_samples, const double *ref_oklab,
                         double scale, const float *wl, const double *cie, const uint8_t *combos, uint32_t count,
                         uint64_t n_combos, const double *weights, const uint32_t *idx, uint64_t n_idx, int exact,
                         int threads, double *out)
{
    if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 4) num_threads(threads)
    for (int64_t c = 0; c < (int64_t)n_combos; ++c)
        orc_evaluate_combo(values, n_bands, n_samples, ref_oklab, scale, wl, cie, combos + (uint64_t)c * count, count,
                           weights + (uint64_t)c * count, idx, n_idx, exact, out + 8 * (uint64_t)c);
}
