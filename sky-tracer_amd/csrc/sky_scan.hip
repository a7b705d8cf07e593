// sky_scan.hip -- sparse-wavelength combo scan (include/sky_scan.h).
//
// The reference's optimizer scores each combo on the CPU, one combo per
// thread, in FP64 (crates/sky-wavelength-optimizer/src/main.rs:414-530).
// Here one lane owns one combo for the whole scan: its band indices and
// coefficients live in registers, the samples stream through LDS in tiles
// shared by the 256 combos of a workgroup, and the per-sample work is three
// cube roots and a quadratic form, with no reduction until the end.
//
// Two reformulations keep the per-sample work small and are exact in real
// arithmetic:
//   * the XYZ -> LMS matrix of Oklab is folded into the per-channel
//     coefficients (LMS = sum_k v_k M1 c_k);
//   * the reference colour is held as LMS^(1/3) and the Oklab distance is
//     |M2 d| = sqrt(d^T (M2^T M2) d) with d the difference of cube roots.
// The refinement pass writes every error of a chunk of combos and selects the
// percentiles exactly with a three-level radix select on the float bits.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/sky_scan.h"
#include "sky_internal.hpp"

namespace {

constexpr int kWG = 256;                // combos per workgroup (one per lane)
constexpr int kTile = 128;              // samples per LDS tile
constexpr int kTileStride = kTile + 1;  // odd: band rows fall on distinct banks
constexpr int kSelectWG = 1024;
constexpr int kBins = 4096;

// Oklab (B. Ottosson, "A perceptual color space for image processing", 2020):
// M1 maps XYZ to LMS, M2 maps the cube roots of LMS to Lab (main.rs:1088-1098).
constexpr double kM1[3][3] = {{0.8189330101, 0.3618667424, -0.1288597137},
                              {0.0329845436, 0.9293118715, 0.0361456387},
                              {0.0482003018, 0.2643662691, 0.6338517070}};
constexpr double kM2[3][3] = {{0.2104542553, 0.7936177850, -0.0040720468},
                              {1.9779984951, -2.4285922050, 0.4505937099},
                              {0.0259040371, 0.7827717662, -0.8086757660}};

struct QuadForm {  // M2^T M2, off-diagonal terms doubled
    float q00, q11, q22, q01, q02, q12;
};

struct ScanArgs {
    const float *rows;       // [n_total][R]: band values, then the reference LMS^(1/3)
    const uint32_t *idx;     // samples to scan (null: 0..n-1)
    uint64_t n;              // number of samples scanned
    const uint8_t *combos;   // [n_combos][N]
    const float *coef;       // [n_combos][N][3] XYZ coefficients
    uint64_t n_combos;
    uint32_t R, B;
    uint64_t n_total;
    QuadForm q;
    double *sum, *sq;
    uint32_t *bad;           // set when a band or sample index is out of range
};

struct ErrArgs {
    const float *cols;       // [R][n_total]
    uint64_t n_total;
    const uint32_t *idx;
    uint64_t n;
    const uint8_t *combos;
    const float *coef;
    uint32_t R, B, blocks_per_combo;
    QuadForm q;
    float *err;              // [chunk][n]
    uint32_t *bad;
};

// LMS coefficients of one channel: M1 c (the XYZ -> LMS matrix folded in)
__device__ __forceinline__ void lms_coef(const float *c, float &l, float &m, float &s)
{
    l = fmaf((float)kM1[0][2], c[2], fmaf((float)kM1[0][1], c[1], (float)kM1[0][0] * c[0]));
    m = fmaf((float)kM1[1][2], c[2], fmaf((float)kM1[1][1], c[1], (float)kM1[1][0] * c[0]));
    s = fmaf((float)kM1[2][2], c[2], fmaf((float)kM1[2][1], c[1], (float)kM1[2][0] * c[0]));
}

__device__ __forceinline__ float cbrt_f(float x)
{
    float r = __builtin_amdgcn_exp2f(__builtin_amdgcn_logf(fabsf(x)) * (1.0f / 3.0f));
    return copysignf(r, x);
}

// Squared Oklab distance from the candidate LMS to the reference cube roots.
// -0 and rounding below zero become +0 (a NaN stays NaN).
__device__ __forceinline__ float dist_sq(float l, float m, float s, float rl, float rm, float rs, const QuadForm &q)
{
    float d0 = cbrt_f(l) - rl, d1 = cbrt_f(m) - rm, d2 = cbrt_f(s) - rs;
    float t0 = fmaf(d2, q.q02, fmaf(d1, q.q01, d0 * q.q00));
    float t1 = fmaf(d2, q.q12, d1 * q.q11);
    float e2 = fmaf(d2, d2 * q.q22, fmaf(d1, t1, d0 * t0));
    return (e2 < 0.0f ? 0.0f : e2) + 0.0f;
}

template <int N>
__global__ __launch_bounds__(kWG) void k_scan(ScanArgs a)
{
    extern __shared__ float tile[];  // [R][kTileStride]
    const uint64_t c = (uint64_t)blockIdx.x * kWG + threadIdx.x;
    const bool live = c < a.n_combos;
    uint32_t addr[N];
    float cl[N], cm[N], cs[N];
#pragma unroll
    for (int k = 0; k < N; ++k) {
        uint32_t b = live ? a.combos[c * N + k] : 0u;
        if (b >= a.B) {
            atomicOr(a.bad, 1u);
            b = 0;
        }
        addr[k] = b * kTileStride;
        cl[k] = cm[k] = cs[k] = 0.0f;
        if (live) lms_coef(a.coef + (c * N + k) * 3, cl[k], cm[k], cs[k]);
    }
    const uint32_t ref = (a.R - 3) * kTileStride;
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    double sum = 0.0, sq = 0.0;
    for (uint64_t base = 0; base < a.n; base += kTile) {
        const uint32_t nt = (uint32_t)min((uint64_t)kTile, a.n - base);
        __syncthreads();
        for (uint32_t r = wave; r < nt; r += kWG / 64) {
            uint64_t s = a.idx ? (uint64_t)a.idx[base + r] : base + r;
            if (s >= a.n_total) {
                atomicOr(a.bad, 2u);
                s = 0;
            }
            const float *row = a.rows + s * a.R;
            for (uint32_t col = lane; col < a.R; col += 64) tile[col * kTileStride + r] = row[col];
        }
        __syncthreads();
        if (live) {
            float ts = 0.0f, tq = 0.0f;
#pragma unroll 4
            for (uint32_t r = 0; r < nt; ++r) {
                float l = 0.0f, m = 0.0f, s = 0.0f;
#pragma unroll
                for (int k = 0; k < N; ++k) {
                    float v = tile[addr[k] + r];
                    l = fmaf(v, cl[k], l);
                    m = fmaf(v, cm[k], m);
                    s = fmaf(v, cs[k], s);
                }
                float e2 = dist_sq(l, m, s, tile[ref + r], tile[ref + kTileStride + r],
                                   tile[ref + 2 * kTileStride + r], a.q);
                ts += __builtin_amdgcn_sqrtf(e2);
                tq += e2;
            }
            sum += ts;
            sq += tq;
        }
    }
    if (live) {
        a.sum[c] = sum;
        a.sq[c] = sq;
    }
}

// Every error of one combo per block row: lanes are samples, the combo's
// indices and coefficients are uniform.
template <int N>
__global__ __launch_bounds__(256) void k_errors(ErrArgs a)
{
    const uint32_t combo = blockIdx.x / a.blocks_per_combo;
    const uint64_t i = (uint64_t)(blockIdx.x - combo * a.blocks_per_combo) * 256 + threadIdx.x;
    if (i >= a.n) return;
    uint64_t s = a.idx ? (uint64_t)a.idx[i] : i;
    if (s >= a.n_total) {
        atomicOr(a.bad, 2u);
        s = 0;
    }
    float l = 0.0f, m = 0.0f, t = 0.0f;
#pragma unroll
    for (int k = 0; k < N; ++k) {
        uint32_t b = a.combos[(uint64_t)combo * N + k];
        if (b >= a.B) {
            atomicOr(a.bad, 1u);
            b = 0;
        }
        float cl, cm, cs;
        lms_coef(a.coef + ((uint64_t)combo * N + k) * 3, cl, cm, cs);
        float v = a.cols[b * a.n_total + s];
        l = fmaf(v, cl, l);
        m = fmaf(v, cm, m);
        t = fmaf(v, cs, t);
    }
    const uint32_t rb = a.R - 3;
    float e2 = dist_sq(l, m, t, a.cols[rb * a.n_total + s], a.cols[(rb + 1) * a.n_total + s],
                       a.cols[(rb + 2) * a.n_total + s], a.q);
    a.err[(uint64_t)combo * a.n + i] = __builtin_amdgcn_sqrtf(e2);
}

// Finds, for each query q, the bin holding the element of rank rank[q]
// (0-based, ascending) of histogram h[q]; rank[q] becomes the rank within the
// bin.  Run by wave 0.
__device__ void find_bins(uint32_t (*h)[kBins], int shared_hist, uint32_t nbins, uint32_t *rank, uint32_t *bin,
                          const bool *valid)
{
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t per = nbins / 64;
    for (int q = 0; q < 4; ++q) {
        if (!valid[q]) continue;
        const uint32_t *hq = h[shared_hist ? 0 : q];
        uint32_t own = 0;
        for (uint32_t j = 0; j < per; ++j) own += hq[lane * per + j];
        uint32_t incl = own;
        for (int d = 1; d < 64; d <<= 1) {
            uint32_t o = __shfl_up(incl, d, 64);
            if (lane >= (uint32_t)d) incl += o;
        }
        const uint32_t target = rank[q];
        const uint64_t mask = __ballot(incl > target);
        const uint32_t owner = (uint32_t)__ffsll((unsigned long long)mask) - 1u;
        uint32_t cum = incl - own;
        uint32_t my_bin = 0, my_rank = 0;
        if (lane == owner) {
            for (uint32_t j = 0; j < per; ++j) {
                uint32_t v = hq[lane * per + j];
                if (cum + v > target) {
                    my_bin = lane * per + j;
                    my_rank = target - cum;
                    break;
                }
                cum += v;
            }
        }
        my_bin = __shfl(my_bin, (int)owner, 64);
        my_rank = __shfl(my_rank, (int)owner, 64);
        if (lane == 0) {
            bin[q] = my_bin;
            rank[q] = my_rank;
        }
    }
}

// One workgroup per combo: sum, sum of squares and maximum, then the four
// percentiles by radix select over the bits of the finite errors (all >= 0,
// so their bit patterns order like their values): bits 30..19, 18..7, 6..0.
__global__ __launch_bounds__(kSelectWG) void k_select(const float *err, uint64_t n, sky_combo_stats *out)
{
    __shared__ uint32_t hist[4][kBins];
    __shared__ double red_s[kSelectWG / 64], red_q[kSelectWG / 64];
    __shared__ float red_m[kSelectWG / 64];
    __shared__ uint32_t red_n[kSelectWG / 64];
    __shared__ uint32_t rank[4], bin1[4], bin2[4], bin3[4];
    __shared__ bool valid[4];
    const float *e = err + (uint64_t)blockIdx.x * n;
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;

    for (uint32_t j = tid; j < kBins; j += kSelectWG) hist[0][j] = 0;
    __syncthreads();
    double s = 0.0, q = 0.0;
    float mx = -INFINITY;
    uint32_t nf = 0;
    for (uint64_t i = tid; i < n; i += kSelectWG) {
        const float v = e[i];
        s += v;
        q += (double)v * v;
        if (!(v != v)) mx = fmaxf(mx, v);
        const uint32_t u = __float_as_uint(v);
        if (u < 0x7F800000u) {
            atomicAdd(&hist[0][u >> 19], 1u);
            ++nf;
        }
    }
    for (int d = 32; d > 0; d >>= 1) {
        s += __shfl_down(s, d, 64);
        q += __shfl_down(q, d, 64);
        mx = fmaxf(mx, __shfl_down(mx, d, 64));
        nf += __shfl_down(nf, d, 64);
    }
    if (lane == 0) {
        red_s[wave] = s;
        red_q[wave] = q;
        red_m[wave] = mx;
        red_n[wave] = nf;
    }
    __syncthreads();
    if (tid == 0) {
        double ts = 0.0, tq = 0.0;
        float tm = -INFINITY;
        uint32_t tn = 0;
        for (int w = 0; w < kSelectWG / 64; ++w) {
            ts += red_s[w];
            tq += red_q[w];
            tm = fmaxf(tm, red_m[w]);
            tn += red_n[w];
        }
        sky_combo_stats &o = out[blockIdx.x];
        o.mean = ts / (double)n;
        o.rmse = sqrt(tq / (double)n);
        o.max = tm;
        o.sample_count = n;
        const double qs[4] = {0.50, 0.90, 0.95, 0.99};
        for (int k = 0; k < 4; ++k) {
            valid[k] = tn > 0;
            rank[k] = tn > 0 ? (uint32_t)round((double)(tn - 1) * qs[k]) : 0u;  // main.rs:1111-1117
        }
    }
    __syncthreads();
    if (wave == 0) find_bins(hist, 1, kBins, rank, bin1, valid);
    __syncthreads();

    for (uint32_t j = tid; j < 4 * kBins; j += kSelectWG) (&hist[0][0])[j] = 0;
    __syncthreads();
    for (uint64_t i = tid; i < n; i += kSelectWG) {
        const uint32_t u = __float_as_uint(e[i]);
        if (u >= 0x7F800000u) continue;
        for (int k = 0; k < 4; ++k)
            if (valid[k] && (u >> 19) == bin1[k]) atomicAdd(&hist[k][(u >> 7) & 4095u], 1u);
    }
    __syncthreads();
    if (wave == 0) find_bins(hist, 0, kBins, rank, bin2, valid);
    __syncthreads();

    for (uint32_t j = tid; j < 4 * kBins; j += kSelectWG) (&hist[0][0])[j] = 0;
    __syncthreads();
    for (uint64_t i = tid; i < n; i += kSelectWG) {
        const uint32_t u = __float_as_uint(e[i]);
        if (u >= 0x7F800000u) continue;
        for (int k = 0; k < 4; ++k)
            if (valid[k] && (u >> 7) == ((bin1[k] << 12) | bin2[k])) atomicAdd(&hist[k][u & 127u], 1u);
    }
    __syncthreads();
    if (wave == 0) find_bins(hist, 0, 128, rank, bin3, valid);
    __syncthreads();
    if (tid == 0) {
        sky_combo_stats &o = out[blockIdx.x];
        double *p[4] = {&o.p50, &o.p90, &o.p95, &o.p99};
        for (int k = 0; k < 4; ++k)
            *p[k] = valid[k] ? (double)__uint_as_float((bin1[k] << 19) | (bin2[k] << 7) | bin3[k]) : (double)NAN;
    }
}

struct Buf {
    void *p = nullptr;
    size_t n = 0;
    hipError_t ensure(size_t bytes)
    {
        if (bytes <= n) return hipSuccess;
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
        hipError_t e = hipMalloc(&p, bytes);
        if (e == hipSuccess) n = bytes;
        return e;
    }
    ~Buf()
    {
        if (p) (void)hipFree(p);
    }
};

sky_status hip_err(hipError_t e, const char *what)
{
    return sky::set_error(e == hipErrorOutOfMemory ? SKY_ERR_OOM : SKY_ERR_SYNC,
                          std::string(what) + ": " + hipGetErrorString(e));
}

#define SCAN_TRY(expr, what)                        \
    do {                                            \
        hipError_t e_ = (expr);                     \
        if (e_ != hipSuccess) return hip_err(e_, what); \
    } while (0)

QuadForm quad_form()
{
    double m[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) m[i][j] = kM2[0][i] * kM2[0][j] + kM2[1][i] * kM2[1][j] + kM2[2][i] * kM2[2][j];
    return QuadForm{(float)m[0][0], (float)m[1][1], (float)m[2][2], (float)(2.0 * m[0][1]), (float)(2.0 * m[0][2]),
                    (float)(2.0 * m[1][2])};
}

void inverse3(const double m[3][3], double o[3][3])
{
    double det = m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) - m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
                 m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
    o[0][0] = (m[1][1] * m[2][2] - m[1][2] * m[2][1]) / det;
    o[0][1] = (m[0][2] * m[2][1] - m[0][1] * m[2][2]) / det;
    o[0][2] = (m[0][1] * m[1][2] - m[0][2] * m[1][1]) / det;
    o[1][0] = (m[1][2] * m[2][0] - m[1][0] * m[2][2]) / det;
    o[1][1] = (m[0][0] * m[2][2] - m[0][2] * m[2][0]) / det;
    o[1][2] = (m[0][2] * m[1][0] - m[0][0] * m[1][2]) / det;
    o[2][0] = (m[1][0] * m[2][1] - m[1][1] * m[2][0]) / det;
    o[2][1] = (m[0][1] * m[2][0] - m[0][0] * m[2][1]) / det;
    o[2][2] = (m[0][0] * m[1][1] - m[0][1] * m[1][0]) / det;
}

}  // namespace

struct sky_scan_dataset {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    uint32_t B = 0, R = 0;
    uint64_t n = 0;
    Buf rows, cols;  // [n][R] and [R][n]
    Buf combos, coef, idx, sum, sq, err, stats, bad;
    QuadForm q{};
    ~sky_scan_dataset()
    {
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

sky_status check_count(uint32_t count)
{
    if (count < 1 || count > SKY_SCAN_MAX_PRIMARIES)
        return sky::set_error(SKY_ERR_INVALID_CONFIG, "combo size must be in 1..=8");
    return SKY_OK;
}

// Validates the host combo list and uploads combos, coefficients and sample
// indices; returns the number of samples scanned.
sky_status upload_combos(sky_scan_dataset *ds, const uint8_t *combos, const float *coef, uint32_t count,
                         uint64_t n_combos, const uint32_t *samples, uint64_t n_samples, uint64_t *n_scan)
{
    if (!ds || !combos || !coef) return sky::set_error(SKY_ERR_INVALID_CONFIG, "null dataset, combos or coefficients");
    if (check_count(count) != SKY_OK) return SKY_ERR_INVALID_CONFIG;
    for (uint64_t i = 0; i < n_combos * count; ++i)
        if (combos[i] >= ds->B) return sky::set_error(SKY_ERR_INVALID_CONFIG, "combo band index out of range");
    if (samples) {
        if (n_samples == 0) return sky::set_error(SKY_ERR_INVALID_CONFIG, "empty sample list");
        for (uint64_t i = 0; i < n_samples; ++i)
            if (samples[i] >= ds->n) return sky::set_error(SKY_ERR_INVALID_CONFIG, "sample index out of range");
    }
    *n_scan = samples ? n_samples : ds->n;
    const size_t coef_bytes = n_combos * count * 3 * sizeof(float);
    SCAN_TRY(ds->combos.ensure(n_combos * count), "alloc combos");
    SCAN_TRY(ds->coef.ensure(coef_bytes), "alloc coefficients");
    SCAN_TRY(hipMemcpyAsync(ds->combos.p, combos, n_combos * count, hipMemcpyHostToDevice, ds->stream), "upload combos");
    SCAN_TRY(hipMemcpyAsync(ds->coef.p, coef, coef_bytes, hipMemcpyHostToDevice, ds->stream), "upload coefficients");
    if (samples) {
        SCAN_TRY(ds->idx.ensure(n_samples * sizeof(uint32_t)), "alloc sample list");
        SCAN_TRY(hipMemcpyAsync(ds->idx.p, samples, n_samples * sizeof(uint32_t), hipMemcpyHostToDevice, ds->stream),
                 "upload sample list");
    }
    return SKY_OK;
}

template <int N>
void launch_scan(const ScanArgs &a, size_t lds, hipStream_t st)
{
    const uint64_t blocks = (a.n_combos + kWG - 1) / kWG;
    hipLaunchKernelGGL(k_scan<N>, dim3((uint32_t)blocks), dim3(kWG), lds, st, a);
}

template <int N>
void launch_errors(const ErrArgs &a, uint32_t chunk, hipStream_t st)
{
    hipLaunchKernelGGL(k_errors<N>, dim3(chunk * a.blocks_per_combo), dim3(256), 0, st, a);
}

// Clears the index-error flag, records ev0 and launches the screening kernel
// on `st` (device pointers throughout), then records ev1.
sky_status launch_screen(sky_scan_dataset *ds, const uint8_t *d_combos, const float *d_coef, uint32_t count,
                         uint64_t n_combos, const uint32_t *d_idx, uint64_t n, double *d_sum, double *d_sq,
                         hipStream_t st)
{
    if ((n_combos + kWG - 1) / kWG >= (1ull << 31)) return sky::set_error(SKY_ERR_INVALID_CONFIG, "too many combos");
    SCAN_TRY(hipMemsetAsync(ds->bad.p, 0, sizeof(uint32_t), st), "clear flag");
    ScanArgs a{(const float *)ds->rows.p, d_idx, n, d_combos, d_coef, n_combos, ds->R, ds->B, ds->n, ds->q,
               d_sum, d_sq, (uint32_t *)ds->bad.p};
    const size_t lds = (size_t)ds->R * kTileStride * sizeof(float);
    SCAN_TRY(hipEventRecord(ds->ev0, st), "record event");
    switch (count) {
    case 1: launch_scan<1>(a, lds, st); break;
    case 2: launch_scan<2>(a, lds, st); break;
    case 3: launch_scan<3>(a, lds, st); break;
    case 4: launch_scan<4>(a, lds, st); break;
    case 5: launch_scan<5>(a, lds, st); break;
    case 6: launch_scan<6>(a, lds, st); break;
    case 7: launch_scan<7>(a, lds, st); break;
    default: launch_scan<8>(a, lds, st); break;
    }
    SCAN_TRY(hipGetLastError(), "launch scan");
    SCAN_TRY(hipEventRecord(ds->ev1, st), "record event");
    return SKY_OK;
}

// After the stream has synchronised: the index-error flag and the timing.
sky_status finish(sky_scan_dataset *ds, uint64_t evaluations, std::chrono::steady_clock::time_point t0,
                  sky_scan_stats *stats)
{
    uint32_t bad = 0;
    SCAN_TRY(hipMemcpy(&bad, ds->bad.p, sizeof bad, hipMemcpyDeviceToHost), "read flag");
    if (bad & 1u) return sky::set_error(SKY_ERR_INVALID_CONFIG, "combo band index out of range");
    if (bad & 2u) return sky::set_error(SKY_ERR_INVALID_CONFIG, "sample index out of range");
    if (stats) {
        float ms = 0.0f;
        SCAN_TRY(hipEventElapsedTime(&ms, ds->ev0, ds->ev1), "event time");
        stats->kernel_ms = ms;
        stats->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        stats->evaluations = evaluations;
    }
    return SKY_OK;
}

}  // namespace

extern "C" {

sky_status sky_scan_dataset_create(int device, const float *band_values, uint32_t n_bands, uint64_t n_samples,
                                   const float *reference_oklab, sky_scan_dataset **out)
{
    if (!out || !band_values || !reference_oklab) return sky::set_error(SKY_ERR_INVALID_CONFIG, "null argument");
    *out = nullptr;
    if (n_bands < 1 || n_bands > SKY_SCAN_MAX_BANDS)
        return sky::set_error(SKY_ERR_INVALID_CONFIG, "band count must be in 1..=64");
    if (n_samples < 1 || n_samples >= (1ull << 32))
        return sky::set_error(SKY_ERR_INVALID_CONFIG, "sample count must be in 1..2^32");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return sky::set_error(SKY_ERR_NO_DEVICE, "no usable HIP device");
    if (device < 0 || device >= ndev) return sky::set_error(SKY_ERR_NO_DEVICE, "device index out of range");
    SCAN_TRY(hipSetDevice(device), "set device");
    auto *ds = new sky_scan_dataset();
    ds->device = device;
    ds->B = n_bands;
    ds->R = n_bands + 3;
    ds->n = n_samples;
    ds->q = quad_form();
    auto bail = [&](sky_status st) {
        delete ds;
        return st;
    };
    hipError_t e = hipStreamCreateWithFlags(&ds->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&ds->ev0);
    if (e == hipSuccess) e = hipEventCreate(&ds->ev1);
    if (e != hipSuccess) return bail(hip_err(e, "create stream"));
    // reference colour as cube roots of LMS: Lab = M2 lms', lms' = M2^-1 Lab
    double inv[3][3];
    inverse3(kM2, inv);
    const uint64_t R = ds->R;
    std::vector<float> rows(n_samples * R), cols(n_samples * R);
    for (uint64_t s = 0; s < n_samples; ++s) {
        for (uint32_t b = 0; b < n_bands; ++b) rows[s * R + b] = band_values[s * n_bands + b];
        const float *lab = reference_oklab + 3 * s;
        for (int r = 0; r < 3; ++r)
            rows[s * R + n_bands + r] = (float)(inv[r][0] * lab[0] + inv[r][1] * lab[1] + inv[r][2] * lab[2]);
        for (uint64_t c = 0; c < R; ++c) cols[c * n_samples + s] = rows[s * R + c];
    }
    if ((e = ds->bad.ensure(sizeof(uint32_t))) != hipSuccess) return bail(hip_err(e, "alloc flag"));
    if ((e = ds->rows.ensure(rows.size() * sizeof(float))) != hipSuccess) return bail(hip_err(e, "alloc samples"));
    if ((e = ds->cols.ensure(cols.size() * sizeof(float))) != hipSuccess) return bail(hip_err(e, "alloc samples"));
    e = hipMemcpy(ds->rows.p, rows.data(), rows.size() * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(ds->cols.p, cols.data(), cols.size() * sizeof(float), hipMemcpyHostToDevice);
    if (e != hipSuccess) return bail(hip_err(e, "upload samples"));
    *out = ds;
    return SKY_OK;
}

void sky_scan_dataset_destroy(sky_scan_dataset *ds)
{
    if (!ds) return;
    (void)hipSetDevice(ds->device);
    delete ds;
}

sky_status sky_scan_combos(sky_scan_dataset *ds, const uint8_t *combos, const float *coef, uint32_t count,
                           uint64_t n_combos, const uint32_t *samples, uint64_t n_samples, double *sum,
                           double *sum_sq, sky_scan_stats *stats)
{
    auto t0 = std::chrono::steady_clock::now();
    if (!sum || !sum_sq) return sky::set_error(SKY_ERR_INVALID_CONFIG, "null output");
    if (n_combos == 0) return SKY_OK;
    if (ds) SCAN_TRY(hipSetDevice(ds->device), "set device");
    uint64_t n = 0;
    sky_status st = upload_combos(ds, combos, coef, count, n_combos, samples, n_samples, &n);
    if (st != SKY_OK) return st;
    SCAN_TRY(ds->sum.ensure(n_combos * sizeof(double)), "alloc sums");
    SCAN_TRY(ds->sq.ensure(n_combos * sizeof(double)), "alloc sums");
    st = launch_screen(ds, (const uint8_t *)ds->combos.p, (const float *)ds->coef.p, count, n_combos,
                       samples ? (const uint32_t *)ds->idx.p : nullptr, n, (double *)ds->sum.p, (double *)ds->sq.p,
                       ds->stream);
    if (st != SKY_OK) return st;
    SCAN_TRY(hipMemcpyAsync(sum, ds->sum.p, n_combos * sizeof(double), hipMemcpyDeviceToHost, ds->stream), "copy sums");
    SCAN_TRY(hipMemcpyAsync(sum_sq, ds->sq.p, n_combos * sizeof(double), hipMemcpyDeviceToHost, ds->stream), "copy sums");
    SCAN_TRY(hipStreamSynchronize(ds->stream), "synchronize");
    return finish(ds, n_combos * n, t0, stats);
}

sky_status sky_scan_combos_device(sky_scan_dataset *ds, const uint8_t *d_combos, const float *d_coef, uint32_t count,
                                  uint64_t n_combos, const uint32_t *d_samples, uint64_t n_samples, double *d_sum,
                                  double *d_sum_sq, void *stream, sky_scan_stats *stats)
{
    auto t0 = std::chrono::steady_clock::now();
    if (!ds || !d_combos || !d_coef || !d_sum || !d_sum_sq)
        return sky::set_error(SKY_ERR_INVALID_CONFIG, "null dataset, combos, coefficients or output");
    if (check_count(count) != SKY_OK) return SKY_ERR_INVALID_CONFIG;
    if (d_samples && n_samples == 0) return sky::set_error(SKY_ERR_INVALID_CONFIG, "empty sample list");
    if (n_combos == 0) return SKY_OK;
    SCAN_TRY(hipSetDevice(ds->device), "set device");
    hipStream_t st = stream ? (hipStream_t)stream : ds->stream;
    const uint64_t n = d_samples ? n_samples : ds->n;
    sky_status s = launch_screen(ds, d_combos, d_coef, count, n_combos, d_samples, n, d_sum, d_sum_sq, st);
    if (s != SKY_OK) return s;
    SCAN_TRY(hipStreamSynchronize(st), "synchronize");
    return finish(ds, n_combos * n, t0, stats);
}

sky_status sky_refine_combos(sky_scan_dataset *ds, const uint8_t *combos, const float *coef, uint32_t count,
                             uint64_t n_combos, const uint32_t *samples, uint64_t n_samples,
                             sky_combo_stats *out, sky_scan_stats *stats)
{
    auto t0 = std::chrono::steady_clock::now();
    if (!out) return sky::set_error(SKY_ERR_INVALID_CONFIG, "null output");
    if (n_combos > 65536) return sky::set_error(SKY_ERR_INVALID_CONFIG, "at most 65536 combos per refinement");
    if (n_combos == 0) return SKY_OK;
    if (ds) SCAN_TRY(hipSetDevice(ds->device), "set device");
    uint64_t n = 0;
    sky_status st = upload_combos(ds, combos, coef, count, n_combos, samples, n_samples, &n);
    if (st != SKY_OK) return st;
    // errors of up to 2^29 floats (2 GiB) at a time
    const uint64_t chunk = std::max<uint64_t>(1, std::min<uint64_t>(n_combos, (1ull << 29) / n));
    SCAN_TRY(ds->err.ensure(chunk * n * sizeof(float)), "alloc errors");
    SCAN_TRY(ds->stats.ensure(n_combos * sizeof(sky_combo_stats)), "alloc statistics");
    const uint32_t bpc = (uint32_t)((n + 255) / 256);
    if ((uint64_t)bpc * chunk >= (1ull << 31)) return sky::set_error(SKY_ERR_INVALID_CONFIG, "refinement too large");
    SCAN_TRY(hipMemsetAsync(ds->bad.p, 0, sizeof(uint32_t), ds->stream), "clear flag");
    SCAN_TRY(hipEventRecord(ds->ev0, ds->stream), "record event");
    for (uint64_t c0 = 0; c0 < n_combos; c0 += chunk) {
        const uint32_t nc = (uint32_t)std::min<uint64_t>(chunk, n_combos - c0);
        ErrArgs a{(const float *)ds->cols.p, ds->n, samples ? (const uint32_t *)ds->idx.p : nullptr, n,
                  (const uint8_t *)ds->combos.p + c0 * count, (const float *)ds->coef.p + c0 * count * 3, ds->R,
                  ds->B, bpc, ds->q, (float *)ds->err.p, (uint32_t *)ds->bad.p};
        switch (count) {
        case 1: launch_errors<1>(a, nc, ds->stream); break;
        case 2: launch_errors<2>(a, nc, ds->stream); break;
        case 3: launch_errors<3>(a, nc, ds->stream); break;
        case 4: launch_errors<4>(a, nc, ds->stream); break;
        case 5: launch_errors<5>(a, nc, ds->stream); break;
        case 6: launch_errors<6>(a, nc, ds->stream); break;
        case 7: launch_errors<7>(a, nc, ds->stream); break;
        default: launch_errors<8>(a, nc, ds->stream); break;
        }
        SCAN_TRY(hipGetLastError(), "launch errors");
        hipLaunchKernelGGL(k_select, dim3(nc), dim3(kSelectWG), 0, ds->stream, (const float *)ds->err.p, n,
                           (sky_combo_stats *)ds->stats.p + c0);
        SCAN_TRY(hipGetLastError(), "launch select");
    }
    SCAN_TRY(hipEventRecord(ds->ev1, ds->stream), "record event");
    SCAN_TRY(hipMemcpyAsync(out, ds->stats.p, n_combos * sizeof(sky_combo_stats), hipMemcpyDeviceToHost, ds->stream),
             "copy statistics");
    SCAN_TRY(hipStreamSynchronize(ds->stream), "synchronize");
    return finish(ds, n_combos * n, t0, stats);
}

}  // extern "C"
