// sky_device.hpp -- device-side building blocks of the spectral sky path tracer
// for gfx950 (CDNA4, wave64).  Shared by the wave-pool integrator, the simple
// per-(pixel, lambda) integrator and the optical-depth kernel.
//
// Precision: geometry is FP32 and never forms r^2 - R^2 by subtracting R^2
// from r^2 (which cancels at 6360 km, the reference's FP32 (r-R)(r+R) at
// sky_trace.wgsl:214 is good to ~0.5 m).  Each vertex carries w = r^2 - R^2
// itself; along a ray from an origin with w_o and b = o.d it is
// w(t) = w_o + t (2b + t), good to ~1 cm of altitude for every ray, and set
// to exactly 0 at a ground hit.  Roots come from q = w - b^2 (the same for
// every point of a line), formed with one rounding.  FP64 is used only where
// the ground's discriminant -q lies within the FP32 error of b of zero: a ray
// grazing the ground within centimetres, whose hit or miss that error could
// flip (ray_setup).
//
// FP32 transcendentals, reciprocals and roots use the gfx950 hardware
// instructions (v_exp/v_log/v_sin/v_cos/v_rcp/v_rsq/v_sqrt_f32, about 1 ulp)
// instead of the correctly rounded library sequences (10-40 instructions
// each); FP64 roots are an FP32 root refined by one Newton step.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sky {

constexpr float kPi = 3.14159265358979323846f;
constexpr float kInvPi = 0.31830988618379067154f;
constexpr float kInv2Pi = 0.15915494309189533577f;
constexpr float kTau = 6.28318530717958647692f;
constexpr float kLn2 = 0.69314718055994530942f;
constexpr float kIsotropicPdf = 0.07957747154594766788f;  // sky_trace.wgsl:10
constexpr float kRayEpsilonKm = 0.005f;                    // sky_trace.wgsl:8
constexpr int kMaxLambda = 64;
constexpr int kMaxLayers = 256;

constexpr uint32_t kFlagPhaseIS = 1u;
constexpr uint32_t kFlagRatio = 2u;
constexpr uint32_t kFlagCarriedTau = 4u;

// Per-wavelength constants: km^-1 coefficients at the profiles' reference
// heights, Planck sun radiance, and cmf(lambda) * weight_nm.
struct DevLambda {
    float br, bo, bas, baa, lsun, wx, wy, wz;
};

// Everything one render needs on the device.
struct DevScene {
    double R, top, dz, obs_y;  // host-side and optical-depth-kernel geometry (FP64)
    float delta_topf;  // r_top^2 - R^2
    float shell_tolf;  // r^2 - R^2 within which an inward origin counts as on a shell boundary
    float obs_yf, obs_wf;  // camera at (0, obs_yf, 0), its r^2 - R^2
    // Shell k's lower boundary h_k (k = 0..L, h_L = top): uniform, or with
    // thicknesses growing by a factor g.  First guess of the shell of
    // altitude h: floor(h c2) (uniform, c1 = 0) or floor(log2(1 + h c1) c2).
    float shell_c1, shell_c2;
    float Rf, R2f, topf, eps;
    float sun[3];
    float omc, cone_pdf;  // 1 - cos(sun radius) (= 2 sin^2(r/2)), 1/(2 pi omc)
    float cone_omega;     // 2 pi omc: the sun cone's solid angle, 1 / cone_pdf
    float inv_hr, inv_ha, hc, inv_hw, g, albedo;
    float inv_2g, inv_dls;  // 1 / (2 g) (0 when |g| < 1e-3), 1 / direct_light_samples
    float l2_hr, l2_ha;   // log2(e) / H: exp(-h / H) = exp2(-h l2_h)
    float inv_w, inv_h;   // 1 / width, 1 / height
    float inv_nlam;       // 1 / nlam
    // Sky-view LUT frame (sky_trace.wgsl:641-681), precomputed on the host.
    float svl_zha, svl_beta, svl_fw[3], svl_side[3];
    uint32_t width, height, spp, sample_offset;
    uint32_t seed_lo, seed_hi;
    uint32_t nlam, L;
    uint32_t dls, projection, max_depth, rr_start, watchdog_limit, max_segments, flags;
    uint32_t walk_cap;  // shell crossings a lane takes per tracking step (0: no cap)
    // Tabulated medium (null tab: the analytic profiles): sigma (R, As, O,
    // Aa) per band at K altitude breakpoints, linear in between; a uniform
    // index of tab_nidx cells gives each altitude's first candidate interval.
    const float4 *tab;         // [lambda * K + k]
    const float *tab_z;        // [K] breakpoints (km), tab_z[0] = 0, tab_z[K-1] = top
    const float *tab_inv;      // [K] 1 / (tab_z[k+1] - tab_z[k])
    const uint16_t *tab_idx;   // [tab_nidx]
    uint32_t tab_k, tab_nidx;
    float tab_inv_dh;
    // One-load interval lookup (null: use tab_idx): per cell of a uniform grid
    // of tab_ncell cells, {k0, z[k0], inv[k0], za} {inv_a, zb, inv_b, 0}; the
    // interval of h in cell j is k0 + (h >= za) + (h >= zb).
    const float4 *tab_cell;    // [2 * tab_ncell]
    uint32_t tab_ncell;
    float tab_cell_inv_dh;
    const DevLambda *lam;
    const float *maj;  // [lambda * L + layer]
    const float *shell_w;  // [L + 1]: r_k^2 - R^2 = h_k (2R + h_k)
    float *film;       // [pixel * nlam + lambda] (may be null)
    float *film_m2;    // same layout (may be null)
    float *xyz;        // [pixel * 3 + c] (may be null)
    uint32_t *diag;    // [0] = watchdog flag
};

// ----------------------------------------------------------- fast math
__device__ __forceinline__ float f_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float f_div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ float f_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float f_rsq(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ float f_exp2(float x) { return __builtin_amdgcn_exp2f(x); }
__device__ __forceinline__ float f_log2(float x) { return __builtin_amdgcn_logf(x); }
// sin / cos of 2 pi x: the hardware takes the angle in revolutions.
__device__ __forceinline__ float f_sin_rev(float x) { return __builtin_amdgcn_sinf(x); }
__device__ __forceinline__ float f_cos_rev(float x) { return __builtin_amdgcn_cosf(x); }

// Free-flight optical depth -ln(max(1 - u, 1e-7)) (sky_trace.wgsl:424).
__device__ __forceinline__ float free_flight(float u)
{
    return -kLn2 * f_log2(fmaxf(1.0f - u, 0.0000001f));
}

// sqrt(x) for x >= 0 in FP64: the FP32 reciprocal root r gives the root
// x r and, halved, the Newton step's 1 / (2 sqrt x); relative error ~1e-13
// (0 for x below 1e-30).
__device__ __forceinline__ double sqrt64(double x)
{
    const float xf = (float)x;
    if (!(xf > 1e-30f)) return 0.0;
    const float r = f_rsq(xf);
    const double s = (double)(xf * r);
    return s + (x - s * s) * (double)(0.5f * r);
}

// 1 / |d| in FP64 for an FP32-normalised d (|d|^2 - 1 ~ 1e-7): third-order
// series of (1 + e)^(-1/2), exact to ~1e-20 there.
__device__ __forceinline__ double inv_norm64(double dx, double dy, double dz)
{
    double e = dx * dx + dy * dy + dz * dz - 1.0;
    return 1.0 + e * (-0.5 + e * (0.375 - 0.3125 * e));
}

// floor(n / d) for n < 2^21 from rd = 1 / d (within an ulp): (n + 1/2) / d
// lies at least 1/(2d) from an integer, and the float product's error is
// below (n + 1/2) / d * 2^-22 < 1/(2d).
__device__ __forceinline__ uint32_t udiv_small(uint32_t n, float rd)
{
    return (uint32_t)(((float)n + 0.5f) * rd);
}

// ---------------------------------------------------------------- RNG
// pcg4d counter hash keyed (pixel ^ seed_lo, sample ^ seed_hi, band, dim)
// (sky_trace.wgsl:156-189).  The first LCG step of the three per-path words
// is done once per path; the dimension word is carried pre-multiplied and
// advanced by the multiplier, so a draw costs 8 integer multiplies.
struct Rng {
    uint32_t x, y, z;  // key * 1664525 + 1013904223
    uint32_t w;        // dim * 1664525 + 1013904223
};

constexpr uint32_t kLcgMul = 1664525u;
constexpr uint32_t kLcgAdd = 1013904223u;

__device__ __forceinline__ Rng rng_make(uint32_t pixel, uint32_t sample, uint32_t band,
                                        const DevScene &s)
{
    Rng r;
    r.x = (pixel ^ s.seed_lo) * kLcgMul + kLcgAdd;
    r.y = (sample ^ s.seed_hi) * kLcgMul + kLcgAdd;
    r.z = band * kLcgMul + kLcgAdd;
    r.w = kLcgAdd;  // dim 0
    return r;
}

__device__ __forceinline__ uint32_t pcg4d_fold(uint32_t x, uint32_t y, uint32_t z, uint32_t w)
{
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
    return x ^ y ^ z ^ w;
}

__device__ __forceinline__ float rng_unit(uint32_t v) { return (float)(v >> 8) * 5.9604644775390625e-8f; }

__device__ __forceinline__ float rng_next(Rng &r)
{
    uint32_t v = pcg4d_fold(r.x, r.y, r.z, r.w);
    r.w += kLcgMul;
    return rng_unit(v);
}

// Value of draw `ahead` positions after the current one, without advancing.
__device__ __forceinline__ float rng_peek(const Rng &r, uint32_t ahead)
{
    return rng_unit(pcg4d_fold(r.x, r.y, r.z, r.w + ahead * kLcgMul));
}

// ------------------------------------------------------------ vectors
struct f3 {
    float x, y, z;
};

__device__ __forceinline__ float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// normalize_or_y, sky_trace.wgsl:191-197.
__device__ __forceinline__ f3 normalize_or_y(f3 v)
{
    float l2 = dot3(v, v);
    if (l2 > 0.0f) {
        float r = f_rsq(l2);
        return f3{v.x * r, v.y * r, v.z * r};
    }
    return f3{0.0f, 1.0f, 0.0f};
}

// direction_from_axis_mu with the branchless orthonormal basis
// (sky_trace.wgsl:512-532): unit axis n, polar cosine mu, polar sine st,
// azimuth phi_rev in revolutions.  The basis vectors of a unit axis are unit
// vectors; only the result is renormalised.
__device__ __forceinline__ f3 frame_dir(f3 n, float mu, float st, float phi_rev)
{
    float sign = n.z >= 0.0f ? 1.0f : -1.0f;
    float a = -f_rcp(sign + n.z);
    float b = n.x * n.y * a;
    f3 t{1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x};
    f3 bt{b, sign + n.y * n.y * a, -n.y};
    float sp = f_sin_rev(phi_rev), cp = f_cos_rev(phi_rev);
    float c1 = st * cp, c2 = st * sp;
    f3 v{t.x * c1 + bt.x * c2 + n.x * mu, t.y * c1 + bt.y * c2 + n.y * mu, t.z * c1 + bt.z * c2 + n.z * mu};
    return normalize_or_y(v);
}

// ------------------------------------------------------------- rays
typedef __attribute__((address_space(3))) const float *LdsShellPtr;
typedef __attribute__((address_space(1))) const float *GlobalShellPtr;

// The shell table r_k^2 - R^2 (k = 0..L) read from LDS or global memory; the
// typed pointers keep each read a plain LDS or global load.
struct ShellTab {
    LdsShellPtr lds;
    GlobalShellPtr glob;
    bool use_lds;
    __device__ __forceinline__ float operator[](int k) const { return use_lds ? lds[k] : glob[k]; }
};

__device__ __forceinline__ ShellTab global_shells(const DevScene &s)
{
    ShellTab t;
    t.lds = (LdsShellPtr) nullptr;
    t.glob = (GlobalShellPtr)s.shell_w;
    t.use_lds = false;
    return t;
}

struct Ray {
    f3 d;          // unit direction
    float bf, qf;  // b = o.d at the origin o, q = w - b^2 (the line's; origin-independent)
    float wo;      // r^2 - R^2 at the origin
    float tmax;
    int kind;     // 0 ground, 1 atmosphere exit (sky_trace.wgsl:89-93)
    int layer;    // majorant shell at t = 0
    int outward;  // moving away from the planet centre at t = 0
    int hit;
};

// r^2 - R^2 at distance t along a ray whose origin has b = o.d and w_o:
// |o + t d|^2 - R^2 = w_o + t (2b + t).
__device__ __forceinline__ float ray_w(float bf, float wo, float t)
{
    return __builtin_fmaf(t, __builtin_fmaf(2.0f, bf, t), wo);
}

// The FP32 error bound of q = w - b0^2 from the error of b0 = p.d: three
// roundings of terms at most |p| <= 6460 km give |db0| <= 3 2^-24 |p| <= 1.2e-3
// km, so |dq| <= 2 |b0| |db0| + b0^2 |(|d|^2 - 1)| (the FP32 unit direction)
// + rounding, below 2.4e-3 |b0| + 2.4e-7 b0^2 + 1e-5.
__device__ __forceinline__ float graze_tol(float b0)
{
    const float ab = fabsf(b0);
    return __builtin_fmaf(ab, __builtin_fmaf(2.4e-7f, ab, 2.4e-3f), 1e-5f);
}

// Ray set-up: next_boundary (sky_trace.wgsl:233-250) and the initial shell of
// the walk (layer_for_altitude, wgsl:252-258, generalised to non-uniform
// shells) for the ray from p + e d along the unit direction d, where the
// vertex p has r^2 - R^2 = w (e: the reference's ray epsilon, wgsl:744, 766,
// or 0).  b^2 - c of wgsl:212-215 is -q, the same for every origin on the
// line; the near ground root is the stable w_o / t1 (the roots' product).
// Where -q is within its FP32 error of zero (graze_tol) b, q and the ground
// roots are formed again in FP64 from the same p, w and d.
template <int COUNT_L = 0>
__device__ __forceinline__ void ray_setup(const DevScene &s, const ShellTab &tab, f3 p, float w, f3 d, float e,
                                          Ray &r)
{
    const float b0 = __builtin_fmaf(p.z, d.z, __builtin_fmaf(p.y, d.y, p.x * d.x));
    float q = __builtin_fmaf(-b0, b0, w);
    float b = b0 + e;
    const float wo = __builtin_fmaf(e, b0 + b, w);
    const float eps = kRayEpsilonKm;
    float tg = -1.0f;
    if (fabsf(q) <= graze_tol(b0)) {
        // grazing the ground: FP64 from the same p, w and (renormalised) d
        const double dx = d.x, dy = d.y, dz = d.z;
        const double b64 = fma((double)p.z, dz, fma((double)p.y, dy, (double)p.x * dx)) * inv_norm64(dx, dy, dz);
        const double q64 = (double)w - b64 * b64;
        const double be = b64 + (double)e;
        if (q64 <= 0.0) {
            const double sg = sqrt64(-q64);
            const double t0 = -be - sg, t1 = sg - be;
            tg = t0 > (double)eps ? (float)t0 : (t1 > (double)eps ? (float)t1 : -1.0f);
        }
        q = (float)q64;
        b = (float)be;
    } else if (q < 0.0f) {
        const float sg = f_sqrt(-q);
        const float t1 = sg - b;
        const float t0 = b < 0.0f ? f_div(wo, t1) : -b - sg;
        tg = t0 > eps ? t0 : (t1 > eps ? t1 : -1.0f);
    }
    r.d = d;
    r.bf = b;
    r.qf = q;
    r.wo = wo;
    const float disc_top = s.delta_topf - q;
    r.hit = disc_top >= 0.0f;
    float t_exit = f_sqrt(fmaxf(disc_top, 0.0f)) - b;
    if (t_exit < eps) t_exit = eps;
    if (tg > eps && tg < t_exit) {
        r.tmax = tg;
        r.kind = 0;
    } else {
        r.tmax = t_exit;
        r.kind = 1;
    }
    const int L = (int)s.L;
    int k = 0;
    if constexpr (COUNT_L > 0) {
        // at most COUNT_L shells (L <= COUNT_L): the shell holding the origin
        // is the number of interior boundaries at or below it (L is the same
        // for the whole wave, so the loop runs exactly L - 1 times)
#pragma unroll 1
        for (int j = 1; j < L; ++j) k += tab[j] <= wo ? 1 : 0;
    } else {
        const float h0 = f_div(wo, f_sqrt(fmaxf(s.R2f + wo, 0.0f)) + s.Rf);
        float kg = s.shell_c1 > 0.0f ? f_log2(fmaxf(1.0f + h0 * s.shell_c1, 1.0f)) * s.shell_c2 : h0 * s.shell_c2;
        k = (int)fminf(fmaxf(kg, 0.0f), (float)(L - 1));
        while (k > 0 && tab[k] > wo) --k;
        while (k < L - 1 && tab[k + 1] <= wo) ++k;
    }
    r.outward = b >= 0.0f;
    if (!r.outward && k > 0 && wo - tab[k] <= s.shell_tolf) k -= 1;
    r.layer = k;
}

// Next crossing of the majorant-shell walk (the segments of
// sky_trace.wgsl:284-294): the root of D_k - q for the boundary the ray meets
// next (an error e of D_k - q puts the crossing within e / 2R of the
// boundary's altitude).
__device__ __forceinline__ float next_crossing(const DevScene &s, const ShellTab &tab, float q, float bf, int k,
                                               int out, int &kn, int &on)
{
    if (!out) {
        float disc = tab[k] - q;
        if (disc > 0.0f) {
            on = 0;
            if (k == 0) {
                kn = 0;
                return __builtin_inff();
            }
            kn = k - 1;
            return -f_sqrt(disc) - bf;
        }
    }
    on = 1;
    if (k + 1 >= (int)s.L) {
        kn = k;
        return __builtin_inff();
    }
    kn = k + 1;
    float disc = tab[k + 1] - q;
    return f_sqrt(fmaxf(disc, 0.0f)) - bf;
}

// The same crossing without branches, for a shell table in LDS: both
// boundaries of shell k are read at once and the root of the selected one is
// taken, so every lane of a wave runs the same instructions.  Equal to
// next_crossing.
__device__ __forceinline__ float next_crossing_sel(const DevScene &s, LdsShellPtr tab, float q, float bf, int k,
                                                   int out, int &kn, int &on)
{
    const float din = tab[k];       // inner boundary of shell k
    const float dout = tab[k + 1];  // outer boundary (the table has L + 1 entries)
    // inward iff the ray reaches the inner boundary: din - q > 0, which for
    // finite operands holds exactly when din > q
    const bool inw = !out && din > q;
    on = inw ? 0 : 1;
    // the next shell, clamped to [0, L - 1]: it stays k exactly at the two
    // edges (inward from shell 0, outward from the top shell)
    kn = min(max(inw ? k - 1 : k + 1, 0), (int)s.L - 1);
    const bool edge = kn == k;
    const float r = f_sqrt(fmaxf((inw ? din : dout) - q, 0.0f));
    const float tc = (inw ? -r : r) - bf;
    return edge ? __builtin_inff() : tc;
}

__device__ __forceinline__ float next_crossing(const DevScene &s, const ShellTab &tab, const Ray &r, int k,
                                               int out, int &kn, int &on)
{
    return next_crossing(s, tab, r.qf, r.bf, k, out, kn, on);
}

// Altitude at distance t: w / (sqrt(R^2 + w) + R) with w = ray_w.
__device__ __forceinline__ float altitude_bw(const DevScene &s, float bf, float wo, float t)
{
    const float w = ray_w(bf, wo, t);
    return f_div(w, f_sqrt(s.R2f + w) + s.Rf);
}

__device__ __forceinline__ float altitude(const DevScene &s, const Ray &r, float t)
{
    return altitude_bw(s, r.bf, r.wo, t);
}

// coefficients_at (sky_trace.wgsl:356-386) for the analytic profiles.
struct Coeffs {
    float sr, sas, so, saa;
};

// Interval k of the breakpoint grid holding altitude h (0 <= h <= top).
__device__ __forceinline__ uint32_t tab_interval(const DevScene &s, float h)
{
    uint32_t j = (uint32_t)(h * s.tab_inv_dh);
    j = j < s.tab_nidx ? j : s.tab_nidx - 1;
    uint32_t k = s.tab_idx[j];
    while (k + 2 < s.tab_k && h >= s.tab_z[k + 1]) ++k;
    while (k > 0 && h < s.tab_z[k]) --k;
    return k;
}

// coefficients_at (medium.rs:28-57) for the tabulated profiles: the merged
// breakpoint grid makes every coefficient linear within an interval.
__device__ __forceinline__ Coeffs tab_medium(const DevScene &s, uint32_t band, float h)
{
    uint32_t k;
    float z0, inv;
    if (s.tab_cell) {  // one record load, then selects: the same k, z[k] and inv[k]
        uint32_t j = (uint32_t)(h * s.tab_cell_inv_dh);
        j = j < s.tab_ncell ? j : s.tab_ncell - 1;
        const float4 c0 = s.tab_cell[2 * j], c1 = s.tab_cell[2 * j + 1];
        const bool ua = h >= c0.w, ub = h >= c1.y;
        k = __float_as_uint(c0.x) + (ua ? 1u : 0u) + (ub ? 1u : 0u);
        z0 = ub ? c1.y : (ua ? c0.w : c0.y);
        inv = ub ? c1.z : (ua ? c1.x : c0.z);
    } else {
        k = tab_interval(s, h);
        z0 = s.tab_z[k];
        inv = s.tab_inv[k];
    }
    float t = (h - z0) * inv;
    t = fminf(fmaxf(t, 0.0f), 1.0f);
    const float4 a = s.tab[band * s.tab_k + k], b = s.tab[band * s.tab_k + k + 1];
    return Coeffs{fmaf(b.x - a.x, t, a.x), fmaf(b.y - a.y, t, a.y), fmaf(b.z - a.z, t, a.z),
                  fmaf(b.w - a.w, t, a.w)};
}

__device__ __forceinline__ Coeffs medium4(const DevScene &s, uint32_t band, float br, float bo, float bas, float baa,
                                          float h)
{
    Coeffs c{0.0f, 0.0f, 0.0f, 0.0f};
    if (h < 0.0f || h > s.topf) return c;
    if (s.tab) return tab_medium(s, band, h);
    float rr = f_exp2(-h * s.l2_hr);
    float ra = f_exp2(-h * s.l2_ha);
    float ro = 1.0f - fabsf(h - s.hc) * s.inv_hw;
    ro = ro < 0.0f ? 0.0f : ro;
    c.sr = br * rr;
    c.sas = bas * ra;
    c.so = bo * ro;
    c.saa = baa * ra;
    return c;
}

__device__ __forceinline__ Coeffs medium(const DevScene &s, uint32_t band, const DevLambda &l, float h)
{
    return medium4(s, band, l.br, l.bo, l.bas, l.baa, h);
}

// ------------------------------------------------------------ phase
__device__ __forceinline__ float phase_rayleigh(float mu)
{
    return 3.0f / (16.0f * kPi) * (1.0f + mu * mu);  // sky_trace.wgsl:477-479
}

__device__ __forceinline__ float phase_hg(float g, float mu)
{
    // (1 - g^2) / (4 pi den^1.5), den^-1.5 as the cube of one reciprocal root
    float den = 1.0f + g * g - 2.0f * g * mu;
    float r = f_rsq(den);
    return (1.0f - g * g) * (0.25f * kInvPi) * (r * r * r);
}

// mixed_phase_value, sky_trace.wgsl:496-510 (Rayleigh + HG aerosol).
__device__ __forceinline__ float mixed_phase(const DevScene &s, float sr, float sas, float mu)
{
    float scat = sr + sas;
    if (scat <= 0.0f) return 0.0f;
    return f_div(sr * phase_rayleigh(mu) + sas * phase_hg(s.g, mu), scat);
}

// invert_rayleigh_cdf (crates/sky-core/src/sampling.rs:182-185): the root of
// mu^3 + 3 mu = 2 y, y = 4 xi - 2, by Cardano: mu = A - 1/A with
// A = cbrt(|y| + sqrt(y^2 + 1)) (odd in y; A in [1, 1.62]).
__device__ __forceinline__ float sample_rayleigh_mu(float xi)
{
    float y = 4.0f * xi - 2.0f;
    float ay = fabsf(y);
    float A = f_exp2(f_log2(ay + f_sqrt(fmaf(ay, ay, 1.0f))) * (1.0f / 3.0f));
    float mu = A - f_rcp(A);
    mu = y < 0.0f ? -mu : mu;
    return fminf(fmaxf(mu, -1.0f), 1.0f);
}

// inv_2g = 1 / (2 g), precomputed on the host.
__device__ __forceinline__ float sample_hg_mu(float g, float inv_2g, float xi)
{
    if (fabsf(g) < 1e-3f) return 1.0f - 2.0f * xi;
    float sv = f_div(1.0f - g * g, 1.0f - g + 2.0f * g * xi);
    float mu = (1.0f + g * g - sv * sv) * inv_2g;
    return fminf(fmaxf(mu, -1.0f), 1.0f);
}

// Sun cone sample (sky_trace.wgsl:541-546) as (mu, st): 1 - cos(theta) is
// formed from 1 - cos(r) = 2 sin^2(r/2) so the 4.7 mrad cone keeps its
// solid angle.
__device__ __forceinline__ void sun_cone_mu(const DevScene &s, float xi0, float &mu, float &st)
{
    float x = fminf(xi0, 0.99999994f) * s.omc;
    mu = 1.0f - x;
    st = f_sqrt(x * (2.0f - x));
}

__device__ __forceinline__ f3 sun_cone_dir(const DevScene &s, float xi0, float xi1)
{
    float mu, st;
    sun_cone_mu(s, xi0, mu, st);
    return frame_dir(f3{s.sun[0], s.sun[1], s.sun[2]}, mu, st, xi1);
}

// direction_in_cone (sky_trace.wgsl:562-564) as |d - s|^2 <= 2 (1 - cos r).
__device__ __forceinline__ bool in_sun_cone_unit(const DevScene &s, f3 n)
{
    float ax = n.x - s.sun[0], ay = n.y - s.sun[1], az = n.z - s.sun[2];
    return ax * ax + ay * ay + az * az <= 2.0f * s.omc;
}

__device__ __forceinline__ bool in_sun_cone(const DevScene &s, f3 d) { return in_sun_cone_unit(s, normalize_or_y(d)); }

// Camera ray direction (sky_trace.wgsl:614-689).
__device__ __forceinline__ f3 camera_dir(const DevScene &s, uint32_t x, uint32_t y, float j1,
                                         float j2)
{
    if (s.projection == 1u) {
        // opaque copies: the terms below are then formed here, not hoisted
        // out of the caller's loops into registers held across them
        float iw = s.inv_w, ih = s.inv_h;
        asm volatile("" : "+v"(iw), "+v"(ih));
        float u = ((float)x + 0.5f) * iw;
        float v = ((float)y + 0.5f) * ih;
        float uu = s.width <= 1 ? 0.5f : f_div(u - 0.5f * iw, 1.0f - iw);
        float vv = s.height <= 1 ? 0.5f : f_div(v - 0.5f * ih, 1.0f - ih);
        uu = fminf(fmaxf(uu, 0.0f), 1.0f);
        vv = fminf(fmaxf(vv, 0.0f), 1.0f);
        float vzc;
        if (vv < 0.75f) {
            float c = vv * (1.0f / 0.75f);
            c = 1.0f - c;
            c = c * c;
            c = 1.0f - c;
            vzc = f_cos_rev(s.svl_zha * c * kInv2Pi);
        } else {
            float c = (vv - 0.75f) * 4.0f;
            c = c * c;
            vzc = f_cos_rev((s.svl_zha + s.svl_beta * c) * kInv2Pi);
        }
        float c = uu * uu;
        float lvc = -(c * 2.0f - 1.0f);
        float sv = f_sqrt(fmaxf(1.0f - vzc * vzc, 0.0f));
        float ss = f_sqrt(fmaxf(1.0f - lvc * lvc, 0.0f));
        f3 v3{sv * (s.svl_fw[0] * lvc + s.svl_side[0] * ss), vzc + sv * (s.svl_fw[1] * lvc + s.svl_side[1] * ss),
              sv * (s.svl_fw[2] * lvc + s.svl_side[2] * ss)};
        return normalize_or_y(v3);
    }
    // azimuth (u - 1/2) 2 pi and elevation (1/2 - v) pi, in revolutions
    float u = ((float)x + j1) * s.inv_w;
    float v = ((float)y + j2) * s.inv_h;
    float az = u - 0.5f;
    float el = (0.5f - v) * 0.5f;
    float se = f_sin_rev(el), ce = f_cos_rev(el);
    float sa = f_sin_rev(az), ca = f_cos_rev(az);
    return normalize_or_y(f3{ce * sa, se, ce * ca});
}

}  // namespace sky
