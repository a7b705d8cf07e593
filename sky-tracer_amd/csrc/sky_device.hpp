// sky_device.hpp -- device-side building blocks of the spectral sky path tracer
// for gfx950 (CDNA4, wave64).  Shared by the wave-pool integrator, the simple
// per-(pixel, lambda) integrator and the optical-depth kernel.
//
// Precision split: ray set-up (origin, b = o.d, q = |o|^2 - b^2 - R^2, boundary
// distances, shell discriminants) is FP64; every per-step quantity is FP32.
// The altitude along a ray is r^2 - R^2 = q + (b + t)^2, which has no
// cancellation at 6360 km (the reference's (r-R)(r+R) at sky_trace.wgsl:214 is
// FP32 and loses ~0.5 m per operand there).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sky {

constexpr float kPi = 3.14159265358979323846f;
constexpr float kInvPi = 0.31830988618379067154f;
constexpr float kTau = 6.28318530717958647692f;
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

// Everything one render needs on the device (kernel argument, by value).
struct DevScene {
    double R, top, dz, delta_top, obs_y;
    float Rf, R2f, topf, eps;
    float sun[3];
    float omc, cone_pdf;  // 1 - cos(sun radius) (= 2 sin^2(r/2)), 1/(2 pi omc)
    float inv_hr, inv_ha, hc, inv_hw, g, albedo;
    // Sky-view LUT frame (sky_trace.wgsl:641-681), precomputed on the host.
    float svl_zha, svl_beta, svl_fw[3], svl_side[3];
    uint32_t width, height, spp, sample_offset;
    uint32_t seed_lo, seed_hi;
    uint32_t nlam, L;
    uint32_t dls, projection, max_depth, rr_start, watchdog_limit, max_segments, flags;
    const DevLambda *lam;
    const float *maj;  // [lambda * L + layer]
    float *film;       // [pixel * nlam + lambda] (may be null)
    float *film_m2;    // same layout (may be null)
    float *xyz;        // [pixel * 3 + c] (may be null)
    uint32_t *diag;    // [0] = watchdog flag
};

// ---------------------------------------------------------------- RNG
// pcg4d counter hash keyed (pixel ^ seed_lo, sample ^ seed_hi, band, dim)
// (sky_trace.wgsl:156-189).  The first LCG step of the three per-path words
// is done once per path; the dimension word is carried pre-multiplied and
// advanced by the multiplier, so a draw costs 8 integer multiplies.
struct Rng {
    uint32_t x, y, z;  // key * 1664525 + 1013904223
    uint32_t w;        // dim * 1664525 + 1013904223
};

__device__ __forceinline__ Rng rng_make(uint32_t pixel, uint32_t sample, uint32_t band,
                                        const DevScene &s)
{
    Rng r;
    r.x = (pixel ^ s.seed_lo) * 1664525u + 1013904223u;
    r.y = (sample ^ s.seed_hi) * 1664525u + 1013904223u;
    r.z = band * 1664525u + 1013904223u;
    r.w = 1013904223u;  // dim 0
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

__device__ __forceinline__ float rng_next(Rng &r)
{
    uint32_t v = pcg4d_fold(r.x, r.y, r.z, r.w);
    r.w += 1664525u;
    return (float)(v >> 8) * 5.9604644775390625e-8f;
}

// Value of draw `ahead` positions after the current one, without advancing.
__device__ __forceinline__ float rng_peek(const Rng &r, uint32_t ahead)
{
    uint32_t v = pcg4d_fold(r.x, r.y, r.z, r.w + ahead * 1664525u);
    return (float)(v >> 8) * 5.9604644775390625e-8f;
}

// ------------------------------------------------------------ vectors
struct f3 {
    float x, y, z;
};

__device__ __forceinline__ float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// normalize_or_y, sky_trace.wgsl:191-197.
__device__ __forceinline__ f3 normalize_or_y(f3 v)
{
    float len = sqrtf(dot3(v, v));
    if (len > 0.0f) return f3{v.x / len, v.y / len, v.z / len};
    return f3{0.0f, 1.0f, 0.0f};
}

// direction_from_axis_mu with the branchless orthonormal basis
// (sky_trace.wgsl:512-532): polar cosine mu, polar sine st, azimuth phi.
__device__ __forceinline__ f3 frame_dir(f3 axis, float mu, float st, float phi)
{
    f3 n = normalize_or_y(axis);
    float sign = n.z >= 0.0f ? 1.0f : -1.0f;
    float a = -1.0f / (sign + n.z);
    float b = n.x * n.y * a;
    f3 t = normalize_or_y(f3{1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x});
    f3 bt = normalize_or_y(f3{b, sign + n.y * n.y * a, -n.y});
    float sp, cp;
    sincosf(phi, &sp, &cp);
    f3 v{t.x * (st * cp) + bt.x * (st * sp) + n.x * mu, t.y * (st * cp) + bt.y * (st * sp) + n.y * mu,
         t.z * (st * cp) + bt.z * (st * sp) + n.z * mu};
    return normalize_or_y(v);
}

// ------------------------------------------------------------- rays
struct Ray {
    double ox, oy, oz;  // origin (km), FP64
    double b, q;        // o.d and |o|^2 - (o.d)^2 - R^2
    f3 d;               // unit direction
    float bf, qf, tmax;
    int kind;     // 0 ground, 1 atmosphere exit (sky_trace.wgsl:89-93)
    int layer;    // majorant shell at t = 0
    int outward;  // moving away from the planet centre at t = 0
    int hit;
};

// Ray set-up: next_boundary (sky_trace.wgsl:233-250) from the impact
// parameter -- b^2 - c of wgsl:212-215 is D - q with D = r_shell^2 - R^2 --
// and the initial shell of the walk.
__device__ __forceinline__ void ray_setup(const DevScene &s, double ox, double oy, double oz, f3 d,
                                          Ray &r)
{
    double dx = d.x, dy = d.y, dz = d.z;
    double inv = 1.0 / sqrt(dx * dx + dy * dy + dz * dz);
    dx *= inv;
    dy *= inv;
    dz *= inv;
    r.ox = ox;
    r.oy = oy;
    r.oz = oz;
    r.d = d;
    double oo = ox * ox + oy * oy + oz * oz;
    double b = ox * dx + oy * dy + oz * dz;
    double q = oo - b * b - s.R * s.R;
    r.b = b;
    r.q = q;
    r.bf = (float)b;
    r.qf = (float)q;
    double disc_top = s.delta_top - q;
    r.hit = disc_top >= 0.0;
    const double eps = (double)kRayEpsilonKm;
    double t_exit = -b + sqrt(disc_top > 0.0 ? disc_top : 0.0);
    double tg = -1.0;
    if (-q >= 0.0) {
        double sg = sqrt(-q);
        double t0 = -b - sg, t1 = -b + sg;
        tg = t0 > eps ? t0 : (t1 > eps ? t1 : -1.0);
    }
    if (t_exit < eps) t_exit = eps;
    if (tg > eps && tg < t_exit) {
        r.tmax = (float)tg;
        r.kind = 0;
    } else {
        r.tmax = (float)t_exit;
        r.kind = 1;
    }
    double h0 = sqrt(oo) - s.R;
    int k = (int)floor(h0 / s.dz);
    r.outward = b >= 0.0;
    if (!r.outward && k > 0 && h0 - k * s.dz <= 1e-9) k -= 1;
    k = k < 0 ? 0 : k;
    k = k > (int)s.L - 1 ? (int)s.L - 1 : k;
    r.layer = k;
}

// Next crossing of the majorant-shell walk (the segments of
// sky_trace.wgsl:284-294): D_k - q in FP64, its root in FP32.
__device__ __forceinline__ float next_crossing(const DevScene &s, const Ray &r, int k, int out,
                                               int &kn, int &on)
{
    if (!out) {
        double hk = k * s.dz;
        double disc = hk * (2.0 * s.R + hk) - r.q;
        if (disc > 0.0) {
            on = 0;
            if (k == 0) {
                kn = 0;
                return __builtin_inff();
            }
            kn = k - 1;
            return -sqrtf((float)disc) - r.bf;
        }
    }
    on = 1;
    if (k + 1 >= (int)s.L) {
        kn = k;
        return __builtin_inff();
    }
    kn = k + 1;
    double hk = (k + 1) * s.dz;
    double disc = hk * (2.0 * s.R + hk) - r.q;
    return sqrtf((float)(disc > 0.0 ? disc : 0.0)) - r.bf;
}

__device__ __forceinline__ float altitude(const DevScene &s, const Ray &r, float t)
{
    float sv = r.bf + t;
    float w = r.qf + sv * sv;
    return w / (sqrtf(s.R2f + w) + s.Rf);
}

// coefficients_at (sky_trace.wgsl:356-386) for the analytic profiles.
struct Coeffs {
    float sr, sas, so, saa;
};

__device__ __forceinline__ Coeffs medium(const DevScene &s, const DevLambda &l, float h)
{
    Coeffs c{0.0f, 0.0f, 0.0f, 0.0f};
    if (h < 0.0f || h > s.topf) return c;
    float rr = expf(-h * s.inv_hr);
    float ra = expf(-h * s.inv_ha);
    float ro = 1.0f - fabsf(h - s.hc) * s.inv_hw;
    ro = ro < 0.0f ? 0.0f : ro;
    c.sr = l.br * rr;
    c.sas = l.bas * ra;
    c.so = l.bo * ro;
    c.saa = l.baa * ra;
    return c;
}

// ------------------------------------------------------------ phase
__device__ __forceinline__ float phase_rayleigh(float mu)
{
    return 3.0f / (16.0f * kPi) * (1.0f + mu * mu);  // sky_trace.wgsl:477-479
}

__device__ __forceinline__ float phase_hg(float g, float mu)
{
    float den = 1.0f + g * g - 2.0f * g * mu;
    return (1.0f - g * g) / (4.0f * kPi * den * sqrtf(den));
}

// mixed_phase_value, sky_trace.wgsl:496-510 (Rayleigh + HG aerosol).
__device__ __forceinline__ float mixed_phase(const DevScene &s, float sr, float sas, float mu)
{
    float scat = sr + sas;
    if (scat <= 0.0f) return 0.0f;
    return sr / scat * phase_rayleigh(mu) + sas / scat * phase_hg(s.g, mu);
}

// invert_rayleigh_cdf, crates/sky-core/src/sampling.rs:182-185.
__device__ __forceinline__ float sample_rayleigh_mu(float xi)
{
    float y = 4.0f * xi - 2.0f;
    float mu = 2.0f * sinhf(asinhf(y) / 3.0f);
    return fminf(fmaxf(mu, -1.0f), 1.0f);
}

__device__ __forceinline__ float sample_hg_mu(float g, float xi)
{
    if (fabsf(g) < 1e-3f) return 1.0f - 2.0f * xi;
    float sv = (1.0f - g * g) / (1.0f - g + 2.0f * g * xi);
    float mu = (1.0f + g * g - sv * sv) / (2.0f * g);
    return fminf(fmaxf(mu, -1.0f), 1.0f);
}

// Sun cone sample (sky_trace.wgsl:541-546), 1 - cos(theta) formed from
// 1 - cos(r) = 2 sin^2(r/2) so the 4.7 mrad cone keeps its solid angle.
__device__ __forceinline__ f3 sun_cone_dir(const DevScene &s, float xi0, float xi1)
{
    float xi = fminf(xi0, 0.99999994f);
    float x = xi * s.omc;
    float mu = 1.0f - x;
    float st = sqrtf(x * (2.0f - x));
    return frame_dir(f3{s.sun[0], s.sun[1], s.sun[2]}, mu, st, kTau * xi1);
}

// direction_in_cone (sky_trace.wgsl:562-564) as |d - s|^2 <= 2 (1 - cos r).
__device__ __forceinline__ bool in_sun_cone(const DevScene &s, f3 d)
{
    f3 n = normalize_or_y(d);
    float ax = n.x - s.sun[0], ay = n.y - s.sun[1], az = n.z - s.sun[2];
    return ax * ax + ay * ay + az * az <= 2.0f * s.omc;
}

// Camera ray direction (sky_trace.wgsl:614-689).
__device__ __forceinline__ f3 camera_dir(const DevScene &s, uint32_t x, uint32_t y, float j1,
                                         float j2)
{
    if (s.projection == 1u) {
        float u = ((float)x + 0.5f) / (float)s.width;
        float v = ((float)y + 0.5f) / (float)s.height;
        float uu = s.width <= 1 ? 0.5f : (u - 0.5f / (float)s.width) * ((float)s.width / ((float)s.width - 1.0f));
        float vv = s.height <= 1 ? 0.5f : (v - 0.5f / (float)s.height) * ((float)s.height / ((float)s.height - 1.0f));
        uu = fminf(fmaxf(uu, 0.0f), 1.0f);
        vv = fminf(fmaxf(vv, 0.0f), 1.0f);
        float vzc;
        if (vv < 0.75f) {
            float c = vv / 0.75f;
            c = 1.0f - c;
            c = c * c;
            c = 1.0f - c;
            vzc = cosf(s.svl_zha * c);
        } else {
            float c = (vv - 0.75f) / 0.25f;
            c = c * c;
            vzc = cosf(s.svl_zha + s.svl_beta * c);
        }
        float c = uu * uu;
        float lvc = -(c * 2.0f - 1.0f);
        float sv = sqrtf(fmaxf(1.0f - vzc * vzc, 0.0f));
        float ss = sqrtf(fmaxf(1.0f - lvc * lvc, 0.0f));
        f3 v3{sv * (s.svl_fw[0] * lvc + s.svl_side[0] * ss), vzc + sv * (s.svl_fw[1] * lvc + s.svl_side[1] * ss),
              sv * (s.svl_fw[2] * lvc + s.svl_side[2] * ss)};
        return normalize_or_y(v3);
    }
    float u = ((float)x + j1) / (float)s.width;
    float v = ((float)y + j2) / (float)s.height;
    float az = (u - 0.5f) * kTau;
    float el = (0.5f - v) * kPi;
    float se, ce, sa, ca;
    sincosf(el, &se, &ce);
    sincosf(az, &sa, &ca);
    return normalize_or_y(f3{ce * sa, se, ce * ca});
}

}  // namespace sky
