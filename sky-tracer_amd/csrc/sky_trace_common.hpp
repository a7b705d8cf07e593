// sky_trace_common.hpp -- tracking and vertex logic shared by the integrators.
//
// Control flow restates crates/sky-spectral-path-tracer/src/sky_trace.wgsl
// (sample_real_collision 405-439, delta_transmittance 441-475,
// direct_sun_at_scatter 566-585, ground_radiance 587-608, russian_roulette
// 691-701, trace_path 703-776) over the analytic medium, with the shell walk
// of sky_device.hpp in place of the per-segment sphere intersections.
#pragma once

#include "sky_device.hpp"

namespace sky {

enum TrackMode { kTrackCollision = 0, kTrackDelta = 1, kTrackRatio = 2 };

// One tracking pass along r up to r.tmax.  Returns true on a real collision
// (kTrackCollision; t_hit and co filled).  For shadow modes T receives the
// transmittance estimate (0/1 for delta tracking, the ratio-tracking weight).
// Reference form: a fresh free flight per majorant segment, discarded on
// overshoot (wgsl:421-436).  kFlagCarriedTau carries the unused majorant
// optical depth into the next segment instead (same law, fewer draws).
__device__ __forceinline__ bool track(const DevScene &s, uint32_t band, const DevLambda &l, const float *maj_row,
                                      const Ray &r, Rng &rng, int mode, float &t_hit, Coeffs &co,
                                      float &T, bool &wd)
{
    const bool carried = (s.flags & kFlagCarriedTau) != 0;
    float t = 0.0f, tau = -1.0f;
    T = 1.0f;
    int k = r.layer, out = r.outward;
    uint32_t guard = 0;
    for (;;) {
        if (t >= r.tmax) return false;
        if (guard > s.max_segments) {
            wd = true;
            T = 0.0f;
            return false;
        }
        guard++;
        int kn, on;
        float tc = next_crossing(s, global_shells(s), r, k, out, kn, on);
        float seg_end = fminf(tc, r.tmax);
        seg_end = fmaxf(seg_end, t);
        float maj = maj_row[k];
        while (t < seg_end) {
            if (carried) {
                if (tau < 0.0f) tau = free_flight(rng_next(rng));
                float seg_tau = maj * (seg_end - t);
                if (tau >= seg_tau) {
                    tau -= seg_tau;
                    break;
                }
                t = t + f_div(tau, maj);
                tau = -1.0f;
                if (t >= seg_end) break;
            } else {
                t = t + f_div(free_flight(rng_next(rng)), maj);
                if (t >= seg_end) break;
            }
            Coeffs c = medium(s, band, l, altitude(s, r, t));
            float st = c.sr + c.sas + c.so + c.saa;
            float ratio = fminf(fmaxf(f_div(st, maj), 0.0f), 1.0f);
            if (mode == kTrackRatio) {
                T *= 1.0f - ratio;
                continue;
            }
            if (rng_next(rng) < ratio) {
                if (mode == kTrackCollision) {
                    t_hit = t;
                    co = c;
                    return true;
                }
                T = 0.0f;
                return false;
            }
        }
        t = seg_end;
        k = kn;
        out = on;
    }
}

__device__ __forceinline__ float transmittance(const DevScene &s, uint32_t band, const DevLambda &l,
                                               const float *maj_row, const Ray &r, Rng &rng, bool &wd)
{
    float T, th;
    Coeffs co;
    track(s, band, l, maj_row, r, rng, (s.flags & kFlagRatio) ? kTrackRatio : kTrackDelta, th, co, T, wd);
    return T;
}

// russian_roulette, sky_trace.wgsl:691-701 (start depth configurable).
__device__ __forceinline__ bool roulette(const DevScene &s, float &tp, uint32_t depth, Rng &rng)
{
    if (depth < s.rr_start) return true;
    float q = fminf(fmaxf(tp, 0.05f), 0.95f);
    if (rng_next(rng) > q) return false;
    tp = tp / q;
    return true;
}

// trace_path (sky_trace.wgsl:703-776) with the camera ray of main
// (wgsl:799-807), the depth cap, the roulette start depth and optional
// phase importance sampling.  A vertex is its FP32 position p and its
// r^2 - R^2 w (sky_device.hpp); new rays start the ray epsilon from it.
__device__ float trace_path(const DevScene &s, uint32_t x, uint32_t y, uint32_t sample,
                            uint32_t band, bool &wd)
{
    const DevLambda l = s.lam[band];
    const float *maj_row = s.maj + (size_t)band * s.L;
    uint32_t pixel = y * s.width + x;
    Rng rng = rng_make(pixel, sample, band, s);
    float j1 = rng_next(rng);
    float j2 = rng_next(rng);
    f3 dir = camera_dir(s, x, y, j1, j2);
    f3 p{0.0f, s.obs_yf, 0.0f};
    float w = s.obs_wf;
    float tp = 1.0f, L = 0.0f;
    uint32_t depth = 0;
    const uint32_t n_light = s.dls ? s.dls : 1u;
    for (;;) {
        if (s.max_depth && depth >= s.max_depth) break;
        if (depth >= s.watchdog_limit) {
            wd = true;
            return L;
        }
        Ray ray;
        const float e = depth > 0 ? kRayEpsilonKm : 0.0f;
        ray_setup(s, global_shells(s), p, w, dir, e, ray);
        // the main ray's origin becomes the vertex
        p = f3{__builtin_fmaf(dir.x, e, p.x), __builtin_fmaf(dir.y, e, p.y), __builtin_fmaf(dir.z, e, p.z)};
        w = ray.wo;
        if (!ray.hit) break;
        float th, Tdummy;
        Coeffs co;
        if (track(s, band, l, maj_row, ray, rng, kTrackCollision, th, co, Tdummy, wd)) {
            p = f3{__builtin_fmaf(ray.d.x, th, p.x), __builtin_fmaf(ray.d.y, th, p.y),
                   __builtin_fmaf(ray.d.z, th, p.z)};
            w = ray_w(ray.bf, ray.wo, th);
            float scat = co.sr + co.sas;
            float ext = scat + co.so + co.saa;
            if (ext <= 0.0f || scat <= 0.0f) break;
            tp = tp * fminf(fmaxf(scat / ext, 0.0f), 1.0f);
            if (tp <= 0.0f) break;
            // direct_sun_at_scatter (wgsl:566-585)
            float sum = 0.0f;
            for (uint32_t i = 0; i < n_light; ++i) {
                float a0 = rng_next(rng);
                float a1 = rng_next(rng);
                f3 ld = sun_cone_dir(s, a0, a1);
                Ray sr;
                ray_setup(s, global_shells(s), p, w, ld, kRayEpsilonKm, sr);
                if (sr.hit && sr.kind == 1) {
                    float T = transmittance(s, band, l, maj_row, sr, rng, wd);
                    float ph = mixed_phase(s, co.sr, co.sas, dot3(ray.d, ld));
                    sum = sum + l.lsun * T * ph / s.cone_pdf;
                }
            }
            L = L + tp * (sum / (float)n_light);
            f3 nd;
            float w;
            if (s.flags & kFlagPhaseIS) {
                float xl = rng_next(rng);
                float xm = rng_next(rng);
                float mu = xl * scat < co.sr ? sample_rayleigh_mu(xm) : sample_hg_mu(s.g, s.inv_2g, xm);
                float st = f_sqrt(fmaxf(1.0f - mu * mu, 0.0f));
                nd = frame_dir(ray.d, mu, st, rng_next(rng));
                w = 1.0f;
            } else {
                // sample_isotropic, wgsl:534-539
                float z = 1.0f - 2.0f * rng_next(rng);
                float rr = f_sqrt(fmaxf(1.0f - z * z, 0.0f));
                float phi = rng_next(rng);
                nd = f3{rr * f_cos_rev(phi), z, rr * f_sin_rev(phi)};
                w = mixed_phase(s, co.sr, co.sas, dot3(ray.d, nd)) / kIsotropicPdf;
            }
            tp = tp * w;
            if (!roulette(s, tp, depth, rng)) break;
            dir = nd;
        } else if (ray.kind == 1) {
            if (depth == 0 && in_sun_cone(s, dir)) L = L + tp * l.lsun;
            break;
        } else {
            p = f3{__builtin_fmaf(ray.d.x, ray.tmax, p.x), __builtin_fmaf(ray.d.y, ray.tmax, p.y),
                   __builtin_fmaf(ray.d.z, ray.tmax, p.z)};
            w = 0.0f;
            f3 nrm = normalize_or_y(p);
            // ground_radiance (wgsl:587-608)
            float sum = 0.0f;
            for (uint32_t i = 0; i < n_light; ++i) {
                float a0 = rng_next(rng);
                float a1 = rng_next(rng);
                f3 ld = sun_cone_dir(s, a0, a1);
                float cs = fmaxf(dot3(nrm, ld), 0.0f);
                if (cs > 0.0f) {
                    Ray sr;
                    ray_setup(s, global_shells(s), p, w, ld, kRayEpsilonKm, sr);
                    if (sr.hit && sr.kind == 1) {
                        float T = transmittance(s, band, l, maj_row, sr, rng, wd);
                        sum = sum + s.albedo * kInvPi * l.lsun * T * cs / s.cone_pdf;
                    }
                }
            }
            L = L + tp * (sum / (float)n_light);
            // sample_cosine_hemisphere, wgsl:548-560
            float u = rng_next(rng);
            float rr = f_sqrt(u);
            float phi = rng_next(rng);
            float ct = f_sqrt(fmaxf(1.0f - u, 0.0f));
            f3 nd = frame_dir(nrm, ct, rr, phi);
            float co_out = fmaxf(dot3(nrm, nd), 0.0f);
            float pdf = co_out * kInvPi;
            if (pdf <= 0.0f || co_out <= 0.0f) break;
            tp = tp * s.albedo * kInvPi * co_out / pdf;
            if (!roulette(s, tp, depth, rng)) break;
            dir = nd;
        }
        if (L < 0.0f || L != L || tp < 0.0f || tp != tp) return 0.0f;
        depth++;
    }
    return fmaxf(L, 0.0f);
}

}  // namespace sky
