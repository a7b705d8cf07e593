// sky_wavepool.hip -- wave-pool integrator for gfx950.
//
// The reference dispatches one invocation per (x, y, band) and loops over a
// chunk of samples in it (sky_trace.wgsl:778-816), so a wave runs as long as
// its longest path.  Here each wave64 owns a work item -- a group of pixels x
// all wavelengths x a range of samples -- and its lanes pull (pixel, band,
// sample) triples from the item as soon as their current path ends.  Every
// path is keyed exactly as the reference keys it (pixel, sample, band, dim),
// so which lane traces it changes nothing but the order of the float sums.
//
// A lane is a small state machine with a single tracking call site:
//   EVENT : vertex/boundary/shadow result (or a new path) waiting for the
//           event handler, which samples directions, sets up the next ray
//           (FP64) and either returns the lane to TRACK or finishes the path;
//   TRACK : one tracking micro-step per loop trip (at most one RNG draw):
//           a majorant-shell crossing, a tentative collision, or its accept
//           test (sample_real_collision / delta_transmittance,
//           sky_trace.wgsl:405-475);
//   NEED  : path finished, waits for the next triple;  IDLE : item drained.
// The event handler runs when at least `event_threshold` lanes wait for it
// (or nothing is tracking); up to `track_steps` micro-steps run between two
// handler passes.  Per-(pixel, band) sums live in LDS (ds_add_f32) and are
// written once per item to a [chunk][pixel][band] partial buffer that
// k_finalize reduces in chunk order.
#include "sky_internal.hpp"
#include "sky_trace_common.hpp"

namespace sky {

constexpr int kWavesPerBlock = 4;
constexpr int kSlotsPerWave = 256;

enum LaneState : int { ST_NEED = 0, ST_EVENT = 1, ST_TRACK = 2, ST_IDLE = 3 };
enum LaneEvent : int { EV_START = 0, EV_HIT = 1, EV_BOUNDARY = 2, EV_SHADOW = 3 };

__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & 63u; }

__device__ __forceinline__ uint32_t mbcnt(uint64_t m)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

// Everything a lane keeps between loop trips.  The main ray starts at the
// current vertex (px, py, pz): the camera position, or the last vertex moved
// by the ray epsilon, exactly as the reference offsets new rays
// (sky_trace.wgsl:744, 766).  During a shadow ray the vertex stays put.
struct LaneCtx {
    int st, ev, mode, vtype, kind, pend;
    uint32_t slot, band;
    float br, bo, bas, baa;  // this band's coefficients (medium evaluation)
    Rng rng;
    float tp, L;
    uint32_t depth;
    double px, py, pz;  // current vertex / main-ray origin (FP64)
    f3 vin;             // incoming direction at the vertex (camera direction at start)
    float nee_sum, nee_factor, rw;
    uint32_t nee_i;
    // ray being tracked
    f3 d;
    float bf, qf, tmax;
    double q;
    float t, tau, seg_end, maj;
    float aux;  // accept ratio pending (delta) / transmittance weight (ratio, shadow result)
    int k, out, kn, on;
    uint32_t guard;
};

// ------------------------------------------------------------ adjoint
// The replay pass of sky_render_adjoint runs the same state machine with
// ADJ = true.  With the majorant fixed, the law of a path depends on the
// parameters only through the accept/reject tests (probability sigma_t / mu),
// the albedo weight sigma_s / sigma_t, the phase function (sampled or used as
// a weight), the ground albedo and the shadow-ray tests, so
//   d E[sum_k C_k] = E[ sum_k C_k (d log C_k + S_k) ],
// where S_k is the summed log-derivative of the decisions before C_k.  Each
// lane carries S (main path), SS (current NEE sample: S plus its own terms)
// and G (this path's gradient), all 10-vectors; Russian roulette is held
// detached.
constexpr int kNP = 10;

struct AdjLane {
    float S[kNP], SS[kNP], G[kNP];
    float w;         // dLoss / dL of this path's (pixel, band), divided by spp
    float vsr, vsas; // sigma_R and sigma_As at the scattering vertex
    float vh;        // altitude of the scattering vertex
};

// partials of sigma_R, sigma_O and the two aerosol coefficients at altitude
// h with respect to the ten parameters (sky_hip.h order).
struct DSig {
    float r0, r1;      // sigma_R: d/d beta_R, d/d H_R
    float o2, o3;      // sigma_O: d/d peak, d/d centre
    float as4, as6, as8;  // sigma_As: d/d beta_As, d/d H_A, d/d angstrom
    float aa5, aa6, aa8;  // sigma_Aa: d/d beta_Aa, d/d H_A, d/d angstrom
};

__device__ __forceinline__ DSig dsig(const DevScene &s, const AdjArgs &A, uint32_t band, float h, float bo,
                                     const Coeffs &co)
{
    DSig d{};
    if (h < 0.0f || h > s.topf) return d;
    const float4 u = A.lam_units[band];
    float rr = expf(-h * s.inv_hr);
    float ra = expf(-h * s.inv_ha);
    float dh = h - s.hc;
    float tent = fmaxf(1.0f - fabsf(dh) * s.inv_hw, 0.0f);
    d.r0 = u.x * rr;
    d.r1 = co.sr * h * A.inv_hr2;
    d.o2 = u.y * tent;
    d.o3 = fabsf(dh) * s.inv_hw < 1.0f ? bo * (dh > 0.0f ? s.inv_hw : -s.inv_hw) : 0.0f;
    d.as4 = u.z * ra;
    d.aa5 = u.z * ra;
    d.as6 = co.sas * h * A.inv_ha2;
    d.aa6 = co.saa * h * A.inv_ha2;
    d.as8 = -co.sas * u.w;
    d.aa8 = -co.saa * u.w;
    return d;
}

// v += f * d sigma_t
__device__ __forceinline__ void add_dsigma_t(float *v, const DSig &d, float f)
{
    v[0] += f * d.r0;
    v[1] += f * d.r1;
    v[2] += f * d.o2;
    v[3] += f * d.o3;
    v[4] += f * d.as4;
    v[5] += f * d.aa5;
    v[6] += f * (d.as6 + d.aa6);
    v[8] += f * (d.as8 + d.aa8);
}

// v += f * d sigma_s
__device__ __forceinline__ void add_dsigma_s(float *v, const DSig &d, float f)
{
    v[0] += f * d.r0;
    v[1] += f * d.r1;
    v[4] += f * d.as4;
    v[6] += f * d.as6;
    v[8] += f * d.as8;
}

// v += d log p(mu), p = (sigma_R p_R + sigma_As p_HG) / sigma_s at the vertex.
__device__ __forceinline__ void add_dlog_phase(const DevScene &s, const AdjArgs &A, const AdjLane &a,
                                               uint32_t band, float mu, float bo, float *v)
{
    Coeffs co{a.vsr, a.vsas, 0.0f, 0.0f};
    // the aerosol absorption share of sigma_As's partials equals sigma_As's
    DSig d = dsig(s, A, band, a.vh, bo, co);
    float pr = phase_rayleigh(mu), ph = phase_hg(s.g, mu);
    float num = a.vsr * pr + a.vsas * ph;
    float scat = a.vsr + a.vsas;
    if (!(num > 0.0f) || !(scat > 0.0f)) return;
    float inum = 1.0f / num, iscat = 1.0f / scat;
    float fr = pr * inum - iscat, fa = ph * inum - iscat;
    v[0] += fr * d.r0;
    v[1] += fr * d.r1;
    v[4] += fa * d.as4;
    v[6] += fa * d.as6;
    v[8] += fa * d.as8;
    float g = s.g, den = 1.0f + g * g - 2.0f * g * mu;
    float dph = ph * (-2.0f * g / (1.0f - g * g) - 3.0f * (g - mu) / den);
    v[7] += a.vsas * dph * inum;
}

// Majorant of (band, shell): from the wave's LDS copy of the job's grid when
// it fits (nlam * L <= kMajLds), else from global memory.  A crossing needs
// it immediately, so the LDS copy removes a global-load latency from every
// shell crossing.
constexpr int kMajLds = 1024;

struct MajSrc {
    const float *lds;  // wave's LDS copy or null
};

__device__ __forceinline__ float majorant(const DevScene &s, const MajSrc &m, uint32_t band, int k)
{
    uint32_t i = band * s.L + (uint32_t)k;
    return m.lds ? m.lds[i] : s.maj[i];
}

__device__ __forceinline__ void seg_begin(const DevScene &s, const MajSrc &m, LaneCtx &c)
{
    Ray r;
    r.q = c.q;
    r.bf = c.bf;
    float tc = next_crossing(s, r, c.k, c.out, c.kn, c.on);
    c.seg_end = fmaxf(fminf(tc, c.tmax), c.t);
    c.maj = majorant(s, m, c.band, c.k);
}

// Sets up a ray from (ox, oy, oz) along d (FP64 set-up, sky_device.hpp) and
// starts tracking it; returns false if it has no boundary.
__device__ __forceinline__ bool ray_begin(const DevScene &s, const MajSrc &m, LaneCtx &c, double ox,
                                          double oy, double oz, f3 d, int mode)
{
    Ray r;
    ray_setup(s, ox, oy, oz, d, r);
    c.d = d;
    c.bf = r.bf;
    c.qf = r.qf;
    c.q = r.q;
    c.tmax = r.tmax;
    c.kind = r.kind;
    c.mode = mode;
    c.t = 0.0f;
    c.tau = -1.0f;
    c.aux = 1.0f;
    c.pend = 0;
    c.guard = 1;
    c.k = r.layer;
    c.out = r.outward;
    seg_begin(s, m, c);
    c.st = ST_TRACK;
    return r.hit != 0;
}

__device__ __forceinline__ float lane_altitude(const DevScene &s, const LaneCtx &c, float t)
{
    float sv = c.bf + t;
    float w = c.qf + sv * sv;
    return w / (sqrtf(s.R2f + w) + s.Rf);
}

__device__ __forceinline__ Coeffs lane_medium(const DevScene &s, const LaneCtx &c, float h)
{
    DevLambda l;
    l.br = c.br;
    l.bo = c.bo;
    l.bas = c.bas;
    l.baa = c.baa;
    return medium(s, l, h);
}

// One tracking micro-step: at most one RNG draw.  An empty segment (its
// end clamped to t) is crossed without a draw, as the reference's inner
// `while t < segment_end` loop does.
template <bool ADJ>
__device__ __forceinline__ void track_step(const DevScene &s, const MajSrc &m, LaneCtx &c, AdjLane &a,
                                           const AdjArgs &A, bool &wd)
{
    const bool carried = (s.flags & kFlagCarriedTau) != 0;
    const bool empty = !(c.t < c.seg_end);
    const bool need_draw = c.pend || (!empty && (!carried || c.tau < 0.0f));
    float u = 0.0f;
    if (need_draw) u = rng_next(c.rng);
    if (c.pend) {
        c.pend = 0;
        if (u < c.aux) {
            c.st = ST_EVENT;
            if (c.mode == kTrackCollision) {
                c.ev = EV_HIT;
            } else {
                c.aux = 0.0f;
                c.ev = EV_SHADOW;
            }
        } else {
            if constexpr (ADJ) {
                // null collision: d log(1 - sigma_t / mu) = -d sigma_t / (mu - sigma_t)
                Coeffs co = lane_medium(s, c, lane_altitude(s, c, c.t));
                float st = co.sr + co.sas + co.so + co.saa;
                float f = -1.0f / fmaxf(c.maj - st, 1e-30f);
                DSig d = dsig(s, A, c.band, lane_altitude(s, c, c.t), c.bo, co);
                add_dsigma_t(c.mode == kTrackCollision ? a.S : a.SS, d, f);
            }
            c.aux = 1.0f;
        }
        return;
    }
    bool cross = true;
    float tn = c.t;
    if (!empty) {
        float ff = 0.0f;
        if (need_draw) ff = -logf(fmaxf(1.0f - u, 0.0000001f));
        if (carried) {
            if (c.tau < 0.0f) c.tau = ff;
            float seg_tau = c.maj * (c.seg_end - c.t);
            cross = c.tau >= seg_tau;
            if (cross) {
                c.tau -= seg_tau;
            } else {
                tn = c.t + c.tau / c.maj;
                c.tau = -1.0f;
                cross = tn >= c.seg_end;
            }
        } else {
            tn = c.t + ff / c.maj;
            cross = tn >= c.seg_end;
        }
    }
    if (cross) {
        c.t = c.seg_end;
        if (c.t >= c.tmax) {
            c.st = ST_EVENT;
            c.ev = c.mode == kTrackCollision ? EV_BOUNDARY : EV_SHADOW;
            return;
        }
        if (c.guard > s.max_segments) {
            wd = true;
            c.st = ST_EVENT;
            c.aux = 0.0f;
            c.ev = c.mode == kTrackCollision ? EV_BOUNDARY : EV_SHADOW;
            c.kind = 2;  // watchdog: ends the path without contribution
            return;
        }
        c.guard++;
        c.k = c.kn;
        c.out = c.on;
        seg_begin(s, m, c);
        return;
    }
    c.t = tn;
    Coeffs co = lane_medium(s, c, lane_altitude(s, c, c.t));
    float st = co.sr + co.sas + co.so + co.saa;
    float ratio = fminf(fmaxf(st / c.maj, 0.0f), 1.0f);
    if (c.mode == kTrackRatio) {
        c.aux *= 1.0f - ratio;
        if constexpr (ADJ) {
            float f = -1.0f / fmaxf(c.maj - st, 1e-30f);
            DSig d = dsig(s, A, c.band, lane_altitude(s, c, c.t), c.bo, co);
            add_dsigma_t(a.SS, d, f);
        }
    } else {
        c.pend = 1;
        c.aux = ratio;
    }
}

// Vertex at distance t along the main ray from (px, py, pz), FP64.
__device__ __forceinline__ void advance_vertex(LaneCtx &c, float t)
{
    double dx = c.d.x, dy = c.d.y, dz = c.d.z;
    double inv = 1.0 / sqrt(dx * dx + dy * dy + dz * dz);
    c.px += dx * inv * (double)t;
    c.py += dy * inv * (double)t;
    c.pz += dz * inv * (double)t;
}

enum Act : int { ACT_NONE = 0, ACT_NEE = 1, ACT_CONT = 2, ACT_MAIN = 3, ACT_FINISH = 4, ACT_TRACKING = 5 };

// Event handler: sky_trace.wgsl:703-776 between two tracking passes.
template <bool ADJ>
__device__ __forceinline__ void handle_event(const DevScene &s, const MajSrc &m, LaneCtx &c, AdjLane &a,
                                             const AdjArgs &A, float *acc_sum, float *acc_sq, float *acc_grad,
                                             bool &wd)
{
    const uint32_t n_light = s.dls ? s.dls : 1u;
    int act = ACT_NONE;
    f3 mdir = c.vin;  // direction of the next main ray
    if (c.ev == EV_START) {
        act = ACT_MAIN;
    } else if (c.ev == EV_BOUNDARY) {
        if (c.kind == 1) {
            // exit: the sun disk is seen only by the camera ray (wgsl:745-749)
            if (c.depth == 0 && in_sun_cone(s, c.d)) {
                float C = c.tp * s.lam[c.band].lsun;
                c.L = c.L + C;
                if constexpr (ADJ) {
                    for (int k = 0; k < kNP; ++k) a.G[k] += a.w * C * a.S[k];
                }
            }
            act = ACT_FINISH;
        } else if (c.kind == 2) {
            act = ACT_FINISH;  // tracking watchdog
        } else {
            advance_vertex(c, c.tmax);
            c.vtype = 1;
            c.nee_i = 0;
            c.nee_sum = 0.0f;
            act = ACT_NEE;
        }
    } else if (c.ev == EV_HIT) {
        const float hv = lane_altitude(s, c, c.t);
        Coeffs co = lane_medium(s, c, hv);
        advance_vertex(c, c.t);
        float scat = co.sr + co.sas;
        float ext = scat + co.so + co.saa;
        act = ACT_FINISH;
        if constexpr (ADJ) {
            // accepted collision times the albedo weight: d log(sigma_s / mu)
            if (scat > 0.0f) add_dsigma_s(a.S, dsig(s, A, c.band, hv, c.bo, co), 1.0f / scat);
            a.vsr = co.sr;
            a.vsas = co.sas;
            a.vh = hv;
        }
        if (ext > 0.0f && scat > 0.0f) {
            c.tp = c.tp * fminf(fmaxf(scat / ext, 0.0f), 1.0f);
            if (c.tp > 0.0f) {
                c.vtype = 0;
                c.vin = c.d;
                c.rw = co.sr / scat;
                c.nee_i = 0;
                c.nee_sum = 0.0f;
                act = ACT_NEE;
            }
        }
    } else {  // EV_SHADOW
        c.nee_sum += c.nee_factor * c.aux;
        if constexpr (ADJ) {
            float C = c.tp * c.nee_factor * c.aux / (float)n_light;
            if (C != 0.0f)
                for (int k = 0; k < kNP; ++k) a.G[k] += a.w * C * a.SS[k];
        }
        c.nee_i++;
        act = c.nee_i < n_light ? ACT_NEE : ACT_CONT;
    }

    // Sun NEE: direct_sun_at_scatter / ground_radiance (wgsl:566-608).
    while (act == ACT_NEE) {
        float a0 = rng_next(c.rng);
        float a1 = rng_next(c.rng);
        f3 ld = sun_cone_dir(s, a0, a1);
        float factor;
        const float lsun = s.lam[c.band].lsun;
        if (c.vtype == 1) {
            double pn = 1.0 / sqrt(c.px * c.px + c.py * c.py + c.pz * c.pz);
            float cs = fmaxf((float)(c.px * pn) * ld.x + (float)(c.py * pn) * ld.y + (float)(c.pz * pn) * ld.z, 0.0f);
            factor = s.albedo * kInvPi * lsun * cs / s.cone_pdf;
            if constexpr (ADJ) {
                for (int k = 0; k < kNP; ++k) a.SS[k] = a.S[k];
                if (s.albedo > 0.0f) a.SS[9] += 1.0f / s.albedo;
            }
            if (!(cs > 0.0f)) {
                c.nee_i++;
                act = c.nee_i < n_light ? ACT_NEE : ACT_CONT;
                continue;
            }
        } else {
            float mu = dot3(c.vin, ld);
            float ph = c.rw * phase_rayleigh(mu) + (1.0f - c.rw) * phase_hg(s.g, mu);
            factor = lsun * ph / s.cone_pdf;
            if constexpr (ADJ) {
                for (int k = 0; k < kNP; ++k) a.SS[k] = a.S[k];
                add_dlog_phase(s, A, a, c.band, mu, c.bo, a.SS);
            }
        }
        bool hit = ray_begin(s, m, c, c.px + (double)ld.x * kRayEpsilonKm, c.py + (double)ld.y * kRayEpsilonKm,
                             c.pz + (double)ld.z * kRayEpsilonKm, ld,
                             (s.flags & kFlagRatio) ? kTrackRatio : kTrackDelta);
        if (!(hit && c.kind == 1)) {
            c.st = ST_EVENT;
            c.nee_i++;
            act = c.nee_i < n_light ? ACT_NEE : ACT_CONT;
            continue;
        }
        c.nee_factor = factor;
        act = ACT_TRACKING;
    }

    if (act == ACT_CONT) {
        c.L = c.L + c.tp * (c.nee_sum / (float)n_light);
        f3 nd;
        float w;
        bool ok = true;
        if (c.vtype == 0) {
            const float rw = c.rw;
            if (s.flags & kFlagPhaseIS) {
                float xl = rng_next(c.rng);
                float xm = rng_next(c.rng);
                float mu = xl < rw ? sample_rayleigh_mu(xm) : sample_hg_mu(s.g, xm);
                float st = sqrtf(fmaxf(1.0f - mu * mu, 0.0f));
                nd = frame_dir(c.vin, mu, st, kTau * rng_next(c.rng));
                w = 1.0f;
                if constexpr (ADJ) add_dlog_phase(s, A, a, c.band, dot3(c.vin, nd), c.bo, a.S);
            } else {
                float z = 1.0f - 2.0f * rng_next(c.rng);
                float rr = sqrtf(fmaxf(1.0f - z * z, 0.0f));
                float phi = kTau * rng_next(c.rng);
                float sp, cp;
                sincosf(phi, &sp, &cp);
                nd = f3{rr * cp, z, rr * sp};
                float mu = dot3(c.vin, nd);
                w = (rw * phase_rayleigh(mu) + (1.0f - rw) * phase_hg(s.g, mu)) / kIsotropicPdf;
                if constexpr (ADJ) add_dlog_phase(s, A, a, c.band, mu, c.bo, a.S);
            }
        } else {
            double pn = 1.0 / sqrt(c.px * c.px + c.py * c.py + c.pz * c.pz);
            f3 nrm{(float)(c.px * pn), (float)(c.py * pn), (float)(c.pz * pn)};
            float u = rng_next(c.rng);
            float rr = sqrtf(u);
            float phi = kTau * rng_next(c.rng);
            float ct = sqrtf(fmaxf(1.0f - u, 0.0f));
            nd = frame_dir(nrm, ct, rr, phi);
            float co_out = fmaxf(dot3(nrm, nd), 0.0f);
            float pdf = co_out * kInvPi;
            ok = pdf > 0.0f && co_out > 0.0f;
            w = ok ? s.albedo * kInvPi * co_out / pdf : 0.0f;
            if constexpr (ADJ) {
                if (s.albedo > 0.0f) a.S[9] += 1.0f / s.albedo;
            }
        }
        act = ACT_FINISH;
        if (ok) {
            c.tp = c.tp * w;
            if (roulette(s, c.tp, c.depth, c.rng)) {
                if (c.L < 0.0f || c.L != c.L || c.tp < 0.0f || c.tp != c.tp) {
                    c.L = 0.0f;
                    if constexpr (ADJ) {
                        for (int k = 0; k < kNP; ++k) a.G[k] = 0.0f;
                    }
                } else {
                    c.depth++;
                    c.px += (double)nd.x * kRayEpsilonKm;
                    c.py += (double)nd.y * kRayEpsilonKm;
                    c.pz += (double)nd.z * kRayEpsilonKm;
                    mdir = nd;
                    act = ACT_MAIN;
                }
            }
        }
    }

    if (act == ACT_MAIN) {
        // loop top of trace_path (wgsl:709-717) plus the depth cap
        act = ACT_FINISH;
        if (!(s.max_depth && c.depth >= s.max_depth)) {
            if (c.depth >= s.watchdog_limit) {
                wd = true;
            } else if (ray_begin(s, m, c, c.px, c.py, c.pz, mdir, kTrackCollision)) {
                act = ACT_TRACKING;
            }
        }
    }

    if (act == ACT_FINISH) {
        float Lf = fmaxf(c.L, 0.0f);
        atomicAdd(&acc_sum[c.slot], Lf);
        atomicAdd(&acc_sq[c.slot], Lf * Lf);
        if constexpr (ADJ) {
            for (int k = 0; k < kNP; ++k)
                if (a.G[k] != 0.0f && a.G[k] == a.G[k]) atomicAdd(&acc_grad[k], a.G[k]);
        }
        c.st = ST_NEED;
    }
}

// Path start (main, sky_trace.wgsl:799-807): triple j of the item.
template <bool ADJ>
__device__ __forceinline__ void start_path(const DevScene &s, LaneCtx &c, AdjLane &a, const AdjArgs &A,
                                           uint32_t j, uint32_t p0, uint32_t np, uint32_t s0)
{
    uint32_t per_sample = np * s.nlam;
    uint32_t si = j / per_sample;
    uint32_t rem = j - si * per_sample;
    uint32_t pl = rem / s.nlam;
    c.band = rem - pl * s.nlam;
    c.slot = rem;
    uint32_t pixel = p0 + pl;
    uint32_t x = pixel % s.width, y = pixel / s.width;
    const DevLambda l = s.lam[c.band];
    c.br = l.br;
    c.bo = l.bo;
    c.bas = l.bas;
    c.baa = l.baa;
    c.rng = rng_make(pixel, s0 + si, c.band, s);
    float j1 = rng_next(c.rng);
    float j2 = rng_next(c.rng);
    c.vin = camera_dir(s, x, y, j1, j2);
    c.px = 0.0;
    c.py = s.obs_y;
    c.pz = 0.0;
    c.tp = 1.0f;
    c.L = 0.0f;
    c.depth = 0;
    c.st = ST_EVENT;
    c.ev = EV_START;
    if constexpr (ADJ) {
        const float *g = A.dl_dxyz + 3 * (size_t)pixel;
        a.w = (g[0] * l.wx + g[1] * l.wy + g[2] * l.wz) * A.inv_spp;
        for (int k = 0; k < kNP; ++k) {
            a.S[k] = 0.0f;
            a.G[k] = 0.0f;
        }
    }
}

// Item i of the launch belongs to the job j with jobs[j].item_base <= i <
// jobs[j+1].item_base (jobs are stored in launch order, largest first).
__device__ __forceinline__ uint32_t find_job(const PoolArgs &a, uint32_t item)
{
    uint32_t lo = 0, hi = a.n_jobs;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (a.jobs[mid].item_base <= item)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

template <bool ADJ>
__device__ __forceinline__ void wavepool_body(const PoolArgs &a, const AdjArgs &A)
{
    __shared__ float lds_sum[kWavesPerBlock][kSlotsPerWave];
    __shared__ float lds_sq[kWavesPerBlock][kSlotsPerWave];
    __shared__ float lds_grad[kWavesPerBlock][ADJ ? 16 : 1];
    __shared__ float lds_maj[kWavesPerBlock][kMajLds];
    const DevJob *loaded = nullptr;  // job whose majorants are in lds_maj
    const uint32_t wid = threadIdx.x >> 6;
    const uint32_t lane = lane_id();
    float *acc_sum = lds_sum[wid];
    float *acc_sq = lds_sq[wid];
    float *acc_grad = lds_grad[wid];
    bool wd = false;
    uint32_t *diag = nullptr;
    LaneCtx c;
    AdjLane al;
    for (;;) {
        uint32_t item = 0;
        if (lane == 0) item = atomicAdd(a.counter, 1u);
        item = __builtin_amdgcn_readfirstlane(item);
        if (item >= a.total_items) break;
        if (item == 0 && lane == 0) *a.t_start = wall_clock64();
        const DevJob &J = a.jobs[find_job(a, item)];
        const DevScene &s = J.s;
        diag = s.diag;
        const uint32_t npx = s.width * s.height;
        const uint32_t local = item - J.item_base;
        const uint32_t g = local % J.groups, ch = local / J.groups;
        const uint32_t p0 = g * J.pixels_per_item;
        const uint32_t np = min(J.pixels_per_item, npx - p0);
        const uint32_t s0 = s.sample_offset + ch * J.samples_per_chunk;
        const uint32_t ns = min(J.samples_per_chunk, s.spp - ch * J.samples_per_chunk);
        const uint32_t nslots = np * s.nlam;
        const uint32_t total = nslots * ns;
        for (uint32_t i = lane; i < nslots; i += 64) {
            acc_sum[i] = 0.0f;
            acc_sq[i] = 0.0f;
        }
        MajSrc msrc{nullptr};
        const uint32_t nmaj = s.nlam * s.L;
        if (nmaj <= (uint32_t)kMajLds) {
            msrc.lds = lds_maj[wid];
            if (loaded != &J) {
                for (uint32_t i = lane; i < nmaj; i += 64) lds_maj[wid][i] = s.maj[i];
                loaded = &J;
            }
        }
        if constexpr (ADJ) {
            if (lane < kNP) acc_grad[lane] = 0.0f;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        uint32_t next = 0;
        c.st = ST_NEED;
        for (;;) {
            uint64_t need = ballot(c.st == ST_NEED);
            if (need) {
                if (c.st == ST_NEED) {
                    uint32_t j = next + mbcnt(need);
                    if (j < total)
                        start_path<ADJ>(s, c, al, A, j, p0, np, s0);
                    else
                        c.st = ST_IDLE;
                }
                next += (uint32_t)__popcll(need);
            }
            uint64_t evm = ballot(c.st == ST_EVENT);
            uint64_t trm = ballot(c.st == ST_TRACK);
            if (!evm && !trm) break;
            if (evm && ((uint32_t)__popcll(evm) >= a.event_threshold || !trm)) {
                if (c.st == ST_EVENT) handle_event<ADJ>(s, msrc, c, al, A, acc_sum, acc_sq, acc_grad, wd);
            }
            for (uint32_t i = 0; i < a.track_steps; ++i) {
                if (!ballot(c.st == ST_TRACK)) break;
                if (c.st == ST_TRACK) track_step<ADJ>(s, msrc, c, al, A, wd);
            }
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        float2 *dst = J.accum + ((size_t)ch * npx + p0) * s.nlam;
        for (uint32_t i = lane; i < nslots; i += 64) dst[i] = make_float2(acc_sum[i], acc_sq[i]);
        if constexpr (ADJ) {
            if (lane < kNP) atomicAdd(&A.grad[lane], (double)acc_grad[lane]);
        }
        __builtin_amdgcn_wave_barrier();
        if (lane == 0) {
            __threadfence();
            uint32_t done = atomicAdd(J.done, 1u);
            if (done + 1 == J.n_items) *J.t_done = wall_clock64();
        }
    }
    if (wd && diag) atomicOr(diag, 1u);
}

// The same integrator under two register budgets: the compiler's own
// allocation (3 waves/SIMD) and a 128-VGPR cap (4 waves/SIMD, with a few
// spills in the event handler).  The autotuner picks one per workload.
__global__ __launch_bounds__(256) void k_wavepool(PoolArgs a) { wavepool_body<false>(a, AdjArgs{}); }

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_wavepool_o4(PoolArgs a)
{
    wavepool_body<false>(a, AdjArgs{});
}

// Path-replay adjoint: the forward state machine plus the score bookkeeping.
__global__ __launch_bounds__(256) void k_wavepool_adjoint(PoolArgs a, AdjArgs A) { wavepool_body<true>(a, A); }

// Per job (blockIdx.y): reduces the chunk partials in chunk order,
// normalises by 1/spp and clamps (integrator.rs:307-310), and forms
// XYZ = sum_band film * cmf * weight.
__global__ __launch_bounds__(256) void k_finalize(const DevJob *jobs)
{
    const DevJob &J = jobs[blockIdx.y];
    const DevScene &s = J.s;
    uint32_t pixel = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t npx = s.width * s.height;
    if (pixel >= npx) return;
    float inv = 1.0f / (float)s.spp;
    float X = 0.0f, Y = 0.0f, Z = 0.0f;
    for (uint32_t b = 0; b < s.nlam; ++b) {
        float sum = 0.0f, sq = 0.0f;
        for (uint32_t ch = 0; ch < J.chunks; ++ch) {
            float2 v = J.accum[((size_t)ch * npx + pixel) * s.nlam + b];
            sum += v.x;
            sq += v.y;
        }
        float f = fmaxf(sum * inv, 0.0f);
        if (s.film) s.film[(size_t)pixel * s.nlam + b] = f;
        if (s.film_m2) s.film_m2[(size_t)pixel * s.nlam + b] = sq * inv;
        DevLambda l = s.lam[b];
        X += f * l.wx;
        Y += f * l.wy;
        Z += f * l.wz;
    }
    if (s.xyz) {
        s.xyz[3 * pixel + 0] = X;
        s.xyz[3 * pixel + 1] = Y;
        s.xyz[3 * pixel + 2] = Z;
    }
}

// Fills n words with zero: used instead of a memset node so that the launch
// sequence can be captured in a HIP graph.
__global__ void k_zero_u32(uint32_t *p, uint32_t n)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

void plan_job(const DevScene &s, const sky_tuning &t, int num_cus, uint32_t min_items, DevJob &j)
{
    const uint32_t npx = s.width * s.height;
    uint32_t cap = kSlotsPerWave / s.nlam;
    uint32_t want = t.pixels_per_item ? t.pixels_per_item : (128u / s.nlam ? 128u / s.nlam : 1u);
    j.pixels_per_item = want < 1 ? 1 : (want > cap ? cap : want);
    j.groups = (npx + j.pixels_per_item - 1) / j.pixels_per_item;
    // split samples only as far as needed for min_items, keeping >= 16 samples per item
    uint64_t chunks = ((uint64_t)min_items + j.groups - 1) / j.groups;
    uint64_t max_chunks = s.spp / 16 ? s.spp / 16 : 1;
    if (chunks > max_chunks) chunks = max_chunks;
    if (chunks < 1) chunks = 1;
    (void)num_cus;
    j.samples_per_chunk = (uint32_t)((s.spp + chunks - 1) / chunks);
    j.chunks = (s.spp + j.samples_per_chunk - 1) / j.samples_per_chunk;
    j.n_items = j.groups * j.chunks;
}

size_t job_accum_bytes(const DevJob &j)
{
    return (size_t)j.chunks * j.s.width * j.s.height * j.s.nlam * sizeof(float2);
}

WavePlan plan_launch(const sky_tuning &t, int num_cus, uint32_t total_items)
{
    WavePlan p;
    uint32_t wpc = t.waves_per_cu ? t.waves_per_cu : 16u;
    uint64_t waves = (uint64_t)(num_cus > 0 ? num_cus : 256) * wpc;
    uint64_t w = waves < total_items ? waves : total_items;
    p.block = 64 * kWavesPerBlock;
    p.grid = (uint32_t)((w + kWavesPerBlock - 1) / kWavesPerBlock);
    p.n_items = total_items;
    p.event_threshold = t.event_threshold ? t.event_threshold : 32u;
    p.track_steps = t.track_steps ? t.track_steps : 8u;
    p.occupancy = t.occupancy;
    return p;
}

uint32_t launch_min_items(const sky_tuning &t, int num_cus)
{
    uint32_t wpc = t.waves_per_cu ? t.waves_per_cu : 16u;
    return 4u * (uint32_t)(num_cus > 0 ? num_cus : 256) * wpc;
}

// meta: [0] item counter, [1..2] launch timestamp, [4 + 2 j ..] per-job done
// counter / pad, then per-job finish timestamps; zeroed by the first kernel.
hipError_t launch_pool(const DevJob *d_jobs, uint32_t n_jobs, uint32_t max_npx, const WavePlan &p,
                       uint32_t *meta, uint32_t meta_words, hipStream_t stream)
{
    hipLaunchKernelGGL(k_zero_u32, dim3((meta_words + 255) / 256), dim3(256), 0, stream, meta, meta_words);
    PoolArgs a;
    a.jobs = d_jobs;
    a.n_jobs = n_jobs;
    a.total_items = p.n_items;
    a.counter = meta;
    a.t_start = (unsigned long long *)(meta + 2);
    a.event_threshold = p.event_threshold;
    a.track_steps = p.track_steps;
    if (p.occupancy == 4)
        hipLaunchKernelGGL(k_wavepool_o4, dim3(p.grid), dim3(p.block), 0, stream, a);
    else
        hipLaunchKernelGGL(k_wavepool, dim3(p.grid), dim3(p.block), 0, stream, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_finalize, dim3((max_npx + 255) / 256, n_jobs), dim3(256), 0, stream, d_jobs);
    return hipGetLastError();
}

hipError_t launch_adjoint(const DevJob *d_jobs, const WavePlan &p, uint32_t *meta, uint32_t meta_words,
                          const AdjArgs &A, hipStream_t stream)
{
    hipLaunchKernelGGL(k_zero_u32, dim3((meta_words + 255) / 256), dim3(256), 0, stream, meta, meta_words);
    PoolArgs a;
    a.jobs = d_jobs;
    a.n_jobs = 1;
    a.total_items = p.n_items;
    a.counter = meta;
    a.t_start = (unsigned long long *)(meta + 2);
    a.event_threshold = p.event_threshold;
    a.track_steps = p.track_steps;
    hipLaunchKernelGGL(k_wavepool_adjoint, dim3(p.grid), dim3(p.block), 0, stream, a, A);
    return hipGetLastError();
}

hipError_t launch_zero(uint32_t *p, uint32_t n, hipStream_t stream)
{
    hipLaunchKernelGGL(k_zero_u32, dim3((n + 255) / 256), dim3(256), 0, stream, p, n);
    return hipGetLastError();
}

}  // namespace sky
