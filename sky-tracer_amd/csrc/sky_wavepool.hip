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
// A lane is a small state machine:
//   TRACK : one tracking step per loop trip: the shell crossings up to the
//           next tentative collision (cheap, looped), then that collision
//           and its accept test (sample_real_collision / delta_transmittance,
//           sky_trace.wgsl:405-475), or the end of the ray;
//   EVENT : a vertex/boundary/shadow result waiting for the event handler.
//           One handler pass resolves the event, draws the step's random
//           numbers and samples ONE direction (sun cone, phase lobe, cosine
//           lobe); lanes whose path ended pull their next triple (camera
//           ray); then ONE ray set-up (FP64) serves every lane that needs a
//           ray, so the lanes of a pass run the same instructions whatever
//           their event was;
//   RAY   : direction chosen, ray set-up pending (inside a handler pass);
//   NEED  : path finished, waits for the next triple;  IDLE : item drained.
// The event handler runs when at least `event_threshold` lanes wait for it
// (or nothing is tracking); up to `track_steps` tracking steps run between
// two handler passes.  Per-(pixel, band) sums live in LDS (ds_add_f32).  An
// item that holds every sample of its pixels writes their film and XYZ itself
// (one band per lane, XYZ by wave-level sums over each pixel's lanes); items
// of a render split into sample chunks write a [chunk][pixel][band] partial
// buffer that k_finalize reduces in chunk order.
#include "sky_internal.hpp"
#include "sky_trace_common.hpp"

namespace sky {

constexpr int kWavesPerBlock = 4;
constexpr int kSlotsPerWave = 256;

enum LaneState : int { ST_NEED = 0, ST_EVENT = 1, ST_TRACK = 2, ST_IDLE = 3, ST_RAY = 4 };
enum LaneEvent : int { EV_HIT = 1, EV_BOUNDARY = 2, EV_SHADOW = 3 };

__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & 63u; }

__device__ __forceinline__ uint32_t mbcnt(uint64_t m)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

// The job table is read-only during a launch.  Reading it through the
// constant address space lets the compiler keep the wave-uniform scene fields
// in scalar registers (scalar-cache loads) instead of re-fetching them with
// vector loads inside divergent code.
typedef const __attribute__((address_space(4))) DevJob *ConstJobPtr;

__device__ __forceinline__ const DevJob &job_at(const DevJob *jobs, uint32_t j)
{
    ConstJobPtr cj = (ConstJobPtr)jobs;
    return *(const DevJob *)(cj + j);
}

// Everything a lane keeps between loop trips.  The current vertex is
// (px, py, pz) with r^2 - R^2 = w carried exactly (sky_device.hpp): the
// camera, a collision or a ground hit, and after a bounce the origin of the
// new main ray (the vertex moved by the ray epsilon, as the reference offsets
// new rays, sky_trace.wgsl:744, 766).  During a shadow ray the vertex stays
// put.
struct LaneCtx {
    int st, ev, mode, vtype, kind;
    uint32_t slot, band;
    float br, bo, bas, baa, lsun;  // this band's coefficients and sun radiance
    Rng rng;
    float tp, L;
    uint32_t depth;
    float px, py, pz, w;  // current vertex / main-ray origin and its r^2 - R^2
    f3 vin;             // incoming direction at the vertex (camera direction at start)
    float nee_sum, nee_factor, rw;
    uint32_t nee_i;
    // ray being tracked
    f3 d;
    float bf, qf, wo, tmax;  // the ray's b, q and origin w (sky_device.hpp)
    float t, tau, seg_end, maj;
    float rm;  // FAST kernels: 1 / maj, from the LDS majorant table
    float aux;  // shadow-ray transmittance (0/1 delta, product for ratio tracking)
    int k, out, kn, on;
    uint32_t guard;  // shell-crossing watchdog (see ray_begin)
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
    float J;         // this path's share of the objective sum dLoss/dXYZ * XYZ
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

// u: the band's A.lam_units entry; rr, ra: the Rayleigh and aerosol density
// factors exp(-h / H) at h.
__device__ __forceinline__ DSig dsig_x(const DevScene &s, const AdjArgs &A, const float4 u, float h, float bo,
                                       const Coeffs &co, float rr, float ra)
{
    DSig d{};
    if (h < 0.0f || h > s.topf) return d;
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

__device__ __forceinline__ DSig dsig(const DevScene &s, const AdjArgs &A, uint32_t band, float h, float bo,
                                     const Coeffs &co)
{
    return dsig_x(s, A, A.lam_units[band], h, bo, co, f_exp2(-h * s.l2_hr), f_exp2(-h * s.l2_ha));
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
// u: the band's A.lam_units entry.
__device__ __forceinline__ void add_dlog_phase(const DevScene &s, const AdjArgs &A, const AdjLane &a,
                                               const float4 u, float mu, float bo, float *v)
{
    Coeffs co{a.vsr, a.vsas, 0.0f, 0.0f};
    // the aerosol absorption share of sigma_As's partials equals sigma_As's
    DSig d = dsig_x(s, A, u, a.vh, bo, co, f_exp2(-a.vh * s.l2_hr), f_exp2(-a.vh * s.l2_ha));
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
// the launch reserved room for it (nlam * L <= kMajLdsMax for every job),
// else from global memory.  A crossing needs it immediately, so the LDS copy
// removes a global-load latency from every shell crossing.
constexpr uint32_t kMajLdsMax = 2048;
// Per-wave LDS layout (floats): sums [256], sums of squares [256], adjoint
// gradient [16 doubles], the ring of path starts [64 x 4], majorant grid
// [maj_lds] (in the FAST kernels as (majorant, reciprocal) pairs), shell
// table [shell_lds].  The paths' gradient contributions cancel heavily
// across paths (scores of both signs), so the item's sum is kept in FP64.
constexpr uint32_t kRingSlots = 64;
constexpr uint32_t kLdsFixed = 2 * kSlotsPerWave + 32 + 4 * kRingSlots;

typedef __attribute__((address_space(3))) const float *LdsFloatPtr;
typedef float v4f __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) const v4f *LdsFloat4Ptr;
typedef __attribute__((address_space(3))) v4f *LdsRingPtr;
typedef float v2f __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(3))) const v2f *LdsFloat2Ptr;
typedef __attribute__((address_space(1))) const float *GlobalFloatPtr;

// The two sources carry their address spaces so that each load is a plain
// LDS or global load (a generic pointer would make every read a flat load,
// which waits on both memory counters).
struct MajSrc {
    LdsFloatPtr lds;      // wave's LDS copy (when use_lds)
    // FAST kernels: the copy as (majorant, its hardware reciprocal) pairs, so
    // that a segment's free-flight arithmetic takes no reciprocal per trip
    LdsFloat2Ptr lds2;
    GlobalFloatPtr glob;  // the job's grid in global memory
    bool use_lds;
    ShellTab shells;      // the job's shell table, LDS copy with the grid
    // FAST kernels: the job's per-band coefficients (beta_R, ozone, beta_As,
    // beta_Aa) and sun radiance, LDS copies read where they are used instead
    // of being carried in registers by every lane
    LdsFloat4Ptr lam4;
    LdsFloatPtr lsun;
    // FAST replays: the job's A.lam_units, an LDS copy (a derivative needs it
    // at every tracking step)
    LdsFloat4Ptr units;
    __device__ __forceinline__ float4 unit(uint32_t band) const
    {
        const v4f u = units[band];
        return make_float4(u.x, u.y, u.z, u.w);
    }
};

__device__ __forceinline__ float majorant(const MajSrc &m, uint32_t L, uint32_t band, int k)
{
    uint32_t i = band * L + (uint32_t)k;
    if (m.use_lds) return m.lds[i];
    return m.glob[i];
}

// MJ: 0 the general kernels; FAST kernels (every job's tables in LDS, no
// global fallback) with the majorant table as single floats (1: the
// ratio-tracked kernels, whose larger state leaves no register for 1 / maj)
// or as (majorant, reciprocal) pairs (2).
template <int MJ>
__device__ __forceinline__ void seg_begin(const DevScene &s, const MajSrc &m, LaneCtx &c)
{
    float tc;
    if constexpr (MJ == 2) {
        const v2f mm = m.lds2[c.band * s.L + (uint32_t)c.k];
        c.maj = mm.x;
        c.rm = mm.y;
        tc = next_crossing_sel(s, m.shells.lds, c.qf, c.bf, c.k, c.out, c.kn, c.on);
    } else if constexpr (MJ == 1) {
        c.maj = m.lds[c.band * s.L + (uint32_t)c.k];
        tc = next_crossing_sel(s, m.shells.lds, c.qf, c.bf, c.k, c.out, c.kn, c.on);
    } else {
        tc = next_crossing(s, m.shells, c.qf, c.bf, c.k, c.out, c.kn, c.on);
        c.maj = majorant(m, s.L, c.band, c.k);
    }
    // min(max(tc, t), tmax) as one median: every caller has t < tmax (a ray
    // starts at 0 < tmax, a crossing is taken only below tmax), and tc is a
    // finite root or +inf
    c.seg_end = __builtin_amdgcn_fmed3f(tc, c.t, c.tmax);
}

// Sets up a ray from the vertex moved by e along d (sky_device.hpp) and
// starts tracking it; returns false if it has no boundary.
template <int MJ>
__device__ __forceinline__ bool ray_begin(const DevScene &s, const MajSrc &m, LaneCtx &c, f3 d, float e, int mode)
{
    Ray r;
    ray_setup<MJ ? (int)kFastMaxShells : 0>(s, m.shells, f3{c.px, c.py, c.pz}, c.w, d, e, r);
    c.d = d;
    c.bf = r.bf;
    c.qf = r.qf;
    c.wo = r.wo;
    c.tmax = r.tmax;
    c.kind = r.kind;
    c.mode = mode;
    c.t = 0.0f;
    c.tau = -1.0f;
    c.aux = 1.0f;
    // crossings made so far (from 1), or in the FAST walk the crossings left
    c.guard = MJ ? s.max_segments : 1u;
    c.k = r.layer;
    c.out = r.outward;
    seg_begin<MJ>(s, m, c);
    c.st = ST_TRACK;
    return r.hit != 0;
}

__device__ __forceinline__ float lane_altitude(const DevScene &s, const LaneCtx &c, float t)
{
    return altitude_bw(s, c.bf, c.wo, t);
}

template <bool FAST>
__device__ __forceinline__ Coeffs lane_medium(const DevScene &s, const MajSrc &m, const LaneCtx &c, float h)
{
    if constexpr (FAST) {
        const v4f b = m.lam4[c.band];
        return medium4(s, c.band, b.x, b.y, b.z, b.w, h);
    } else {
        return medium4(s, c.band, c.br, c.bo, c.bas, c.baa, h);
    }
}

// One tracking step.  The walk consumes shell crossings (no draw for an
// empty segment, as the reference's inner `while t < segment_end` loop; with
// carried optical depth, a crossing uses up part of the current draw) until
// a tentative collision lies inside the current segment or the ray ends.
// The collision is then evaluated and accepted or rejected in the same step.
template <bool ADJ>
__device__ __forceinline__ void track_step(const DevScene &s, const MajSrc &m, LaneCtx &c, AdjLane &a,
                                           const AdjArgs &A, uint32_t &wd)
{
    float tn = 0.0f;
    const bool carried = (s.flags & kFlagCarriedTau) != 0;
    uint32_t walked = 0;
    for (;;) {
        if (c.t < c.seg_end) {
            if (c.tau < 0.0f) c.tau = free_flight(rng_next(c.rng));
            const float seg_tau = c.maj * (c.seg_end - c.t);
            if (carried && c.tau >= seg_tau) {
                c.tau -= seg_tau;
            } else {
                tn = c.t + f_div(c.tau, c.maj);
                c.tau = -1.0f;
                if (tn < c.seg_end) break;
            }
        }
        c.t = c.seg_end;
        if (c.t >= c.tmax) {
            c.st = ST_EVENT;
            c.ev = c.mode == kTrackCollision ? EV_BOUNDARY : EV_SHADOW;
            return;
        }
        if (c.guard > s.max_segments) {
            wd = 1u;
            c.st = ST_EVENT;
            c.aux = 0.0f;
            c.ev = c.mode == kTrackCollision ? EV_BOUNDARY : EV_SHADOW;
            c.kind = 2;  // watchdog: ends the path without contribution
            return;
        }
        c.guard++;
        c.k = c.kn;
        c.out = c.on;
        seg_begin<0>(s, m, c);
        // at most walk_cap crossings per step: the lane goes on from here in
        // the next step (same draws, same order), so the wave is not held by
        // its longest walk
        if (++walked == s.walk_cap) return;
    }
    c.t = tn;
    const float h = lane_altitude(s, c, tn);
    Coeffs co = lane_medium<false>(s, m, c, h);
    float st = co.sr + co.sas + co.so + co.saa;
    float ratio = fminf(fmaxf(f_div(st, c.maj), 0.0f), 1.0f);
    if (c.mode == kTrackRatio) {
        c.aux *= 1.0f - ratio;
        if constexpr (ADJ) {
            float f = -1.0f / fmaxf(c.maj - st, 1e-30f);
            add_dsigma_t(a.SS, dsig(s, A, c.band, h, c.bo, co), f);
        }
        return;
    }
    float u = rng_next(c.rng);
    if (u < ratio) {
        c.st = ST_EVENT;
        if (c.mode == kTrackCollision) {
            c.ev = EV_HIT;
            // the vertex's albedo weight and Rayleigh share (wgsl:722-731);
            // a zero throughput ends the path at the handler
            float scat = co.sr + co.sas;
            if constexpr (ADJ) {
                // accepted collision times the albedo weight: d log(sigma_s / mu)
                if (scat > 0.0f) add_dsigma_s(a.S, dsig(s, A, c.band, h, c.bo, co), 1.0f / scat);
                a.vsr = co.sr;
                a.vsas = co.sas;
                a.vh = h;
            }
            if (st > 0.0f && scat > 0.0f) {
                c.tp = c.tp * fminf(fmaxf(f_div(scat, st), 0.0f), 1.0f);
                c.rw = f_div(co.sr, scat);
            } else {
                c.tp = 0.0f;
            }
        } else {
            c.aux = 0.0f;
            c.ev = EV_SHADOW;
        }
    } else {
        if constexpr (ADJ) {
            // null collision: d log(1 - sigma_t / mu) = -d sigma_t / (mu - sigma_t)
            float f = -1.0f / fmaxf(c.maj - st, 1e-30f);
            add_dsigma_t(c.mode == kTrackCollision ? a.S : a.SS, dsig(s, A, c.band, h, c.bo, co), f);
        }
    }
}

// The tracking step of the FAST kernels (tables in LDS; CARRIED: carried
// optical depth, else a fresh draw per shell as the reference; RATIO: ratio-
// tracked shadow rays, else delta-tracked): the same walk, draws and
// arithmetic as track_step, written for SIMT.  The step's free flight is drawn before the walk (the first draw of a
// step is made in its first non-empty segment either way, so the random
// sequence is unchanged); each trip of the walk either finds the collision
// inside the current segment or crosses into the next one, with selects in
// place of branches and one exit per outcome; the collision test is in select
// form too.
template <bool ADJ, bool CARRIED, bool RATIO>
__device__ __forceinline__ void track_step_fast(const DevScene &s, const MajSrc &m, LaneCtx &c, AdjLane &a,
                                                const AdjArgs &A, uint32_t &wd)
{
    float tn = 0.0f;
    if (c.tau < 0.0f && c.t < c.seg_end) c.tau = free_flight(rng_next(c.rng));
    // the walk stops at a boundary: the ray's end, the watchdog (no crossings
    // left) or, in the ratio-tracked kernels, this step's share of crossings
    // (walk_cap: the lane resumes at the boundary in the next step, with the
    // same draws in the same order; compiled only there, where it pays)
    bool stop = false;
    uint32_t floor = 0u;
    if constexpr (RATIO) floor = s.walk_cap && c.guard > s.walk_cap ? c.guard - s.walk_cap : 0u;
    // 1 / majorant of the current segment (the table's f_rcp(maj)); the loop
    // leaves with the segment of the collision or of the ray's end, whose
    // ratio below reuses it
    float rm;
    for (;;) {
        const bool nonempty = c.t < c.seg_end;
        rm = RATIO ? f_rcp(c.maj) : c.rm;
        tn = c.t + c.tau * rm;
        if constexpr (CARRIED) {
            const float seg_tau = c.maj * (c.seg_end - c.t);
            const bool cross = !nonempty || c.tau >= seg_tau;
            if (!cross && tn < c.seg_end) break;
            // a draw that lands at or past the segment end by rounding is spent
            c.tau = cross ? (nonempty ? c.tau - seg_tau : c.tau) : -1.0f;
        } else {
            // every non-empty segment spends its own draw (sky_trace.wgsl:425-434)
            if (nonempty && tn < c.seg_end) break;
            c.tau = nonempty ? -1.0f : c.tau;
        }
        c.t = c.seg_end;
        if ((c.t >= c.tmax) | (c.guard <= floor)) {
            stop = true;
            break;
        }
        c.guard--;
        c.k = c.kn;
        c.out = c.on;
        seg_begin<RATIO ? 1 : 2>(s, m, c);
        if (c.tau < 0.0f && c.t < c.seg_end) c.tau = free_flight(rng_next(c.rng));
    }
    // The end of the ray, or the tentative collision and its accept test
    // (with the albedo weight of an accepted main-ray collision), in
    // select form: the collision's draw is taken only by lanes that have one.
    // A ratio-tracked shadow ray instead multiplies its transmittance by
    // 1 - sigma_t / majorant and goes on (sky_trace.wgsl:441-475).
    const bool paused = RATIO && stop && c.t < c.tmax && c.guard != 0u;
    const bool ended = stop && !paused;
    const bool col = !stop;
    const bool main = c.mode == kTrackCollision;
    const bool rt = RATIO && c.mode == kTrackRatio;
    // (an ended lane evaluates the medium at its end point, unused)
    const float h = lane_altitude(s, c, col ? tn : c.t);
    Coeffs co;
    float rr = 0.0f, ra = 0.0f;
    if constexpr (ADJ) {
        // the replays run the analytic medium only (the adjoint refuses
        // tables): medium4's arithmetic, keeping its two density factors for
        // the derivatives below
        const v4f b = m.lam4[c.band];
        rr = f_exp2(-h * s.l2_hr);
        ra = f_exp2(-h * s.l2_ha);
        float ro = 1.0f - fabsf(h - s.hc) * s.inv_hw;
        ro = ro < 0.0f ? 0.0f : ro;
        co = (h < 0.0f || h > s.topf) ? Coeffs{0.0f, 0.0f, 0.0f, 0.0f}
                                      : Coeffs{b.x * rr, b.z * ra, b.y * ro, b.w * ra};
    } else {
        co = lane_medium<true>(s, m, c, h);
    }
    const float st = co.sr + co.sas + co.so + co.saa;
    const float ratio = fminf(fmaxf(st * rm, 0.0f), 1.0f);
    const float u = rng_peek(c.rng, 0);
    c.rng.w += col && !rt ? kLcgMul : 0u;
    const bool acc = col && !rt && u < ratio;
    const bool dog = ended && c.t < c.tmax;  // watchdog: ends the path without contribution
    wd |= dog ? 1u : 0u;
    c.t = col ? tn : c.t;
    c.tau = col ? -1.0f : c.tau;
    c.st = (ended || acc) ? ST_EVENT : c.st;
    c.ev = ended ? (main ? EV_BOUNDARY : EV_SHADOW) : (acc ? (main ? EV_HIT : EV_SHADOW) : c.ev);
    // an accepted main-ray collision keeps sigma_R, sigma_As and sigma_t in
    // rw, nee_factor and aux (free during a main ray): the handler forms the
    // vertex's albedo weight and Rayleigh share from them
    const bool vert = acc && main;
    c.rw = vert ? co.sr : c.rw;
    c.nee_factor = vert ? co.sas : c.nee_factor;
    c.aux = vert ? st : (dog || (acc && !main) ? 0.0f : (rt && col ? c.aux * (1.0f - ratio) : c.aux));
    c.kind = dog ? 2 : c.kind;
    if constexpr (ADJ) {
        // accepted main-ray collision: d log(sigma_s / mu) into S; null
        // collision: d log(1 - sigma_t / mu) into S (main ray) or SS (shadow)
        const DSig d = dsig_x(s, A, m.unit(c.band), h, m.lam4[c.band].y, co, rr, ra);
        const bool rej = col && !acc;
        const float scat = co.sr + co.sas;
        const float fs = vert && scat > 0.0f ? 1.0f / scat : 0.0f;
        const float fr = rej ? -1.0f / fmaxf(c.maj - st, 1e-30f) : 0.0f;
        const float fm = main ? fr : 0.0f, fh = main ? 0.0f : fr;
        a.S[0] += (fs + fm) * d.r0;
        a.S[1] += (fs + fm) * d.r1;
        a.S[2] += fm * d.o2;
        a.S[3] += fm * d.o3;
        a.S[4] += (fs + fm) * d.as4;
        a.S[5] += fm * d.aa5;
        a.S[6] += fs * d.as6 + fm * (d.as6 + d.aa6);
        a.S[8] += fs * d.as8 + fm * (d.as8 + d.aa8);
        add_dsigma_t(a.SS, d, fh);
        a.vsr = vert ? co.sr : a.vsr;
        a.vsas = vert ? co.sas : a.vsas;
        a.vh = vert ? h : a.vh;
    }
}

// Vertex at distance t along the main ray from (px, py, pz): r^2 - R^2 is
// the ray's w at t, or exactly 0 on the ground.
__device__ __forceinline__ void advance_vertex(LaneCtx &c, float t, bool ground)
{
    c.px = __builtin_fmaf(c.d.x, t, c.px);
    c.py = __builtin_fmaf(c.d.y, t, c.py);
    c.pz = __builtin_fmaf(c.d.z, t, c.pz);
    c.w = ground ? 0.0f : ray_w(c.bf, c.wo, t);
}

enum Next : int { NX_NEE = 1, NX_CONT = 2, NX_MAIN = 3, NX_FINISH = 4 };



// Path end: the path's radiance and its square into the wave's LDS sums.
template <bool ADJ>
__device__ __forceinline__ void finish_path(LaneCtx &c, AdjLane &a, float *acc_sum, float *acc_sq,
                                            double *acc_grad)
{
    float Lf = fmaxf(c.L, 0.0f);
    atomicAdd(&acc_sum[c.slot], Lf);
    atomicAdd(&acc_sq[c.slot], Lf * Lf);
    if constexpr (ADJ) {
        for (int k = 0; k < kNP; ++k)
            if (a.G[k] != 0.0f && a.G[k] == a.G[k]) atomicAdd(&acc_grad[k], (double)a.G[k]);
        if (a.J != 0.0f && a.J == a.J) atomicAdd(&acc_grad[kNP], (double)a.J);
    }
    c.st = ST_NEED;
}

// Event handler: one step of trace_path (sky_trace.wgsl:703-776) between two
// tracking passes.  NEE = one sun sample of direct_sun_at_scatter /
// ground_radiance (wgsl:566-608); CONT = the NEE average, the bounce
// direction, its weight and the roulette (wgsl:691-701, 751-770); MAIN = the
// next main ray.  This is the first half of a handler pass, everything up to
// the ray set-up: it leaves the lane in RAY (c.d, c.mode and, for a shadow
// ray, c.nee_factor set), EVENT (a zero sun sample, for the next pass) or
// NEED (path finished and accumulated).
template <bool ADJ>
__device__ __forceinline__ void handle_event(const DevScene &s, const MajSrc &m, LaneCtx &c, AdjLane &a,
                                             const AdjArgs &A, float *acc_sum, float *acc_sq, double *acc_grad,
                                             uint32_t &wd)
{
    const uint32_t n_light = s.dls;
    const bool phase_is = (s.flags & kFlagPhaseIS) != 0;
    int nx = NX_FINISH;

    // 1. the event; a ground hit or an accepted collision moves the vertex
    const bool at_ground = c.ev == EV_BOUNDARY && c.kind == 0;
    if (at_ground || c.ev == EV_HIT) advance_vertex(c, at_ground ? c.tmax : c.t, at_ground);
    if (c.ev == EV_BOUNDARY) {
        if (c.kind == 1) {
            // exit: the sun disk is seen only by the camera ray (wgsl:745-749)
            if (c.depth == 0 && in_sun_cone(s, c.d)) {
                float C = c.tp * c.lsun;
                c.L = c.L + C;
                if constexpr (ADJ) {
                    for (int k = 0; k < kNP; ++k) a.G[k] += a.w * C * a.S[k];
                    a.J += a.w * C;
                }
            }
        } else if (c.kind == 0) {
            c.vtype = 1;
            c.nee_i = 0;
            c.nee_sum = 0.0f;
            nx = NX_NEE;
        } else if (c.kind == 3) {
            wd = 1u;
        }
        // kind 2: tracking watchdog or a main ray without boundary, the path ends
    } else if (c.ev == EV_HIT) {
        // tp and rw were updated when the collision was accepted
        if (c.tp > 0.0f) {
            c.vtype = 0;
            c.vin = c.d;
            c.nee_i = 0;
            c.nee_sum = 0.0f;
            nx = NX_NEE;
        }
    } else {  // EV_SHADOW
        c.nee_sum += c.nee_factor * c.aux;
        if constexpr (ADJ) {
            float C = c.tp * c.nee_factor * c.aux / (float)n_light;
            if (C != 0.0f) {
                for (int k = 0; k < kNP; ++k) a.G[k] += a.w * C * a.SS[k];
                a.J += a.w * C;
            }
        }
        c.nee_i++;
        nx = c.nee_i < n_light ? NX_NEE : NX_CONT;
    }

    // 2. this pass's random numbers, in the reference's order: sun cone (2);
    //    bounce direction (3 with phase sampling, else 2), then the roulette
    //    draw from rr_start_depth on.
    const bool scatter_is = c.vtype == 0 && phase_is;
    uint32_t ndraw = 0;
    if (nx == NX_NEE) ndraw = 2;
    if (nx == NX_CONT) ndraw = (scatter_is ? 3u : 2u) + (c.depth >= s.rr_start ? 1u : 0u);
    float u0 = 0.0f, u1 = 0.0f, u2 = 0.0f, u3 = 0.0f;
    if (ndraw > 0) u0 = rng_peek(c.rng, 0);
    if (ndraw > 1) u1 = rng_peek(c.rng, 1);
    if (ndraw > 2) u2 = rng_peek(c.rng, 2);
    if (ndraw > 3) u3 = rng_peek(c.rng, 3);
    c.rng.w += ndraw * kLcgMul;

    // 3. one direction sample: sun cone (NEE), phase lobe or isotropic
    //    (sample_isotropic, wgsl:534-539, as a frame around +y with the
    //    azimuth negated), cosine lobe (sample_cosine_hemisphere, wgsl:548-560)
    f3 dir = c.vin;
    f3 nrm{0.0f, 1.0f, 0.0f};
    if (c.vtype == 1) nrm = normalize_or_y(f3{c.px, c.py, c.pz});
    if (nx == NX_NEE || nx == NX_CONT) {
        f3 axis;
        float mu, st, phi;
        if (nx == NX_NEE) {
            sun_cone_mu(s, u0, mu, st);
            axis = f3{s.sun[0], s.sun[1], s.sun[2]};
            phi = u1;
        } else if (c.vtype == 0) {
            if (phase_is) {
                mu = u0 < c.rw ? sample_rayleigh_mu(u1) : sample_hg_mu(s.g, s.inv_2g, u1);
                axis = c.vin;
                phi = u2;
            } else {
                mu = 1.0f - 2.0f * u0;
                axis = f3{0.0f, 1.0f, 0.0f};
                phi = -u1;
            }
            st = f_sqrt(fmaxf(1.0f - mu * mu, 0.0f));
        } else {
            mu = f_sqrt(fmaxf(1.0f - u0, 0.0f));
            st = f_sqrt(u0);
            axis = nrm;
            phi = u1;
        }
        dir = frame_dir(axis, mu, st, phi);
    }

    // 4. weights
    bool need_ray = false;
    int mode = kTrackCollision;
    float factor = 0.0f;
    if (nx == NX_NEE) {
        if (c.vtype == 1) {
            float cs = fmaxf(dot3(nrm, dir), 0.0f);
            factor = s.albedo * kInvPi * c.lsun * cs / s.cone_pdf;
            need_ray = cs > 0.0f;
            if constexpr (ADJ) {
                for (int k = 0; k < kNP; ++k) a.SS[k] = a.S[k];
                if (s.albedo > 0.0f) a.SS[9] += 1.0f / s.albedo;
            }
        } else {
            float mu = dot3(c.vin, dir);
            float ph = c.rw * phase_rayleigh(mu) + (1.0f - c.rw) * phase_hg(s.g, mu);
            factor = c.lsun * ph / s.cone_pdf;
            need_ray = true;
            if constexpr (ADJ) {
                for (int k = 0; k < kNP; ++k) a.SS[k] = a.S[k];
                add_dlog_phase(s, A, a, A.lam_units[c.band], mu, c.bo, a.SS);
            }
        }
        mode = (s.flags & kFlagRatio) ? kTrackRatio : kTrackDelta;
        if (!need_ray) {
            // ground facing away from the sun: a zero sample
            c.st = ST_EVENT;
            c.ev = EV_SHADOW;
            c.nee_factor = 0.0f;
            c.aux = 0.0f;
        }
    } else if (nx == NX_CONT) {
        c.L = c.L + c.tp * (c.nee_sum / (float)n_light);
        float w;
        bool ok = true;
        if (c.vtype == 0) {
            float mu = dot3(c.vin, dir);
            w = phase_is ? 1.0f
                         : (c.rw * phase_rayleigh(mu) + (1.0f - c.rw) * phase_hg(s.g, mu)) * (1.0f / kIsotropicPdf);
            if constexpr (ADJ) add_dlog_phase(s, A, a, A.lam_units[c.band], mu, c.bo, a.S);
        } else {
            // albedo/pi cos / (cos/pi)
            float co_out = fmaxf(dot3(nrm, dir), 0.0f);
            ok = co_out > 0.0f;
            w = s.albedo;
            if constexpr (ADJ) {
                if (s.albedo > 0.0f) a.S[9] += 1.0f / s.albedo;
            }
        }
        nx = NX_FINISH;
        if (ok) {
            c.tp = c.tp * w;
            bool live = true;
            if (c.depth >= s.rr_start) {
                float q = fminf(fmaxf(c.tp, 0.05f), 0.95f);
                float urr = scatter_is ? u3 : u2;
                if (urr > q)
                    live = false;
                else
                    c.tp = c.tp / q;
            }
            if (live) {
                if (c.L < 0.0f || c.L != c.L || c.tp < 0.0f || c.tp != c.tp) {
                    c.L = 0.0f;
                    if constexpr (ADJ) {
                        for (int k = 0; k < kNP; ++k) a.G[k] = 0.0f;
                        a.J = 0.0f;
                    }
                } else {
                    // the ray stage starts the new main ray the ray epsilon
                    // from the vertex (depth > 0)
                    c.depth++;
                    nx = NX_MAIN;
                }
            }
        }
    }
    if (nx == NX_MAIN) {
        // loop top of trace_path (wgsl:709-717) plus the depth cap
        nx = NX_FINISH;
        if (!(s.max_depth && c.depth >= s.max_depth)) {
            if (c.depth >= s.watchdog_limit) {
                wd = 1u;
            } else {
                need_ray = true;
                mode = kTrackCollision;
                nx = NX_MAIN;
            }
        }
    }

    if (need_ray) {
        c.st = ST_RAY;
        c.d = dir;
        c.mode = mode;
        if (nx == NX_NEE) c.nee_factor = factor;
    }
    if (nx == NX_FINISH) finish_path<ADJ>(c, a, acc_sum, acc_sq, acc_grad);
}

// handle_event for the FAST kernels (IS: Rayleigh/HG lobe sampling, else
// isotropic directions weighted by the phase function as the reference;
// RATIO: ratio-tracked shadow rays, else delta-tracked): the same steps with every branch that has no side effect
// turned into selects, so a pass costs one evaluation of each piece instead of
// the exec-mask bookkeeping of nested branches.  Same draws and arithmetic,
// except that the four weight divisions multiply by the hardware reciprocal
// (within 1 ulp of the quotient).
template <bool ADJ, bool IS, bool RATIO>
__device__ __forceinline__ void handle_event_fast(const DevScene &s, const MajSrc &m, LaneCtx &c, AdjLane &a,
                                                  const AdjArgs &A, float *acc_sum, float *acc_sq, double *acc_grad,
                                                  uint32_t &wd)
{
    const uint32_t n_light = s.dls;
    const float lsun = m.lsun[c.band];

    // 1. the event
    const bool bnd = c.ev == EV_BOUNDARY, hit = c.ev == EV_HIT, shd = !bnd && !hit;
    const bool at_ground = bnd && c.kind == 0;
    {
        // an accepted collision: the albedo weight and the vertex's Rayleigh
        // share (wgsl:722-731) from sigma_R, sigma_As and sigma_t kept by the
        // tracking step; a zero throughput ends the path below
        const float sr = c.rw, st = c.aux, scat = sr + c.nee_factor;
        const bool pos = st > 0.0f && scat > 0.0f;
        c.tp = hit ? (pos ? c.tp * fminf(fmaxf(f_div(scat, st), 0.0f), 1.0f) : 0.0f) : c.tp;
        c.rw = hit && pos ? f_div(sr, scat) : c.rw;
    }
    {
        // advance_vertex for a ground hit or an accepted collision (a lane
        // that does not move adds d * 0, which leaves its position exact)
        const float tm = at_ground ? c.tmax : (hit ? c.t : 0.0f);
        c.px = __builtin_fmaf(c.d.x, tm, c.px);
        c.py = __builtin_fmaf(c.d.y, tm, c.py);
        c.pz = __builtin_fmaf(c.d.z, tm, c.pz);
        c.w = at_ground ? 0.0f : (hit ? ray_w(c.bf, c.wo, c.t) : c.w);
    }
    // exit: the sun disk is seen only by the camera ray (wgsl:745-749); the
    // ray directions are unit vectors already (camera_dir, frame_dir)
    const bool sun_seen = bnd && c.kind == 1 && c.depth == 0 && in_sun_cone_unit(s, c.d);
    if constexpr (ADJ) {
        // the sun disk's contribution with the path's score, a sun sample's
        // with the sample's score
        const float Csun = c.tp * lsun;
        const float Csh = c.tp * c.nee_factor * c.aux / (float)n_light;
        const bool gh = shd && Csh != 0.0f;
        for (int k = 0; k < kNP; ++k)
            a.G[k] += (sun_seen ? a.w * Csun * a.S[k] : 0.0f) + (gh ? a.w * Csh * a.SS[k] : 0.0f);
        a.J += (sun_seen ? a.w * Csun : 0.0f) + (gh ? a.w * Csh : 0.0f);
    }
    c.L = sun_seen ? c.L + c.tp * lsun : c.L;
    wd |= bnd && c.kind == 3 ? 1u : 0u;
    const bool scatter_vertex = hit && c.tp > 0.0f;
    const bool nee_start = at_ground || scatter_vertex;
    c.vtype = at_ground ? 1 : (scatter_vertex ? 0 : c.vtype);
    c.vin.x = scatter_vertex ? c.d.x : c.vin.x;
    c.vin.y = scatter_vertex ? c.d.y : c.vin.y;
    c.vin.z = scatter_vertex ? c.d.z : c.vin.z;
    // a shadow result adds its sun sample
    const float nee_sum = shd ? c.nee_sum + c.nee_factor * c.aux : c.nee_sum;
    const uint32_t nee_i = shd ? c.nee_i + 1u : c.nee_i;
    c.nee_sum = nee_start ? 0.0f : nee_sum;
    c.nee_i = nee_start ? 0u : nee_i;
    int nx = nee_start ? NX_NEE : (shd ? (nee_i < n_light ? NX_NEE : NX_CONT) : NX_FINISH);
    const bool is_nee = nx == NX_NEE, is_cont = nx == NX_CONT;
    const bool scat_v = c.vtype == 0;

    // 2. this pass's random numbers in the reference's order: sun cone (2);
    //    lobe choice, mu, azimuth (3), isotropic direction (2) or cosine lobe
    //    (2), then the roulette draw
    const bool rr = c.depth >= s.rr_start;
    const bool lobe = IS && scat_v;
    const uint32_t ndraw = is_nee ? 2u : (is_cont ? (lobe ? 3u : 2u) + (rr ? 1u : 0u) : 0u);
    const float u0 = rng_peek(c.rng, 0);
    const float u1 = rng_peek(c.rng, 1);
    float u2 = 0.0f, u3 = 0.0f;
    if (ndraw > 2) u2 = rng_peek(c.rng, 2);
    if (ndraw > 3) u3 = rng_peek(c.rng, 3);
    c.rng.w += ndraw * kLcgMul;

    // 3. one direction: sun cone (NEE), Rayleigh/HG lobe (IS) or isotropic
    //    (sample_isotropic, wgsl:534-539, as a frame around +y with the
    //    azimuth negated) at a scattering vertex, cosine lobe at the ground
    const f3 nrm = normalize_or_y(f3{c.px, c.py, c.pz});
    // sun cone (sun_cone_mu): mu = 1 - x, sin = sqrt(x (2 - x))
    const float x_sun = fminf(u0, 0.99999994f) * s.omc;
    float mu_v0, phi_v;
    f3 axis_v;
    if constexpr (IS) {
        mu_v0 = u0 < c.rw ? sample_rayleigh_mu(u1) : sample_hg_mu(s.g, s.inv_2g, u1);
        axis_v = c.vin;
        phi_v = u2;
    } else {
        mu_v0 = 1.0f - 2.0f * u0;
        axis_v = f3{0.0f, 1.0f, 0.0f};
        phi_v = -u1;
    }
    const float mu_cos = f_sqrt(fmaxf(1.0f - u0, 0.0f));
    const float mu = is_nee ? 1.0f - x_sun : (scat_v ? mu_v0 : mu_cos);
    // one root for the three polar sines: x (2 - x), 1 - mu^2 (clamped), u0
    const float st = f_sqrt(is_nee ? x_sun * (2.0f - x_sun) : (scat_v ? fmaxf(1.0f - mu_v0 * mu_v0, 0.0f) : u0));
    const float phi = (!is_nee && scat_v) ? phi_v : u1;
    const f3 sun{s.sun[0], s.sun[1], s.sun[2]};
    const f3 axis = is_nee ? sun : (scat_v ? axis_v : (c.vtype == 1 ? nrm : f3{0.0f, 1.0f, 0.0f}));
    const f3 dir = (is_nee || is_cont) ? frame_dir(axis, mu, st, phi) : c.vin;

    // 4. weights.  NEE: the sun sample's factor; a ground facing away from
    //    the sun is a zero sample without a ray.
    const f3 gn = c.vtype == 1 ? nrm : f3{0.0f, 1.0f, 0.0f};
    const float cs = fmaxf(dot3(gn, dir), 0.0f);
    const float mu_v = dot3(c.vin, dir);
    const float ph = c.rw * phase_rayleigh(mu_v) + (1.0f - c.rw) * phase_hg(s.g, mu_v);
    const float factor = (c.vtype == 1 ? s.albedo * kInvPi * lsun * cs : lsun * ph) * s.cone_omega;
    const bool nee_ray = c.vtype == 1 ? cs > 0.0f : true;
    // CONT: the NEE average, the bounce weight (1 for lobe sampling, albedo
    // for the cosine lobe) and the roulette (wgsl:691-701, 751-770)
    const float Lc = c.L + c.tp * (c.nee_sum * s.inv_dls);
    const bool ok = scat_v || cs > 0.0f;
    // bounce weight: 1 for a lobe sample, phase / isotropic pdf for an
    // isotropic one, the albedo for the cosine lobe
    const float tp1 = c.tp * (scat_v ? (IS ? 1.0f : ph * (1.0f / kIsotropicPdf)) : s.albedo);
    const float q = fminf(fmaxf(tp1, 0.05f), 0.95f);
    const bool killed = rr && (lobe ? u3 : u2) > q;
    const float tp2 = rr ? f_div(tp1, q) : tp1;
    const bool live = ok && !killed;
    const bool bad = Lc < 0.0f || Lc != Lc || tp2 < 0.0f || tp2 != tp2;
    const bool bounce = is_cont && live && !bad;
    if constexpr (ADJ) {
        // NEE: the sun sample's score starts from the path's plus its phase
        // (scattering vertex) or albedo (ground) term; CONT: the bounce adds
        // the same terms to the path's score
        float D[kNP];
        for (int k = 0; k < kNP; ++k) D[k] = 0.0f;
        add_dlog_phase(s, A, a, m.unit(c.band), mu_v, m.lam4[c.band].y, D);
        const float alb = s.albedo > 0.0f ? 1.0f / s.albedo : 0.0f;
        for (int k = 0; k < kNP; ++k) {
            const float t = scat_v ? D[k] : (k == 9 ? alb : 0.0f);
            a.SS[k] = is_nee ? a.S[k] + t : a.SS[k];
            a.S[k] = is_cont ? a.S[k] + t : a.S[k];
            a.G[k] = is_cont && live && bad ? 0.0f : a.G[k];
        }
        a.J = is_cont && live && bad ? 0.0f : a.J;
    }
    if (is_cont) {
        c.L = live && bad ? 0.0f : Lc;
        c.tp = ok ? (killed ? tp1 : tp2) : c.tp;
    }
    // (the ray stage starts the bounce's main ray the ray epsilon from the
    // vertex: depth > 0)
    c.depth = bounce ? c.depth + 1u : c.depth;
    nx = is_cont ? (bounce ? NX_MAIN : NX_FINISH) : nx;
    // loop top of trace_path (wgsl:709-717) plus the depth cap
    bool main_ray = false;
    if (nx == NX_MAIN) {
        const bool capped = s.max_depth && c.depth >= s.max_depth;
        const bool dog = !capped && c.depth >= s.watchdog_limit;
        wd |= dog ? 1u : 0u;
        main_ray = !capped && !dog;
        nx = main_ray ? NX_MAIN : NX_FINISH;
    }
    const bool need_ray = (is_nee && nee_ray) || main_ray;
    if (is_nee && !nee_ray) {
        c.st = ST_EVENT;
        c.ev = EV_SHADOW;
        c.nee_factor = 0.0f;
        c.aux = 0.0f;
    }
    if (need_ray) {
        c.st = ST_RAY;
        c.d = dir;
        c.mode = is_nee ? (RATIO ? kTrackRatio : kTrackDelta) : kTrackCollision;
        c.nee_factor = is_nee ? factor : c.nee_factor;
    }
    if (nx == NX_FINISH) finish_path<ADJ>(c, a, acc_sum, acc_sq, acc_grad);
}

// Second half of a handler pass: one ray set-up for every lane in RAY -- a
// shadow ray, or a main ray after a bounce, from the vertex moved by the ray
// epsilon (wgsl:744, 766), or the camera ray from the camera (depth 0).  A
// main ray's origin becomes the path's vertex.  A shadow ray that does not
// leave the atmosphere is a zero sun sample; a main ray with no boundary
// ends the path at the next pass.
template <int MJ>
__device__ __forceinline__ void ray_pass(const DevScene &s, const MajSrc &m, LaneCtx &c)
{
    const bool shadow = c.mode != kTrackCollision;
    const f3 dir = c.d;
    const float e = shadow || c.depth > 0u ? kRayEpsilonKm : 0.0f;
    bool hit = ray_begin<MJ>(s, m, c, dir, e, c.mode);
    {
        const float em = shadow ? 0.0f : e;
        c.px = __builtin_fmaf(dir.x, em, c.px);
        c.py = __builtin_fmaf(dir.y, em, c.py);
        c.pz = __builtin_fmaf(dir.z, em, c.pz);
        c.w = shadow ? c.w : c.wo;
    }
    if (shadow) {
        if (!(hit && c.kind == 1)) {
            c.st = ST_EVENT;
            c.ev = EV_SHADOW;
            c.aux = 0.0f;
        }
    } else if (!hit) {
        c.st = ST_EVENT;
        c.ev = EV_BOUNDARY;
        c.kind = 2;
    }
}

// Work-item constants shared by the lanes of a wave.
struct ItemInfo {
    uint32_t p0, x0, y0, nslots, s0;
    float r_nslots, r_nlam, r_width;
};

// Path start (main, sky_trace.wgsl:799-807) in two parts.  A start is triple
// j = si * nslots + rem of the item, ordered (sample, pixel, band) with the
// band fastest.  make_start forms its camera ray -- the two jitter draws and
// the direction -- ahead of need: the whole wave makes 64 consecutive starts
// at once into the wave's ring, {dir.x, dir.y, dir.z, rem | si << 8}
// (rem < nslots <= 256, si < 2^24).  start_path gives a lane whose path
// ended the next start of the ring.  Lanes take starts in triple order either
// way, so the ring changes no path and no summation order.
__device__ __forceinline__ v4f make_start(const DevScene &s, uint32_t si, uint32_t rem, const ItemInfo &it)
{
    const uint32_t pl = udiv_small(rem, it.r_nlam);
    const uint32_t band = rem - pl * s.nlam;
    // pixel p0 + pl, from the item's first pixel (x0, y0); pl < 256
    uint32_t x = it.x0 + pl, y = it.y0;
    if (s.width >= 256u) {
        if (x >= s.width) {
            x -= s.width;
            y++;
        }
    } else {
        uint32_t dy = udiv_small(x, it.r_width);
        x -= dy * s.width;
        y += dy;
    }
    const Rng r = rng_make(it.p0 + pl, it.s0 + si, band, s);
    const float j1 = rng_peek(r, 0);
    const float j2 = rng_peek(r, 1);
    const f3 d = camera_dir(s, x, y, j1, j2);
    return v4f{d.x, d.y, d.z, __uint_as_float(rem | (si << 8))};
}

template <bool ADJ, bool FAST>
__device__ __forceinline__ void start_path(const DevScene &s, LaneCtx &c, AdjLane &a, const AdjArgs &A,
                                           const v4f e, const ItemInfo &it)
{
    const uint32_t key = __float_as_uint(e.w);
    const uint32_t rem = key & 255u, si = key >> 8;
    const uint32_t pl = udiv_small(rem, it.r_nlam);
    c.band = rem - pl * s.nlam;
    c.slot = rem;
    const uint32_t pixel = it.p0 + pl;
    if constexpr (!FAST) {
        const DevLambda l = s.lam[c.band];
        c.br = l.br;
        c.bo = l.bo;
        c.bas = l.bas;
        c.baa = l.baa;
        c.lsun = l.lsun;
    }
    // the key words of rng_make, past the camera's two jitter draws
    c.rng = rng_make(pixel, it.s0 + si, c.band, s);
    c.rng.w += 2u * kLcgMul;
    c.vin = f3{e.x, e.y, e.z};
    c.px = 0.0f;
    c.py = s.obs_yf;
    c.pz = 0.0f;
    c.w = s.obs_wf;
    c.tp = 1.0f;
    c.L = 0.0f;
    c.depth = 0;
    c.vtype = 0;
    // main (wgsl:799-807): the camera ray, set up by the handler pass's ray
    // stage.  At the loop top of trace_path only the depth watchdog can end
    // a path at depth 0 (a max_depth of 0 means no cap).
    c.d = c.vin;
    c.mode = kTrackCollision;
    if (s.watchdog_limit == 0) {
        c.st = ST_EVENT;
        c.ev = EV_BOUNDARY;
        c.kind = 3;  // depth watchdog before the first ray
    } else {
        c.st = ST_RAY;
    }
    if constexpr (ADJ) {
        const float *g = A.dl_dxyz + 3 * (size_t)pixel;
        const DevLambda l = s.lam[c.band];
        a.w = (g[0] * l.wx + g[1] * l.wy + g[2] * l.wz) * A.inv_spp;
        for (int k = 0; k < kNP; ++k) {
            a.S[k] = 0.0f;
            a.G[k] = 0.0f;
        }
        a.J = 0.0f;
    }
}

// Sum of x over each aligned group of nl lanes (nl a power of two <= 64):
// log2(nl) butterfly exchanges, after which every lane of a group holds the
// group's sum.  All 64 lanes must take part.
__device__ __forceinline__ float group_sum(float x, uint32_t nl)
{
    for (uint32_t o = nl >> 1; o > 0; o >>= 1) x += __shfl_xor(x, (int)o);
    return x;
}

// End of a work item that holds every sample of its np pixels: the film
// (1/spp, clamped at 0, integrator.rs:307-310), its second moment and
// XYZ = sum_band film * cmf * weight, straight from the wave's LDS sums.  The
// film rows of the item's pixels are contiguous ([pixel][band]), so the
// stores coalesce.  XYZ: with a power-of-two band count of at least 4, one
// band per lane and a wave-level sum over each pixel's lanes (64 slots at a
// time); lane k then takes pixel k's sums and writes them.  Other band counts
// sum a pixel's bands in band order on one lane.
__device__ __forceinline__ void item_outputs(const DevScene &s, uint32_t p0, uint32_t np, uint32_t nslots,
                                             const float *sum, const float *sq, uint32_t lane)
{
    const float inv = 1.0f / (float)s.spp;
    const size_t base = (size_t)p0 * s.nlam;
    if (s.film)
        for (uint32_t i = lane; i < nslots; i += 64) s.film[base + i] = fmaxf(sum[i] * inv, 0.0f);
    if (s.film_m2)
        for (uint32_t i = lane; i < nslots; i += 64) s.film_m2[base + i] = sq[i] * inv;
    if (!s.xyz) return;
    const uint32_t nl = s.nlam;
    if (nl >= 4 && (nl & (nl - 1)) == 0) {
        // nslots <= 256, so np <= 64: one lane per pixel at the end
        const uint32_t ppc = 64u / nl;  // pixels per 64 slots
        float X = 0.0f, Y = 0.0f, Z = 0.0f;
        for (uint32_t c = 0; c * 64u < nslots; ++c) {
            const uint32_t i = c * 64u + lane;
            float x = 0.0f, y = 0.0f, z = 0.0f;
            if (i < nslots) {
                const float f = fmaxf(sum[i] * inv, 0.0f);
                const DevLambda l = s.lam[lane & (nl - 1)];
                x = f * l.wx;
                y = f * l.wy;
                z = f * l.wz;
            }
            x = group_sum(x, nl);
            y = group_sum(y, nl);
            z = group_sum(z, nl);
            const int src = (int)((lane % ppc) * nl);  // first lane of this lane's pixel in chunk c
            const float gx = __shfl(x, src), gy = __shfl(y, src), gz = __shfl(z, src);
            if (lane / ppc == c) {
                X = gx;
                Y = gy;
                Z = gz;
            }
        }
        if (lane < np) {
            float *o = s.xyz + 3 * (size_t)(p0 + lane);
            o[0] = X;
            o[1] = Y;
            o[2] = Z;
        }
    } else {
        // (one 48-byte store per item instead of three strided ones measured
        // more HBM traffic: the item's bytes share cache lines with the
        // neighbouring items, which other XCDs write)
        for (uint32_t pl = lane; pl < np; pl += 64) {
            float X = 0.0f, Y = 0.0f, Z = 0.0f;
            for (uint32_t b = 0; b < s.nlam; ++b) {
                const float f = fmaxf(sum[pl * s.nlam + b] * inv, 0.0f);
                const DevLambda l = s.lam[b];
                X += f * l.wx;
                Y += f * l.wy;
                Z += f * l.wz;
            }
            float *o = s.xyz + 3 * (size_t)(p0 + pl);
            o[0] = X;
            o[1] = Y;
            o[2] = Z;
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
        if (job_at(a.jobs, mid).item_base <= item)
            lo = mid;
        else
            hi = mid;
    }
    return __builtin_amdgcn_readfirstlane(lo);
}

// FK: 0 the general kernels; 1 FAST with carried optical depth and lobe
// sampling; 2 FAST with the reference's estimator (a fresh draw per shell,
// isotropic directions); 3 as 1 with ratio-tracked shadow rays.
template <bool ADJ, int FK>
__device__ __forceinline__ void wavepool_body(const PoolArgs &a, const AdjArgs &A)
{
    constexpr bool FAST = FK != 0;
    constexpr int MJ = FK == 0 ? 0 : (FK == 3 ? 1 : 2);  // majorant table form (seg_begin)
    extern __shared__ float dyn_lds[];
    uint32_t loaded = 0xFFFFFFFFu;  // job whose majorants are in the wave's LDS copy
    // wave index, read into a scalar register so that the wave's LDS bases are
    // scalars (the replay kernels, short of scalar registers, keep it per lane)
    const uint32_t wid = ADJ ? threadIdx.x >> 6 : __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    float *wave_lds = dyn_lds + wid * (kLdsFixed + a.maj_lds + a.shell_lds + a.lam_lds);
    float *acc_sum = wave_lds;
    float *acc_sq = wave_lds + kSlotsPerWave;
    double *acc_grad = (double *)(wave_lds + 2 * kSlotsPerWave);
    const LdsRingPtr ring = (LdsRingPtr)(wave_lds + 2 * kSlotsPerWave + 32);
    float *lds_maj = wave_lds + kLdsFixed;
    float *lds_shell = lds_maj + a.maj_lds;
    // FAST: [band] 4-float coefficients, then [band] sun radiance (kMaxLambda bands each)
    v4f *lds_lam4 = (v4f *)(lds_maj + a.maj_lds + a.shell_lds);
    float *lds_lsun = (float *)(lds_lam4 + (uint32_t)kMaxLambda);
    // FAST replays: [band] A.lam_units after the sun radiance
    v4f *lds_units = (v4f *)(lds_lsun + (uint32_t)kMaxLambda);
    uint32_t wd = 0;  // a watchdog fired on this lane
    uint32_t *diag = nullptr;
    LaneCtx c;
    AdjLane al;
    for (;;) {
        // the lane index, opaque to the compiler for each item: the per-lane
        // LDS addresses of the item's set-up and output loops are then formed
        // again per item instead of being held (or spilled) across the
        // integration loop
        uint32_t lane = lane_id();
        asm volatile("" : "+v"(lane));
        uint32_t item = 0;
        if (lane == 0) item = atomicAdd(a.counter, 1u);
        item = __builtin_amdgcn_readfirstlane(item);
        if (item >= a.total_items) break;
        if (item == 0 && lane == 0) *a.t_start = wall_clock64();
        const uint32_t jid = find_job(a, item);
        const DevJob &J = job_at(a.jobs, jid);
        const DevScene &s = J.s;
        diag = s.diag;
        const uint32_t npx = s.width * s.height;
        const uint32_t local = item - J.item_base;
        const uint32_t g = local % J.groups, ch = local / J.groups;
        ItemInfo it;
        it.p0 = g * J.pixels_per_item;
        const uint32_t np = min(J.pixels_per_item, npx - it.p0);
        it.x0 = it.p0 % s.width;
        it.y0 = it.p0 / s.width;
        it.s0 = s.sample_offset + ch * J.samples_per_chunk;
        const uint32_t ns = min(J.samples_per_chunk, s.spp - ch * J.samples_per_chunk);
        it.nslots = np * s.nlam;
        // reciprocals for udiv_small (within an ulp): the hardware reciprocal,
        // kept in a scalar register, and the scene's host-side ones
        it.r_nslots = __uint_as_float(__builtin_amdgcn_readfirstlane(__float_as_uint(f_rcp((float)it.nslots))));
        it.r_nlam = s.inv_nlam;
        it.r_width = s.inv_w;
        for (uint32_t i = lane; i < it.nslots; i += 64) {
            acc_sum[i] = 0.0f;
            acc_sq[i] = 0.0f;
        }
        MajSrc msrc;
        msrc.lds = (LdsFloatPtr)lds_maj;
        msrc.lds2 = (LdsFloat2Ptr)lds_maj;
        msrc.glob = (GlobalFloatPtr)s.maj;
        const uint32_t nmaj = s.nlam * s.L;
        msrc.use_lds = FAST || (nmaj <= a.maj_lds && s.L + 1 <= a.shell_lds);
        msrc.shells.lds = (LdsShellPtr)lds_shell;
        msrc.shells.glob = (GlobalShellPtr)s.shell_w;
        msrc.shells.use_lds = msrc.use_lds;
        msrc.lam4 = (LdsFloat4Ptr)lds_lam4;
        msrc.lsun = (LdsFloatPtr)lds_lsun;
        msrc.units = (LdsFloat4Ptr)lds_units;
        if (msrc.use_lds) {
            if (loaded != jid) {
                if constexpr (MJ == 2) {
                    for (uint32_t i = lane; i < nmaj; i += 64) {
                        const float mj = s.maj[i];
                        ((v2f *)lds_maj)[i] = v2f{mj, f_rcp(mj)};
                    }
                } else {
                    for (uint32_t i = lane; i < nmaj; i += 64) lds_maj[i] = s.maj[i];
                }
                for (uint32_t i = lane; i <= s.L; i += 64) lds_shell[i] = s.shell_w[i];
                if constexpr (FAST) {
                    for (uint32_t i = lane; i < s.nlam; i += 64) {
                        const DevLambda l = s.lam[i];
                        lds_lam4[i] = v4f{l.br, l.bo, l.bas, l.baa};
                        lds_lsun[i] = l.lsun;
                        if constexpr (ADJ) {
                            const float4 u = A.lam_units[i];
                            lds_units[i] = v4f{u.x, u.y, u.z, u.w};
                        }
                    }
                }
                loaded = jid;
            }
        }
        if constexpr (ADJ) {
            if (lane <= kNP) acc_grad[lane] = 0.0;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        // Wave-uniform: the next triple to make a start for, fill_si * nslots +
        // fill_rem, and the ring's unused starts, `avail` from slot `head` on
        // (slots mod kRingSlots).
        uint32_t fill_si = 0, fill_rem = 0, head = 0, avail = 0;
        c.st = ST_NEED;
        for (;;) {
            const bool more = avail != 0 || fill_si < ns;
            const uint64_t evm = ballot(c.st == ST_EVENT);
            const uint64_t trm = ballot(c.st == ST_TRACK);
            const uint64_t ndm = more ? ballot(c.st == ST_NEED) : 0ull;
            if (!evm && !trm && !ndm) break;
            if ((evm || ndm) && ((uint32_t)__popcll(evm | ndm) >= a.event_threshold || !trm)) {
                // handler pass: events, refill of finished lanes, ray set-ups
                if constexpr (FAST) {
                    if (c.st == ST_EVENT)
                        handle_event_fast<ADJ, FK == 1 || FK == 3, FK == 3>(s, msrc, c, al, A, acc_sum, acc_sq,
                                                                            acc_grad, wd);
                } else {
                    if (c.st == ST_EVENT) handle_event<ADJ>(s, msrc, c, al, A, acc_sum, acc_sq, acc_grad, wd);
                }
                const uint64_t need = ballot(c.st == ST_NEED);
                if (need && more) {
                    // the k-th lane in NEED takes ring slot head + k: first the
                    // unused starts, then, if they do not cover every lane, the
                    // next 64 starts, made by the whole wave into the slots after
                    // them (the slots just read are free again)
                    const uint32_t n_need = (uint32_t)__popcll(need);
                    if (c.st == ST_NEED) {
                        const uint32_t rank = mbcnt(need);
                        if (rank < avail) start_path<ADJ, FAST>(s, c, al, A, ring[(head + rank) & (kRingSlots - 1)], it);
                    }
                    if (n_need > avail) {
                        const uint32_t base = head + avail;
                        uint32_t have = 0;  // starts made (triples past the item's end are not)
                        if (fill_si < ns) {
                            const uint32_t r = fill_rem + lane;  // < nslots + 64
                            const uint32_t dq = udiv_small(r, it.r_nslots);
                            const uint32_t si = fill_si + dq;
                            ring[(base + lane) & (kRingSlots - 1)] = make_start(s, si, r - dq * it.nslots, it);
                            have = (uint32_t)__popcll(ballot(si < ns));
                            const uint32_t adv = fill_rem + kRingSlots;
                            const uint32_t dqa = __builtin_amdgcn_readfirstlane(udiv_small(adv, it.r_nslots));
                            fill_si += dqa;
                            fill_rem = adv - dqa * it.nslots;
                            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                            __builtin_amdgcn_wave_barrier();
                            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                        }
                        // the lanes still in NEED, in the same order
                        const uint64_t rest = ballot(c.st == ST_NEED);
                        if (c.st == ST_NEED) {
                            const uint32_t rank = mbcnt(rest);
                            if (rank < have)
                                start_path<ADJ, FAST>(s, c, al, A, ring[(base + rank) & (kRingSlots - 1)], it);
                            else
                                c.st = ST_IDLE;
                        }
                        const uint32_t used = n_need - avail < have ? n_need - avail : have;
                        head = base + used;
                        avail = have - used;
                    } else {
                        head += n_need;
                        avail -= n_need;
                    }
                }
                if (c.st == ST_RAY) ray_pass<MJ>(s, msrc, c);
            }
            for (uint32_t i = 0; i < a.track_steps; ++i) {
                if (!ballot(c.st == ST_TRACK)) break;
                if (c.st == ST_TRACK) {
                    if constexpr (FAST)
                        track_step_fast<ADJ, FK == 1 || FK == 3, FK == 3>(s, msrc, c, al, A, wd);
                    else
                        track_step<ADJ>(s, msrc, c, al, A, wd);
                }
            }
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        if constexpr (ADJ) {
            if (lane <= kNP) atomicAdd(&A.grad[lane], acc_grad[lane]);
        } else if (J.chunks == 1) {
            // the item holds every sample of its pixels: final values now
            item_outputs(s, it.p0, np, it.nslots, acc_sum, acc_sq, lane);
        } else {
            float2 *dst = J.accum + ((size_t)ch * npx + it.p0) * s.nlam;
            for (uint32_t i = lane; i < it.nslots; i += 64) dst[i] = make_float2(acc_sum[i], acc_sq[i]);
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

// The same integrator under several register budgets: the compiler's own
// allocation and caps for at least 4, 5, 6 or 8 waves per SIMD (512 VGPRs per
// lane / waves, 8-register granules: 128, 96, 80, 64).  The autotuner picks
// one per workload.  The _fast kernels serve launches whose jobs all use
// carried optical depth, lobe sampling and delta-tracked shadow rays, the
// _fastref kernels launches whose jobs all use the reference's estimator
// (a fresh draw per shell, isotropic directions, delta-tracked shadow rays),
// the _fastrt kernels carried optical depth + lobe sampling with ratio-
// tracked shadow rays, all with at most kFastMaxShells shells and their
// tables in LDS.
__global__ __launch_bounds__(256) void k_wavepool(PoolArgs a) { wavepool_body<false, 0>(a, AdjArgs{}); }
__global__ __launch_bounds__(256) void k_wavepool_fast(PoolArgs a) { wavepool_body<false, 1>(a, AdjArgs{}); }
__global__ __launch_bounds__(256) void k_wavepool_fastref(PoolArgs a) { wavepool_body<false, 2>(a, AdjArgs{}); }
__global__ __launch_bounds__(256) void k_wavepool_fastrt(PoolArgs a) { wavepool_body<false, 3>(a, AdjArgs{}); }

#define SKY_WAVEPOOL_OCC(N, NAME, FK)                                                                        \
    __global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(N))) void NAME##_o##N(PoolArgs a) \
    {                                                                                                        \
        wavepool_body<false, FK>(a, AdjArgs{});                                                              \
    }
#define SKY_WAVEPOOL_OCC3(N)                \
    SKY_WAVEPOOL_OCC(N, k_wavepool, 0)      \
    SKY_WAVEPOOL_OCC(N, k_wavepool_fast, 1) \
    SKY_WAVEPOOL_OCC(N, k_wavepool_fastref, 2) \
    SKY_WAVEPOOL_OCC(N, k_wavepool_fastrt, 3)
SKY_WAVEPOOL_OCC3(4)
SKY_WAVEPOOL_OCC3(5)
SKY_WAVEPOOL_OCC3(6)
SKY_WAVEPOOL_OCC3(8)
#undef SKY_WAVEPOOL_OCC3
#undef SKY_WAVEPOOL_OCC

// Path-replay adjoint: the forward state machine plus the score bookkeeping.
// Capped at 4 waves per SIMD (128 VGPRs, a few spills; uncapped the replays
// take 137-144 VGPRs and run 3 waves per SIMD): 7% less replay time at
// 256x128 px x 32 bands.  A 5-wave cap (96 VGPRs) spills heavily and is slower.
#define SKY_ADJOINT(NAME, FK)                                                                                \
    __global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void NAME(PoolArgs a, AdjArgs A) \
    {                                                                                                        \
        wavepool_body<true, FK>(a, A);                                                                       \
    }
SKY_ADJOINT(k_wavepool_adjoint, 0)
SKY_ADJOINT(k_wavepool_adjoint_fast, 1)
SKY_ADJOINT(k_wavepool_adjoint_fastref, 2)
SKY_ADJOINT(k_wavepool_adjoint_fastrt, 3)
#undef SKY_ADJOINT

// Per job (blockIdx.y) whose samples were split into several chunks: reduces
// the chunk partials in chunk order, normalises by 1/spp and clamps
// (integrator.rs:307-310), and forms XYZ = sum_band film * cmf * weight.  A
// block takes P = 256 / nlam whole pixels: thread t owns entry t of their
// [pixel][band] rows, so every chunk's partials are read in one coalesced
// sweep; the block's film values then sit in LDS for the per-pixel XYZ sums
// (band order, as item_outputs).  Jobs done in one chunk wrote their outputs
// at the end of their items.
__global__ __launch_bounds__(256) void k_finalize(const DevJob *jobs)
{
    __shared__ float sf[256];
    const DevJob &J = job_at(jobs, blockIdx.y);
    if (J.chunks <= 1) return;
    const DevScene &s = J.s;
    const uint32_t npx = s.width * s.height, nl = s.nlam;
    const uint32_t P = 256u / nl;
    const uint32_t px0 = blockIdx.x * P;
    if (px0 >= npx) return;
    const uint32_t npix = min(P, npx - px0);
    const uint32_t t = threadIdx.x;
    const float inv = 1.0f / (float)s.spp;
    float f = 0.0f;
    if (t < npix * nl) {
        const size_t e = (size_t)px0 * nl + t;
        float sum = 0.0f, sq = 0.0f;
        for (uint32_t ch = 0; ch < J.chunks; ++ch) {
            const float2 v = J.accum[(size_t)ch * npx * nl + e];
            sum += v.x;
            sq += v.y;
        }
        f = fmaxf(sum * inv, 0.0f);
        sf[t] = f;
        if (s.film) s.film[e] = f;
        if (s.film_m2) s.film_m2[e] = sq * inv;
    }
    if (s.xyz && nl >= 4 && (nl & (nl - 1)) == 0) {
        // the same wave-level sums as the work items' outputs: entry t is
        // band t % nl of pixel t / nl, groups of nl lanes lie inside a wave
        const DevLambda l = s.lam[t & (nl - 1)];
        const float X = group_sum(f * l.wx, nl), Y = group_sum(f * l.wy, nl), Z = group_sum(f * l.wz, nl);
        if (t < npix * nl && (t & (nl - 1)) == 0) {
            float *o = s.xyz + 3 * (size_t)(px0 + t / nl);
            o[0] = X;
            o[1] = Y;
            o[2] = Z;
        }
        return;
    }
    __syncthreads();
    if (s.xyz && t < npix) {
        float X = 0.0f, Y = 0.0f, Z = 0.0f;
        for (uint32_t b = 0; b < nl; ++b) {
            const float f = sf[t * nl + b];
            const DevLambda l = s.lam[b];
            X += f * l.wx;
            Y += f * l.wy;
            Z += f * l.wz;
        }
        float *o = s.xyz + 3 * (size_t)(px0 + t);
        o[0] = X;
        o[1] = Y;
        o[2] = Z;
    }
}

// Blocks of k_finalize for a job (0 when the job needs none).
uint32_t finalize_blocks(const DevJob &j)
{
    if (j.chunks <= 1) return 0;
    const uint32_t P = 256u / j.s.nlam;
    return (j.s.width * j.s.height + P - 1) / P;
}

// Fills n words with zero: used instead of a memset node so that the launch
// sequence can be captured in a HIP graph.
__global__ void k_zero_u32(uint32_t *p, uint32_t n)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// Launch shape for zero tuning fields: what the autotuner picks on MI355X for
// large renders (64 and 32 bands, both estimators); worth 1.3-1.6x over the
// earlier 16 waves per CU, threshold 32 and 8 steps for every shell set.
constexpr uint32_t kDefaultWavesPerCu = 24, kDefaultEventThreshold = 8, kDefaultTrackSteps = 2;

// Triples of one item are numbered j < pixels x bands x samples, which the
// kernel splits with float reciprocals: exact below 2^24.
constexpr uint32_t kMaxItemTriples = 1u << 24;

// Default pixels per work item: the fewest that give an item of all spp
// samples at least 256 paths per lane (16384 triples).  Longer items spend
// less of their time draining (the lanes idle while the item's last paths
// finish); shorter ones leave less idle time at the end of the launch, when
// the last items finish at different times.
constexpr uint32_t kItemTriplesTarget = 16384;

void plan_job(const DevScene &s, const sky_tuning &t, int num_cus, uint32_t min_items, DevJob &j)
{
    const uint32_t npx = s.width * s.height;
    uint32_t cap = kSlotsPerWave / s.nlam;
    const uint64_t per_pixel = (uint64_t)s.nlam * s.spp;
    uint32_t want = t.pixels_per_item ? t.pixels_per_item
                                      : (uint32_t)((kItemTriplesTarget + per_pixel - 1) / per_pixel > cap
                                                       ? cap
                                                       : (kItemTriplesTarget + per_pixel - 1) / per_pixel);
    j.pixels_per_item = want < 1 ? 1 : (want > cap ? cap : want);
    j.groups = (npx + j.pixels_per_item - 1) / j.pixels_per_item;
    // split samples only as far as needed for min_items, keeping >= 16 samples per item
    uint64_t chunks = ((uint64_t)min_items + j.groups - 1) / j.groups;
    uint64_t max_chunks = s.spp / 16 ? s.spp / 16 : 1;
    if (chunks > max_chunks) chunks = max_chunks;
    if (chunks < 1) chunks = 1;
    const uint64_t slots = (uint64_t)j.pixels_per_item * s.nlam;
    const uint64_t max_spc = kMaxItemTriples / slots;
    const uint64_t min_chunks = (s.spp + max_spc - 1) / max_spc;
    if (chunks < min_chunks) chunks = min_chunks;
    (void)num_cus;
    j.samples_per_chunk = (uint32_t)((s.spp + chunks - 1) / chunks);
    j.chunks = (s.spp + j.samples_per_chunk - 1) / j.samples_per_chunk;
    j.n_items = j.groups * j.chunks;
}

// Partial-sum buffer of a job: only jobs split into several sample chunks
// need one (a single-chunk item writes its pixels' outputs itself).
size_t job_accum_bytes(const DevJob &j)
{
    if (j.chunks <= 1) return 0;
    return (size_t)j.chunks * j.s.width * j.s.height * j.s.nlam * sizeof(float2);
}

WavePlan plan_launch(const sky_tuning &t, int num_cus, uint32_t total_items, uint32_t max_maj,
                     uint32_t max_layers, uint32_t flags_all)
{
    WavePlan p;
    p.maj_lds = max_maj <= kMajLdsMax ? (max_maj + 63u) & ~63u : 0u;
    const uint32_t est = flags_all & (kFlagCarriedTau | kFlagPhaseIS | kFlagRatio);
    const bool fits = p.maj_lds != 0 && max_layers <= kFastMaxShells && flags_all != kFlagMixed;
    p.fast = !fits                                                  ? 0
             : est == (kFlagCarriedTau | kFlagPhaseIS)                ? 1
             : est == 0                                               ? 2
             : est == (kFlagCarriedTau | kFlagPhaseIS | kFlagRatio) ? 3
                                                                      : 0;
    p.shell_lds = p.maj_lds ? ((max_layers + 1u + 7u) & ~7u) : 0u;
    if (p.fast == 1 || p.fast == 2) p.maj_lds *= 2;  // (majorant, reciprocal) pairs
    p.lam_lds = p.fast ? 5u * (uint32_t)kMaxLambda : 0u;
    p.lds_bytes = kWavesPerBlock * (kLdsFixed + p.maj_lds + p.shell_lds + p.lam_lds) * (uint32_t)sizeof(float);
    uint32_t wpc = t.waves_per_cu ? t.waves_per_cu : kDefaultWavesPerCu;
    uint64_t waves = (uint64_t)(num_cus > 0 ? num_cus : 256) * wpc;
    uint64_t w = waves < total_items ? waves : total_items;
    p.block = 64 * kWavesPerBlock;
    p.grid = (uint32_t)((w + kWavesPerBlock - 1) / kWavesPerBlock);
    p.n_items = total_items;
    p.event_threshold = t.event_threshold ? t.event_threshold : kDefaultEventThreshold;
    p.track_steps = t.track_steps ? t.track_steps : kDefaultTrackSteps;
    p.occupancy = t.occupancy;
    return p;
}

uint32_t launch_min_items(const sky_tuning &t, int num_cus)
{
    uint32_t wpc = t.waves_per_cu ? t.waves_per_cu : kDefaultWavesPerCu;
    return 4u * (uint32_t)(num_cus > 0 ? num_cus : 256) * wpc;
}

// meta: [0] item counter, [1..2] launch timestamp, [4 + 2 j ..] per-job done
// counter / pad, then per-job finish timestamps; zeroed by the first kernel.
hipError_t launch_pool(const DevJob *d_jobs, uint32_t n_jobs, uint32_t fin_blocks, const WavePlan &p,
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
    a.maj_lds = p.maj_lds;
    a.shell_lds = p.shell_lds;
    a.lam_lds = p.lam_lds;
    typedef void (*PoolKernel)(PoolArgs);
    static const PoolKernel kernels[4][5] = {
        {k_wavepool, k_wavepool_o4, k_wavepool_o5, k_wavepool_o6, k_wavepool_o8},
        {k_wavepool_fast, k_wavepool_fast_o4, k_wavepool_fast_o5, k_wavepool_fast_o6, k_wavepool_fast_o8},
        {k_wavepool_fastref, k_wavepool_fastref_o4, k_wavepool_fastref_o5, k_wavepool_fastref_o6,
         k_wavepool_fastref_o8},
        {k_wavepool_fastrt, k_wavepool_fastrt_o4, k_wavepool_fastrt_o5, k_wavepool_fastrt_o6, k_wavepool_fastrt_o8}};
    const int oi = p.occupancy == 4 ? 1 : p.occupancy == 5 ? 2 : p.occupancy == 6 ? 3 : p.occupancy == 8 ? 4 : 0;
    const PoolKernel k = kernels[p.fast][oi];
    hipLaunchKernelGGL(k, dim3(p.grid), dim3(p.block), p.lds_bytes, stream, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (fin_blocks == 0) return hipSuccess;
    hipLaunchKernelGGL(k_finalize, dim3(fin_blocks, n_jobs), dim3(256), 0, stream, d_jobs);
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
    a.maj_lds = p.maj_lds;
    a.shell_lds = p.shell_lds;
    // the FAST replays keep A.lam_units in LDS beside the per-band coefficients
    const uint32_t units_lds = p.fast ? 4u * (uint32_t)kMaxLambda : 0u;
    a.lam_lds = p.lam_lds + units_lds;
    const uint32_t lds_bytes = p.lds_bytes + kWavesPerBlock * units_lds * (uint32_t)sizeof(float);
    typedef void (*AdjKernel)(PoolArgs, AdjArgs);
    static const AdjKernel kernels[4] = {k_wavepool_adjoint, k_wavepool_adjoint_fast, k_wavepool_adjoint_fastref,
                                         k_wavepool_adjoint_fastrt};
    hipLaunchKernelGGL(kernels[p.fast], dim3(p.grid), dim3(p.block), lds_bytes, stream, a, A);
    return hipGetLastError();
}

hipError_t launch_zero(uint32_t *p, uint32_t n, hipStream_t stream)
{
    hipLaunchKernelGGL(k_zero_u32, dim3((n + 255) / 256), dim3(256), 0, stream, p, n);
    return hipGetLastError();
}

}  // namespace sky
