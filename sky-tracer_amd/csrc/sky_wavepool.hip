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

struct WorkDesc {
    uint32_t pixels_per_item;  // P, P * nlam <= kSlotsPerWave
    uint32_t groups;           // ceil(npx / P)
    uint32_t chunks;           // sample chunks C
    uint32_t samples_per_chunk;
    uint32_t n_items;          // groups * chunks
    uint32_t event_threshold;
    uint32_t track_steps;
    float2 *accum;             // [chunk][pixel][band] (sum, sum of squares)
    uint32_t *counter;
};

__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & 63u; }

__device__ __forceinline__ uint32_t mbcnt(uint64_t m)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

// Everything a lane keeps between loop trips.
struct LaneCtx {
    int st, ev;
    uint32_t slot, band;
    DevLambda l;
    Rng rng;
    float tp, L;
    uint32_t depth;
    // current vertex
    double px, py, pz;
    f3 vin;
    int vtype;  // 0 scattering, 1 ground
    float nee_sum, nee_factor;
    float rw;  // sigma_R / sigma_s at a scattering vertex (phase mixture weight)
    uint32_t nee_i;
    // pending main ray (origin, direction) between vertices
    f3 dir;
    // ray being tracked
    Ray ray;
    int mode;  // TrackMode
    float t, tau, seg_end, maj, T, ratio;
    int k, out, kn, on, pend;
    uint32_t guard;
};

__device__ __forceinline__ void seg_begin(const DevScene &s, LaneCtx &c)
{
    float tc = next_crossing(s, c.ray, c.k, c.out, c.kn, c.on);
    c.seg_end = fmaxf(fminf(tc, c.ray.tmax), c.t);
    c.maj = s.maj[c.band * s.L + c.k];
}

__device__ __forceinline__ void track_begin(const DevScene &s, LaneCtx &c, int mode)
{
    c.mode = mode;
    c.t = 0.0f;
    c.tau = -1.0f;
    c.T = 1.0f;
    c.pend = 0;
    c.guard = 1;
    c.k = c.ray.layer;
    c.out = c.ray.outward;
    seg_begin(s, c);
    c.st = ST_TRACK;
}

// One tracking micro-step: at most one RNG draw.  An empty segment (its
// end clamped to t) is crossed without a draw, as the reference's inner
// `while t < segment_end` loop does.
__device__ __forceinline__ void track_step(const DevScene &s, LaneCtx &c, bool &wd)
{
    const bool carried = (s.flags & kFlagCarriedTau) != 0;
    const bool empty = !(c.t < c.seg_end);
    const bool need_draw = c.pend || (!empty && (!carried || c.tau < 0.0f));
    float u = 0.0f;
    if (need_draw) u = rng_next(c.rng);
    if (c.pend) {
        c.pend = 0;
        if (u < c.ratio) {
            c.st = ST_EVENT;
            if (c.mode == kTrackCollision) {
                c.ev = EV_HIT;
            } else {
                c.T = 0.0f;
                c.ev = EV_SHADOW;
            }
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
        if (c.t >= c.ray.tmax) {
            c.st = ST_EVENT;
            c.ev = c.mode == kTrackCollision ? EV_BOUNDARY : EV_SHADOW;
            return;
        }
        if (c.guard > s.max_segments) {
            wd = true;
            c.st = ST_EVENT;
            c.T = 0.0f;
            c.ev = c.mode == kTrackCollision ? EV_BOUNDARY : EV_SHADOW;
            c.ray.kind = 2;  // watchdog: ends the path without contribution
            return;
        }
        c.guard++;
        c.k = c.kn;
        c.out = c.on;
        seg_begin(s, c);
        return;
    }
    c.t = tn;
    Coeffs co = medium(s, c.l, altitude(s, c.ray, c.t));
    float st = co.sr + co.sas + co.so + co.saa;
    float ratio = fminf(fmaxf(st / c.maj, 0.0f), 1.0f);
    if (c.mode == kTrackRatio) {
        c.T *= 1.0f - ratio;
    } else {
        c.pend = 1;
        c.ratio = ratio;
    }
}

__device__ __forceinline__ void vertex_pos(const Ray &r, float t, double &px, double &py, double &pz)
{
    double dx = r.d.x, dy = r.d.y, dz = r.d.z;
    double inv = 1.0 / sqrt(dx * dx + dy * dy + dz * dz);
    px = r.ox + dx * inv * (double)t;
    py = r.oy + dy * inv * (double)t;
    pz = r.oz + dz * inv * (double)t;
}

enum Act : int { ACT_NONE = 0, ACT_NEE = 1, ACT_CONT = 2, ACT_MAIN = 3, ACT_FINISH = 4, ACT_TRACKING = 5 };

// Event handler: sky_trace.wgsl:703-776 between two tracking passes.
__device__ __forceinline__ void handle_event(const DevScene &s, LaneCtx &c, float *acc_sum,
                                             float *acc_sq, bool &wd)
{
    const uint32_t n_light = s.dls ? s.dls : 1u;
    int act = ACT_NONE;
    if (c.ev == EV_START) {
        act = ACT_MAIN;
    } else if (c.ev == EV_BOUNDARY) {
        if (c.ray.kind == 1) {
            // exit: the sun disk is seen only by the camera ray (wgsl:745-749)
            if (c.depth == 0 && in_sun_cone(s, c.ray.d)) c.L = c.L + c.tp * c.l.lsun;
            act = ACT_FINISH;
        } else if (c.ray.kind == 2) {
            act = ACT_FINISH;  // tracking watchdog
        } else {
            vertex_pos(c.ray, c.ray.tmax, c.px, c.py, c.pz);
            c.vtype = 1;
            c.nee_i = 0;
            c.nee_sum = 0.0f;
            act = ACT_NEE;
        }
    } else if (c.ev == EV_HIT) {
        vertex_pos(c.ray, c.t, c.px, c.py, c.pz);
        Coeffs co = medium(s, c.l, altitude(s, c.ray, c.t));
        float scat = co.sr + co.sas;
        float ext = scat + co.so + co.saa;
        act = ACT_FINISH;
        if (ext > 0.0f && scat > 0.0f) {
            c.tp = c.tp * fminf(fmaxf(scat / ext, 0.0f), 1.0f);
            if (c.tp > 0.0f) {
                c.vtype = 0;
                c.vin = c.ray.d;
                c.rw = co.sr / scat;
                c.nee_i = 0;
                c.nee_sum = 0.0f;
                act = ACT_NEE;
            }
        }
    } else {  // EV_SHADOW
        c.nee_sum += c.nee_factor * c.T;
        c.nee_i++;
        act = c.nee_i < n_light ? ACT_NEE : ACT_CONT;
    }

    // Sun NEE: direct_sun_at_scatter / ground_radiance (wgsl:566-608).
    while (act == ACT_NEE) {
        float a0 = rng_next(c.rng);
        float a1 = rng_next(c.rng);
        f3 ld = sun_cone_dir(s, a0, a1);
        float factor;
        double nx = 0, ny = 0, nz = 0;
        if (c.vtype == 1) {
            double pn = 1.0 / sqrt(c.px * c.px + c.py * c.py + c.pz * c.pz);
            nx = c.px * pn;
            ny = c.py * pn;
            nz = c.pz * pn;
            float cs = fmaxf((float)nx * ld.x + (float)ny * ld.y + (float)nz * ld.z, 0.0f);
            factor = s.albedo * kInvPi * c.l.lsun * cs / s.cone_pdf;
            if (!(cs > 0.0f)) {
                c.nee_i++;
                act = c.nee_i < n_light ? ACT_NEE : ACT_CONT;
                continue;
            }
        } else {
            float mu = dot3(c.vin, ld);
            float ph = c.rw * phase_rayleigh(mu) + (1.0f - c.rw) * phase_hg(s.g, mu);
            factor = c.l.lsun * ph / s.cone_pdf;
        }
        ray_setup(s, c.px + (double)ld.x * kRayEpsilonKm, c.py + (double)ld.y * kRayEpsilonKm,
                  c.pz + (double)ld.z * kRayEpsilonKm, ld, c.ray);
        if (!(c.ray.hit && c.ray.kind == 1)) {
            c.nee_i++;
            act = c.nee_i < n_light ? ACT_NEE : ACT_CONT;
            continue;
        }
        c.nee_factor = factor;
        track_begin(s, c, (s.flags & kFlagRatio) ? kTrackRatio : kTrackDelta);
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
            } else {
                float z = 1.0f - 2.0f * rng_next(c.rng);
                float rr = sqrtf(fmaxf(1.0f - z * z, 0.0f));
                float phi = kTau * rng_next(c.rng);
                float sp, cp;
                sincosf(phi, &sp, &cp);
                nd = f3{rr * cp, z, rr * sp};
                float mu = dot3(c.vin, nd);
                w = (rw * phase_rayleigh(mu) + (1.0f - rw) * phase_hg(s.g, mu)) / kIsotropicPdf;
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
        }
        act = ACT_FINISH;
        if (ok) {
            c.tp = c.tp * w;
            if (roulette(s, c.tp, c.depth, c.rng)) {
                if (c.L < 0.0f || c.L != c.L || c.tp < 0.0f || c.tp != c.tp) {
                    c.L = 0.0f;
                } else {
                    c.depth++;
                    c.px += (double)nd.x * kRayEpsilonKm;
                    c.py += (double)nd.y * kRayEpsilonKm;
                    c.pz += (double)nd.z * kRayEpsilonKm;
                    c.dir = nd;
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
            } else {
                ray_setup(s, c.px, c.py, c.pz, c.dir, c.ray);
                if (c.ray.hit) {
                    track_begin(s, c, kTrackCollision);
                    act = ACT_TRACKING;
                }
            }
        }
    }

    if (act == ACT_FINISH) {
        float Lf = fmaxf(c.L, 0.0f);
        atomicAdd(&acc_sum[c.slot], Lf);
        atomicAdd(&acc_sq[c.slot], Lf * Lf);
        c.st = ST_NEED;
    }
}

// Path start (main, sky_trace.wgsl:799-807): triple j of the item.
__device__ __forceinline__ void start_path(const DevScene &s, LaneCtx &c, uint32_t j, uint32_t p0,
                                           uint32_t np, uint32_t s0)
{
    uint32_t per_sample = np * s.nlam;
    uint32_t si = j / per_sample;
    uint32_t rem = j - si * per_sample;
    uint32_t pl = rem / s.nlam;
    c.band = rem - pl * s.nlam;
    c.slot = rem;
    uint32_t pixel = p0 + pl;
    uint32_t x = pixel % s.width, y = pixel / s.width;
    c.l = s.lam[c.band];
    c.rng = rng_make(pixel, s0 + si, c.band, s);
    float j1 = rng_next(c.rng);
    float j2 = rng_next(c.rng);
    c.dir = camera_dir(s, x, y, j1, j2);
    c.px = 0.0;
    c.py = s.obs_y;
    c.pz = 0.0;
    c.tp = 1.0f;
    c.L = 0.0f;
    c.depth = 0;
    c.st = ST_EVENT;
    c.ev = EV_START;
}

__global__ __launch_bounds__(256) void k_wavepool(DevScene s, WorkDesc w)
{
    __shared__ float lds_sum[kWavesPerBlock][kSlotsPerWave];
    __shared__ float lds_sq[kWavesPerBlock][kSlotsPerWave];
    const uint32_t wid = threadIdx.x >> 6;
    const uint32_t lane = lane_id();
    float *acc_sum = lds_sum[wid];
    float *acc_sq = lds_sq[wid];
    const uint32_t npx = s.width * s.height;
    bool wd = false;
    LaneCtx c;
    for (;;) {
        uint32_t item = 0;
        if (lane == 0) item = atomicAdd(w.counter, 1u);
        item = __builtin_amdgcn_readfirstlane(item);
        if (item >= w.n_items) break;
        const uint32_t g = item % w.groups, ch = item / w.groups;
        const uint32_t p0 = g * w.pixels_per_item;
        const uint32_t np = min(w.pixels_per_item, npx - p0);
        const uint32_t s0 = s.sample_offset + ch * w.samples_per_chunk;
        const uint32_t ns = min(w.samples_per_chunk, s.spp - ch * w.samples_per_chunk);
        const uint32_t nslots = np * s.nlam;
        const uint32_t total = nslots * ns;
        for (uint32_t i = lane; i < nslots; i += 64) {
            acc_sum[i] = 0.0f;
            acc_sq[i] = 0.0f;
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
                        start_path(s, c, j, p0, np, s0);
                    else
                        c.st = ST_IDLE;
                }
                next += (uint32_t)__popcll(need);
            }
            uint64_t evm = ballot(c.st == ST_EVENT);
            uint64_t trm = ballot(c.st == ST_TRACK);
            if (!evm && !trm) break;
            if (evm && ((uint32_t)__popcll(evm) >= w.event_threshold || !trm)) {
                if (c.st == ST_EVENT) handle_event(s, c, acc_sum, acc_sq, wd);
            }
            for (uint32_t i = 0; i < w.track_steps; ++i) {
                if (!ballot(c.st == ST_TRACK)) break;
                if (c.st == ST_TRACK) track_step(s, c, wd);
            }
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        float2 *dst = w.accum + ((size_t)ch * npx + p0) * s.nlam;
        for (uint32_t i = lane; i < nslots; i += 64) dst[i] = make_float2(acc_sum[i], acc_sq[i]);
        __builtin_amdgcn_wave_barrier();
    }
    if (wd) atomicOr(s.diag, 1u);
}

// Reduces the chunk partials in chunk order, normalises by 1/spp and clamps
// (integrator.rs:307-310), and forms XYZ = sum_band film * cmf * weight.
__global__ __launch_bounds__(256) void k_finalize(DevScene s, const float2 *accum, uint32_t chunks)
{
    uint32_t pixel = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t npx = s.width * s.height;
    if (pixel >= npx) return;
    float inv = 1.0f / (float)s.spp;
    float X = 0.0f, Y = 0.0f, Z = 0.0f;
    for (uint32_t b = 0; b < s.nlam; ++b) {
        float sum = 0.0f, sq = 0.0f;
        for (uint32_t ch = 0; ch < chunks; ++ch) {
            float2 v = accum[((size_t)ch * npx + pixel) * s.nlam + b];
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

__global__ void k_zero_u32(uint32_t *p, uint32_t n)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

WavePlan plan_wavepool(const DevScene &s, const sky_tuning &t, int num_cus)
{
    WavePlan p;
    const uint32_t npx = s.width * s.height;
    uint32_t cap = kSlotsPerWave / s.nlam;
    uint32_t want = t.pixels_per_item ? t.pixels_per_item : (128u / s.nlam ? 128u / s.nlam : 1u);
    p.pixels_per_item = want < 1 ? 1 : (want > cap ? cap : want);
    p.groups = (npx + p.pixels_per_item - 1) / p.pixels_per_item;
    uint32_t wpc = t.waves_per_cu ? t.waves_per_cu : 16u;
    uint64_t waves = (uint64_t)(num_cus > 0 ? num_cus : 256) * wpc;
    // enough items to balance the persistent waves, but >= 16 samples per item
    uint64_t chunks = (4 * waves + p.groups - 1) / p.groups;
    uint64_t max_chunks = s.spp / 16 ? s.spp / 16 : 1;
    if (chunks > max_chunks) chunks = max_chunks;
    if (chunks < 1) chunks = 1;
    p.samples_per_chunk = (uint32_t)((s.spp + chunks - 1) / chunks);
    p.chunks = (s.spp + p.samples_per_chunk - 1) / p.samples_per_chunk;
    p.n_items = p.groups * p.chunks;
    uint64_t w = waves < p.n_items ? waves : p.n_items;
    p.block = 64 * kWavesPerBlock;
    p.grid = (uint32_t)((w + kWavesPerBlock - 1) / kWavesPerBlock);
    p.accum_bytes = (size_t)p.chunks * npx * s.nlam * sizeof(float2);
    p.event_threshold = t.event_threshold ? t.event_threshold : 32u;
    p.track_steps = t.track_steps ? t.track_steps : 8u;
    return p;
}

hipError_t launch_wavepool(const DevScene &s, const WavePlan &p, void *accum, uint32_t *counter,
                           hipStream_t stream)
{
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, stream, counter, 1u);
    WorkDesc w;
    w.pixels_per_item = p.pixels_per_item;
    w.groups = p.groups;
    w.chunks = p.chunks;
    w.samples_per_chunk = p.samples_per_chunk;
    w.n_items = p.n_items;
    w.event_threshold = p.event_threshold;
    w.track_steps = p.track_steps;
    w.accum = (float2 *)accum;
    w.counter = counter;
    hipLaunchKernelGGL(k_wavepool, dim3(p.grid), dim3(p.block), 0, stream, s, w);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    uint32_t npx = s.width * s.height;
    hipLaunchKernelGGL(k_finalize, dim3((npx + 255) / 256), dim3(256), 0, stream, s,
                       (const float2 *)accum, p.chunks);
    return hipGetLastError();
}

}  // namespace sky
