// sky_api.cpp -- C ABI of libsky_hip.so (declared in include/sky_hip.h).
//
// render() of crates/sky-spectral-path-tracer/src/integrator.rs:125-317
// re-designed for one MI355X: a long-lived context owns the device, a stream
// and grow-only device buffers (the reference builds a wgpu device, pipeline
// and buffers on every call, integrator.rs:171-250); one wave-pool launch
// replaces the 2-spp dispatch loop (integrator.rs:252-283); normalisation and
// XYZ happen on the device, and only the requested outputs are copied back.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "sky_internal.hpp"

namespace {

thread_local std::string g_last_error;

// Working memory of sky_render_device_noalloc: one copy per GPU, held until
// the process ends.  Host code never passes `scratch` itself to a kernel or
// a copy (in host code the name is only a stand-in for the device symbol):
// each context looks up its device's address with hipGetSymbolAddress when
// it is created and uses that pointer.
static __device__ __align__(256) uint8_t scratch[SKY_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
uint8_t *g_scratch_ptr[kMaxDevices] = {};

// The device's scratch address, looked up once per device with that device
// current (the caller has made it current).
hipError_t scratch_address(int device, uint8_t **out)
{
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[device] = (uint8_t *)p;
    }
    *out = g_scratch_ptr[device];
    return hipSuccess;
}

sky_status fail(sky_status st, const std::string &msg)
{
    g_last_error = msg;
    return st;
}

struct DevBuf {
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
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
};

}  // namespace

struct sky_context {
    int device = 0;
    uint8_t *scratch = nullptr;  // this device's sky_render_device_noalloc scratch (device address)
    int num_cus = 256;
    double clock_khz = 100000.0;  // wall_clock64 rate
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    sky_tuning tuning{};
    DevBuf lam, maj, shells, film, film_m2, xyz, accum, meta, diag, rays, tau, lamd, jobs, adj_units, adj_dl,
        adj_grad, tab, b_tab;
    std::vector<sky::DevJob> host_jobs;  // staging, alive until the stream syncs
    std::vector<uint8_t> host_tab, host_btab;  // medium tables, alive until the stream syncs
    // ragged batches
    DevBuf b_lam, b_maj, b_shells, b_xyz, b_film, b_accum, b_meta, b_jobs;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    uint64_t graph_key = 0;
};

namespace {

sky_status hip_fail(hipError_t e, const char *what, sky_status st = SKY_ERR_SYNC)
{
    if (e == hipErrorOutOfMemory) st = SKY_ERR_OOM;
    return fail(st, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIP_TRY(expr, what, st)                                      \
    do {                                                             \
        hipError_t e_ = (expr);                                      \
        if (e_ != hipSuccess) return hip_fail(e_, what, st);         \
    } while (0)

// Device buffers of one render.
struct RenderBufs {
    void *lam = nullptr, *maj = nullptr, *shells = nullptr, *diag = nullptr, *meta = nullptr, *job = nullptr,
         *tab = nullptr, *accum = nullptr;
};

// Carves the buffers of a render from the allocation-free entry point's
// scratch array (every offset 256-byte aligned); false if they do not fit.
bool carve_scratch(uint8_t *base, const sky::HostScene &hs, size_t accum_bytes, size_t tab_bytes, RenderBufs &b)
{
    size_t off = 0;
    auto take = [&](size_t bytes) {
        void *p = base + off;
        off += (bytes + 255) & ~size_t(255);
        return p;
    };
    b.meta = take(64 * sizeof(uint32_t));
    b.diag = take(16 * sizeof(uint32_t));
    b.job = take(sizeof(sky::DevJob));
    b.lam = take(hs.lam.size() * sizeof(sky::DevLambda));
    b.maj = take(hs.maj.size() * sizeof(float));
    b.shells = take(hs.shell_w.size() * sizeof(float));
    b.tab = tab_bytes ? take(tab_bytes) : nullptr;
    b.accum = accum_bytes ? take(accum_bytes) : nullptr;
    return off <= SKY_NOALLOC_SCRATCH_BYTES;
}

// Runs one render into device buffers; outputs already set in hs.dev.  The
// working buffers are the context's grow-only device buffers, or with
// `scratch` (the allocation-free entry point) carved from that device
// array: then nothing is allocated on the device, the work split is the
// usual one when its partial sums fit, else one sample chunk per pixel group.
sky_status run_render(sky_context *ctx, sky::HostScene &hs, hipStream_t stream, sky_render_stats *stats,
                      uint8_t *scratch = nullptr)
{
    sky::DevScene &s = hs.dev;
    const bool simple = (s.flags & SKY_FLAG_SIMPLE_KERNEL) != 0;
    const uint32_t npx = s.width * s.height;
    sky::DevJob j{};
    if (!simple) {
        j.s = s;
        sky::plan_job(s, ctx->tuning, ctx->num_cus, sky::launch_min_items(ctx->tuning, ctx->num_cus), j);
    }
    if (!hs.tab.empty()) hs.tab.pack(ctx->host_tab);
    const size_t tab_bytes = hs.tab.empty() ? 0 : ctx->host_tab.size();
    RenderBufs b;
    if (scratch) {
        if (simple) return fail(SKY_ERR_INVALID_CONFIG, "the allocation-free entry point runs the wave-pool integrator");
        if (!carve_scratch(scratch, hs, sky::job_accum_bytes(j), tab_bytes, b)) {
            sky::plan_job(s, ctx->tuning, ctx->num_cus, 0, j);  // one chunk per pixel group when spp allows
            if (!carve_scratch(scratch, hs, sky::job_accum_bytes(j), tab_bytes, b))
                return fail(SKY_ERR_INVALID_CONFIG,
                            "render needs more than the allocation-free entry point's " +
                                std::to_string(SKY_NOALLOC_SCRATCH_BYTES >> 20) +
                                " MiB of scratch (tables plus sample-chunk partials; fewer samples per pixel, "
                                "wavelengths or medium breakpoints fit)");
        }
    } else {
        HIP_TRY(ctx->lam.ensure(hs.lam.size() * sizeof(sky::DevLambda)), "alloc lambda table", SKY_ERR_OOM);
        HIP_TRY(ctx->maj.ensure(hs.maj.size() * sizeof(float)), "alloc majorants", SKY_ERR_OOM);
        HIP_TRY(ctx->shells.ensure(hs.shell_w.size() * sizeof(float)), "alloc shell table", SKY_ERR_OOM);
        HIP_TRY(ctx->diag.ensure(64), "alloc diagnostics", SKY_ERR_OOM);
        HIP_TRY(ctx->meta.ensure(64), "alloc work counters", SKY_ERR_OOM);
        HIP_TRY(ctx->jobs.ensure(sizeof(sky::DevJob)), "alloc job table", SKY_ERR_OOM);
        if (tab_bytes) HIP_TRY(ctx->tab.ensure(tab_bytes), "alloc medium table", SKY_ERR_OOM);
        if (!simple) HIP_TRY(ctx->accum.ensure(sky::job_accum_bytes(j)), "alloc partial sums", SKY_ERR_OOM);
        if (simple && !s.film) {
            HIP_TRY(ctx->film.ensure((size_t)npx * s.nlam * sizeof(float)), "alloc film", SKY_ERR_OOM);
            s.film = (float *)ctx->film.p;
        }
        b.lam = ctx->lam.p;
        b.maj = ctx->maj.p;
        b.shells = ctx->shells.p;
        b.diag = ctx->diag.p;
        b.meta = ctx->meta.p;
        b.job = ctx->jobs.p;
        b.tab = ctx->tab.p;
        b.accum = ctx->accum.p;
    }
    HIP_TRY(hipMemcpyAsync(b.lam, hs.lam.data(), hs.lam.size() * sizeof(sky::DevLambda), hipMemcpyHostToDevice,
                           stream),
            "upload lambda table", SKY_ERR_COPY);
    HIP_TRY(hipMemcpyAsync(b.maj, hs.maj.data(), hs.maj.size() * sizeof(float), hipMemcpyHostToDevice, stream),
            "upload majorants", SKY_ERR_COPY);
    HIP_TRY(hipMemcpyAsync(b.shells, hs.shell_w.data(), hs.shell_w.size() * sizeof(float), hipMemcpyHostToDevice,
                           stream),
            "upload shell table", SKY_ERR_COPY);
    HIP_TRY(sky::launch_zero((uint32_t *)b.diag, 16, stream), "clear diagnostics", SKY_ERR_SYNC);
    if (tab_bytes)
        HIP_TRY(hipMemcpyAsync(b.tab, ctx->host_tab.data(), tab_bytes, hipMemcpyHostToDevice, stream),
                "upload medium table", SKY_ERR_COPY);
    hs.tab.attach(s, b.tab);
    s.lam = (const sky::DevLambda *)b.lam;
    s.maj = (const float *)b.maj;
    s.shell_w = (const float *)b.shells;
    s.diag = (uint32_t *)b.diag;
    sky::LaunchShape shape;
    sky::WavePlan plan;
    if (!simple) {
        j.s = s;
        j.item_base = 0;
        j.accum = (float2 *)b.accum;
        j.done = (uint32_t *)b.meta + 4;
        j.t_done = (unsigned long long *)((uint32_t *)b.meta + 6);
        ctx->host_jobs.assign(1, j);
        HIP_TRY(hipMemcpyAsync(b.job, ctx->host_jobs.data(), sizeof(j), hipMemcpyHostToDevice, stream), "upload job",
                SKY_ERR_COPY);
        plan = sky::plan_launch(ctx->tuning, ctx->num_cus, j.n_items, s.nlam * s.L, s.L, s.flags);
    }
    HIP_TRY(hipEventRecord(ctx->ev0, stream), "event record", SKY_ERR_SYNC);
    if (simple) {
        HIP_TRY(sky::launch_simple(s, stream, shape), "launch simple integrator", SKY_ERR_SYNC);
        if (s.xyz) HIP_TRY(sky::launch_film_to_xyz(s, stream), "launch xyz", SKY_ERR_SYNC);
    } else {
        HIP_TRY(sky::launch_pool((const sky::DevJob *)b.job, 1, sky::finalize_blocks(j), plan, (uint32_t *)b.meta, 8,
                                 stream),
                "launch wave-pool integrator", SKY_ERR_SYNC);
        shape.grid = plan.grid;
        shape.block = plan.block;
        shape.items = plan.n_items;
    }
    HIP_TRY(hipEventRecord(ctx->ev1, stream), "event record", SKY_ERR_SYNC);
    uint32_t wd = 0;
    HIP_TRY(hipMemcpyAsync(&wd, b.diag, sizeof(uint32_t), hipMemcpyDeviceToHost, stream), "read diagnostics",
            SKY_ERR_COPY);
    HIP_TRY(hipStreamSynchronize(stream), "synchronize", SKY_ERR_SYNC);
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
    if (stats) {
        stats->kernel_ms = ms;
        stats->path_samples = (uint64_t)s.width * s.height * s.spp * s.nlam;
        stats->watchdog = wd;
        stats->workgroups = shape.grid;
        stats->block = shape.block;
        stats->items = shape.items;
        stats->adjoint_value = 0.0;
    }
    if (wd) return fail(SKY_ERR_WATCHDOG, "GPU path watchdog was reached");  // integrator.rs:301-304
    return SKY_OK;
}

}  // namespace

sky_status sky::set_error(sky_status st, const std::string &msg) { return fail(st, msg); }

extern "C" {

const char *sky_last_error(void) { return g_last_error.c_str(); }

const char *sky_version(void) { return "sky_hip 1 (gfx950 wave-pool integrator)"; }

void sky_default_planet(sky_planet *p)
{
    p->ground_radius_km = 6360.0;
    p->atmosphere_radius_km = 6460.0;
}

void sky_default_atmosphere(sky_atmosphere *a)
{
    a->rayleigh_sca_550_m_inv = 1.3558e-5;
    a->rayleigh_h_km = 8.0;
    a->ozone_peak_m_inv = 1.9e-6;
    a->ozone_center_km = 25.0;
    a->ozone_half_width_km = 15.0;
    a->ozone_lambda0_nm = 600.0;
    a->ozone_sigma_nm = 80.0;
    a->aerosol_sca_550_m_inv = 3.996e-6;
    a->aerosol_abs_550_m_inv = 4.4e-6;
    a->aerosol_h_km = 1.2;
    a->aerosol_g = 0.8;
    a->angstrom = 1.3;
    a->ground_albedo = 0.3;
    a->sun_temperature_k = 5778.0;
    a->sun_angular_radius_rad = 0.2675 * 3.14159265358979323846 / 180.0;
}

// RenderConfig::default (config.rs:38-52) plus the extension defaults.
void sky_default_config(sky_render_config *c)
{
    std::memset(c, 0, sizeof(*c));
    c->width = 1024;
    c->height = 512;
    c->spp = 256;
    c->seed = 0x5EC720260430ull;
    c->sun_elevation_deg = 0.0;
    c->sun_azimuth_deg = 0.0;
    c->observer_altitude_km = 0.2;
    c->direct_light_samples = 1;
    c->projection = SKY_PROJECTION_PANORAMA;
    c->max_depth = 0;
    c->rr_start_depth = 4;
    c->watchdog_limit = 1024;
    c->max_segments = 100000;
    c->majorant_layers = 64;
    c->flags = 0;
    c->majorant_margin = 1.05;
    c->shell_growth = 1.0;
}

sky_status sky_context_create(int device, sky_context **out)
{
    if (!out) return fail(SKY_ERR_INVALID_CONFIG, "null context pointer");
    *out = nullptr;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0)
        return fail(SKY_ERR_NO_DEVICE, "no usable HIP device");  // RenderError::NoAdapter
    if (device < 0 || device >= n) return fail(SKY_ERR_NO_DEVICE, "device index out of range");
    sky_context *c = new sky_context();
    c->device = device;
    if ((e = hipSetDevice(device)) != hipSuccess ||
        (e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipEventCreate(&c->ev0)) != hipSuccess || (e = hipEventCreate(&c->ev1)) != hipSuccess) {
        delete c;
        return hip_fail(e, "device initialisation", SKY_ERR_DEVICE_INIT);
    }
    if ((e = scratch_address(device, &c->scratch)) != hipSuccess) {
        sky_context_destroy(c);
        return hip_fail(e, "scratch address", SKY_ERR_DEVICE_INIT);
    }
    int cus = 0, khz = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && cus > 0)
        c->num_cus = cus;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device) == hipSuccess && khz > 0)
        c->clock_khz = khz;
    *out = c;
    return SKY_OK;
}

void sky_context_destroy(sky_context *c)
{
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    for (DevBuf *b : {&c->lam, &c->maj, &c->shells, &c->b_shells, &c->film, &c->film_m2, &c->xyz, &c->accum, &c->meta, &c->diag,
                      &c->rays, &c->tau, &c->lamd, &c->jobs, &c->b_lam, &c->b_maj, &c->b_xyz, &c->b_film,
                      &c->b_accum, &c->b_meta, &c->b_jobs, &c->adj_units, &c->adj_dl, &c->adj_grad})
        b->release();
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

sky_status sky_context_set_tuning(sky_context *c, const sky_tuning *t)
{
    if (!c || !t) return fail(SKY_ERR_INVALID_CONFIG, "null context or tuning");
    if (t->event_threshold > 64 || t->pixels_per_item > 256 || t->waves_per_cu > 64 || t->track_steps > 4096 ||
        (t->occupancy != 0 && t->occupancy != 4 && t->occupancy != 5 && t->occupancy != 6 && t->occupancy != 8))
        return fail(SKY_ERR_INVALID_CONFIG, "tuning value out of range");
    c->tuning = *t;
    return SKY_OK;
}

sky_status sky_context_get_tuning(const sky_context *c, sky_tuning *t)
{
    if (!c || !t) return fail(SKY_ERR_INVALID_CONFIG, "null context or tuning");
    *t = c->tuning;
    return SKY_OK;
}

sky_status sky_wavelength_table(const sky_atmosphere *atm, const sky_wavelengths *wl, float *out,
                                uint32_t n_out)
{
    if (!atm || !wl || !out || !wl->center_nm || wl->count == 0 || wl->count > 64)
        return fail(SKY_ERR_INVALID_CONFIG, "invalid wavelength table request");
    if (n_out < 8 * wl->count) return fail(SKY_ERR_INVALID_CONFIG, "output too small");
    std::string err;
    if (atm->tables && !sky::validate_tables(atm, wl, err)) return fail(SKY_ERR_INVALID_CONFIG, err);
    std::vector<sky::DevLambda> rows;
    sky::fill_wavelength_rows(atm, wl, rows);
    std::memcpy(out, rows.data(), rows.size() * sizeof(sky::DevLambda));
    return SKY_OK;
}

sky_status sky_majorant_grid(const sky_planet *planet, const sky_atmosphere *atm,
                             const sky_wavelengths *wl, uint32_t layers, double margin,
                             double shell_growth, float *out, uint32_t n_out)
{
    const double g = shell_growth == 0.0 ? 1.0 : shell_growth;
    if (!planet || !atm || !wl || !out || !wl->center_nm || wl->count == 0 || wl->count > 64 ||
        layers == 0 || layers > 256 || !(g >= 1.0 && g <= 16.0) || layers * std::log2(g) > 1000.0)
        return fail(SKY_ERR_INVALID_CONFIG, "invalid majorant grid request");
    if (n_out < layers * wl->count) return fail(SKY_ERR_INVALID_CONFIG, "output too small");
    std::string err;
    if (atm->tables && !sky::validate_tables(atm, wl, err)) return fail(SKY_ERR_INVALID_CONFIG, err);
    std::vector<float> maj;
    sky::fill_majorants(planet, atm, wl, layers, margin, g, maj);
    std::memcpy(out, maj.data(), maj.size() * sizeof(float));
    return SKY_OK;
}

sky_status sky_validate_config(const sky_planet *planet, const sky_atmosphere *atm,
                               const sky_wavelengths *wl, const sky_render_config *cfg)
{
    sky::HostScene hs;
    std::string err;
    sky_status st = sky::build_scene(planet, atm, wl, cfg, hs, err);
    return st == SKY_OK ? SKY_OK : fail(st, err);
}

sky_status sky_render_device(sky_context *ctx, const sky_planet *planet, const sky_atmosphere *atm,
                             const sky_wavelengths *wl, const sky_render_config *cfg, float *d_film,
                             float *d_film_m2, float *d_xyz, void *stream, sky_render_stats *stats)
{
    auto t0 = std::chrono::steady_clock::now();
    // validate before touching the device (integrator.rs:126-127)
    sky::HostScene hs;
    std::string err;
    sky_status st = sky::build_scene(planet, atm, wl, cfg, hs, err);
    if (st != SKY_OK) return fail(st, err);
    if (!ctx) return fail(SKY_ERR_NO_DEVICE, "no device context");
    HIP_TRY(hipSetDevice(ctx->device), "set device", SKY_ERR_DEVICE_INIT);
    hs.dev.film = d_film;
    hs.dev.film_m2 = d_film_m2;
    hs.dev.xyz = d_xyz;
    hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
    st = run_render(ctx, hs, s, stats);
    if (stats)
        stats->total_ms =
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return st;
}

sky_status sky_render_device_noalloc(sky_context *ctx, const sky_planet *planet, const sky_atmosphere *atm,
                                     const sky_wavelengths *wl, const sky_render_config *cfg, float *d_film,
                                     float *d_film_m2, float *d_xyz, void *stream, sky_render_stats *stats)
{
    auto t0 = std::chrono::steady_clock::now();
    sky::HostScene hs;
    std::string err;
    sky_status st = sky::build_scene(planet, atm, wl, cfg, hs, err);
    if (st != SKY_OK) return fail(st, err);
    if (!ctx) return fail(SKY_ERR_NO_DEVICE, "no device context");
    if (!ctx->scratch) return fail(SKY_ERR_DEVICE_INIT, "context has no scratch address");
    HIP_TRY(hipSetDevice(ctx->device), "set device", SKY_ERR_DEVICE_INIT);
    hs.dev.film = d_film;
    hs.dev.film_m2 = d_film_m2;
    hs.dev.xyz = d_xyz;
    st = run_render(ctx, hs, stream ? (hipStream_t)stream : ctx->stream, stats, ctx->scratch);
    if (stats)
        stats->total_ms =
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return st;
}

sky_status sky_render(sky_context *ctx, const sky_planet *planet, const sky_atmosphere *atm,
                      const sky_wavelengths *wl, const sky_render_config *cfg, float *film,
                      float *film_m2, float *xyz, sky_render_stats *stats)
{
    auto t0 = std::chrono::steady_clock::now();
    // validate before touching the device (integrator.rs:126-127)
    sky::HostScene hs;
    std::string err;
    sky_status st = sky::build_scene(planet, atm, wl, cfg, hs, err);
    if (st != SKY_OK) return fail(st, err);
    if (!ctx) return fail(SKY_ERR_NO_DEVICE, "no device context");
    HIP_TRY(hipSetDevice(ctx->device), "set device", SKY_ERR_DEVICE_INIT);
    const size_t npx = (size_t)cfg->width * cfg->height;
    const size_t nf = npx * wl->count;
    if (film) HIP_TRY(ctx->film.ensure(nf * sizeof(float)), "alloc film", SKY_ERR_OOM);
    if (film_m2) HIP_TRY(ctx->film_m2.ensure(nf * sizeof(float)), "alloc film_m2", SKY_ERR_OOM);
    if (xyz) HIP_TRY(ctx->xyz.ensure(npx * 3 * sizeof(float)), "alloc xyz", SKY_ERR_OOM);
    hs.dev.film = film ? (float *)ctx->film.p : nullptr;
    hs.dev.film_m2 = film_m2 ? (float *)ctx->film_m2.p : nullptr;
    hs.dev.xyz = xyz ? (float *)ctx->xyz.p : nullptr;
    st = run_render(ctx, hs, ctx->stream, stats);
    if (st != SKY_OK && st != SKY_ERR_WATCHDOG) return st;
    sky_status st_copy = SKY_OK;
    auto copy = [&](void *dst, const void *src, size_t bytes) {
        if (!dst || st_copy != SKY_OK) return;
        hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream);
        if (e != hipSuccess) st_copy = hip_fail(e, "copy result to host", SKY_ERR_COPY);
    };
    copy(film, hs.dev.film, nf * sizeof(float));
    copy(film_m2, hs.dev.film_m2, nf * sizeof(float));
    copy(xyz, hs.dev.xyz, npx * 3 * sizeof(float));
    if (st_copy == SKY_OK) {
        hipError_t e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) st_copy = hip_fail(e, "synchronize", SKY_ERR_SYNC);
    }
    if (stats)
        stats->total_ms =
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (st_copy != SKY_OK) return st_copy;
    if (st == SKY_ERR_WATCHDOG) return fail(SKY_ERR_WATCHDOG, "GPU path watchdog was reached");
    return SKY_OK;
}

}  // extern "C"

namespace {

// Replay of the render's paths with the score bookkeeping (k_wavepool_adjoint).
sky_status adjoint_impl(sky_context *ctx, const sky_planet *planet, const sky_atmosphere *atm,
                        const sky_wavelengths *wl, const sky_render_config *cfg, const float *d_dl,
                        double *grad, hipStream_t stream, sky_render_stats *stats)
{
    sky::HostScene hs;
    std::string err;
    sky_status st = sky::build_scene(planet, atm, wl, cfg, hs, err);
    if (st != SKY_OK) return fail(st, err);
    sky::DevScene &s = hs.dev;
    HIP_TRY(ctx->lam.ensure(hs.lam.size() * sizeof(sky::DevLambda)), "alloc lambda table", SKY_ERR_OOM);
    HIP_TRY(ctx->maj.ensure(hs.maj.size() * sizeof(float)), "alloc majorants", SKY_ERR_OOM);
    HIP_TRY(ctx->shells.ensure(hs.shell_w.size() * sizeof(float)), "alloc shell table", SKY_ERR_OOM);
    HIP_TRY(ctx->diag.ensure(64), "alloc diagnostics", SKY_ERR_OOM);
    HIP_TRY(ctx->meta.ensure(64), "alloc work counters", SKY_ERR_OOM);
    HIP_TRY(ctx->jobs.ensure(sizeof(sky::DevJob)), "alloc job table", SKY_ERR_OOM);
    HIP_TRY(ctx->adj_units.ensure(wl->count * sizeof(float4)), "alloc adjoint table", SKY_ERR_OOM);
    HIP_TRY(ctx->adj_grad.ensure((SKY_NUM_PARAMS + 1) * sizeof(double)), "alloc gradient", SKY_ERR_OOM);
    std::vector<float4> units(wl->count);
    for (uint32_t i = 0; i < wl->count; ++i) {
        double l = wl->center_nm[i], r = 550.0 / l, dl = (l - atm->ozone_lambda0_nm) / atm->ozone_sigma_nm;
        units[i] = make_float4((float)(1000.0 * r * r * r * r), (float)(1000.0 * std::exp(-0.5 * dl * dl)),
                               (float)(1000.0 * std::pow(l / 550.0, -atm->angstrom)), (float)std::log(l / 550.0));
    }
    HIP_TRY(hipMemcpyAsync(ctx->lam.p, hs.lam.data(), hs.lam.size() * sizeof(sky::DevLambda),
                           hipMemcpyHostToDevice, stream), "upload lambda table", SKY_ERR_COPY);
    HIP_TRY(hipMemcpyAsync(ctx->maj.p, hs.maj.data(), hs.maj.size() * sizeof(float), hipMemcpyHostToDevice,
                           stream), "upload majorants", SKY_ERR_COPY);
    HIP_TRY(hipMemcpyAsync(ctx->shells.p, hs.shell_w.data(), hs.shell_w.size() * sizeof(float),
                           hipMemcpyHostToDevice, stream), "upload shell table", SKY_ERR_COPY);
    HIP_TRY(hipMemcpyAsync(ctx->adj_units.p, units.data(), units.size() * sizeof(float4), hipMemcpyHostToDevice,
                           stream), "upload adjoint table", SKY_ERR_COPY);
    HIP_TRY(sky::launch_zero((uint32_t *)ctx->diag.p, 16, stream), "clear diagnostics", SKY_ERR_SYNC);
    HIP_TRY(sky::launch_zero((uint32_t *)ctx->adj_grad.p, 2 * (SKY_NUM_PARAMS + 1), stream), "clear gradient",
            SKY_ERR_SYNC);
    s.lam = (const sky::DevLambda *)ctx->lam.p;
    s.maj = (const float *)ctx->maj.p;
    s.shell_w = (const float *)ctx->shells.p;
    s.diag = (uint32_t *)ctx->diag.p;
    s.film = s.film_m2 = s.xyz = nullptr;
    ctx->host_jobs.assign(1, sky::DevJob{});
    sky::DevJob &j = ctx->host_jobs[0];
    j.s = s;
    sky::plan_job(s, ctx->tuning, ctx->num_cus, sky::launch_min_items(ctx->tuning, ctx->num_cus), j);
    j.item_base = 0;
    j.accum = nullptr;  // the replay sums only the gradient
    j.done = (uint32_t *)ctx->meta.p + 4;
    j.t_done = (unsigned long long *)((uint32_t *)ctx->meta.p + 6);
    HIP_TRY(hipMemcpyAsync(ctx->jobs.p, &j, sizeof(j), hipMemcpyHostToDevice, stream), "upload job", SKY_ERR_COPY);
    sky::WavePlan plan = sky::plan_launch(ctx->tuning, ctx->num_cus, j.n_items, s.nlam * s.L, s.L, s.flags);
    sky::AdjArgs A;
    A.dl_dxyz = d_dl;
    A.lam_units = (const float4 *)ctx->adj_units.p;
    A.inv_hr2 = (float)(1.0 / (atm->rayleigh_h_km * atm->rayleigh_h_km));
    A.inv_ha2 = (float)(1.0 / (atm->aerosol_h_km * atm->aerosol_h_km));
    A.inv_spp = 1.0f / (float)s.spp;
    A.grad = (double *)ctx->adj_grad.p;
    HIP_TRY(hipEventRecord(ctx->ev0, stream), "event record", SKY_ERR_SYNC);
    HIP_TRY(sky::launch_adjoint((const sky::DevJob *)ctx->jobs.p, plan, (uint32_t *)ctx->meta.p, 8, A, stream),
            "launch adjoint", SKY_ERR_SYNC);
    HIP_TRY(hipEventRecord(ctx->ev1, stream), "event record", SKY_ERR_SYNC);
    uint32_t wd = 0;
    double host_grad[SKY_NUM_PARAMS + 1];
    HIP_TRY(hipMemcpyAsync(host_grad, ctx->adj_grad.p, (SKY_NUM_PARAMS + 1) * sizeof(double), hipMemcpyDeviceToHost,
                           stream),
            "copy gradient", SKY_ERR_COPY);
    HIP_TRY(hipMemcpyAsync(&wd, ctx->diag.p, sizeof(uint32_t), hipMemcpyDeviceToHost, stream), "read diagnostics",
            SKY_ERR_COPY);
    HIP_TRY(hipStreamSynchronize(stream), "synchronize", SKY_ERR_SYNC);
    std::memcpy(grad, host_grad, SKY_NUM_PARAMS * sizeof(double));
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
    if (stats) {
        stats->kernel_ms = ms;
        stats->path_samples = (uint64_t)s.width * s.height * s.spp * s.nlam;
        stats->watchdog = wd;
        stats->workgroups = plan.grid;
        stats->block = plan.block;
        stats->items = plan.n_items;
        stats->adjoint_value = host_grad[SKY_NUM_PARAMS];
    }
    if (wd) return fail(SKY_ERR_WATCHDOG, "GPU path watchdog was reached");
    return SKY_OK;
}

}  // namespace

extern "C" {

sky_status sky_render_adjoint(sky_context *ctx, const sky_planet *planet, const sky_atmosphere *atm,
                              const sky_wavelengths *wl, const sky_render_config *cfg, const float *dl_dxyz,
                              double *grad, sky_render_stats *stats)
{
    auto t0 = std::chrono::steady_clock::now();
    sky_status st = sky_validate_config(planet, atm, wl, cfg);
    if (st != SKY_OK) return st;
    if (atm->tables) return fail(SKY_ERR_INVALID_CONFIG, "the adjoint differentiates the analytic medium only");
    if (!dl_dxyz || !grad) return fail(SKY_ERR_INVALID_CONFIG, "null cotangent or gradient buffer");
    if (!ctx) return fail(SKY_ERR_NO_DEVICE, "no device context");
    HIP_TRY(hipSetDevice(ctx->device), "set device", SKY_ERR_DEVICE_INIT);
    size_t bytes = (size_t)cfg->width * cfg->height * 3 * sizeof(float);
    HIP_TRY(ctx->adj_dl.ensure(bytes), "alloc cotangent", SKY_ERR_OOM);
    HIP_TRY(hipMemcpyAsync(ctx->adj_dl.p, dl_dxyz, bytes, hipMemcpyHostToDevice, ctx->stream), "upload cotangent",
            SKY_ERR_COPY);
    st = adjoint_impl(ctx, planet, atm, wl, cfg, (const float *)ctx->adj_dl.p, grad, ctx->stream, stats);
    if (stats)
        stats->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return st;
}

sky_status sky_render_adjoint_device(sky_context *ctx, const sky_planet *planet, const sky_atmosphere *atm,
                                     const sky_wavelengths *wl, const sky_render_config *cfg,
                                     const float *d_dl_dxyz, double *grad, void *stream, sky_render_stats *stats)
{
    auto t0 = std::chrono::steady_clock::now();
    sky_status st = sky_validate_config(planet, atm, wl, cfg);
    if (st != SKY_OK) return st;
    if (atm->tables) return fail(SKY_ERR_INVALID_CONFIG, "the adjoint differentiates the analytic medium only");
    if (!d_dl_dxyz || !grad) return fail(SKY_ERR_INVALID_CONFIG, "null cotangent or gradient buffer");
    if (!ctx) return fail(SKY_ERR_NO_DEVICE, "no device context");
    HIP_TRY(hipSetDevice(ctx->device), "set device", SKY_ERR_DEVICE_INIT);
    st = adjoint_impl(ctx, planet, atm, wl, cfg, d_dl_dxyz, grad, stream ? (hipStream_t)stream : ctx->stream,
                      stats);
    if (stats)
        stats->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return st;
}

sky_status sky_render_batch(sky_context *ctx, const sky_batch_item *items, uint32_t n, uint32_t flags,
                            double *finish_ms, sky_batch_stats *stats)
{
    auto t0 = std::chrono::steady_clock::now();
    if (!items || n == 0) return fail(SKY_ERR_INVALID_CONFIG, "empty batch");
    std::vector<sky::HostScene> scenes(n);
    for (uint32_t i = 0; i < n; ++i) {
        std::string err;
        sky_status st = sky::build_scene(items[i].planet, items[i].atm, items[i].wl, items[i].cfg, scenes[i], err);
        if (st != SKY_OK) return fail(st, "batch item " + std::to_string(i) + ": " + err);
        if (scenes[i].dev.flags & SKY_FLAG_SIMPLE_KERNEL)
            return fail(SKY_ERR_INVALID_CONFIG, "batch item " + std::to_string(i) + ": batches use the wave-pool kernel");
    }
    if (!ctx) return fail(SKY_ERR_NO_DEVICE, "no device context");
    HIP_TRY(hipSetDevice(ctx->device), "set device", SKY_ERR_DEVICE_INIT);
    hipStream_t stream = ctx->stream;

    // Largest renders first (longest-processing-time order), so the launch
    // ends on small items.
    std::vector<uint32_t> order(n);
    for (uint32_t i = 0; i < n; ++i) order[i] = i;
    auto cost = [&](uint32_t i) {
        const sky::DevScene &d = scenes[i].dev;
        return (double)d.width * d.height * d.nlam * d.spp;
    };
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return cost(a) > cost(b); });

    // Table offsets.
    size_t n_lam = 0, n_maj = 0, n_shl = 0, n_xyz = 0, n_film = 0, n_acc = 0, n_tab = 0;
    uint32_t max_npx = 0, total_items = 0, max_maj = 0, max_layers = 0, flags_all = 0, fin_blocks = 0;
    std::vector<size_t> o_lam(n), o_maj(n), o_shl(n), o_xyz(n), o_film(n), o_acc(n), o_tab(n);
    ctx->host_jobs.assign(n, sky::DevJob{});
    const uint32_t min_items = 0;  // the batch balances across renders; no sample split needed
    uint64_t samples = 0;
    for (uint32_t k = 0; k < n; ++k) {
        uint32_t i = order[k];
        sky::HostScene &h = scenes[i];
        uint32_t npx = h.dev.width * h.dev.height;
        o_lam[i] = n_lam;
        o_maj[i] = n_maj;
        o_shl[i] = n_shl;
        o_xyz[i] = n_xyz;
        o_film[i] = n_film;
        o_tab[i] = n_tab;
        if (!h.tab.empty()) n_tab += h.tab.bytes();
        n_lam += h.lam.size();
        n_maj += h.maj.size();
        n_shl += h.shell_w.size();
        n_xyz += (size_t)npx * 3;
        if (items[i].film) n_film += (size_t)npx * h.nlam;
        sky::DevJob &j = ctx->host_jobs[k];
        j.s = h.dev;
        sky::plan_job(h.dev, ctx->tuning, ctx->num_cus, min_items, j);
        j.item_base = total_items;
        total_items += j.n_items;
        o_acc[i] = n_acc;
        n_acc += sky::job_accum_bytes(j) / sizeof(float2);
        max_npx = npx > max_npx ? npx : max_npx;
        fin_blocks = std::max(fin_blocks, sky::finalize_blocks(j));
        max_maj = h.maj.size() > max_maj ? (uint32_t)h.maj.size() : max_maj;
        max_layers = h.dev.L > max_layers ? h.dev.L : max_layers;
        flags_all = k == 0 || h.dev.flags == flags_all ? h.dev.flags : sky::kFlagMixed;
        samples += (uint64_t)npx * h.dev.spp * h.dev.nlam;
    }
    const uint32_t meta_words = 8 + 4 * n;
    HIP_TRY(ctx->b_lam.ensure(n_lam * sizeof(sky::DevLambda)), "alloc batch lambda tables", SKY_ERR_OOM);
    HIP_TRY(ctx->b_maj.ensure(n_maj * sizeof(float)), "alloc batch majorants", SKY_ERR_OOM);
    HIP_TRY(ctx->b_shells.ensure(n_shl * sizeof(float)), "alloc batch shell tables", SKY_ERR_OOM);
    HIP_TRY(ctx->b_xyz.ensure(n_xyz * sizeof(float)), "alloc batch xyz", SKY_ERR_OOM);
    if (n_film) HIP_TRY(ctx->b_film.ensure(n_film * sizeof(float)), "alloc batch film", SKY_ERR_OOM);
    HIP_TRY(ctx->b_accum.ensure(n_acc * sizeof(float2)), "alloc batch partial sums", SKY_ERR_OOM);
    HIP_TRY(ctx->b_meta.ensure(meta_words * sizeof(uint32_t)), "alloc batch counters", SKY_ERR_OOM);
    HIP_TRY(ctx->b_jobs.ensure(n * sizeof(sky::DevJob)), "alloc batch job table", SKY_ERR_OOM);
    if (n_tab) HIP_TRY(ctx->b_tab.ensure(n_tab), "alloc batch medium tables", SKY_ERR_OOM);
    ctx->host_btab.assign(n_tab, 0);
    HIP_TRY(ctx->diag.ensure(64), "alloc diagnostics", SKY_ERR_OOM);
    std::vector<sky::DevLambda> all_lam(n_lam);
    std::vector<float> all_maj(n_maj);
    std::vector<float> all_shl(n_shl);
    uint32_t *meta = (uint32_t *)ctx->b_meta.p;
    for (uint32_t k = 0; k < n; ++k) {
        uint32_t i = order[k];
        sky::HostScene &h = scenes[i];
        std::copy(h.lam.begin(), h.lam.end(), all_lam.begin() + o_lam[i]);
        std::copy(h.maj.begin(), h.maj.end(), all_maj.begin() + o_maj[i]);
        std::copy(h.shell_w.begin(), h.shell_w.end(), all_shl.begin() + o_shl[i]);
        sky::DevJob &j = ctx->host_jobs[k];
        j.s.lam = (const sky::DevLambda *)ctx->b_lam.p + o_lam[i];
        j.s.maj = (const float *)ctx->b_maj.p + o_maj[i];
        j.s.shell_w = (const float *)ctx->b_shells.p + o_shl[i];
        if (!h.tab.empty()) {
            std::vector<uint8_t> blob;
            h.tab.pack(blob);
            std::copy(blob.begin(), blob.end(), ctx->host_btab.begin() + o_tab[i]);
        }
        h.tab.attach(j.s, (const uint8_t *)ctx->b_tab.p + o_tab[i]);
        j.s.xyz = (float *)ctx->b_xyz.p + o_xyz[i];
        j.s.film = items[i].film ? (float *)ctx->b_film.p + o_film[i] : nullptr;
        j.s.film_m2 = nullptr;
        j.s.diag = (uint32_t *)ctx->diag.p;
        j.accum = (float2 *)ctx->b_accum.p + o_acc[i];
        j.done = meta + 8 + 2 * k;
        j.t_done = (unsigned long long *)(meta + 8 + 2 * n + 2 * k);
    }
    HIP_TRY(hipMemcpyAsync(ctx->b_lam.p, all_lam.data(), n_lam * sizeof(sky::DevLambda), hipMemcpyHostToDevice,
                           stream), "upload batch lambda tables", SKY_ERR_COPY);
    HIP_TRY(hipMemcpyAsync(ctx->b_maj.p, all_maj.data(), n_maj * sizeof(float), hipMemcpyHostToDevice, stream),
            "upload batch majorants", SKY_ERR_COPY);
    HIP_TRY(hipMemcpyAsync(ctx->b_shells.p, all_shl.data(), n_shl * sizeof(float), hipMemcpyHostToDevice, stream),
            "upload batch shell tables", SKY_ERR_COPY);
    HIP_TRY(hipMemcpyAsync(ctx->b_jobs.p, ctx->host_jobs.data(), n * sizeof(sky::DevJob), hipMemcpyHostToDevice,
                           stream), "upload batch job table", SKY_ERR_COPY);
    if (n_tab)
        HIP_TRY(hipMemcpyAsync(ctx->b_tab.p, ctx->host_btab.data(), n_tab, hipMemcpyHostToDevice, stream),
                "upload batch medium tables", SKY_ERR_COPY);
    HIP_TRY(sky::launch_zero((uint32_t *)ctx->diag.p, 16, stream), "clear diagnostics", SKY_ERR_SYNC);
    sky::WavePlan plan = sky::plan_launch(ctx->tuning, ctx->num_cus, total_items, max_maj, max_layers, flags_all);

    // The graph holds only kernel nodes (zero, integrate, finalize); the
    // tables it reads are uploaded above on every call.
    const bool use_graph = !(flags & SKY_BATCH_NO_GRAPH);
    uint64_t key = 1469598103934665603ull;
    auto mix = [&](uint64_t v) { key = (key ^ v) * 1099511628211ull; };
    mix(n);
    mix(total_items);
    mix(max_npx);
    mix(fin_blocks);
    mix((uint64_t)(uintptr_t)ctx->b_jobs.p);
    mix((uint64_t)(uintptr_t)ctx->b_meta.p);
    mix(plan.grid);
    mix(plan.event_threshold);
    mix(plan.track_steps);
    mix(plan.occupancy);
    mix(plan.lds_bytes);
    mix(plan.fast);
    bool reused = false;
    HIP_TRY(hipEventRecord(ctx->ev0, stream), "event record", SKY_ERR_SYNC);
    if (use_graph) {
        if (!ctx->graph_exec || ctx->graph_key != key) {
            if (ctx->graph_exec) (void)hipGraphExecDestroy(ctx->graph_exec);
            if (ctx->graph) (void)hipGraphDestroy(ctx->graph);
            ctx->graph_exec = nullptr;
            ctx->graph = nullptr;
            HIP_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "begin capture", SKY_ERR_SYNC);
            hipError_t le = sky::launch_pool((const sky::DevJob *)ctx->b_jobs.p, n, fin_blocks, plan, meta,
                                             meta_words, stream);
            hipError_t ce = hipStreamEndCapture(stream, &ctx->graph);
            if (le != hipSuccess) return hip_fail(le, "capture batch launch", SKY_ERR_SYNC);
            if (ce != hipSuccess) return hip_fail(ce, "end capture", SKY_ERR_SYNC);
            HIP_TRY(hipGraphInstantiate(&ctx->graph_exec, ctx->graph, nullptr, nullptr, 0), "instantiate graph",
                    SKY_ERR_SYNC);
            ctx->graph_key = key;
        } else {
            reused = true;
        }
        HIP_TRY(hipGraphLaunch(ctx->graph_exec, stream), "launch graph", SKY_ERR_SYNC);
    } else {
        HIP_TRY(sky::launch_pool((const sky::DevJob *)ctx->b_jobs.p, n, fin_blocks, plan, meta, meta_words, stream),
                "launch batch", SKY_ERR_SYNC);
    }
    HIP_TRY(hipEventRecord(ctx->ev1, stream), "event record", SKY_ERR_SYNC);
    std::vector<uint32_t> host_meta(meta_words);
    uint32_t wd = 0;
    HIP_TRY(hipMemcpyAsync(host_meta.data(), meta, meta_words * sizeof(uint32_t), hipMemcpyDeviceToHost, stream),
            "read batch counters", SKY_ERR_COPY);
    HIP_TRY(hipMemcpyAsync(&wd, ctx->diag.p, sizeof(uint32_t), hipMemcpyDeviceToHost, stream), "read diagnostics",
            SKY_ERR_COPY);
    for (uint32_t i = 0; i < n; ++i) {
        const sky::HostScene &h = scenes[i];
        size_t npx = (size_t)h.dev.width * h.dev.height;
        if (items[i].xyz)
            HIP_TRY(hipMemcpyAsync(items[i].xyz, (float *)ctx->b_xyz.p + o_xyz[i], npx * 3 * sizeof(float),
                                   hipMemcpyDeviceToHost, stream), "copy batch xyz", SKY_ERR_COPY);
        if (items[i].film)
            HIP_TRY(hipMemcpyAsync(items[i].film, (float *)ctx->b_film.p + o_film[i], npx * h.nlam * sizeof(float),
                                   hipMemcpyDeviceToHost, stream), "copy batch film", SKY_ERR_COPY);
    }
    HIP_TRY(hipStreamSynchronize(stream), "synchronize", SKY_ERR_SYNC);
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
    // per-render completion times from the device clock
    unsigned long long t_start = (unsigned long long)host_meta[2] | ((unsigned long long)host_meta[3] << 32);
    std::vector<double> lat(n);
    for (uint32_t k = 0; k < n; ++k) {
        uint32_t i = order[k];
        size_t w = 8 + 2 * n + 2 * k;
        unsigned long long td = (unsigned long long)host_meta[w] | ((unsigned long long)host_meta[w + 1] << 32);
        lat[i] = td >= t_start ? (double)(td - t_start) / ctx->clock_khz : 0.0;
        if (finish_ms) finish_ms[i] = lat[i];
    }
    if (stats) {
        std::vector<double> sorted = lat;
        std::sort(sorted.begin(), sorted.end());
        auto pct = [&](double q) {
            size_t idx = (size_t)std::ceil(q * sorted.size()) - 1;
            return sorted[idx < sorted.size() ? idx : sorted.size() - 1];
        };
        stats->kernel_ms = ms;
        stats->path_samples = samples;
        stats->renders = n;
        stats->items = total_items;
        stats->watchdog = wd;
        stats->graph_reused = reused;
        stats->latency_p50_ms = pct(0.50);
        stats->latency_p99_ms = pct(0.99);
        stats->latency_max_ms = sorted.back();
        stats->total_ms =
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    if (wd) return fail(SKY_ERR_WATCHDOG, "GPU path watchdog was reached");
    return SKY_OK;
}

sky_status sky_optical_depth(sky_context *ctx, const sky_planet *planet, const sky_atmosphere *atm,
                             const sky_wavelengths *wl, const double *rays, uint32_t n, double *tau)
{
    if (!ctx) return fail(SKY_ERR_NO_DEVICE, "null context");
    if (!rays || !tau) return fail(SKY_ERR_INVALID_CONFIG, "null ray or output buffer");
    sky_render_config cfg;
    sky_default_config(&cfg);
    cfg.width = cfg.height = cfg.spp = 1;
    sky::HostScene hs;
    std::string err;
    sky_status st = sky::build_scene(planet, atm, wl, &cfg, hs, err);
    if (st != SKY_OK) return fail(st, err);
    if (n == 0) return SKY_OK;
    for (uint32_t i = 0; i < n; ++i) {
        const double *r = rays + 7 * (size_t)i;
        bool ok = true;
        for (int j = 0; j < 7; ++j) ok = ok && std::isfinite(r[j]);
        double dn = r[3] * r[3] + r[4] * r[4] + r[5] * r[5];
        if (!ok || !(dn > 0.0) || !(r[6] >= 0.0 && r[6] <= 1.0e5))
            return fail(SKY_ERR_INVALID_CONFIG, "rays need finite origins, nonzero directions and 0 <= t_max <= 1e5 km");
    }
    HIP_TRY(hipSetDevice(ctx->device), "set device", SKY_ERR_DEVICE_INIT);
    std::vector<double> lamd(4 * (size_t)wl->count + 4 + 16);
    for (uint32_t i = 0; i < wl->count; ++i) {
        // FP64 coefficients for the quadrature (same formulas as the scene build)
        double l = wl->center_nm[i];
        double r = 550.0 / l, dl = (l - atm->ozone_lambda0_nm) / atm->ozone_sigma_nm;
        double ang = std::pow(l / 550.0, -atm->angstrom);
        lamd[4 * i + 0] = atm->rayleigh_sca_550_m_inv * 1000.0 * r * r * r * r;
        lamd[4 * i + 1] = atm->ozone_peak_m_inv * 1000.0 * std::exp(-0.5 * dl * dl);
        lamd[4 * i + 2] = atm->aerosol_sca_550_m_inv * 1000.0 * ang;
        lamd[4 * i + 3] = atm->aerosol_abs_550_m_inv * 1000.0 * ang;
    }
    lamd[4 * wl->count + 0] = atm->rayleigh_h_km;
    lamd[4 * wl->count + 1] = atm->aerosol_h_km;
    lamd[4 * wl->count + 2] = atm->ozone_center_km;
    lamd[4 * wl->count + 3] = atm->ozone_half_width_km;
    // 16-point Gauss-Legendre rule by Newton iteration on P_16
    for (int i = 0; i < 8; ++i) {
        double x = std::cos(3.14159265358979323846 * (i + 0.75) / 16.5), dp = 0.0;
        for (int it = 0; it < 100; ++it) {
            double p0 = 1.0, p1 = x;
            for (int k = 2; k <= 16; ++k) {
                double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            dp = 16.0 * (x * p1 - p0) / (x * x - 1.0);
            double dx = p1 / dp;
            x -= dx;
            if (std::fabs(dx) < 1e-16) break;
        }
        lamd[4 * wl->count + 4 + (7 - i)] = x;
        lamd[4 * wl->count + 12 + (7 - i)] = 2.0 / ((1.0 - x * x) * dp * dp);
    }
    size_t rb = (size_t)n * 7 * sizeof(double), tb = (size_t)n * wl->count * sizeof(double);
    HIP_TRY(ctx->rays.ensure(rb), "alloc rays", SKY_ERR_OOM);
    HIP_TRY(ctx->tau.ensure(tb), "alloc tau", SKY_ERR_OOM);
    HIP_TRY(ctx->lamd.ensure(lamd.size() * sizeof(double)), "alloc coefficients", SKY_ERR_OOM);
    HIP_TRY(hipMemcpyAsync(ctx->rays.p, rays, rb, hipMemcpyHostToDevice, ctx->stream), "upload rays", SKY_ERR_COPY);
    HIP_TRY(hipMemcpyAsync(ctx->lamd.p, lamd.data(), lamd.size() * sizeof(double), hipMemcpyHostToDevice,
                           ctx->stream),
            "upload coefficients", SKY_ERR_COPY);
    if (!hs.tab.empty()) {
        hs.tab.pack(ctx->host_tab);
        HIP_TRY(ctx->tab.ensure(ctx->host_tab.size()), "alloc medium table", SKY_ERR_OOM);
        HIP_TRY(hipMemcpyAsync(ctx->tab.p, ctx->host_tab.data(), ctx->host_tab.size(), hipMemcpyHostToDevice,
                               ctx->stream),
                "upload medium table", SKY_ERR_COPY);
    }
    hs.tab.attach(hs.dev, ctx->tab.p);
    HIP_TRY(sky::launch_optical_depth(hs.dev, (const double *)ctx->rays.p, n, (const double *)ctx->lamd.p,
                                      (double *)ctx->tau.p, ctx->stream),
            "launch optical depth", SKY_ERR_SYNC);
    HIP_TRY(hipMemcpyAsync(tau, ctx->tau.p, tb, hipMemcpyDeviceToHost, ctx->stream), "copy tau", SKY_ERR_COPY);
    HIP_TRY(hipStreamSynchronize(ctx->stream), "synchronize", SKY_ERR_SYNC);
    return SKY_OK;
}

}  // extern "C"
