// sky.cpp -- C++ host mirror of the reference integrator (see sky.hpp).
#include "sky.hpp"

#include <mutex>

namespace sky {

namespace {

RenderError to_error(sky_status st)
{
    using K = RenderError::Kind;
    K k = K::Poll;
    switch (st) {
    case SKY_ERR_INVALID_CONFIG: k = K::InvalidConfig; break;
    case SKY_ERR_NO_DEVICE: k = K::NoAdapter; break;
    case SKY_ERR_DEVICE_INIT: k = K::RequestDevice; break;
    case SKY_ERR_COPY: k = K::BufferMap; break;
    case SKY_ERR_WATCHDOG: k = K::Watchdog; break;
    case SKY_ERR_OOM: k = K::OutOfMemory; break;
    default: break;
    }
    return RenderError(k, sky_last_error());
}

// One context per process, created on first use (the reference creates a
// device per call, integrator.rs:171-190; reusing it keeps buffers warm).
sky_context *shared_context()
{
    static std::once_flag once;
    static sky_context *ctx = nullptr;
    static sky_status st = SKY_OK;
    std::call_once(once, [] { st = sky_context_create(0, &ctx); });
    if (st != SKY_OK) throw to_error(st);
    return ctx;
}

}  // namespace

SceneData load_scene_data(float sun_elevation_deg, float sun_azimuth_deg, uint32_t band_count)
{
    SceneData s;
    sky_default_planet(&s.planet);
    sky_default_atmosphere(&s.atmosphere);
    s.sun_elevation_deg = sun_elevation_deg;
    s.sun_azimuth_deg = sun_azimuth_deg;
    if (band_count == 0) {
        // the reference's fixed grid: 41 bands at 380, 390, ..., 780 nm, 10 nm wide (spectrum.rs:5, 166-175)
        for (uint32_t i = 0; i < 41; ++i) {
            s.wavelengths_nm.push_back(380.0 + 10.0 * i);
            s.weights_nm.push_back(10.0);
        }
        return s;
    }
    double step = 400.0 / band_count;
    for (uint32_t i = 0; i < band_count; ++i) s.wavelengths_nm.push_back(380.0 + step * (i + 0.5));
    return s;
}

Film render(const SceneData &scene, const RenderConfig &config)
{
    sky_render_config c;
    sky_default_config(&c);
    c.width = config.width;
    c.height = config.height;
    c.spp = config.spp;
    c.seed = config.seed;
    c.sun_elevation_deg = scene.sun_elevation_deg;  // the sun comes from the scene (integrator.rs:343-345)
    c.sun_azimuth_deg = scene.sun_azimuth_deg;
    c.observer_altitude_km = config.observer_altitude_km;
    c.direct_light_samples = config.direct_light_samples;
    c.projection = (uint32_t)config.output_projection;
    c.sample_offset = config.sample_offset;
    c.max_depth = config.max_depth;
    c.rr_start_depth = config.rr_start_depth;
    c.majorant_layers = config.majorant_layers;
    c.shell_growth = config.shell_growth;
    c.flags = config.flags;
    sky_wavelengths wl;
    wl.count = (uint32_t)scene.wavelengths_nm.size();
    wl.center_nm = scene.wavelengths_nm.data();
    wl.weight_nm = scene.weights_nm.empty() ? nullptr : scene.weights_nm.data();
    sky_status st = sky_validate_config(&scene.planet, &scene.atmosphere, &wl, &c);
    if (st != SKY_OK) throw to_error(st);
    size_t npx = (size_t)c.width * c.height;
    std::vector<float> film(npx * wl.count), xyz(npx * 3);
    sky_render_stats stats{};
    st = sky_render(shared_context(), &scene.planet, &scene.atmosphere, &wl, &c, film.data(), nullptr, xyz.data(),
                    &stats);
    if (st != SKY_OK) throw to_error(st);
    return Film(c.width, c.height, wl.count, std::move(film), std::move(xyz));
}

}  // namespace sky
