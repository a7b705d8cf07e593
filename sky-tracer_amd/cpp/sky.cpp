// sky.cpp -- C++ host mirror of the reference integrator (see sky.hpp).
#include "sky.hpp"

#include <algorithm>
#include <fstream>
#include <mutex>
#include <sstream>

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

namespace {

std::vector<std::vector<std::string>> csv_rows(const std::string &path)  // data.rs:154-167
{
    std::ifstream f(path);
    if (!f) throw std::runtime_error("cannot open " + path);
    std::vector<std::vector<std::string>> rows;
    std::string line;
    auto trim = [](std::string v) {
        size_t a = v.find_first_not_of(" \t\r\n"), b = v.find_last_not_of(" \t\r\n");
        return a == std::string::npos ? std::string() : v.substr(a, b - a + 1);
    };
    while (std::getline(f, line)) {
        line = trim(line);
        if (line.empty() || line[0] == '#') continue;
        std::vector<std::string> cols;
        std::stringstream ss(line);
        std::string c;
        while (std::getline(ss, c, ',')) cols.push_back(trim(c));
        rows.push_back(cols);
    }
    return rows;
}

double num(const std::string &v, const std::string &path)
{
    char *end = nullptr;
    double d = std::strtod(v.c_str(), &end);
    if (v.empty() || *end) throw std::runtime_error(path + ": not a number: " + v);
    return d;
}

}  // namespace

SceneData load_scene_data(const std::string &dir, float sun_elevation_deg, float sun_azimuth_deg)
{
    SceneData s;
    sky_default_planet(&s.planet);
    s.planet.atmosphere_radius_km = 6480.0;  // geometry.rs:15
    sky_default_atmosphere(&s.atmosphere);
    s.atmosphere.ground_albedo = 0.18;               // atmosphere.rs:9
    s.atmosphere.sun_angular_radius_rad = 0.00465047;  // atmosphere.rs:10
    s.sun_elevation_deg = sun_elevation_deg;
    s.sun_azimuth_deg = sun_azimuth_deg;
    auto m = std::make_shared<MediumData>();
    const std::string pb = dir + "/bands.csv";
    for (const auto &r : csv_rows(pb)) {  // data.rs:53-73
        if (r.size() != 6) throw std::runtime_error(pb + ": expected 6 columns");
        s.wavelengths_nm.push_back(num(r[1], pb));
        s.weights_nm.push_back(1.0);  // band-integrated values: cmf weights alone (spectrum.rs:226-241)
        m->solar_irradiance_w_m2.push_back(num(r[4], pb));
        m->ozone_cross_section_cm2.push_back(num(r[5], pb));
    }
    if (s.wavelengths_nm.size() != 41)
        throw std::runtime_error("expected 41 bands, found " + std::to_string(s.wavelengths_nm.size()));
    auto profile = [&](const std::string &name, size_t cols, std::vector<double> &out, std::vector<size_t> keep) {
        const std::string p = dir + "/" + name;
        std::vector<std::vector<double>> rows;
        for (const auto &r : csv_rows(p)) {
            if (r.size() != cols) throw std::runtime_error(p + ": expected " + std::to_string(cols) + " columns");
            std::vector<double> v;
            for (size_t c : keep) v.push_back(num(r[c], p));
            rows.push_back(v);
        }
        std::stable_sort(rows.begin(), rows.end(), [](const auto &a, const auto &b) { return a[0] < b[0]; });
        for (const auto &v : rows) out.insert(out.end(), v.begin(), v.end());
    };
    profile("atmosphere_profile.csv", 4, m->atmosphere, {0, 2, 3});  // data.rs:75-90
    profile("aerosol_profile.csv", 5, m->aerosol, {0, 1, 2, 3, 4});   // data.rs:92-112
    const char *species[4] = {"inso", "waso", "soot", "suso"};       // atmosphere.rs:8
    m->aerosol_optics.assign(41 * 4 * 2, 0.0);
    std::vector<bool> seen(41 * 4, false);
    const std::string po = dir + "/aerosol_optics.csv";
    for (const auto &r : csv_rows(po)) {  // data.rs:114-134
        if (r.size() != 4) throw std::runtime_error(po + ": expected 4 columns");
        int sp = -1;
        for (int k = 0; k < 4; ++k)
            if (r[0] == species[k]) sp = k;
        const double b = num(r[1], po);
        if (sp < 0 || b < 0 || b >= 41 || b != (double)(size_t)b)
            throw std::runtime_error(po + ": bad species or band " + r[0] + "," + r[1]);
        const size_t i = (size_t)b * 4 + sp;
        m->aerosol_optics[2 * i] = num(r[2], po);
        m->aerosol_optics[2 * i + 1] = num(r[3], po);
        seen[i] = true;
    }
    for (bool v : seen)
        if (!v) throw std::runtime_error(po + ": missing entries");
    s.medium = m;
    return s;
}

namespace {
sky_tuning g_tuning{};  // zero fields: the library's defaults
}

void set_launch_tuning(const sky_tuning &tuning) { g_tuning = tuning; }

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
    sky_medium_tables tables{};
    const sky_atmosphere atm = scene.atmosphere_c(tables);
    sky_status st = sky_validate_config(&scene.planet, &atm, &wl, &c);
    if (st != SKY_OK) throw to_error(st);
    size_t npx = (size_t)c.width * c.height;
    std::vector<float> film(npx * wl.count), xyz(npx * 3);
    sky_render_stats stats{};
    sky_context *ctx = shared_context();
    st = sky_context_set_tuning(ctx, &g_tuning);
    if (st != SKY_OK) throw to_error(st);
    st = sky_render(ctx, &scene.planet, &atm, &wl, &c, film.data(), nullptr, xyz.data(), &stats);
    if (st != SKY_OK) throw to_error(st);
    return Film(c.width, c.height, wl.count, std::move(film), std::move(xyz));
}

}  // namespace sky
