// sky_render -- command-line driver mirroring sky-reference-render
// (crates/sky-reference-render/src/main.rs:22-138): the same flags for the
// render itself, the load / render / output / total timing lines, and the
// film written as raw float32 files plus a small JSON manifest (the EXR/PNG
// writers of film_output.rs are not part of this build).
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "sky.hpp"

namespace {

std::string format_duration(double s)  // main.rs:230-241
{
    char buf[64];
    if (s >= 1.0)
        std::snprintf(buf, sizeof buf, "%.3fs", s);
    else
        std::snprintf(buf, sizeof buf, "%.1fms", s * 1e3);
    return buf;
}

void usage()
{
    std::puts(
        "sky_render [--width N] [--height N] [--spp N] [--seed N] [--sun-elevation-deg F]\n"
        "           [--sun-azimuth-deg F] [--observer-altitude-km F] [--direct-light-samples N]\n"
        "           [--sky-view-lut] [--bands N] [--max-depth N] [--fast] [--out DIR]");
}

}  // namespace

int main(int argc, char **argv)
{
    auto total0 = std::chrono::steady_clock::now();
    sky::RenderConfig cfg;
    bool lut = false;
    bool width_set = false, height_set = false;
    uint32_t bands = 16;
    std::string out = "out";
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        auto next = [&]() -> const char * {
            if (i + 1 >= argc) {
                usage();
                std::exit(2);
            }
            return argv[++i];
        };
        if (a == "--width") cfg.width = (uint32_t)std::strtoul(next(), nullptr, 10), width_set = true;
        else if (a == "--height") cfg.height = (uint32_t)std::strtoul(next(), nullptr, 10), height_set = true;
        else if (a == "--spp") cfg.spp = (uint32_t)std::strtoul(next(), nullptr, 10);
        else if (a == "--seed") cfg.seed = std::strtoull(next(), nullptr, 0);
        else if (a == "--sun-elevation-deg") cfg.sun_elevation_deg = std::strtof(next(), nullptr);
        else if (a == "--sun-azimuth-deg") cfg.sun_azimuth_deg = std::strtof(next(), nullptr);
        else if (a == "--observer-altitude-km") cfg.observer_altitude_km = std::strtof(next(), nullptr);
        else if (a == "--direct-light-samples") cfg.direct_light_samples = (uint32_t)std::strtoul(next(), nullptr, 10);
        else if (a == "--sky-view-lut") lut = true;
        else if (a == "--bands") bands = (uint32_t)std::strtoul(next(), nullptr, 10);
        else if (a == "--max-depth") cfg.max_depth = (uint32_t)std::strtoul(next(), nullptr, 10);
        else if (a == "--fast") cfg.flags |= SKY_FLAG_CARRIED_TAU | SKY_FLAG_PHASE_IS;
        else if (a == "--out") out = next();
        else {
            usage();
            return 2;
        }
    }
    // resolve_dimensions, main.rs:140-154
    if (!width_set) cfg.width = lut ? 256 : 2048;
    if (!height_set) cfg.height = lut ? 256 : 1024;
    cfg.output_projection = lut ? sky::OutputProjection::SkyViewLut : sky::OutputProjection::Panorama;

    auto load0 = std::chrono::steady_clock::now();
    sky::SceneData scene = sky::load_scene_data(cfg.sun_elevation_deg, cfg.sun_azimuth_deg, bands);
    double load_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - load0).count();
    std::printf("config: gpu-integrator output=%s direct-light-samples=%u\n", sky::label(cfg.output_projection),
                cfg.direct_light_samples);
    auto render0 = std::chrono::steady_clock::now();
    try {
        sky::Film film = sky::render(scene, cfg);
        double render_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - render0).count();
        auto out0 = std::chrono::steady_clock::now();
        std::string mk = "mkdir -p '" + out + "'";
        if (std::system(mk.c_str()) != 0) throw std::runtime_error("cannot create " + out);
        FILE *f = std::fopen((out + "/film_f32.raw").c_str(), "wb");
        FILE *g = std::fopen((out + "/xyz_f32.raw").c_str(), "wb");
        FILE *m = std::fopen((out + "/asset.json").c_str(), "w");
        if (!f || !g || !m) throw std::runtime_error("cannot write outputs in " + out);
        std::fwrite(film.values().data(), sizeof(float), film.values().size(), f);
        std::fwrite(film.xyz().data(), sizeof(float), film.xyz().size(), g);
        std::fprintf(m, "{\"width\": %u, \"height\": %u, \"bands\": %u, \"spp\": %u, \"projection\": \"%s\", "
                        "\"film\": \"film_f32.raw\", \"xyz\": \"xyz_f32.raw\", \"wavelengths_nm\": [",
                     film.width(), film.height(), film.band_count(), cfg.spp, sky::label(cfg.output_projection));
        for (size_t i = 0; i < scene.wavelengths_nm.size(); ++i)
            std::fprintf(m, "%s%.6f", i ? ", " : "", scene.wavelengths_nm[i]);
        std::fprintf(m, "]}\n");
        std::fclose(f);
        std::fclose(g);
        std::fclose(m);
        double out_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - out0).count();
        double total_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - total0).count();
        std::printf("wrote %ux%u %s with %u spp to %s\n", film.width(), film.height(),
                    sky::label(cfg.output_projection), cfg.spp, out.c_str());
        std::printf("timing: load=%s render=%s output=%s total=%s\n", format_duration(load_s).c_str(),
                    format_duration(render_s).c_str(), format_duration(out_s).c_str(),
                    format_duration(total_s).c_str());
    } catch (const sky::RenderError &e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
