// sky.hpp -- C++ host interface mirroring the reference integrator crate.
//
//   sky::SceneData scene = sky::load_scene_data(elevation_deg, azimuth_deg);   // data.rs:15-51
//   sky::Film film = sky::render(scene, config);                               // integrator.rs:125-128
//
// RenderConfig mirrors config.rs:2-52, Film mirrors film.rs:2-62 and
// RenderError mirrors integrator.rs:19-39 (thrown instead of returned in a
// Result).  Everything runs through the C ABI of libsky_hip.so.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/sky_hip.h"

namespace sky {

enum class OutputProjection { Panorama = 0, SkyViewLut = 1 };  // config.rs:14-36

inline const char *label(OutputProjection p)
{
    return p == OutputProjection::Panorama ? "panorama" : "sky-view LUT";
}

struct RenderConfig {  // config.rs:2-12, defaults config.rs:38-52
    uint32_t width = 1024, height = 512, spp = 256;
    uint64_t seed = 0x5EC720260430ull;
    float sun_elevation_deg = 0.0f, sun_azimuth_deg = 0.0f;
    float observer_altitude_km = 0.2f;
    uint32_t direct_light_samples = 1;
    OutputProjection output_projection = OutputProjection::Panorama;
    // extensions (defaults reproduce the reference)
    uint32_t sample_offset = 0, max_depth = 0, rr_start_depth = 4;
    uint32_t majorant_layers = 64;
    uint32_t flags = 0;
    double shell_growth = 1.0;  // majorant shell thickness growth (1: uniform)
};

struct SceneData {  // analytic counterpart of sky_core::atmosphere::SceneData
    sky_planet planet{};
    sky_atmosphere atmosphere{};
    float sun_elevation_deg = 0.0f, sun_azimuth_deg = 0.0f;
    std::vector<double> wavelengths_nm;
    std::vector<double> weights_nm;  // empty: uniform spacing
};

class Film {  // film.rs:2-62
public:
    Film(uint32_t w, uint32_t h, uint32_t b, std::vector<float> values, std::vector<float> xyz = {})
        : w_(w), h_(h), b_(b), values_(std::move(values)), xyz_(std::move(xyz))
    {
    }
    uint32_t width() const { return w_; }
    uint32_t height() const { return h_; }
    uint32_t band_count() const { return b_; }
    const std::vector<float> &values() const { return values_; }
    const std::vector<float> &xyz() const { return xyz_; }
    const float *pixel_spectrum(size_t pixel) const { return values_.data() + pixel * b_; }

private:
    uint32_t w_, h_, b_;
    std::vector<float> values_, xyz_;
};

class RenderError : public std::runtime_error {  // integrator.rs:19-39
public:
    enum class Kind { InvalidConfig, NoAdapter, RequestDevice, BufferMap, Poll, Watchdog, OutOfMemory };
    RenderError(Kind k, const std::string &msg) : std::runtime_error(msg), kind_(k) {}
    Kind kind() const { return kind_; }

private:
    Kind kind_;
};

// load_scene_data for the analytic atmosphere: band_count midpoint bands over
// 380-780 nm, or with band_count 0 the reference's 41 bands at 380..780 nm
// every 10 nm.
SceneData load_scene_data(float sun_elevation_deg, float sun_azimuth_deg, uint32_t band_count = 16);

// integrator::render: validates, renders on device 0 (one long-lived
// context per process), returns the normalised film (and its XYZ image).
Film render(const SceneData &scene, const RenderConfig &config);

}  // namespace sky
