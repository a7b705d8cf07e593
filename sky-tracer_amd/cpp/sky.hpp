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
#include <memory>
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

// The reference's tabulated medium (data.rs:53-152) in the layout of
// sky_medium_tables.
struct MediumData {
    std::vector<double> atmosphere;      // [n][3]: altitude_km, air_cm3, ozone_cm3 (ascending)
    std::vector<double> aerosol;         // [n][5]: altitude_km, mass_g_m3 of inso, waso, soot, suso
    std::vector<double> aerosol_optics;  // [bands][4][2]: scattering, absorption per g/m^3
    std::vector<double> ozone_cross_section_cm2, solar_irradiance_w_m2;  // [bands]
};

struct SceneData {  // sky_core::atmosphere::SceneData
    sky_planet planet{};
    sky_atmosphere atmosphere{};
    float sun_elevation_deg = 0.0f, sun_azimuth_deg = 0.0f;
    std::vector<double> wavelengths_nm;
    std::vector<double> weights_nm;  // empty: uniform spacing
    std::shared_ptr<const MediumData> medium;  // null: the analytic profiles
    // The atmosphere with the tables attached; `storage` must outlive its use.
    sky_atmosphere atmosphere_c(sky_medium_tables &storage) const
    {
        sky_atmosphere a = atmosphere;
        a.tables = nullptr;
        if (medium) {
            storage.atmosphere_points = (uint32_t)(medium->atmosphere.size() / 3);
            storage.aerosol_points = (uint32_t)(medium->aerosol.size() / 5);
            storage.atmosphere = medium->atmosphere.data();
            storage.aerosol = medium->aerosol.data();
            storage.aerosol_optics = medium->aerosol_optics.data();
            storage.ozone_cross_section_cm2 = medium->ozone_cross_section_cm2.data();
            storage.rayleigh_cross_section_m2 = nullptr;
            storage.solar_irradiance_w_m2 = medium->solar_irradiance_w_m2.data();
            a.tables = &storage;
        }
        return a;
    }
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

// load_scene_data (data.rs:15-51) from the reference's data directory:
// bands.csv, atmosphere_profile.csv, aerosol_profile.csv and
// aerosol_optics.csv; planet 6360 / 6480 km, ground albedo 0.18, sun radius
// 0.00465047 rad (atmosphere.rs:9-10, geometry.rs:15).  The aerosol phase is
// Henyey-Greenstein (g 0.8): the reference's mie_phase.csv is not in its
// checkout.  Throws std::runtime_error naming the file on a missing or
// malformed table.
SceneData load_scene_data(const std::string &data_dir, float sun_elevation_deg, float sun_azimuth_deg);

// load_scene_data for the analytic atmosphere: band_count midpoint bands over
// 380-780 nm, or with band_count 0 the reference's 41 bands at 380..780 nm
// every 10 nm.
SceneData load_scene_data(float sun_elevation_deg, float sun_azimuth_deg, uint32_t band_count = 16);

// integrator::render: validates, renders on device 0 (one long-lived
// context per process), returns the normalised film (and its XYZ image).
Film render(const SceneData &scene, const RenderConfig &config);

// Launch shape of the wave-pool integrator for later render() calls, applied
// to the process-wide context after a render's configuration is validated
// (sky_context_set_tuning; an invalid shape makes render() throw).
void set_launch_tuning(const sky_tuning &tuning);

}  // namespace sky
