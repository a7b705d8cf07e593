// sky_internal.hpp -- host-side interfaces between the C ABI (sky_api.cpp) and
// the kernel translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/sky_hip.h"
#include "sky_device.hpp"

namespace sky {

// Host copy of one render's packed scene (integrator.rs:319-418 PackedScene):
// the device struct plus the per-wavelength and majorant tables it points to.
struct HostScene {
    DevScene dev{};
    std::vector<DevLambda> lam;
    std::vector<float> maj;
    uint32_t nlam = 0;
};

// Validates (integrator.rs:130-163, generalised) and packs one render.
sky_status build_scene(const sky_planet *planet, const sky_atmosphere *atm,
                       const sky_wavelengths *wl, const sky_render_config *cfg, HostScene &out,
                       std::string &err);

void fill_wavelength_rows(const sky_atmosphere *atm, const sky_wavelengths *wl,
                          std::vector<DevLambda> &rows);
void fill_majorants(const sky_planet *planet, const sky_atmosphere *atm,
                    const sky_wavelengths *wl, uint32_t layers, double margin,
                    std::vector<float> &out);

struct LaunchShape {
    uint32_t block = 256;
    uint32_t grid = 0;
    uint32_t items = 0;
};

// Kernel launchers (sky_kernels.hip).  Device pointers in `s` must be set.
hipError_t launch_simple(const DevScene &s, hipStream_t stream, LaunchShape &shape);
hipError_t launch_film_to_xyz(const DevScene &s, hipStream_t stream);
// Work split of the wave-pool integrator (sky_wavepool.hip).
struct WavePlan {
    uint32_t pixels_per_item = 1, groups = 0, chunks = 1, samples_per_chunk = 0, n_items = 0;
    uint32_t grid = 0, block = 256, event_threshold = 32, track_steps = 8;
    size_t accum_bytes = 0;
};
WavePlan plan_wavepool(const DevScene &s, const sky_tuning &t, int num_cus);
// accum: plan.accum_bytes of device memory; counter: one device uint32.
hipError_t launch_wavepool(const DevScene &s, const WavePlan &plan, void *accum,
                           uint32_t *counter, hipStream_t stream);
hipError_t launch_optical_depth(const DevScene &s, const double *rays, uint32_t n,
                                const double *dev_lambda_d, double *tau, hipStream_t stream);

}  // namespace sky
