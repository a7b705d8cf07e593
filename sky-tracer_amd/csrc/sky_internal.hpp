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
// Tabulated medium on the merged breakpoint grid of both profiles (every
// coefficient is linear between consecutive breakpoints), packed for the
// device as one blob: sigma [band][K][4] | z [K] | 1/len [K] | index [nidx].
struct MediumTable {
    std::vector<float> sig, z, inv;
    std::vector<uint16_t> idx;
    float inv_dh = 0.0f;
    // Interval records of a uniform cell grid, 8 floats per cell (empty when
    // some cell would need more than two breakpoints): see build_cells.
    std::vector<float> cell;
    float cell_inv_dh = 0.0f;
    bool empty() const { return z.empty(); }
    size_t off_z() const { return (sig.size() * 4 + 15) & ~size_t(15); }
    size_t off_inv() const { return off_z() + ((z.size() * 4 + 15) & ~size_t(15)); }
    size_t off_idx() const { return off_inv() + ((inv.size() * 4 + 15) & ~size_t(15)); }
    // 64-byte aligned, so a 32-byte record never straddles a cache line; every
    // blob size is a multiple of 64, so blobs packed back to back stay aligned.
    size_t off_cell() const { return (off_idx() + idx.size() * 2 + 63) & ~size_t(63); }
    size_t bytes() const { return (off_cell() + cell.size() * 4 + 63) & ~size_t(63); }
    void pack(std::vector<uint8_t> &blob) const;
    // Points the scene's tab* fields into a device copy of the blob.
    void attach(DevScene &s, const void *dev) const;
};

// The reference's coefficients_at (medium.rs:28-57) in FP64: sigma R, As,
// O, Aa (km^-1) of band b at altitude h from the raw profiles.
void table_coeffs(const sky_medium_tables *t, uint32_t band, double lambda_nm, double h, double out[4]);
// Checks the tables of atm (when present) against the wavelength set.
bool validate_tables(const sky_atmosphere *atm, const sky_wavelengths *wl, std::string &err);

struct HostScene {
    DevScene dev{};
    std::vector<DevLambda> lam;
    std::vector<float> maj;
    std::vector<float> shell_w;  // [L + 1] r_k^2 - R^2 of the shell boundaries
    MediumTable tab;              // empty: analytic medium
    uint32_t nlam = 0;
};

// Validates (integrator.rs:130-163, generalised) and packs one render.
sky_status build_scene(const sky_planet *planet, const sky_atmosphere *atm,
                       const sky_wavelengths *wl, const sky_render_config *cfg, HostScene &out,
                       std::string &err);

void fill_wavelength_rows(const sky_atmosphere *atm, const sky_wavelengths *wl,
                          std::vector<DevLambda> &rows);
double shell_height(double top, uint32_t layers, double growth, uint32_t k);
void fill_majorants(const sky_planet *planet, const sky_atmosphere *atm,
                    const sky_wavelengths *wl, uint32_t layers, double margin, double growth,
                    std::vector<float> &out);

struct LaunchShape {
    uint32_t block = 256;
    uint32_t grid = 0;
    uint32_t items = 0;
};

// Kernel launchers (sky_kernels.hip).  Device pointers in `s` must be set.
hipError_t launch_simple(const DevScene &s, hipStream_t stream, LaunchShape &shape);
hipError_t launch_film_to_xyz(const DevScene &s, hipStream_t stream);
// One render of a wave-pool launch: its scene, its work split and where its
// partial sums and completion record go (device memory).
struct DevJob {
    DevScene s;
    uint32_t pixels_per_item, groups, chunks, samples_per_chunk, n_items, item_base;
    float2 *accum;              // [chunk][pixel][band] (sum, sum of squares); only when chunks > 1
    uint32_t *done;             // items finished
    unsigned long long *t_done; // device clock when the last item finished
};

struct PoolArgs {
    const DevJob *jobs;
    uint32_t n_jobs, total_items;
    uint32_t *counter;
    unsigned long long *t_start;
    uint32_t event_threshold, track_steps;
    uint32_t maj_lds;  // floats per wave of the LDS majorant copy (0: read the grids from global memory)
    uint32_t shell_lds;  // floats per wave of the LDS shell table
    uint32_t lam_lds;    // floats per wave of the LDS per-band coefficients (FAST kernels)
};

// Inputs of the adjoint replay (sky_render_adjoint).
struct AdjArgs {
    const float *dl_dxyz;     // [pixel * 3 + c]
    const float4 *lam_units;  // per band: (beta_R, ozone, aerosol) per unit parameter, ln(l / 550)
    float inv_hr2, inv_ha2;   // 1 / H^2
    float inv_spp;
    double *grad;             // [10] gradient, [10] the objective (sum dLoss/dXYZ * XYZ), atomically
};

// Launch-wide settings of the wave-pool integrator (sky_wavepool.hip).
struct WavePlan {
    uint32_t n_items = 0, grid = 0, block = 256, event_threshold = 32, track_steps = 8, occupancy = 0;
    uint32_t maj_lds = 0;   // floats per wave for the majorant grid in LDS (0: global)
    uint32_t shell_lds = 0; // floats per wave for the shell table in LDS (0 with maj_lds 0)
    uint32_t lds_bytes = 0; // dynamic LDS per block
    int fast = 0;           // 1: the _fast, 2: the _fastref, 3: the _fastrt kernels (see plan_launch)
    uint32_t lam_lds = 0;   // floats per wave for the per-band coefficients in LDS (fast kernels)
};

void plan_job(const DevScene &s, const sky_tuning &t, int num_cus, uint32_t min_items, DevJob &j);
size_t job_accum_bytes(const DevJob &j);
uint32_t finalize_blocks(const DevJob &j);
uint32_t launch_min_items(const sky_tuning &t, int num_cus);
// max_maj: the largest nlam x majorant_layers over the launch's jobs;
// max_layers: the largest majorant_layers; flags_all: the estimator flags if
// every job has the same ones, else kFlagMixed.  The _fast kernels serve
// carried optical depth + lobe sampling + delta-tracked shadow rays, the
// _fastrt kernels the same with ratio-tracked shadow rays, the _fastref
// kernels the reference's estimator (fresh draw per shell, isotropic
// directions, delta-tracked shadow rays), all with the tables in LDS and at
// most kFastMaxShells majorant shells.
constexpr uint32_t kFlagMixed = 0x80000000u;
constexpr uint32_t kFastMaxShells = 8;
WavePlan plan_launch(const sky_tuning &t, int num_cus, uint32_t total_items, uint32_t max_maj,
                     uint32_t max_layers, uint32_t flags_all);
// fin_blocks: the largest finalize_blocks over the launch's jobs (0: no
// job was split into sample chunks, no finalize launch).
hipError_t launch_pool(const DevJob *d_jobs, uint32_t n_jobs, uint32_t fin_blocks, const WavePlan &p,
                       uint32_t *meta, uint32_t meta_words, hipStream_t stream);
hipError_t launch_zero(uint32_t *p, uint32_t n, hipStream_t stream);
hipError_t launch_adjoint(const DevJob *d_jobs, const WavePlan &p, uint32_t *meta, uint32_t meta_words,
                          const AdjArgs &A, hipStream_t stream);
hipError_t launch_optical_depth(const DevScene &s, const double *rays, uint32_t n,
                                const double *dev_lambda_d, double *tau, hipStream_t stream);

// Records the message returned by sky_last_error() and passes st through.
sky_status set_error(sky_status st, const std::string &msg);

}  // namespace sky
