// film_output.hpp -- on-disk outputs of a render, mirroring
// crates/sky-reference-render/src/film_output.rs and the asset manifest of
// crates/sky-core/src/asset.rs:
//   <out>/bands/sky_NNNnm.exr   one grey RGB EXR per band (film_output.rs:83-97)
//   <out>/sky_rgb.exr           linear sRGB (film_output.rs:99-110)
//   <out>/sky_rgb.png           exposure tonemap + sRGB 8-bit (film_output.rs:112-137)
//   <out>/asset.json            SpectralAssetManifest with colorimetry (main.rs:156-228)
// Colour: XYZ from the band values and the per-band cmf weights, Bradford
// adaptation of the sun's white to D65, then XYZ -> linear sRGB
// (spectrum.rs:84-164).  EXRs are single-part scanline files, uncompressed,
// float channels B, G, R; PNGs are RGB8 with stored (uncompressed) deflate.
#pragma once

#include <string>
#include <vector>

#include "sky.hpp"

namespace sky {

// Solar-white-to-D65 Bradford adaptation for a band set (spectrum.rs:84-106, 134-149).
struct WhiteBalance {
    float source_white_xyz_y1[3];
    float target_white_xyz_y1[3];
    float xyz_from_xyz[3][3];
};

struct SpectralRgb {
    std::vector<float> band_xyz;  // [band][3]: cmf(lambda) * weight_nm
    WhiteBalance wb;
    void to_linear_srgb(const float *values, float rgb[3]) const;
};

SpectralRgb spectral_rgb_solar_d65(const SceneData &scene);

// Band file name: "sky_%03.0fnm.exr" of the band centre (film_output.rs:85-88).
std::string band_exr_name(double center_nm);

void write_outputs(const Film &film, const SceneData &scene, const std::string &out_dir, float png_exposure);
void write_rgb_outputs(const Film &film, const SceneData &scene, const std::string &out_dir, float png_exposure);
Film read_film_from_band_exrs(const std::string &out_dir, const SceneData &scene);
void write_asset_manifest(const RenderConfig &config, const SceneData &scene, const std::string &out_dir);
// Rewrites the colorimetry of an existing <out>/asset.json (main.rs:199-212);
// false when there is no manifest.
bool update_asset_colorimetry(const std::string &out_dir, const SceneData &scene);

// Single-part scanline EXR with float B, G, R channels ([pixel][3] = r, g, b).
void write_exr_rgb(const std::string &path, uint32_t width, uint32_t height, const std::vector<float> &rgb);
// Reads the R channel of a single-part EXR: scanline or tiled (full-resolution
// level), compression NONE, RLE, ZIPS or ZIP, R stored as half or float.  The
// band images other tools write (e.g. RLE-compressed 64x64 tiles) read the same
// as write_exr_rgb's uncompressed scanlines.  Other compressions, deep and
// multi-part files raise std::runtime_error.
std::vector<float> read_exr_r(const std::string &path, uint32_t &width, uint32_t &height);
void write_png_rgb8(const std::string &path, uint32_t width, uint32_t height, const std::vector<uint8_t> &rgb);

}  // namespace sky
