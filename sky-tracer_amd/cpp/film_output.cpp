// film_output.cpp -- see film_output.hpp.
#include "film_output.hpp"

#include <zlib.h>

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <sys/stat.h>

namespace sky {

namespace {

// Published constants: Bradford cone response (Lam 1985), the D65 white the
// reference adapts to (spectrum.rs:13) and the XYZ -> linear sRGB (D65)
// matrix of IEC 61966-2-1.
constexpr double kBradford[3][3] = {{0.8951, 0.2664, -0.1614}, {-0.7502, 1.7135, 0.0367}, {0.0389, -0.0685, 1.0296}};
constexpr float kD65WhiteY1[3] = {0.9504559f, 1.0f, 1.0890578f};
constexpr float kSrgbFromXyz[3][3] = {{3.2404542f, -1.5371385f, -0.4985314f},
                                      {-0.9692660f, 1.8760108f, 0.0415560f},
                                      {0.0556434f, -0.2040259f, 1.0572252f}};

void inverse3(const double m[3][3], double out[3][3])
{
    double a = m[0][0], b = m[0][1], c = m[0][2], d = m[1][0], e = m[1][1], f = m[1][2], g = m[2][0], h = m[2][1],
           i = m[2][2];
    double det = a * (e * i - f * h) - b * (d * i - f * g) + c * (d * h - e * g);
    out[0][0] = (e * i - f * h) / det;
    out[0][1] = (c * h - b * i) / det;
    out[0][2] = (b * f - c * e) / det;
    out[1][0] = (f * g - d * i) / det;
    out[1][1] = (a * i - c * g) / det;
    out[1][2] = (c * d - a * f) / det;
    out[2][0] = (d * h - e * g) / det;
    out[2][1] = (b * g - a * h) / det;
    out[2][2] = (a * e - b * d) / det;
}

void normalize_y1(const double xyz[3], double out[3])
{
    if (std::fabs(xyz[1]) <= 1.1920929e-7) {  // spectrum.rs:58-64
        for (int k = 0; k < 3; ++k) out[k] = kD65WhiteY1[k];
        return;
    }
    out[0] = xyz[0] / xyz[1];
    out[1] = 1.0;
    out[2] = xyz[2] / xyz[1];
}

void make_dir(const std::string &p)
{
    if (::mkdir(p.c_str(), 0777) != 0 && errno != EEXIST) throw std::runtime_error("cannot create directory " + p);
}

void make_dirs(const std::string &p)
{
    for (size_t i = 1; i <= p.size(); ++i)
        if (i == p.size() || p[i] == '/') make_dir(p.substr(0, i));
}

struct File {
    FILE *f;
    explicit File(const std::string &path, const char *mode) : f(std::fopen(path.c_str(), mode))
    {
        if (!f) throw std::runtime_error("cannot open " + path);
    }
    ~File()
    {
        if (f) std::fclose(f);
    }
};

void put_bytes(std::vector<uint8_t> &b, const void *p, size_t n)
{
    const uint8_t *c = (const uint8_t *)p;
    b.insert(b.end(), c, c + n);
}
template <class T> void put_le(std::vector<uint8_t> &b, T v)
{
    for (size_t i = 0; i < sizeof(T); ++i) b.push_back((uint8_t)((uint64_t)v >> (8 * i)));
}
void put_f32(std::vector<uint8_t> &b, float v)
{
    uint32_t u;
    std::memcpy(&u, &v, 4);
    put_le<uint32_t>(b, u);
}
void put_attr(std::vector<uint8_t> &b, const char *name, const char *type, const std::vector<uint8_t> &value)
{
    put_bytes(b, name, std::strlen(name) + 1);
    put_bytes(b, type, std::strlen(type) + 1);
    put_le<int32_t>(b, (int32_t)value.size());
    b.insert(b.end(), value.begin(), value.end());
}

float half_to_float(uint16_t h)
{
    uint32_t s = (h >> 15) & 1, e = (h >> 10) & 31, m = h & 1023;
    float v;
    if (e == 0)
        v = std::ldexp((float)m, -24);
    else if (e == 31)
        v = m ? NAN : INFINITY;
    else
        v = std::ldexp((float)(m | 1024), (int)e - 25);
    return s ? -v : v;
}

uint32_t crc32(const uint8_t *p, size_t n, uint32_t crc = 0)
{
    static uint32_t table[256];
    static bool init = false;
    if (!init) {
        for (uint32_t i = 0; i < 256; ++i) {
            uint32_t c = i;
            for (int k = 0; k < 8; ++k) c = (c & 1) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
            table[i] = c;
        }
        init = true;
    }
    crc = ~crc;
    for (size_t i = 0; i < n; ++i) crc = table[(crc ^ p[i]) & 255] ^ (crc >> 8);
    return ~crc;
}

void put_be32(std::vector<uint8_t> &b, uint32_t v)
{
    for (int i = 3; i >= 0; --i) b.push_back((uint8_t)(v >> (8 * i)));
}

void png_chunk(std::vector<uint8_t> &out, const char *type, const std::vector<uint8_t> &data)
{
    put_be32(out, (uint32_t)data.size());
    size_t start = out.size();
    put_bytes(out, type, 4);
    out.insert(out.end(), data.begin(), data.end());
    put_be32(out, crc32(out.data() + start, out.size() - start));
}

float tonemap(float c, float exposure) { return 1.0f - std::exp(-std::fmax(c, 0.0f) * exposure); }  // film_output.rs:164-169

uint8_t to_srgb8(float linear)  // film_output.rs:171-179
{
    float x = std::fmin(std::fmax(linear, 0.0f), 1.0f);
    float s = x <= 0.0031308f ? 12.92f * x : 1.055f * std::pow(x, 1.0f / 2.4f) - 0.055f;
    s = std::fmin(std::fmax(s, 0.0f), 1.0f);
    return (uint8_t)(s * 255.0f + 0.5f);
}

}  // namespace

SpectralRgb spectral_rgb_solar_d65(const SceneData &scene)
{
    const uint32_t nb = (uint32_t)scene.wavelengths_nm.size();
    sky_wavelengths wl{nb, scene.wavelengths_nm.data(), scene.weights_nm.empty() ? nullptr : scene.weights_nm.data()};
    std::vector<float> rows(8 * (size_t)nb);
    sky_medium_tables tables{};
    const sky_atmosphere atm = scene.atmosphere_c(tables);
    if (sky_wavelength_table(&atm, &wl, rows.data(), (uint32_t)rows.size()) != SKY_OK)
        throw RenderError(RenderError::Kind::InvalidConfig, sky_last_error());
    SpectralRgb c;
    c.band_xyz.resize(3 * (size_t)nb);
    double sun[3] = {0.0, 0.0, 0.0};
    for (uint32_t b = 0; b < nb; ++b) {
        for (int k = 0; k < 3; ++k) {
            c.band_xyz[3 * b + k] = rows[8 * b + 5 + k];
            sun[k] += (double)rows[8 * b + 4] * rows[8 * b + 5 + k];
        }
    }
    // Bradford adaptation of the sun's white to D65 (spectrum.rs:84-106)
    double src[3], dst[3], dst_in[3] = {kD65WhiteY1[0], kD65WhiteY1[1], kD65WhiteY1[2]};
    normalize_y1(sun, src);
    normalize_y1(dst_in, dst);
    double inv[3][3], ls[3], ld[3];
    inverse3(kBradford, inv);
    for (int r = 0; r < 3; ++r) {
        ls[r] = kBradford[r][0] * src[0] + kBradford[r][1] * src[1] + kBradford[r][2] * src[2];
        ld[r] = kBradford[r][0] * dst[0] + kBradford[r][1] * dst[1] + kBradford[r][2] * dst[2];
    }
    for (int r = 0; r < 3; ++r)
        for (int col = 0; col < 3; ++col) {
            double v = 0.0;
            for (int k = 0; k < 3; ++k) v += inv[r][k] * (ls[k] != 0.0 ? ld[k] / ls[k] : 1.0) * kBradford[k][col];
            c.wb.xyz_from_xyz[r][col] = (float)v;
        }
    for (int k = 0; k < 3; ++k) {
        c.wb.source_white_xyz_y1[k] = (float)src[k];
        c.wb.target_white_xyz_y1[k] = (float)dst[k];
    }
    return c;
}

void SpectralRgb::to_linear_srgb(const float *values, float rgb[3]) const
{
    float xyz[3] = {0.0f, 0.0f, 0.0f}, w[3];
    const size_t nb = band_xyz.size() / 3;
    for (size_t b = 0; b < nb; ++b)
        for (int k = 0; k < 3; ++k) xyz[k] += values[b] * band_xyz[3 * b + k];
    for (int r = 0; r < 3; ++r) w[r] = wb.xyz_from_xyz[r][0] * xyz[0] + wb.xyz_from_xyz[r][1] * xyz[1] + wb.xyz_from_xyz[r][2] * xyz[2];
    bool finite = true;
    for (int r = 0; r < 3; ++r) {
        rgb[r] = kSrgbFromXyz[r][0] * w[0] + kSrgbFromXyz[r][1] * w[1] + kSrgbFromXyz[r][2] * w[2];
        finite = finite && std::isfinite(rgb[r]);
    }
    if (!finite) rgb[0] = rgb[1] = rgb[2] = 0.0f;  // finite_or_black
}

std::string band_exr_name(double center_nm)
{
    char buf[64];
    std::snprintf(buf, sizeof buf, "sky_%03.0fnm.exr", center_nm);
    return buf;
}

void write_exr_rgb(const std::string &path, uint32_t width, uint32_t height, const std::vector<float> &rgb)
{
    if (rgb.size() != 3 * (size_t)width * height) throw std::runtime_error("EXR: wrong pixel count");
    std::vector<uint8_t> b;
    const uint8_t magic[4] = {0x76, 0x2f, 0x31, 0x01};
    put_bytes(b, magic, 4);
    put_le<uint32_t>(b, 2u);  // version 2, single-part scanline
    std::vector<uint8_t> v;
    for (const char *ch : {"B", "G", "R"}) {  // channels sorted by name
        put_bytes(v, ch, 2);
        put_le<int32_t>(v, 2);  // FLOAT
        put_le<uint32_t>(v, 0u);  // pLinear + reserved
        put_le<int32_t>(v, 1);
        put_le<int32_t>(v, 1);
    }
    v.push_back(0);
    put_attr(b, "channels", "chlist", v);
    put_attr(b, "compression", "compression", std::vector<uint8_t>{0});
    v.clear();
    put_le<int32_t>(v, 0);
    put_le<int32_t>(v, 0);
    put_le<int32_t>(v, (int32_t)width - 1);
    put_le<int32_t>(v, (int32_t)height - 1);
    put_attr(b, "dataWindow", "box2i", v);
    put_attr(b, "displayWindow", "box2i", v);
    put_attr(b, "lineOrder", "lineOrder", std::vector<uint8_t>{0});
    v.clear();
    put_f32(v, 1.0f);
    put_attr(b, "pixelAspectRatio", "float", v);
    v.clear();
    put_f32(v, 0.0f);
    put_f32(v, 0.0f);
    put_attr(b, "screenWindowCenter", "v2f", v);
    v.clear();
    put_f32(v, 1.0f);
    put_attr(b, "screenWindowWidth", "float", v);
    b.push_back(0);  // end of header
    const size_t line_bytes = 3 * (size_t)width * 4;
    uint64_t off = b.size() + 8 * (uint64_t)height;
    for (uint32_t y = 0; y < height; ++y) {
        put_le<uint64_t>(b, off);
        off += 8 + line_bytes;
    }
    for (uint32_t y = 0; y < height; ++y) {
        put_le<int32_t>(b, (int32_t)y);
        put_le<int32_t>(b, (int32_t)line_bytes);
        for (int c = 2; c >= 0; --c)  // B, G, R
            for (uint32_t x = 0; x < width; ++x) put_f32(b, rgb[3 * ((size_t)y * width + x) + c]);
    }
    File f(path, "wb");
    if (std::fwrite(b.data(), 1, b.size(), f.f) != b.size()) throw std::runtime_error("cannot write " + path);
}

// OpenEXR chunk decoding (file-layout document, "Compression"): RLE and ZIP
// streams hold the block's bytes after two reversible steps -- the bytes split
// into even and odd positions (first half, second half), then each byte
// replaced by its difference to the previous one plus 128.  RLE: a signed
// count byte c; c < 0 copies -c literal bytes, c >= 0 repeats the next byte
// c + 1 times.  ZIP/ZIPS: a zlib stream.  A chunk whose stored size equals the
// uncompressed size is stored raw.
namespace {

bool rle_decode(const uint8_t *in, size_t n, std::vector<uint8_t> &out, size_t want)
{
    out.clear();
    size_t i = 0;
    while (i < n) {
        const int c = (int8_t)in[i++];
        if (c < 0) {
            const size_t k = (size_t)(-c);
            if (i + k > n || out.size() + k > want) return false;
            out.insert(out.end(), in + i, in + i + k);
            i += k;
        } else {
            if (i >= n || out.size() + (size_t)c + 1 > want) return false;
            out.insert(out.end(), (size_t)c + 1, in[i++]);
        }
    }
    return out.size() == want;
}

void unpredict_deinterleave(std::vector<uint8_t> &t, std::vector<uint8_t> &out)
{
    const size_t n = t.size();
    for (size_t i = 1; i < n; ++i) t[i] = (uint8_t)(t[i - 1] + t[i] - 128);
    out.resize(n);
    const size_t half = (n + 1) / 2;
    for (size_t i = 0; i < half; ++i) {
        out[2 * i] = t[i];
        if (2 * i + 1 < n) out[2 * i + 1] = t[half + i];
    }
}

}  // namespace

std::vector<float> read_exr_r(const std::string &path, uint32_t &width, uint32_t &height)
{
    File f(path, "rb");
    std::vector<uint8_t> b;
    uint8_t buf[65536];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof buf, f.f)) > 0) b.insert(b.end(), buf, buf + n);
    auto fail = [&](const std::string &why) { return std::runtime_error("EXR " + path + ": " + why); };
    if (b.size() < 8 || b[0] != 0x76 || b[1] != 0x2f || b[2] != 0x31 || b[3] != 0x01) throw fail("not an EXR file");
    // version 2; flags: 0x200 tiled, 0x400 long names, 0x800 deep data, 0x1000 multi-part
    if (b[4] != 2 || (b[5] & 0x18)) throw fail("only single-part, non-deep files are read");
    const bool tiled = (b[5] & 0x02) != 0;
    size_t p = 8;
    auto cstr = [&]() {
        size_t s = p;
        while (p < b.size() && b[p]) ++p;
        if (p >= b.size()) throw fail("truncated header");
        return std::string((const char *)&b[s], (p++) - s);
    };
    auto i32 = [&](size_t at) {
        if (at + 4 > b.size()) throw fail("truncated file");
        int32_t v;
        std::memcpy(&v, &b[at], 4);
        return v;
    };
    struct Ch {
        std::string name;
        int type;
    };
    std::vector<Ch> chans;
    int compression = -1;
    int32_t xmin = 0, ymin = 0, xmax = -1, ymax = -1;
    uint32_t tile_w = 0, tile_h = 0;
    for (;;) {
        if (p >= b.size()) throw fail("truncated header");
        if (b[p] == 0) {
            ++p;
            break;
        }
        std::string name = cstr(), type = cstr();
        int32_t size = i32(p);
        p += 4;
        if (size < 0) throw fail("negative attribute size");
        size_t end = p + (size_t)size;
        if (end > b.size()) throw fail("truncated attribute");
        if (name == "channels") {
            size_t q = p;
            while (q < end && b[q]) {
                size_t s = q;
                while (q < end && b[q]) ++q;
                if (q + 17 > end) throw fail("truncated channel list");
                std::string cn((const char *)&b[s], q - s);
                ++q;
                int32_t t;
                std::memcpy(&t, &b[q], 4);
                if (t < 0 || t > 2) throw fail("unknown channel type");
                int32_t xs, ys;
                std::memcpy(&xs, &b[q + 8], 4);
                std::memcpy(&ys, &b[q + 12], 4);
                if (xs != 1 || ys != 1) throw fail("subsampled channels are not read");
                chans.push_back({cn, t});
                q += 16;
            }
        } else if (name == "compression") {
            if (size < 1) throw fail("truncated attribute");
            compression = b[p];
        } else if (name == "dataWindow") {
            if (size < 16) throw fail("truncated attribute");
            xmin = i32(p);
            ymin = i32(p + 4);
            xmax = i32(p + 8);
            ymax = i32(p + 12);
        } else if (name == "tiles") {
            if (size < 9) throw fail("truncated attribute");
            std::memcpy(&tile_w, &b[p], 4);
            std::memcpy(&tile_h, &b[p + 4], 4);
        }
        p = end;
    }
    if (compression < 0 || compression > 3)
        throw fail("compression " + std::to_string(compression) + " is not read (NONE, RLE, ZIPS and ZIP are)");
    if (xmax < xmin || ymax < ymin || (int64_t)xmax - xmin >= (1 << 24) || (int64_t)ymax - ymin >= (1 << 24))
        throw fail("bad data window");
    width = (uint32_t)(xmax - xmin + 1);
    height = (uint32_t)(ymax - ymin + 1);
    if (tiled && (tile_w == 0 || tile_h == 0 || tile_w > (1u << 16) || tile_h > (1u << 16)))
        throw fail("bad tile description");
    size_t pixel_bytes = 0, r_off = (size_t)-1;
    int r_type = -1;
    for (const Ch &c : chans) {
        if (c.name == "R") {
            r_off = pixel_bytes;
            r_type = c.type;
        }
        pixel_bytes += c.type == 1 ? 2 : 4;  // HALF; UINT and FLOAT are 4 bytes
    }
    if (r_off == (size_t)-1) throw fail("no R channel");
    if (r_type == 0) throw fail("R channel is UINT");
    // blocks: scanline chunks of 1 (NONE, RLE, ZIPS) or 16 (ZIP) lines, or level-0 tiles
    const uint32_t lines_per_block = compression == 3 ? 16u : 1u;
    const uint32_t bw = tiled ? tile_w : width, bh = tiled ? tile_h : lines_per_block;
    const uint32_t nbx = (width + bw - 1) / bw, nby = (height + bh - 1) / bh;
    const size_t nblocks = (size_t)nbx * nby;
    if (p + 8 * nblocks > b.size()) throw fail("truncated offset table");
    std::vector<float> out((size_t)width * height);
    std::vector<uint8_t> raw, tmp;
    for (size_t k = 0; k < nblocks; ++k) {
        uint64_t off;
        std::memcpy(&off, &b[p + 8 * k], 8);
        if (off >= b.size()) throw fail("offset out of range");
        uint32_t bx, by;
        size_t data;
        if (tiled) {
            const int32_t tx = i32(off), ty = i32(off + 4), lx = i32(off + 8), ly = i32(off + 12);
            if (lx != 0 || ly != 0) continue;  // not the full-resolution level
            if (tx < 0 || ty < 0 || (uint32_t)tx >= nbx || (uint32_t)ty >= nby) throw fail("bad tile coordinates");
            bx = (uint32_t)tx;
            by = (uint32_t)ty;
            data = (size_t)off + 16;
        } else {
            const int32_t y = i32(off);
            if (y < ymin || y > ymax || (uint32_t)(y - ymin) % bh) throw fail("bad scanline block");
            bx = 0;
            by = (uint32_t)(y - ymin) / bh;
            data = (size_t)off + 4;
        }
        const int32_t packed = i32(data);
        data += 4;
        if (packed < 0 || data + (size_t)packed > b.size()) throw fail("truncated block");
        const uint32_t x0 = bx * bw, y0 = by * bh;
        const uint32_t cw = std::min(bw, width - x0), chh = std::min(bh, height - y0);
        const size_t want = pixel_bytes * cw * chh;
        const uint8_t *src = &b[data];
        if ((size_t)packed != want) {
            if (compression == 0) throw fail("block size does not match the data window");
            if (compression == 1) {
                if (!rle_decode(src, (size_t)packed, tmp, want)) throw fail("bad RLE block");
            } else {
                tmp.resize(want);
                uLongf dst = (uLongf)want;
                if (uncompress(tmp.data(), &dst, src, (uLong)packed) != Z_OK || dst != want)
                    throw fail("bad ZIP block");
            }
            unpredict_deinterleave(tmp, raw);
            src = raw.data();
        }
        // per line of the block: every channel's cw samples in channel order
        for (uint32_t yy = 0; yy < chh; ++yy) {
            const uint8_t *line = src + (size_t)yy * pixel_bytes * cw + r_off * cw;
            float *dst = &out[(size_t)(y0 + yy) * width + x0];
            for (uint32_t x = 0; x < cw; ++x) {
                if (r_type == 1) {
                    uint16_t h;
                    std::memcpy(&h, line + 2 * x, 2);
                    dst[x] = half_to_float(h);
                } else {
                    std::memcpy(&dst[x], line + 4 * x, 4);
                }
            }
        }
    }
    return out;
}

void write_png_rgb8(const std::string &path, uint32_t width, uint32_t height, const std::vector<uint8_t> &rgb)
{
    std::vector<uint8_t> raw;
    raw.reserve((size_t)height * (3 * width + 1));
    for (uint32_t y = 0; y < height; ++y) {
        raw.push_back(0);  // filter: none
        raw.insert(raw.end(), rgb.begin() + 3 * (size_t)y * width, rgb.begin() + 3 * (size_t)(y + 1) * width);
    }
    std::vector<uint8_t> z = {0x78, 0x01};
    size_t i = 0;
    do {  // stored deflate blocks of at most 65535 bytes, the last one flagged
        const size_t n = std::min<size_t>(65535, raw.size() - i);
        z.push_back(i + n >= raw.size() ? 1 : 0);
        put_le<uint16_t>(z, (uint16_t)n);
        put_le<uint16_t>(z, (uint16_t)~n);
        z.insert(z.end(), raw.begin() + i, raw.begin() + i + n);
        i += n;
    } while (i < raw.size());
    uint32_t a = 1, bsum = 0;  // Adler-32
    for (uint8_t c : raw) {
        a = (a + c) % 65521u;
        bsum = (bsum + a) % 65521u;
    }
    put_be32(z, (bsum << 16) | a);
    std::vector<uint8_t> out = {0x89, 'P', 'N', 'G', 0x0D, 0x0A, 0x1A, 0x0A};
    std::vector<uint8_t> ihdr;
    put_be32(ihdr, width);
    put_be32(ihdr, height);
    ihdr.insert(ihdr.end(), {8, 2, 0, 0, 0});
    png_chunk(out, "IHDR", ihdr);
    png_chunk(out, "IDAT", z);
    png_chunk(out, "IEND", {});
    File f(path, "wb");
    if (std::fwrite(out.data(), 1, out.size(), f.f) != out.size()) throw std::runtime_error("cannot write " + path);
}

void write_rgb_outputs(const Film &film, const SceneData &scene, const std::string &out_dir, float png_exposure)
{
    make_dirs(out_dir);
    SpectralRgb conv = spectral_rgb_solar_d65(scene);
    const size_t npx = (size_t)film.width() * film.height();
    std::vector<float> rgb(3 * npx);
    std::vector<uint8_t> png(3 * npx);
    for (size_t i = 0; i < npx; ++i) {
        conv.to_linear_srgb(film.pixel_spectrum(i), &rgb[3 * i]);
        for (int k = 0; k < 3; ++k) png[3 * i + k] = to_srgb8(tonemap(rgb[3 * i + k], png_exposure));
    }
    write_exr_rgb(out_dir + "/sky_rgb.exr", film.width(), film.height(), rgb);
    write_png_rgb8(out_dir + "/sky_rgb.png", film.width(), film.height(), png);
}

void write_outputs(const Film &film, const SceneData &scene, const std::string &out_dir, float png_exposure)
{
    make_dirs(out_dir + "/bands");
    const size_t npx = (size_t)film.width() * film.height();
    std::vector<float> grey(3 * npx);
    for (uint32_t b = 0; b < film.band_count(); ++b) {
        for (size_t i = 0; i < npx; ++i) grey[3 * i] = grey[3 * i + 1] = grey[3 * i + 2] = film.pixel_spectrum(i)[b];
        write_exr_rgb(out_dir + "/bands/" + band_exr_name(scene.wavelengths_nm[b]), film.width(), film.height(), grey);
    }
    write_rgb_outputs(film, scene, out_dir, png_exposure);
}

Film read_film_from_band_exrs(const std::string &out_dir, const SceneData &scene)
{
    const uint32_t nb = (uint32_t)scene.wavelengths_nm.size();
    if (nb == 0) throw std::runtime_error("no spectral bands to rebuild");
    uint32_t w = 0, h = 0;
    std::vector<float> values;
    for (uint32_t b = 0; b < nb; ++b) {
        std::string path = out_dir + "/bands/" + band_exr_name(scene.wavelengths_nm[b]);
        uint32_t bw, bh;
        std::vector<float> img = read_exr_r(path, bw, bh);
        if (b == 0) {
            w = bw;
            h = bh;
            values.assign((size_t)w * h * nb, 0.0f);
        } else if (bw != w || bh != h) {
            throw std::runtime_error("band EXR " + path + " has other dimensions than the first band");
        }
        for (size_t i = 0; i < img.size(); ++i) values[i * nb + b] = img[i];
    }
    return Film(w, h, nb, std::move(values));
}

namespace {

// Shortest decimal that reads back as the same float (serde's f32 form).
std::string fmt_f32(float v)
{
    char buf[32];
    for (int prec = 6; prec <= 9; ++prec) {
        std::snprintf(buf, sizeof buf, "%.*g", prec, (double)v);
        if (std::strtof(buf, nullptr) == v) break;
    }
    return buf;
}

std::string fmt_vec3(const float v[3])
{
    return "[" + fmt_f32(v[0]) + ", " + fmt_f32(v[1]) + ", " + fmt_f32(v[2]) + "]";
}

// The "colorimetry" value of the manifest (main.rs:215-228), indented as a
// member of the top-level object.
std::string colorimetry_json(const SceneData &scene)
{
    const WhiteBalance wb = spectral_rgb_solar_d65(scene).wb;
    std::string j = "{\n";
    j += "    \"cmf\": \"CIE 1931 2 degree standard observer\",\n";
    j += "    \"cmf_source\": \"Wyman, Sloan, Shirley 2013 multi-lobe fit\",\n";
    j += "    \"rgb_color_space\": \"linear sRGB D65\",\n";
    j += "    \"white_balance\": {\n";
    j += "      \"method\": \"Bradford solar-white-to-D65\",\n";
    j += "      \"source_white_xyz_y1\": " + fmt_vec3(wb.source_white_xyz_y1) + ",\n";
    j += "      \"target_white_xyz_y1\": " + fmt_vec3(wb.target_white_xyz_y1) + ",\n";
    j += "      \"xyz_from_xyz\": [" + fmt_vec3(wb.xyz_from_xyz[0]) + ", " + fmt_vec3(wb.xyz_from_xyz[1]) + ", " +
         fmt_vec3(wb.xyz_from_xyz[2]) + "]\n";
    j += "    }\n  }";
    return j;
}

void write_text(const std::string &path, const std::string &text)
{
    File f(path, "w");
    if (std::fwrite(text.data(), 1, text.size(), f.f) != text.size()) throw std::runtime_error("cannot write " + path);
}

}  // namespace

void write_asset_manifest(const RenderConfig &config, const SceneData &scene, const std::string &out_dir)
{
    make_dirs(out_dir);
    const bool lut = config.output_projection == OutputProjection::SkyViewLut;
    std::string j = "{\n  \"version\": 0,\n";
    j += std::string("  \"kind\": \"") + (lut ? "spectral_sky_view_lut_v0" : "spectral_panorama_v0") + "\",\n";
    j += "  \"dimensions\": [" + std::to_string(config.width) + ", " + std::to_string(config.height) + "],\n";
    j += "  \"spp\": " + std::to_string(config.spp) + ",\n";
    j += "  \"seed\": " + std::to_string((unsigned long long)config.seed) + ",\n";
    j += "  \"sun_elevation_deg\": " + fmt_f32(config.sun_elevation_deg) + ",\n";
    j += "  \"sun_azimuth_deg\": " + fmt_f32(config.sun_azimuth_deg) + ",\n";
    j += "  \"observer_altitude_km\": " + fmt_f32(config.observer_altitude_km) + ",\n";
    j += "  \"band_centers_nm\": [";
    for (size_t i = 0; i < scene.wavelengths_nm.size(); ++i)
        j += (i ? ", " : "") + fmt_f32((float)scene.wavelengths_nm[i]);
    j += "],\n  \"data_hash\": null,\n";
    j += "  \"colorimetry\": " + colorimetry_json(scene) + ",\n";
    j += "  \"files\": {\n    \"rgb_exr\": \"sky_rgb.exr\",\n    \"rgb_png\": \"sky_rgb.png\",\n";
    j += "    \"band_exrs\": [";
    for (size_t i = 0; i < scene.wavelengths_nm.size(); ++i)
        j += std::string(i ? ", " : "") + "\"bands/" + band_exr_name(scene.wavelengths_nm[i]) + "\"";
    j += "]\n  }\n}\n";
    write_text(out_dir + "/asset.json", j);
}

bool update_asset_colorimetry(const std::string &out_dir, const SceneData &scene)
{
    const std::string path = out_dir + "/asset.json";
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) return false;  // main.rs:203-205: no manifest, nothing to update
    std::string j;
    char buf[4096];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof buf, f)) > 0) j.append(buf, n);
    std::fclose(f);
    // Replace the value of the top-level "colorimetry" member (or insert one
    // before "files"); strings in this manifest hold no braces.
    const std::string key = "\"colorimetry\"";
    size_t k = j.find(key);
    if (k != std::string::npos) {
        size_t open = j.find('{', k + key.size());
        if (open == std::string::npos) throw std::runtime_error(path + ": malformed colorimetry");
        size_t close = open;
        for (int depth = 0; close < j.size(); ++close) {
            if (j[close] == '{') ++depth;
            else if (j[close] == '}' && --depth == 0) break;
        }
        if (close >= j.size()) throw std::runtime_error(path + ": malformed colorimetry");
        j.replace(open, close + 1 - open, colorimetry_json(scene));
    } else {
        size_t files = j.find("\"files\"");
        if (files == std::string::npos) throw std::runtime_error(path + ": no files member");
        j.insert(files, key + ": " + colorimetry_json(scene) + ",\n  ");
    }
    write_text(path, j);
    return true;
}

}  // namespace sky
