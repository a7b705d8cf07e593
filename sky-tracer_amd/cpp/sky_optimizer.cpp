// sky_wavelength_optimizer -- the reference's sparse-wavelength sky-view LUT
// primary search (crates/sky-wavelength-optimizer/src/main.rs) with every
// combo scored on the GPU through include/sky_scan.h:
//   load the sky-view LUT assets (asset.json + band EXRs), reference colour
//   per LUT sample, p99 luminance scale                      main.rs:255-351
//   screen all combos on a deterministic sample subset, or a
//   beam search with local refinement above 5M combos        main.rs:156-199, 560-691
//   refine the best candidates on all samples with exact
//   percentiles, rank by RMSE + 0.25 P95                     main.rs:200-241
//   ranking.csv, ranking_by_elevation.csv, report.md         main.rs:1203-1402
// Same flags and outputs as the reference CLI.  Assets are read from the
// uncompressed EXRs that sky_render writes.
#include <algorithm>
#include <cerrno>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <set>
#include <stdexcept>
#include <string>
#include <sys/stat.h>
#include <thread>
#include <vector>

#include "../../include/sky_scan.h"
#include "film_output.hpp"

namespace {

constexpr uint32_t kBandCount = 41;                    // spectrum.rs:5
constexpr uint64_t kExhaustiveLimit = 5000000;         // main.rs:41
constexpr uint64_t kRefineChunk = 65536;

// ------------------------------------------------------------------ JSON

struct Json {
    enum Type { Null, Bool, Num, Str, Arr, Obj } type = Null;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<Json> arr;
    std::vector<std::pair<std::string, Json>> obj;
    const Json *get(const char *k) const
    {
        for (const auto &kv : obj)
            if (kv.first == k) return &kv.second;
        return nullptr;
    }
};

class JsonReader {
public:
    explicit JsonReader(const std::string &t) : s_(t) {}
    Json parse()
    {
        Json v = value();
        ws();
        if (p_ != s_.size()) fail("trailing characters");
        return v;
    }

private:
    const std::string &s_;
    size_t p_ = 0;
    [[noreturn]] void fail(const char *why) { throw std::runtime_error(std::string("JSON: ") + why); }
    char peek() const { return p_ < s_.size() ? s_[p_] : '\0'; }
    char take()
    {
        if (p_ >= s_.size()) fail("unexpected end");
        return s_[p_++];
    }
    void ws()
    {
        while (p_ < s_.size() && std::strchr(" \t\r\n", s_[p_])) ++p_;
    }
    bool lit(const char *w)
    {
        size_t n = std::strlen(w);
        if (s_.compare(p_, n, w) != 0) return false;
        p_ += n;
        return true;
    }
    std::string string()
    {
        if (peek() != '"') fail("expected string");
        ++p_;
        std::string out;
        while (p_ < s_.size() && s_[p_] != '"') {
            char c = s_[p_++];
            if (c == '\\') {
                char e = take();
                switch (e) {
                case 'n': out += '\n'; break;
                case 't': out += '\t'; break;
                case 'r': out += '\r'; break;
                case 'b': out += '\b'; break;
                case 'f': out += '\f'; break;
                case 'u': {
                    if (p_ + 4 > s_.size()) fail("bad escape");
                    unsigned cp = (unsigned)std::strtoul(s_.substr(p_, 4).c_str(), nullptr, 16);
                    p_ += 4;
                    if (cp < 0x80) {
                        out += (char)cp;
                    } else if (cp < 0x800) {
                        out += (char)(0xC0 | (cp >> 6));
                        out += (char)(0x80 | (cp & 63));
                    } else {
                        out += (char)(0xE0 | (cp >> 12));
                        out += (char)(0x80 | ((cp >> 6) & 63));
                        out += (char)(0x80 | (cp & 63));
                    }
                    break;
                }
                default: out += e;
                }
            } else {
                out += c;
            }
        }
        if (p_ >= s_.size()) fail("unterminated string");
        ++p_;
        return out;
    }
    Json value()
    {
        ws();
        if (p_ >= s_.size()) fail("unexpected end");
        Json v;
        char c = peek();
        if (c == '{') {
            v.type = Json::Obj;
            ++p_;
            ws();
            if (peek() == '}') {
                ++p_;
                return v;
            }
            for (;;) {
                ws();
                std::string k = string();
                ws();
                if (take() != ':') fail("expected ':'");
                v.obj.emplace_back(k, value());
                ws();
                if (peek() == ',') {
                    ++p_;
                    continue;
                }
                if (take() != '}') fail("expected '}'");
                return v;
            }
        }
        if (c == '[') {
            v.type = Json::Arr;
            ++p_;
            ws();
            if (peek() == ']') {
                ++p_;
                return v;
            }
            for (;;) {
                v.arr.push_back(value());
                ws();
                if (peek() == ',') {
                    ++p_;
                    continue;
                }
                if (take() != ']') fail("expected ']'");
                return v;
            }
        }
        if (c == '"') {
            v.type = Json::Str;
            v.str = string();
            return v;
        }
        if (lit("true")) {
            v.type = Json::Bool;
            v.b = true;
            return v;
        }
        if (lit("false")) {
            v.type = Json::Bool;
            return v;
        }
        if (lit("null")) return v;
        char *end = nullptr;
        v.num = std::strtod(s_.c_str() + p_, &end);
        if (end == s_.c_str() + p_) fail("bad value");
        p_ = (size_t)(end - s_.c_str());
        v.type = Json::Num;
        return v;
    }
};

std::string read_text(const std::string &path)
{
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) throw std::runtime_error("cannot open " + path);
    std::string t;
    char buf[65536];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof buf, f)) > 0) t.append(buf, n);
    std::fclose(f);
    return t;
}

// ------------------------------------------------------------------ options

enum Strategy : uint32_t { Voronoi = SKY_WEIGHTS_VORONOI, CmfMass = SKY_WEIGHTS_CMF_MASS, CmfMoments = SKY_WEIGHTS_CMF_MOMENTS };

const char *strategy_label(uint32_t s)
{
    return s == CmfMass ? "cmf-mass" : s == CmfMoments ? "cmf-moments" : "voronoi";
}

const char *strategy_description(uint32_t s)  // main.rs:64-77
{
    if (s == CmfMass) return "Non-negative fixed-sum least squares matching integrated CIE XYZ response for a flat spectrum.";
    if (s == CmfMoments)
        return "Non-negative fixed-sum least squares matching integrated CIE XYZ response times normalized 1, "
               "lambda, and lambda^2 moments.";
    return "Wavelength-axis Voronoi widths. Each sampled wavelength represents the nearest interval in nm.";
}

struct Options {  // main.rs:17-36
    std::vector<std::string> asset_dirs;
    uint32_t primary_count = 3;
    uint64_t screen_samples = 65536;
    uint64_t refine_top = 512;
    uint64_t top = 11;
    uint64_t threads = 0;
    uint64_t beam_width = 4096;
    uint32_t strategy = Voronoi;
    std::string out_dir = "target/primary_search";
    int device = 0;
};

void usage()
{
    std::puts(
        "sky_wavelength_optimizer --asset-dir DIR [--asset-dir DIR ...] [--primary-count N] [--screen-samples N]\n"
        "                         [--refine-top N] [--top N] [--threads N] [--beam-width N]\n"
        "                         [--weight-strategy voronoi|cmf-mass|cmf-moments] [--out-dir DIR] [--device N]\n"
        "Search physically weighted sparse-wavelength sky-view LUT primaries (combos scored on the GPU).");
}

// ------------------------------------------------------------------ dataset

struct Asset {
    std::string path;
    uint64_t width = 0, height = 0, spp = 0;
    float sun_elevation_deg = 0.0f;
    uint64_t start = 0, count = 0;
};

struct Dataset {
    std::vector<Asset> assets;
    std::vector<float> wl;
    std::vector<double> width, cie;  // [band], [band][3]
    std::vector<float> values;       // [sample][band]
    std::vector<double> ref_y;
    std::vector<float> ref_oklab;    // [sample][3]
    double y_p99 = 0.0, scale = 1.0;
    uint64_t n() const { return ref_y.size(); }
    uint32_t bands() const { return (uint32_t)wl.size(); }
};

double percentile_sorted(const std::vector<double> &v, double q)  // main.rs:1111-1117
{
    if (v.empty()) return NAN;
    q = std::min(1.0, std::max(0.0, q));
    return v[(size_t)std::llround((double)(v.size() - 1) * q)];
}

double percentile(std::vector<double> v, double q)  // main.rs:1102-1109
{
    v.erase(std::remove_if(v.begin(), v.end(), [](double x) { return !std::isfinite(x); }), v.end());
    std::sort(v.begin(), v.end());
    return percentile_sorted(v, q);
}

void xyz_to_oklab(const double *xyz, double *lab)  // main.rs:1088-1098 (Ottosson 2020)
{
    const double l = std::cbrt(0.8189330101 * xyz[0] + 0.3618667424 * xyz[1] - 0.1288597137 * xyz[2]);
    const double m = std::cbrt(0.0329845436 * xyz[0] + 0.9293118715 * xyz[1] + 0.0361456387 * xyz[2]);
    const double s = std::cbrt(0.0482003018 * xyz[0] + 0.2643662691 * xyz[1] + 0.6338517070 * xyz[2]);
    lab[0] = 0.2104542553 * l + 0.7936177850 * m - 0.0040720468 * s;
    lab[1] = 1.9779984951 * l - 2.4285922050 * m + 0.4505937099 * s;
    lab[2] = 0.0259040371 * l + 0.7827717662 * m - 0.8086757660 * s;
}

void check(sky_status st)
{
    if (st != SKY_OK) throw std::runtime_error(sky_last_error());
}

Dataset load_dataset(const std::vector<std::string> &dirs)  // main.rs:255-351
{
    if (dirs.empty()) throw std::runtime_error("at least one --asset-dir is required");
    Dataset d;
    std::vector<std::vector<float>> bands;
    std::vector<double> xyz;
    for (const std::string &dir : dirs) {
        const std::string mpath = dir + "/asset.json";
        Json m = JsonReader(read_text(mpath)).parse();
        const Json *version = m.get("version"), *kind = m.get("kind"), *dims = m.get("dimensions"),
                   *centers = m.get("band_centers_nm"), *files = m.get("files");
        if (!version || !kind || !dims || !centers || !files || dims->arr.size() != 2)
            throw std::runtime_error(mpath + ": missing manifest fields");
        if (version->num != 0.0) throw std::runtime_error(mpath + ": expected version 0");
        if (kind->str != "spectral_sky_view_lut_v0")
            throw std::runtime_error(mpath + ": expected kind spectral_sky_view_lut_v0, found " + kind->str);
        if (centers->arr.size() != kBandCount)
            throw std::runtime_error(mpath + ": expected 41 bands, found " + std::to_string(centers->arr.size()));
        const Json *exrs = files->get("band_exrs");
        if (!exrs || exrs->arr.size() != centers->arr.size())
            throw std::runtime_error(mpath + ": band file count mismatch");
        std::vector<float> wl;
        for (const Json &c : centers->arr) wl.push_back((float)c.num);
        if (d.wl.empty()) {
            d.wl = wl;
            bands.assign(wl.size(), {});
        } else if (d.wl != wl) {
            throw std::runtime_error(mpath + ": wavelength grid differs from first asset");
        }
        Asset a;
        a.path = dir;
        a.width = (uint64_t)dims->arr[0].num;
        a.height = (uint64_t)dims->arr[1].num;
        if (const Json *spp = m.get("spp")) a.spp = (uint64_t)spp->num;
        if (const Json *el = m.get("sun_elevation_deg")) a.sun_elevation_deg = (float)el->num;
        const uint64_t npx = a.width * a.height;
        for (size_t b = 0; b < exrs->arr.size(); ++b) {
            const std::string p = dir + "/" + exrs->arr[b].str;
            uint32_t w = 0, h = 0;
            std::vector<float> img = sky::read_exr_r(p, w, h);
            if (w != a.width || h != a.height)
                throw std::runtime_error(p + ": dimensions " + std::to_string(w) + "x" + std::to_string(h) +
                                         " differ from manifest " + std::to_string(a.width) + "x" +
                                         std::to_string(a.height));
            bands[b].insert(bands[b].end(), img.begin(), img.end());
        }
        a.start = d.ref_y.size();
        a.count = npx;
        if (d.cie.empty()) {
            d.cie.resize(3 * d.wl.size());
            check(sky_cie_1931_2deg(d.wl.data(), (uint32_t)d.wl.size(), d.cie.data()));
        }
        for (uint64_t px = 0; px < npx; ++px) {
            double s[3] = {0.0, 0.0, 0.0};
            for (size_t b = 0; b < d.wl.size(); ++b) {
                const double v = (double)bands[b][a.start + px];
                for (int c = 0; c < 3; ++c) s[c] += v * d.cie[3 * b + c];
            }
            xyz.insert(xyz.end(), s, s + 3);
            d.ref_y.push_back(s[1]);
        }
        d.assets.push_back(a);
    }
    const uint64_t n = d.ref_y.size();
    const uint32_t nb = d.bands();
    d.values.resize(n * nb);
    for (uint32_t b = 0; b < nb; ++b)
        for (uint64_t s = 0; s < n; ++s) d.values[s * nb + b] = bands[b][s];
    d.y_p99 = percentile(d.ref_y, 0.99);
    d.scale = std::max(d.y_p99, 1.0e-12);
    d.ref_oklab.resize(3 * n);
    for (uint64_t s = 0; s < n; ++s) {
        double x[3] = {xyz[3 * s] / d.scale, xyz[3 * s + 1] / d.scale, xyz[3 * s + 2] / d.scale}, lab[3];
        xyz_to_oklab(x, lab);
        for (int c = 0; c < 3; ++c) d.ref_oklab[3 * s + c] = (float)lab[c];
    }
    d.width.resize(nb);
    check(sky_band_widths(d.wl.data(), nb, d.width.data()));
    return d;
}

std::vector<uint32_t> screening_indices(uint64_t count, uint64_t requested)  // main.rs:1096-1105
{
    std::vector<uint32_t> idx;
    if (requested == 0 || requested >= count) {
        idx.resize(count);
        std::iota(idx.begin(), idx.end(), 0u);
        return idx;
    }
    const double stride = (double)count / (double)requested;
    for (uint64_t i = 0; i < requested; ++i)
        idx.push_back((uint32_t)std::min<uint64_t>((uint64_t)std::floor(((double)i + 0.5) * stride), count - 1));
    return idx;
}

uint64_t combo_count(uint64_t len, uint64_t count)  // main.rs:550-560
{
    if (count > len) return 0;
    count = std::min(count, len - count);
    uint64_t r = 1;
    for (uint64_t i = 0; i < count; ++i) r = r * (len - i) / (i + 1);
    return r;
}

// ------------------------------------------------------------------ scoring

struct Scored {  // ComboResult, main.rs:97-110
    std::vector<uint8_t> combo;
    double weights[SKY_SCAN_MAX_PRIMARIES] = {};
    double mean = NAN, rmse = NAN, p50 = NAN, p90 = NAN, p95 = NAN, p99 = NAN, max = NAN, score = NAN;
    uint64_t samples = 0;
};

// total order with NaN last (f64::total_cmp for non-negative scores)
bool score_less(double a, double b)
{
    if (std::isnan(a)) return false;
    if (std::isnan(b)) return true;
    return a < b;
}

struct Scorer {
    const Dataset &d;
    sky_scan_dataset *ds;
    uint32_t strategy;
    uint32_t threads;
    double kernel_ms = 0.0;
    uint64_t evaluations = 0;

    std::vector<double> weights(const std::vector<uint8_t> &combos, uint32_t count) const
    {
        std::vector<double> w(combos.size());
        check(sky_quadrature_weights(d.wl.data(), d.bands(), d.cie.data(), combos.data(), count,
                                     combos.size() / count, strategy, threads, w.data()));
        return w;
    }
    std::vector<float> coefficients(const std::vector<uint8_t> &combos, const std::vector<double> &w) const
    {
        std::vector<float> c(3 * combos.size());
        for (size_t i = 0; i < combos.size(); ++i) {
            const uint8_t b = combos[i];
            const double f = w[i] / d.width[b] / d.scale;
            for (int k = 0; k < 3; ++k) c[3 * i + k] = (float)(f * d.cie[3 * b + k]);
        }
        return c;
    }
    // Screening score (RMSE) of each combo on the sample subset.
    std::vector<double> screen(const std::vector<uint8_t> &combos, uint32_t count, const std::vector<uint32_t> &idx)
    {
        const uint64_t nc = combos.size() / count;
        std::vector<double> sum(nc), sq(nc), rmse(nc);
        if (nc == 0) return rmse;
        const std::vector<float> coef = coefficients(combos, weights(combos, count));
        sky_scan_stats st{};
        check(sky_scan_combos(ds, combos.data(), coef.data(), count, nc, idx.data(), idx.size(), sum.data(), sq.data(),
                              &st));
        kernel_ms += st.kernel_ms;
        evaluations += st.evaluations;
        for (uint64_t i = 0; i < nc; ++i) rmse[i] = std::sqrt(sq[i] / (double)idx.size());
        return rmse;
    }
    // Exact statistics on the given samples (all when idx is empty).
    std::vector<Scored> refine(const std::vector<uint8_t> &combos, uint32_t count, const std::vector<uint32_t> &idx)
    {
        const uint64_t nc = combos.size() / count;
        std::vector<Scored> out(nc);
        if (nc == 0) return out;
        const std::vector<double> w = weights(combos, count);
        const std::vector<float> coef = coefficients(combos, w);
        std::vector<sky_combo_stats> stats(nc);
        for (uint64_t c0 = 0; c0 < nc; c0 += kRefineChunk) {
            const uint64_t m = std::min(kRefineChunk, nc - c0);
            sky_scan_stats st{};
            check(sky_refine_combos(ds, combos.data() + c0 * count, coef.data() + 3 * c0 * count, count, m,
                                    idx.empty() ? nullptr : idx.data(), idx.size(), stats.data() + c0, &st));
            kernel_ms += st.kernel_ms;
            evaluations += st.evaluations;
        }
        for (uint64_t i = 0; i < nc; ++i) {
            Scored &r = out[i];
            r.combo.assign(combos.begin() + i * count, combos.begin() + (i + 1) * count);
            for (uint32_t k = 0; k < count; ++k) r.weights[k] = w[i * count + k];
            const sky_combo_stats &s = stats[i];
            r.mean = s.mean;
            r.rmse = s.rmse;
            r.p50 = s.p50;
            r.p90 = s.p90;
            r.p95 = s.p95;
            r.p99 = s.p99;
            r.max = s.max;
            r.samples = s.sample_count;
            r.score = r.rmse + 0.25 * r.p95;
        }
        return out;
    }
};

std::vector<uint8_t> all_combos(uint32_t len, uint32_t count)  // main.rs:542-548, 693-712
{
    std::vector<uint8_t> out;
    std::vector<uint8_t> cur(count);
    for (uint32_t k = 0; k < count; ++k) cur[k] = (uint8_t)k;
    if (count > len) return out;
    for (;;) {
        out.insert(out.end(), cur.begin(), cur.end());
        int k = (int)count - 1;
        while (k >= 0 && cur[k] == len - count + k) --k;
        if (k < 0) break;
        ++cur[k];
        for (uint32_t j = k + 1; j < count; ++j) cur[j] = (uint8_t)(cur[j - 1] + 1);
    }
    return out;
}

std::vector<size_t> order_by(const std::vector<double> &score)
{
    std::vector<size_t> o(score.size());
    std::iota(o.begin(), o.end(), 0);
    std::stable_sort(o.begin(), o.end(), [&](size_t a, size_t b) { return score_less(score[a], score[b]); });
    return o;
}

// Beam search with local refinement (main.rs:562-691); candidates are scored a
// whole depth (or a whole local-search round) per GPU call.
std::vector<uint8_t> beam_search(Scorer &sc, uint32_t n_bands, uint32_t count, const std::vector<uint32_t> &idx,
                                 uint64_t beam_width, uint64_t keep_count, std::vector<double> *scores)
{
    beam_width = std::max<uint64_t>(1, beam_width);
    keep_count = std::max(keep_count, beam_width);
    std::vector<std::vector<uint8_t>> beam;
    for (uint32_t i = 0; i < n_bands; ++i) beam.push_back({(uint8_t)i});
    for (uint32_t depth = 2; depth <= count; ++depth) {
        std::set<std::vector<uint8_t>> seen;
        std::vector<uint8_t> cand;
        for (const auto &c : beam)
            for (uint32_t i = 0; i < n_bands; ++i) {
                if (std::find(c.begin(), c.end(), (uint8_t)i) != c.end()) continue;
                std::vector<uint8_t> e = c;
                e.push_back((uint8_t)i);
                std::sort(e.begin(), e.end());
                if (seen.insert(e).second) cand.insert(cand.end(), e.begin(), e.end());
            }
        const std::vector<double> s = sc.screen(cand, depth, idx);
        const std::vector<size_t> o = order_by(s);
        const size_t keep = std::min<size_t>(beam_width, o.size());
        std::printf("beam depth %u: evaluated %zu, keeping %zu\n", depth, o.size(), keep);
        std::fflush(stdout);
        std::vector<std::vector<uint8_t>> next;
        for (size_t i = 0; i < keep; ++i) next.emplace_back(cand.begin() + o[i] * depth, cand.begin() + (o[i] + 1) * depth);
        beam.swap(next);
    }
    // local search from every beam member (main.rs:641-691), in lock step
    std::vector<std::vector<uint8_t>> cur = beam;
    std::vector<double> cur_score;
    {
        std::vector<uint8_t> flat;
        for (const auto &c : cur) flat.insert(flat.end(), c.begin(), c.end());
        cur_score = sc.screen(flat, count, idx);
    }
    std::vector<bool> active(cur.size(), true);
    for (;;) {
        std::vector<uint8_t> flat;
        std::vector<std::pair<size_t, size_t>> range(cur.size(), {0, 0});  // candidate range per member
        size_t nc = 0;
        for (size_t m = 0; m < cur.size(); ++m) {
            if (!active[m]) continue;
            range[m].first = nc;
            const auto &c = cur[m];
            for (uint32_t slot = 0; slot < count; ++slot)
                for (uint32_t i = 0; i < n_bands; ++i) {
                    bool clash = false;
                    for (uint32_t k = 0; k < count; ++k) clash = clash || (k != slot && c[k] == i);
                    if (clash) continue;
                    std::vector<uint8_t> e = c;
                    e[slot] = (uint8_t)i;
                    std::sort(e.begin(), e.end());
                    if (e == c) continue;
                    flat.insert(flat.end(), e.begin(), e.end());
                    ++nc;
                }
            range[m].second = nc;
        }
        if (nc == 0) break;
        const std::vector<double> s = sc.screen(flat, count, idx);
        bool moved = false;
        for (size_t m = 0; m < cur.size(); ++m) {
            if (!active[m]) continue;
            size_t best = (size_t)-1;
            double best_s = cur_score[m];
            for (size_t j = range[m].first; j < range[m].second; ++j)
                if (s[j] < best_s) {
                    best_s = s[j];
                    best = j;
                }
            if (best == (size_t)-1) {
                active[m] = false;
                continue;
            }
            cur[m].assign(flat.begin() + best * count, flat.begin() + (best + 1) * count);
            cur_score[m] = best_s;
            moved = true;
        }
        if (!moved) break;
    }
    std::set<std::vector<uint8_t>> seen;
    std::vector<uint8_t> improved;
    for (const auto &c : cur)
        if (seen.insert(c).second) improved.insert(improved.end(), c.begin(), c.end());
    const std::vector<double> s = sc.screen(improved, count, idx);
    const std::vector<size_t> o = order_by(s);
    const size_t keep = std::min<size_t>(keep_count, o.size());
    std::vector<uint8_t> out;
    scores->clear();
    for (size_t i = 0; i < keep; ++i) {
        out.insert(out.end(), improved.begin() + o[i] * count, improved.begin() + (o[i] + 1) * count);
        scores->push_back(s[o[i]]);
    }
    return out;
}

// ------------------------------------------------------------------ output

std::string fnum(double v, int prec)  // Rust's {:.prec}
{
    if (std::isnan(v)) return "NaN";
    if (std::isinf(v)) return v > 0 ? "inf" : "-inf";
    char b[64];
    std::snprintf(b, sizeof b, "%.*f", prec, v);
    return b;
}

std::string fexp(double v, int prec)  // Rust's {:.prec e}: mantissa, 'e', exponent without sign padding
{
    if (!std::isfinite(v)) return fnum(v, prec);
    char b[64];
    std::snprintf(b, sizeof b, "%.*e", prec, v);
    std::string s(b);
    size_t e = s.find('e');
    int ex = std::atoi(s.c_str() + e + 1);
    return s.substr(0, e) + "e" + std::to_string(ex);
}

std::string join(const Scored &r, const Dataset &d, const char *sep, bool weights, int prec)
{
    std::string s;
    for (size_t k = 0; k < r.combo.size(); ++k) {
        if (k) s += sep;
        s += weights ? fnum(r.weights[k], prec) : fnum(d.wl[r.combo[k]], 0);
    }
    return s;
}

std::string headers(const char *prefix, const char *suffix, size_t n)
{
    std::string s;
    for (size_t i = 0; i < n; ++i) s += (i ? "," : "") + std::string(prefix) + std::to_string(i) + suffix;
    return s;
}

FILE *open_out(const std::string &p)
{
    FILE *f = std::fopen(p.c_str(), "w");
    if (!f) throw std::runtime_error("cannot write " + p);
    return f;
}

std::string stats_cells(const Scored &r)
{
    return fnum(r.score, 9) + "," + fnum(r.mean, 9) + "," + fnum(r.rmse, 9) + "," + fnum(r.p50, 9) + "," +
           fnum(r.p90, 9) + "," + fnum(r.p95, 9) + "," + fnum(r.p99, 9) + "," + fnum(r.max, 9) + "," +
           std::to_string(r.samples);
}

void make_dirs(const std::string &p)
{
    for (size_t i = 1; i <= p.size(); ++i)
        if (i == p.size() || p[i] == '/') {
            std::string s = p.substr(0, i);
            if (::mkdir(s.c_str(), 0777) != 0 && errno != EEXIST) throw std::runtime_error("cannot create " + s);
        }
}

double since_ms(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

int main(int argc, char **argv)
{
    const auto t_start = std::chrono::steady_clock::now();
    Options o;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        auto next = [&]() -> std::string {
            if (i + 1 >= argc) {
                usage();
                std::exit(2);
            }
            return argv[++i];
        };
        auto num = [&]() { return (uint64_t)std::strtoull(next().c_str(), nullptr, 10); };
        if (a == "--asset-dir") o.asset_dirs.push_back(next());
        else if (a == "--primary-count") o.primary_count = (uint32_t)num();
        else if (a == "--screen-samples") o.screen_samples = num();
        else if (a == "--refine-top") o.refine_top = num();
        else if (a == "--top") o.top = num();
        else if (a == "--threads") o.threads = num();
        else if (a == "--beam-width") o.beam_width = num();
        else if (a == "--out-dir") o.out_dir = next();
        else if (a == "--device") o.device = (int)num();
        else if (a == "--weight-strategy") {
            std::string s = next();
            if (s == "voronoi") o.strategy = Voronoi;
            else if (s == "cmf-mass") o.strategy = CmfMass;
            else if (s == "cmf-moments") o.strategy = CmfMoments;
            else {
                usage();
                return 2;
            }
        } else {
            usage();
            return 2;
        }
    }
    if (o.asset_dirs.empty()) {
        usage();
        return 2;
    }
    sky_scan_dataset *ds = nullptr;
    try {
        if (o.primary_count < 1 || o.primary_count > SKY_SCAN_MAX_PRIMARIES)  // main.rs:243-248
            throw std::runtime_error("--primary-count must be in 1..=8");
        make_dirs(o.out_dir);
        Dataset d = load_dataset(o.asset_dirs);
        const std::vector<uint32_t> idx = screening_indices(d.n(), o.screen_samples);
        const uint64_t ncombo = combo_count(d.bands(), o.primary_count);
        const uint32_t threads =
            o.threads ? (uint32_t)o.threads : std::max(1u, std::thread::hardware_concurrency());
        std::printf("loaded %zu assets, %llu samples, %u bands, %llu %u-wavelength combos\n", d.assets.size(),
                    (unsigned long long)d.n(), d.bands(), (unsigned long long)ncombo, o.primary_count);
        std::printf("screening with %zu deterministic samples, refining top %llu on all samples, threads=%u, "
                    "weight_strategy=%s\n",
                    idx.size(), (unsigned long long)o.refine_top, threads, strategy_label(o.strategy));
        std::fflush(stdout);
        check(sky_scan_dataset_create(o.device, d.values.data(), d.bands(), d.n(), d.ref_oklab.data(), &ds));
        Scorer sc{d, ds, o.strategy, threads};
        const uint32_t count = o.primary_count;

        const auto t_screen = std::chrono::steady_clock::now();
        std::vector<uint8_t> screened;  // ordered by screening score
        if (ncombo <= kExhaustiveLimit) {
            const std::vector<uint8_t> combos = all_combos(d.bands(), count);
            const std::vector<double> s = sc.screen(combos, count, idx);
            for (size_t i : order_by(s)) screened.insert(screened.end(), combos.begin() + i * count, combos.begin() + (i + 1) * count);
        } else {
            std::printf("combo count exceeds exhaustive in-memory limit %llu; using deterministic beam_width=%llu\n",
                        (unsigned long long)kExhaustiveLimit, (unsigned long long)o.beam_width);
            std::vector<double> s;
            screened = beam_search(sc, d.bands(), count, idx, o.beam_width, o.refine_top, &s);
        }
        const double screen_ms = since_ms(t_screen);

        const auto t_refine = std::chrono::steady_clock::now();
        const uint64_t refine_count = std::min<uint64_t>(o.refine_top, screened.size() / count);
        std::vector<Scored> refined =
            sc.refine(std::vector<uint8_t>(screened.begin(), screened.begin() + refine_count * count), count, {});
        std::stable_sort(refined.begin(), refined.end(),
                         [](const Scored &a, const Scored &b) { return score_less(a.score, b.score); });
        const size_t top = std::min<size_t>(o.top, refined.size());
        refined.resize(top);

        // per-asset statistics of every ranked combo (main.rs:1237-1275)
        std::vector<uint8_t> top_flat;
        for (const Scored &r : refined) top_flat.insert(top_flat.end(), r.combo.begin(), r.combo.end());
        std::vector<std::vector<Scored>> per_asset;
        for (const Asset &a : d.assets) {
            std::vector<uint32_t> ai(a.count);
            std::iota(ai.begin(), ai.end(), (uint32_t)a.start);
            per_asset.push_back(sc.refine(top_flat, count, ai));
        }
        const double refine_ms = since_ms(t_refine);

        FILE *f = open_out(o.out_dir + "/ranking.csv");  // main.rs:1203-1235
        std::fprintf(f, "rank,%s,%s,score,mean,rmse,p50,p90,p95,p99,max,sample_count\n",
                     headers("lambda", "_nm", count).c_str(), headers("weight", "_nm", count).c_str());
        for (size_t r = 0; r < top; ++r)
            std::fprintf(f, "%zu,%s,%s,%s\n", r + 1, join(refined[r], d, ",", false, 0).c_str(),
                         join(refined[r], d, ",", true, 6).c_str(), stats_cells(refined[r]).c_str());
        std::fclose(f);

        f = open_out(o.out_dir + "/ranking_by_elevation.csv");
        std::fprintf(f, "rank,elevation_deg,%s,score,mean,rmse,p50,p90,p95,p99,max,sample_count\n",
                     headers("lambda", "_nm", count).c_str());
        for (size_t r = 0; r < top; ++r)
            for (size_t a = 0; a < d.assets.size(); ++a)
                std::fprintf(f, "%zu,%s,%s,%s\n", r + 1, fnum(d.assets[a].sun_elevation_deg, 6).c_str(),
                             join(refined[r], d, ",", false, 0).c_str(), stats_cells(per_asset[a][r]).c_str());
        std::fclose(f);

        f = open_out(o.out_dir + "/report.md");  // main.rs:1277-1402
        std::fprintf(f, "# Sky-View %u-Wavelength Search\n\n## Dataset\n\n", count);
        std::fprintf(f, "- Assets: %zu\n- Total LUT samples: %llu\n", d.assets.size(), (unsigned long long)d.n());
        std::fprintf(f, "- Wavelength grid: %u bands, %s..%s nm\n", d.bands(), fnum(d.wl.front(), 0).c_str(),
                     fnum(d.wl.back(), 0).c_str());
        std::fprintf(f, "- Weight strategy: `%s`\n- Weight strategy description: %s\n", strategy_label(o.strategy),
                     strategy_description(o.strategy));
        std::fprintf(f, "- Reference Y p99 scale: %s\n", fexp(d.y_p99, 9).c_str());
        std::fprintf(f, "- Screening samples: %zu\n- Full-resolution refined candidates: %llu\n\n", idx.size(),
                     (unsigned long long)refine_count);
        std::fprintf(f, "| Elevation | SPP | Dimensions | Samples | Asset |\n| ---: | ---: | ---: | ---: | --- |\n");
        for (const Asset &a : d.assets)
            std::fprintf(f, "| %s | %llu | %llux%llu | %llu | `%s` |\n", fnum(a.sun_elevation_deg, 2).c_str(),
                         (unsigned long long)a.spp, (unsigned long long)a.width, (unsigned long long)a.height,
                         (unsigned long long)a.count, a.path.c_str());
        std::fprintf(f, "\n## Ranking\n\nScreening ranks all %u-wavelength combos by deterministic-subsample RMSE. "
                        "Final score is `RMSE_OK + 0.25 * P95_OK`, computed on all loaded LUT samples after direct "
                        "physical XYZ quadrature and Oklab conversion.\n\n",
                     count);
        std::fprintf(f, "| Rank | Wavelengths nm | Weights nm | Score | RMSE | Mean | P95 | P99 | Max |\n"
                        "| ---: | --- | --- | ---: | ---: | ---: | ---: | ---: | ---: |\n");
        for (size_t r = 0; r < top; ++r) {
            const Scored &s = refined[r];
            std::fprintf(f, "| %zu | %s | %s | %s | %s | %s | %s | %s | %s |\n", r + 1, join(s, d, ", ", false, 0).c_str(),
                         join(s, d, ", ", true, 1).c_str(), fnum(s.score, 6).c_str(), fnum(s.rmse, 6).c_str(),
                         fnum(s.mean, 6).c_str(), fnum(s.p95, 6).c_str(), fnum(s.p99, 6).c_str(), fnum(s.max, 6).c_str());
        }
        if (top > 0) {
            std::fprintf(f, "\n## Best Combo By Elevation\n\nBest overall combo: `%s nm`.\n\n",
                         join(refined[0], d, ", ", false, 0).c_str());
            std::fprintf(f, "| Elevation | Score | RMSE | Mean | P95 | P99 | Max |\n| ---: | ---: | ---: | ---: | ---: | ---: | ---: |\n");
            for (size_t a = 0; a < d.assets.size(); ++a) {
                const Scored &s = per_asset[a][0];
                std::fprintf(f, "| %s | %s | %s | %s | %s | %s | %s |\n", fnum(d.assets[a].sun_elevation_deg, 2).c_str(),
                             fnum(s.score, 6).c_str(), fnum(s.rmse, 6).c_str(), fnum(s.mean, 6).c_str(),
                             fnum(s.p95, 6).c_str(), fnum(s.p99, 6).c_str(), fnum(s.max, 6).c_str());
            }
            std::fprintf(f, "\nFull per-elevation metrics for every ranked combo are in `ranking_by_elevation.csv`. "
                            "Rank 2 through rank 11 are the top 10 runners-up when 11 results are requested.\n");
        }
        std::fclose(f);

        std::printf("top %zu:\n", top);
        for (size_t r = 0; r < top; ++r) {
            const Scored &s = refined[r];
            char rank[32];
            std::snprintf(rank, sizeof rank, "#%-2zu", r + 1);
            std::printf("%s %19snm score=%s rmse=%s p95=%s max=%s\n", rank, join(s, d, "/", false, 0).c_str(),
                        fnum(s.score, 6).c_str(), fnum(s.rmse, 6).c_str(), fnum(s.p95, 6).c_str(),
                        fnum(s.max, 6).c_str());
        }
        std::printf("wrote %s\n", o.out_dir.c_str());
        std::printf("timing: screen=%.1fms refine=%.1fms total=%.1fms gpu=%.1fms evaluations=%llu\n", screen_ms,
                    refine_ms, since_ms(t_start), sc.kernel_ms, (unsigned long long)sc.evaluations);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        if (ds) sky_scan_dataset_destroy(ds);
        return 1;
    }
    sky_scan_dataset_destroy(ds);
    return 0;
}
