// Host-side check of the tabulated medium's interval lookups (built and run by
// tests/test_tabulated_medium.py; no GPU needed).  For merged breakpoint grids
// built by the library's own build_medium_table, the one-load cell lookup
// (the kernel's tab_medium, restated on the host as interval_by_cell), the
// index-and-walk search (tab_interval) and the counting search must give the
// same interval, start and reciprocal width at random altitudes and at every
// breakpoint and cell boundary with their 64 nearest FP32 neighbours on each
// side.  Prints one line per table: name, breakpoints, cells, probes,
// mismatches.
#include "../sky-tracer_amd/csrc/sky_scene.cpp"

#include <cstdio>
#include <random>
#include <vector>

using namespace sky;

// tab_interval (sky_device.hpp) on the host.
static uint32_t interval_by_index(const MediumTable &m, float h)
{
    const uint32_t K = (uint32_t)m.z.size(), n = (uint32_t)m.idx.size();
    uint32_t j = (uint32_t)(h * m.inv_dh);
    j = j < n ? j : n - 1;
    uint32_t k = m.idx[j];
    while (k + 2 < K && h >= m.z[k + 1]) ++k;
    while (k > 0 && h < m.z[k]) --k;
    return k;
}

static void check(const char *name, const std::vector<double> &alt, double top)
{
    // one atmosphere row per altitude, one aerosol row at 0 km
    std::vector<double> atm, aer{0.0, 1e-6, 1e-6, 1e-6, 1e-6};
    for (double z : alt) {
        atm.push_back(z);
        atm.push_back(2.5e19);
        atm.push_back(1e12);
    }
    const double optics[8] = {1, 0.5, 1, 0.5, 1, 0.5, 1, 0.5}, o3xs = 1e-21, lam = 550.0, w = 1.0;
    sky_medium_tables t{};
    t.atmosphere_points = (uint32_t)alt.size();
    t.aerosol_points = 1;
    t.atmosphere = atm.data();
    t.aerosol = aer.data();
    t.aerosol_optics = optics;
    t.ozone_cross_section_cm2 = &o3xs;
    sky_atmosphere a{};
    a.tables = &t;
    sky_wavelengths wl{1, &lam, &w};
    MediumTable m;
    build_medium_table(&a, &wl, top, m);

    const float topf = (float)top;
    const size_t n = m.cell.size() / 8;
    std::vector<float> probe;
    std::mt19937 rng(12345);
    std::uniform_real_distribution<float> U(0.0f, topf);
    for (int i = 0; i < 200000; ++i) probe.push_back(U(rng));
    std::vector<float> centres(m.z.begin(), m.z.end());
    for (size_t j = 0; j <= n; ++j) centres.push_back((float)(j * (top / (double)n)));
    for (size_t j = 0; j <= m.idx.size(); ++j) centres.push_back((float)(j * (top / (double)m.idx.size())));
    for (float c : centres) {
        float up = c, dn = c;
        for (int s = 0; s < 64; ++s) {
            probe.push_back(up);
            probe.push_back(dn);
            up = std::nextafter(up, INFINITY);
            dn = std::nextafter(dn, -INFINITY);
        }
    }
    long tested = 0, bad = 0;
    for (float h : probe) {
        if (!(h >= 0.0f && h <= topf)) continue;
        ++tested;
        const uint32_t kc = interval_by_count(m.z, h), ki = interval_by_index(m, h);
        bool ok = ki == kc;
        if (n) {
            float z0, inv;
            const uint32_t k = interval_by_cell(m, h, z0, inv);
            ok = ok && k == kc && z0 == m.z[kc] && inv == m.inv[kc];
        }
        bad += ok ? 0 : 1;
    }
    std::printf("%s K=%zu cells=%zu tested=%ld bad=%ld\n", name, m.z.size(), n, tested, bad);
}

int main()
{
    // the reference's merged grid: AFGL every km to 25, then 2.5 and 5 km steps
    std::vector<double> ref;
    for (int i = 0; i <= 25; ++i) ref.push_back(i);
    for (double z = 27.5; z <= 50.0; z += 2.5) ref.push_back(z);
    for (double z = 55.0; z <= 120.0; z += 5.0) ref.push_back(z);
    check("reference", ref, 120.0);  // atmosphere radius 6480 km over a 6360 km ground
    check("single_point", {0.0}, 100.0);
    // every 1024-cell boundary is a breakpoint; fl(1024 / top) rounds down for
    // top 100 and up for top 120, so altitudes next to a boundary land on both
    // sides of it
    for (double top : {100.0, 120.0}) {
        std::vector<double> b;
        for (int j = 0; j <= 1024; ++j) b.push_back(j * (top / 1024));
        check(top == 100.0 ? "cell_boundaries_top_100" : "cell_boundaries_top_120", b, top);
    }
    std::vector<double> f;  // 25 m spacing: needs 2048 cells
    for (int j = 0; j <= 4000; ++j) f.push_back(j * 0.025);
    check("fine_25m", f, 100.0);
    check("dense_cluster", {0.0, 10.0, 10.0001, 10.0002, 10.0003, 50.0}, 100.0);
    check("irregular_top_120", {0.0, 0.3, 7.77, 33.3333333, 64.0, 99.99999, 119.9}, 120.0);
    return 0;
}
