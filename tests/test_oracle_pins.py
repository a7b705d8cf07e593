"""Pins the CPU oracle against the reference's known answers, golden vectors
and closed forms (CPU only).

Each check is chosen so that a plausible slip in the restatement (a dropped
term, a swapped operand, a wrong unit) fails it:
  * RNG: golden vectors from an independent restatement of sky_trace.wgsl:156-189;
  * CIE: the reference's own CIE 1931 table (data/CIE_xyz_1931_2deg.csv);
  * Planck: Wien's displacement law and the Stefan-Boltzmann integral;
  * geometry: the reference's KATs (geometry.rs:101-117, top moved to 6460 km),
    the literal reference boundary/segment logic, closed forms;
  * medium/majorants: closed forms and the invariant of data.rs:253-276;
  * optical depth: vertical closed forms, mpmath quadrature on slant rays;
  * sampling: the KATs of sampling.rs:234-274 and density checks.
"""
import json
import math
import os

import numpy as np
import pytest

import oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _json(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


# ------------------------------------------------------------------ RNG

def test_pcg4d_golden_vectors():
    g = _json("rng_vectors.json")
    for v in g["pcg4d"]:
        assert list(O.pcg4d(v["in"])) == v["out"]
    for v in g["rng"]:
        assert O.rng_u32(v["pixel"], v["sample"], v["band"], v["dim"], v["seed"]) == v["u32"]
        assert O.rng_f32(v["pixel"], v["sample"], v["band"], v["dim"], v["seed"]) == v["f32"]


def test_rng_f32_range_and_mean():
    vals = np.array([O.rng_f32(p, 3, 1, d, 1) for p in range(64) for d in range(64)])
    assert vals.min() >= 0.0 and vals.max() < 1.0
    assert abs(vals.mean() - 0.5) < 0.02  # 4096 draws: sd 0.0045


# ------------------------------------------------------------ colorimetry

def test_cmf_fit_matches_reference_cie_table():
    c = _json("cie_1931.json")
    rows = np.array(c["every_5nm"])
    fit = np.array([O.cmf_xyz(l) for l in rows[:, 0]])
    err = np.abs(fit - rows[:, 1:])
    # published accuracy of the multi-lobe fit is ~1e-2 absolute
    assert err.max(axis=0).max() < 0.03
    np.testing.assert_allclose(fit.sum(0) * 5, rows[:, 1:].sum(0) * 5, rtol=5e-3)
    for key in ("grid16", "grid64"):
        g = np.array(c[key])
        fg = np.array([O.cmf_xyz(l) for l in g[:, 0]])
        assert np.abs(fg - g[:, 1:]).max() < 0.03


def test_cmf_y555_is_one():
    """spectrum.rs:374-381: y(555) = 1 +- 0.01."""
    assert abs(_json("cie_1931.json")["y_555"] - 1.0) < 0.01
    assert abs(O.cmf_xyz(555.0)[1] - 1.0) < 0.01


def test_planck_wien_and_stefan_boltzmann():
    T = 5778.0
    lam = np.linspace(200.0, 3000.0, 28001)
    B = np.array([O.planck(l, T) for l in lam])
    peak = lam[B.argmax()]
    assert abs(peak - 2.897771955e6 / T) < 0.2  # Wien, nm
    # integral over all wavelengths = sigma T^4 / pi (W m^-2 sr^-1)
    from scipy.integrate import quad

    total, _ = quad(lambda l: O.planck(l, T), 10.0, 1e6, limit=500, points=[500.0, 2000.0])
    sigma = 5.670374419e-8
    assert abs(total / (sigma * T ** 4 / math.pi) - 1.0) < 1e-6


def test_lambda_table_units():
    p = O.Params(lambda_nm=[550.0, 600.0])
    t = O.lambda_table(p)
    # km^-1 = 1000 * m^-1 at the reference wavelength/height
    assert t[0, 0] == pytest.approx(1.3558e-2, rel=1e-12)
    assert t[1, 1] == pytest.approx(1.9e-3, rel=1e-12)  # ozone peak at 600 nm
    assert t[0, 2] == pytest.approx(3.996e-3, rel=1e-12)
    assert t[0, 3] == pytest.approx(4.4e-3, rel=1e-12)
    assert t[1, 2] / t[0, 2] == pytest.approx((600 / 550) ** -1.3, rel=1e-12)
    assert t[0, 4] == pytest.approx(O.planck(550.0, 5778.0), rel=1e-12)


# ---------------------------------------------------------------- geometry

def _obs(h=0.2):
    return np.array([0.0, 6360.0 + h, 0.0])


def test_reference_geometry_kats():
    """geometry.rs:101-117 with the 6460 km top: up exits at 99.8 km, down
    hits the ground at 0.2 km."""
    p = O.Params()
    hit, t, kind = O.next_boundary_reference(p, _obs(), [0, 1, 0])
    assert hit and kind == 1 and abs(t - 99.8) < 0.01
    hit, t, kind = O.next_boundary_reference(p, _obs(), [0, -1, 0])
    assert hit and kind == 0 and abs(t - 0.2) < 0.01


def test_walk_segments_match_reference_segments():
    """The shell walk gives the segments of sky_trace.wgsl:284-294 (up to the
    reference's 5 m guard) on random and grazing rays."""
    p = O.Params()
    rng = np.random.default_rng(11)
    dirs = [rng.normal(size=3) for _ in range(300)]
    # grazing: nearly horizontal from several heights
    for e in (-0.02, -0.005, 0.0, 0.003, 0.03):
        dirs.append(np.array([math.cos(e), math.sin(e), 0.0]))
    checked = 0
    for i, d in enumerate(dirs):
        d = d / np.linalg.norm(d)
        h = [0.2, 1.0, 24.9, 37.3, 80.0][i % 5]
        o = _obs(h)
        hit, tmax, kind = O.next_boundary_reference(p, o, d)
        e_ref, l_ref = O.segments(p, o, d, tmax, reference=True)
        e_w, l_w = O.segments(p, o, d, tmax, reference=False)
        # drop degenerate walk segments shorter than the reference guard
        keep = np.diff(np.concatenate([[0.0], e_w])) > 0.006
        e_w, l_w = e_w[keep], l_w[keep]
        keep_r = np.diff(np.concatenate([[0.0], e_ref])) > 0.006
        e_ref, l_ref = e_ref[keep_r], l_ref[keep_r]
        assert len(e_ref) == len(e_w), (i, e_ref, e_w)
        np.testing.assert_allclose(e_w, e_ref, rtol=1e-9, atol=1e-7)
        assert np.array_equal(l_w, l_ref)
        checked += 1
    assert checked == len(dirs)


def test_grazing_boundary_closed_form():
    """Horizontal ray at height h: exit at sqrt(Rt^2 - (R+h)^2), no ground hit
    (FP64)."""
    p = O.Params()
    for h in (0.0001, 0.2, 5.0, 50.0):
        hit, t, kind = O.next_boundary_reference(p, _obs(h), [1, 0, 0])
        assert kind == 1
        assert t == pytest.approx(math.sqrt(6460.0 ** 2 - (6360.0 + h) ** 2), rel=1e-12)


def _exact_setup(p, o, d, e):
    """next_boundary and the walk's first shell in double for the FP32
    tracer's view of a vertex: the position rounded to FP32, r^2 - R^2 of the
    double position rounded to FP32, the FP32 direction renormalised."""
    R = p.ground_radius_km
    Rt = p.atmosphere_radius_km
    p32 = np.asarray(o, np.float32).astype(np.float64)
    w = float(np.float32(np.dot(o, o) - R * R))
    d32 = np.asarray(d, np.float32).astype(np.float64)
    d64 = d32 / np.linalg.norm(d32)
    b0 = float(np.dot(p32, d64))
    q = w - b0 * b0
    b = b0 + e
    wo = w + e * (2.0 * b0 + e)
    eps = 0.005
    tg = -1.0
    if q <= 0.0:
        t0, t1 = -b - math.sqrt(-q), -b + math.sqrt(-q)
        tg = t0 if t0 > eps else (t1 if t1 > eps else -1.0)
    t_exit = max(-b + math.sqrt((Rt - R) * (Rt + R) - q), eps)
    kind, t = (0, tg) if (tg > eps and tg < t_exit) else (1, t_exit)
    H = O.shell_heights(p)
    D = H * (2.0 * R + H)
    layer = int(np.sum(D[1:-1] <= wo))
    return dict(kind=kind, t_max=t, layer=layer, b=b, q=q, wo=wo)


def test_fp32_ray_setup_matches_double_geometry():
    """The FP32 tracer's ray set-up (b from the FP32 vertex, q = w - b^2 from
    the carried r^2 - R^2, the stable near ground root) against the same
    geometry in double: b within its stated rounding bound 3 2^-24 |p|, q
    within the grazing tolerance, the same kind and shell, roots within a
    metre, and the near ground root from a vertex metres above the ground to
    1e-5 relative (the unstable -b - sqrt(-q) is off by several % there)."""
    p = O.Params()
    R = p.ground_radius_km
    rng = np.random.default_rng(5)
    n_checked = 0
    for h in (0.0, 0.005, 0.2, 1.0, 5.3, 25.7, 80.1, 99.9):
        for i in range(60):
            n = rng.normal(size=3)
            o = n / np.linalg.norm(n) * (R + h)
            d = rng.normal(size=3)
            if i % 4 == 0:  # steeply down
                d = -o / np.linalg.norm(o) + 0.05 * d
            d = d / np.linalg.norm(d)
            for e in (0.0, 0.005):
                got, graze = O.ray_setup_f32(p, o, d, e)
                ref = _exact_setup(p, o, d, e)
                bound = 3 * 2.0 ** -24 * np.linalg.norm(o)
                assert abs(got["b"] - ref["b"]) <= bound + 1e-9, (h, i, e)
                ab = abs(ref["b"] - e)
                assert abs(got["q"] - ref["q"]) <= ab * (2.4e-7 * ab + 2.4e-3) + 1e-5, (h, i, e)
                assert got["kind"] == ref["kind"] and got["layer"] == ref["layer"], (h, i, e, got, ref)
                assert abs(got["t_max"] - ref["t_max"]) <= 1e-6 * ref["t_max"] + 1e-3, (h, i, e, got, ref)
                assert not graze
                n_checked += 1
    assert n_checked == 8 * 60 * 2
    # near ground root from vertices 7 and 20 m up (beyond the 5 m ray
    # epsilon): t0 = w_o / (sqrt(-q) - b)
    for h in (0.007, 0.02):
        for mu in (-1.0, -0.7, -0.3, -0.1):
            o = np.array([0.0, R + h, 0.0])
            d = np.array([math.sqrt(1.0 - mu * mu), mu, 0.0])
            got, _ = O.ray_setup_f32(p, o, d)
            ref = _exact_setup(p, o, d, 0.0)
            assert got["kind"] == 0 and abs(got["t_max"] / ref["t_max"] - 1.0) < 1e-5, (h, mu, got, ref)


def test_fp32_ray_setup_is_exact_at_ground_tangency():
    """Rays within centimetres of tangency to the ground take the FP64 branch,
    and hit or miss exactly as the double geometry says; the FP32 q alone
    (error ~2|b| 1.2 m) would decide about half of them wrongly."""
    p = O.Params()
    R = p.ground_radius_km
    rng = np.random.default_rng(9)
    n_graze = 0
    for h in (0.3, 2.0, 10.0, 40.0):
        for k in range(12):
            phi = rng.uniform(0, 2 * math.pi)
            up = np.array([0.0, 1.0, 0.0])
            o = (R + h) * np.array([math.sin(phi) * 0.6, math.cos(phi), math.sin(phi) * 0.8])
            u = o / np.linalg.norm(o)
            t = np.cross(u, up if abs(u[1]) < 0.9 else np.array([1.0, 0.0, 0.0]))
            t /= np.linalg.norm(t)
            # dip below the horizontal by the tangent angle, nudged by a few
            # centimetres of miss distance either way
            dip = math.acos(R / (R + h)) + rng.uniform(-1.0, 1.0) * 2e-5 / (R + h)
            d = math.cos(dip) * t - math.sin(dip) * u
            d32 = np.asarray(d, np.float32).astype(np.float64)
            ref = _exact_setup(p, o, d32, 0.0)
            if abs(ref["q"]) < 2e-3:
                continue  # too close to call in double from the rounded inputs
            got, graze = O.ray_setup_f32(p, o, d32)
            assert graze, (h, k, ref["q"])
            assert got["kind"] == ref["kind"], (h, k, ref["q"], got, ref)
            n_graze += 1
    assert n_graze >= 30


# ---------------------------------------------------------- medium, majorants

def test_medium_closed_form():
    p = O.Params(lambda_nm=[450.0])
    lam = O.lambda_table(p)[0]
    for h in (0.0, 1.2, 8.0, 10.0, 25.0, 33.0, 40.0, 99.0):
        sr, sas, so, saa = O.coefficients(p, 0, h)
        assert sr == pytest.approx(lam[0] * math.exp(-h / 8.0), rel=1e-14)
        assert sas == pytest.approx(lam[2] * math.exp(-h / 1.2), rel=1e-14)
        assert saa == pytest.approx(lam[3] * math.exp(-h / 1.2), rel=1e-14)
        assert so == pytest.approx(lam[1] * max(0.0, 1.0 - abs(h - 25.0) / 15.0), abs=1e-18)
        assert sr + sas + so + saa >= sr + sas  # medium.rs:211-212
    assert O.coefficients(p, 0, -0.01).sum() == 0.0 and O.coefficients(p, 0, 100.01).sum() == 0.0


def test_shell_heights():
    """Uniform shells are k top / L; growing shells are geometric with ratio
    g between consecutive thicknesses and still span [0, top]."""
    h = O.shell_heights(O.Params(majorant_layers=8))
    np.testing.assert_allclose(h, np.arange(9) * 12.5, rtol=0, atol=1e-12)
    h = O.shell_heights(O.Params(majorant_layers=5, shell_growth=2.5))
    assert h[0] == 0.0 and h[-1] == pytest.approx(100.0)
    np.testing.assert_allclose(np.diff(h)[1:] / np.diff(h)[:-1], 2.5, rtol=1e-12)


@pytest.mark.parametrize("layers,growth", [(64, 1.0), (4, 3.0), (7, 1.6)])
def test_majorants_dominate_extinction(layers, growth):
    """data.rs:253-276 on a dense grid: sigma_t <= majorant of its layer."""
    for kw in (dict(), dict(ozone_center_km=23.3), dict(aerosol_h_km=0.7, rayleigh_h_km=6.0)):
        p = O.Params(lambda_nm=list(O.midpoint_grid(64)), majorant_layers=layers, shell_growth=growth, **kw)
        maj = O.majorant_grid(p)
        lam = O.lambda_table(p)
        h = np.linspace(0.0, 100.0, 24001)
        edges = O.shell_heights(p)
        layer = np.clip(np.searchsorted(edges, h, side="right") - 1, 0, layers - 1)
        ro = np.maximum(0.0, 1.0 - np.abs(h - p.ozone_center_km) / p.ozone_half_width_km)
        for b in range(64):
            st = lam[b, 0] * np.exp(-h / p.rayleigh_h_km) + (lam[b, 2] + lam[b, 3]) * np.exp(-h / p.aerosol_h_km) + lam[b, 1] * ro
            assert np.all(st <= maj[b, layer] * (1 + 1e-12))
            # and the margin is the reference's 5 %: the bound is attained
            assert np.max(st / maj[b, layer]) > 1 / 1.05 - 1e-3


# ------------------------------------------------------------ optical depth

def test_optical_depth_vertical_closed_form():
    p = O.Params(lambda_nm=[400.0, 550.0, 700.0])
    lam = O.lambda_table(p)
    o = _obs(0.0)
    for b in range(3):
        tau = O.optical_depth(p, b, o, [0, 1, 0], 100.0)
        ref = lam[b, 0] * 8.0 * (1 - math.exp(-100 / 8.0)) + (lam[b, 2] + lam[b, 3]) * 1.2 * (1 - math.exp(-100 / 1.2)) \
            + lam[b, 1] * 15.0  # tent area = peak x half width x 2 / 2
        assert tau == pytest.approx(ref, rel=1e-10)


def test_optical_depth_slant_mpmath():
    mpmath = pytest.importorskip("mpmath")
    mpmath.mp.dps = 30
    p = O.Params(lambda_nm=[430.0])
    lam = O.lambda_table(p)[0]
    R = 6360.0
    for (h, elev) in ((0.2, 0.0), (0.2, 5.0), (3.0, -0.5), (12.0, 30.0), (0.2, -3.0)):
        e = math.radians(elev)
        o = np.array([0.0, R + h, 0.0])
        d = np.array([math.cos(e), math.sin(e), 0.0])
        hit, tmax, kind = O.next_boundary_reference(p, o, d)

        def sigma(t):
            x = mpmath.mpf(o[0]) + t * d[0]
            y = mpmath.mpf(o[1]) + t * d[1]
            alt = mpmath.sqrt(x * x + y * y) - R
            if alt < 0 or alt > 100:
                return mpmath.mpf(0)
            ro = max(0, 1 - abs(alt - 25) / 15)
            return lam[0] * mpmath.e ** (-alt / 8) + (lam[2] + lam[3]) * mpmath.e ** (-alt / 1.2) + lam[1] * ro

        # split at the tent kinks for the adaptive rule
        cuts = [0.0, tmax]
        b = float(np.dot(o, d))
        q = float(np.dot(o, o) - b * b - R * R)
        for hk in (10.0, 25.0, 40.0):
            disc = hk * (2 * R + hk) - q
            if disc > 0:
                for t in (-math.sqrt(disc) - b, math.sqrt(disc) - b):
                    if 0 < t < tmax:
                        cuts.append(t)
        cuts = sorted(cuts)
        ref = float(mpmath.quad(sigma, cuts))
        tau = O.optical_depth(p, 0, o, d, tmax)
        assert tau == pytest.approx(ref, rel=1e-7), (h, elev)


# ------------------------------------------------------------- sampling

def test_rayleigh_inverse_cdf_endpoints():
    """sampling.rs:269-274."""
    assert O.sample_rayleigh_mu(0.0) == pytest.approx(-1.0, abs=1e-6)
    assert O.sample_rayleigh_mu(0.5) == pytest.approx(0.0, abs=1e-6)
    assert O.sample_rayleigh_mu(1.0) == pytest.approx(1.0, abs=1e-6)


def test_rayleigh_inverse_is_the_cdf_inverse():
    # CDF of 3/8 (1 + mu^2) on [-1, 1]: (mu^3 + 3 mu + 4) / 8
    for xi in np.linspace(0.01, 0.99, 50):
        mu = O.sample_rayleigh_mu(xi)
        assert (mu ** 3 + 3 * mu + 4) / 8 == pytest.approx(xi, abs=1e-12)


def test_phase_functions_normalised():
    from scipy.integrate import quad

    assert 2 * math.pi * quad(O.phase_rayleigh, -1, 1)[0] == pytest.approx(1.0, rel=1e-10)
    for g in (-0.5, 0.0, 0.3, 0.8, 0.95):
        val = 2 * math.pi * quad(lambda m: O.phase_hg(g, m), -1, 1, points=[0.99], limit=200)[0]
        assert val == pytest.approx(1.0, rel=1e-8)
    assert O.phase_rayleigh(-1) > 0 and O.phase_rayleigh(0) > 0 and O.phase_rayleigh(1) > 0  # phase.rs:155-160


def test_hg_inverse_cdf_matches_density():
    """Inverse CDF against the HG CDF in closed form."""
    g = 0.8
    for xi in np.linspace(0.02, 0.98, 25):
        mu = O.sample_hg_mu(g, xi)
        # CDF of p(mu)*2pi on [-1, mu]
        cdf = (1 - g * g) / (2 * g) * (1 / math.sqrt(1 + g * g - 2 * g * mu) - 1 / (1 + g))
        assert cdf == pytest.approx(xi, abs=1e-10)


def test_sun_cone_sample_kats():
    """sampling.rs:234-249: sampled direction inside the cone, pdf finite > 0,
    pdf = 1 / solid angle."""
    p = O.Params(sun_elevation_deg=10.0)
    s = O.sun_direction(p)
    r = p.sun_angular_radius_rad
    for xi0, xi1 in ((0.25, 0.75), (0.0, 0.0), (0.999999, 0.5), (0.5, 0.1)):
        d, pdf, inside = O.sun_cone_dir(p, xi0, xi1)
        assert inside and np.dot(d, s) >= math.cos(r) - 1e-12
        assert pdf == pytest.approx(1 / (2 * math.pi * (1 - math.cos(r))), rel=1e-9)
    # xi0 -> 1 reaches the rim
    d, _, _ = O.sun_cone_dir(p, 0.99999994, 0.3)
    assert math.acos(min(1.0, np.dot(d, s))) == pytest.approx(r, rel=1e-4)


def test_cosine_hemisphere_kats():
    """sampling.rs:251-258: above the surface, pdf = cos / pi."""
    for n in ([0, 1, 0], [0.3, 0.2, -0.93], [0, 0, -1]):
        n = np.array(n) / np.linalg.norm(n)
        for u, xi in ((0.1, 0.2), (0.5, 0.9), (0.99, 0.01)):
            d, pdf = O.cosine_dir(n, u, xi)
            assert np.dot(d, n) > 0
            assert pdf == pytest.approx(np.dot(d, n) / math.pi, abs=1e-12)
            assert np.dot(d, n) == pytest.approx(math.sqrt(1 - u), abs=1e-12)


def test_frame_is_orthonormal():
    rng = np.random.default_rng(5)
    for _ in range(50):
        a = rng.normal(size=3)
        a /= np.linalg.norm(a)
        mu = rng.uniform(-1, 1)
        d = O.frame_dir(a, mu, math.sqrt(1 - mu * mu), rng.uniform(0, 2 * math.pi))
        assert np.linalg.norm(d) == pytest.approx(1.0, abs=1e-12)
        assert np.dot(d, a) == pytest.approx(mu, abs=1e-12)


def test_panorama_camera_closed_form():
    """sky_trace.wgsl:614-623: row centre v=0.5 looks at the horizon, v=0 at
    the zenith; u=0.5 looks along +z."""
    p = O.Params()
    assert np.allclose(O.camera_dir(p, 0.5, 0.5), [0, 0, 1], atol=1e-15)
    assert np.allclose(O.camera_dir(p, 0.25, 0.0), [0, 1, 0], atol=1e-12)
    assert np.allclose(O.camera_dir(p, 0.75, 0.5), [1, 0, 0], atol=1e-12)


@pytest.mark.parametrize("h_obs", [0.2, 5.0])
@pytest.mark.parametrize("azimuth", [0.0, 37.0, 200.0])
def test_sky_view_lut_camera_closed_form(h_obs, azimuth):
    """sky_view_lut_dir (sky_trace.wgsl:625-681) pinned by its geometry.
    Pixel centres map to uv = (x/(W-1), y/(H-1)); with H = 9 the rows are
    v = 0, 1/8, ..., 1:
      * row 0 looks at the zenith and the last row at the nadir;
      * row 6 (v = 0.75, the sky/ground split) is the geometric horizon: the
        ray is tangent to the ground sphere (impact parameter R);
      * rows above it see the sky (the ray leaves the atmosphere), rows below
        hit the ground, and the zenith angle grows down the rows;
      * column u = 0 faces the sun's azimuth, u = 1 faces away, and
        u = 1/sqrt(2) (cos = 1 - 2u^2 = 0) faces side = up x forward;
      * the non-linear parametrisation: zenith angle (pi - beta)(1 - (1 - v/0.75)^2)
        in the sky, pi - beta + beta ((v - 0.75)/0.25)^2 below, azimuth
        acos(1 - 2 u^2) from the sun."""
    W, H = 3, 9
    p = O.Params(width=W, height=H, projection=O.PROJ_SKY_VIEW_LUT, observer_altitude_km=h_obs,
                 sun_azimuth_deg=azimuth, sun_elevation_deg=10.0)
    R = p.ground_radius_km
    vh = R + h_obs
    o = np.array([0.0, vh, 0.0])
    a = math.radians(azimuth)
    fw = np.array([math.sin(a), 0.0, math.cos(a)])
    side = np.array([math.cos(a), 0.0, -math.sin(a)])  # (0,1,0) x fw

    def d(x, y, u=None):
        return O.camera_dir(p, (x + 0.5) / W if u is None else u, (y + 0.5) / H)

    for x in range(W):
        assert np.allclose(d(x, 0), [0, 1, 0], atol=1e-12)
        assert np.allclose(d(x, H - 1), [0, -1, 0], atol=1e-12)
        hz = d(x, 6)
        assert np.linalg.norm(np.cross(o, hz)) == pytest.approx(R, rel=1e-12)
        assert hz[1] == pytest.approx(-math.sqrt(vh * vh - R * R) / vh, rel=1e-12)
    zen = [math.acos(np.clip(d(1, y)[1], -1, 1)) for y in range(H)]
    assert all(z1 < z2 for z1, z2 in zip(zen[:-1], zen[1:]))
    for y in (1, 2, 3, 4, 5, 7):  # row 6 is tangent: checked above
        hit, t, kind = O.next_boundary_reference(p, o, d(1, y))
        assert hit and kind == (1 if y < 6 else 0), (y, kind)
    # a dense sweep of v: sky exactly above the 0.75 split, ground below it
    for vv in np.linspace(0.005, 0.995, 199):
        if abs(vv - 0.75) < 1e-3:
            continue
        hit, t, kind = O.next_boundary_reference(p, o, O.camera_dir(p, 0.5, vv * (H - 1) / H + 0.5 / H))
        assert hit and kind == (1 if vv < 0.75 else 0), (vv, kind)
    beta = math.acos(math.sqrt(vh * vh - R * R) / vh)
    assert zen[3] == pytest.approx((math.pi - beta) * 0.75, rel=1e-12)  # v = 0.375
    assert zen[7] == pytest.approx(math.pi - beta + beta * 0.25, rel=1e-12)  # v = 0.875
    for y in (2, 7):
        toward, away = d(0, y), d(W - 1, y)
        hor_t = toward - toward[1] * np.array([0, 1, 0])
        hor_a = away - away[1] * np.array([0, 1, 0])
        assert np.allclose(hor_t / np.linalg.norm(hor_t), fw, atol=1e-12)
        assert np.allclose(hor_a / np.linalg.norm(hor_a), -fw, atol=1e-12)
        u_side = (1 / math.sqrt(2)) * (W - 1) / W + 0.5 / W
        s = d(0, y, u=u_side)
        hor_s = s - s[1] * np.array([0, 1, 0])
        assert np.allclose(hor_s / np.linalg.norm(hor_s), side, atol=1e-9)
        mid = d(1, y)  # u = 0.5: cos(azimuth from the sun) = 0.5
        hor_m = mid - mid[1] * np.array([0, 1, 0])
        hor_m /= np.linalg.norm(hor_m)
        assert hor_m @ fw == pytest.approx(0.5, abs=1e-12)
        assert hor_m @ side == pytest.approx(math.sqrt(0.75), abs=1e-12)
