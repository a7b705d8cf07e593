"""Tabulated medium (the reference's SceneData tables, data.rs:53-152,
medium.rs:28-155) on the CPU: the oracle's coefficients, optical depth and
majorants, and the library's host-side tables, against what pins them.

Pins:
  * the reference's own unit test (medium.rs:170-214): single-point profiles
    give sigma_R = 2.5e19 * 5.8e-31 * 1e9, sigma_O = 1e12 * 1e-21 * 1e5, and
    four species of 1e-6 g/m^3 at (1.0, 0.5) km^-1 per g/m^3;
  * numpy's np.interp (an independent piecewise-linear interpolator with the
    same end clamping) for the coefficients at any altitude;
  * the trapezoid rule, exact for a function linear between breakpoints, for
    the vertical optical depth;
  * the ozone column of the committed AFGL profile in Dobson units, and the
    sun radiance E / (2 pi (1 - cos r)) of data.rs:30-33;
  * the counting search (number of breakpoints at or below h) for the
    kernel's two interval lookups, on the library's own host-side tables.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle as O
import pysky
from pysky import tables as T
from pysky.api import InvalidConfig

DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_data")


@pytest.fixture(scope="module")
def ref():
    centers, tab = T.load_tables(DATA)
    return centers, tab


def _params(centers, tab, **kw):
    return O.Params(lambda_nm=list(centers), weight_nm=[1.0] * len(centers), tables=tab,
                    atmosphere_radius_km=6480.0, ground_albedo=0.18, sun_angular_radius_rad=0.00465047, **kw)


def _np_coeffs(centers, tab, b, h):
    """The medium from the raw tables with np.interp."""
    a, q = tab.atmosphere, tab.aerosol
    air = np.interp(h, a[:, 0], a[:, 1])
    o3 = np.interp(h, a[:, 0], a[:, 2])
    m = np.array([np.interp(h, q[:, 0], q[:, 1 + s]) for s in range(4)])
    xs = 5.8e-31 * (550.0 / centers[b]) ** 4
    return np.array([air * xs * 1e9, m @ tab.aerosol_optics[b, :, 0], o3 * tab.ozone_cross_section_cm2[b] * 1e5,
                     m @ tab.aerosol_optics[b, :, 1]])


def test_loader_reads_the_reference_tables(ref):
    centers, tab = ref
    np.testing.assert_array_equal(centers, np.arange(380, 781, 10))
    assert tab.atmosphere.shape == (50, 3) and tab.aerosol.shape == (14, 5)
    assert np.all(np.diff(tab.atmosphere[:, 0]) > 0) and np.all(np.diff(tab.aerosol[:, 0]) > 0)
    assert tab.atmosphere[0, 0] == 0.0 and tab.atmosphere[-1, 0] == 120.0
    assert tab.aerosol_optics.shape == (41, 4, 2) and np.all(tab.aerosol_optics >= 0)
    assert tab.solar_irradiance_w_m2[0] == pytest.approx(11.1116966)


def test_reference_unit_test_values():
    tab = T.MediumTables(atmosphere=np.array([[0.0, 2.5e19, 1.0e12]]),
                         aerosol=np.array([[0.0, 1e-6, 1e-6, 1e-6, 1e-6]]),
                         aerosol_optics=np.array([[[1.0, 0.5]] * 4]), ozone_cross_section_cm2=np.array([1e-21]))
    p = O.Params(lambda_nm=[550.0], tables=tab, atmosphere_radius_km=6480.0)
    sr, sas, so, saa = O.coefficients(p, 0, 0.0)
    assert sr == pytest.approx(2.5e19 * 5.8e-31 * 1e9, rel=1e-12)
    assert so == pytest.approx(1e12 * 1e-21 * 1e5, rel=1e-12)
    assert sas == pytest.approx(4e-6, rel=1e-12) and saa == pytest.approx(2e-6, rel=1e-12)
    assert sr + sas + so + saa >= sr + sas  # extinction >= scattering
    # single points clamp: the same values anywhere in the atmosphere
    np.testing.assert_allclose(O.coefficients(p, 0, 77.0), [sr, sas, so, saa], rtol=1e-12)
    # the library's host table (sea-level row) agrees
    scene = pysky.SceneData(atmosphere_radius_km=6480.0, atmosphere=pysky.AtmosphereParams(tables=tab),
                            wavelengths_nm=[550.0])
    row = pysky.wavelength_table(scene)[0]
    np.testing.assert_allclose(row[[0, 2, 1, 3]], [sr, sas, so, saa], rtol=1e-6)


def test_coefficients_match_np_interp(ref):
    centers, tab = ref
    p = _params(centers, tab)
    rng = np.random.default_rng(2)
    for h in np.concatenate([rng.uniform(0, 120, 40), tab.atmosphere[:10, 0], tab.aerosol[:, 0], [0.0, 120.0]]):
        for b in (0, 17, 40):
            np.testing.assert_allclose(O.coefficients(p, b, h), _np_coeffs(centers, tab, b, h), rtol=1e-12,
                                       atol=1e-30)
    assert np.all(O.coefficients(p, 3, 121.0) == 0) and np.all(O.coefficients(p, 3, -0.1) == 0)


def test_vertical_optical_depth_is_the_trapezoid_sum(ref):
    centers, tab = ref
    p = _params(centers, tab)
    z = np.unique(np.concatenate([tab.atmosphere[:, 0], tab.aerosol[:, 0]]))
    z = z[(z >= 0) & (z <= 120)]
    for b in (0, 20, 40):
        sig = np.array([_np_coeffs(centers, tab, b, h).sum() for h in z])
        trap = float(np.sum(0.5 * (sig[1:] + sig[:-1]) * np.diff(z)))
        tau = O.optical_depth(p, b, [0.0, 6360.0, 0.0], [0.0, 1.0, 0.0], 120.0)
        assert tau == pytest.approx(trap, rel=1e-10)
    # the ozone column of the AFGL profile, in Dobson units (2.687e16 cm^-2)
    o3 = np.sum(0.5 * (tab.atmosphere[1:, 2] + tab.atmosphere[:-1, 2]) * np.diff(tab.atmosphere[:, 0])) * 1e5
    assert 250 < o3 / 2.687e16 < 450


def test_majorants_dominate_and_match(ref):
    centers, tab = ref
    for layers, growth in ((64, 1.0), (3, 4.0)):
        p = _params(centers, tab, majorant_layers=layers, shell_growth=growth)
        want = O.majorant_grid(p)
        scene = T.load_reference_scene(DATA, 5.0)
        got = pysky.majorant_grid(scene, layers=layers, shell_growth=growth)
        np.testing.assert_allclose(got, want.reshape(got.shape), rtol=1e-6)
        hs = O.shell_heights(p)
        for b in (0, 40):
            for k in range(layers):
                for h in np.linspace(hs[k], hs[k + 1], 41):
                    assert O.coefficients(p, b, h).sum() <= want.reshape(41, layers)[b, k] * (1 + 1e-12)


def test_sun_radiance_is_irradiance_over_solid_angle(ref):
    centers, tab = ref
    scene = T.load_reference_scene(DATA, 5.0)
    rows = pysky.wavelength_table(scene)
    omega = 2 * np.pi * (1 - np.cos(0.00465047))
    np.testing.assert_allclose(rows[:, 4], tab.solar_irradiance_w_m2 / omega, rtol=1e-6)
    lam = O.lambda_table(_params(centers, tab))
    np.testing.assert_allclose(lam[:, 4], tab.solar_irradiance_w_m2 / omega, rtol=1e-9)  # 1 - cos r cancels


def test_invalid_tables_are_rejected(ref):
    centers, tab = ref
    import dataclasses
    bad = dataclasses.replace(tab, atmosphere=tab.atmosphere[::-1].copy())
    scene = T.load_reference_scene(DATA, 5.0)
    scene.atmosphere.tables = bad
    with pytest.raises(InvalidConfig):
        pysky.majorant_grid(scene)
    opt = tab.aerosol_optics.copy()
    opt[3, 1, 0] = -1.0
    scene.atmosphere.tables = dataclasses.replace(tab, aerosol_optics=opt)
    with pytest.raises(InvalidConfig):
        pysky.wavelength_table(scene)


def test_interval_lookups_match_the_counting_search(tmp_path):
    """The kernel's one-load cell lookup and its index-and-walk search, run on
    the host over tables built by the library's build_medium_table
    (tests/cell_lookup_check.cpp), pick the interval the counting search picks,
    with the same start and reciprocal width, at random altitudes and next to
    every breakpoint and cell boundary.  Tables whose breakpoints sit on cell
    boundaries, with FP32 n / top rounded either way, catch a cell range built
    without its rounding margin (the host's own check then drops the cells,
    which the cell counts catch)."""
    hipcc = os.environ.get("HIPCC") or shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not found")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = str(tmp_path / "cell_lookup_check")
    subprocess.run([hipcc, "-O2", "-std=c++17", "-w", "-o", exe,
                    os.path.join(root, "tests", "cell_lookup_check.cpp")], check=True)
    out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    rows = {}
    for line in out.strip().splitlines():
        name, *fields = line.split()
        rows[name] = {k: int(v) for k, v in (f.split("=") for f in fields)}
    expect_cells = {"reference": 1024, "single_point": 1024, "cell_boundaries_top_100": 1024,
                    "cell_boundaries_top_120": 1024, "fine_25m": 2048, "dense_cluster": 0,
                    "irregular_top_120": 1024}
    assert set(rows) == set(expect_cells)
    for name, r in rows.items():
        assert r["bad"] == 0, (name, r)
        assert r["cells"] == expect_cells[name], (name, r)
        assert r["tested"] > 300000, (name, r)
    assert rows["reference"]["K"] == 50
