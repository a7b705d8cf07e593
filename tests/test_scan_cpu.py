"""Sparse-wavelength combo scan, CPU side: the oracle and the library's host
functions (band widths, quadrature weights, CIE table) against what pins them.

Pins:
  * the reference's own test vectors (main.rs:1437-1479): Voronoi weights
    (75, 75, 95, 165) for bands (4, 10, 19, 29) of the 41-band grid; CMF-fit
    weights non-negative with fixed sum; unused slots zero;
  * the weights the reference printed in its committed optimizer reports
    (tests/golden/optimizer_weights.json: 55 combos over Voronoi, CMF-mass
    and CMF-moments) to the printed 6 decimals;
  * Oklab: the XYZ -> Oklab pairs published with the colour space
    (B. Ottosson, 2020);
  * the CIE table against the reference's CSV (tests/golden/cie_1931.json);
  * percentiles against sorting (index round((n-1) q) over finite values);
  * the all-band combo with Voronoi weights reproduces the reference colour
    exactly (error 0), so a dropped width, weight or scale term fails.
"""
import json
import os

import numpy as np
import pytest

import oracle as O
from pysky import scan

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
GRID = np.arange(380, 781, 10).astype(np.float32)
STRAT = {"voronoi": 0, "cmf-mass": 1, "cmf-moments": 2}


def _cie_fixture(grid):
    rows = np.array(json.load(open(os.path.join(GOLD, "cie_1931.json")))["every_5nm"])
    return np.array([rows[rows[:, 0] == l][0, 1:] for l in grid]).astype(np.float32).astype(np.float64)


def _golden_rows():
    return json.load(open(os.path.join(GOLD, "optimizer_weights.json")))["rows"]


def test_reference_voronoi_vector():
    for w in (O.quadrature_weights(GRID, [4, 10, 19, 29], O.SCAN_VORONOI),
              scan.quadrature_weights(GRID, [[4, 10, 19, 29]], scan.VORONOI)[0]):
        assert list(w) == [75.0, 75.0, 95.0, 165.0]


@pytest.mark.parametrize("strategy", [1, 2])
def test_reference_cmf_fit_properties(strategy):
    cie = _cie_fixture(GRID)
    total = 820.0 - 370.0 - 40.0  # 375 .. 785 nm
    for w in (O.quadrature_weights(GRID, [3, 10, 18, 25], strategy, cie),
              scan.quadrature_weights(GRID, [[3, 10, 18, 25]], strategy, cie)[0]):
        assert np.all(w >= -1e-9) and abs(w.sum() - total) < 1e-6
    w = scan.quadrature_weights(GRID, [[4, 12, 29]], strategy, cie)[0]
    assert w.size == 3 and abs(w.sum() - total) < 1e-6


def test_weights_match_reference_reports():
    cie = _cie_fixture(GRID)
    rows = _golden_rows()
    assert len(rows) == 55 and {r["strategy"] for r in rows} == set(STRAT)
    for r in rows:
        idx = [int(np.where(GRID == l)[0][0]) for l in r["lambda_nm"]]
        want = np.array(r["weights_nm"])
        got_o = O.quadrature_weights(GRID, idx, STRAT[r["strategy"]], cie)
        got_p = scan.quadrature_weights(GRID, [idx], STRAT[r["strategy"]], cie)[0]
        np.testing.assert_allclose(got_o, want, atol=6e-7, err_msg=str(r))
        np.testing.assert_allclose(got_p, want, atol=6e-7, err_msg=str(r))


def test_product_weights_match_oracle_on_random_combos():
    rng = np.random.default_rng(4)
    cie = scan.cie_1931_2deg(GRID)
    for count in (1, 2, 3, 5, 8):
        combos = np.array([np.sort(rng.choice(41, count, replace=False)) for _ in range(40)], np.uint8)
        for s in (0, 1, 2):
            got = scan.quadrature_weights(GRID, combos, s, cie, threads=3)
            want = np.array([O.quadrature_weights(GRID, c, s, cie) for c in combos])
            np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-9)
            if s == 0:
                np.testing.assert_allclose(got.sum(1), 410.0, rtol=1e-12)


def test_band_widths():
    np.testing.assert_array_equal(scan.band_widths(GRID), np.full(41, 10.0))
    wl = np.array([400.0, 410.0, 430.0, 470.0], np.float32)
    np.testing.assert_allclose(scan.band_widths(wl), [10.0, 15.0, 30.0, 40.0])
    np.testing.assert_array_equal(scan.band_widths(wl), O.band_widths(wl))


def test_cie_table_matches_reference_csv():
    rows = np.array(json.load(open(os.path.join(GOLD, "cie_1931.json")))["every_5nm"])
    got = scan.cie_1931_2deg(rows[:, 0])
    np.testing.assert_array_equal(got, rows[:, 1:].astype(np.float32).astype(np.float64))
    # linear between table points, clamped outside
    mid = scan.cie_1931_2deg([552.5, 300.0, 900.0])
    a, b = rows[rows[:, 0] == 550][0, 1:], rows[rows[:, 0] == 555][0, 1:]
    np.testing.assert_allclose(mid[0], 0.5 * (a + b), rtol=1e-6)
    np.testing.assert_array_equal(mid[1], rows[0, 1:].astype(np.float32))
    np.testing.assert_array_equal(mid[2], rows[-1, 1:].astype(np.float32))


def test_oklab_published_pairs():
    pairs = [([0.950, 1.000, 1.089], [1.000, 0.000, 0.000]), ([1.0, 0.0, 0.0], [0.450, 1.236, -0.019]),
             ([0.0, 1.0, 0.0], [0.922, -0.671, 0.263]), ([0.0, 0.0, 1.0], [0.153, -1.415, -0.449])]
    for xyz, lab in pairs:
        np.testing.assert_allclose(O.xyz_to_oklab(xyz), lab, atol=6e-4)
        np.testing.assert_allclose(scan.xyz_to_oklab(xyz), O.xyz_to_oklab(xyz), rtol=1e-14, atol=1e-15)


def test_percentile_definition():
    rng = np.random.default_rng(1)
    v = rng.exponential(size=1001)
    v[[3, 70]] = np.nan
    v[500] = np.inf
    fin = np.sort(v[np.isfinite(v)])
    for q in (0.5, 0.9, 0.95, 0.99, 0.0, 1.0):
        k = int(np.floor((fin.size - 1) * q + 0.5))  # round half away from zero, positive
        assert O.percentile(v, q) == fin[k]
    assert np.isnan(O.percentile([np.nan, np.inf], 0.5))


def _dataset(n=600, bands=41, seed=0):
    rng = np.random.default_rng(seed)
    # smooth positive spectra with per-sample slope and level
    lam = (GRID - 580.0) / 200.0
    level = rng.lognormal(0.0, 1.0, size=(n, 1))
    slope = rng.normal(0.0, 0.8, size=(n, 1))
    curve = rng.normal(0.0, 0.5, size=(n, 1))
    v = level * np.exp(slope * lam + curve * lam ** 2) * (1 + 0.05 * rng.standard_normal((n, bands)))
    return np.abs(v).astype(np.float32)


def test_all_band_voronoi_combo_is_exact():
    v = _dataset(200)
    cie = scan.cie_1931_2deg(GRID)
    ref, scale = O.reference_oklab(v, cie)
    combo = np.arange(41, dtype=np.uint8)
    w = O.band_widths(GRID)  # Voronoi cells of the full grid are the band widths
    out = O.evaluate_combos(v, ref, scale, GRID, cie, combo, w, exact=True)[0]
    assert out[0] < 1e-14 and out[6] < 1e-13
    # a sparse combo is not exact, and the score adds a quarter of p95
    sparse = np.array([4, 10, 19, 29], np.uint8)
    o2 = O.evaluate_combos(v, ref, scale, GRID, cie, sparse, O.quadrature_weights(GRID, sparse, 0), exact=True)[0]
    assert o2[0] > 1e-3 and abs(o2[7] - (o2[1] + 0.25 * o2[4])) < 1e-15


def test_reference_scale_is_p99_of_y():
    v = _dataset(300)
    cie = scan.cie_1931_2deg(GRID)
    ref, scale = O.reference_oklab(v, cie)
    y = v.astype(np.float64) @ cie[:, 1]
    assert abs(scale / O.percentile(y, 0.99) - 1.0) < 1e-13
    np.testing.assert_allclose(ref[7], O.xyz_to_oklab((v[7].astype(np.float64) @ cie) / scale), rtol=1e-12, atol=1e-15)


def test_screening_indices_and_combos():
    idx = scan.screening_indices(655360, 65536)
    assert idx.size == 65536 and idx[0] == 5 and idx[-1] == 655355 and np.all(np.diff(idx) == 10)
    assert scan.screening_indices(100, 0).size == 100 and scan.screening_indices(100, 200).size == 100
    c = scan.all_combos(41, 3)
    assert c.shape == (10660, 3) and list(c[0]) == [0, 1, 2] and list(c[-1]) == [38, 39, 40]
    assert np.all(np.diff(c.astype(int), axis=1) > 0)


def test_invalid_weight_requests():
    from pysky.api import InvalidConfig
    with pytest.raises(InvalidConfig):
        scan.quadrature_weights(GRID, [[0, 41]], 0)
    with pytest.raises(InvalidConfig):
        scan.quadrature_weights(GRID, [[0, 1]], 7)
    with pytest.raises(InvalidConfig):
        scan.quadrature_weights(GRID, np.zeros((1, 9), np.uint8), 0)
