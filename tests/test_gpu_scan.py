"""Sparse-wavelength combo scan on the GPU against the FP64 oracle
(evaluate_combo, crates/sky-wavelength-optimizer/src/main.rs:414-488).

The GPU evaluates each sample in FP32 (FP64 sums); the oracle in FP64.  The
tolerance 2e-4 relative covers FP32 rounding of the XYZ quadrature, the cube
roots and the difference of two Oklab colours of similar size."""
import numpy as np
import pytest

import oracle as O
from pysky import scan
from pysky.api import InvalidConfig

pytestmark = pytest.mark.gpu

GRID = np.arange(380, 781, 10).astype(np.float32)
RTOL = 2e-4


def _dataset(n, seed):
    rng = np.random.default_rng(seed)
    lam = (GRID - 580.0) / 200.0
    level = rng.lognormal(0.0, 1.0, size=(n, 1))
    slope = rng.normal(0.0, 0.8, size=(n, 1))
    curve = rng.normal(0.0, 0.5, size=(n, 1))
    v = level * np.exp(slope * lam + curve * lam ** 2) * (1 + 0.05 * rng.standard_normal((n, GRID.size)))
    return np.abs(v).astype(np.float32)


@pytest.fixture(scope="module")
def data():
    v = _dataset(4096, 3)
    cie = scan.cie_1931_2deg(GRID)
    ref, scale = O.reference_oklab(v, cie)
    print(f"[gpu] scan dataset {v.shape} scale {scale:.4g}", flush=True)
    ds = scan.ScanDataset(v, ref)
    yield v, cie, ref, scale, ds
    ds.close()


def _combos(count, n, seed):
    rng = np.random.default_rng(seed)
    return np.array([np.sort(rng.choice(GRID.size, count, replace=False)) for _ in range(n)], np.uint8)


def _coef(combos, strategy, cie, scale):
    w = scan.quadrature_weights(GRID, combos, strategy, cie)
    return w, scan.combo_coefficients(combos, w, scan.band_widths(GRID), cie, scale)


@pytest.mark.parametrize("count", [1, 2, 3, 4, 5, 6, 7, 8])
@pytest.mark.parametrize("subset", [False, True])
def test_scan_matches_oracle(data, count, subset):
    v, cie, ref, scale, ds = data
    combos = _combos(count, 300, count)
    w, coef = _coef(combos, scan.CMF_MASS if count % 2 else scan.VORONOI, cie, scale)
    idx = scan.screening_indices(v.shape[0], 1000) if subset else None
    print(f"[gpu] scan count={count} subset={subset}", flush=True)
    s, q, st = ds.scan(combos, coef, idx)
    n = 1000 if subset else v.shape[0]
    assert st.evaluations == 300 * n
    want = O.evaluate_combos(v, ref, scale, GRID, cie, combos, w, idx=idx)
    np.testing.assert_allclose(s / n, want[:, 0], rtol=RTOL)
    np.testing.assert_allclose(np.sqrt(q / n), want[:, 1], rtol=RTOL)
    # the ranking by RMSE agrees wherever the oracle's scores are apart by more than the tolerance
    order = np.argsort(np.sqrt(q / n))
    top = want[order[:10], 1]
    assert np.all(np.diff(top) > -2 * RTOL * top[1:])


@pytest.mark.parametrize("count", [2, 4, 8])
def test_refine_matches_oracle(data, count):
    v, cie, ref, scale, ds = data
    combos = _combos(count, 40, 10 + count)
    w, coef = _coef(combos, scan.CMF_MOMENTS, cie, scale)
    print(f"[gpu] refine count={count}", flush=True)
    got, st = ds.refine(combos, coef)
    want = O.evaluate_combos(v, ref, scale, GRID, cie, combos, w, exact=True)
    for k, f in enumerate(scan.STAT_FIELDS):
        np.testing.assert_allclose(got[f], want[:, k], rtol=RTOL, err_msg=f)
    assert np.all(got["sample_count"] == v.shape[0])
    # the screening sums of the same combos give the same mean and RMSE
    s, q, _ = ds.scan(combos, coef)
    np.testing.assert_allclose(s / v.shape[0], got["mean"], rtol=1e-6)
    np.testing.assert_allclose(np.sqrt(q / v.shape[0]), got["rmse"], rtol=1e-6)


def test_refine_on_sample_subset(data):
    v, cie, ref, scale, ds = data
    combos = _combos(3, 16, 99)
    w, coef = _coef(combos, scan.VORONOI, cie, scale)
    idx = np.arange(1000, 3048, dtype=np.uint32)  # one "asset" range
    got, _ = ds.refine(combos, coef, idx)
    want = O.evaluate_combos(v, ref, scale, GRID, cie, combos, w, idx=idx, exact=True)
    for k, f in enumerate(scan.STAT_FIELDS):
        np.testing.assert_allclose(got[f], want[:, k], rtol=RTOL, err_msg=f)


def test_combo_that_spans_the_signal_is_exact():
    """Samples that are zero outside bands (5, 12, 20, 33): the combo of those
    bands with weights equal to the band widths reproduces the reference
    colour, so every error is ~0."""
    rng = np.random.default_rng(8)
    bands = np.array([5, 12, 20, 33], np.uint8)
    v = np.zeros((2000, GRID.size), np.float32)
    v[:, bands] = rng.lognormal(0.0, 0.7, size=(2000, 4)).astype(np.float32)
    cie = scan.cie_1931_2deg(GRID)
    ref, scale = O.reference_oklab(v, cie)
    ds = scan.ScanDataset(v, ref)
    coef = scan.combo_coefficients(bands[None], np.full((1, 4), 10.0), scan.band_widths(GRID), cie, scale)
    got, _ = ds.refine(bands[None], coef)
    assert got["max"][0] < 2e-5 and got["rmse"][0] < 5e-6
    ds.close()


def test_non_finite_samples_follow_the_reference(data):
    """A NaN band value makes that sample's error NaN: the mean and RMSE are
    NaN, the percentiles skip it and the maximum ignores it (main.rs:459-476)."""
    v, cie, ref, scale, _ = data
    v2 = v[:512].copy()
    v2[17, 9] = np.nan
    ds = scan.ScanDataset(v2, ref[:512])
    combos = np.array([[3, 9, 20], [1, 2, 3]], np.uint8)
    w, coef = _coef(combos, scan.VORONOI, cie, scale)
    got, _ = ds.refine(combos, coef)
    want = O.evaluate_combos(v2, ref[:512], scale, GRID, cie, combos, w, exact=True)
    assert np.isnan(got["mean"][0]) and np.isnan(want[0, 0])
    for k, f in enumerate(scan.STAT_FIELDS[2:], start=2):
        np.testing.assert_allclose(got[f], want[:, k], rtol=RTOL, err_msg=f)
    np.testing.assert_allclose(got["mean"][1], want[1, 0], rtol=RTOL)
    ds.close()


def test_deterministic(data):
    v, cie, ref, scale, ds = data
    combos = _combos(5, 1000, 5)
    _, coef = _coef(combos, scan.VORONOI, cie, scale)
    a = ds.scan(combos, coef)[:2]
    b = ds.scan(combos, coef)[:2]
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_invalid_requests(data):
    v, cie, ref, scale, ds = data
    coef = np.ones((1, 2, 3), np.float32)
    with pytest.raises(InvalidConfig):
        ds.scan(np.array([[0, 41]], np.uint8), coef)
    with pytest.raises(InvalidConfig):
        ds.scan(np.array([[0, 1]], np.uint8), coef, np.array([0, 4096], np.uint32))
    with pytest.raises(InvalidConfig):
        ds.scan(np.zeros((1, 9), np.uint8), np.ones((1, 9, 3), np.float32))


def test_device_pointer_variant(data):
    import torch
    v, cie, ref, scale, ds = data
    combos = _combos(6, 700, 6)
    _, coef = _coef(combos, scan.VORONOI, cie, scale)
    idx = scan.screening_indices(v.shape[0], 777)
    s, q, _ = ds.scan(combos, coef, idx)
    dev = torch.device("cuda:0")
    tc = torch.from_numpy(combos).to(dev)
    tf = torch.from_numpy(coef).to(dev)
    ti = torch.from_numpy(idx.astype(np.int32)).to(dev)
    ts = torch.zeros(700, dtype=torch.float64, device=dev)
    tq = torch.zeros(700, dtype=torch.float64, device=dev)
    st = ds.scan_device(tc, tf, ts, tq, samples=ti)
    assert st.evaluations == 700 * 777
    assert np.array_equal(ts.cpu().numpy(), s) and np.array_equal(tq.cpu().numpy(), q)
    # out-of-range indices are clamped on the device and reported
    bad = tc.clone()
    bad[3, 2] = 41
    with pytest.raises(InvalidConfig):
        ds.scan_device(bad, tf, ts, tq)
    bad_i = ti.clone()
    bad_i[5] = v.shape[0]
    with pytest.raises(InvalidConfig):
        ds.scan_device(tc, tf, ts, tq, samples=bad_i)
    # and the dataset still works afterwards
    ds.scan_device(tc, tf, ts, tq, samples=ti)
    assert np.array_equal(ts.cpu().numpy(), s)
