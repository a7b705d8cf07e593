"""sky_wavelength_optimizer end to end on the GPU (main.rs:146-241): the
ranked combos, their weights and statistics against the FP64 oracle's
evaluate_combo on the same LUT samples, for the exhaustive search and for the
beam search with local refinement that takes over above 5M combos."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
import pysky
from optimizer_assets import GRID, read_ranking, spectra, write_asset
from pysky import scan

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(pysky.LIB_PATH), "..", "bin", "sky_wavelength_optimizer")
RTOL = 2e-4
STRAT = {"voronoi": 0, "cmf-mass": 1, "cmf-moments": 2}


@pytest.fixture(scope="module")
def assets(tmp_path_factory):
    root = tmp_path_factory.mktemp("lut")
    vals = []
    for k, el in enumerate((2.0, 20.0)):
        v = spectra(24 * 24, 30 + k).reshape(24, 24, 41)
        write_asset(str(root / f"elev_{k}"), v, el)
        vals.append(v.reshape(-1, 41))
    return root, np.concatenate(vals)


def _run(root, out, *args):
    cmd = [EXE, "--asset-dir", str(root / "elev_0"), "--asset-dir", str(root / "elev_1"), "--out-dir", str(out), *args]
    print("[gpu] " + " ".join(cmd[1:]), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    print(r.stdout[-2000:])
    assert r.returncode == 0, r.stderr
    return r.stdout


def _oracle_rows(values, rows, strategy):
    cie = scan.cie_1931_2deg(GRID)
    ref, scale = O.reference_oklab(values, cie)
    out = []
    for r in rows:
        lam = [float(r[k]) for k in sorted(r) if k.startswith("lambda")]
        combo = np.array([int(np.where(GRID == l)[0][0]) for l in lam], np.uint8)
        w = O.quadrature_weights(GRID, combo, STRAT[strategy], cie)
        out.append((combo, w, O.evaluate_combos(values, ref, scale, GRID, cie, combo, w, exact=True)[0]))
    return out, ref, scale, cie


@pytest.mark.parametrize("strategy", ["voronoi", "cmf-mass"])
def test_exhaustive_search_matches_oracle(assets, tmp_path, strategy):
    root, values = assets
    stdout = _run(root, tmp_path, "--primary-count", "3", "--screen-samples", "300", "--refine-top", "40",
                  "--top", "6", "--weight-strategy", strategy)
    assert "loaded 2 assets, 1152 samples, 41 bands, 10660 3-wavelength combos" in stdout
    rows = read_ranking(tmp_path / "ranking.csv")
    assert len(rows) == 6
    checked, ref, scale, cie = _oracle_rows(values, rows, strategy)
    for r, (combo, w, o) in zip(rows, checked):
        np.testing.assert_allclose([float(r[f"weight{k}_nm"]) for k in range(3)], w, atol=1e-6)
        for k, f in enumerate(("mean", "rmse", "p50", "p90", "p95", "p99", "max", "score")):
            np.testing.assert_allclose(float(r[f]), o[k], rtol=RTOL, err_msg=f)
        assert int(r["sample_count"]) == values.shape[0]
    # ranks follow the score
    scores = [float(r["score"]) for r in rows]
    assert scores == sorted(scores)
    # screening: the 40 refined candidates are the oracle's 40 best RMSE on the same subset
    combos = scan.all_combos(41, 3)
    weights = np.array([O.quadrature_weights(GRID, c, STRAT[strategy], cie) for c in combos])
    idx = scan.screening_indices(values.shape[0], 300)
    s = O.evaluate_combos(values, ref, scale, GRID, cie, combos, weights, idx=idx)[:, 1]
    best40 = np.sort(s)[39]
    for combo, _, _ in checked:
        i = int(np.where((combos == combo).all(axis=1))[0][0])
        assert s[i] <= best40 * (1 + RTOL)
    per = read_ranking(tmp_path / "ranking_by_elevation.csv")
    assert len(per) == 12 and {float(p["elevation_deg"]) for p in per} == {2.0, 20.0}
    report = open(tmp_path / "report.md").read()
    assert "# Sky-View 3-Wavelength Search" in report and f"Weight strategy: `{strategy}`" in report


def test_beam_search_above_exhaustive_limit(assets, tmp_path):
    root, values = assets
    stdout = _run(root, tmp_path, "--primary-count", "7", "--screen-samples", "200", "--beam-width", "6",
                  "--refine-top", "6", "--top", "3")
    assert "using deterministic beam_width=6" in stdout and "beam depth 7: evaluated" in stdout
    rows = read_ranking(tmp_path / "ranking.csv")
    assert len(rows) == 3
    checked, *_ = _oracle_rows(values, rows, "voronoi")
    for r, (combo, w, o) in zip(rows, checked):
        assert len(set(combo.tolist())) == 7
        for k, f in enumerate(("mean", "rmse", "p95", "score")):
            np.testing.assert_allclose(float(r[f]), o[[0, 1, 4, 7][k]], rtol=RTOL, err_msg=f)


def test_rendered_luts_end_to_end(tmp_path):
    """sky_render --sky-view-lut (41 bands) for two elevations, then the optimizer on them."""
    render = os.path.join(os.path.dirname(EXE), "sky_render")
    for k, el in enumerate((0.0, 10.0)):
        r = subprocess.run([render, "--sky-view-lut", "--width", "24", "--height", "24", "--spp", "16",
                            "--sun-elevation-deg", str(el), "--out", str(tmp_path / f"elev_{k}")],
                           capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
    stdout = _run(tmp_path, tmp_path / "out", "--primary-count", "4", "--screen-samples", "256",
                  "--refine-top", "16", "--top", "4")
    assert "101270 4-wavelength combos" in stdout
    rows = read_ranking(tmp_path / "out" / "ranking.csv")
    assert len(rows) == 4 and all(float(r["rmse"]) > 0 for r in rows)
