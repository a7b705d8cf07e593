"""sky_wavelength_optimizer argument and dataset validation (main.rs:243-248,
353-386): every failure here is reported before a GPU is touched."""
import os
import subprocess

import numpy as np
import pytest

import pysky
from optimizer_assets import spectra, write_asset

EXE = os.path.join(os.path.dirname(pysky.LIB_PATH), "..", "bin", "sky_wavelength_optimizer")


def _run(*args):
    return subprocess.run([EXE, *args], capture_output=True, text=True, timeout=120)


def test_usage_without_assets():
    r = _run()
    assert r.returncode == 2 and "--asset-dir" in r.stdout


def test_primary_count_range(tmp_path):
    r = _run("--asset-dir", str(tmp_path), "--primary-count", "9", "--out-dir", str(tmp_path / "o"))
    assert r.returncode == 1 and "--primary-count must be in 1..=8" in r.stderr


def test_rejects_panorama_assets(tmp_path):
    write_asset(tmp_path / "a", spectra(16, 0).reshape(4, 4, 41), 5.0, kind="spectral_panorama_v0")
    r = _run("--asset-dir", str(tmp_path / "a"), "--out-dir", str(tmp_path / "o"))
    assert r.returncode == 1 and "expected kind spectral_sky_view_lut_v0" in r.stderr


def test_rejects_other_band_counts(tmp_path):
    v = spectra(16, 0).reshape(4, 4, 41)[:, :, :16]
    write_asset(tmp_path / "a", v, 5.0, centers=np.linspace(392.5, 767.5, 16))
    r = _run("--asset-dir", str(tmp_path / "a"), "--out-dir", str(tmp_path / "o"))
    assert r.returncode == 1 and "expected 41 bands, found 16" in r.stderr


def test_rejects_mismatched_grids_and_sizes(tmp_path):
    write_asset(tmp_path / "a", spectra(16, 0).reshape(4, 4, 41), 5.0)
    write_asset(tmp_path / "b", spectra(16, 1).reshape(4, 4, 41), 10.0, centers=np.arange(381, 782, 10))
    r = _run("--asset-dir", str(tmp_path / "a"), "--asset-dir", str(tmp_path / "b"), "--out-dir", str(tmp_path / "o"))
    assert r.returncode == 1 and "wavelength grid differs from first asset" in r.stderr
    import json
    m = json.load(open(tmp_path / "a" / "asset.json"))
    m["dimensions"] = [5, 4]
    json.dump(m, open(tmp_path / "a" / "asset.json", "w"))
    r = _run("--asset-dir", str(tmp_path / "a"), "--out-dir", str(tmp_path / "o"))
    assert r.returncode == 1 and "differ from manifest 5x4" in r.stderr
