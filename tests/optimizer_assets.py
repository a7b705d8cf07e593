"""Synthetic sky-view LUT assets (asset.json + 41 band EXRs, the layout
sky_render writes) for the wavelength-optimizer tests."""
import json
import os

import numpy as np

from exr_png import write_exr

GRID = np.arange(380, 781, 10).astype(np.float32)


def spectra(n, seed):
    """Smooth positive spectra with per-sample level, slope and curvature."""
    rng = np.random.default_rng(seed)
    lam = (GRID - 580.0) / 200.0
    level = rng.lognormal(0.0, 1.0, size=(n, 1))
    slope = rng.normal(0.0, 0.8, size=(n, 1))
    curve = rng.normal(0.0, 0.5, size=(n, 1))
    v = level * np.exp(slope * lam + curve * lam ** 2) * (1 + 0.05 * rng.standard_normal((n, GRID.size)))
    return np.abs(v).astype(np.float32)


def write_asset(out, values, elevation, kind="spectral_sky_view_lut_v0", spp=64, centers=None):
    """values: [height, width, bands]."""
    os.makedirs(os.path.join(out, "bands"), exist_ok=True)
    h, w, nb = values.shape
    centers = list(GRID[:nb]) if centers is None else list(centers)
    files = []
    for b in range(nb):
        name = "bands/sky_%03.0fnm.exr" % centers[b]
        v = values[:, :, b]
        write_exr(os.path.join(out, name), {"R": v, "G": v, "B": v})
        files.append(name)
    manifest = {"version": 0, "kind": kind, "dimensions": [w, h], "spp": spp, "seed": 1,
                "sun_elevation_deg": float(elevation), "sun_azimuth_deg": 0.0, "observer_altitude_km": 0.2,
                "band_centers_nm": [float(c) for c in centers], "data_hash": None,
                "files": {"rgb_exr": "sky_rgb.exr", "rgb_png": "sky_rgb.png", "band_exrs": files}}
    with open(os.path.join(out, "asset.json"), "w") as f:
        json.dump(manifest, f, indent=2)


def read_ranking(path):
    import csv
    with open(path) as f:
        return list(csv.DictReader(f))
