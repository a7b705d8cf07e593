"""Edge geometries and shapes against the oracle, path for path at 1 spp
(sky_trace.wgsl keys every path by (pixel, sample, band, dim), so each film
entry of a 1-spp render is one path that the GPU and the FP32 oracle trace
alike up to float rounding):

* observer altitudes from the ground to just below the top of the atmosphere
  (config.rs:47 defaults to 0.2 km);
* the sun at the zenith, below the horizon and just above it at oblique
  azimuths;
* the sky-view LUT row on the horizon, whose camera rays graze the ground
  (the ray set-up's FP64 branch);
* one-pixel and odd-sized images with 1, 2, 4, 5 and 64 bands, whose XYZ takes
  the wave-level band sums (power-of-two band counts from 4) or the band-order
  sum on one lane (the others), both for renders that hold every sample in one
  work item and for renders split into sample chunks (k_finalize).
"""
import numpy as np
import pytest

import oracle as O
import pysky
from helpers import case, paired_stats

pytestmark = pytest.mark.gpu

# carried optical depth + lobe sampling in 3 shells growing 4x: the specialised kernels
FAST = dict(mode="carried_is", layers=3, shell_growth=4.0)


def _pair(gpu_context, scene, cfg, op):
    print(f"[gpu] {cfg.width}x{cfg.height} nl={scene.band_count} spp={cfg.spp} sun={scene.sun_elevation_deg} "
          f"h_obs={cfg.observer_altitude_km}", flush=True)
    g = pysky.render(scene, cfg, context=gpu_context)
    ref, _, _, _, rc = O.render(op, "f32")
    assert rc == 0
    film = g.values()
    assert np.all(np.isfinite(film)) and np.all(film >= 0)
    frac, bad = paired_stats(film, ref.ravel())
    print(f"paired fraction {frac:.5f} ({len(bad)} of {film.size} differ), items {g.stats.items}", flush=True)
    return g, frac, bad


@pytest.mark.parametrize("h", [0.0, 30.0, 99.0])
def test_observer_altitudes_pair_with_oracle(gpu_context, h):
    scene, cfg, op = case(width=64, height=32, n_lambda=8, spp=1, observer_altitude_km=h, **FAST)
    _, frac, _ = _pair(gpu_context, scene, cfg, op)
    assert frac >= 0.995


@pytest.mark.parametrize("h", [1.0, 20.0])
def test_sky_view_horizon_row_grazes_the_ground(gpu_context, h):
    """A 33-row sky-view LUT has its v = 0.75 row (y = 24) on the horizon
    (sky_trace.wgsl:625-681): camera rays tangent to the ground, the rays the
    ray set-up forms in FP64 (the oracle counts exactly the 256 of that row).
    The other rows pair path for path.  On the tangent row the GPU's and the
    oracle's FP32 camera directions (hardware against library cosines, ~1 ulp
    apart) can fall on either side of tangency, so that row is checked finite
    and non-negative only."""
    scene, cfg, op = case(width=32, height=33, n_lambda=8, spp=1, projection=1, observer_altitude_km=h, **FAST)
    print(f"[gpu] sky-view 32x33 nl=8 h_obs={h}", flush=True)
    g = pysky.render(scene, cfg, context=gpu_context)
    ref, _, _, cnt, rc = O.render(op, "f32")
    assert rc == 0 and cnt["graze64"] == 32 * 8
    film = g.values().reshape(33, 32, 8)
    assert np.all(np.isfinite(film)) and np.all(film >= 0)
    rows = np.arange(33) != 24
    frac, bad = paired_stats(film[rows], ref[rows])
    print(f"paired fraction off the horizon row {frac:.5f} ({len(bad)} differ)", flush=True)
    assert frac >= 0.995


@pytest.mark.parametrize("sun,azimuth", [(90.0, 0.0), (-6.0, 135.0), (0.3, 270.0)])
def test_sun_positions_pair_with_oracle(gpu_context, sun, azimuth):
    scene, cfg, op = case(width=64, height=32, n_lambda=8, spp=1, sun=sun, azimuth=azimuth, **FAST)
    _, frac, _ = _pair(gpu_context, scene, cfg, op)
    assert frac >= 0.995


@pytest.mark.parametrize("w,h,nl", [(1, 1, 1), (3, 5, 5), (7, 3, 2), (5, 2, 4), (2, 3, 64)])
@pytest.mark.parametrize("spp", [1, 32])
def test_small_shapes_pair_and_weigh_xyz(gpu_context, w, h, nl, spp):
    scene, cfg, op = case(width=w, height=h, n_lambda=nl, spp=spp, **FAST)
    g, _, bad = _pair(gpu_context, scene, cfg, op)
    # at most one entry in 200 (and at least one allowed) may differ by a rounding-flipped decision
    assert len(bad) <= max(1, g.values().size // 200)
    film = g.values().reshape(w * h, nl).astype(np.float64)
    xyz = film @ O.lambda_table(op)[:, 5:8]
    np.testing.assert_allclose(g.xyz.reshape(xyz.shape), xyz, rtol=2e-5, atol=1e-6 * max(xyz.max(), 1e-30))
