"""The benchmark configurations against the oracle, with the launch and the
estimator the benchmark times (shells, occupancy and work split from the
tuning cache; the specialised kernels the library picks for them).

  * twilight / grazing (512x256, 64 bands, sun -4 deg, aerosol x4, albedo 0.8,
    depth <= 64): at 1 spp path for path with the FP32 oracle on a 64x32
    sub-grid; at 2048 spp per 4x4-pixel block against the FP64 oracle;
  * ragged batches: items of one launch (3, 4, 8, 16 and 64 wavelengths drawn
    in 400-700 nm, suns from -6 to 90 deg) path for path with the oracle;
  * the full 256-render ragged batch (256 spp): finite, non-negative, and
    equal to the same renders issued one by one on a subset.
"""
import dataclasses
import math

import numpy as np
import pytest

import oracle as O
import pysky
from helpers import oracle_flags, paired_stats
from pysky import autotune

pytestmark = pytest.mark.gpu

C3 = dict(width=512, height=256, n_lambda=64, max_depth=64, sun=-4.0, albedo=0.8, aerosol=4.0)


def _c3(spp, seed=1):
    atm = pysky.AtmosphereParams(ground_albedo=C3["albedo"])
    atm.aerosol_sca_550_m_inv *= C3["aerosol"]
    atm.aerosol_abs_550_m_inv *= C3["aerosol"]
    scene = pysky.SceneData(atmosphere=atm, sun_elevation_deg=C3["sun"],
                            wavelengths_nm=list(pysky.midpoint_grid(C3["n_lambda"])))
    cfg = pysky.RenderConfig(width=C3["width"], height=C3["height"], spp=spp, seed=seed, max_depth=C3["max_depth"],
                             rr_start_depth=4, carried_tau=True, phase_importance_sampling=True)
    ctx = pysky.Context(0)
    params = autotune.tuned(scene, cfg)
    cfg = autotune.apply(ctx, cfg, params)
    return scene, cfg, ctx, params


def oracle_params(scene, cfg, **over):
    """The oracle's parameters for one pysky render."""
    a = scene.atmosphere
    kw = dict(width=cfg.width, height=cfg.height, spp=cfg.spp, seed=cfg.seed, sample_offset=cfg.sample_offset,
              sun_elevation_deg=scene.sun_elevation_deg, sun_azimuth_deg=scene.sun_azimuth_deg,
              max_depth=cfg.max_depth, rr_start_depth=cfg.rr_start_depth, majorant_layers=cfg.majorant_layers,
              shell_growth=cfg.shell_growth, projection=int(cfg.output_projection),
              ground_albedo=a.ground_albedo, aerosol_sca_550_m_inv=a.aerosol_sca_550_m_inv,
              aerosol_abs_550_m_inv=a.aerosol_abs_550_m_inv, aerosol_g=a.aerosol_g,
              lambda_nm=list(scene.wavelengths_nm), weight_nm=scene.weights_nm,
              flags=oracle_flags(phase_importance_sampling=cfg.phase_importance_sampling,
                                 ratio_tracking=cfg.ratio_tracking, carried_tau=cfg.carried_tau))
    kw.update(over)
    return O.Params(**kw)


def _grid(step, w, h):
    xs, ys = np.arange(step // 2, w, step), np.arange(step // 2, h, step)
    return (ys[:, None] * w + xs[None, :]).ravel(), len(ys), len(xs)


def test_c3_paths_pair_with_oracle():
    scene, cfg, ctx, params = _c3(spp=1)
    print(f"[gpu] c3 1 spp, tuning {params}", flush=True)
    f = pysky.render(scene, cfg, want_film=True, want_xyz=False, context=ctx)
    pix, _, _ = _grid(8, cfg.width, cfg.height)
    ref, _, _, rc = O.render_pixels(oracle_params(scene, cfg), pix, "f32")
    assert rc == 0
    got = f.values().reshape(-1, C3["n_lambda"])[pix]
    frac, _ = paired_stats(got, ref)
    print(f"[gpu] c3: {frac:.4%} of {ref.size} paths paired", flush=True)
    assert frac >= 0.99


def test_c3_matches_fp64_oracle():
    """2048 spp (the configuration's count) against the FP64 oracle's 256 spp
    on independent streams.  Twilight path radiance is heavy-tailed (24%
    image noise at these counts), so the z-scores are formed per 4x4 block of
    the 64x32 sub-grid (16 pixels x 3 channels each)."""
    scene, cfg, ctx, _ = _c3(spp=2048)
    f = pysky.render(scene, cfg, want_film=True, want_m2=True, want_xyz=True, context=ctx)
    print(f"[gpu] c3 2048 spp: kernel {f.stats.kernel_ms:.1f} ms", flush=True)
    pix, ny, nx = _grid(8, cfg.width, cfg.height)
    nb = C3["n_lambda"]
    op = oracle_params(scene, cfg, spp=256, seed=(7 * 65536 << 32) | 1)
    fb, mb, _, rc = O.render_pixels(op, pix, "f64")
    assert rc == 0
    w = O.lambda_table(op)[:, 5:8]
    fa = f.values().reshape(-1, nb)[pix].astype(np.float64)
    ma = f.m2.reshape(-1, nb)[pix].astype(np.float64)
    np.testing.assert_allclose(f.xyz.reshape(-1, 3)[pix], fa @ w, rtol=1e-4, atol=1e-6 * (fa @ w).max())
    va = (np.maximum(ma - fa ** 2, 0) / cfg.spp) @ (w ** 2)
    vb = (np.maximum(mb - fb ** 2, 0) / op.spp) @ (w ** 2)

    def blocks(x):
        return x.reshape(ny // 4, 4, nx // 4, 4, 3).sum(axis=(1, 3)).reshape(-1, 3)

    z = (blocks(fa @ w) - blocks(fb @ w)) / np.sqrt(blocks(va) + blocks(vb))
    z_tot = ((fa @ w).sum() - (fb @ w).sum()) / math.sqrt(va.sum() + vb.sum())
    print(f"[gpu] c3 vs FP64: block z mean {z.mean():+.3f} std {z.std():.3f} max {np.abs(z).max():.2f}, "
          f"total z {z_tot:+.2f}", flush=True)
    assert abs(z_tot) < 4.0
    assert abs(z.mean()) < 5.0 / math.sqrt(z.size)
    assert 0.85 < z.std() < 1.15
    assert np.abs(z).max() < 5.0


def _batch_items(spp):
    """One item per wavelength-set size of the ragged workload, at small
    resolutions, with the tuned shells of the batch benchmark."""
    rng = np.random.default_rng(11)
    out = []
    for i, (nb, el, (w, h)) in enumerate(zip((3, 4, 8, 16, 64, 3, 8),
                                             (-5.0, 12.0, 45.0, -1.5, 5.0, 88.0, 2.0),
                                             ((64, 32), (128, 64), (64, 32), (128, 64), (64, 32), (64, 32),
                                              (128, 64)))):
        lam = np.sort(rng.uniform(400.0, 700.0, nb))
        scene = pysky.SceneData(sun_elevation_deg=el, wavelengths_nm=list(lam), weights_nm=list(np.full(nb, 300.0 / nb)))
        cfg = pysky.RenderConfig(width=w, height=h, spp=spp, seed=1000003 + i, max_depth=32, rr_start_depth=4,
                                 carried_tau=True, phase_importance_sampling=True, majorant_layers=3, shell_growth=4.0)
        out.append((scene, cfg))
    return out


def test_ragged_batch_items_pair_with_oracle():
    batch = _batch_items(spp=1)
    ctx = pysky.Context(0)
    autotune.apply(ctx, pysky.RenderConfig(), autotune.tuned(batch[4][0], pysky.RenderConfig(
        width=1024, height=512, carried_tau=True, phase_importance_sampling=True)))
    res = pysky.render_batch(batch, want_film=True, context=ctx)
    for (scene, cfg), film in zip(batch, res.films):
        op = oracle_params(scene, cfg)
        ref, _, xyz, _, rc = O.render(op, "f32")
        assert rc == 0
        frac, _ = paired_stats(film.values(), ref.ravel())
        print(f"[gpu] batch item {scene.band_count} bands sun {scene.sun_elevation_deg}: {frac:.4%} paired", flush=True)
        assert frac >= 0.99
        # XYZ with the item's own quadrature weights
        np.testing.assert_allclose(film.xyz.reshape(-1, 3), film.values().reshape(-1, scene.band_count).astype(
            np.float64) @ O.lambda_table(op)[:, 5:8], rtol=1e-4, atol=1e-6 * max(float(film.xyz.max()), 1e-30))


def test_full_ragged_batch():
    batch = pysky.ragged_batch(n=256, seed=1, spp=256, max_depth=32)
    batch = [(sc, dataclasses.replace(cfg, majorant_layers=3, shell_growth=4.0)) for sc, cfg in batch]
    ctx = pysky.Context(0)
    autotune.apply(ctx, pysky.RenderConfig(), autotune.tuned(pysky.SceneData(
        wavelengths_nm=list(pysky.midpoint_grid(64))), pysky.RenderConfig(width=1024, height=512, carried_tau=True,
                                                                          phase_importance_sampling=True)))
    res = pysky.render_batch(batch, context=ctx)
    print(f"[gpu] c4: {res.stats.renders} renders, {res.stats.path_samples:.3e} samples in "
          f"{res.stats.kernel_ms:.0f} ms, p50 {res.stats.latency_p50_ms:.0f} p99 {res.stats.latency_p99_ms:.0f} ms",
          flush=True)
    assert res.stats.renders == 256 and res.stats.watchdog == 0
    for f in res.films:
        assert np.all(np.isfinite(f.xyz)) and np.all(f.xyz >= 0)
    assert sum(float(f.xyz.sum()) > 0 for f in res.films) >= 250  # a few deep-twilight renders may be dark
    cost = [cfg.width * cfg.height * cfg.spp * sc.band_count for sc, cfg in batch]
    order = np.argsort(cost)
    for i in list(order[:3]) + list(order[-3:]):
        scene, cfg = batch[i]
        single = pysky.render(scene, cfg, want_film=False, context=ctx)
        np.testing.assert_allclose(res.films[i].xyz, single.xyz, rtol=1e-4, atol=1e-6 * max(single.xyz.max(), 1e-30))


def test_sky_view_sweep_batch_pairs_with_oracle():
    """The reference's sky-view bake (scripts/bake_skyview_elevations.ps1:
    ten sun elevations -2..90 deg, sky-view LUT projection, 41 bands of 10 nm
    from 380 nm, uncapped depth with roulette from depth 4) as one batch of ten
    renders at 32x32 and 1 spp: every item path for path with the oracle."""
    lam = [380.0 + 10.0 * i for i in range(41)]
    batch = []
    for el in (-2.0, -1.0, 0.0, 1.0, 2.0, 5.0, 10.0, 20.0, 45.0, 90.0):
        scene = pysky.SceneData(sun_elevation_deg=el, wavelengths_nm=lam, weights_nm=[10.0] * 41)
        cfg = pysky.RenderConfig(width=32, height=32, spp=1, seed=0x5EC720260430, max_depth=0, rr_start_depth=4,
                                 output_projection=pysky.OutputProjection.SkyViewLut, carried_tau=True,
                                 phase_importance_sampling=True, majorant_layers=3, shell_growth=4.0)
        batch.append((scene, cfg))
    ctx = pysky.Context(0)
    res = pysky.render_batch(batch, want_film=True, context=ctx)
    for (scene, cfg), film in zip(batch, res.films):
        ref, _, _, _, rc = O.render(oracle_params(scene, cfg), "f32")
        assert rc == 0
        frac, _ = paired_stats(film.values(), ref.ravel())
        print(f"[gpu] sky-view sweep, sun {scene.sun_elevation_deg:+.0f} deg: {frac:.4%} paired", flush=True)
        assert frac >= 0.99
