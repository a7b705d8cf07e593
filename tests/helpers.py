"""Shared helpers: one configuration description mapped onto both the product
(pysky) and the oracle, so parity tests compare like with like."""
from __future__ import annotations

import numpy as np

import oracle as O
import pysky

FLAGS = {
    "reference": dict(),
    "carried": dict(carried_tau=True),
    "carried_is": dict(carried_tau=True, phase_importance_sampling=True),
    "carried_is_ratio": dict(carried_tau=True, phase_importance_sampling=True, ratio_tracking=True),
}


def oracle_flags(phase_importance_sampling=False, ratio_tracking=False, carried_tau=False, **_):
    f = 0
    if phase_importance_sampling:
        f |= O.FLAG_PHASE_IS
    if ratio_tracking:
        f |= O.FLAG_RATIO_TRACKING
    if carried_tau:
        f |= O.FLAG_CARRIED_TAU
    return f


def case(width=128, height=64, n_lambda=16, spp=64, sun=10.0, max_depth=8, seed=1, mode="reference",
         layers=64, projection=0, sample_offset=0, albedo=0.3, aerosol_scale=1.0, azimuth=0.0,
         simple_kernel=False, wavelengths=None, **extra):
    """Returns (scene, config, oracle params) for one configuration."""
    lam = list(pysky.midpoint_grid(n_lambda)) if wavelengths is None else list(wavelengths)
    atm = pysky.AtmosphereParams(ground_albedo=albedo)
    atm.aerosol_sca_550_m_inv *= aerosol_scale
    atm.aerosol_abs_550_m_inv *= aerosol_scale
    scene = pysky.SceneData(atmosphere=atm, sun_elevation_deg=sun, sun_azimuth_deg=azimuth,
                            wavelengths_nm=lam)
    fl = dict(FLAGS[mode])
    cfg = pysky.RenderConfig(width=width, height=height, spp=spp, seed=seed, max_depth=max_depth,
                             majorant_layers=layers, output_projection=pysky.OutputProjection(projection),
                             sample_offset=sample_offset, simple_kernel=simple_kernel, **fl, **extra)
    op = O.Params(width=width, height=height, spp=spp, seed=seed, sun_elevation_deg=sun,
                  sun_azimuth_deg=azimuth, max_depth=max_depth, majorant_layers=layers,
                  projection=projection, sample_offset=sample_offset, ground_albedo=albedo,
                  aerosol_sca_550_m_inv=atm.aerosol_sca_550_m_inv,
                  aerosol_abs_550_m_inv=atm.aerosol_abs_550_m_inv, lambda_nm=lam,
                  flags=oracle_flags(**fl),
                  **{k: v for k, v in extra.items() if k in ("direct_light_samples", "rr_start_depth",
                                                             "watchdog_limit", "max_segments",
                                                             "observer_altitude_km", "shell_growth")})
    return scene, cfg, op


def paired_stats(gpu: np.ndarray, ref: np.ndarray, rtol=1e-3, atol_frac=1e-6):
    """Fraction of entries that agree to rtol (relative to the entry or to
    atol_frac of the largest reference value)."""
    gpu = np.asarray(gpu, dtype=np.float64).ravel()
    ref = np.asarray(ref, dtype=np.float64).ravel()
    scale = max(np.abs(ref).max(), 1e-30)
    tol = np.maximum(rtol * np.abs(ref), atol_frac * scale)
    ok = np.abs(gpu - ref) <= tol
    return ok.mean(), np.where(~ok)[0]


def zscores(mean_a, m2_a, n_a, mean_b, m2_b, n_b):
    """Per-cell z of two independent Monte Carlo means from first/second moments."""
    va = np.maximum(m2_a - mean_a ** 2, 0.0) / n_a
    vb = np.maximum(m2_b - mean_b ** 2, 0.0) / n_b
    v = va + vb
    mask = v > 0
    z = np.zeros_like(mean_a)
    z[mask] = (mean_a[mask] - mean_b[mask]) / np.sqrt(v[mask])
    return z, mask
