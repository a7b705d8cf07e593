"""The specialised integrator kernels (k_wavepool_fast*: carried optical depth
and lobe sampling; k_wavepool_fastrt*: the same with ratio-tracked shadow rays;
k_wavepool_fastref*: the reference's estimator, a fresh draw per shell and
isotropic directions; all with at most 8 majorant shells held in LDS) against
the general kernels on the same paths.

A single render with one of those estimators launches a specialised kernel; a
batch whose jobs use different estimator flags launches the general one.  Paths are
keyed the same way in both, so the films agree up to the order of the float
sums and the specialised handler's reciprocal weight divisions (within 1 ulp
of the quotients): entry by entry to 1e-4, except where a roulette or accept
decision sits within an ulp of its threshold."""
import dataclasses
import os

import numpy as np
import pytest

import pysky
from helpers import paired_stats
from pysky import tables as T

pytestmark = pytest.mark.gpu

DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_data")


def _fast_and_general(ctx, scene, cfg):
    other = dataclasses.replace(cfg, carried_tau=not cfg.carried_tau, phase_importance_sampling=False, width=8,
                                height=4, spp=1)
    general = pysky.render_batch([(scene, cfg), (scene, other)], want_film=True, use_graph=False,
                                 context=ctx).films[0].values()
    fast = pysky.render(scene, cfg, context=ctx).values()
    return fast, general


ESTIMATORS = {
    "carried_is": dict(carried_tau=True, phase_importance_sampling=True),
    "carried_is_ratio": dict(carried_tau=True, phase_importance_sampling=True, ratio_tracking=True),
    "reference": dict(),
}


def _analytic(sun, albedo, layers, growth, estimator="carried_is"):
    atm = pysky.AtmosphereParams(ground_albedo=albedo)
    scene = pysky.SceneData(atmosphere=atm, sun_elevation_deg=sun, wavelengths_nm=list(pysky.midpoint_grid(16)))
    cfg = pysky.RenderConfig(width=64, height=32, spp=32, seed=3, max_depth=16, rr_start_depth=4,
                             majorant_layers=layers, shell_growth=growth, **ESTIMATORS[estimator])
    return scene, cfg


@pytest.mark.parametrize("estimator", sorted(ESTIMATORS))
@pytest.mark.parametrize("sun,albedo,layers,growth", [(5.0, 0.3, 3, 4.0), (-4.0, 0.8, 8, 1.0), (30.0, 0.3, 1, 1.0)])
def test_fast_kernel_traces_the_general_kernels_paths(gpu_context, sun, albedo, layers, growth, estimator):
    scene, cfg = _analytic(sun, albedo, layers, growth, estimator)
    print(f"[gpu] fast vs general: sun {sun} albedo {albedo} shells {layers} growth {growth} {estimator}",
          flush=True)
    fast, general = _fast_and_general(gpu_context, scene, cfg)
    # A ratio-tracked shadow ray's transmittance is a product of 1 - sigma_t /
    # majorant over every tentative collision; on grazing rays (sun below the
    # horizon) last-bit differences move it by more than 1e-4, as they move the
    # FP32 oracle against the FP64 one, so it is paired as the oracle tests do.
    ratio = estimator == "carried_is_ratio"
    frac, bad = paired_stats(fast, general, rtol=1e-2 if ratio else 1e-4)
    print(f"paired fraction {frac:.6f} ({len(bad)} of {fast.size} differ)")
    assert frac >= (0.995 if ratio else 0.999)
    assert np.all(np.isfinite(fast)) and np.any(fast > 0)
    np.testing.assert_allclose(fast.sum(), general.sum(), rtol=1e-3)


def test_fast_kernel_on_the_reference_tables(gpu_context):
    scene = T.load_reference_scene(DATA, 5.0)
    cfg = pysky.RenderConfig(width=48, height=24, spp=16, seed=4, max_depth=8, majorant_layers=3,
                             shell_growth=6.0, carried_tau=True, phase_importance_sampling=True)
    print("[gpu] fast vs general on the reference tables", flush=True)
    fast, general = _fast_and_general(gpu_context, scene, cfg)
    frac, bad = paired_stats(fast, general, rtol=1e-4)
    print(f"paired fraction {frac:.6f} ({len(bad)} of {fast.size} differ)")
    assert frac >= 0.999
