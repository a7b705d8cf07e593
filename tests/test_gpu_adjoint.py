"""Path-replay adjoint against central finite differences with common random
numbers (same seed, same fixed majorant grid) on the GPU.

For the ground albedo and the aerosol asymmetry g, with isotropic phase
sampling and no roulette (max_depth 4 < rr_start 4 + 1), no sampling decision
depends on the parameter: each path's contribution is a smooth function of
it, so FD with common random numbers and the adjoint must agree to float
precision.  For the medium parameters, whose accept/reject tests move with the
parameter, the two are compared as estimators of the same derivative over
seeds with random streams of their own (pysky.independent_seed), with no
weight on the pixels around the sun disk: there the derivative of the sun's
transmitted radiance reaches FD only through rare flips of a camera ray's
accept decisions, which leaves the FD estimate heavy-tailed (a few seeds
carry it), while the adjoint sees it on every path through its score."""
import dataclasses

import numpy as np
import pytest

import pysky
from pysky import PARAM_NAMES
from pysky.adjoint import fixed_majorant, params_vector, render_adjoint, with_params

pytestmark = pytest.mark.gpu


def _setup(spp=64, mode=None, seed=1, w=64, h=32, nl=8, max_depth=4, mask_sun=False):
    scene = pysky.SceneData(sun_elevation_deg=10.0, wavelengths_nm=list(pysky.midpoint_grid(nl)))
    est = mode if mode is not None else dict(carried_tau=True)
    cfg = pysky.RenderConfig(width=w, height=h, spp=spp, seed=seed, max_depth=max_depth, **est)
    cfg.majorant_table = fixed_majorant(scene, cfg, 1.25)
    rng = np.random.default_rng(42)
    cot = rng.uniform(0.5, 1.5, size=(h, w, 3)).astype(np.float32)
    if mask_sun:
        # the panorama puts azimuth a at x = (a / 2 pi + 1/2) w, elevation e at
        # y = (1/2 - e / pi) h (sky_trace.wgsl:615-621)
        xs = int((np.radians(scene.sun_azimuth_deg) / (2 * np.pi) + 0.5) * w) % w
        ys = int((0.5 - np.radians(scene.sun_elevation_deg) / np.pi) * h)
        cot[max(ys - 1, 0):ys + 2, max(xs - 1, 0):xs + 2] = 0.0
    return scene, cfg, cot


def _J(scene, cfg, cot, theta, ctx):
    sc = dataclasses.replace(scene, atmosphere=with_params(scene.atmosphere, theta))
    film = pysky.render(sc, cfg, want_film=False, context=ctx)
    return float(np.sum(cot.astype(np.float64) * film.xyz.reshape(cot.shape)))


def _fd(scene, cfg, cot, k, rel, ctx):
    th = params_vector(scene.atmosphere)
    hstep = rel * th[k]
    tp, tm = th.copy(), th.copy()
    tp[k] += hstep
    tm[k] -= hstep
    return (_J(scene, cfg, cot, tp, ctx) - _J(scene, cfg, cot, tm, ctx)) / (2 * hstep)


def test_zero_cotangent_gives_zero_gradient(gpu_context):
    scene, cfg, cot = _setup(spp=4)
    g, _ = render_adjoint(scene, cfg, np.zeros_like(cot), context=gpu_context)
    assert np.all(g == 0.0)


@pytest.mark.parametrize("k", [PARAM_NAMES.index("ground_albedo"), PARAM_NAMES.index("aerosol_g")])
def test_smooth_parameters_match_fd_exactly(gpu_context, k):
    scene, cfg, cot = _setup(spp=32)
    g, st = render_adjoint(scene, cfg, cot, context=gpu_context)
    fd = _fd(scene, cfg, cot, k, 1e-3, gpu_context)
    print(f"[gpu] {PARAM_NAMES[k]}: adjoint {g[k]:.6e} fd {fd:.6e} ({st.kernel_ms:.1f} ms)", flush=True)
    assert g[k] == pytest.approx(fd, rel=2e-3)


def test_medium_parameters_match_fd_statistically(gpu_context):
    seeds = range(1, 9)
    adj, fd = [], []
    ks = [0, 1, 2, 3, 4, 5, 6, 8]
    for sd in seeds:
        scene, cfg, cot = _setup(spp=64, seed=pysky.independent_seed(sd), mask_sun=True)
        g, _ = render_adjoint(scene, cfg, cot, context=gpu_context)
        adj.append(g[ks])
        fd.append([_fd(scene, cfg, cot, k, 1e-2, gpu_context) for k in ks])
    adj, fd = np.array(adj), np.array(fd)
    diff = adj - fd
    se = diff.std(axis=0, ddof=1) / np.sqrt(len(seeds))
    for i, k in enumerate(ks):
        print(f"[gpu] {PARAM_NAMES[k]}: adjoint {adj[:, i].mean():.5e} fd {fd[:, i].mean():.5e} "
              f"diff/se {diff[:, i].mean() / max(se[i], 1e-300):.2f}", flush=True)
    assert np.all(np.abs(diff.mean(axis=0)) <= 4 * se + 1e-3 * np.abs(fd.mean(axis=0)))


FAST_ESTIMATORS = {
    # the C5 estimator (carried optical depth, lobe sampling), 3 shells growing 4x
    "carried_is": dict(carried_tau=True, phase_importance_sampling=True, majorant_layers=3, shell_growth=4.0),
    # the same with ratio-tracked shadow rays
    "carried_is_ratio": dict(carried_tau=True, phase_importance_sampling=True, ratio_tracking=True,
                             majorant_layers=4, shell_growth=2.5),
    # the reference's estimator (a fresh draw per shell, isotropic directions)
    "reference": dict(majorant_layers=2, shell_growth=6.0),
}


@pytest.mark.parametrize("estimator", sorted(FAST_ESTIMATORS))
def test_fast_adjoint_matches_fd_statistically(gpu_context, estimator):
    """The specialised replay kernels against central FD over seeds.  With
    lobe sampling the direction draws move with g and the scattering
    coefficients, so every parameter is compared as an estimator; with the
    reference's isotropic directions g only weights the paths (smooth, so
    checked against FD with a small step in test_fast_adjoint_smooth_g)."""
    est = FAST_ESTIMATORS[estimator]
    adj, fd = [], []
    ks = [k for k in range(10) if not (estimator == "reference" and PARAM_NAMES[k] == "aerosol_g")]
    for sd in range(1, 9):
        scene, cfg, cot = _setup(spp=64, seed=pysky.independent_seed(sd), mode=est, mask_sun=True)
        g, _ = render_adjoint(scene, cfg, cot, context=gpu_context)
        adj.append(g[ks])
        fd.append([_fd(scene, cfg, cot, k, 1e-2, gpu_context) for k in ks])
    adj, fd = np.array(adj), np.array(fd)
    diff = adj - fd
    se = diff.std(axis=0, ddof=1) / np.sqrt(len(adj))
    for i, k in enumerate(ks):
        print(f"[gpu] fast {estimator} {PARAM_NAMES[k]}: adjoint {adj[:, i].mean():.5e} fd {fd[:, i].mean():.5e} "
              f"diff/se {diff[:, i].mean() / max(se[i], 1e-300):.2f}", flush=True)
    assert np.all(np.abs(diff.mean(axis=0)) <= 4 * se + 1e-3 * np.abs(fd.mean(axis=0)))


def test_fast_adjoint_smooth_g(gpu_context):
    """The reference estimator's replay (isotropic directions, no roulette at
    depth 4): g enters only through the phase weights, so FD with common
    random numbers and a small step agrees with the adjoint up to the FP32
    rounding of the two renders."""
    scene, cfg, cot = _setup(spp=32, mode=FAST_ESTIMATORS["reference"])
    k = PARAM_NAMES.index("aerosol_g")
    g, _ = render_adjoint(scene, cfg, cot, context=gpu_context)
    fd = _fd(scene, cfg, cot, k, 1e-3, gpu_context)
    print(f"[gpu] fast reference aerosol_g: adjoint {g[k]:.6e} fd {fd:.6e}", flush=True)
    assert g[k] == pytest.approx(fd, rel=3e-3)


def test_autograd_function(gpu_context):
    torch = pytest.importorskip("torch")
    from pysky.adjoint import sky_render

    scene, cfg, cot = _setup(spp=16)
    theta = torch.tensor(params_vector(scene.atmosphere), dtype=torch.float64, requires_grad=True)
    xyz = sky_render(theta, scene, cfg, gpu_context)
    loss = (xyz * torch.as_tensor(cot, device=xyz.device)).sum()
    loss.backward()
    g, _ = render_adjoint(scene, cfg, cot, context=gpu_context)
    np.testing.assert_allclose(theta.grad.numpy(), g, rtol=1e-6, atol=0)


def test_inverse_fit_moves_toward_truth(gpu_context):
    """Adam with adjoint gradients (the inverse-fit workload, small): from
    +/-20% perturbed parameters the block-mean relative loss drops and the
    best-identified parameters (Rayleigh scale and height, ground albedo)
    move toward their true values."""
    torch = pytest.importorskip("torch")
    from pysky import fit
    from pysky.torch_ops import render_device

    scene = pysky.SceneData(sun_elevation_deg=10.0, wavelengths_nm=list(pysky.midpoint_grid(16)))
    cfg = pysky.RenderConfig(width=128, height=64, spp=256, seed=7, max_depth=8, carried_tau=True,
                             phase_importance_sampling=True, majorant_layers=8)
    target = torch.zeros((64, 128, 3), dtype=torch.float32, device="cuda:0")
    render_device(scene, cfg, xyz=target, context=gpu_context)
    truth = params_vector(scene.atmosphere)
    theta0 = fit.perturbed(scene)
    res = fit.fit(scene, cfg, target, theta0, steps=60, lr=0.01, context=gpu_context)
    h = res["history"]
    err0 = np.abs(theta0 / truth - 1.0)
    err = np.abs(res["theta"] / truth - 1.0)
    print(f"[gpu] fit loss {h[0]:.3e} -> {np.mean(h[-10:]):.3e}; errors "
          + " ".join(f"{n}={e:.3f}" for n, e in zip(PARAM_NAMES, err)), flush=True)
    assert np.mean(h[-10:]) < 0.2 * h[0]
    for name in ("rayleigh_sca_550_m_inv", "rayleigh_h_km", "ground_albedo"):
        k = PARAM_NAMES.index(name)
        assert err[k] < 0.5 * err0[k], name
