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
    """The objective sum cot * XYZ of the render at theta, from the replay's
    FP64 sum over the render's own paths (stats.adjoint_value): the FP32 XYZ
    image rounds J at ~1e-7 relative, which at h = 1e-3 is ~1e-3 of a
    finite difference (test_replay_objective_is_the_render_objective)."""
    sc = dataclasses.replace(scene, atmosphere=with_params(scene.atmosphere, theta))
    _, st = render_adjoint(sc, cfg, cot, context=ctx)
    return st.adjoint_value


def _fd(scene, cfg, cot, k, rel, ctx):
    th = params_vector(scene.atmosphere)
    hstep = rel * th[k]
    tp, tm = th.copy(), th.copy()
    tp[k] += hstep
    tm[k] -= hstep
    return (_J(scene, cfg, cot, tp, ctx) - _J(scene, cfg, cot, tm, ctx)) / (2 * hstep)


@pytest.mark.parametrize("estimator", ["general", "carried_is", "reference"])
def test_replay_objective_is_the_render_objective(gpu_context, estimator):
    """The replay traces the render's paths: its FP64 objective equals
    sum cot * XYZ of the FP32 image up to that image's rounding."""
    mode = None if estimator == "general" else FAST_ESTIMATORS[estimator]
    scene, cfg, cot = _setup(spp=64, mode=mode, max_depth=8)
    film = pysky.render(scene, cfg, want_film=False, context=gpu_context)
    j32 = float(np.sum(cot.astype(np.float64) * film.xyz.reshape(cot.shape)))
    _, st = render_adjoint(scene, cfg, cot, context=gpu_context)
    print(f"[gpu] {estimator}: J image {j32:.9e} replay {st.adjoint_value:.9e} "
          f"rel {st.adjoint_value / j32 - 1:+.1e}", flush=True)
    assert st.adjoint_value == pytest.approx(j32, rel=2e-6)


def test_zero_cotangent_gives_zero_gradient(gpu_context):
    scene, cfg, cot = _setup(spp=4)
    g, _ = render_adjoint(scene, cfg, np.zeros_like(cot), context=gpu_context)
    assert np.all(g == 0.0)


@pytest.mark.parametrize("k", [PARAM_NAMES.index("ground_albedo"), PARAM_NAMES.index("aerosol_g")])
def test_smooth_parameters_match_fd_exactly(gpu_context, k):
    scene, cfg, cot = _setup(spp=32)
    g, st = render_adjoint(scene, cfg, cot, context=gpu_context)
    fd = _fd(scene, cfg, cot, k, 1e-3, gpu_context)
    print(f"[gpu] {PARAM_NAMES[k]}: adjoint {g[k]:.6e} fd {fd:.6e} rel {g[k] / fd - 1:+.2e} "
          f"({st.kernel_ms:.1f} ms)", flush=True)
    assert g[k] == pytest.approx(fd, rel=1e-3)


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
    print(f"[gpu] fast reference aerosol_g: adjoint {g[k]:.6e} fd {fd:.6e} rel {g[k] / fd - 1:+.2e}", flush=True)
    assert g[k] == pytest.approx(fd, rel=1e-3)


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


# ------------------------------------------------- the C5 configuration
# 256x128, 32 bands, 256 spp, sun 10 deg (the adjoint benchmark's render)

@pytest.mark.parametrize("estimator,k", [("carried", "ground_albedo"), ("carried", "aerosol_g"),
                                         ("reference", "ground_albedo"), ("reference", "aerosol_g")])
def test_c5_smooth_parameters_match_fd(gpu_context, estimator, k):
    """At the full C5 size, without roulette (depth 4) and with isotropic
    directions, the albedo and g only weight each path: FD with common random
    numbers (h = 1e-3 relative) and the adjoint agree to 1e-3."""
    est = dict(carried_tau=True) if estimator == "carried" else FAST_ESTIMATORS["reference"]
    scene, cfg, cot = _setup(spp=256, mode=est, w=256, h=128, nl=32, max_depth=4)
    ki = PARAM_NAMES.index(k)
    g, st = render_adjoint(scene, cfg, cot, context=gpu_context)
    fd = _fd(scene, cfg, cot, ki, 1e-3, gpu_context)
    print(f"[gpu] c5 {estimator} {k}: adjoint {g[ki]:.6e} fd {fd:.6e} rel {g[ki] / fd - 1:+.2e} "
          f"(replay {st.kernel_ms:.1f} ms)", flush=True)
    assert g[ki] == pytest.approx(fd, rel=1e-3)


def test_c5_medium_parameters_fd_agreement(gpu_context):
    """The C5 estimator (carried optical depth, lobe sampling, 3 shells
    growing 4x, the specialised replay) at depth <= 8 with roulette from
    depth 4, so roulette is active in the replay: adjoint against central FD
    (h = 1e-2) over 8 renders with independent streams, sun disk masked.
    Prints each parameter's relative difference and its standard error."""
    est = FAST_ESTIMATORS["carried_is"]
    adj, fd = [], []
    for sd in range(1, 9):
        scene, cfg, cot = _setup(spp=256, seed=pysky.independent_seed(sd), mode=est, w=256, h=128, nl=32,
                                 max_depth=8, mask_sun=True)
        assert cfg.rr_start_depth == 4
        g, _ = render_adjoint(scene, cfg, cot, context=gpu_context)
        adj.append(g)
        fd.append([_fd(scene, cfg, cot, k, 1e-2, gpu_context) for k in range(10)])
    adj, fd = np.array(adj), np.array(fd)
    diff = adj - fd
    se = diff.std(axis=0, ddof=1) / np.sqrt(len(adj))
    for k in range(10):
        m = fd[:, k].mean()
        print(f"[gpu] c5 {PARAM_NAMES[k]}: adjoint {adj[:, k].mean():+.5e} fd {m:+.5e} "
              f"relative difference {diff[:, k].mean() / m:+.2e} +- {se[k] / abs(m):.1e}", flush=True)
    assert np.all(np.abs(diff.mean(axis=0)) <= 4 * se + 1e-3 * np.abs(fd.mean(axis=0)))


def test_sun_disk_adjoint_is_transmittance_derivative(gpu_context):
    """With the cotangent on the 3x3 pixels around the sun only (1024x512, so
    the 0.27 deg disk covers a few of them), J is the disk seen through the
    atmosphere, L_sun exp(-tau), plus in-scattered light J_in.  J'/J is then
    -d tau / d theta along the sun's direction, per band weighted by the
    band's share of J: the adjoint's score of the camera rays' accept/reject
    decisions against FD of the deterministic FP64 optical depth (the
    oracle's Gauss-Legendre quadrature).  J_in is measured from the ring of
    16 disk-free pixels around the block (the g = 0.8 aureole is flat over
    a few degrees) and bounds the remainder: |d J_in / d log theta| / J <=
    10 J_in / J (the forward HG lobe's d log p / d log g is 8.5; every
    other elasticity of in-scattered light is below 2)."""
    import oracle as O

    est = FAST_ESTIMATORS["carried_is"]
    w, h, nl = 1024, 512, 32
    ratios, shares, jin = [], None, []
    for sd in range(1, 9):
        scene, cfg, _ = _setup(spp=64, seed=pysky.independent_seed(sd), mode=est, w=w, h=h, nl=nl, max_depth=8)
        xs = int((np.radians(scene.sun_azimuth_deg) / (2 * np.pi) + 0.5) * w) % w
        ys = int((0.5 - np.radians(scene.sun_elevation_deg) / np.pi) * h)
        cot = np.zeros((h, w, 3), np.float32)
        cot[ys - 1:ys + 2, xs - 1:xs + 2] = 1.0
        f = pysky.render(scene, cfg, want_film=True, context=gpu_context)
        img = f.xyz.reshape(h, w, 3).astype(np.float64).sum(axis=2)
        J = float(img[ys - 1:ys + 2, xs - 1:xs + 2].sum())
        ring = float(img[ys - 2:ys + 3, xs - 2:xs + 3].sum()) - J
        jin.append(9.0 * ring / 16.0 / J)
        g, _ = render_adjoint(scene, cfg, cot, context=gpu_context)
        ratios.append(g / J)
        wsum = pysky.wavelength_table(scene)[:, 5:8].astype(np.float64).sum(axis=1)
        sh = f.values().reshape(h, w, nl)[ys - 1:ys + 2, xs - 1:xs + 2].astype(np.float64).sum(axis=(0, 1)) * wsum
        shares = sh if shares is None else shares + sh
    shares /= shares.sum()
    ratios = np.array(ratios)
    rem = 10.0 * float(np.mean(jin))
    # -d tau_b / d theta_k along the sun's centre ray, FD of the FP64 quadrature
    theta0 = params_vector(scene.atmosphere)
    sun = O.sun_direction(O.Params(sun_elevation_deg=scene.sun_elevation_deg))
    o = np.array([0.0, 6360.0 + cfg.observer_altitude_km, 0.0])
    b = float(o @ sun)
    t_top = -b + np.sqrt(b * b - (o @ o - 6460.0 ** 2))

    def tau(theta):
        a = with_params(scene.atmosphere, theta)
        op = O.Params(lambda_nm=list(scene.wavelengths_nm), **{n: getattr(a, n) for n in PARAM_NAMES})
        return np.array([O.optical_depth(op, bnd, o, sun, t_top) for bnd in range(nl)])

    target = np.zeros(10)
    for k in range(10):
        tp, tm = theta0.copy(), theta0.copy()
        hk = 1e-4 * abs(theta0[k])
        tp[k] += hk
        tm[k] -= hk
        target[k] = -float(shares @ ((tau(tp) - tau(tm)) / (2 * hk)))
    mean = ratios.mean(axis=0)
    se = ratios.std(axis=0, ddof=1) / np.sqrt(len(ratios))
    print(f"[gpu] sun disk: in-scattered share of J {np.mean(jin):.2e}", flush=True)
    for k in range(10):
        print(f"[gpu] sun disk {PARAM_NAMES[k]}: d log J / d log theta {mean[k] * theta0[k]:+.5e} "
              f"+- {se[k] * theta0[k]:.1e}, -d tau / d log theta {target[k] * theta0[k]:+.5e}", flush=True)
    assert np.mean(jin) < 2e-4
    for k in range(10):
        err = abs(mean[k] - target[k]) * theta0[k]
        assert err <= 4 * se[k] * theta0[k] + 1e-2 * abs(target[k] * theta0[k]) + rem, PARAM_NAMES[k]
