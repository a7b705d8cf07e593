"""Differentiable rendering: the path-replay adjoint of libsky_hip.so.

sky_render_adjoint replays the paths of a render (same RNG keys) with the
majorant grid held fixed and returns d/dtheta of sum(dl_dxyz * XYZ) for ten
atmosphere parameters (PARAM_NAMES).  SkyRender wraps forward + adjoint as a
torch.autograd.Function over a float64 parameter tensor.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
from typing import Optional

import numpy as np

from . import _lib as L
from .api import (PARAM_NAMES, AtmosphereParams, Context, RenderConfig, SceneData, check, config_c,
                  default_context, majorant_grid)

L_NP = len(PARAM_NAMES)


def params_vector(atm: AtmosphereParams) -> np.ndarray:
    return np.array([getattr(atm, n) for n in PARAM_NAMES], dtype=np.float64)


def with_params(atm: AtmosphereParams, theta) -> AtmosphereParams:
    vals = {n: float(v) for n, v in zip(PARAM_NAMES, np.asarray(theta, dtype=np.float64))}
    return dataclasses.replace(atm, **vals)


def fixed_majorant(scene: SceneData, config: RenderConfig, margin: float = 1.25) -> np.ndarray:
    """A majorant grid for the scene's current parameters with extra margin,
    to hold fixed while the parameters move (FD, adjoint, fitting)."""
    return majorant_grid(scene, config.majorant_layers, margin, config.shell_growth)


def render_adjoint(scene: SceneData, config: RenderConfig, dl_dxyz: np.ndarray,
                   context: Optional[Context] = None):
    """Host cotangent [H, W, 3] -> (gradient [10], stats)."""
    g = np.ascontiguousarray(dl_dxyz, dtype=np.float32).ravel()
    if g.size != 3 * config.width * config.height:
        raise ValueError("cotangent must have H*W*3 entries")
    grad = np.zeros(L_NP, dtype=np.float64)
    planet, atm, wl = scene.planet_c(), scene.atmosphere.to_c(), scene.wavelengths_c()
    cfg = config_c(scene, config)
    stats = L.RenderStats()
    ctx = (context or default_context()).handle
    check(L.lib().sky_render_adjoint(ctx, C.byref(planet), C.byref(atm), C.byref(wl), C.byref(cfg),
                                     g.ctypes.data_as(C.POINTER(C.c_float)),
                                     grad.ctypes.data_as(C.POINTER(C.c_double)), C.byref(stats)))
    return grad, stats


def render_adjoint_device(scene: SceneData, config: RenderConfig, dl_dxyz, context: Optional[Context] = None):
    """Device cotangent (float32 CUDA tensor, H*W*3 elements) -> (gradient [10], stats)."""
    import torch

    assert dl_dxyz.is_cuda and dl_dxyz.dtype == torch.float32 and dl_dxyz.is_contiguous()
    if dl_dxyz.numel() != 3 * config.width * config.height:
        raise ValueError("cotangent must have H*W*3 entries")
    grad = np.zeros(L_NP, dtype=np.float64)
    planet, atm, wl = scene.planet_c(), scene.atmosphere.to_c(), scene.wavelengths_c()
    cfg = config_c(scene, config)
    stats = L.RenderStats()
    ctx = (context or default_context(torch.cuda.current_device())).handle
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    check(L.lib().sky_render_adjoint_device(ctx, C.byref(planet), C.byref(atm), C.byref(wl), C.byref(cfg),
                                            C.c_void_p(dl_dxyz.data_ptr()),
                                            grad.ctypes.data_as(C.POINTER(C.c_double)), stream, C.byref(stats)))
    return grad, stats


def _make_function():
    import torch

    from .torch_ops import render_device

    class SkyRender(torch.autograd.Function):
        """theta (float64 [10]) -> XYZ image (float32 CUDA [H, W, 3]).

        The backward pass replays the forward's paths (same seed and fixed
        majorant), so gradients use common random numbers with the image."""

        @staticmethod
        def forward(ctx, theta, scene: SceneData, config: RenderConfig, context=None):
            th = theta.detach().cpu().numpy()
            sc = dataclasses.replace(scene, atmosphere=with_params(scene.atmosphere, th))
            xyz = torch.zeros((config.height, config.width, 3), dtype=torch.float32,
                              device=torch.device("cuda", torch.cuda.current_device()))
            render_device(sc, config, xyz=xyz, context=context)
            ctx.scene, ctx.config, ctx.context = sc, config, context
            ctx.theta_device, ctx.theta_dtype = theta.device, theta.dtype
            return xyz

        @staticmethod
        def backward(ctx, grad_xyz):
            g = grad_xyz.detach().to(torch.float32).contiguous()
            grad, _ = render_adjoint_device(ctx.scene, ctx.config, g, context=ctx.context)
            return torch.as_tensor(grad, dtype=ctx.theta_dtype, device=ctx.theta_device), None, None, None

    return SkyRender


_sky_render_fn = None


def sky_render(theta, scene: SceneData, config: RenderConfig, context: Optional[Context] = None):
    """Differentiable render: XYZ image as a function of the parameter tensor."""
    global _sky_render_fn
    if _sky_render_fn is None:
        _sky_render_fn = _make_function()
    return _sky_render_fn.apply(theta, scene, config, context)
