"""Inverse fit of the atmosphere parameters from an XYZ image.

The target is a render at the true parameters (its own seed); the fit starts
from perturbed parameters and runs Adam on the relative squared error of
block-averaged images,
with gradients from the path-replay adjoint (SkyRender: forward + replay with
common random numbers).  Each step renders with random streams of its own
(independent_seed), so the gradient is an unbiased estimate of the expected
loss's gradient; Adam works
in log-parameter space because the parameters span many orders of magnitude
(m^-1 coefficients ~1e-6, heights ~1-25 km, g and albedo ~0.1-1).  The
majorant grid is rebuilt from the current parameters every step (forward and
replay of one step share it).
"""
from __future__ import annotations

import dataclasses
import math
import time
from typing import Callable, Optional

import numpy as np

from .adjoint import params_vector, sky_render, with_params
from .api import Context, RenderConfig, SceneData, independent_seed


def _blocks(x, b: int):
    """[H, W, 3] -> [H/b, W/b, 3] block means (noise / b, detail kept at b px)."""
    h, w, c = x.shape
    return x.reshape(h // b, b, w // b, b, c).mean(dim=(1, 3))


def fit(scene: SceneData, config: RenderConfig, target_xyz, theta0, steps: int = 200, lr: float = 0.01,
        seed0: int = 1000, block: int = 4, context: Optional[Context] = None,
        log: Optional[Callable[[str], None]] = None) -> dict:
    """Returns dict(theta, history, seconds).  target_xyz: float32 CUDA
    tensor [H, W, 3]; theta0: starting parameters (sky_hip.h order).  The
    loss is the mean squared relative error of block means outside the sun
    disk, so the direct sun (10^4 times brighter than the sky) does not
    outweigh the sky."""
    import torch

    dev = target_xyz.device
    tb = _blocks(target_xyz.double(), block)
    scale = tb + 1e-3 * float(tb.mean())
    # blocks that contain the sun disk: a rare, huge direct term whose Monte
    # Carlo noise would drown the sky's signal; they are left out of the loss
    lum = tb[..., 1]
    keep = (lum < 20.0 * lum.median()).to(torch.float64)[..., None]
    n_keep = float(keep.sum()) * tb.shape[-1]
    u = torch.tensor(np.log(np.abs(np.asarray(theta0, dtype=np.float64))), dtype=torch.float64, device=dev,
                     requires_grad=True)
    sign = torch.tensor(np.sign(np.asarray(theta0, dtype=np.float64)), dtype=torch.float64, device=dev)
    opt = torch.optim.Adam([u], lr=lr)
    # cosine decay to 10% of lr; the estimate is the mean of log theta over
    # the last fifth of the steps (the iterates' noise averages out)
    sched = torch.optim.lr_scheduler.LambdaLR(opt, lambda i: 0.1 + 0.45 * (1.0 + math.cos(math.pi * i / steps)))
    history = []
    tail = []
    t0 = time.perf_counter()
    for it in range(steps):
        cfg = dataclasses.replace(config, seed=independent_seed(seed0 + it), majorant_table=None)
        opt.zero_grad()
        theta = sign * torch.exp(u)
        xyz = sky_render(theta, scene, cfg, context)
        loss = ((((_blocks(xyz.double(), block) - tb) / scale) ** 2) * keep).sum() / n_keep
        loss.backward()
        opt.step()
        sched.step()
        history.append(float(loss.detach()))
        if it >= steps - max(steps // 5, 1):
            tail.append(u.detach().clone())
        if log and (it % 20 == 0 or it == steps - 1):
            log(f"[fit] step {it}: loss {history[-1]:.4e}")
    torch.cuda.synchronize()
    theta = (sign * torch.exp(torch.stack(tail).mean(dim=0))).cpu().numpy()
    return dict(theta=theta, history=history, seconds=time.perf_counter() - t0)


def perturbed(scene: SceneData, rel: float = 0.2, seed: int = 11) -> np.ndarray:
    """True parameters scaled by independent factors 1 +/- rel (random sign)."""
    rng = np.random.default_rng(seed)
    theta = params_vector(scene.atmosphere)
    return theta * (1.0 + rel * rng.choice([-1.0, 1.0], size=theta.size))


__all__ = ["fit", "perturbed", "with_params"]
