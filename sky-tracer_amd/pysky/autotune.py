"""Autotuner of the wave-pool integrator's launch parameters.

Knobs (none changes the estimator's expectation; each is checked by the
parity tests): majorant shells L and their thickness growth g (delta
tracking samples the exact free-flight law for any majorant),
event-batching threshold, tracking
micro-steps per handler pass, register budget (occupancy variant),
persistent waves per CU, pixels per work item (lambda packing: pixels x
wavelengths per wave).  Coordinate descent over short renders of the actual
workload, timed with the library's HIP events; results are cached per shape
class in tuning_cache.json next to this file.
"""
from __future__ import annotations

import dataclasses
import json
import os
import time
from typing import Dict, Optional, Tuple

from .api import Context, RenderConfig, SceneData, render

CACHE_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning_cache.json")

# the library's defaults for the launch shape (zero tuning fields), 16 uniform shells
DEFAULT = dict(majorant_layers=16, event_threshold=8, track_steps=2, occupancy=0, waves_per_cu=24,
               pixels_per_item=0, shell_growth=1.0)
SPACE = {
    "majorant_layers": [2, 3, 4, 5, 6, 7, 8, 12, 16, 24, 32, 64],
    "shell_growth": [1.0, 1.5, 2.0, 2.5, 3.0, 4.0, 5.0, 6.0, 8.0],
    "event_threshold": [8, 16, 24, 32, 40, 48, 56],
    "track_steps": [1, 2, 4, 8, 16, 32],
    "occupancy": [0, 4, 5, 6, 8],
    "waves_per_cu": [8, 12, 16, 24, 32],
    "pixels_per_item": [0, 1, 2, 4, 8],  # 0: the library's spp-aware default
}


def shape_class(scene: SceneData, config: RenderConfig) -> str:
    nb = scene.band_count
    npx = config.width * config.height
    nb_cls = 3 if nb <= 3 else 4 if nb <= 4 else 8 if nb <= 8 else 16 if nb <= 16 else 32 if nb <= 32 else 64
    px_cls = "small" if npx <= 8192 else "medium" if npx <= 131072 else "large"
    est = f"is{int(config.phase_importance_sampling)}ct{int(config.carried_tau)}rt{int(config.ratio_tracking)}"
    medium = ":tab" if getattr(scene.atmosphere, "tables", None) is not None else ""
    return f"nl{nb_cls}:{px_cls}:{est}{medium}"


def load_cache() -> Dict[str, dict]:
    if os.path.exists(CACHE_PATH):
        with open(CACHE_PATH) as f:
            return json.load(f)
    return {}


def tuned(scene: SceneData, config: RenderConfig) -> dict:
    """Cached tuning for the configuration's shape class (defaults if none)."""
    return dict(DEFAULT, **load_cache().get(shape_class(scene, config), {}).get("params", {}))


def apply(ctx: Context, config: RenderConfig, params: dict) -> RenderConfig:
    ctx.set_tuning(waves_per_cu=params["waves_per_cu"], pixels_per_item=params["pixels_per_item"],
                   event_threshold=params["event_threshold"], track_steps=params["track_steps"],
                   occupancy=params["occupancy"])
    return dataclasses.replace(config, majorant_layers=params["majorant_layers"],
                               shell_growth=params.get("shell_growth", 1.0), majorant_table=None)


def _time(ctx: Context, scene: SceneData, config: RenderConfig, params: dict, reps: int = 2) -> float:
    cfg = apply(ctx, config, params)
    best = float("inf")
    for _ in range(reps):
        film = render(scene, cfg, want_film=False, want_xyz=True, context=ctx)
        best = min(best, film.stats.kernel_ms)
    return best


def autotune(scene: SceneData, config: RenderConfig, ctx: Context, probe_spp: int = 16,
             rounds: int = 2, log=print, save: bool = True, split_spp: int = 0) -> Tuple[dict, float]:
    """Coordinate descent from the cached or default point; returns (params,
    best kernel ms at probe_spp).  With split_spp, the work split (pixels
    per item, 0 = the library's spp-aware default) is then chosen again at
    that many samples per pixel: an item's length decides how much of it
    drains with idle lanes and how long the launch's last items run, so the
    split depends on the spp the renders actually use."""
    probe = dataclasses.replace(config, spp=probe_spp, sample_offset=0)
    cur = tuned(scene, config)
    best = _time(ctx, scene, probe, cur)
    log(f"[autotune] start {cur} -> {best:.1f} ms")
    for _ in range(rounds):
        improved = False
        for knob, values in SPACE.items():
            for v in values:
                if v == cur[knob]:
                    continue
                cand = dict(cur, **{knob: v})
                t0 = time.perf_counter()
                ms = _time(ctx, scene, probe, cand)
                log(f"[autotune] {knob}={v}: {ms:.1f} ms ({time.perf_counter() - t0:.1f} s)")
                if ms < best * 0.995:
                    best, cur, improved = ms, cand, True
        if not improved:
            break
    split = None
    if split_spp:
        wide = dataclasses.replace(config, spp=split_spp, sample_offset=0)
        times = {}
        for v in SPACE["pixels_per_item"]:
            times[v] = _time(ctx, scene, wide, dict(cur, pixels_per_item=v), reps=1)
            log(f"[autotune] pixels_per_item={v} at {split_spp} spp: {times[v]:.1f} ms")
        pick = min(times, key=times.get)
        if 0 in times and times[0] <= times[pick] * 1.005:
            pick = 0  # the spp-aware default, within 0.5% of the best here, adapts to other spp
        if times[pick] < times[cur["pixels_per_item"]] * 0.995:
            cur = dict(cur, pixels_per_item=pick)
        split = {"spp": split_spp, "ms": {str(k): round(v, 1) for k, v in times.items()}}
    log(f"[autotune] best {cur} -> {best:.1f} ms")
    if save:
        cache = load_cache()
        cache[shape_class(scene, config)] = {"params": cur, "probe_ms": best, "probe_spp": probe_spp,
                                             "workload": f"{config.width}x{config.height}x{scene.band_count}"}
        if split:
            cache[shape_class(scene, config)]["split_probe"] = split
        with open(CACHE_PATH, "w") as f:
            json.dump(cache, f, indent=1, sort_keys=True)
    apply(ctx, config, cur)
    return cur, best
