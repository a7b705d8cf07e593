"""Ragged batches: many independent renders in one persistent launch.

The reference bakes its sky-view set with one process and one device per
render (scripts/bake_skyview_elevations.ps1:13-35); sky_render_batch runs
the whole set as one wave-pool launch (captured in a HIP graph).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import _lib as L
from .api import (AtmosphereParams, Context, Film, RenderConfig, SceneData, check, config_c,
                  default_context)

RESOLUTIONS = ((64, 32), (128, 64), (256, 128), (512, 256))
BAND_COUNTS = (3, 4, 8, 16, 64)


@dataclass
class BatchResult:
    films: List[Film]
    finish_ms: np.ndarray
    stats: L.BatchStats


def ragged_batch(n: int = 256, seed: int = 1, spp: int = 256, max_depth: int = 32,
                 estimator: Optional[dict] = None) -> List[Tuple[SceneData, RenderConfig]]:
    """The ragged workload: sun elevation U[-6, 90] deg, resolution uniform
    over 64x32..512x256, 3..64 wavelengths drawn U[400, 700] nm (sorted, with
    equal quadrature weights over that range), `spp` samples each; render i
    uses seed `seed * 1000003 + i` so every render has its own RNG stream."""
    rng = np.random.default_rng(seed)
    out = []
    est = estimator or dict(carried_tau=True, phase_importance_sampling=True)
    for i in range(n):
        el = float(rng.uniform(-6.0, 90.0))
        w, h = RESOLUTIONS[int(rng.integers(0, len(RESOLUTIONS)))]
        nb = BAND_COUNTS[int(rng.integers(0, len(BAND_COUNTS)))]
        lam = np.sort(rng.uniform(400.0, 700.0, nb))
        scene = SceneData(atmosphere=AtmosphereParams(), sun_elevation_deg=el, wavelengths_nm=list(lam),
                          weights_nm=list(np.full(nb, 300.0 / nb)))
        cfg = RenderConfig(width=w, height=h, spp=spp, seed=seed * 1000003 + i, max_depth=max_depth, **est)
        out.append((scene, cfg))
    return out


def render_batch(batch: Sequence[Tuple[SceneData, RenderConfig]], *, want_film: bool = False,
                 use_graph: bool = True, context: Optional[Context] = None) -> BatchResult:
    n = len(batch)
    items = (L.BatchItem * n)()
    keep = []
    films = []
    for i, (scene, cfg) in enumerate(batch):
        planet, atm, wl = scene.planet_c(), scene.atmosphere.to_c(), scene.wavelengths_c()
        cc = config_c(scene, cfg)
        npx = cfg.width * cfg.height
        xyz = np.zeros(npx * 3, dtype=np.float32)
        film = np.zeros(npx * scene.band_count, dtype=np.float32) if want_film else None
        keep.append((planet, atm, wl, cc, xyz, film))
        items[i].planet = C.pointer(planet)
        items[i].atm = C.pointer(atm)
        items[i].wl = C.pointer(wl)
        items[i].cfg = C.pointer(cc)
        items[i].xyz = xyz.ctypes.data_as(C.POINTER(C.c_float))
        items[i].film = film.ctypes.data_as(C.POINTER(C.c_float)) if film is not None else None
        films.append(Film(cfg.width, cfg.height, scene.band_count, film, xyz=xyz))
    finish = np.zeros(n, dtype=np.float64)
    stats = L.BatchStats()
    ctx = (context or default_context()).handle
    flags = 0 if use_graph else L.BATCH_NO_GRAPH
    check(L.lib().sky_render_batch(ctx, items, n, flags, finish.ctypes.data_as(C.POINTER(C.c_double)),
                                   C.byref(stats)))
    return BatchResult(films, finish, stats)
