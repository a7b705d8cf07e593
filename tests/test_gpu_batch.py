"""Ragged batches in one persistent launch (GPU)."""
import numpy as np
import pytest

import pysky

pytestmark = pytest.mark.gpu


def _small_batch(n=12, spp=16, seed=5):
    b = pysky.ragged_batch(n=n, seed=seed, spp=spp, max_depth=8)
    # keep the test light: clamp resolutions to <= 128x64
    for scene, cfg in b:
        if cfg.width > 128:
            cfg.width, cfg.height = 128, 64
    return b


def test_batch_matches_individual_renders(gpu_context):
    batch = _small_batch()
    res = pysky.render_batch(batch, want_film=True, context=gpu_context)
    print(f"[gpu] batch of {len(batch)}: kernel {res.stats.kernel_ms:.1f} ms items {res.stats.items}", flush=True)
    for (scene, cfg), film in zip(batch, res.films):
        single = pysky.render(scene, cfg, context=gpu_context)
        np.testing.assert_allclose(film.values(), single.values(), rtol=1e-4, atol=1e-6 * single.values().max())
        np.testing.assert_allclose(film.xyz, single.xyz, rtol=1e-4, atol=1e-6 * single.xyz.max())


def test_graph_replays_match_eager(gpu_context):
    batch = _small_batch(n=8, seed=9)
    eager = pysky.render_batch(batch, use_graph=False, context=gpu_context)
    ref = [f.xyz.copy() for f in eager.films]
    reused = []
    for rep in range(3):
        # eager work between replays
        pysky.render(batch[0][0], batch[0][1], context=gpu_context)
        res = pysky.render_batch(batch, use_graph=True, context=gpu_context)
        reused.append(res.stats.graph_reused)
        for a, b in zip(res.films, ref):
            assert np.array_equal(a.xyz, b), f"replay {rep} differs"
    assert reused[1] and reused[2]


def test_batch_latencies(gpu_context):
    batch = _small_batch(n=16, seed=3)
    res = pysky.render_batch(batch, context=gpu_context)
    assert np.all(res.finish_ms > 0)
    assert res.finish_ms.max() <= res.stats.kernel_ms * 1.05 + 0.5
    assert res.stats.latency_p50_ms <= res.stats.latency_p99_ms <= res.stats.latency_max_ms


def test_batch_invalid_item_reports_index(gpu_context):
    batch = _small_batch(n=3)
    batch[1][1].spp = 0
    with pytest.raises(pysky.InvalidConfig, match="batch item 1"):
        pysky.render_batch(batch, context=gpu_context)
