"""sky_render_device_noalloc (GPU): the allocation-free render against
sky_render_device on the same inputs, through the C ABI.

Two calls in a row on a fresh context, the second smaller: the first
(128x64, 16 bands, 64 spp) would split its samples into chunks whose
partials exceed the scratch array, so it runs one chunk per pixel group
(same paths, another float summation order); the second (64x32, 8 bands)
runs the usual work split and must be bit-identical.  Neither call changes
the device's free memory.  A render that cannot fit is refused before any
launch."""
import ctypes as C

import numpy as np
import pytest

import pysky
from helpers import case
from pysky import _lib as L
from pysky.api import config_c

pytestmark = pytest.mark.gpu


def _noalloc(scene, cfg, ctx, xyz, film, m2):
    """Outputs are device tensors of the render's full size, or None."""
    import torch

    npx, nb = cfg.width * cfg.height, scene.band_count
    for t, n in ((xyz, 3 * npx), (film, nb * npx), (m2, nb * npx)):
        assert t is None or (t.is_cuda and t.numel() == n and t.dtype == torch.float32)
    ptr = lambda t: None if t is None else C.c_void_p(t.data_ptr())  # noqa: E731
    planet, atm, wl = scene.planet_c(), scene.atmosphere.to_c(), scene.wavelengths_c()
    cc = config_c(scene, cfg)
    stats = L.RenderStats()
    st = L.lib().sky_render_device_noalloc(ctx.handle, C.byref(planet), C.byref(atm), C.byref(wl), C.byref(cc),
                                           ptr(film), ptr(m2), ptr(xyz),
                                           C.c_void_p(torch.cuda.current_stream().cuda_stream), C.byref(stats))
    return st, stats


def test_noalloc_matches_render_device():
    import torch

    from pysky.torch_ops import render_device

    dev = torch.device("cuda:0")
    ref_ctx = pysky.Context(0)
    ctx = pysky.Context(0)
    calls = [case(width=128, height=64, n_lambda=16, spp=64, mode="carried_is", layers=3, shell_growth=4.0),
             case(width=64, height=32, n_lambda=8, spp=32, mode="reference", layers=8)]
    outs = []
    for scene, cfg, _ in calls:
        npx, nb = cfg.width * cfg.height, scene.band_count
        # [xyz, film, film_m2] of sky_render_device, then of the allocation-free call
        outs.append([torch.zeros(n, dtype=torch.float32, device=dev) for n in (3 * npx, nb * npx, nb * npx) * 2])
    refs = []
    for (scene, cfg, _), o in zip(calls, outs):
        refs.append(render_device(scene, cfg, xyz=o[0], film=o[1], film_m2=o[2], context=ref_ctx))
    torch.cuda.synchronize()
    for i, ((scene, cfg, _), o) in enumerate(zip(calls, outs)):
        print(f"[gpu] noalloc call {i}: {cfg.width}x{cfg.height} x {scene.band_count} bands x {cfg.spp} spp",
              flush=True)
        free0 = torch.cuda.mem_get_info()[0]
        st, stats = _noalloc(scene, cfg, ctx, o[3], o[4], o[5])
        torch.cuda.synchronize()
        free1 = torch.cuda.mem_get_info()[0]
        assert st == L.SKY_OK, L.lib().sky_last_error().decode()
        assert free1 == free0, (free0 - free1)
        assert stats.path_samples == cfg.width * cfg.height * cfg.spp * scene.band_count
        a = [t.cpu().numpy().astype(np.float64) for t in o[:3]]
        b = [t.cpu().numpy().astype(np.float64) for t in o[3:]]
        assert np.all(np.isfinite(b[0])) and b[0].max() > 0
        if i == 0:
            assert stats.items < refs[0].items  # one chunk per pixel group
            for x, y in zip(a, b):
                np.testing.assert_allclose(y, x, rtol=2e-5, atol=1e-6 * x.max())
        else:
            assert stats.items == refs[1].items  # the usual split
            for x, y in zip(a, b):
                np.testing.assert_array_equal(y, x)


def test_noalloc_refuses_what_does_not_fit():
    """4 pixels x 64 bands per item: 256 triples per sample, so an item holds
    at most 2^24 / 256 = 65536 samples and 70000 spp need two chunks, whose
    partials (2 x 1024 x 512 x 64 x 8 bytes) exceed the scratch array.  No
    outputs are passed, so nothing could be written even if it ran."""
    scene, cfg, _ = case(width=1024, height=512, n_lambda=64, spp=70000, mode="carried_is", layers=3,
                         shell_growth=4.0)
    ctx = pysky.Context(0)
    ctx.set_tuning(pixels_per_item=4)
    st, _ = _noalloc(scene, cfg, ctx, None, None, None)
    assert st == L.SKY_ERR_INVALID_CONFIG
    assert "allocation-free" in L.lib().sky_last_error().decode()
