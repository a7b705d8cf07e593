"""Output forms of the integrator (GPU).

A work item holding every sample of its pixels writes their film, second
moment and XYZ straight from its LDS sums; a render whose samples are split
into chunks (small images) writes [chunk][pixel][band] partials that
k_finalize reduces in chunk order.  Both must give the same values, XYZ must
be the weighted film, and an XYZ-only render of the 1024x512 x 64-band shape
must not allocate the partial buffer (268 MB at one chunk per pixel group).
"""
import dataclasses

import numpy as np
import pytest

import oracle as O
import pysky
from helpers import case

pytestmark = pytest.mark.gpu


def _render(scene, cfg, ctx, **kw):
    return pysky.render(scene, cfg, context=ctx, **kw)


@pytest.mark.parametrize("n_lambda,spp", [(16, 64), (3, 48), (64, 32)])
def test_single_chunk_and_chunked_renders_agree(n_lambda, spp):
    scene, cfg, op = case(width=64, height=32, n_lambda=n_lambda, spp=spp, mode="carried_is", layers=3,
                          shell_growth=4.0)
    npx = cfg.width * cfg.height
    chunked = pysky.Context(0)  # default launch: 4 x 256 x 24 items wanted -> samples split in chunks
    whole = pysky.Context(0)
    whole.set_tuning(waves_per_cu=1, pixels_per_item=1)  # 1024 items wanted, one per pixel: one chunk
    a = _render(scene, cfg, chunked, want_m2=True)
    b = _render(scene, cfg, whole, want_m2=True)
    # the library's default split: the fewest pixels per item giving 16384 paths, at most 256 / bands
    ppi_a = min(-(-16384 // (n_lambda * spp)), 256 // n_lambda)
    assert a.stats.items > (npx + ppi_a - 1) // ppi_a  # more items than pixel groups: chunked
    assert b.stats.items == npx  # one item per pixel: single chunk
    fa, fb = a.values().astype(np.float64), b.values().astype(np.float64)
    assert fa.size == npx * n_lambda
    scale = fa.max()
    # the same sums in another order
    np.testing.assert_allclose(fb, fa, rtol=2e-5, atol=1e-6 * scale)
    np.testing.assert_allclose(b.m2.astype(np.float64), a.m2.astype(np.float64), rtol=2e-5,
                               atol=1e-6 * a.m2.max())
    w = O.lambda_table(op)[:, 5:8]
    for f in (a, b):
        xyz = f.values().reshape(-1, n_lambda).astype(np.float64) @ w
        np.testing.assert_allclose(f.xyz.reshape(xyz.shape), xyz, rtol=2e-5, atol=1e-6 * xyz.max())


def test_xyz_only_render_matches_film_render():
    """The headline shape at 8 spp with the tuned launch (one chunk per pixel
    group): the XYZ of an XYZ-only render is bit-identical to the XYZ of a
    render that also returns the film, and equals the weighted film."""
    from pysky import autotune

    scene = pysky.SceneData(sun_elevation_deg=5.0, wavelengths_nm=list(pysky.midpoint_grid(64)))
    cfg = pysky.RenderConfig(width=1024, height=512, spp=8, seed=1, max_depth=32, rr_start_depth=4,
                             carried_tau=True, phase_importance_sampling=True)
    ctx = pysky.Context(0)
    cfg = autotune.apply(ctx, cfg, autotune.tuned(scene, cfg))
    a = _render(scene, cfg, ctx, want_film=False, want_xyz=True)
    b = _render(scene, cfg, ctx, want_film=True, want_xyz=True)
    assert a.stats.items == 1024 * 512 // 4  # one chunk
    np.testing.assert_array_equal(a.xyz, b.xyz)
    lam = pysky.wavelength_table(scene)
    xyz = b.values().reshape(-1, 64).astype(np.float64) @ lam[:, 5:8].astype(np.float64)
    np.testing.assert_allclose(b.xyz.reshape(-1, 3), xyz, rtol=2e-5, atol=1e-6 * xyz.max())


def test_xyz_only_render_allocates_no_partials():
    import torch

    from pysky import autotune
    from pysky.torch_ops import render_device

    scene = pysky.SceneData(sun_elevation_deg=5.0, wavelengths_nm=list(pysky.midpoint_grid(64)))
    cfg = pysky.RenderConfig(width=1024, height=512, spp=1, seed=1, max_depth=32, carried_tau=True,
                             phase_importance_sampling=True)
    xyz = torch.zeros((512, 1024, 3), dtype=torch.float32, device="cuda:0")
    # a first render loads the kernels and sizes the runtime's scratch space
    warm = pysky.Context(0)
    render_device(scene, autotune.apply(warm, cfg, autotune.tuned(scene, cfg)), xyz=xyz, context=warm)
    torch.cuda.synchronize()
    ctx = pysky.Context(0)
    cfg = autotune.apply(ctx, cfg, autotune.tuned(scene, cfg))
    free0, _ = torch.cuda.mem_get_info()
    st = render_device(scene, cfg, xyz=xyz, context=ctx)
    torch.cuda.synchronize()
    free1, _ = torch.cuda.mem_get_info()
    assert st.items == 1024 * 512 // 4
    assert float(xyz.sum()) > 0
    # tables, counters and the job record only (the partials would be 268 MB)
    assert free0 - free1 < 32 << 20, (free0 - free1) / 2**20
