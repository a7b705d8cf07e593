"""C-ABI checks that need no GPU: the library loads, exports every function
include/*.h declares, validates like integrator.rs:130-163, and its host
scene build (per-wavelength table, majorant grid) matches the oracle."""
import ctypes as C
import glob
import os
import re

import numpy as np
import pytest

import oracle as O
import pysky
from pysky import _lib as L


def _declared_functions():
    """Functions declared by every header under include/."""
    names = []
    inc = os.path.dirname(L.HEADER_PATH)
    for h in sorted(glob.glob(os.path.join(inc, "*.h"))):
        text = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        names += re.findall(r"^\s*(?:const\s+)?[a-z_]+\s*\**\s*(sky_[a-z0-9_]+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(L.LIB_PATH)
    names = _declared_functions()
    assert len(names) >= 21 and "sky_scan_combos" in names and "sky_render" in names
    for n in names:
        assert hasattr(lib, n), n


def test_struct_sizes_match_header_layout():
    # the Python mirrors must have the C layout (no padding surprises)
    assert C.sizeof(L.RenderConfigC) == 4 * 4 + 8 + 3 * 8 + 8 * 4 + 8 + 8 + 8
    assert C.sizeof(L.Atmosphere) == 15 * 8 + 8
    assert C.sizeof(L.MediumTablesC) == 8 + 6 * 8
    assert C.sizeof(L.Tuning) == 5 * 4
    assert C.sizeof(L.RenderStats) == 2 * 8 + 8 + 4 * 4 + 8


def test_defaults_match_reference():
    cfg = L.RenderConfigC()
    L.lib().sky_default_config(C.byref(cfg))
    # config.rs:38-52
    assert (cfg.width, cfg.height, cfg.spp) == (1024, 512, 256)
    assert cfg.seed == 0x5EC7_2026_0430
    assert cfg.observer_altitude_km == pytest.approx(0.2)
    assert cfg.direct_light_samples == 1 and cfg.projection == 0
    # integrator.rs:14, sky_trace.wgsl:412/692, data.rs:48, medium.rs:77
    assert cfg.watchdog_limit == 1024 and cfg.max_segments == 100000
    assert cfg.rr_start_depth == 4 and cfg.majorant_layers == 64
    assert cfg.majorant_margin == pytest.approx(1.05)
    py = pysky.RenderConfig()
    assert (py.width, py.height, py.spp, py.seed) == (1024, 512, 256, 0x5EC7_2026_0430)


@pytest.mark.parametrize("n", [3, 16, 41, 64])
def test_wavelength_table_matches_oracle(n):
    lam = list(np.sort(np.random.default_rng(n).uniform(390, 770, n))) if n in (3, 41) else list(pysky.midpoint_grid(n))
    scene = pysky.SceneData(wavelengths_nm=lam)
    got = pysky.wavelength_table(scene)
    ref = O.lambda_table(O.Params(lambda_nm=lam))
    np.testing.assert_allclose(got, ref, rtol=2e-7)


@pytest.mark.parametrize("layers,growth", [(8, 1.0), (16, 1.0), (64, 1.0), (200, 1.0), (3, 3.0), (6, 1.7)])
def test_majorant_grid_matches_oracle(layers, growth):
    scene = pysky.SceneData(wavelengths_nm=list(pysky.midpoint_grid(16)))
    got = pysky.majorant_grid(scene, layers, shell_growth=growth)
    ref = O.majorant_grid(O.Params(majorant_layers=layers, shell_growth=growth))
    np.testing.assert_allclose(got, ref, rtol=2e-7)


@pytest.mark.parametrize("bad", [
    dict(width=0), dict(height=0), dict(spp=0), dict(majorant_layers=0), dict(majorant_layers=300),
    dict(shell_growth=0.5), dict(shell_growth=20.0), dict(majorant_layers=256, shell_growth=16.0),
    dict(majorant_margin=0.9), dict(output_projection=2), dict(width=70000, height=70000),
])
def test_invalid_config_is_reported_before_device(bad):
    scene = pysky.load_scene(10.0)
    cfg = pysky.RenderConfig(width=4, height=2, spp=1)
    for k, v in bad.items():
        setattr(cfg, k, v)
    with pytest.raises(pysky.InvalidConfig):
        pysky.render(scene, cfg, context=_null_context())


@pytest.mark.parametrize("bad", [dict(wavelengths_nm=[]), dict(wavelengths_nm=[500.0] * 65),
                                 dict(wavelengths_nm=[-1.0])])
def test_invalid_wavelengths(bad):
    scene = pysky.load_scene(10.0, **{})
    scene.wavelengths_nm = bad["wavelengths_nm"]
    with pytest.raises(pysky.InvalidConfig):
        pysky.render(scene, pysky.RenderConfig(width=4, height=2, spp=1), context=_null_context())


@pytest.mark.parametrize("field,value", [("aerosol_g", 1.0), ("ground_albedo", 1.5), ("rayleigh_h_km", 0.0),
                                         ("sun_angular_radius_rad", 0.0)])
def test_invalid_atmosphere(field, value):
    scene = pysky.load_scene(10.0)
    setattr(scene.atmosphere, field, value)
    with pytest.raises(pysky.InvalidConfig):
        pysky.render(scene, pysky.RenderConfig(width=4, height=2, spp=1), context=_null_context())


@pytest.mark.parametrize("growth", [1.0, 4.0])
def test_majorant_table_must_bound_the_extinction(growth):
    """A caller's fixed majorant below some shell's maximum extinction would
    bias delta tracking (the accept ratio is clamped at 1), so it is refused
    before the device; the margin-1 grid itself is the smallest valid one."""
    scene = pysky.load_scene(10.0)
    cfg = pysky.RenderConfig(width=4, height=2, spp=1, majorant_layers=3, shell_growth=growth)
    exact = pysky.majorant_grid(scene, 3, 1.0, growth)
    for table in (exact, pysky.majorant_grid(scene, 3, 1.25, growth)):
        cfg.majorant_table = table
        with pytest.raises(pysky.NoAdapter):
            pysky.render(scene, cfg, context=_null_context())
    for band, shell in ((0, 0), (scene.band_count - 1, 2)):
        low = np.array(exact, dtype=np.float32).reshape(scene.band_count, 3).copy()
        low[band, shell] *= 0.999
        cfg.majorant_table = low
        with pytest.raises(pysky.InvalidConfig, match="below the shell's maximum extinction"):
            pysky.render(scene, cfg, context=_null_context())


class _NullCtx:
    handle = None


def _null_context():
    return _NullCtx()


def test_valid_config_without_device_is_no_adapter():
    """A valid render with no device context maps to RenderError::NoAdapter
    (integrator.rs:179), which smoke.rs:18 treats as a skip."""
    with pytest.raises(pysky.NoAdapter):
        pysky.render(pysky.load_scene(5.0), pysky.RenderConfig(width=4, height=2, spp=1), context=_null_context())


def test_context_create_status():
    h = C.c_void_p()
    st = L.lib().sky_context_create(0, C.byref(h))
    assert st in (L.SKY_OK, L.SKY_ERR_NO_DEVICE)
    if st == L.SKY_OK:
        L.lib().sky_context_destroy(h)
    st = L.lib().sky_context_create(-1, C.byref(h))
    assert st == L.SKY_ERR_NO_DEVICE


def test_film_api_mirrors_reference():
    """film.rs:2-62."""
    v = np.arange(2 * 3 * 4, dtype=np.float32)
    f = pysky.Film(3, 2, 4, v.copy())
    assert (f.width(), f.height(), f.band_count()) == (3, 2, 4)
    np.testing.assert_array_equal(f.pixel_spectrum(1), [4, 5, 6, 7])
    f.set_pixel_spectrum(5, [9, 9, 9, 9])
    np.testing.assert_array_equal(f.values()[20:24], [9, 9, 9, 9])
    assert f.as_array().shape == (2, 3, 4)
    assert pysky.OutputProjection.Panorama.label() == "panorama"
    assert pysky.OutputProjection.SkyViewLut.label() == "sky-view LUT"


def test_cli_reports_invalid_config(tmp_path):
    """sky_render (the sky-reference-render counterpart) validates before it
    needs a device."""
    import os
    import subprocess

    exe = os.path.join(os.path.dirname(L.LIB_PATH), "..", "bin", "sky_render")
    r = subprocess.run([exe, "--width", "0", "--height", "4", "--spp", "1", "--out", str(tmp_path)],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 1
    assert "greater than zero" in r.stderr


def test_independent_seeds_give_disjoint_sample_keys():
    """A path's sample key is sample ^ seed_hi (sky_trace.wgsl:799-800): the
    seeds of pysky.independent_seed give every render sample keys of its own
    for up to 65536 samples per pixel, and leave the pixel keys alone."""
    spp = 65536
    keys = []
    for k in range(1, 5):
        seed = pysky.independent_seed(k)
        assert seed & 0xFFFFFFFF == 0
        hi = seed >> 32
        keys.append({s ^ hi for s in range(0, spp, 997)} | {(spp - 1) ^ hi})
    for i in range(len(keys)):
        for j in range(i + 1, len(keys)):
            assert not keys[i] & keys[j]
    # seeds 1 and 2, by contrast, hand the same set of pixel keys to a 64-pixel image
    assert {p ^ 1 for p in range(64)} == {p ^ 2 for p in range(64)}


def test_small_quotient_from_a_rounded_reciprocal():
    """The integrator splits a triple index into (sample, pixel, band) and a
    pixel into (x, y) with floor((n + 1/2) * rd), rd the correctly rounded
    float 1 / d, in float32 arithmetic: (n + 1/2) / d lies at least 1/(2d)
    from an integer and the product's error is below (n + 1/2) / d * 2^-22,
    so the floor is exact for n < 2^21.  Checked exhaustively where the
    kernel uses it (n < nslots + 64 <= 320, x < width + 256) and on random
    n up to 2^21 for divisors up to 2^16."""
    d = np.arange(1, 513, dtype=np.float32)
    rd = (np.float32(1.0) / d).astype(np.float32)
    n = np.arange(0, 1024, dtype=np.float32)
    q = np.floor((n[None, :] + np.float32(0.5)) * rd[:, None]).astype(np.int64)
    np.testing.assert_array_equal(q, np.arange(1024)[None, :] // np.arange(1, 513)[:, None])
    rng = np.random.default_rng(3)
    nn = rng.integers(0, 1 << 21, size=400000)
    dd = rng.integers(1, 1 << 16, size=nn.size)
    rdd = (np.float32(1.0) / dd.astype(np.float32)).astype(np.float32)
    qq = np.floor((nn.astype(np.float32) + np.float32(0.5)) * rdd).astype(np.int64)
    np.testing.assert_array_equal(qq, nn // dd)
