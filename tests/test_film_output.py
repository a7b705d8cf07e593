"""On-disk outputs of the sky_render CLI (film_output.rs, asset.rs,
main.rs:85-104 rebuild path), checked on the CPU through --rebuild-rgb-only,
which reads band EXRs and writes the RGB EXR, the PNG and the manifest's
colorimetry without touching a GPU.

Pins: a pixel whose spectrum is the sun's comes out neutral (r = g = b to 5e-4) after
the solar-white-to-D65 adaptation; the adaptation maps the source white onto
D65; the linear-sRGB matrix maps D65 onto (1, 1, 1) (to 3e-4); files are read back by
independent readers (tests/exr_png.py) and the PNG's CRCs and Adler-32 are
verified by zlib."""
import json
import os
import struct
import subprocess

import numpy as np
import pytest

import oracle as O
import pysky
from exr_png import read_exr, read_png_rgb8, write_exr

EXE = os.path.join(os.path.dirname(pysky.LIB_PATH), "..", "bin", "sky_render")
D65 = np.array([0.9504559, 1.0, 1.0890578])
BRADFORD = np.array([[0.8951, 0.2664, -0.1614], [-0.7502, 1.7135, 0.0367], [0.0389, -0.0685, 1.0296]])
SRGB_FROM_XYZ = np.array([[3.2404542, -1.5371385, -0.4985314], [-0.9692660, 1.8760108, 0.0415560],
                          [0.0556434, -0.2040259, 1.0572252]])


def _bands(n=16):
    return list(pysky.midpoint_grid(n))


def _band_name(nm):
    return "sky_%03.0fnm.exr" % nm


def _run(*args):
    r = subprocess.run([EXE, *args], capture_output=True, text=True, timeout=120)
    return r


def _rebuild(out, *args):
    return _run("--rebuild-rgb-only", "--bands", "16", "--out", str(out), *args)


def _write_bands(out, film, **encoding):
    os.makedirs(out / "bands", exist_ok=True)
    for b, nm in enumerate(_bands(film.shape[2])):
        v = film[:, :, b]
        write_exr(out / "bands" / _band_name(nm), {"R": v, "G": v, "B": v}, **encoding)


def _film(h=12, w=20, nb=16, seed=5):
    rng = np.random.default_rng(seed)
    film = rng.gamma(2.0, 0.5, size=(h, w, nb)).astype(np.float32)
    lam = O.lambda_table(O.Params(lambda_nm=_bands(nb)))
    film[0, 0] = (lam[:, 4] * 0.01).astype(np.float32)  # the sun's spectrum
    film[0, 1] = 0.0
    return film, lam


def _expected(film, lam, exposure=0.05):
    band_xyz = lam[:, 5:8]
    sun = lam[:, 4] @ band_xyz
    src = sun / sun[1]
    lms_s, lms_d = BRADFORD @ src, BRADFORD @ D65
    M = np.linalg.inv(BRADFORD) @ np.diag(lms_d / lms_s) @ BRADFORD
    rgb = film.astype(np.float64) @ band_xyz @ M.T @ SRGB_FROM_XYZ.T
    t = 1.0 - np.exp(-np.maximum(rgb, 0.0) * exposure)
    x = np.clip(t, 0.0, 1.0)
    s = np.where(x <= 0.0031308, 12.92 * x, 1.055 * x ** (1 / 2.4) - 0.055)
    png = np.floor(np.clip(s, 0, 1) * 255.0 + 0.5).astype(np.int32)
    return rgb, png, M, src


def test_math_pins():
    """The reference constants themselves: sRGB(D65) = (1, 1, 1) to 3e-4 (the
    IEC matrix is built on the 5-digit D65 white 0.95047 / 1.08883, the white
    here is the CMF-integrated 0.9504559 / 1.0890578), and a Bradford
    adaptation maps the source white onto D65."""
    np.testing.assert_allclose(SRGB_FROM_XYZ @ D65, [1, 1, 1], atol=3e-4)
    np.testing.assert_allclose(SRGB_FROM_XYZ @ [0.95047, 1.0, 1.08883], [1, 1, 1], atol=2e-6)
    _, lam = _film()
    _, _, M, src = _expected(np.zeros((1, 1, 16), np.float32), lam)
    np.testing.assert_allclose(M @ src, D65, rtol=1e-12)


def test_rebuild_rgb_from_band_exrs(tmp_path):
    film, lam = _film()
    _write_bands(tmp_path, film)
    old = {"version": 0, "kind": "spectral_panorama_v0", "dimensions": [20, 12], "spp": 7, "seed": 3,
           "sun_elevation_deg": 0.0, "sun_azimuth_deg": 0.0, "observer_altitude_km": 0.2,
           "band_centers_nm": _bands(), "data_hash": None,
           "colorimetry": {"cmf": "x", "cmf_source": "y", "rgb_color_space": "z",
                           "white_balance": {"method": "identity", "source_white_xyz_y1": [1, 1, 1],
                                             "target_white_xyz_y1": [1, 1, 1],
                                             "xyz_from_xyz": [[1, 0, 0], [0, 1, 0], [0, 0, 1]]}},
           "files": {"rgb_exr": "sky_rgb.exr", "rgb_png": "sky_rgb.png", "band_exrs": []}}
    (tmp_path / "asset.json").write_text(json.dumps(old, indent=2))
    r = _rebuild(tmp_path)
    assert r.returncode == 0, r.stderr
    assert "rebuilt rgb outputs from 16 band EXRs" in r.stdout and "timing: load=" in r.stdout

    rgb_want, png_want, M, src = _expected(film, lam)
    exr = read_exr(tmp_path / "sky_rgb.exr")
    assert sorted(exr) == ["B", "G", "R"]
    rgb = np.stack([exr["R"], exr["G"], exr["B"]], axis=-1)
    np.testing.assert_allclose(rgb, rgb_want, rtol=2e-5, atol=1e-6 * np.abs(rgb_want).max())
    sun_px = rgb[0, 0]
    # the sun's own spectrum is neutral, up to the sRGB matrix's 3e-4 white offset
    assert np.ptp(sun_px) < 5e-4 * sun_px.mean()
    assert np.all(rgb[0, 1] == 0)

    png = read_png_rgb8(tmp_path / "sky_rgb.png").astype(np.int32)
    assert png.shape == (12, 20, 3)
    assert np.abs(png - png_want).max() <= 1

    m = json.loads((tmp_path / "asset.json").read_text())
    assert m["spp"] == 7 and m["dimensions"] == [20, 12] and m["files"]["rgb_exr"] == "sky_rgb.exr"
    wb = m["colorimetry"]["white_balance"]
    assert wb["method"] == "Bradford solar-white-to-D65"
    assert m["colorimetry"]["rgb_color_space"] == "linear sRGB D65"
    np.testing.assert_allclose(wb["xyz_from_xyz"], M, rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(wb["source_white_xyz_y1"], src, rtol=1e-6)
    np.testing.assert_allclose(wb["target_white_xyz_y1"], D65, rtol=1e-6)


def test_rebuild_reads_half_float_bands(tmp_path):
    film, lam = _film(h=4, w=6)
    film = film.astype(np.float16).astype(np.float32)
    _write_bands(tmp_path, film, half=True)
    r = _rebuild(tmp_path)
    assert r.returncode == 0, r.stderr
    assert not (tmp_path / "asset.json").exists()  # no manifest: nothing is created
    exr = read_exr(tmp_path / "sky_rgb.exr")
    rgb_want, _, _, _ = _expected(film, lam)
    np.testing.assert_allclose(exr["G"], rgb_want[..., 1], rtol=2e-5, atol=1e-6 * np.abs(rgb_want).max())


ENCODINGS = [
    dict(compression="rle", tile=(64, 64)),  # what exr 1.74's write_rgb_file emits (RLE, 64x64 tiles)
    dict(compression="rle"),
    dict(compression="zips"),
    dict(compression="zip"),
    dict(compression="zip", tile=(8, 5), shuffle=3),
    dict(compression="rle", tile=(7, 4), mipmap=True),
    dict(compression="none", tile=(6, 6)),
    dict(compression="zips", half=True),
]


@pytest.mark.parametrize("encoding", ENCODINGS, ids=lambda e: "-".join(f"{k}={v}" for k, v in e.items()))
def test_rebuild_reads_compressed_and_tiled_bands(tmp_path, encoding):
    """Band EXRs written as RLE/ZIPS/ZIP scanlines or tiles (ragged edge
    tiles, chunks out of order, a mipmap chain whose smaller levels must be
    skipped) rebuild the same bytes as the uncompressed scanline bands.
    37 lines leave a ragged 5-line ZIP block; zero and constant regions give
    the run-length coder runs."""
    film, _ = _film(h=37, w=23)
    film[5:9] = 0.0
    film[12, :, :] = 1.0
    half = encoding.get("half", False)
    if half:
        film = film.astype(np.float16).astype(np.float32)
    plain, coded = tmp_path / "plain", tmp_path / "coded"
    _write_bands(plain, film, half=half)
    _write_bands(coded, film, **encoding)
    size = lambda d: sum(f.stat().st_size for f in (d / "bands").iterdir())
    if encoding["compression"] != "none" and not encoding.get("mipmap"):
        assert size(coded) < size(plain)  # some chunks are stored compressed
    for d in (plain, coded):
        r = _rebuild(d)
        assert r.returncode == 0, r.stderr
    for name in ("sky_rgb.exr", "sky_rgb.png"):
        assert (coded / name).read_bytes() == (plain / name).read_bytes(), name


def _patched(path, old, new):
    b = path.read_bytes()
    assert b.count(old) == 1
    path.write_bytes(b.replace(old, new))


def test_rebuild_rejects_unread_encodings(tmp_path):
    film, _ = _film(h=4, w=6)
    _write_bands(tmp_path, film)
    band = tmp_path / "bands" / _band_name(_bands()[2])
    # PIZ (4) is named and refused, not misread
    _patched(band, b"compression\0compression\0\x01\0\0\0\0", b"compression\0compression\0\x01\0\0\0\x04")
    r = _rebuild(tmp_path)
    assert r.returncode == 1 and "compression 4 is not read" in r.stderr and band.name in r.stderr
    _write_bands(tmp_path, film)
    # multi-part flag (0x1000 in the version field)
    _patched(band, struct.pack("<ii", 20000630, 2), struct.pack("<ii", 20000630, 0x1002))
    r = _rebuild(tmp_path)
    assert r.returncode == 1 and "single-part" in r.stderr


def test_rebuild_png_exposure(tmp_path):
    film, lam = _film(h=3, w=5)
    _write_bands(tmp_path, film)
    r = _rebuild(tmp_path, "--png-exposure", "0.4")
    assert r.returncode == 0, r.stderr
    _, png_want, _, _ = _expected(film, lam, exposure=0.4)
    assert np.abs(read_png_rgb8(tmp_path / "sky_rgb.png").astype(np.int32) - png_want).max() <= 1


def test_rebuild_rejects_mismatched_band_sizes(tmp_path):
    film, _ = _film(h=4, w=6)
    _write_bands(tmp_path, film)
    v = np.zeros((5, 6), np.float32)
    write_exr(tmp_path / "bands" / _band_name(_bands()[3]), {"R": v, "G": v, "B": v})
    r = _rebuild(tmp_path)
    assert r.returncode == 1 and "other dimensions" in r.stderr


def test_rebuild_missing_band(tmp_path):
    r = _rebuild(tmp_path)
    assert r.returncode == 1 and "sky_392nm.exr" in r.stderr
    # without --bands the reference's 41-band grid (380..780 nm every 10 nm) is expected
    r = _run("--rebuild-rgb-only", "--out", str(tmp_path))
    assert r.returncode == 1 and "sky_380nm.exr" in r.stderr


def test_unknown_flag_prints_usage():
    r = _run("--no-such-flag")
    assert r.returncode == 2 and "--rebuild-rgb-only" in r.stdout
