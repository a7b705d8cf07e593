"""The FP64 accuracy fixture of the 1024x512 / 64-wavelength workload
(tests/golden/c2_fp64_reference.npz, written by make_c2_reference.py with the
lowest-variance estimator: carried optical depth, lobe sampling, ratio-tracked
shadow rays, 16 shells) checked against the oracle's reference-form estimator
(fresh free flight per shell, isotropic bounces, delta-tracked shadow rays,
64 shells -- the reference's own form), rendered here independently on a
subset of the fixture's pixels.  Both estimate the same integral, so the
differences must be noise."""
import os

import numpy as np

import oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "c2_fp64_reference.npz")


def test_fixture_layout():
    ref = np.load(GOLDEN)
    xs = np.arange(8, 1024, 16)
    ys = np.arange(8, 512, 16)
    assert np.array_equal(ref["pixels"], (ys[:, None] * 1024 + xs[None, :]).ravel())
    assert int(ref["spp"]) == 65536
    assert np.allclose(ref["wavelengths_nm"], O.midpoint_grid(64))
    assert np.all(np.isfinite(ref["xyz"])) and np.all(ref["xyz"] > 0)
    assert np.all(ref["xyz_se"] > 0) and np.all(ref["xyz_se"] < 0.01 * ref["xyz"] + 1e-6)


def test_fixture_matches_reference_form_estimator():
    ref = np.load(GOLDEN)
    sel = np.arange(0, 2048, 64)  # 32 pixels spread over the image rows
    pix = ref["pixels"][sel]
    spp = 1024
    p = O.Params(width=1024, height=512, spp=spp, seed=77, sun_elevation_deg=5.0, max_depth=32,
                 rr_start_depth=4, lambda_nm=list(O.midpoint_grid(64)), majorant_layers=64, flags=0)
    film, m2, _, rc = O.render_pixels(p, pix, "f64", threads=8)
    assert rc == 0
    w = O.lambda_table(p)[:, 5:8]
    xyz = film @ w
    var = (np.maximum(m2 - film ** 2, 0.0) / spp) @ (w ** 2)
    z = (xyz - ref["xyz"][sel]) / np.sqrt(var + ref["xyz_se"][sel] ** 2)
    print(f"z mean {z.mean():+.3f} std {z.std():.3f} max|z| {np.abs(z).max():.2f}")
    assert abs(z.mean()) < 5.0 / np.sqrt(z.size)
    assert np.abs(z).max() < 5.0
    assert 0.6 < z.std() < 1.4
