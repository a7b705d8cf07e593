"""FP64 accuracy reference for the 1024x512 / 64-wavelength workload.

The double-precision oracle renders the 64x32 pixel grid x = 16 i + 8,
y = 16 j + 8 of the 1024x512 panorama (sun 5 deg, depth <= 32, roulette from
depth 4, 64 midpoint wavelengths 383.125 + 6.25 i nm) at 65536 samples per
pixel.  It uses the lowest-variance estimator of the same expectation the
oracle offers (carried optical depth, Rayleigh/HG lobe sampling, ratio-tracked
shadow rays, 16 majorant shells).  Written: c2_fp64_reference.npz with the
XYZ means of the grid pixels and their standard errors (from per-wavelength
second moments; wavelengths are independent streams).

Resumable: progress (double sums per sample chunk) is kept in a state file
(--state, default c2_reference_state.npz in the system temp directory).
Run from the repository root:
    python tests/golden/make_c2_reference.py [--threads 6] [--spp 65536] [--state PATH]
"""
import argparse
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import numpy as np  # noqa: E402

import oracle as O  # noqa: E402

W, H, NL = 1024, 512, 64
CHUNK = 2048


def params(spp, offset):
    return O.Params(width=W, height=H, spp=spp, sample_offset=offset, seed=1, sun_elevation_deg=5.0,
                    max_depth=32, rr_start_depth=4, lambda_nm=list(O.midpoint_grid(NL)), majorant_layers=16,
                    flags=O.FLAG_CARRIED_TAU | O.FLAG_PHASE_IS | O.FLAG_RATIO_TRACKING)


def grid_pixels():
    xs = np.arange(8, W, 16)
    ys = np.arange(8, H, 16)
    return (ys[:, None] * W + xs[None, :]).ravel()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=6)
    ap.add_argument("--spp", type=int, default=65536)
    ap.add_argument("--state", default=os.path.join(tempfile.gettempdir(), "c2_reference_state.npz"),
                    help="resumable progress file")
    args = ap.parse_args()
    state_path = args.state
    pix = grid_pixels()
    if os.path.exists(state_path):
        st = np.load(state_path)
        s1, s2, done = st["s1"], st["s2"], int(st["done"])
    else:
        s1 = np.zeros((len(pix), NL))
        s2 = np.zeros((len(pix), NL))
        done = 0
    while done < args.spp:
        n = min(CHUNK, args.spp - done)
        t0 = time.time()
        f, m2, _, rc = O.render_pixels(params(n, done), pix, "f64", threads=args.threads)
        assert rc == 0
        s1 += f * n
        s2 += m2 * n
        done += n
        np.savez(state_path, s1=s1, s2=s2, done=done)
        print(f"{done}/{args.spp} spp ({time.time() - t0:.0f} s per chunk)", flush=True)
    lam = O.lambda_table(params(1, 0))
    wx = lam[:, 5:8]  # [NL, 3]
    mean = s1 / done
    var = np.maximum(s2 / done - mean ** 2, 0.0) / done
    xyz = mean @ wx
    xyz_se = np.sqrt(var @ (wx ** 2))
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "c2_fp64_reference.npz")
    np.savez_compressed(out, pixels=pix, xyz=xyz, xyz_se=xyz_se, spp=done,
                        wavelengths_nm=np.asarray(O.midpoint_grid(NL)))
    rel = np.linalg.norm(xyz_se) / np.linalg.norm(xyz)
    print(f"wrote {out}: {len(pix)} px, {done} spp, relative standard error of the image {rel:.2e}")


if __name__ == "__main__":
    main()
