"""Freezes the algorithmic-flop model used by bench.py's roofline.

F_unit(config, mode) = sum_e  n_e(config, mode) * c_e (per majorant shell count
L and shell growth g), where n_e is the mean
count per path x lambda sample of event e measured by the CPU oracle at the
configuration's seed on its subsampled pixel grid, and c_e the fixed FP32 cost
of the event's arithmetic (FMA = 2, add/mul/cmp/min/max/cvt = 1, each
transcendental or sqrt/div = 1).  pcg4d integer work is reported separately.

Run from the repository root:  python tests/golden/make_flop_model.py [--missing]
Writes sky-tracer_amd/pysky/flop_model.json (--missing: adds the configurations
the committed file lacks, leaving its entries as they are).
"""
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import numpy as np  # noqa: E402

import oracle as O  # noqa: E402

# FP32 operations per event of the integrator's arithmetic.
COSTS = {
    "paths": 26,             # jitter, equirect direction (2 sincos), normalise, XYZ weight
    "boundary_queries": 40,  # ray set-up: b, q, top/ground roots, initial shell
    "segments": 11,          # shell crossing: D_k - q, root, segment optical depth
    "free_flights": 6,       # -log(1 - u), distance
    "density_evals": 26,     # altitude from (b, q), two exponentials, tent, ratio
    "scatter_vertices": 110, # FP64 position, coefficients, albedo, phase sample, frame
    "nee_samples": 80,       # sun cone sample, frame, phase or cosine factor
    "ground_vertices": 85,   # position, normal, cosine sample, frame, roulette
}
# pcg4d per draw with the per-path words pre-multiplied: 8 mul, 8 add,
# 4 shift + 4 xor, 3 xor fold, shift, convert, scale.
INT_OPS_PER_DRAW = 31

WORKLOADS = {
    # name: (width, height, n_lambda, spp_subsample, max_depth, sun, albedo, aerosol_scale, grid_step)
    "c1": (128, 64, 16, 16, 8, 10.0, 0.3, 1.0, 1),
    "c2": (1024, 512, 64, 16, 32, 5.0, 0.3, 1.0, 16),
    "c3": (512, 256, 64, 8, 64, -4.0, 0.8, 4.0, 8),
}
MODES = {
    "reference": 0,
    "carried": O.FLAG_CARRIED_TAU,
    "carried_is": O.FLAG_CARRIED_TAU | O.FLAG_PHASE_IS,
    "carried_is_ratio": O.FLAG_CARRIED_TAU | O.FLAG_PHASE_IS | O.FLAG_RATIO_TRACKING,
}


def key(workload, mode, layers, growth=1.0):
    k = f"{workload}:{mode}:L{layers}"
    return k if growth == 1.0 else f"{k}:g{growth:g}"


def measure(workload, mode, layers, growth=1.0):
    w, h, nl, spp, depth, sun, alb, aer, step = WORKLOADS[workload]
    p = O.Params(width=w, height=h, spp=spp, seed=1, sun_elevation_deg=sun, max_depth=depth,
                 ground_albedo=alb, aerosol_sca_550_m_inv=3.996e-6 * aer,
                 aerosol_abs_550_m_inv=4.4e-6 * aer, lambda_nm=list(O.midpoint_grid(nl)),
                 flags=MODES[mode], majorant_layers=layers, shell_growth=growth)
    xs = np.arange(step // 2, w, step)
    ys = np.arange(step // 2, h, step)
    pixels = (ys[:, None] * w + xs[None, :]).ravel()
    _, _, cnt, rc = O.render_pixels(p, pixels, "f32")
    assert rc == 0
    n = cnt["paths"]
    ev = {k: cnt[k] / n for k in cnt if k != "paths"}
    flops = COSTS["paths"] + sum(COSTS[k] * ev.get(k, 0.0) for k in COSTS if k != "paths")
    return {
        "events_per_sample": ev,
        "flops_per_sample": flops,
        "rng_int_ops_per_sample": INT_OPS_PER_DRAW * ev["rng_draws"],
        "sample": f"{len(pixels)} px (every {step}th) x {nl} lambda x {spp} spp, seed 1",
    }


def main(missing_only=False):
    """Measures every configuration; with missing_only (--missing) keeps the
    committed entries as they are and measures only configurations they lack."""
    dst = os.path.join(ROOT, "sky-tracer_amd", "pysky", "flop_model.json")
    out = {"costs": COSTS, "int_ops_per_rng_draw": INT_OPS_PER_DRAW, "configs": {}}
    if missing_only:
        with open(dst) as f:
            out = json.load(f)
    for wl in ("c1", "c2", "c3"):
        for mode in ("reference", "carried", "carried_is", "carried_is_ratio"):
            if mode in ("carried_is", "carried_is_ratio"):
                combos = [(L, 1.0) for L in (4, 5, 6, 7, 8, 12, 16, 24, 32, 64)]
                combos += [(L, g) for L in (2, 3, 4, 5, 6, 7, 8) for g in (1.5, 2.0, 2.5, 3.0, 4.0, 5.0, 6.0)]
                if mode == "carried_is_ratio":
                    combos = [(L, g) for L, g in combos if L <= 8]
            elif mode == "reference":
                combos = [(16, 1.0), (64, 1.0)] + [(L, g) for L in (2, 3) for g in (4.0, 5.0, 6.0)]
            else:
                combos = [(16, 1.0), (64, 1.0)]
            for layers, g in combos:
                k = key(wl, mode, layers, g)
                if missing_only and k in out["configs"]:
                    continue
                out["configs"][k] = measure(wl, mode, layers, g)
                print(k, round(out["configs"][k]["flops_per_sample"], 1), flush=True)
    with open(dst, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print("wrote", dst)


if __name__ == "__main__":
    main(missing_only="--missing" in sys.argv[1:])
