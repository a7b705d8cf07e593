#!/usr/bin/env python3
"""Throughput benchmark of the MI355X spectral sky path tracer.

One step = one full render of the workload (every pixel x wavelength x
sample traced to termination) through the C ABI, outputs in device memory.
Prints one JSON line: path x lambda samples/s on one GPU, the roofline of the
integrator kernel (algorithmic FP32 flops per sample x samples / kernel time)
and the CPU oracle's throughput on the host cores for the same estimator.

    python bench.py [--steps K] [--warmup W] [--workload c2] [--mode carried_is]
"""
from __future__ import annotations

import argparse
import dataclasses
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "sky-tracer_amd"))

WORKLOADS = {
    # name: width, height, n_lambda, spp, max_depth, sun elevation, albedo, aerosol scale
    "c1": dict(width=128, height=64, n_lambda=16, spp=64, max_depth=8, sun=10.0, albedo=0.3, aerosol=1.0),
    "c2": dict(width=1024, height=512, n_lambda=64, spp=1024, max_depth=32, sun=5.0, albedo=0.3, aerosol=1.0),
    "c3": dict(width=512, height=256, n_lambda=64, spp=2048, max_depth=64, sun=-4.0, albedo=0.8, aerosol=4.0),
}
MODES = {
    "reference": dict(),
    "carried": dict(carried_tau=True),
    "carried_is": dict(carried_tau=True, phase_importance_sampling=True),
    "carried_is_ratio": dict(carried_tau=True, phase_importance_sampling=True, ratio_tracking=True),
}
# how oracle/Makefile builds the CPU baseline (gcc -O3 -march=x86-64-v3 measured no faster on
# this path: 10.2 against 9.9 M samples/s on 8 threads)
ORACLE_BUILD = "gcc -O2 -std=c11 -fopenmp (scalar, no -march), OpenMP dynamic schedule over pixels"
PEAK_FP32_TFLOPS = 157.3  # MI355X published FP32 vector peak
# HBM bytes per integrator launch from rocprofv3 FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0 (scales from
# a 256 MiB copy): profiles/k_wavepool_c2_1024x512x64_counters.md.  Each work item writes its
# pixels' XYZ (the 6.3 MB image, as partial cache lines shared with items on other XCDs) and
# reads only the scene tables, so the figure does not depend on spp at the tuned work split.
TRAFFIC_BYTES = {"c2": (2.0 * 101.44 + 16576.1) * 1024}
# Same for the combo scan of all 6-band combos x 65536 samples:
# profiles/k_scan_6band_combos_x_65536_samples_counters.md.
SCAN_TRAFFIC_BYTES = {6: (2.0 * 5189454.0 + 70256.1) * 1024}


def pool_kernel(mode, params, n_lambda):
    """Name of the integrator kernel a render launches: the library picks the
    _fast (_fastrt) variants for carried optical depth + lobe sampling with
    delta- (ratio-) tracked shadow rays and the _fastref variants for the
    reference's estimator, all with at most 8 majorant shells whose grid fits
    its LDS copy
    (n_lambda x shells <= 2048), and the register-budget variant from the
    tuning's occupancy."""
    layers = params["majorant_layers"]
    fits = layers <= 8 and n_lambda * layers <= 2048
    kinds = {"carried_is": "_fast", "carried_is_ratio": "_fastrt", "reference": "_fastref"}
    kind = kinds.get(mode, "") if fits else ""
    occ = params.get("occupancy", 0)
    return "k_wavepool" + kind + (f"_o{occ}" if occ else "")


def make_scene(w):
    import pysky

    atm = pysky.AtmosphereParams(ground_albedo=w["albedo"])
    atm.aerosol_sca_550_m_inv *= w["aerosol"]
    atm.aerosol_abs_550_m_inv *= w["aerosol"]
    return pysky.SceneData(atmosphere=atm, sun_elevation_deg=w["sun"],
                           wavelengths_nm=list(pysky.midpoint_grid(w["n_lambda"])))


def make_config(w, mode, layers, seed=1):
    import pysky

    return pysky.RenderConfig(width=w["width"], height=w["height"], spp=w["spp"], seed=seed,
                              max_depth=w["max_depth"], rr_start_depth=4, majorant_layers=layers,
                              **MODES[mode])


def cpu_baseline(wname, w, mode, layers, shell_growth, target_s=15.0):
    """The oracle (C restatement of the same estimator, FP32) on the host
    cores: every sample of the workload for an evenly spread subset of its
    pixels, the subset sized from a short probe so the run takes ~target_s.
    It traces the GPU run's paths: same estimator flags, majorant shells
    (count and thickness growth), depth cap and roulette."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np

    import oracle as O

    flags = {"reference": 0, "carried": O.FLAG_CARRIED_TAU,
             "carried_is": O.FLAG_CARRIED_TAU | O.FLAG_PHASE_IS,
             "carried_is_ratio": O.FLAG_CARRIED_TAU | O.FLAG_PHASE_IS | O.FLAG_RATIO_TRACKING}[mode]
    threads = max(1, min(os.cpu_count() or 1, 16))
    npx = w["width"] * w["height"]

    def run(pixels, spp):
        p = O.Params(width=w["width"], height=w["height"], spp=spp, seed=1, sun_elevation_deg=w["sun"],
                     max_depth=w["max_depth"], ground_albedo=w["albedo"],
                     aerosol_sca_550_m_inv=3.996e-6 * w["aerosol"], aerosol_abs_550_m_inv=4.4e-6 * w["aerosol"],
                     lambda_nm=list(O.midpoint_grid(w["n_lambda"])), flags=flags, majorant_layers=layers,
                     shell_growth=shell_growth, rr_start_depth=4)
        t0 = time.perf_counter()
        O.render_pixels(p, pixels, "f32", threads=threads)
        return time.perf_counter() - t0, len(pixels) * w["n_lambda"] * spp

    probe = np.linspace(0, npx - 1, min(npx, 1024)).astype(np.int64)
    dt, n = run(probe, min(w["spp"], 16))
    rate = n / max(dt, 1e-6)
    want = int(target_s * rate / (w["n_lambda"] * w["spp"]))
    pixels = np.linspace(0, npx - 1, max(64, min(npx, want))).astype(np.int64)
    dt, n = run(pixels, w["spp"])
    return {
        "value": n / dt,
        "unit": "path x lambda samples/s",
        "cores": threads,
        "kind": "port",
        "sample": f"oracle FP32, {len(pixels)} of {w['width']}x{w['height']} px (evenly spread) x "
                  f"{w['n_lambda']} lambda x {w['spp']} spp, {layers} shells growing {shell_growth:g}x, "
                  f"{dt:.1f} s on {threads} threads",
        "build": ORACLE_BUILD,
    }


def copy_calibration(mib):
    """Three plain device-to-device copies of `mib` MiB: under rocprofv3 --pmc
    their FETCH_SIZE / WRITE_SIZE per dispatch, against the known mib x 1024
    KiB, give the scale factors applied to the integrator's counters."""
    import torch

    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_kib_per_dispatch": mib * 1024, "dispatches": 3}))


SKYVIEW_ELEVATIONS = (-2.0, -1.0, 0.0, 1.0, 2.0, 5.0, 10.0, 20.0, 45.0, 90.0)


def skyview_sweep(spp, mode, data_dir=None):
    """The reference's sky-view LUT bake (scripts/bake_skyview_elevations.ps1:2-35):
    10 sun elevations, 256x256 sky-view LUT projection, 41 bands 380..780 nm
    (10 nm wide), observer 0.2 km, one sun sample, uncapped depth.  With
    `data_dir`, the scene is the reference's own tables (tabulated medium)."""
    import pysky
    from pysky import tables

    lam = [380.0 + 10.0 * i for i in range(41)]
    out = []
    for el in SKYVIEW_ELEVATIONS:
        if data_dir:
            scene = tables.load_reference_scene(data_dir, el)
        else:
            scene = pysky.SceneData(sun_elevation_deg=el, wavelengths_nm=lam, weights_nm=[10.0] * 41)
        cfg = pysky.RenderConfig(width=256, height=256, spp=spp, seed=0x5EC720260430, max_depth=0,
                                 output_projection=pysky.OutputProjection.SkyViewLut, **MODES[mode])
        out.append((scene, cfg))
    return out


def bench_batch(args, ctx):
    """Ragged batch (256 renders: sun U[-6, 90] deg, 64x32..512x256, 3..64
    wavelengths in 400-700 nm, 256 spp) -- or the reference's sky-view LUT
    elevation sweep -- in one persistent launch replayed from a HIP graph,
    against the same renders issued one by one."""
    import numpy as np
    import torch

    from pysky.batch import ragged_batch, render_batch
    from pysky.torch_ops import render_device

    if args.workload == "skyview":
        data_dir = None if args.medium == "analytic" else args.data_dir
        batch = skyview_sweep(args.spp or 4096, args.mode, data_dir)
        what = (f"skyview: the reference's LUT bake, 10 sun elevations -2..90 deg, 256x256 sky-view LUT, "
                f"41 bands 380-780 nm, {args.spp or 4096} spp, uncapped depth, one launch, HIP graph, "
                + ("analytic atmosphere" if data_dir is None else
                   "the reference's tables (AFGL + OPAC, HG aerosol phase)"))
    else:
        batch = ragged_batch(n=256, seed=1, spp=args.spp or 256, max_depth=32)
        what = ("c4: 256 ragged renders (sun U[-6,90] deg, 64x32..512x256, 3..64 lambda in "
                "400-700 nm, 256 spp, depth<=32), one launch, HIP graph")
    layers, growth = args.layers, args.shell_growth
    batch = [(sc, dataclasses.replace(cfg, majorant_layers=layers, shell_growth=growth)) for sc, cfg in batch]
    samples = sum(cfg.width * cfg.height * cfg.spp * sc.band_count for sc, cfg in batch)
    print(f"[bench] batch of {len(batch)} renders, {samples:.3e} samples, layers={layers} growth={growth}",
          file=sys.stderr, flush=True)
    for i in range(args.warmup):
        r = render_batch(batch, context=ctx)
        print(f"[bench] warmup {i}: kernel {r.stats.kernel_ms:.1f} ms", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    kms, p50, p99 = [], [], []
    t0 = time.perf_counter()
    for i in range(args.steps):
        r = render_batch(batch, context=ctx)
        kms.append(r.stats.kernel_ms)
        p50.append(r.stats.latency_p50_ms)
        p99.append(r.stats.latency_p99_ms)
        print(f"[bench] step {i}: kernel {r.stats.kernel_ms:.1f} ms graph_reused={r.stats.graph_reused} "
              f"p50 {r.stats.latency_p50_ms:.1f} p99 {r.stats.latency_p99_ms:.1f} ms", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / args.steps
    # the same renders one at a time (device outputs, no host copies)
    dev = torch.device("cuda:0")
    outs = [torch.zeros(cfg.height * cfg.width * 3, dtype=torch.float32, device=dev) for _, cfg in batch]
    for (sc, cfg), o in zip(batch[:8], outs[:8]):
        render_device(sc, cfg, xyz=o, context=ctx)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    single_kernel_ms = 0.0
    for (sc, cfg), o in zip(batch, outs):
        single_kernel_ms += render_device(sc, cfg, xyz=o, context=ctx).kernel_ms
    torch.cuda.synchronize()
    single_wall = time.perf_counter() - t1
    value = samples / wall
    single_rate = samples / single_wall
    return {
        "metric": "batch path x lambda samples/s on 1 MI355X",
        "value": value,
        "unit": "path x lambda samples/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32 (f64 only at grazing ground intersections)",
        "data": ("the reference's scene tables (data/*.csv), HG aerosol phase"
                 if args.workload == "skyview" and args.medium == "tables" else "synthetic (analytic atmosphere)"),
        "config": {"workload": what, "estimator": args.mode if args.workload == "skyview" else "carried_is",
                   "majorant_layers": layers, "shell_growth": growth},
        "kernel_ms": float(np.mean(kms)),
        "latency_p50_ms": float(np.mean(p50)), "latency_p99_ms": float(np.mean(p99)),
        "single_render_rate": single_rate, "single_render_kernel_ms_sum": single_kernel_ms,
        "batch_over_single": value / single_rate,
    }


def bench_adjoint(args, ctx):
    """Adjoint at 256x128, 32 lambda, 256 spp (sun 10 deg, depth <= 8): one
    gradient evaluation = forward render + path replay of d mean(log XYZ) /
    d theta.  The gradient's relative RMSE at 256 spp is that of 8 renders
    with independent random streams against a 16x-spp reference gradient,
    taken over the gradient with respect to log theta (theta_k dL/dtheta_k,
    so that parameters in m^-1, km and plain numbers weigh alike); the spp
    that reaches 1% follows from RMSE ~ spp^-1/2."""
    import numpy as np
    import torch

    import pysky
    from pysky.adjoint import params_vector, render_adjoint_device
    from pysky.torch_ops import render_device

    spp = args.spp or 256
    scene = pysky.SceneData(sun_elevation_deg=10.0, wavelengths_nm=list(pysky.midpoint_grid(32)))
    base = pysky.RenderConfig(width=256, height=128, spp=spp, seed=7, max_depth=8, rr_start_depth=4,
                              majorant_layers=args.layers, shell_growth=args.shell_growth, **MODES[args.mode])
    dev = torch.device("cuda:0")
    xyz = torch.zeros(128 * 256 * 3, dtype=torch.float32, device=dev)
    # loss = mean(log XYZ) over the image, so every pixel weighs by its relative
    # change (an XYZ^2 loss would be the sun disk's transmittance alone);
    # its cotangent 1 / (n XYZ) comes from an independent 4x-spp render
    render_device(scene, dataclasses.replace(base, spp=4 * spp, seed=pysky.independent_seed(99)), xyz=xyz,
                  context=ctx)
    cot = (1.0 / (xyz.numel() * torch.clamp(xyz, min=1e-3 * float(xyz.mean())))).contiguous()
    grads, f_ms, a_ms = [], [], []
    for i in range(max(args.warmup, 1) + 8):
        cfg = dataclasses.replace(base, seed=pysky.independent_seed(100 + i))
        st = render_device(scene, cfg, xyz=xyz, context=ctx)
        g, ast = render_adjoint_device(scene, cfg, cot, context=ctx)
        if i >= max(args.warmup, 1):
            grads.append(g)
            f_ms.append(st.kernel_ms)
            a_ms.append(ast.kernel_ms)
    g_ref, _ = render_adjoint_device(scene, dataclasses.replace(base, spp=16 * spp, seed=pysky.independent_seed(98)),
                                     cot, context=ctx)
    torch.cuda.synchronize()
    G = np.array(grads)
    mean = G.mean(axis=0)
    theta = params_vector(scene.atmosphere)
    # RMSE against the reference (whose own noise, 1/16 of a render's variance, is included)
    rel_rmse = float(np.sqrt(np.mean(np.sum(((G - g_ref) * theta) ** 2, axis=1))) / np.linalg.norm(g_ref * theta))
    rel_rmse_raw = float(np.sqrt(np.mean(np.sum((G - g_ref) ** 2, axis=1))) / np.linalg.norm(g_ref))
    per_param = np.sqrt(G.var(axis=0, ddof=1)) / np.maximum(np.abs(mean), 1e-300)
    # spp that reaches 1% (RMSE ~ spp^-1/2), never below the configuration's own
    spp_1pct = max(spp, int(np.ceil(spp * (rel_rmse / 0.01) ** 2)))
    t_eval = (np.mean(f_ms) + np.mean(a_ms)) / 1e3
    evals_1pct = 1.0 / (t_eval * spp_1pct / spp)
    return {
        "metric": "adjoint grad evals/s at 1% RMSE on 1 MI355X",
        "value": evals_1pct,
        "unit": "gradient evaluations/s",
        "n_gpus": 1, "steps": 8, "warmup": max(args.warmup, 1), "ms_per_step": t_eval * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32 (f64 only at grazing ground intersections)",
        "data": "synthetic (analytic atmosphere)",
        "config": {"workload": f"c5: 256x128, 32 lambda, {spp} spp, depth<=8, sun 10 deg, 10-parameter adjoint",
                   "estimator": args.mode, "majorant_layers": base.majorant_layers,
                   "shell_growth": base.shell_growth},
        "forward_ms": float(np.mean(f_ms)), "adjoint_ms": float(np.mean(a_ms)),
        "adjoint_over_forward": float(np.mean(a_ms) / np.mean(f_ms)),
        "grad_rel_rmse_at_spp": rel_rmse, "grad_rel_rmse_at_spp_unscaled": rel_rmse_raw,
        "spp_for_1pct": spp_1pct,
        "grad_rel_std_per_param": [round(float(v), 5) for v in per_param],
        "grad_evals_per_s_at_spp": 1.0 / t_eval,
    }


def run_fit(args, ctx):
    """Recover +/-20%-perturbed parameters from a seed-7 target render by Adam
    (200 steps) on the C5 configuration."""
    import numpy as np
    import torch

    import pysky
    from pysky import fit
    from pysky.adjoint import params_vector
    from pysky.torch_ops import render_device

    scene = pysky.SceneData(sun_elevation_deg=10.0, wavelengths_nm=list(pysky.midpoint_grid(32)))
    cfg = pysky.RenderConfig(width=256, height=128, spp=args.spp or 256, seed=7, max_depth=8, rr_start_depth=4,
                             carried_tau=True, phase_importance_sampling=True, majorant_layers=args.layers,
                             shell_growth=args.shell_growth)
    target = torch.zeros((128, 256, 3), dtype=torch.float32, device="cuda:0")
    render_device(scene, cfg, xyz=target, context=ctx)
    truth = params_vector(scene.atmosphere)
    theta0 = fit.perturbed(scene)
    res = fit.fit(scene, cfg, target, theta0, steps=200, context=ctx,
                  log=lambda m: print(m, file=sys.stderr, flush=True))
    err0 = np.abs(theta0 / truth - 1.0)
    err = np.abs(res["theta"] / truth - 1.0)
    return {
        "metric": "inverse fit (Adam, 200 steps, adjoint gradients) on 1 MI355X",
        "value": res["seconds"], "unit": "s", "higher_is_better": False, "n_gpus": 1,
        "config": {"workload": "c5 fit: 256x128, 32 lambda, 256 spp, target seed 7, start +/-20%"},
        "loss_first": res["history"][0], "loss_last": float(np.mean(res["history"][-10:])),
        "param_names": list(pysky.PARAM_NAMES),
        "rel_error_start": [round(float(v), 4) for v in err0],
        "rel_error_end": [round(float(v), 4) for v in err],
    }


# FP32 flops of one (combo, sample) evaluation in k_scan<6>'s inner loop, counted from its gfx950
# ISA (the loop is unrolled 4x): 23 FMAs (x2), 13 other FP32 adds/muls/subs and 7 transcendentals
# (3 log2, 3 exp2, 1 sqrt; x1 each) per evaluation -- 57.4 VALU wave-instructions per wave64 of
# evaluations by SQ_INSTS_VALU (profiles/k_scan_6band_combos_x_65536_samples_counters.md).
SCAN_KERNEL_FLOPS_PER_EVAL = {6: 66}


def scan_reference_ops_per_eval(primaries):
    """The reference's FP64 operations per evaluation of evaluate_combo
    (main.rs:443-458): per channel v * w / width and three multiply-adds
    (8 per band); XYZ / scale (3); XYZ -> LMS (15); three cube roots (3);
    LMS' -> Lab (15); the distance (9); the two running sums (3)."""
    return 8 * primaries + 3 + 15 + 3 + 15 + 9 + 3


def scan_dataset(spp):
    """The sky-view LUT bake of 10 elevations x 256^2 x 41 bands (655,360
    samples) rendered at `spp` samples, with the reference colour of every
    sample (main.rs:276-341)."""
    import numpy as np

    import pysky
    from pysky import scan

    ctx = pysky.Context(0)
    vals = []
    for scene, cfg in skyview_sweep(spp, "carried_is"):
        vals.append(pysky.render(scene, cfg, want_xyz=False, context=ctx).values().reshape(-1, 41))
    v = np.concatenate(vals).astype(np.float32)
    wl = np.array([380.0 + 10.0 * i for i in range(41)], np.float32)
    cie = scan.cie_1931_2deg(wl)
    xyz = v.astype(np.float64) @ cie
    y = np.sort(xyz[:, 1][np.isfinite(xyz[:, 1])])
    scale = max(float(y[int(np.floor((y.size - 1) * 0.99 + 0.5))]), 1e-12)
    return v, wl, cie, scale, scan.xyz_to_oklab(xyz / scale)


def bench_scan(args):
    """The wavelength optimizer's screening pass (main.rs:156-199): every
    combo of `primaries` bands from the 41-band grid scored on 65,536 evenly
    spread samples of the 10-elevation sky-view LUT bake, in one launch with
    every operand resident in HBM."""
    import numpy as np
    import torch

    from pysky import scan

    P = args.primaries
    v, wl, cie, scale, ref = scan_dataset(args.spp or 64)
    ds = scan.ScanDataset(v, ref)
    import itertools
    combos = np.fromiter(itertools.chain.from_iterable(itertools.combinations(range(41), P)), np.uint8).reshape(-1, P)
    t0 = time.perf_counter()
    w = scan.quadrature_weights(wl, combos, scan.VORONOI)
    coef = scan.combo_coefficients(combos, w, scan.band_widths(wl), cie, scale)
    prep_s = time.perf_counter() - t0
    idx = scan.screening_indices(v.shape[0], 65536)
    dev = torch.device("cuda:0")
    tc = torch.from_numpy(combos).to(dev)
    tf = torch.from_numpy(coef).to(dev)
    ti = torch.from_numpy(idx.astype(np.int32)).to(dev)
    ts = torch.zeros(len(combos), dtype=torch.float64, device=dev)
    tq = torch.zeros_like(ts)
    evals = len(combos) * idx.size
    print(f"[bench] scan: {len(combos)} {P}-band combos x {idx.size} samples = {evals:.3e} evaluations "
          f"(weights+coefficients on the host {prep_s:.1f} s)", file=sys.stderr, flush=True)
    for i in range(args.warmup):
        st = ds.scan_device(tc, tf, ts, tq, samples=ti)
        print(f"[bench] warmup {i}: kernel {st.kernel_ms:.1f} ms", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    kernel_ms = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        st = ds.scan_device(tc, tf, ts, tq, samples=ti)
        kernel_ms.append(st.kernel_ms)
        print(f"[bench] step {i}: kernel {st.kernel_ms:.1f} ms", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    value = evals * args.steps / elapsed
    avg_s = sum(kernel_ms) / len(kernel_ms) / 1e3
    f_eval = SCAN_KERNEL_FLOPS_PER_EVAL.get(P)
    achieved = f_eval * evals / avg_s / 1e12 if f_eval else None
    rmse = np.sqrt(tq.cpu().numpy() / idx.size)
    best = combos[int(np.argmin(rmse))]
    cpu = None
    if not args.no_cpu:
        cpu = scan_cpu_baseline(v, wl, cie, scale, ref, combos, w, idx)
    out = {
        "metric": "combo x sample evaluations/s (wavelength-optimizer screening) on 1 MI355X",
        "value": value, "unit": "evaluations/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 (f64 sums)", "data": "synthetic (a sky-view LUT bake rendered by libsky_hip)",
        "config": {"workload": f"scan: all {len(combos)} {P}-band combos of the 41-band grid, Voronoi weights, "
                               f"65536 of 655360 sky-view LUT samples ({args.spp or 64}-spp bake), one launch",
                   "parallelism": "1 GPU"},
        "roofline": None if achieved is None else {
            "bound": "valu", "achieved": round(achieved, 3), "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
            "frac": round(achieved / PEAK_FP32_TFLOPS, 5), "traffic": SCAN_TRAFFIC_BYTES.get(P),
            "flops_per_eval": f_eval, "reference_fp64_ops_per_eval": scan_reference_ops_per_eval(P),
            "kernel": f"k_scan<{P}>", "kernel_ms_avg": round(avg_s * 1e3, 2)},
        "precision": "FP32 per-sample colour (XYZ, cube roots, distance), FP64 sums over samples; "
                     "sums within 2e-4 of the FP64 oracle (tests/test_gpu_scan.py)",
        "cpu_baseline": cpu,
        "best_screened_combo_nm": [float(wl[b]) for b in best],
    }
    if cpu:
        out["speedup_vs_cpu"] = round(value / cpu["value"], 1)
    return out


def scan_cpu_baseline(v, wl, cie, scale, ref, combos, w, idx, target_s=15.0):
    """The FP64 oracle's evaluate_combo on the host cores over a bounded,
    evenly spread subset of the combos, on the same screening samples."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np

    import oracle as O

    threads = max(1, min(os.cpu_count() or 1, 16))

    def run(sel):
        t0 = time.perf_counter()
        O.evaluate_combos(v, ref, scale, wl, cie, combos[sel], w[sel], idx=idx, threads=threads)
        return time.perf_counter() - t0

    probe = np.linspace(0, len(combos) - 1, 64).astype(np.int64)
    rate = 64 / max(run(probe), 1e-6)
    sel = np.linspace(0, len(combos) - 1, max(64, min(len(combos), int(target_s * rate)))).astype(np.int64)
    dt = run(sel)
    return {"value": len(sel) * idx.size / dt, "unit": "evaluations/s", "cores": threads, "kind": "port",
            "sample": f"oracle FP64 evaluate_combo, {len(sel)} of {len(combos)} combos (evenly spread) x "
                      f"{idx.size} samples, {dt:.1f} s on {threads} threads"}


def accuracy_vs_fp64(xyz, w, args):
    """Image relative RMSE (over the 64x32 grid x = 16i+8, y = 16j+8) of the
    rendered XYZ against the FP64 oracle's 65536-spp render of the same scene.
    The figure includes the FP64 render's own Monte Carlo noise (`ref_noise`)."""
    import numpy as np

    if args.workload != "c2":
        raise SystemExit("--accuracy compares the 1024x512 sun-5-degree render only")
    ref = np.load(os.path.join(ROOT, "tests", "golden", "c2_fp64_reference.npz"))
    got = xyz.reshape(-1, 3).cpu().numpy().astype(np.float64)[ref["pixels"]]
    rms = float(np.sqrt(np.mean(ref["xyz"] ** 2)))
    return {"rel_rmse": round(float(np.sqrt(np.mean((got - ref["xyz"]) ** 2))) / rms, 6),
            "ref_noise": round(float(np.sqrt(np.mean(ref["xyz_se"] ** 2))) / rms, 6),
            "ref_spp": int(ref["spp"]), "spp": w["spp"], "pixels": int(ref["pixels"].size)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c2", choices=sorted(WORKLOADS) + ["c4", "c5", "c5fit", "skyview", "scan"])
    ap.add_argument("--mode", default="carried_is", choices=sorted(MODES))
    ap.add_argument("--layers", type=int, default=0, help="majorant shells (0: tuned)")
    ap.add_argument("--shell-growth", type=float, default=0.0, help="shell thickness growth (0: tuned)")
    ap.add_argument("--autotune", action="store_true", help="re-tune launch parameters first (untimed)")
    ap.add_argument("--spp", type=int, default=0, help="override samples per pixel")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--primaries", type=int, default=6, help="scan: bands per combo (6: all 4,496,388 combos)")
    ap.add_argument("--medium", default="analytic", choices=["analytic", "tables"],
                    help="skyview: the analytic atmosphere or the reference's scene tables")
    ap.add_argument("--data-dir", default=os.path.join(ROOT, "tests", "golden", "reference_data"),
                    help="directory of the reference's scene tables (bands.csv, profiles, optics)")
    ap.add_argument("--accuracy", action="store_true",
                    help="c2 only: relative RMSE of the last render's XYZ against the 65536-spp FP64 render "
                         "of tests/golden/c2_fp64_reference.npz")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run 3 device copies of MIB MiB (scale for the FETCH/WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the integrator shards by sample range with no "
              "exchange but the final XYZ sum, and only the 1-GPU figure is reported.", file=sys.stderr)
        sys.exit(2)

    import torch

    import pysky
    from pysky.torch_ops import render_device

    from pysky import autotune

    if args.workload == "scan":
        print(json.dumps(bench_scan(args)))
        return
    if args.workload in ("c4", "c5", "c5fit", "skyview"):
        ctx = pysky.Context(0)
        if args.workload in ("c4", "skyview"):
            p = autotune.tuned(pysky.SceneData(wavelengths_nm=list(pysky.midpoint_grid(64))),
                               pysky.RenderConfig(width=1024, height=512, carried_tau=True,
                                                  phase_importance_sampling=True))
            autotune.apply(ctx, pysky.RenderConfig(), p)
            if args.workload == "skyview" and args.medium == "tables":
                # shells tuned for the reference's tables (aerosol concentrated near the ground)
                from pysky import tables
                pt = autotune.tuned(tables.load_reference_scene(args.data_dir, 5.0),
                                    pysky.RenderConfig(width=256, height=256, carried_tau=True,
                                                       phase_importance_sampling=True))
                p = dict(p, majorant_layers=pt["majorant_layers"], shell_growth=pt.get("shell_growth", 1.0))
            args.layers = args.layers or p["majorant_layers"]
            args.shell_growth = p.get("shell_growth", 1.0) if args.shell_growth <= 0 else args.shell_growth
            out = bench_batch(args, ctx)
        else:
            # the adjoint and the fit use the large-render tuning (shells, threshold,
            # steps, occupancy) with 40 waves per CU: with the replay kernels at 4
            # waves per SIMD, forward + replay at 256x128 x 32 bands take 57.7 ms
            # at 40 waves per CU, 58.3 at 32 and 60.0 at 24
            p = autotune.tuned(pysky.SceneData(wavelengths_nm=list(pysky.midpoint_grid(64))),
                               pysky.RenderConfig(width=1024, height=512, carried_tau=True,
                                                  phase_importance_sampling=True))
            p = dict(p, waves_per_cu=40)
            autotune.apply(ctx, pysky.RenderConfig(), p)
            args.layers = args.layers or p["majorant_layers"]
            args.shell_growth = p.get("shell_growth", 1.0) if args.shell_growth <= 0 else args.shell_growth
            out = bench_adjoint(args, ctx) if args.workload == "c5" else run_fit(args, ctx)
        print(json.dumps(out))
        return

    w = dict(WORKLOADS[args.workload])
    if args.spp:
        w["spp"] = args.spp
    scene = make_scene(w)
    # the FP64 reference render used seed 1: an accuracy run draws other random numbers
    cfg = make_config(w, args.mode, 64, seed=2 if args.accuracy else 1)
    ctx = pysky.Context(0)
    if args.autotune:
        autotune.autotune(scene, cfg, ctx, log=lambda m: print(m, file=sys.stderr, flush=True))
    params = autotune.tuned(scene, cfg)
    if args.layers:
        params["majorant_layers"] = args.layers
    if args.shell_growth > 0:
        params["shell_growth"] = args.shell_growth
    cfg = autotune.apply(ctx, cfg, params)
    args.layers = cfg.majorant_layers
    dev = torch.device("cuda:0")
    xyz = torch.zeros((w["height"], w["width"], 3), dtype=torch.float32, device=dev)
    samples = w["width"] * w["height"] * w["n_lambda"] * w["spp"]
    print(f"[bench] {args.workload} {w} mode={args.mode} layers={args.layers} samples/step={samples:.3e}",
          file=sys.stderr, flush=True)
    for i in range(args.warmup):
        st = render_device(scene, cfg, xyz=xyz, context=ctx)
        print(f"[bench] warmup {i}: kernel {st.kernel_ms:.1f} ms", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    kernel_ms = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        st = render_device(scene, cfg, xyz=xyz, context=ctx)
        kernel_ms.append(st.kernel_ms)
        print(f"[bench] step {i}: kernel {st.kernel_ms:.1f} ms items={st.items} grid={st.workgroups}",
              file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    value = samples * args.steps / elapsed
    avg_kernel_s = sum(kernel_ms) / len(kernel_ms) / 1e3

    roofline = None
    fm_path = os.path.join(ROOT, "sky-tracer_amd", "pysky", "flop_model.json")
    key = f"{args.workload}:{args.mode}:L{args.layers}"
    if cfg.shell_growth != 1.0:
        key += f":g{cfg.shell_growth:g}"
    if os.path.exists(fm_path):
        fm = json.load(open(fm_path))
        if key in fm["configs"]:
            f_unit = fm["configs"][key]["flops_per_sample"]
            achieved = f_unit * samples / avg_kernel_s / 1e12
            roofline = {"bound": "valu", "achieved": round(achieved, 3), "peak": PEAK_FP32_TFLOPS,
                        "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32_TFLOPS, 5),
                        "traffic": TRAFFIC_BYTES.get(args.workload),
                        "flops_per_sample": round(f_unit, 1),
                        "rng_int_tops": round(fm["configs"][key]["rng_int_ops_per_sample"] * samples
                                              / avg_kernel_s / 1e12, 3),
                        "kernel": pool_kernel(args.mode, params, w["n_lambda"]),
                        "kernel_ms_avg": round(avg_kernel_s * 1e3, 2)}
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(args.workload, w, args.mode, args.layers, cfg.shell_growth)
    out = {
        "metric": "spectral path samples/s (path x lambda) on 1 MI355X",
        "value": value,
        "unit": "path x lambda samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (f64 only at grazing ground intersections)",
        "data": "synthetic (analytic atmosphere, no external data)",
        "config": {"workload": f"{args.workload}: {w['width']}x{w['height']}, {w['n_lambda']} lambda, "
                               f"{w['spp']} spp, depth<={w['max_depth']}, sun {w['sun']} deg",
                   "estimator": args.mode, "majorant_layers": args.layers, "parallelism": "1 GPU",
                   "tuning": params},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    if cpu:
        out["speedup_vs_cpu"] = round(value / cpu["value"], 1)
    if args.accuracy:
        out["accuracy"] = accuracy_vs_fp64(xyz, w, args)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
