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
}
PEAK_FP32_TFLOPS = 157.3  # MI355X FP32 (packed vector FMA); unpacked vector issue is half
# HBM bytes per integrator launch from rocprofv3 FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0 (scales from
# a 256 MiB copy): profiles/k_wavepool_c2_1024x512x64_counters.md.  The launch writes one
# [pixel][band] float2 partial per pixel group and reads only the scene tables, so the figure
# does not depend on spp at the tuned work split.
TRAFFIC_BYTES = {"c2": (2.0 * 105.9 + 270528.0) * 1024}


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


def cpu_baseline(wname, w, mode, layers, target_s=15.0):
    """The oracle (C restatement of the same estimator, FP32) on the host
    cores: every sample of the workload for an evenly spread subset of its
    pixels, the subset sized from a short probe so the run takes ~target_s."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np

    import oracle as O

    flags = {"reference": 0, "carried": O.FLAG_CARRIED_TAU,
             "carried_is": O.FLAG_CARRIED_TAU | O.FLAG_PHASE_IS}[mode]
    threads = max(1, min(os.cpu_count() or 1, 16))
    npx = w["width"] * w["height"]

    def run(pixels, spp):
        p = O.Params(width=w["width"], height=w["height"], spp=spp, seed=1, sun_elevation_deg=w["sun"],
                     max_depth=w["max_depth"], ground_albedo=w["albedo"],
                     aerosol_sca_550_m_inv=3.996e-6 * w["aerosol"], aerosol_abs_550_m_inv=4.4e-6 * w["aerosol"],
                     lambda_nm=list(O.midpoint_grid(w["n_lambda"])), flags=flags, majorant_layers=layers)
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
                  f"{w['n_lambda']} lambda x {w['spp']} spp, {dt:.1f} s on {threads} threads",
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c2", choices=sorted(WORKLOADS))
    ap.add_argument("--mode", default="carried_is", choices=sorted(MODES))
    ap.add_argument("--layers", type=int, default=0, help="majorant shells (0: tuned)")
    ap.add_argument("--autotune", action="store_true", help="re-tune launch parameters first (untimed)")
    ap.add_argument("--spp", type=int, default=0, help="override samples per pixel")
    ap.add_argument("--no-cpu", action="store_true")
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

    w = dict(WORKLOADS[args.workload])
    if args.spp:
        w["spp"] = args.spp
    scene = make_scene(w)
    cfg = make_config(w, args.mode, 64)
    ctx = pysky.Context(0)
    if args.autotune:
        autotune.autotune(scene, cfg, ctx, log=lambda m: print(m, file=sys.stderr, flush=True))
    params = autotune.tuned(scene, cfg)
    if args.layers:
        params["majorant_layers"] = args.layers
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
    if os.path.exists(fm_path):
        fm = json.load(open(fm_path))
        if key in fm["configs"]:
            f_unit = fm["configs"][key]["flops_per_sample"]
            achieved = f_unit * samples / avg_kernel_s / 1e12
            roofline = {"bound": "valu", "achieved": round(achieved, 3), "peak": PEAK_FP32_TFLOPS,
                        "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32_TFLOPS, 5),
                        "traffic": TRAFFIC_BYTES.get(args.workload),
                        "flops_per_sample": round(f_unit, 1),
                        "frac_nonpacked": round(achieved / (PEAK_FP32_TFLOPS / 2), 5),
                        "rng_int_tops": round(fm["configs"][key]["rng_int_ops_per_sample"] * samples
                                              / avg_kernel_s / 1e12, 3),
                        "kernel": "k_wavepool", "kernel_ms_avg": round(avg_kernel_s * 1e3, 2)}
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(args.workload, w, args.mode, args.layers)
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
        "dtype": "f32 (f64 ray set-up)",
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
    print(json.dumps(out))


if __name__ == "__main__":
    main()
