"""Benchmark: Mrays/s of the primary DE raymarch on 1080p Mandelbulb-8, one MI355X.

A step is one full frame (1920x1080, 1 ray per pixel: march + binary search + 6-tap normal +
Lambert shading) rendered by libmbx.so from constants already resident in HBM.  The JSON line
also carries the FP32 VALU roofline of the render kernel (algorithmic flops per launch over the
launch's device time, HIP events on the launch stream) and the CPU oracle timed on this host.

    python bench.py [--gpus 1] [--steps 20] [--warmup 3]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mrays/s primary DE raymarch, 1080p Mandelbulb-8, 1 GPU; % FP32 VALU peak"
FP32_PEAK_TFLOPS = 157.3  # MI355X published FP32 vector peak
# Algorithmic flop counts (fixed, independent of the instruction mix): per DE evaluation
# (ray point, threshold, DE finish, step) and per Mandelbulb-8 iteration.
F_DE = 25
F_IT = {"mandelbulb8": 48}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--cpu-seconds", type=float, default=12.0,
                    help="approximate CPU time budget of the CPU baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    return ap.parse_args()


def cpu_baseline(p, budget_s):
    """The FP64 oracle (built with the reference's release flags) on this host's cores, on
    whole frames of the same workload until about budget_s has elapsed."""
    from oracle import oracle
    threads = min(16, os.cpu_count() or 1)
    rays = 0
    frames = 0
    t0 = time.time()
    while True:
        oracle.render(p, threads=threads, fast=True)
        rays += p.imageWidth * p.imageHeight
        frames += 1
        if time.time() - t0 >= budget_s or frames >= 64:
            break
    dt = time.time() - t0
    return {"value": rays / dt / 1e6, "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": "%d full %dx%d frame(s) of the same scene, FP64 oracle built -O3 -ffast-math "
                      "-fopenmp, %.1f s" % (frames, p.imageWidth, p.imageHeight, dt)}


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build measures one GPU only (rays are independent; the multi-GPU "
              "sharding is a design note, not a measured path). Run with --gpus 1.",
              file=sys.stderr)
        sys.exit(1)
    import torch
    from mandelbulber2_amd import runtime, scene

    p = scene.mandelbulb8(args.width, args.height)
    stream = torch.cuda.current_stream(0)
    ctx = runtime.Context(0, stream)
    ctx.upload(p)
    bufs = ctx.alloc_outputs()
    for _ in range(args.warmup):
        ctx.render(bufs)
    torch.cuda.synchronize()

    starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        starts[i].record(stream)
        ctx.render(bufs)
        ends[i].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    launch_ms = [s.elapsed_time(e) for s, e in zip(starts, ends)]
    st = ctx.stats()  # counters of the last launch (identical every frame)
    ctx.close()

    rays = p.imageWidth * p.imageHeight
    ms_per_step = wall * 1e3 / args.steps
    avg_launch_s = sum(launch_ms) / len(launch_ms) / 1e3
    flops = F_DE * st["de_evals"] + F_IT["mandelbulb8"] * st["iterations"]
    achieved = flops / avg_launch_s / 1e12
    out = {
        "metric": METRIC,
        "value": rays / (ms_per_step / 1e3) / 1e6,
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "mandelbulb8_%dx%d" % (p.imageWidth, p.imageHeight),
                   "formula": "mandelbulb power 8, analytic log DE, bailout 10, maxiter 250",
                   "camera": "distance 3.5, fov 53 deg", "spp": 1,
                   "shading": "6-tap normals + Lambert, 1 directional light, no shadows/AO/GI"},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": FP32_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / FP32_PEAK_TFLOPS, "traffic": None,
                     "kernel": "render_kernel<kBulb8>", "avg_launch_ms": avg_launch_s * 1e3,
                     "flops_per_launch": flops},
        "de_evals_per_s": st["de_evals"] / avg_launch_s,
        "iterations_per_s": st["iterations"] / avg_launch_s,
        "de_evals_per_frame": st["de_evals"],
        "iterations_per_frame": st["iterations"],
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(p, args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
