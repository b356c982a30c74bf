"""Benchmark: Mrays/s of the primary DE raymarch on 1080p Mandelbulb-8, one MI355X.

A step renders 8 frames of 1920x1080 Mandelbulb-8 (1 ray per pixel: march + binary search +
6-tap normal + Lambert shading) along a camera orbit, 3 degrees per frame, in one grid launch of
libmbx.so (one 64-lane workgroup per 8x8 tile of every frame), from constants already resident
in HBM; `--frames 1` makes the step the single-frame launch.  A single frame's launch is bounded
by its few longest rays, which graze the surface for hundreds of march steps and finish nearly
alone; the tiles of the other frames of a batch fill the GPU meanwhile.  The JSON line also
carries the FP32 VALU roofline of the batch launch (algorithmic flops over the launch's device
time, HIP events on the launch stream), the same frames rendered one plain launch per frame (the
reference engine's job-per-frame shape: `plain_kernel.batch_speedup`), the persistent tile-queue
kernel, the single-frame launch, and the CPU oracle timed on this host.

    python bench.py [--gpus 1] [--steps 20] [--warmup 3] [--frames 8]
                    [--workload mandelbulb8|material|mandelbox|menger|hybrid|animation64|deepzoom|adjoint512]

`--workload material` times the same 1080p Mandelbulb-8 with the reference's default material:
palette colouring from the default surface gradient, plastic and metallic highlights, glow.
`--workload mandelbox` times the Mandelbox scene instead (scale 2, delta DE by the reference's
finite differences in FP64, hard shadow, 16-ray AO, seeded camera; 1920x1080).
`--workload animation64` times a 64-frame batch (power 2 -> 12, mixed resolutions, 4 spp with
GI depth 2) in one grid launch against one plain launch per frame.
`--workload hybrid` times the reference's hybrid002 example scene (Mandelbox, Menger sponge and
Mandelbulb power 2 in one hybrid sequence) at 1920x1080 on the general FP64 formula kernel.
`--workload deepzoom` times the 1024x1024 deep-zoom Mandelbox (camera 1e-9 from the surface).
`--workload adjoint512` times a differentiable render plus its adjoint (dL/dtheta) at 512x512.
"""
import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP32_PEAK_TFLOPS = 157.3  # MI355X published FP32 vector peak (packed)
FP64_PEAK_TFLOPS = 78.6   # MI355X published FP64 peak
# FP64 FLOPs of the adjoint kernel per pixel gradient (issued lanes; 9.28e9 per 512x512 launch)
ADJOINT_FP64_FLOP_PER_PIXEL = 35400.0
# Algorithmic flop counts (fixed, independent of the instruction mix): per DE evaluation
# (ray point, threshold, DE finish, step) and per formula iteration.
F_DE = 25
F_IT = {"mandelbulb8": 48, "material": 48, "mandelbox": 29, "animation64": 32, "deepzoom": 29,
        # the hybrid's cycle: 2 Mandelbox (29) + 3 Menger sponge (20) + 1 Mandelbulb power 2 (30)
        # iterations, each with the 4D c addition and length: (2*29 + 3*20 + 30) / 6
        "hybrid": 25,
        # Menger-mode KIFS iteration: scale and offset (10), DE scale (1), |z|^2 and sqrt (6),
        # the three compare-swaps and abs folds counted as 3
        "menger": 20}
# A far-field step (no DE evaluation: the step is provably the clamped 3) does the ray point (6),
# |p|^2 (5) and the scan update (2).
F_FAR = 13

WORKLOADS = {
    "mandelbulb8": {
        "metric": "Mrays/s primary DE raymarch, 1080p Mandelbulb-8, 1 GPU; % FP32 VALU peak",
        "kernel": "batch_kernel<kBulb8>", "dtype": "f32", "peak": FP32_PEAK_TFLOPS,
        "config": {"formula": "mandelbulb power 8, analytic log DE, bailout 10, maxiter 250",
                   "camera": "distance 3.5, fov 53 deg", "spp": 1,
                   "shading": "6-tap normals + Lambert, 1 directional light, no shadows/AO/GI"},
    },
    "material": {
        "metric": "Mrays/s primary DE raymarch with the default material (palette colouring, "
                  "plastic + metallic highlights, glow), 1080p Mandelbulb-8, 1 GPU; % FP32 VALU peak",
        "kernel": "batch_kernel<kBulb8> (full variant)", "dtype": "f32", "peak": FP32_PEAK_TFLOPS,
        "config": {"formula": "mandelbulb power 8, analytic log DE, bailout 10, maxiter 250",
                   "camera": "distance 3.5, fov 53 deg", "spp": 1,
                   "shading": "6-tap normals, Lambert + plastic and metallic highlights (roughness "
                              "0.1), palette colour from a 4 x maxiter colouring orbit, glow over "
                              "the march steps, 1 directional light, no shadows/AO/GI",
                   "flops": "the DE model only: colouring orbit, highlights and glow not counted"},
    },
    "animation64": {
        "metric": "Mrays/s (samples) of a 64-frame animation batch in one launch, "
                  "1 GPU; speedup over per-frame plain launches",
        "kernel": "batch_kernel<kBulbTrig>", "dtype": "f32", "peak": FP32_PEAK_TFLOPS,
        "config": {"formula": "mandelbulb power 2 -> 12 over 64 frames (trigonometric form), "
                              "analytic log DE, bailout 10, maxiter 250",
                   "camera": "seed 3, random walk on the radius 2.5-4 shell",
                   "resolutions": "per frame from 1920x1080 / 1280x720 / 640x480 (seeded)",
                   "spp": 4, "shading": "Monte Carlo GI depth 2, Lambert, 1 directional light"},
    },
    "deepzoom": {
        "metric": "Mrays/s deep-zoom Mandelbox (camera 1e-9 from the surface, DE threshold 1e-11), "
                  "1024x1024, 1 GPU; % FP64 peak",
        "kernel": "batch_kernel<kBoxAnalytic>", "dtype": "f64", "peak": FP64_PEAK_TFLOPS,
        "config": {"formula": "mandelbox scale 2, analytic linear DE, maxiter 50",
                   "camera": "seed 5: 1e-9 back from an FP64 surface-search point",
                   "threshold": "constant 1e-11", "spp": 1,
                   "shading": "6-tap normals + Lambert, 1 directional light"},
    },
    "adjoint512": {
        "metric": "Mpixels/s of one differentiable render + adjoint (dL/dtheta, 10 parameters) "
                  "at 512x512, 1 GPU",
        "kernel": "render_diff_kernel + adjoint_kernel", "dtype": "f64", "peak": FP64_PEAK_TFLOPS,
        "config": {"formula": "mandelbulb, seeded power near 8 (trigonometric form), node rotation",
                   "camera": "seed 0, distance 2.6-3.4", "loss": "L2 against a target at "
                   "theta * (1 +- U(0, 0.05))", "parameters": "power, rotation xyz, camera xyz, "
                   "target xyz", "shading": "Lambert, analytic normals, FP64 Newton-refined hits"},
    },
    "menger": {
        "metric": "Mrays/s DE raymarch of the Menger-mode kaleidoscopic IFS with the default material "
                  "and a hard shadow, 1080p, 1 GPU; % FP32 VALU peak",
        "kernel": "split_march_kernel<kIfsAnalytic> + split_shade_kernel<kIfsAnalytic> (split render, "
                  "one timed launch of both)", "plain_kernel": "render_kernel<kIfsAnalytic> (full variant)",
        "dtype": "f32", "peak": FP32_PEAK_TFLOPS,
        "config": {"formula": "kaleidoscopic IFS, Menger mode, scale 3, offset (1,1,1), abs folds, "
                              "maxiter 30, analytic IFS DE (as IFS_001.fract)",
                   "camera": "(2.2, -2.6, 1.7) to the origin, fov 53 deg", "spp": 1,
                   "shading": "default material (palette, plastic + metallic highlights, glow), "
                              "hard shadow from the main light"},
    },
    "hybrid": {
        "metric": "Mrays/s DE raymarch of the hybrid002 example scene (Mandelbox x2, Menger sponge "
                  "x3, Mandelbulb power 2 per cycle), 1080p, 1 GPU; % FP64 peak",
        "kernel": "batch_kernel<kGeneric>", "dtype": "f64", "peak": FP64_PEAK_TFLOPS,
        "config": {"formula": "hybrid sequence from tests/golden/fract/hybrid002.fract: mandelbox "
                              "(scale -0.4) 2 iterations, menger sponge 3, mandelbulb power 2 1; "
                              "analytic linear DE, maxiter 250, FP64 4D orbit",
                   "camera": "the scene's own", "spp": 1,
                   "shading": "the scene's material and main light; effects the path does not "
                              "render (lightmap AO, aux lights, post effects) off"},
    },
    "mandelbox": {
        "metric": "Mrays/s DE raymarch with hard shadow + 16-ray AO, 1080p Mandelbox scale 2, "
                  "1 GPU; % FP64 peak",
        "kernel": "batch_kernel<kBoxDelta>", "dtype": "f64", "peak": FP64_PEAK_TFLOPS,
        "config": {"formula": "mandelbox scale 2, fold 1/2, radii 0.5/1, bailout 100, maxiter 50, "
                              "delta DE (4-orbit finite difference), linear",
                   "camera": "seed 1, radius U(6,10), target origin", "spp": 1,
                   "shading": "hard penetrating shadow, 16 AO rays (Fibonacci hemisphere)"},
    },
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="mandelbulb8")
    ap.add_argument("--frames", type=int, default=8,
                    help="mandelbulb8: frames of the 1080p orbit rendered per step in one launch "
                         "(1 = the single-frame launch)")
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--cpu-seconds", type=float, default=12.0,
                    help="approximate CPU time budget of the CPU baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true", help="skip timing the plain kernel")
    ap.add_argument("--far-field", type=int, default=None, choices=[0, 1],
                    help="MBX_TUNE_FAR_FIELD for the run (default: the library's default)")
    ap.add_argument("--tile-width", type=int, default=None, choices=[4, 8, 16, 32],
                    help="MBX_TUNE_TILE_WIDTH for the run (default: the library's, 8)")
    ap.add_argument("--autotune", action="store_true",
                    help="run the launch-shape autotuner (tile shape x wave count) before timing "
                         "the mandelbulb8 workload instead of the library's default tile shape; "
                         "its choice follows timing noise between shapes within ~3%% of each other")
    ap.add_argument("--fma-probe", action="store_true",
                    help="measure the unpacked FP32 FMA rate (v_fma_f32, the instructions this "
                         "build's kernels may use) with a probe kernel compiled by hipcc, and exit")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run a 256 MiB device-to-device copy, whose FETCH_SIZE/WRITE_SIZE "
                         "counters calibrate the profiler's byte counts")
    return ap.parse_args()


def measured_traffic(workload, w, h):
    """HBM bytes per launch from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary of this
    workload (copy-calibrated), or None when there is none for this size."""
    name = {"image": "compile_image_kernel_1080p.md",
            "mandelbulb8": "batch_kernel_bulb8_mandelbulb8_1080p.md",
            "mandelbulb8_orbit8": "batch_kernel_bulb8_orbit8_1080p.md",
            "material": "batch_kernel_bulb8_material_1080p.md",
            "mandelbox": "batch_kernel_boxdelta_mandelbox_1080p.md",
            "hybrid": "batch_kernel_generic_hybrid002_1080p.md",
            "menger": "split_kernels_ifs_menger_1080p.md"}.get(workload)
    if not name or (w, h) != (1920, 1080):
        return None, None
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None, None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return int(line.split(":")[1]), "profiles/" + name
    return None, None


def make_scene(name, w, h):
    from mandelbulber2_amd import scene
    if name == "mandelbulb8":
        return scene.mandelbulb8(w, h)
    if name == "material":
        return scene.reference_material(scene.mandelbulb8(w, h))
    if name == "animation64":
        return scene.animation64(seed=3)
    if name == "deepzoom":
        return scene.deep_zoom_mandelbox(1024, 1024, seed=5)
    if name == "menger":
        p = scene.reference_material(scene.menger_ifs(w, h))
        p.light.castShadows = 1
        return p
    if name == "hybrid":
        from mandelbulber2_amd import abi, runtime
        with open(os.path.join(ROOT, "tests", "golden", "fract", "hybrid002.fract"), "rb") as f:
            p, _ = runtime.load_fract(f.read())
        p.imageWidth, p.imageHeight = w, h
        if p.ambientOcclusionEnabled and p.ambientOcclusionMode != abi.MBX_AO_FAST:
            p.ambientOcclusionEnabled = 0   # multi-ray AO weighs its rays by a lightmap texture
        return p
    return scene.mandelbox_shadow_ao(w, h, seed=1, scale=2.0)


def host_cpu():
    """Threads for the CPU baseline and a description of the host.  The baseline runs one
    thread per core this process may use: the CPUs of its affinity mask, capped by the cgroup's
    CPU quota and by OMP_NUM_THREADS when either is set (a shared host grants a job a share of
    its cores)."""
    affinity = len(os.sched_getaffinity(0))
    total = os.cpu_count() or affinity
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, period = f.read().split()[:2]
        if q != "max":
            quota = max(1, int(-(-int(q) // int(period))))
    except (OSError, ValueError):
        pass
    omp = None
    try:
        omp = int(os.environ.get("OMP_NUM_THREADS", "")) or None
    except ValueError:
        pass
    threads = min(v for v in (affinity, quota, omp) if v)
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return threads, {"host_cores": total, "affinity_cores": affinity, "cgroup_cpu_quota": quota,
                     "omp_num_threads": omp, "cpu_model": model}


def cpu_baseline(p, budget_s):
    """The FP64 oracle (built with the reference's release flags) on this host's cores, on
    rows of the same frame until about budget_s has elapsed."""
    from oracle import oracle
    threads, host = host_cpu()
    H = p.imageHeight
    band = max(8, H // 16)
    rays = 0
    t0 = time.time()
    y = 0
    while True:
        y1 = min(H, y + band)
        oracle.render(p, threads=threads, rows=(y, y1), fast=True)
        rays += (y1 - y) * p.imageWidth
        y = 0 if y1 >= H else y1
        if time.time() - t0 >= budget_s or rays >= 64 * H * p.imageWidth:
            break
    dt = time.time() - t0
    return dict({"value": rays / dt / 1e6, "unit": "Mrays/s", "cores": threads, "kind": "port",
                 "sample": "%d rays (row bands of the same %dx%d frame), FP64 oracle built -O3 "
                           "-ffast-math -fopenmp, %d threads, %.1f s"
                           % (rays, p.imageWidth, H, threads, dt)}, **host)


def time_kernel(torch, ctx, bufs, stream, steps, warmup, step_fn=None):
    step_fn = step_fn or (lambda: ctx.render(bufs))
    for _ in range(warmup):
        step_fn()
    torch.cuda.synchronize()
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        starts[i].record(stream)
        step_fn()
        ends[i].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    launch_ms = [s.elapsed_time(e) for s, e in zip(starts, ends)]
    return wall, launch_ms


def cpu_baseline_batch(frames, budget_s):
    """The FP64 oracle on row bands of the batch's first and last frames (power 2 and 12) until
    about budget_s has elapsed; Mrays/s counts samples."""
    from oracle import oracle
    threads, host = host_cpu()
    rays = 0
    t0 = time.time()
    k = 0
    while True:
        p = frames[0] if k % 2 == 0 else frames[-1]
        y = (k // 2 * 8) % p.imageHeight
        y1 = min(p.imageHeight, y + 8)
        oracle.render(p, threads=threads, rows=(y, y1), fast=True)
        rays += (y1 - y) * p.imageWidth * p.DOFSamples
        k += 1
        if time.time() - t0 >= budget_s:
            break
    dt = time.time() - t0
    return dict({"value": rays / dt / 1e6, "unit": "Mrays/s", "cores": threads, "kind": "port",
                 "sample": "%d samples (8-row bands of frames 0 and 63, 4 spp, GI depth 2), FP64 "
                           "oracle built -O3 -ffast-math -fopenmp, %d threads, %.1f s"
                           % (rays, threads, dt)}, **host)


def cpu_baseline_gradient(p, budget_s):
    """The same gradient by central differences of the FP64 oracle render (21 renders per
    gradient, the only way the reference's CPU renderer gives one), on a row band of the frame
    scaled down to about budget_s; Mpixels/s per gradient."""
    from oracle import oracle
    from mandelbulber2_amd import diff
    threads, host = host_cpu()
    theta = diff.theta_of(p)
    rows = 8
    y0 = p.imageHeight // 2 - rows // 2
    t0 = time.time()
    grads = 0
    while True:
        for k in range(-1, 10):
            for sgn in ((0.0,) if k < 0 else (1e-5, -1e-5)):
                th = theta.copy()
                if k >= 0:
                    th[k] += sgn
                oracle.render_diff(diff.with_theta(p, th), threads=threads, rows=(y0, y0 + rows))
        grads += 1
        if time.time() - t0 >= budget_s:
            break
    dt = time.time() - t0
    px = grads * rows * p.imageWidth
    return dict({"value": px / dt / 1e6, "unit": "Mpixels/s", "cores": threads, "kind": "port",
                 "sample": "%d gradient(s) of the middle %d rows of the 512x512 frame by 21 FP64 oracle "
                           "renders each (central differences), %d threads, %.1f s"
                           % (grads, rows, threads, dt)}, **host)


def bench_adjoint(args, torch, runtime, wl):
    """One step = differentiable render at theta + loss gradient + adjoint, 512x512."""
    from mandelbulber2_amd import diff, scene
    import numpy as np
    p = scene.inverse_rendering(512, 512, seed=0)
    theta = diff.theta_of(p)
    rs = np.random.default_rng(100)
    theta_t = theta * (1.0 + rs.choice([-1.0, 1.0], size=theta.size) * rs.uniform(0.0, 0.05, size=theta.size))
    r = diff.DiffRenderer(p)
    tgt_bufs, _ = r.render(theta_t)
    target = tgt_bufs["rgb"].clone()
    stream = torch.cuda.current_stream(0)

    def step():
        bufs, ck = r.render(theta)
        g = r.adjoint(theta, ck, (2.0 * (bufs["rgb"] - target)).contiguous())
        return g
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    fwd_ms, adj_ms = [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        bufs, ck = r.render(theta)
        fwd_ms.append(r.ctx.stats()["kernel_ms"])
        g = r.adjoint(theta, ck, (2.0 * (bufs["rgb"] - target)).contiguous())
        adj_ms.append(r.ctx.stats()["kernel_ms"])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    r.close()
    n = p.imageWidth * p.imageHeight
    ms_per_step = wall * 1e3 / args.steps
    out = {
        "metric": wl["metric"], "value": n / (ms_per_step / 1e3) / 1e6, "unit": "Mpixels/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": wl["dtype"],
        "data": "synthetic",
        "config": dict({"workload": "adjoint512", "width": 512, "height": 512}, **wl["config"]),
        # the adjoint kernel against the FP64 peak: its FP64 FLOPs per pixel gradient counted by
        # SQ_INSTS_VALU_FLOPS_FP64 x 64 lanes (the counter's unit, calibrated on an FMA chain of
        # known size); the render half is a march whose time is set by its longest rays
        "roofline": {"bound": "valu", "kernel": "adjoint_kernel",
                     "achieved": ADJOINT_FP64_FLOP_PER_PIXEL * n / (statistics.median(adj_ms) * 1e-3) / 1e12,
                     "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": ADJOINT_FP64_FLOP_PER_PIXEL * n / (statistics.median(adj_ms) * 1e-3) / 1e12
                     / FP64_PEAK_TFLOPS, "traffic": None,
                     "flop_per_pixel": ADJOINT_FP64_FLOP_PER_PIXEL},
        "kernels_ms": {"render_diff": statistics.median(fwd_ms), "adjoint": statistics.median(adj_ms)},
        "gradient": [float(v) for v in g],
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline_gradient(p, args.cpu_seconds)
    print(json.dumps(out), flush=True)


def bench_orbit(args, torch, runtime, wl):
    """One step = `args.frames` 1920x1080 Mandelbulb-8 frames along a camera orbit (3 degrees
    per frame) rendered in one grid launch.  The same frames are also timed one plain launch per
    frame (the reference engine's job-per-frame shape) and in the persistent kernel, and the
    orbit's first frame alone (the single-frame launch)."""
    from mandelbulber2_amd import scene
    n = args.frames
    frames = scene.mandelbulb8_orbit(args.width, args.height, n)
    stream = torch.cuda.current_stream(0)
    ctx = runtime.Context(0, stream)
    if args.far_field is not None:
        ctx.tune(runtime.MBX_TUNE_FAR_FIELD, args.far_field)
    ctx.tune(runtime.MBX_TUNE_TIMING, 0)
    bufs = [ctx.alloc_outputs(p) for p in frames]
    ctx.tune(runtime.MBX_TUNE_KERNEL, runtime.KERNEL_GRID)
    wall, launch_ms = time_kernel(torch, ctx, None, stream, args.steps, args.warmup,
                                  lambda: ctx.render_batch(frames, bufs))
    st = ctx.stats()
    legs = {}
    if not args.no_plain:
        psteps = max(3, args.steps // 2)
        ctx.tune(runtime.MBX_TUNE_KERNEL, runtime.KERNEL_PLAIN)
        pwall, pms = time_kernel(torch, ctx, None, stream, psteps, 1, lambda: ctx.render_batch(frames, bufs))
        legs["plain_kernel"] = {"kernel": "render_kernel<kBulb8>, one launch per frame",
                                "ms_per_step": pwall * 1e3 / psteps,
                                "mrays_per_s": n * args.width * args.height / (pwall / psteps) / 1e6}
        ctx.tune(runtime.MBX_TUNE_KERNEL, runtime.KERNEL_PERSISTENT)
        qwall, _ = time_kernel(torch, ctx, None, stream, psteps, 1, lambda: ctx.render_batch(frames, bufs))
        legs["persistent_kernel"] = {"kernel": "persistent_kernel<kBulb8>, tile queues",
                                     "ms_per_step": qwall * 1e3 / psteps}
        ctx.tune(runtime.MBX_TUNE_KERNEL, runtime.KERNEL_GRID)
        ctx.upload(frames[0])
        swall, sms = time_kernel(torch, ctx, bufs[0], stream, args.steps, args.warmup)
        s1 = ctx.stats()
        sflops = F_DE * s1["de_evals"] + F_IT["mandelbulb8"] * s1["iterations"] + F_FAR * s1["far_steps"]
        savg = sum(sms) / len(sms) / 1e3
        legs["single_frame"] = {"kernel": "batch_kernel<kBulb8>, one frame per launch",
                                "ms_per_step": swall * 1e3 / args.steps,
                                "mrays_per_s": args.width * args.height / (swall / args.steps) / 1e6,
                                "avg_launch_ms": savg * 1e3, "flops_per_launch": sflops,
                                "roofline_frac": sflops / savg / 1e12 / wl["peak"]}
    image = None
    if not args.no_plain:
        p = frames[0]
        rgb16 = torch.empty_like(bufs[0]["rgb"], dtype=torch.uint16)
        rgb8 = torch.empty_like(bufs[0]["rgb"], dtype=torch.uint8)
        iwall, ims = time_kernel(torch, ctx, None, stream, args.steps, args.warmup,
                                 lambda: ctx.compile_image(bufs[0]["rgb"], p.image, rgb16, rgb8))
        ibytes = 21 * p.imageWidth * p.imageHeight
        iavg = sum(ims) / len(ims) / 1e3
        itraffic, isrc = measured_traffic("image", p.imageWidth, p.imageHeight)
        image = {"kernel": "compile_image_kernel", "bound": "hbm", "avg_launch_ms": iavg * 1e3,
                 "achieved": ibytes / iavg / 1e9, "peak": 8000.0, "unit": "GB/s",
                 "frac": ibytes / iavg / 1e9 / 8000.0, "bytes_per_launch": ibytes,
                 "traffic": itraffic, "traffic_source": isrc}
    ctx.close()
    if args.calibrate_copy:
        a = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    rays = n * args.width * args.height
    ms_per_step = wall * 1e3 / args.steps
    avg_launch_s = sum(launch_ms) / len(launch_ms) / 1e3
    flops = F_DE * st["de_evals"] + F_IT["mandelbulb8"] * st["iterations"] + F_FAR * st["far_steps"]
    achieved = flops / avg_launch_s / 1e12
    traffic, traffic_src = measured_traffic("mandelbulb8_orbit%d" % n, args.width, args.height)
    out = {
        "metric": wl["metric"], "value": rays / (ms_per_step / 1e3) / 1e6, "unit": "Mrays/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": wl["dtype"],
        "data": "synthetic",
        "config": dict({"workload": "mandelbulb8_%dx%d_orbit%d" % (args.width, args.height, n),
                        "frames_per_step": n,
                        "orbit": "camera at distance 3.5 in the z = 0 plane, 3 degrees per frame",
                        "launch": "one grid launch per step (8x8 tiles of every frame)"}, **wl["config"]),
        "roofline": {"bound": "valu", "achieved": achieved, "peak": wl["peak"], "unit": "TFLOP/s",
                     "frac": achieved / wl["peak"], "traffic": traffic, "traffic_source": traffic_src,
                     "kernel": wl["kernel"], "avg_launch_ms": avg_launch_s * 1e3,
                     "median_launch_ms": statistics.median(launch_ms), "flops_per_launch": flops},
        "de_evals_per_s": st["de_evals"] / avg_launch_s,
        "iterations_per_s": st["iterations"] / avg_launch_s,
        "de_evals_per_step": st["de_evals"], "iterations_per_step": st["iterations"],
        "far_steps_per_step": st["far_steps"],
    }
    out.update(legs)
    if "plain_kernel" in legs:
        out["plain_kernel"]["batch_speedup"] = legs["plain_kernel"]["ms_per_step"] / ms_per_step
    if image is not None:
        out["image_compile"] = image
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(frames[0], args.cpu_seconds)
    print(json.dumps(out), flush=True)


def bench_batch(args, torch, runtime, wl):
    """One step = the whole 64-frame batch in one grid launch; the
    comparison renders the same frames one plain launch per frame."""
    frames = make_scene(args.workload, 0, 0)
    stream = torch.cuda.current_stream(0)
    ctx = runtime.Context(0, stream)
    if args.far_field is not None:
        ctx.tune(runtime.MBX_TUNE_FAR_FIELD, args.far_field)
    bufs = [ctx.alloc_outputs(p) for p in frames]
    ctx.tune(runtime.MBX_TUNE_KERNEL, runtime.KERNEL_GRID)
    wall, launch_ms = time_kernel(torch, ctx, None, stream, args.steps, args.warmup,
                                  lambda: ctx.render_batch(frames, bufs))
    st = ctx.stats()
    plain = None
    if not args.no_plain:
        ctx.tune(runtime.MBX_TUNE_KERNEL, runtime.KERNEL_PLAIN)

        def per_frame():
            for p, b in zip(frames, bufs):
                ctx.upload(p)
                ctx.render(b)
        psteps = max(1, args.steps // 2)
        pwall, pms = time_kernel(torch, ctx, None, stream, psteps, 1, per_frame)
        plain = {"kernel": wl["kernel"].replace("batch_kernel", "render_kernel") +
                 " x 64 launches", "ms_per_step": pwall * 1e3 / psteps,
                 "median_launch_ms": statistics.median(pms)}
    ctx.close()
    ms_per_step = wall * 1e3 / args.steps
    avg_launch_s = sum(launch_ms) / len(launch_ms) / 1e3
    flops = F_DE * st["de_evals"] + F_IT[args.workload] * st["iterations"] + F_FAR * st["far_steps"]
    achieved = flops / avg_launch_s / 1e12
    out = {
        "metric": wl["metric"], "value": st["rays"] / (ms_per_step / 1e3) / 1e6, "unit": "Mrays/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": wl["dtype"],
        "data": "synthetic",
        "config": dict({"workload": "animation64", "frames": len(frames),
                        "samples_per_batch": st["rays"]}, **wl["config"]),
        "roofline": {"bound": "valu", "achieved": achieved, "peak": wl["peak"], "unit": "TFLOP/s",
                     "frac": achieved / wl["peak"], "traffic": None, "kernel": wl["kernel"],
                     "avg_launch_ms": avg_launch_s * 1e3,
                     "median_launch_ms": statistics.median(launch_ms), "flops_per_launch": flops},
        "de_evals_per_s": st["de_evals"] / avg_launch_s,
        "iterations_per_s": st["iterations"] / avg_launch_s,
    }
    if plain is not None:
        plain["batch_speedup"] = plain["ms_per_step"] / ms_per_step
        out["plain_kernel"] = plain
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline_batch(frames, args.cpu_seconds)
    print(json.dumps(out), flush=True)


FMA_PROBE_SRC = r"""
#include <hip/hip_runtime.h>
#include <cstdio>
// eight independent FMA chains per lane; 8192 workgroups of 256 lanes (8 waves per SIMD)
__global__ __launch_bounds__(256) void fma_kernel(float *out, int iters, float a, float b)
{
	float x0 = threadIdx.x, x1 = x0 + 1, x2 = x0 + 2, x3 = x0 + 3, x4 = x0 + 4, x5 = x0 + 5, x6 = x0 + 6, x7 = x0 + 7;
	for (int i = 0; i < iters; i++)
	{
		x0 = __builtin_fmaf(x0, a, b); x1 = __builtin_fmaf(x1, a, b); x2 = __builtin_fmaf(x2, a, b);
		x3 = __builtin_fmaf(x3, a, b); x4 = __builtin_fmaf(x4, a, b); x5 = __builtin_fmaf(x5, a, b);
		x6 = __builtin_fmaf(x6, a, b); x7 = __builtin_fmaf(x7, a, b);
	}
	out[blockIdx.x * 256 + threadIdx.x] = x0 + x1 + x2 + x3 + x4 + x5 + x6 + x7;
}
int main()
{
	const int blocks = 8192, iters = 20000;
	float *out = nullptr;
	if (hipMalloc(&out, sizeof(float) * blocks * 256) != hipSuccess) return 1;
	hipEvent_t e0, e1;
	if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return 1;
	fma_kernel<<<blocks, 256>>>(out, 100, 0.999f, 0.001f);
	if (hipDeviceSynchronize() != hipSuccess) return 1;
	float best = 1e30f;
	for (int rep = 0; rep < 5; rep++)
	{
		(void)hipEventRecord(e0);
		fma_kernel<<<blocks, 256>>>(out, iters, 0.999f, 0.001f);
		(void)hipEventRecord(e1);
		if (hipEventSynchronize(e1) != hipSuccess) return 1;
		float ms = 0.0f;
		(void)hipEventElapsedTime(&ms, e0, e1);
		best = ms < best ? ms : best;
	}
	printf("%.6f\n", 2.0 * 8.0 * iters * (double)blocks * 256 / best / 1e9);
	return 0;
}
"""


def fma_probe():
    """Unpacked FP32 FMA throughput of the GPU (TFLOP/s), from a probe kernel built with the same
    flags as the library (packed FP32 instructions off)."""
    import subprocess
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        src, exe = os.path.join(d, "fma.hip"), os.path.join(d, "fma")
        with open(src, "w") as f:
            f.write(FMA_PROBE_SRC)
        subprocess.check_call(["/opt/rocm/bin/hipcc", "-O3", "--offload-arch=gfx950", "-Xclang", "-target-feature",
                               "-Xclang", "-packed-fp32-ops", src, "-o", exe], stderr=subprocess.DEVNULL)
        tflops = float(subprocess.check_output([exe], timeout=120).decode().split()[-1])
    print(json.dumps({"metric": "unpacked FP32 FMA rate (v_fma_f32)", "value": tflops, "unit": "TFLOP/s",
                      "published_fp32_peak": FP32_PEAK_TFLOPS, "fraction_of_published": tflops / FP32_PEAK_TFLOPS}),
          flush=True)


def main():
    args = parse()
    if args.fma_probe:
        return fma_probe()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build measures one GPU only (rays are independent; the multi-GPU "
              "sharding is a design note, not a measured path). Run with --gpus 1.",
              file=sys.stderr)
        sys.exit(1)
    import torch
    from mandelbulber2_amd import runtime

    wl = WORKLOADS[args.workload]
    if args.workload == "animation64":
        return bench_batch(args, torch, runtime, wl)
    if args.workload == "adjoint512":
        return bench_adjoint(args, torch, runtime, wl)
    if args.workload == "mandelbulb8" and args.frames > 1:
        return bench_orbit(args, torch, runtime, wl)
    p = make_scene(args.workload, args.width, args.height)
    stream = torch.cuda.current_stream(0)
    ctx = runtime.Context(0, stream)
    if args.far_field is not None:
        ctx.tune(runtime.MBX_TUNE_FAR_FIELD, args.far_field)
    ctx.upload(p)
    bufs = ctx.alloc_outputs()
    ctx.tune(runtime.MBX_TUNE_KERNEL, runtime.KERNEL_GRID)
    launch_shape = None
    if args.tile_width:
        ctx.tune(runtime.MBX_TUNE_TILE_WIDTH, args.tile_width)
        launch_shape = {"waves": ctx.get_tune(runtime.MBX_TUNE_WAVES), "tile_width": args.tile_width,
                        "source": "--tile-width"}
    elif args.workload == "mandelbulb8" and args.autotune:
        waves, _ = ctx.autotune(bufs, reps=2)
        launch_shape = {"waves": waves, "tile_width": ctx.get_tune(runtime.MBX_TUNE_TILE_WIDTH),
                        "source": "mbx_autotune"}
    else:
        launch_shape = {"waves": ctx.get_tune(runtime.MBX_TUNE_WAVES),
                        "tile_width": ctx.get_tune(runtime.MBX_TUNE_TILE_WIDTH), "source": "default"}
    # the bench brackets each launch with its own events; the library's pair would add a second
    # pair of stream markers per frame
    ctx.tune(runtime.MBX_TUNE_TIMING, 0)
    wall, launch_ms = time_kernel(torch, ctx, bufs, stream, args.steps, args.warmup)
    st = ctx.stats()  # counters of the last launch (identical every frame)

    one_kernel = None
    if args.workload == "menger" and not args.no_plain:
        # the same frame as one launch of the full-variant batch kernel (the split render off)
        ctx.tune(runtime.MBX_TUNE_SPLIT, 0)
        owall, oms = time_kernel(torch, ctx, bufs, stream, args.steps, args.warmup)
        ctx.tune(runtime.MBX_TUNE_SPLIT, 1)
        one_kernel = {"kernel": "batch_kernel<kIfsAnalytic> (full variant, one launch)",
                      "ms_per_step": owall * 1e3 / args.steps, "avg_launch_ms": sum(oms) / len(oms),
                      "split_speedup": owall / wall}
    plain = None
    if not args.no_plain:
        ctx.tune(runtime.MBX_TUNE_KERNEL, runtime.KERNEL_PLAIN)
        psteps = max(3, args.steps // 2)
        pwall, pms = time_kernel(torch, ctx, bufs, stream, psteps, 1)
        ctx.tune(runtime.MBX_TUNE_TILE_WIDTH, 8)
        plain = {"kernel": wl.get("plain_kernel", wl["kernel"].replace("batch_kernel", "render_kernel")),
                 "ms_per_step": pwall * 1e3 / psteps, "median_launch_ms": statistics.median(pms)}
        ctx.tune(runtime.MBX_TUNE_KERNEL, runtime.KERNEL_PERSISTENT)
        qwall, qms = time_kernel(torch, ctx, bufs, stream, psteps, 1)
        plain["persistent_kernel_ms_per_step"] = qwall * 1e3 / psteps
        ctx.tune(runtime.MBX_TUNE_KERNEL, runtime.KERNEL_GRID)
    image = None
    if not args.no_plain:
        # the image post-processing of the rendered frame (cImage::CompileImage +
        # ConvertTo8bitChar): 12 bytes read, 6 + 3 written per pixel, HBM-bound
        rgb16 = torch.empty_like(bufs["rgb"], dtype=torch.uint16)
        rgb8 = torch.empty_like(bufs["rgb"], dtype=torch.uint8)
        iwall, ims = time_kernel(torch, ctx, bufs, stream, args.steps, args.warmup,
                                 lambda: ctx.compile_image(bufs["rgb"], p.image, rgb16, rgb8))
        ibytes = 21 * p.imageWidth * p.imageHeight
        iavg = sum(ims) / len(ims) / 1e3
        itraffic, isrc = measured_traffic("image", p.imageWidth, p.imageHeight)
        image = {"kernel": "compile_image_kernel", "bound": "hbm", "avg_launch_ms": iavg * 1e3,
                 "achieved": ibytes / iavg / 1e9, "peak": 8000.0, "unit": "GB/s",
                 "frac": ibytes / iavg / 1e9 / 8000.0, "bytes_per_launch": ibytes,
                 "traffic": itraffic, "traffic_source": isrc}
    ctx.close()
    if args.calibrate_copy:
        a = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        del a, b

    rays = p.imageWidth * p.imageHeight
    ms_per_step = wall * 1e3 / args.steps
    avg_launch_s = sum(launch_ms) / len(launch_ms) / 1e3
    flops = F_DE * st["de_evals"] + F_IT[args.workload] * st["iterations"] + F_FAR * st["far_steps"]
    traffic, traffic_src = measured_traffic(args.workload, p.imageWidth, p.imageHeight)
    achieved = flops / avg_launch_s / 1e12
    out = {
        "metric": wl["metric"],
        "value": rays / (ms_per_step / 1e3) / 1e6,
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": wl["dtype"],
        "data": "synthetic",
        "config": dict({"workload": "%s_%dx%d" % (args.workload, p.imageWidth, p.imageHeight),
                        "launch": launch_shape}, **wl["config"]),
        "roofline": {"bound": "valu", "achieved": achieved, "peak": wl["peak"],
                     "unit": "TFLOP/s", "frac": achieved / wl["peak"], "traffic": traffic,
                     "traffic_source": traffic_src,
                     "kernel": wl["kernel"], "avg_launch_ms": avg_launch_s * 1e3,
                     "median_launch_ms": statistics.median(launch_ms),
                     "flops_per_launch": flops},
        "de_evals_per_s": st["de_evals"] / avg_launch_s,
        "iterations_per_s": st["iterations"] / avg_launch_s,
        "de_evals_per_frame": st["de_evals"],
        "iterations_per_frame": st["iterations"],
        "far_steps_per_frame": st["far_steps"],
    }
    if plain is not None:
        plain["batch_speedup"] = plain["ms_per_step"] / ms_per_step
        out["plain_kernel"] = plain
    if one_kernel is not None:
        out["one_kernel"] = one_kernel
    if image is not None:
        out["image_compile"] = image
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(p, args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
