#!/usr/bin/env python3
"""Benchmark of the MI355X render-pass executor.

One step = one render pass of the workload (clear, the draw, the final
colour/depth store) over device-resident inputs.  Default workload: the
textured 1080p single draw -- sphere 256x128 (65,024 triangles, displaced by
seeded value noise), 1024x1024 sRGB checker texture sampled bilinear with
Repeat, perspective-correct uv, 1920x1080 Rgba8Unorm + Depth32Float Less.

Prints one JSON line with Mfragments/s, the frame time, the dominant kernel's
HBM roofline fraction (event-timed on the device stream) and the CPU
oracle's rate on the same pass (single thread).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "wgpu-cpu_amd"))

HBM_PEAK_GBS = 8000.0
KERNELS = ["k_raster", "k_tiles", "k_setup", "k_clip", "k_clear", "k_end_submission"]


def make_spec(name: str):
    from wgc import scenes
    if name == "c2":
        return scenes.config2_textured(extended=True), "textured 1080p: 1920x1080, sphere 256x128 (65,024 tris), " \
                                                          "1024^2 sRGB checker, bilinear+Repeat, perspective uv, Depth32Float Less"
    if name == "c2ref":
        return scenes.config2_textured(extended=False), "textured 1080p, reference semantics (linear uv, nearest, no sRGB decode)"
    if name == "c1":
        return scenes.config1_plain_small(), "plain 256x256 sphere 32x16 (960 tris), vertex colour, Depth32Float Less"
    if name == "c3":
        return scenes.config3_overdraw(), "overdraw 200k tris over 2560x2440, overwrite in API order, no depth"
    if name == "c3blend":
        return scenes.config3_overdraw(blend=True), "overdraw 200k tris over 2560x2440, src-over in API order"
    raise SystemExit(f"unknown config {name}")


def time_gpu(dev, pp, steps: int, warmup: int):
    for _ in range(warmup):
        pp.submit()
    dev.poll()
    t0 = time.perf_counter()
    dev.timestamp(0)
    for _ in range(steps):
        pp.submit()
    dev.timestamp(1)
    dev.poll()
    wall = time.perf_counter() - t0
    ev_ms = dev.elapsed_ms(0, 1)
    return ev_ms, wall * 1e3


def kernel_profile(dev, pp, reps: int, warmup: int = 5):
    """Average event-timed duration of each kernel over `reps` submissions,
    after `warmup` untimed ones in the same (kernel-timing) launch mode."""
    dev.set_kernel_timing(True)
    for _ in range(warmup):
        pp.submit()
    dev.poll()
    dev.set_kernel_timing(True)   # clears what the warm-up launches recorded
    for _ in range(reps):
        pp.submit()
    dev.poll()
    out = {}
    for k in KERNELS:
        ms, n = dev.kernel_time(k)
        if n:
            out[k] = {"avg_us": 1e3 * ms / n, "launches": n}
    dev.set_kernel_timing(False)
    return out


def cpu_baseline(spec, budget_s: float, max_reps: int):
    """Times the CPU oracle on the pass; returns the median seconds, the
    fragment count, the repetitions and the last rendered image (the checker
    of the GPU's image)."""
    from oracle import bridge
    times = []
    r = None
    t_start = time.perf_counter()
    while len(times) < max_reps and (time.perf_counter() - t_start) < budget_s or len(times) < 1:
        t0 = time.perf_counter()
        r = bridge.render(spec)
        times.append(time.perf_counter() - t0)
    times.sort()
    med = times[len(times) // 2]
    return med, r.fragments, len(times), r


def parity_of(got, ref) -> dict:
    """Byte comparison of the GPU's final attachments with the oracle's."""
    import numpy as np
    out = {"fragments": bool(got.fragments == ref.fragments)}
    if ref.color is not None:
        diff = (got.color != ref.color).any(-1)
        out["color_px_differing"] = int(diff.sum())
        out["color_max_lsb"] = int(np.abs(got.color.astype(int) - ref.color.astype(int)).max())
    if ref.depth is not None:
        out["depth_px_differing"] = int((got.depth.view(np.uint32) != ref.depth.view(np.uint32)).sum())
    out["parity"] = bool(out["fragments"] and out.get("color_px_differing", 0) == 0
                         and out.get("depth_px_differing", 0) == 0)
    return out


def run_config(dev_cache, name, steps, warmup, wgsl=False):
    import wgc
    from wgc import runner, scenes
    spec, desc = make_spec(name)
    if wgsl:
        # the same pass with the workload's WGSL shaders, compiled into the
        # kernels at pipeline creation, in place of the built-in programs
        spec = scenes.as_wgsl(spec)
        desc += "; WGSL shaders (wgpu-cpu_amd/shaders) through the WGSL front end"
    sem = spec.semantics
    if sem not in dev_cache:
        dev_cache[sem] = wgc.Device(0, sem)
    dev = dev_cache[sem]
    pp = runner.PreparedPass(dev, spec)
    pp.submit()
    dev.poll()
    frags = dev.last_pass_stats().fragments
    ev_ms, wall_ms = time_gpu(dev, pp, steps, warmup)
    return spec, desc, dev, pp, frags, ev_ms, wall_ms


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--config", default="c2", choices=["c2", "c2ref", "c1", "c3", "c3blend"])
    ap.add_argument("--shaders", default="builtin", choices=["builtin", "wgsl"],
                    help="built-in programs or the same shaders as WGSL modules")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-seconds", type=float, default=10.0,
                    help="length of each serving run (the full workload is 60 s)")
    ap.add_argument("--serve-rate", type=float, default=2000.0, help="Poisson arrival rate, requests/s")
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB (HBM counter calibration for profiles)")
    ap.add_argument("--profile-batch", type=int, default=0, metavar="N",
                    help="only run N submissions of the batched workload (for counter profiles)")
    ap.add_argument("--raster-tile", type=int, default=0, choices=range(8),
                    help="raster tile shape of --profile-batch (wgc.RASTER_TILES: 6 is 32x32 tiles of 512 threads, 7 of 64)")
    args = ap.parse_args()

    if args.calibrate_copy:
        import torch
        n = args.calibrate_copy << 20
        src = torch.empty(n, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calibration_copy_bytes": n, "reads": n, "writes": n}), flush=True)
        return

    if args.profile_batch:
        profile_batch(args.profile_batch, args.raster_tile)
        return

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus > 1 or world > 1:
        print("bench.py: this build runs on one GPU; passes and requests shard across GPUs "
              "without exchange (see the design notes), but no multi-GPU run is part of it.", file=sys.stderr)
        sys.exit(2)

    import wgc
    from wgc import scenes
    wgc.load_library()  # fails loudly when libwgc.so is missing

    devs = {}
    spec, desc, dev, pp, frags, ev_ms, wall_ms = run_config(devs, args.config, args.steps, args.warmup,
                                                            wgsl=args.shaders == "wgsl")
    ms_per_step = ev_ms / args.steps
    value = frags * args.steps / (ev_ms * 1e-3) / 1e6

    prof = kernel_profile(dev, pp, reps=max(10, min(50, args.steps)))
    dominant = max(prof, key=lambda k: prof[k]["avg_us"] * prof[k]["launches"]) if prof else None
    alg_bytes = scenes.algorithmic_bytes(spec)
    roofline = None
    if dominant:
        # the dominant kernel's own bytes, not the whole pass's
        kbytes = kernel_bytes(spec, dominant)
        avg_s = prof[dominant]["avg_us"] * 1e-6
        achieved = kbytes / avg_s / 1e9
        roofline = {"bound": "hbm", "kernel": dominant, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                    "traffic": measured_traffic(args.config, dominant),
                    "algorithmic_bytes_per_launch": kbytes,
                    "kernel_avg_us": round(prof[dominant]["avg_us"], 2)}
    pass_bytes_rate = alg_bytes / (ms_per_step * 1e-3) / 1e9
    others = {}
    for k in prof:
        if k != dominant and k in ("k_setup", "k_clip", "k_tiles", "k_raster"):
            kb = kernel_bytes(spec, k)
            gbs = kb / (prof[k]["avg_us"] * 1e-6) / 1e9
            others[k] = {"algorithmic_bytes_per_launch": kb, "avg_us": round(prof[k]["avg_us"], 2),
                         "achieved_GBps": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)}

    cpu = None
    parity = None
    if not args.no_cpu:
        med, cfrags, reps, ref_img = cpu_baseline(spec, args.cpu_budget, 50)
        # the image the timed steps left in the attachments, against the oracle's
        parity = parity_of(pp.result(), ref_img)
        cpu_value = cfrags / med / 1e6
        cpu = {"value": round(cpu_value, 3), "unit": "Mfragments/s", "cores": 1, "kind": "port",
               "sample": f"{reps} full passes of the same workload on the CPU oracle (C restatement of the "
                         f"reference executor, one thread), median {med * 1e3:.1f} ms/pass",
               "frame_ms": round(med * 1e3, 3), "cpu_model": cpu_model(), "nproc": os.cpu_count()}
        if cfrags != frags:
            cpu["warning"] = f"fragment counts differ: cpu {cfrags} gpu {frags}"

    extra = {}
    if not args.no_extra:
        for name in ["c2ref", "c1", "c3", "c3blend"]:
            if name == args.config:
                continue
            s2, d2, dv2, pp2, f2, e2, w2 = run_config(devs, name, max(10, args.steps // 4), 3)
            n2 = max(10, args.steps // 4)
            extra[name] = {"workload": d2, "Mfragments_per_s": round(f2 * n2 / (e2 * 1e-3) / 1e6, 2),
                           "frame_ms": round(e2 / n2, 4), "fragments": f2}
        # the headline pass with the other shader form
        other = args.shaders != "wgsl"
        s2, d2, dv2, pp2, f2, e2, w2 = run_config(devs, args.config, args.steps, 3, wgsl=other)
        extra[args.config + ("_wgsl" if other else "_builtin")] = {
            "workload": d2, "Mfragments_per_s": round(f2 * args.steps / (e2 * 1e-3) / 1e6, 2),
            "frame_ms": round(e2 / args.steps, 4), "fragments": f2}

    readback = None
    if not args.no_extra:
        readback = run_readback(dev, args.steps)

    batched = None
    if not args.no_batch:
        batched = run_batched(devs, args)

    serve = None
    if not args.no_serve:
        serve = run_serving(devs, args)

    line = {
        "metric": "Mfragments/s (render pass throughput; 1080p frame ms beside it)",
        "value": round(value, 2),
        "unit": "Mfragments/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded scene generators)",
        "config": {"workload": desc, "pass": args.config, "fragments_per_pass": frags,
                   "wall_ms_per_step": round(wall_ms / args.steps, 4)},
        "frame_ms": round(ms_per_step, 4),
        "pass_algorithmic_GBps": round(pass_bytes_rate, 1),
        "kernels_us": {k: round(v["avg_us"], 2) for k, v in prof.items()},
        "roofline": roofline,
        "other_kernels_hbm": others,
        "cpu_baseline": cpu,
        "parity": parity["parity"] if parity else None,
        "parity_detail": parity,
        "speedup_vs_cpu": round(value / cpu["value"], 1) if cpu else None,
        "extra": extra,
        "readback": readback,
        "batched": batched,
        "serving": serve,
    }
    print(json.dumps(line), flush=True)


def run_readback(dev, steps):
    """wgpu-style readback of a 1080p RGBA8 target: copy_texture_to_buffer
    (7680-byte rows, already 256-aligned) submitted `steps` times; device
    time per copy and the bytes it moves (read + write)."""
    import wgc
    w, h = 1920, 1080
    tex = dev.create_texture(w, h, wgc.RGBA8_UNORM)
    buf = dev.create_buffer(w * h * 4)
    enc = dev.create_command_encoder()
    for _ in range(10):
        enc.copy_texture_to_buffer(tex, (0, 0), buf, 0, w * 4, w, h)
    cb = enc.finish()
    # two untimed submissions first (with OPTION_GRAPHS the first captures
    # and instantiates the command buffer's graph, outside the timed window)
    for _ in range(2):
        dev.submit([cb])
    dev.poll()
    reps = max(10, steps // 2)
    dev.timestamp(0)
    for _ in range(reps):
        dev.submit([cb])
    dev.timestamp(1)
    dev.poll()
    ms = dev.elapsed_ms(0, 1) / (reps * 10)
    nbytes = 2 * w * h * 4
    return {"op": "copy_texture_to_buffer 1920x1080 RGBA8", "us_per_copy": round(ms * 1e3, 2),
            "GBps": round(nbytes / (ms * 1e-3) / 1e9, 1), "hbm_frac": round(nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 3)}


def run_batched(devs, args):
    """256 independent passes (W, H log-uniform in [64, 1024], 100-50k
    triangles, half textured) in one submission: shared launches, raster tile
    shape autotuned; against the same passes launched one launch set at a
    time."""
    import wgc
    from wgc import autotune, runner, scenes
    seed = scenes.SEED_BASE + 4
    specs = [scenes.batched_pass(seed, i, *a) for i, a in enumerate(scenes.batched_sizes(seed, 256))]
    if wgc.SEMANTICS_REFERENCE not in devs:
        devs[wgc.SEMANTICS_REFERENCE] = wgc.Device(0, wgc.SEMANTICS_REFERENCE)
    dev = devs[wgc.SEMANTICS_REFERENCE]
    pp = runner.PreparedPasses(dev, specs)
    pp.submit()
    dev.poll()
    frags = pp.total_fragments()
    tune = autotune.tune(pp, reps=3)
    best = dev.get_option(wgc.OPTION_RASTER_TILE)
    reps = max(3, args.steps // 20)
    ms_batched = autotune.time_submissions(pp, reps)
    prof = kernel_profile(dev, pp, reps=2)
    dev.set_option(wgc.OPTION_BATCHING, 0)
    ms_plain = autotune.time_submissions(pp, max(2, reps // 2), warmup=1)
    dev.set_option(wgc.OPTION_BATCHING, 1)
    alg = batched_algorithmic_bytes(specs)
    raster = prof.get("k_tiles")
    return {
        "workload": "256 passes, W,H log-uniform [64,1024], spheres of 100-50k triangles, half textured "
                    "(one shared 1024^2 texture), Rgba8Unorm + Depth32Float Less",
        "fragments_per_submission": frags,
        "ms_per_submission": round(ms_batched, 4),
        "Mfragments_per_s": round(frags / (ms_batched * 1e-3) / 1e6, 2),
        "plain_ms_per_submission": round(ms_plain, 4),
        "plain_Mfragments_per_s": round(frags / (ms_plain * 1e-3) / 1e6, 2),
        "speedup_vs_plain": round(ms_plain / ms_batched, 2),
        "algorithmic_GBps": round(alg / (ms_batched * 1e-3) / 1e9, 1),
        "hbm_frac_of_submission": round(alg / (ms_batched * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "tile_kernel_hbm_frac": round(alg / (raster["avg_us"] * 1e-6 * raster["launches"] / 2) / 1e9 / HBM_PEAK_GBS, 4)
        if raster else None,
        "autotune_ms": {wgc.RASTER_TILES[k].__repr__(): round(v, 4) for k, v in tune.items()},
        "raster_tile": wgc.RASTER_TILES[best],
        "kernels_us": {k: round(v["avg_us"], 2) for k, v in prof.items()},
    }


def run_serving(devs, args):
    """Poisson arrivals (default 2000 req/s) of requests of the batched
    workload's distribution against 32 resident meshes, targets and uniforms
    sub-allocated from a 64 GB heap (the HBM budget): continuous batching
    against plain per-request launches of the same trace; then the batched
    form at a rate beyond what the plain form sustains."""
    import wgc
    from wgc import scenes, serving
    if wgc.SEMANTICS_REFERENCE not in devs:
        devs[wgc.SEMANTICS_REFERENCE] = wgc.Device(0, wgc.SEMANTICS_REFERENCE)
    dev = devs[wgc.SEMANTICS_REFERENCE]
    dev.set_option(wgc.OPTION_RASTER_TILE, 0)
    pool = serving.MeshPool(dev, 32, seed=scenes.SEED_BASE + 5)
    budget = 64 << 30
    trace = serving.make_trace(scenes.SEED_BASE + 5, args.serve_rate, args.serve_seconds, len(pool.meshes))
    warm = serving.make_trace(scenes.SEED_BASE + 6, args.serve_rate, 0.5, len(pool.meshes))
    out = {"workload": f"Poisson {args.serve_rate:g} req/s for {args.serve_seconds:g} s, W,H log-uniform "
                       f"[64,1024], 32 resident spheres of 100-50k triangles (half textured), "
                       f"64 GB heap budget", "requests": len(trace)}
    for mode, batched in (("batched", True), ("plain", False)):
        srv = serving.Server(dev, pool, budget, max_batch=256, max_inflight=2)
        srv.run(warm, batched=batched)
        out[mode] = srv.run(trace, batched=batched).summary()
        srv.heap.release()
        print(f"bench.py: serving {mode} done: {out[mode]}", file=sys.stderr, flush=True)
    # saturation: 10x the rate for a shorter run
    hi = serving.make_trace(scenes.SEED_BASE + 7, args.serve_rate * 10, max(1.0, args.serve_seconds / 5),
                            len(pool.meshes))
    for mode, batched in (("batched_10x", True), ("plain_10x", False)):
        srv = serving.Server(dev, pool, budget, max_batch=256, max_inflight=2)
        out[mode] = srv.run(hi, batched=batched).summary()
        srv.heap.release()
        print(f"bench.py: serving {mode} done: {out[mode]}", file=sys.stderr, flush=True)
    return out


def batched_specs():
    from wgc import scenes
    seed = scenes.SEED_BASE + 4
    return [scenes.batched_pass(seed, i, *a) for i, a in enumerate(scenes.batched_sizes(seed, 256))]


def batched_algorithmic_bytes(specs) -> int:
    """Stores of every pass plus vertices and indices, and the shared 4 MiB
    texture once per submission."""
    from wgc import scenes
    return sum(scenes.algorithmic_bytes(s, count_texture=False) for s in specs) + 4 * 1024 * 1024


def profile_batch(n: int, raster_tile: int = 0):
    """The batched workload alone: n submissions with the given raster tile
    shape (wgc.RASTER_TILES), for kernel-trace and counter runs."""
    import wgc
    from wgc import runner
    specs = batched_specs()
    dev = wgc.Device(0, wgc.SEMANTICS_REFERENCE)
    dev.set_option(wgc.OPTION_RASTER_TILE, raster_tile)
    pp = runner.PreparedPasses(dev, specs)
    for _ in range(2):
        pp.submit()
    dev.poll()
    dev.timestamp(0)
    for _ in range(n):
        pp.submit()
    dev.timestamp(1)
    dev.poll()
    ms = dev.elapsed_ms(0, 1) / n
    print(json.dumps({"batched_submissions": n, "raster_tile": wgc.RASTER_TILES[raster_tile],
                      "ms_per_submission": round(ms, 4),
                      "fragments_per_submission": pp.total_fragments(),
                      "algorithmic_bytes": batched_algorithmic_bytes(specs)}), flush=True)


def _primitives(d) -> int:
    n = d.count
    per = {0: n, 1: n // 2, 2: max(n - 1, 0), 3: n // 3, 4: max(n - 2, 0)}[d.topology]
    return per * d.instance_count


def setup_preclears(spec) -> bool:
    """A single draw with LoadOp::Clear whose setup grid has a thread for every
    32 pixels has setup store the clear values over the whole target."""
    px = spec.width * spec.height
    clears = (spec.has_color and spec.clear_color is not None) or (spec.has_depth and spec.clear_depth is not None)
    prims = _primitives(spec.draws[0]) if len(spec.draws) == 1 else 0
    return bool(clears and prims > 0 and -(-prims // 256) * 256 * 32 >= px)


def kernel_bytes(spec, kernel: str) -> int:
    """Algorithmic bytes of one launch of `kernel` in the pass.

    k_setup: the vertex and index reads and the uniform, plus the whole
    target's colour + depth when setup stores the clear values.
    k_tiles / k_raster: the texture, and 4 B of colour and 4 B of depth for
    every pixel the kernel must store: the pixels the draw covers when setup
    stored the clear (counted on the GPU with fragment counting, on a device
    of its own), else the whole target."""
    import wgc
    from wgc import runner
    per_px = (4 if spec.has_color else 0) + (4 if spec.has_depth else 0)
    px = spec.width * spec.height
    pre = setup_preclears(spec)
    if kernel in ("k_setup", "k_clip"):
        b = sum(len(d.vertex_data) + (len(d.index_data) if d.index_data else 0) + (64 if d.uniform else 0)
                for d in spec.draws)
        return b + (px * per_px if pre and kernel == "k_setup" else 0)
    tex = sum(len(d.texture[2]) for d in spec.draws if d.texture)
    if not pre:
        return px * per_px + tex
    dev = wgc.Device(0, spec.semantics)
    r = runner.render(spec, dev, fragment_counts=True)
    covered = int((r.frag_count > 0).sum())
    del dev
    return covered * per_px + tex


def measured_traffic(workload: str, kernel: str):
    """HBM bytes per launch of `kernel` on `workload` from the committed
    rocprofv3 counter summary (calibrated FETCH_SIZE + WRITE_SIZE), or None."""
    path = ROOT / "profiles" / "hbm_traffic.txt"
    try:
        for line in path.read_text().splitlines():
            f = line.split()
            if len(f) >= 3 and not line.startswith("#") and f[0] == workload and f[1] == kernel:
                return int(f[2])
    except OSError:
        pass
    return None


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


if __name__ == "__main__":
    main()
