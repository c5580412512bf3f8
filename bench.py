#!/usr/bin/env python3
"""Benchmark of the render hot path on one MI355X.

A step is one frame of the workload rendered by the FP32 wavefront path
tracer (pixel loop of render::render, src/render.rs:1587-1747) with inputs
(scene, BVH) resident in HBM and the framebuffer left in HBM.  Prints one
JSON line: Mpaths/s (paths = W*H*ceil(sqrt(spp))^2), the plain FP64
one-path-per-thread GPU kernel on the same frame, the image error between
them, the roofline of the dominant kernel (extend: BVH traversal) and the
CPU oracle timed on a row band of the same frame.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mpaths/s FP32-batched vs plain GPU kernel at RMSE≤1e-3; adjoint grads/s"
HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def workload_scene(name):
    from luminara_amd import workloads
    if name == "random_spheres":
        return "random_spheres_1920x1080_spp128_d50", workloads.random_spheres(1920, 1080, 128, 50, 42)
    if name == "cornell":
        return "cornell_256x256_spp64_d8", workloads.cornell_box(256, 256, 64, 8, 42)
    raise SystemExit(f"unknown workload {name}")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="random_spheres")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--no-fp64", action="store_true", help="skip the plain FP64 kernel comparison")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--stack-lds", type=int, default=0)
    ap.add_argument("--pool", type=int, default=0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU only: every GPU box here has a single "
                                   "MI355X; run with --gpus 1"}))
        sys.exit(2)

    import numpy as np
    import torch
    import luminara_amd as lm
    from luminara_amd import _abi

    wl_name, toml = workload_scene(args.workload)
    cfg, cam, world = lm.load_scene(toml)
    t0 = time.time()
    bvh = world.bvh_info()
    log(f"[bench] {wl_name}: {len(world.prims)} prims, BVH {bvh} built+uploaded in {time.time() - t0:.2f}s")
    h, w = lm.output_shape(cfg)
    out = torch.empty((h, w, 3), dtype=torch.float32, device="cuda")
    params = lm.make_params(cfg, path="wavefront", pool_paths=args.pool, stack_lds=args.stack_lds,
                            flags=_abi.FLAG_STAGE_TIMING)

    for _ in range(args.warmup):
        lm.render_device(world, params, out.data_ptr())
    torch.cuda.synchronize()
    stats = []
    t_start = time.perf_counter()
    for _ in range(args.steps):
        stats.append(lm.render_device(world, params, out.data_ptr()))
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    paths = stats[0]["paths"]
    ms_per_step = elapsed * 1e3 / args.steps
    value = paths * args.steps / elapsed / 1e6
    img32 = out.double().cpu().numpy()
    log(f"[bench] wavefront FP32: {ms_per_step:.1f} ms/frame, {value:.1f} Mpaths/s, stats {stats[-1]}")

    # roofline of the extend kernel: algorithmic bytes from an instrumented
    # run of the same traversal on the same frame, time from HIP events
    counts = lm.profile_extend(world, lm.make_params(cfg, path="wavefront", pool_paths=args.pool,
                                                     stack_lds=args.stack_lds), out.data_ptr())
    ext_ms = sum(s["extend_ms"] for s in stats) / args.steps
    launches = sum(s["extend_launches"] for s in stats) / args.steps
    io_bytes = 48  # per ray: origin+beta, direction+beta (2 x 16 B) read, hit record 16 B written
    alg_bytes = (counts["node_fetches"] * counts["node_bytes"] + counts["prim_tests"] * counts["prim_bytes"]
                 + counts["rays"] * io_bytes)
    bytes_per_launch = alg_bytes / launches
    avg_launch_s = ext_ms / launches / 1e3
    achieved = bytes_per_launch / avg_launch_s
    roofline = {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                "frac": achieved / HBM_PEAK, "traffic": None, "kernel": "k_extend",
                "avg_launch_us": avg_launch_s * 1e6, "launches_per_step": launches,
                "bytes_per_launch": bytes_per_launch,
                "per_ray": {"node_fetches": counts["node_fetches"] / counts["rays"],
                            "prim_tests": counts["prim_tests"] / counts["rays"]}}
    log(f"[bench] extend: {counts}, {ext_ms:.2f} ms/frame over {launches:.0f} launches")

    extra = {}
    if not args.no_fp64:
        p64 = lm.make_params(cfg, path="plain64")
        out64 = torch.empty_like(out)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        s64 = lm.render_device(world, p64, out64.data_ptr())
        torch.cuda.synchronize()
        t64 = time.perf_counter() - t1
        img64 = out64.double().cpu().numpy()
        rel = float(np.linalg.norm(img32 - img64) / max(np.linalg.norm(img64), 1e-300))
        fp64_rate = paths / t64 / 1e6
        extra = {"plain_fp64_mpaths_per_s": fp64_rate, "speedup_vs_plain_fp64": value / fp64_rate,
                 "relrmse_fp32_vs_fp64": rel}
        log(f"[bench] plain FP64: {t64 * 1e3:.1f} ms, relRMSE(FP32 wavefront, FP64) = {rel:.3e}")

    cpu = None
    if not args.no_cpu:
        import oracle
        threads = min(16, os.cpu_count() or 1)
        band = 2
        _, _, st = oracle.render(world, cfg, rows=(0, band), threads=threads, want_var=False)
        per_row = st["render_secs"] / band
        band = int(max(1, min(h, args.cpu_seconds / max(per_row, 1e-6))))
        _, _, st = oracle.render(world, cfg, rows=(0, band), threads=threads, want_var=False)
        cpu_paths = st["camera_rays"]
        cpu = {"value": cpu_paths / st["render_secs"] / 1e6, "unit": "Mpaths/s", "cores": st["threads"],
               "kind": "port",
               "sample": f"rows 0..{band} of the same {w}x{h} frame ({cpu_paths} paths, "
                         f"{st['render_secs']:.1f} s), CPU FP64 oracle, counter RNG"}
        log(f"[bench] cpu oracle: {cpu}")

    line = {
        "metric": METRIC, "value": value, "unit": "Mpaths/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": wl_name, "width": cfg.width, "height": cfg.height,
                   "spp": cfg.samples_per_pixel, "max_depth": cfg.max_depth, "paths_per_step": paths,
                   "prims": len(world.prims), "bvh_nodes": bvh["nodes"]},
        "roofline": roofline, "cpu_baseline": cpu,
        "mrays_per_s": (stats[-1]["camera_rays"] + stats[-1]["extension_rays"] + stats[-1]["shadow_rays"])
                       / (ms_per_step / 1e3) / 1e6,
    }
    line.update(extra)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
