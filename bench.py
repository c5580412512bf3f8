#!/usr/bin/env python3
"""Benchmark of the render hot path on one MI355X.

A step renders the whole workload once with the FP32 batched wavefront
path tracer (pixel loop of render::render, src/render.rs:1587-1747) with
inputs (scenes, BVHs) resident in HBM and framebuffers left in HBM.  The
default workload is the 64-scene batch: seeded displaced geodesic spheres
with 1k-200k triangles plus random spheres, 128^2-512^2, 16/64/256 spp,
depth 8, NEE on 1-4 quad lights -- all scenes in one wavefront.

Prints one JSON line: Mpaths/s (paths = sum of W*H*ceil(sqrt(spp))^2), the
plain FP64 one-path-per-thread GPU kernel on the same scenes and the
relative RMSE between the two images, the roofline of the dominant kernel
(extend: BVH traversal), the CPU oracle timed on a bounded sample, and the
inverse-rendering leg: adjoint gradient evaluations per second (a 32-spp
forward render plus a 32-spp path-replay adjoint at 512^2).
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mpaths/s FP32-batched vs plain GPU kernel at RMSE≤1e-3; adjoint grads/s"
HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def measured_traffic(workload, kernel="k_extend"):
    """HBM-side bytes per launch of `kernel` on `workload`, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (scaled by the copy-kernel
    calibration recorded there), with the summary's path and its record (L2
    hit rate, VALU lane use), or Nones when no summary covers it: PMC
    counters need a profiler run of their own and cannot be read here."""
    import glob
    import re
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", f"{kernel}_*.md"))):
        m = re.search(r"<!-- traffic (\{.*\}) -->", open(path).read())
        if not m:
            continue
        rec = json.loads(m.group(1))
        if rec.get("workload") == workload and rec.get("kernel") == kernel:
            return rec["bytes_per_launch"], os.path.relpath(path, ROOT), rec
    return None, None, None


def build_workload(name, n_scenes=64):
    """Returns (workload name, [(config, world)], description)."""
    import luminara_amd as lm
    from luminara_amd import workloads
    if name == "batch64":
        scenes, tris = [], 0
        for i in range(n_scenes):
            toml, add_mesh, info = workloads.batch_scene(i, seed=42)
            cfg, _, world = lm.load_scene(toml)
            tris += add_mesh(world)
            scenes.append((cfg, world))
        return (f"batch{n_scenes}_displaced_spheres_1k-200k_tris_depth8_nee", scenes,
                {"scenes": n_scenes, "triangles": tris})
    if name == "random_spheres":
        cfg, _, world = lm.load_scene(workloads.random_spheres(1920, 1080, 128, 50, 42))
        return "random_spheres_1920x1080_spp128_d50", [(cfg, world)], {"scenes": 1}
    if name == "mesh_heavy":
        toml, add_mesh = workloads.mesh_heavy()
        cfg, _, world = lm.load_scene(toml)
        tris = add_mesh(world)
        return "mesh_heavy_1920x1080_spp64_d12_1.31M_tris", [(cfg, world)], {"scenes": 1, "triangles": tris}
    if name == "cornell":
        cfg, _, world = lm.load_scene(workloads.cornell_box(256, 256, 64, 8, 42))
        return "cornell_256x256_spp64_d8", [(cfg, world)], {"scenes": 1}
    raise SystemExit(f"unknown workload {name}")


def cpu_baseline(scenes, seconds, threads):
    """Times the CPU FP64 oracle on evenly spaced rows of every scene (every
    k-th row, the same k for all, so the sample spans sky, objects and
    ground like the whole frame), sized to about `seconds` of pixel-loop time."""
    import oracle
    import luminara_amd as lm

    def run(step):
        paths, secs = 0, 0.0
        for i, (cfg, world) in enumerate(scenes):
            h, _ = lm.output_shape(cfg)
            _, _, st = oracle.render(world, cfg, rows=(step // 2 % h, h), row_step=step, threads=threads,
                                     want_var=False, scene_key=i)
            paths += st["camera_rays"]
            secs += st["render_secs"]
        return paths, secs

    total = 0
    for cfg, _ in scenes:
        h, w = lm.output_shape(cfg)
        total += h * w * math.ceil(math.sqrt(cfg.samples_per_pixel)) ** 2
    step = 256
    for _ in range(4):  # calibrate: thread start-up dominates the smallest samples
        paths, secs = run(step)
        if secs >= 0.7 * seconds or step == 1:
            break
        frac = min(1.0, seconds * paths / max(secs, 1e-6) / total)
        step = max(1, min(step - 1, int(round(1.0 / max(frac, 1e-9)))))
    return paths, secs, 1.0 / step


def adjoint_leg(steps, warmup, res=512, spp=32, tex_res=256, n_spheres=32):
    """Inverse-rendering step on the Cornell geometry: a 32-spp FP32 forward
    render, dL/dI of the L2 loss against a target rendered with hidden
    parameters, and a 32-spp FP32 path-replay adjoint for the back wall's
    albedo texels, the spheres' GGX roughness and the light's emission."""
    import torch
    import luminara_amd as lm
    from luminara_amd import _abi, workloads
    toml, setup = workloads.inverse_rendering_scene(res, res, spp, 8, 42, tex_res=tex_res, n_spheres=n_spheres)
    cfg, _, w = lm.load_scene(toml)
    info = setup(w)
    ttoml, tsetup = workloads.inverse_rendering_scene(res, res, spp, 8, 42, tex_res=tex_res, n_spheres=n_spheres,
                                                      hidden=True)
    tcfg, _, tw = lm.load_scene(ttoml)
    tsetup(tw)
    shape = lm.output_shape(cfg) + (3,)
    target = torch.empty(shape, dtype=torch.float32, device="cuda")
    tp = lm.make_params(tcfg, path="plain64")
    lm.render_device(tw, tp, target.data_ptr())
    img = torch.empty_like(target)
    dl = torch.empty_like(target)
    n_mat = w.desc_struct.n_materials
    gtex = torch.zeros(tex_res * tex_res * 3, dtype=torch.float32, device="cuda")
    grough = torch.zeros(n_mat, dtype=torch.float32, device="cuda")
    p = lm.make_params(cfg, path="wavefront")
    spec = _abi.AdjointSpec(info["texture"], info["emission_material"], 0, 0)
    n_params = tex_res * tex_res * 3 + len(info["sphere_materials"]) + 1

    def step():
        st = lm.render_device(w, p, img.data_ptr())
        torch.sub(img, target, out=dl)
        dl.mul_(2.0)
        ad = lm.adjoint_device(w, p, spec, dl.data_ptr(), gtex.data_ptr(), grough.data_ptr())
        return st, ad

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    fwd_ms = adj_ms = 0.0
    t0 = time.perf_counter()
    for _ in range(steps):
        st, ad = step()
        fwd_ms += st["render_ms"]
        adj_ms += ad["adjoint_ms"]
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    loss = float(((img - target) ** 2).sum())
    paths = ad["paths"]
    return {"workload": f"cornell_inverse_{res}x{res}_spp{spp}_tex{tex_res}_spheres{n_spheres}_depth8",
            "grads_per_s": steps / el, "unit": "full gradient evaluations (forward + adjoint) per second",
            "parameters": n_params, "parameter_grads_per_s": n_params * steps / el,
            "ms_per_step": el * 1e3 / steps, "forward_ms": fwd_ms / steps, "adjoint_ms": adj_ms / steps,
            "adjoint_mpaths_per_s": paths / (adj_ms / steps / 1e3) / 1e6, "loss": loss, "dtype": "f32"}


def image_leg(scenes, worlds, frame, n_pix, reps):
    """The image outputs of the batch, timed with HIP events on the stream
    their kernels run on: the default post chain (exposure, ACES, sRGB, 8-bit;
    12 B read + 4 B written per pixel) over every framebuffer of the batch in
    one launch, the Radiance RGBE pixels of the exposed HDR (12 B read + 4 B
    written), and the AOV passes (one pixel-centre ray per pixel, per scene)."""
    import torch
    import luminara_amd as lm
    s = torch.cuda.current_stream()
    rgba = torch.empty(n_pix * 4, dtype=torch.uint8, device="cuda")
    hx = torch.empty(n_pix * 3, dtype=torch.float32, device="cuda")
    rgbe = torch.empty(n_pix * 4, dtype=torch.uint8, device="cuda")
    cfg = lm.RenderConfig()

    def timed(fn):
        fn()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(s)
        for _ in range(reps):
            fn()
        e1.record(s)
        e1.synchronize()
        return e0.elapsed_time(e1) / reps

    post_ms = timed(lambda: lm.post_device(frame.data_ptr(), n_pix, 1, cfg, 0, rgba.data_ptr(), 0, s.cuda_stream))
    lm.post_device(frame.data_ptr(), n_pix, 1, cfg, 0, rgba.data_ptr(), hx.data_ptr(), s.cuda_stream)
    rgbe_ms = timed(lambda: lm.encode_rgbe_device(hx.data_ptr(), n_pix, rgbe.data_ptr(), s.cuda_stream))
    bufs = []
    for c, _ in scenes:
        h, w = lm.output_shape(c)
        bufs.append((torch.empty((h, w), dtype=torch.float32, device="cuda"),
                     torch.empty((h, w, 3), dtype=torch.uint8, device="cuda"),
                     torch.empty((h, w, 3), dtype=torch.uint8, device="cuda")))
    aparams = [lm.make_params(c, scene_key=i, stream=s.cuda_stream) for i, (c, _) in enumerate(scenes)]

    def aov_all():
        for w, p, (d, n, a) in zip(worlds, aparams, bufs):
            lm.aov_device(w, p, d.data_ptr(), n.data_ptr(), a.data_ptr())
    aov_ms = timed(aov_all)
    hbm = 8.0e12
    return {"pixels": n_pix,
            "post_chain": {"kernel": "k_post", "ms": post_ms, "bytes_per_pixel": 16,
                           "achieved_gb_s": n_pix * 16 / (post_ms / 1e3) / 1e9,
                           "frac_of_hbm_peak": n_pix * 16 / (post_ms / 1e3) / hbm},
            "rgbe": {"kernel": "k_rgbe", "ms": rgbe_ms, "bytes_per_pixel": 16,
                     "achieved_gb_s": n_pix * 16 / (rgbe_ms / 1e3) / 1e9,
                     "frac_of_hbm_peak": n_pix * 16 / (rgbe_ms / 1e3) / hbm},
            "aov": {"kernel": "k_aov", "ms": aov_ms, "launches": len(scenes),
                    "mrays_per_s": n_pix / (aov_ms / 1e3) / 1e6}}


def ingest_leg(freq=256):
    """Scene ingestion of the mesh-heavy scene's 1.31M-triangle displaced
    icosphere through the TOML `[[mesh]]` path: one ASCII PLY file, loaded
    from its text (the sidecar written on the way), loaded again from the
    binary sidecar, and the scene built (SAH BVH, 4-wide collapse, upload)."""
    import shutil
    import tempfile
    import luminara_amd as lm
    from luminara_amd import workloads
    v, f = workloads.displaced_sphere(freq, [0.0, 1.75, 0.0], 1.6, 42)
    tmp = tempfile.mkdtemp(prefix="lmn_ingest_")
    try:
        path = os.path.join(tmp, "icosphere.ply")
        workloads.write_ply(path, v, f)
        toml = (f'[[mesh]]\nfile = "{path}"\n'
                'material = { type = "pbr", color = [0.7, 0.7, 0.7], roughness = 0.3 }\n')
        old = os.environ.get("LMN_MESH_SIDECAR")
        os.environ["LMN_MESH_SIDECAR"] = "auto"
        try:
            t0 = time.perf_counter()
            _, _, w_text = lm.load_scene(toml)
            t_text = time.perf_counter() - t0
            t0 = time.perf_counter()
            _, _, w_side = lm.load_scene(toml)
            t_side = time.perf_counter() - t0
        finally:
            if old is None:
                del os.environ["LMN_MESH_SIDECAR"]
            else:
                os.environ["LMN_MESH_SIDECAR"] = old
        n_tri = len(w_side.prims)
        t0 = time.perf_counter()
        w_side.device_scene()
        t_build = time.perf_counter() - t0
        res = {"triangles": n_tri, "vertices": len(v), "ply_bytes": os.path.getsize(path),
               "sidecar_bytes": os.path.getsize(path + ".lmnmesh"),
               "text_load_s": t_text, "sidecar_load_s": t_side, "scene_create_s": t_build,
               "build_threads": int(os.environ.get("LMN_BUILD_THREADS", "0")) or min(64, os.cpu_count() or 1)}
        del w_text, w_side
        return res
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def host_cores():
    """Cores this process may run on: its CPU affinity, capped by a cgroup v2
    CPU quota when one is set (a container's share of the host) and by
    OMP_NUM_THREADS when the job's scheduler sets it."""
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count() or 1
    try:
        quota, period = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if quota != "max":
            n = min(n, max(1, int(float(quota) / float(period))))
    except (OSError, ValueError):
        pass
    try:  # a job's CPU share set by its scheduler (the thread-pool convention)
        n = min(n, max(1, int(os.environ["OMP_NUM_THREADS"])))
    except (KeyError, ValueError):
        pass
    return n


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def counter_limiter(rec, hbm_frac):
    """What holds the kernel below its HBM roofline, read off the committed
    PMC counters: HBM-side traffic near peak means bandwidth; otherwise VALU
    instruction issue and the texture-data (load return) unit's busy share
    tell an issue-bound kernel from a latency-bound one, and the VALU lane
    use how much of that issue divergence wastes."""
    if hbm_frac is None:
        return None
    if hbm_frac >= 0.6:
        return "hbm bandwidth"
    rec = rec or {}
    lanes, issue, td = rec.get("valu_lane_util"), rec.get("valu_issue"), rec.get("td_busy")
    parts = [f"HBM-side {hbm_frac:.1%} of peak"]
    if issue is not None:
        parts.append(f"VALU issue {issue:.0%} of cycles")
    if td is not None:
        parts.append(f"load return (TD) {td:.0%} busy")
    if lanes is not None:
        parts.append(f"VALU lanes {lanes:.0%} active")
    busy = [n for n, v in (("VALU issue", issue), ("load return", td)) if v is not None and v >= 0.75]
    if busy:
        return " and ".join(busy) + " with divergent lanes (" + ", ".join(parts) + ")"
    if lanes is not None and lanes < 0.5:
        return "latency and lane divergence (" + ", ".join(parts) + ")"
    return "latency (" + ", ".join(parts) + ")"


def measure(wl_name, scenes, desc, args, want_images, tag):
    """Times one workload: the FP32 batched wavefront (graph launches; a
    second region with stage events), the extend kernel's roofline, the plain
    FP64 kernel's builds and the FP32-vs-FP64 error."""
    import numpy as np
    import torch
    import luminara_amd as lm
    from luminara_amd import _abi

    t0 = time.time()
    worlds = [w for _, w in scenes]
    nodes = 0
    for w in worlds:
        nodes += w.bvh_info()["nodes"]
    log(f"[bench:{tag}] {wl_name}: {desc}, {nodes} BVH nodes, scenes built and uploaded in {time.time() - t0:.1f}s")
    # every scene's framebuffer is a view into one device buffer, so the post
    # chain can run over the whole batch's pixels in one launch
    shapes = [lm.output_shape(c) for c, _ in scenes]
    n_pix = sum(h * w for h, w in shapes)
    frame = torch.empty(n_pix * 3, dtype=torch.float32, device="cuda")
    outs, off = [], 0
    for h, w in shapes:
        outs.append(frame[off:off + h * w * 3].view(h, w, 3))
        off += h * w * 3
    ptrs = [o.data_ptr() for o in outs]

    def params(path, flags=0):
        return [lm.make_params(c, path=path, scene_key=i, pool_paths=args.pool, stack_lds=args.stack_lds,
                               block_size=args.block, grid_blocks=args.grid, flags=flags)
                for i, (c, _) in enumerate(scenes)]

    # launch configuration (path pool, workgroup size, LDS stack depth, grid)
    # chosen by the library's autotuner on probe renders of this batch; the
    # timed renders below leave those knobs at 0 and take the tuned values
    tuning = None
    if not (args.pool or args.stack_lds or args.block or args.grid or args.no_tune):
        tuning = lm.autotune(worlds, params("wavefront"), ptrs)
        log(f"[bench:{tag}] autotune: {tuning}")
    # Timed region 1 (`value`): the product path, blocks of iterations
    # replayed as hipGraphs.  Timed region 2: the same renders launched
    # directly with HIP events between the stages (events recorded by graph
    # nodes cannot be timed), for the per-stage times and k_extend's average
    # launch duration on its stream.
    p_graph = params("wavefront")
    p32 = params("wavefront", _abi.FLAG_STAGE_TIMING)
    for _ in range(args.warmup):
        lm.render_batch_device(worlds, p_graph, ptrs)
        lm.render_batch_device(worlds, p32, ptrs)
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for _ in range(args.steps):
        lm.render_batch_device(worlds, p_graph, ptrs)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    stats = []
    t_start = time.perf_counter()
    for _ in range(args.steps):
        stats.append(lm.render_batch_device(worlds, p32, ptrs))
    torch.cuda.synchronize()
    elapsed_direct = time.perf_counter() - t_start
    paths = stats[0]["paths"]
    ms_per_step = elapsed * 1e3 / args.steps
    value = paths * args.steps / elapsed / 1e6
    value_direct = paths * args.steps / elapsed_direct / 1e6
    imgs32 = [o.double().cpu().numpy() for o in outs]
    log(f"[bench:{tag}] wavefront FP32: {ms_per_step:.1f} ms/step, {value:.1f} Mpaths/s (graph launches); "
        f"{value_direct:.1f} Mpaths/s launched directly with stage events; stats {stats[-1]}")

    ext_ms = sum(s["extend_ms"] for s in stats) / args.steps
    launches = sum(s["extend_launches"] for s in stats) / args.steps
    roofline = None
    if not args.no_count:
        # algorithmic bytes of the extend kernel: node and primitive fetches
        # counted by an instrumented run of the same traversal on the same
        # work, plus the kernel's own reads and writes per ray; time from HIP
        # events on its stream
        counts = lm.profile_extend(worlds, params("wavefront"), ptrs)
        rays, hits = counts["rays"], counts["hits"]
        # a stored (bounce) ray reads its origin and direction quarters (32 B;
        # a camera ray is generated in the kernel from the launch tables, one
        # per path); a hit writes the 32-B surface record, a miss its 16-B half
        io_bytes = 32 * max(rays - paths, 0) + 32 * hits + 16 * (rays - hits)
        alg_bytes = counts["node_fetches"] * counts["node_bytes"] + counts["prim_tests"] * counts["prim_bytes"] + io_bytes
        bytes_per_launch = alg_bytes / launches
        avg_launch_s = ext_ms / launches / 1e3
        achieved = bytes_per_launch / avg_launch_s
        traffic, traffic_src, rec = measured_traffic(wl_name)
        hbm_side = None
        if traffic:
            hbm_side = {"achieved": traffic / avg_launch_s / 1e9, "frac": traffic / avg_launch_s / HBM_PEAK,
                        "bytes_per_launch": traffic, "source": traffic_src}
        roofline = {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                    "frac": achieved / HBM_PEAK, "traffic": traffic, "traffic_unit": "bytes per launch",
                    "traffic_source": traffic_src, "l2_hit_rate": rec.get("l2_hit_rate") if rec else None,
                    "hbm_side": hbm_side,
                    "limiter": counter_limiter(rec, hbm_side["frac"] if hbm_side else None),
                    "valu_lane_util": rec.get("valu_lane_util") if rec else None,
                    "kernel": "k_extend", "avg_launch_us": avg_launch_s * 1e6, "launches_per_step": launches,
                    "bytes_per_launch": bytes_per_launch,
                    "bytes_model": f"{counts['node_bytes']} B per 4-wide BVH node fetched + 64 B per primitive record tested "
                                   "+ 32 B read per stored (bounce) ray + 32 B (hit) or 16 B (miss) record written",
                    "per_ray": {"node_fetches": counts["node_fetches"] / rays, "prim_tests": counts["prim_tests"] / rays,
                                "hit_fraction": hits / rays}}
        log(f"[bench:{tag}] extend: {counts}, {ext_ms:.2f} ms/step over {launches:.0f} launches, "
            f"{achieved / 1e9:.0f} GB/s algorithmic; HBM side {hbm_side}")

    res = {"workload": wl_name, "value": value, "unit": "Mpaths/s", "ms_per_step": ms_per_step,
           "paths_per_step": paths, "bvh_nodes": nodes, "launch": tuning, "roofline": roofline,
           "value_direct_launch": value_direct, "max_depth": scenes[0][0].max_depth,
           "mrays_per_s": (stats[-1]["camera_rays"] + stats[-1]["extension_rays"] + stats[-1]["shadow_rays"])
                          / (ms_per_step / 1e3) / 1e6,
           "stage_ms": {"extend": ext_ms, "shade": sum(s["shade_ms"] for s in stats) / args.steps,
                        "shadow": sum(s["shadow_ms"] for s in stats) / args.steps,
                        "regen": sum(s["regen_ms"] for s in stats) / args.steps,
                        "marched_kinds": sum(s["march_ms"] for s in stats) / args.steps}}
    res.update(desc)
    if not args.no_fp64:
        # the plain one-path-per-thread FP64 kernel, every path its own
        # thread; built for 4, 3 and 2 waves per SIMD (register budgets),
        # each warmed up, then timed twice: the fastest build is the
        # denominator.  All builds give the same image.
        outs64 = [torch.empty_like(o) for o in outs]
        p64 = [o.data_ptr() for o in outs64]
        builds = {}
        for waves in (4, 3, 2):
            pp = params("plain64", waves << _abi.FLAG_PLAIN_WAVES_SHIFT)
            lm.render_batch_device(worlds, pp, p64)
            torch.cuda.synchronize()
            best = math.inf
            for _ in range(2):
                t1 = time.perf_counter()
                lm.render_batch_device(worlds, pp, p64)
                torch.cuda.synchronize()
                best = min(best, time.perf_counter() - t1)
            builds[waves] = paths / best / 1e6
        fp64_rate = max(builds.values())
        imgs64 = [o.double().cpu().numpy() for o in outs64]
        num = sum(float(np.sum((a - b) ** 2)) for a, b in zip(imgs32, imgs64))
        den = sum(float(np.sum(b ** 2)) for b in imgs64)
        per = [float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)) for a, b in zip(imgs32, imgs64)]
        rel = (num / den) ** 0.5
        res.update({"plain_fp64_mpaths_per_s": fp64_rate,
                    "plain_fp64_builds": {f"{w}_waves_per_simd": v for w, v in builds.items()},
                    "speedup_vs_plain_fp64": value / fp64_rate, "relrmse_fp32_vs_fp64": rel,
                    "relrmse_fp32_vs_fp64_max_scene": max(per),
                    "rmse_gate_1e-3": {"aggregate": rel <= 1e-3, "every_scene": max(per) <= 1e-3}})
        if len(per) > 1:
            res["relrmse_per_scene"] = [float(f"{x:.3g}") for x in per]
        log(f"[bench:{tag}] plain FP64 builds {builds} Mpaths/s; relRMSE(FP32, FP64) = {rel:.3e} "
            f"(worst scene {max(per):.3e})")
    if want_images and not args.no_images:
        images = image_leg(scenes, worlds, frame, n_pix, max(args.steps, 3) * 10)
        log(f"[bench:{tag}] image outputs: {images}")
        res["image_outputs"] = images
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="batch64")
    ap.add_argument("--scenes", type=int, default=64)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--no-fp64", action="store_true", help="skip the plain FP64 kernel comparison")
    ap.add_argument("--no-count", action="store_true", help="skip the instrumented traversal count")
    ap.add_argument("--cpu-seconds", type=float, default=16.0)
    ap.add_argument("--stack-lds", type=int, default=0)
    ap.add_argument("--pool", type=int, default=0, help="path-pool size (with any launch knob: no autotuning)")
    ap.add_argument("--block", type=int, default=0, help="workgroup size of the traversal kernels (128 or 256)")
    ap.add_argument("--grid", type=int, default=0, help="workgroups of the grid-stride kernels")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the inverse-rendering adjoint leg")
    ap.add_argument("--no-tune", action="store_true", help="skip the autotuner (library defaults or cache)")
    ap.add_argument("--no-images", action="store_true", help="skip the post chain / RGBE / AOV timings")
    ap.add_argument("--no-mesh-heavy", action="store_true",
                    help="skip the second leg: the 1.31M-triangle mesh-heavy scene (1080p, 64 spp, depth 12)")
    ap.add_argument("--no-ingest", action="store_true",
                    help="skip timing the 1.31M-triangle PLY load (text, binary sidecar) and scene build")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB (5 times): the reference kernel that "
                         "scales the FETCH_SIZE / WRITE_SIZE counters to bytes")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = args.copy_calibration << 18  # float32 elements
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4, "launches": 5}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU only: every GPU box here has a single "
                                   "MI355X; run with --gpus 1"}))
        sys.exit(2)

    # torch before the library: liblmn.so then binds to the HIP runtime torch
    # loaded (one runtime per process), not a second copy of it
    import torch  # noqa: F401
    wl_name, scenes, desc = build_workload(args.workload, args.scenes)
    main_res = measure(wl_name, scenes, desc, args, True, args.workload)

    cpu = None
    if not args.no_cpu:
        threads = host_cores()
        cpaths, csecs, frac = cpu_baseline(scenes, args.cpu_seconds, threads)
        cpu = {"value": cpaths / csecs / 1e6, "unit": "Mpaths/s", "cores": threads, "kind": "port",
               "cpu_model": cpu_model(), "cores_visible": os.cpu_count(),
               "sample": f"every k-th row of every scene, k = {round(1 / frac)} ({cpaths} paths, {csecs:.1f} s), "
                         f"CPU FP64 oracle (restatement of the reference pixel loop), counter RNG, one thread per "
                         f"core this job may use (affinity, cgroup quota, OMP_NUM_THREADS)"}
        log(f"[bench] cpu oracle: {cpu}")
    del scenes

    mesh = None
    if args.workload != "mesh_heavy" and not args.no_mesh_heavy:
        mname, mscenes, mdesc = build_workload("mesh_heavy")
        mesh = measure(mname, mscenes, mdesc, args, False, "mesh_heavy")
        del mscenes

    if mesh is not None and not args.no_ingest:
        mesh["ingest"] = ingest_leg()
        log(f"[bench] ingest: {mesh['ingest']}")

    adj = None
    if not args.no_adjoint:
        adj = adjoint_leg(max(args.steps, 3), 1)
        log(f"[bench] adjoint: {adj}")

    r = main_res
    line = {
        "metric": METRIC, "value": r["value"], "unit": "Mpaths/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": r["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": dict({"workload": wl_name, "paths_per_step": r["paths_per_step"], "bvh_nodes": r["bvh_nodes"],
                        "max_depth": r["max_depth"], "launch": r["launch"]}, **desc),
        "roofline": r["roofline"], "cpu_baseline": cpu, "adjoint": adj,
    }
    for k, v in r.items():
        if k not in ("workload", "value", "unit", "ms_per_step", "paths_per_step", "bvh_nodes", "launch",
                     "roofline", "max_depth") and k not in desc:
            line[k] = v
    if mesh is not None:
        line["mesh_heavy"] = mesh
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
