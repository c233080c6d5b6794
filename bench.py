"""Benchmark: Msamples/s of the MI355X wavefront path tracer on the Disney-grid
configuration (8x8 level-6 icospheres = 5,242,882 triangles on a ground quad,
2048x1024 sky+sun envmap, 512x512, 256 spp, max depth 12), plain path_mis.

A step renders the configuration's whole job once: spp_per_step (default 256)
samples per pixel of the 512x512 view in one persistent render call; sample
offsets advance from step to step, so every step draws new samples.
Inputs are resident in HBM before the timed region; the timed region is K
render calls, each ending in a stream synchronisation.  After it, a second leg
measures the GPU time to relMSE <= 1e-3 (validation_tests/compare_exr.py:102)
of plain and guided rendering against a high-spp plain reference.

Prints one JSON line (see the keys below).  --gpus N > 1 is not supported: this
build runs on one GPU.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "pbrt-renderer-_amd"))

METRIC = "Msamples/s; time-to-relMSE≤1e-3, guided vs plain, on 1 MI355X"
HBM_PEAK_GBS = 8000.0
NODE_BYTES, TRI_BYTES = 80, 48
# per extension ray besides the traversal: ray read 32 B, hit write 16 B, queue index 4 B
EXT_IO_BYTES = 32 + 16 + 4
# per shadow ray besides the traversal: origin/direction read 32 B
SHADOW_IO_BYTES = 32


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(scene, cam, spp, depth):
    """The CPU oracle (path_mis restated on the CPU, Nori BVH2) on a bounded sample."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import orc
    threads = min(os.cpu_count() or 1, 16)
    t0 = time.time()
    o = orc.OracleScene(scene)
    log(f"[bench] oracle scene (Nori BVH2) built in {time.time() - t0:.1f}s")
    t0 = time.time()
    _, stats = o.render(cam, spp, depth, seed=1, nthreads=threads)
    dt = time.time() - t0
    return {"value": float(stats[0]) / dt / 1e6, "unit": "Msamples/s", "cores": threads, "kind": "port",
            "sample": f"{cam.width}x{cam.height} at {spp} spp (first {spp} samples of every pixel), "
                      f"{int(stats[0])} samples in {dt:.2f}s"}


def traffic_from_profiles():
    """HBM bytes per extension ray of the closest-hit kernel from the committed rocprof counter summary
    (its line `hbm_bytes_per_ray: N`: FETCH_SIZE + WRITE_SIZE per launch, scaled by the copy calibration,
    divided by the rays of those launches)."""
    path = os.path.join(ROOT, "profiles", "k_extend_disney_grid_traffic.md")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith("hbm_bytes_per_ray:"):
                return float(line.split(":")[1])
    return None


def rel_mse(img, ref):
    """mean over pixels and channels of (img - ref)^2 / (ref^2 + 1e-2)
    (validation_tests/compare_exr.py:102)."""
    return float(np.mean((img - ref) ** 2 / (ref * ref + 1e-2)))


def calibrate_copy(nbytes):
    """A plain device copy of known size, for scaling the FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": nbytes, "launches": 3}), flush=True)


def time_to_rel_mse(gpu, ng, cam, depth, ref, start_spp, growth, max_spp, thresh, guided, seed):
    """Progressive render, checking relMSE against `ref` after 1, growth, growth^2, ...
    times start_spp samples per pixel, until relMSE <= thresh.  Returns the GPU
    milliseconds of the renders (every kernel, guiding inference and training
    included), the spp reached and the convergence curve."""
    gpu.set_option("guiding", 1 if guided else 0)
    if guided:
        gpu.set_option("guide_reset", 1)
    gpu.reset_film()
    ms, spp, curve, e = 0.0, 0, [], float("inf")
    target = start_spp
    while spp < max_spp:
        chunk = min(max(target - spp, 1), max_spp - spp)
        img = gpu.render(ng.make_view(cam, chunk, depth, sample_offset=spp), seed=seed)
        ms += gpu.stats().ms_total
        spp += chunk
        e = rel_mse(img[:, :3] / img[:, 3:4], ref)
        curve.append((spp, round(ms, 2), float(f"{e:.4g}")))
        if e <= thresh:
            break
        target = int(target * growth)
    gpu.set_option("guiding", 0)
    return {"ms": round(ms, 2), "spp": spp, "rel_mse": e, "reached": e <= thresh, "curve": curve}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--spp-per-step", type=int, default=256)
    ap.add_argument("--depth", type=int, default=12)
    ap.add_argument("--subdiv", type=int, default=6)
    ap.add_argument("--batch", type=int, default=1 << 21)
    ap.add_argument("--no-autotune", action="store_true", help="keep --batch and --chunk-iters")
    ap.add_argument("--chunk-iters", type=int, default=8, help="pool iterations per done-flag check (--no-autotune)")
    ap.add_argument("--batched", action="store_true",
                    help="also render the batched textured scene (16 views, 256^2 .. 1920x1080, 16 .. 256 spp) in "
                         "one persistent render call and report its throughput")
    ap.add_argument("--guided", action="store_true",
                    help="time guided rendering (online-trained network) instead of plain path_mis")
    ap.add_argument("--tune-spp", type=int, default=0,
                    help="spp of the autotuner's calibration renders (0: one step's spp, so the candidates run "
                         "the timed steps' workload)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-spp", type=int, default=64)
    ap.add_argument("--no-time-to-mse", action="store_true", help="skip the time-to-relMSE leg")
    ap.add_argument("--mse-res", type=int, default=64, help="image side of the time-to-relMSE leg")
    ap.add_argument("--mse-ref-spp", type=int, default=1 << 20)
    ap.add_argument("--mse-start-spp", type=int, default=256)
    ap.add_argument("--mse-growth", type=float, default=1.25)
    ap.add_argument("--mse-max-spp", type=int, default=1 << 18)
    ap.add_argument("--mse-threshold", type=float, default=1e-3)
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="only run a device copy of BYTES bytes (counter calibration) and exit")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus} is not supported; bench.py measures one GPU (use --gpus 1)",
              file=sys.stderr)
        sys.exit(2)

    import nori_gpu as ng

    t0 = time.time()
    scene = ng.generate_scene(2, subdiv=args.subdiv)
    cam = scene.camera
    log(f"[bench] scene: {len(scene.indices)} triangles, envmap {scene.env_rgb.shape[1]}x{scene.env_rgb.shape[0]}, "
        f"{time.time() - t0:.1f}s")

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(scene, cam, args.cpu_spp, args.depth)
        log(f"[bench] cpu baseline {cpu['value']:.3f} Msamples/s on {cpu['cores']} threads")

    gpu = ng.NoriGPU(0)
    t0 = time.time()
    gpu.upload(scene)
    info = gpu.bvh_info()
    log(f"[bench] upload: BVH8 {info.nodes} nodes depth {info.max_depth} ({info.node_bytes / 1e6:.1f} MB nodes, "
        f"{info.tri_bytes / 1e6:.1f} MB triangles), build {info.build_ms:.0f} ms, total {time.time() - t0:.1f}s")
    gpu.set_option("batch", args.batch)
    gpu.set_option("chunk_iters", args.chunk_iters)
    spp = args.spp_per_step
    tune = None
    if not args.no_autotune:
        # Untimed: pool size and chunk length for this view, from short renders.
        tspp = args.tune_spp or spp
        r = gpu.autotune(ng.make_view(cam, tspp, args.depth), calib_spp=tspp)
        tune = {"pool": r.best_pool, "chunk_iters": r.best_chunk,
                "candidates": [[r.pool[k], r.chunk[k], round(r.msamples_per_s[k], 1)] for k in range(r.count)]}
        log(f"[bench] autotune: pool {r.best_pool}, chunk {r.best_chunk} "
            f"({r.best_msamples_per_s:.1f} Msamples/s on the calibration renders)")

    # Traversal counts of this workload (instrumented kernels, untimed): average
    # BVH8 nodes and triangles per extension and per shadow ray.
    gpu.set_option("count_traversal", 1)
    gpu.reset_film()
    gpu.render(ng.make_view(cam, 4, args.depth, sample_offset=10 ** 6), seed=99, readback=False)
    cs = gpu.stats()
    gpu.set_option("count_traversal", 0)
    ext_nodes = cs.ext_nodes / max(cs.extension_rays, 1)
    ext_tris = cs.ext_tris / max(cs.extension_rays, 1)
    sh_nodes = cs.shadow_nodes / max(cs.shadow_rays, 1)
    sh_tris = cs.shadow_tris / max(cs.shadow_rays, 1)
    log(f"[bench] per extension ray {ext_nodes:.2f} nodes {ext_tris:.2f} tris; per shadow ray {sh_nodes:.2f} nodes "
        f"{sh_tris:.2f} tris")

    if args.guided:
        gpu.set_option("guiding", 1)
    gpu.reset_film()
    for w in range(args.warmup):
        gpu.render(ng.make_view(cam, spp, args.depth, sample_offset=(10 ** 6) + 64 + w * spp), seed=7,
                   readback=False)
        log(f"[bench] warmup step {w + 1}/{args.warmup} done")
    gpu.reset_film()

    # Timed region: K render calls (graph-replayed bounce loop, no per-kernel events).
    acc = {"samples": 0, "ext": 0, "sh": 0, "ms_ext": 0.0, "ms_sh": 0.0, "ms_shade": 0.0, "launch_ext": 0,
           "ms_gpu": 0.0}
    t0 = time.perf_counter()
    for k in range(args.steps):
        gpu.render(ng.make_view(cam, spp, args.depth, sample_offset=k * spp), seed=7, readback=False)
        s = gpu.stats()
        acc["samples"] += s.samples
        acc["ext"] += s.extension_rays
        acc["sh"] += s.shadow_rays
        acc["ms_gpu"] += s.ms_total
        log(f"[bench] step {k + 1}/{args.steps}: {s.samples / s.ms_total / 1e3:.1f} Msamples/s")
    wall = time.perf_counter() - t0

    # Kernel times: the same K steps again with HIP events around every launch
    # on the renderer's stream (eager launches; events inside a graph are not used).
    gpu.set_option("timing_mode", 1)
    gpu.reset_film()
    acc["ext_timed"] = 0
    for k in range(args.steps):
        gpu.render(ng.make_view(cam, spp, args.depth, sample_offset=k * spp), seed=7, readback=False)
        s = gpu.stats()
        acc["ext_timed"] += s.extension_rays
        acc["ms_ext"] += s.ms_trace_closest
        acc["ms_sh"] += s.ms_trace_shadow
        acc["ms_shade"] += s.ms_shade
        acc["launch_ext"] += s.launches_closest
        log(f"[bench] timing pass {k + 1}/{args.steps}: extend {s.ms_trace_closest:.1f} ms, shadow "
            f"{s.ms_trace_shadow:.1f} ms, shade {s.ms_shade:.1f} ms, total {s.ms_total:.1f} ms")
    gpu.set_option("timing_mode", 0)
    gpu.set_option("guiding", 0)

    value = acc["samples"] / wall / 1e6

    ext_bytes = acc["ext_timed"] * (NODE_BYTES * ext_nodes + TRI_BYTES * ext_tris + EXT_IO_BYTES)
    achieved = ext_bytes / (acc["ms_ext"] / 1e3) / 1e9 if acc["ms_ext"] > 0 else None
    traffic_per_ray = traffic_from_profiles()
    rays_per_launch = acc["ext_timed"] / max(acc["launch_ext"], 1)
    traffic = None if traffic_per_ray is None else round(traffic_per_ray * rays_per_launch)

    ttm = None
    if not args.no_time_to_mse:
        # The same view at mse_res x mse_res: a plain reference at mse_ref_spp
        # (independent seed), then plain and guided (network trained online from
        # its initialisation) progressive renders until relMSE <= threshold.
        # Plain path tracing of this scene is dominated by rare high-energy paths
        # (sun caustics off glossy spheres), so 1e-3 takes ~1e5 spp; the leg runs
        # at reduced resolution to keep the reference converged and the run short.
        gpu.set_option("timing_mode", 0)
        mcam = ng.lookat_camera_like(cam, args.mse_res, args.mse_res)
        gpu.reset_film()
        t0 = time.time()
        refimg = gpu.render(ng.make_view(mcam, args.mse_ref_spp, args.depth, sample_offset=1 << 24), seed=1234)
        ref = refimg[:, :3] / refimg[:, 3:4]
        log(f"[bench] relMSE reference: {args.mse_res}x{args.mse_res} at {args.mse_ref_spp} spp in "
            f"{time.time() - t0:.1f}s")
        legs = {}
        for name, guided, seed in (("plain", False, 77), ("guided", True, 78)):
            legs[name] = time_to_rel_mse(gpu, ng, mcam, args.depth, ref, args.mse_start_spp, args.mse_growth,
                                         args.mse_max_spp, args.mse_threshold, guided, seed)
            r = legs[name]
            log(f"[bench] {name}: relMSE {r['rel_mse']:.3g} at {r['spp']} spp after {r['ms']:.0f} ms GPU time")
        # Where the threshold is not reached within mse_max_spp, the time to it
        # is extrapolated from the last point with relMSE ~ 1/N (labelled).
        for r in legs.values():
            r["ms_to_threshold"] = r["ms"] if r["reached"] else round(r["ms"] * r["rel_mse"] / args.mse_threshold, 1)
            r["extrapolated"] = not r["reached"]
        ttm = {"threshold": args.mse_threshold, "image": f"{args.mse_res}x{args.mse_res}", "ref_spp": args.mse_ref_spp,
               "plain": legs["plain"], "guided": legs["guided"],
               "speedup_guided_over_plain": round(legs["plain"]["ms_to_threshold"] / legs["guided"]["ms_to_threshold"],
                                                  3)}

    result = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"Disney icosphere grid 8x8, level {args.subdiv}, {len(scene.indices)} triangles, "
                        f"2048x1024 sky+sun envmap, {cam.width}x{cam.height}, {spp} spp per step, "
                        f"max depth {args.depth}, {'guided' if args.guided else 'plain'} path_mis",
            "spp_per_step": spp,
            "path_pool": tune["pool"] if tune else args.batch,
            "autotune": tune,
        },
        "mrays_per_s": round((acc["ext"] + acc["sh"]) / wall / 1e6, 2),
        "gpu_ms": {"extend": round(acc["ms_ext"], 2), "shadow": round(acc["ms_sh"], 2),
                   "shade_raygen": round(acc["ms_shade"], 2), "render_total": round(acc["ms_gpu"], 2)},
        "traversal": {"ext_nodes_per_ray": round(ext_nodes, 3), "ext_tris_per_ray": round(ext_tris, 3),
                      "shadow_nodes_per_ray": round(sh_nodes, 3), "shadow_tris_per_ray": round(sh_tris, 3)},
        "roofline": {
            "kernel": "k_extend (BVH8 closest-hit traversal)",
            "bound": "hbm",
            "achieved": None if achieved is None else round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": None if achieved is None else round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "bytes_per_launch": round(ext_bytes / max(acc["launch_ext"], 1)),
            "avg_launch_ms": round(acc["ms_ext"] / max(acc["launch_ext"], 1), 4),
        },
        "cpu_baseline": cpu,
        "time_to_rel_mse": ttm,
    }
    if cpu:
        result["gpu_over_cpu"] = round(value / cpu["value"], 1)
    if args.batched:
        result["batched"] = batched_leg(gpu, args, log)
    print(json.dumps(result), flush=True)
    gpu.close()


def batched_leg(gpu, args, log):
    """The batched textured scene (9.4 M triangles, 128 1024^2 RGBA8 textures,
    alpha-tested cards, thin lens): sixteen views of 256^2 .. 1920x1080 at
    16 .. 256 spp in one persistent render call, after an untimed autotune on
    the same views at 8 spp each."""
    import nori_gpu as ng
    t0 = time.time()
    sc = ng.generate_scene(5)
    log(f"[bench] batched scene: {len(sc.indices)} triangles, {len(sc.textures)} textures, "
        f"{time.time() - t0:.1f}s")
    gpu.upload(sc)
    sizes = [(256, 256), (320, 240), (400, 300), (480, 270), (512, 512), (640, 360), (640, 480), (800, 450),
             (800, 600), (960, 540), (1024, 576), (1024, 768), (1280, 720), (1366, 768), (1600, 900), (1920, 1080)]
    spps = list(range(16, 257, 16))
    views = [ng.make_view(ng.lookat_camera_like(sc.camera, w, h), spp, args.depth) for (w, h), spp in zip(sizes, spps)]
    r = gpu.autotune(views, calib_spp=8)
    samples = sum(v.camera.width * v.camera.height * v.spp for v in views)
    gpu.reset_film()
    gpu.render(views, seed=3, readback=False)  # warm-up (graph capture)
    t = time.perf_counter()
    gpu.render([ng.make_view(v.camera, v.spp, v.max_depth, sample_offset=v.spp) for v in views], seed=3,
               readback=False)
    ms = (time.perf_counter() - t) * 1e3
    log(f"[bench] batched: {samples / ms / 1e3:.1f} Msamples/s over {samples} samples in {ms:.0f} ms")
    return {"workload": "batched textured scene, 9.4 M triangles, 128 x 1024^2 textures, 16 views 256^2 .. "
                        "1920x1080 at 16 .. 256 spp, one render call", "samples": samples, "ms": round(ms, 1),
            "msamples_per_s": round(samples / ms / 1e3, 2), "path_pool": r.best_pool, "chunk_iters": r.best_chunk}


if __name__ == "__main__":
    main()
