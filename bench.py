"""Benchmark: Msamples/s of the MI355X wavefront path tracer on the batched
large scene, and time to relMSE <= 1e-3 (validation_tests/compare_exr.py:102),
guided vs plain, on the Disney grid.

Headline (`value`): the batched textured scene, the largest workload that fits
one GPU: 10,000,448 triangles (a 2048^2-cell heightfield terrain, 199 displaced
icospheres, 20,000 alpha-tested hair cards), 128 procedural 1024^2 RGBA8
textures (64 albedo, 64 tangent-space normal maps) plus the cards' alpha map,
mixed Disney / Lambert / dielectric materials, a thin-lens camera (aperture
radius 0.05, focus distance 5), sixteen views from 256^2 to 1920x1080 at 16 to
256 spp rendered in one persistent render call, max depth 12, plain path_mis.
A step renders all sixteen views once (1.85 G samples); sample offsets advance
from step to step, so every step draws new samples.  Inputs are resident in
HBM before the timed region; the timed region is K render calls, each ending
in a stream synchronisation.

Secondary legs: the Disney icosphere grid (5,242,882 triangles, 2048x1024
sky+sun envmap, 512x512, 256 spp, max depth 12) plain; the time to relMSE
<= 1e-3 of plain and guided rendering of that grid at its 512x512 resolution
and of the sun room (guide_scenes.py, a scene lit through a wall patch that
next-event estimation cannot reach) at 256x256; and the throughput of the Cornell box, the 50,000-strand hair scene (lit and
white furnace) and the 256^3 heterogeneous medium (g = 0, 0.3, 0.8), each at
its full size.

Prints one JSON line.  --gpus N > 1 is not supported: this build runs on one GPU.
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
# per extension ray besides the traversal: ray read 32 B, hit write 16 B, class-queue entry 4 B
EXT_IO_BYTES = 32 + 16 + 4
# per alpha-tested candidate: its 32-B alpha record (3 UVs, texture level 0) and a 2x2 footprint of
# decoded (red, alpha) texels, 8 B each
ALPHA_TEST_BYTES = 32 + 4 * 8
BATCHED_SIZES = [(256, 256), (320, 240), (400, 300), (480, 270), (512, 512), (640, 360), (640, 480), (800, 450),
                 (800, 600), (960, 540), (1024, 576), (1024, 768), (1280, 720), (1366, 768), (1600, 900),
                 (1920, 1080)]
BATCHED_SPP = list(range(16, 257, 16))


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_cpu():
    """Host cores this process may use, the machine's count and CPU model.  A
    host shared by several GPUs grants each one a share of its cores
    (OMP_NUM_THREADS, where set); the CPU baseline runs on every core of that
    share."""
    nproc = os.cpu_count() or 1
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = nproc
    share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    threads = min(usable, share) if share > 0 else usable
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return threads, {"nproc": nproc, "affinity": usable, "share": share or None, "cpu_model": model}


def cpu_baseline(scene, views, depth, budget_s):
    """The CPU oracle (path_mis restated on the CPU over Nori's BVH2, 32x32
    blocks on a std::thread pool) on a bounded sample of the same workload:
    the views in order at 1 spp each (the first sample of every pixel) until
    budget_s of wall time has been spent."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import orc
    threads, info = host_cpu()
    t0 = time.time()
    o = orc.OracleScene(scene)
    log(f"[bench] oracle scene (Nori BVH2) built in {time.time() - t0:.1f}s")
    samples, dt, done = 0, 0.0, []
    for k, v in enumerate(views):
        t0 = time.time()
        _, stats = o.render(v.camera, 1, depth, seed=1, nthreads=threads)
        dt += time.time() - t0
        samples += int(stats[0])
        done.append(f"{v.camera.width}x{v.camera.height}")
        if dt > budget_s:
            break
    return {"value": samples / dt / 1e6, "unit": "Msamples/s", "cores": threads, "kind": "port",
            "sample": f"views {', '.join(done)} at 1 spp (the first sample of every pixel), {samples} samples in "
                      f"{dt:.2f}s", **info}


def traffic_from_profiles(name):
    """HBM bytes per extension ray of the closest-hit kernel from a committed
    rocprof counter summary (its line `hbm_bytes_per_ray: N`: FETCH_SIZE +
    WRITE_SIZE per launch, scaled by the copy calibration, divided by the rays
    of those launches)."""
    path = os.path.join(ROOT, "profiles", name)
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


def measure(gpu, ng, make_views, steps, warmup, tune, tune_views, pool, chunk, traversal, waves, label, timing_steps):
    """Untimed autotune (or the given pool/chunk), an untimed traversal-count
    pass, `warmup` untimed steps, `steps` timed steps (graph-replayed bounce
    loop), then the first `timing_steps` of those steps again with HIP events
    around every launch on the renderer's stream (kernel times)."""
    gpu.set_option("batch", pool)
    gpu.set_option("chunk_iters", chunk)
    gpu.set_option("traversal", traversal)
    gpu.set_option("trace_waves", waves)
    tuned = None
    if tune:
        r = gpu.autotune(tune_views, calib_spp=tune_views[0].spp)
        tuned = {"pool": r.best_pool, "chunk_iters": r.best_chunk, "traversal": r.best_traversal,
                 "trace_waves": r.best_trace_waves,
                 "candidates [pool, chunk, traversal, waves per SIMD cap, Msamples/s]":
                     [[r.pool[k], r.chunk[k], r.traversal[k], r.trace_waves[k], round(r.msamples_per_s[k], 1)]
                      for k in range(r.count)]}
        log(f"[bench] {label} autotune: pool {r.best_pool}, chunk {r.best_chunk}, traversal {r.best_traversal}, "
            f"waves cap {r.best_trace_waves} ({r.best_msamples_per_s:.1f} Msamples/s on the calibration renders)")
    gpu.set_option("count_traversal", 1)
    gpu.reset_film()
    gpu.render(make_views(10 ** 6, True), seed=99, readback=False)
    cs = gpu.stats()
    gpu.set_option("count_traversal", 0)
    per = lambda a, b: a / max(b, 1)  # noqa: E731
    trav = {"ext_nodes_per_ray": per(cs.ext_nodes, cs.extension_rays), "ext_tris_per_ray": per(cs.ext_tris,
                                                                                                 cs.extension_rays),
            "ext_alpha_tests_per_ray": per(cs.ext_alpha_tests, cs.extension_rays),
            "shadow_nodes_per_ray": per(cs.shadow_nodes, cs.shadow_rays),
            "shadow_tris_per_ray": per(cs.shadow_tris, cs.shadow_rays),
            "shadow_alpha_tests_per_ray": per(cs.shadow_alpha_tests, cs.shadow_rays)}
    log(f"[bench] {label} per extension ray {trav['ext_nodes_per_ray']:.2f} nodes {trav['ext_tris_per_ray']:.2f} "
        f"tris {trav['ext_alpha_tests_per_ray']:.3f} alpha tests; per shadow ray {trav['shadow_nodes_per_ray']:.2f} "
        f"nodes {trav['shadow_tris_per_ray']:.2f} tris")
    gpu.reset_film()
    for w in range(warmup):
        gpu.render(make_views((10 ** 6) + 64 + w * 256, False), seed=7, readback=False)
        log(f"[bench] {label} warmup step {w + 1}/{warmup} done")
    gpu.reset_film()
    acc = {"samples": 0, "ext": 0, "sh": 0, "ms_gpu": 0.0}
    t0 = time.perf_counter()
    for k in range(steps):
        gpu.render(make_views(k * 256, False), seed=7, readback=False)
        s = gpu.stats()
        acc["samples"] += s.samples
        acc["ext"] += s.extension_rays
        acc["sh"] += s.shadow_rays
        acc["ms_gpu"] += s.ms_total
        log(f"[bench] {label} step {k + 1}/{steps}: {s.samples / s.ms_total / 1e3:.1f} Msamples/s, "
            f"{s.extension_rays} extension rays in {s.launches_closest} launches")
    wall = time.perf_counter() - t0
    gpu.set_option("timing_mode", 1)
    gpu.reset_film()
    tm = {"ext": 0, "ms_ext": 0.0, "ms_sh": 0.0, "ms_shade": 0.0, "launch_ext": 0, "ms_total": 0.0}
    for k in range(timing_steps):
        gpu.render(make_views(k * 256, False), seed=7, readback=False)
        s = gpu.stats()
        tm["ext"] += s.extension_rays
        tm["ms_ext"] += s.ms_trace_closest
        tm["ms_sh"] += s.ms_trace_shadow
        tm["ms_shade"] += s.ms_shade
        tm["launch_ext"] += s.launches_closest
        tm["ms_total"] += s.ms_total
        log(f"[bench] {label} timing pass {k + 1}/{timing_steps}: extend {s.ms_trace_closest:.1f} ms, shadow "
            f"{s.ms_trace_shadow:.1f} ms, shade {s.ms_shade:.1f} ms, total {s.ms_total:.1f} ms")
    gpu.set_option("timing_mode", 0)
    ext_b = NODE_BYTES * trav["ext_nodes_per_ray"] + TRI_BYTES * trav["ext_tris_per_ray"] + EXT_IO_BYTES + \
        ALPHA_TEST_BYTES * trav["ext_alpha_tests_per_ray"]
    bytes_launch = tm["ext"] * ext_b / max(tm["launch_ext"], 1)
    ms_launch = tm["ms_ext"] / max(tm["launch_ext"], 1)
    achieved = bytes_launch / (ms_launch / 1e3) / 1e9 if ms_launch > 0 else None
    return {
        "msamples_per_s": acc["samples"] / wall / 1e6, "ms_per_step": wall / steps * 1e3,
        "samples_per_step": acc["samples"] // max(steps, 1),
        "mrays_per_s": (acc["ext"] + acc["sh"]) / wall / 1e6,
        "gpu_ms": {"extend": round(tm["ms_ext"], 2), "shadow": round(tm["ms_sh"], 2),
                   "shade_raygen": round(tm["ms_shade"], 2), "render_total_timed_pass": round(tm["ms_total"], 2),
                   "render_total": round(acc["ms_gpu"], 2)},
        "traversal": {k: round(v, 4) for k, v in trav.items()},
        "ext_bytes_per_ray": ext_b, "rays_per_launch": tm["ext"] / max(tm["launch_ext"], 1),
        "bytes_per_launch": bytes_launch, "avg_launch_ms": ms_launch, "achieved_gbs": achieved,
        "autotune": tuned,
    }


def roofline(m, profile, kernel):
    per_ray = traffic_from_profiles(profile)
    return {"kernel": kernel, "bound": "hbm",
            "achieved": None if m["achieved_gbs"] is None else round(m["achieved_gbs"], 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": None if m["achieved_gbs"] is None else round(m["achieved_gbs"] / HBM_PEAK_GBS, 4),
            "traffic": None if per_ray is None else round(per_ray * m["rays_per_launch"]),
            "bytes_per_ray": round(m["ext_bytes_per_ray"], 1), "bytes_per_launch": round(m["bytes_per_launch"]),
            "avg_launch_ms": round(m["avg_launch_ms"], 4)}


def batched_views(ng, cam):
    return [ng.make_view(ng.lookat_camera_like(cam, w, h), spp, 12) for (w, h), spp in zip(BATCHED_SIZES, BATCHED_SPP)]


def time_to_rel_mse(gpu, ng, cam, depth, ref, ref_err, start_spp, max_spp, thresh, guided, seed, train=()):
    """Progressive render against `ref` at 1, 2, 4, ... times start_spp
    samples per pixel, until the relMSE (less the reference's own error
    ref_err) reaches thresh or max_spp.  Guided: the network starts from its
    initialisation and trains online during the passes listed in `train`
    ((spp, Adam learning rate x 1e6) each, 32 steps per pool round), which
    count toward the image and the time; it is frozen afterwards.  Returns the
    GPU milliseconds of the renders (every kernel, guiding inference and
    training included), the spp reached and the curve.  The time to thresh
    is interpolated between the two points around the crossing (a straight
    line in log relMSE over log time); past the last point it is
    extrapolated with relMSE ~ 1/N (labelled)."""
    gpu.set_option("guiding", 1 if guided else 0)
    gpu.reset_film()
    ms, spp, curve, e = 0.0, 0, [], float("inf")

    def rmse(img):
        return max(rel_mse(img[:, :3] / img[:, 3:4], ref) - ref_err, 0.0)

    steps = 0
    if guided:
        gpu.set_option("guide_reset", 1)
        gpu.set_option("guide_train_steps", 32)
        for chunk, lr in train:
            gpu.set_option("guide_lr_micro", lr)
            img = gpu.render(ng.make_view(cam, chunk, depth, sample_offset=spp), seed=seed)
            st = gpu.stats()
            ms += st.ms_total
            steps += st.guide_train_steps
            spp += chunk
        gpu.set_option("guide_train_steps", 0)
        if spp:
            e = rmse(img)
            curve.append((spp, round(ms, 2), float(f"{e:.4g}")))
    target = max(start_spp, 2 * spp)
    while spp < max_spp and e > thresh:
        chunk = min(target - spp, max_spp - spp)
        img = gpu.render(ng.make_view(cam, chunk, depth, sample_offset=spp), seed=seed)
        ms += gpu.stats().ms_total
        spp += chunk
        e = rmse(img)
        curve.append((spp, round(ms, 2), float(f"{e:.4g}")))
        target *= 2
    gpu.set_option("guiding", 0)
    gpu.set_option("guide_train_steps", 4)
    gpu.set_option("guide_lr_micro", 2000)
    reached = e <= thresh
    if not reached:
        t_thr = ms * e / thresh
    elif len(curve) >= 2 and curve[-2][2] > thresh and e > 0.0:
        # crossing between the last two points, on the straight line through
        # them in log relMSE over log time
        (_, t1, e1), (_, t2, e2) = curve[-2], curve[-1]
        t_thr = t1 * (t2 / t1) ** (np.log(e1 / thresh) / np.log(e1 / e2))
    else:
        t_thr = ms
    out = {"ms": round(ms, 2), "spp": spp, "rel_mse": e, "reached": reached,
           "ms_to_threshold": round(float(t_thr), 1), "extrapolated": not reached,
           "curve": curve}
    if guided:
        out["train_steps"] = steps
        out["train_passes"] = [list(t) for t in train]
    return out


def time_to_rel_mse_legs(gpu, ng, sc, args):
    """Plain and guided time to relMSE <= threshold on the Disney grid at
    512x512, depth 12, once with the reference's envmap sampling weights
    (texel luminance) and once with the 3 x 3 footprint weights.  Per mode the
    reference is two independent plain renders of mse_ref_spp/2 each; their
    mutual relMSE is four times the full reference's own error, which is
    subtracted from each leg's measured relMSE."""
    cam = sc.camera
    train = [(64, 2000)] * 8 + [(64, 500)] * 4
    modes = {}
    for mode, name in ((0, "texel"), (1, "footprint")):
        gpu.set_option("env_sampling", mode)
        gpu.upload(sc)
        gpu.set_option("guiding", 0)
        half = args.mse_ref_spp // 2
        refs = []
        t0 = time.time()
        for k in range(2):
            gpu.reset_film()
            img = gpu.render(ng.make_view(cam, half, 12, sample_offset=(1 << 24) + k * half), seed=1234 + k)
            refs.append(img[:, :3] / img[:, 3:4])
        ref = 0.5 * (refs[0] + refs[1])
        ref_err = rel_mse(refs[0], refs[1]) / 4
        log(f"[bench] relMSE reference ({name} envmap sampling): 512x512 at {args.mse_ref_spp} spp in "
            f"{time.time() - t0:.1f}s, own relMSE {ref_err:.3g}")
        tt = {"ref_rel_mse": ref_err}
        for leg, guided, seed in (("plain", False, 77), ("guided", True, 78)):
            tt[leg] = time_to_rel_mse(gpu, ng, cam, 12, ref, ref_err, args.mse_start_spp, args.mse_max_spp,
                                      args.mse_threshold, guided, seed, train)
            r = tt[leg]
            log(f"[bench] {name} {leg}: relMSE {r['rel_mse']:.3g} at {r['spp']} spp after {r['ms']:.0f} ms GPU time")
        tt["speedup_guided_over_plain"] = round(tt["plain"]["ms_to_threshold"] / tt["guided"]["ms_to_threshold"], 3)
        modes[name] = tt
    gpu.set_option("env_sampling", 0)  # for later uploads
    t, f = modes["texel"], modes["footprint"]
    return {"threshold": args.mse_threshold, "image": "512x512", "depth": 12, "ref_spp": args.mse_ref_spp,
            "envmap_sampling": modes,
            # the network's own gain, both legs with the same envmap sampling
            "speedup_guided_over_plain": f["speedup_guided_over_plain"],
            "speedup_guided_over_plain_texel_sampling": t["speedup_guided_over_plain"],
            # footprint weights alone, plain against plain
            "speedup_footprint_over_texel_plain": round(t["plain"]["ms_to_threshold"] /
                                                        f["plain"]["ms_to_threshold"], 3),
            "speedup_footprint_guided_over_texel_plain": round(t["plain"]["ms_to_threshold"] /
                                                               f["guided"]["ms_to_threshold"], 3)}


def sun_room_legs(gpu, ng, args):
    """Plain and guided time to relMSE <= threshold on the sun room (a closed
    room lit through a roof opening, the sun reaching most surfaces only via a
    sunlit wall patch that next-event estimation cannot see), 256x256, depth 8,
    footprint envmap weights.  The guided leg trains the network from its
    initialisation online for four passes of 64 spp (32 Adam steps per pool
    round, learning rate 2e-3, the last pass 5e-4), which count toward the
    image and the time, then renders with the network frozen.  Reference: two
    independent plain renders of 32768 spp, own error subtracted as in the
    Disney grid legs."""
    from guide_scenes import sun_room

    sc = sun_room(width=256, height=256)
    cam, depth, half = sc.camera, 8, 1 << 15
    gpu.set_option("env_sampling", 1)
    gpu.upload(sc)
    gpu.set_option("guiding", 0)
    refs = []
    t0 = time.time()
    for k in range(2):
        gpu.reset_film()
        img = gpu.render(ng.make_view(cam, half, depth, sample_offset=(1 << 24) + k * half), seed=2234 + k)
        refs.append(img[:, :3] / img[:, 3:4])
    ref = 0.5 * (refs[0] + refs[1])
    ref_err = rel_mse(refs[0], refs[1]) / 4
    log(f"[bench] relMSE reference (sun room): 256x256 at {2 * half} spp in {time.time() - t0:.1f}s, "
        f"own relMSE {ref_err:.3g}")
    train = [(64, 2000)] * 3 + [(64, 500)]
    tt = {"threshold": args.mse_threshold, "image": "256x256", "depth": depth, "ref_spp": 2 * half,
          "envmap_sampling": "footprint", "ref_rel_mse": ref_err}
    for leg, guided, seed in (("plain", False, 87), ("guided", True, 88)):
        tt[leg] = time_to_rel_mse(gpu, ng, cam, depth, ref, ref_err, 1024, 1 << 14, args.mse_threshold, guided,
                                  seed, train)
        r = tt[leg]
        log(f"[bench] sun room {leg}: relMSE {r['rel_mse']:.3g} at {r['spp']} spp after {r['ms']:.0f} ms GPU time")
    tt["speedup_guided_over_plain"] = round(tt["plain"]["ms_to_threshold"] / tt["guided"]["ms_to_threshold"], 3)
    gpu.set_option("env_sampling", 0)  # for later uploads
    return tt


def guide_frame(gpu, ng):
    """Kernel times of one frame of the sun room (128x128, 1024 spp, depth 8,
    footprint envmap weights, default pool), plain and with a network trained
    for four guided passes of 64 spp and then frozen: each frame once as a
    graph-replayed render and once with HIP events around every launch."""
    from guide_scenes import sun_room

    sc = sun_room(width=128, height=128)
    cam = sc.camera
    gpu.set_option("env_sampling", 1)
    gpu.upload(sc)
    gpu.set_option("guide_reset", 1)
    gpu.set_option("guide_train_steps", 32)
    gpu.set_option("guiding", 1)
    for k in range(4):
        gpu.render(ng.make_view(cam, 64, 8, sample_offset=k * 64), seed=5, readback=False)
    gpu.set_option("guide_train_steps", 0)
    out = {}
    for name, guided in (("plain", 0), ("guided", 1)):
        gpu.set_option("guiding", guided)
        gpu.reset_film()
        gpu.render(ng.make_view(cam, 1024, 8, sample_offset=4096), seed=7, readback=False)
        graph_ms = gpu.stats().ms_total
        gpu.set_option("timing_mode", 1)
        gpu.reset_film()
        gpu.render(ng.make_view(cam, 1024, 8, sample_offset=4096), seed=7, readback=False)
        s = gpu.stats()
        gpu.set_option("timing_mode", 0)
        out[name] = {"graph_ms": round(graph_ms, 2), "extend_ms": round(s.ms_trace_closest, 2),
                     "shadow_ms": round(s.ms_trace_shadow, 2), "shade_ms": round(s.ms_shade, 2),
                     "guide_infer_ms": round(s.ms_guide_infer, 2), "timed_total_ms": round(s.ms_total, 2),
                     "extension_rays": int(s.extension_rays)}
        log(f"[bench] guide frame {name}: {out[name]}")
    gpu.set_option("guiding", 0)
    gpu.set_option("guide_train_steps", 4)
    gpu.set_option("env_sampling", 0)  # for later uploads
    return out


CONFIG_LEGS = [
    # name, generator config, generator options, spp, max depth, what it is
    ("cornell", 1, {}, 64, 8, "Cornell box, 2 level-4 icospheres, 10,252 triangles, quad light, 256x256, 64 spp, "
                              "max depth 8, RR from 3"),
    ("hair_lit", 3, {}, 256, 32, "50,000 hair strands x 16 segments (1.6 M ribbon triangles), Chiang BCSDF, three "
                                 "quad lights, 512x512, 256 spp, max depth 32"),
    ("hair_furnace", 3, {"variant": 1}, 256, 32, "50,000 hair strands, sigma_a = 0, uniform unit envmap (white "
                                                 "furnace), 512x512, 256 spp, max depth 32"),
    ("volume_g0", 4, {"variant": 0}, 512, 64, "256^3 fBm density, max sigma_t 20, albedo 0.9, HG g = 0, 16^3 "
                                              "brick majorants, diffuse box, sky+sun envmap, 512x512, 512 spp, "
                                              "max depth 64"),
    ("volume_g0.3", 4, {"variant": 1}, 512, 64, "as volume_g0 with HG g = 0.3"),
    ("volume_g0.8", 4, {"variant": 2}, 512, 64, "as volume_g0 with HG g = 0.8"),
]


def config_legs(gpu, ng, steps, pool):
    """Msamples/s of the other workload configurations, each at its full size:
    one untimed frame, then `steps` timed frames (one render call each, new
    sample offsets every frame), plain path_mis, at the pool and chunk length
    the autotuner picks on the two large scenes."""
    out = {}
    for name, cfg, kw, spp, depth, what in CONFIG_LEGS:
        t0 = time.time()
        sc = ng.generate_scene(cfg, **kw)
        gpu.upload(sc)
        cam = sc.camera
        samples = cam.width * cam.height * spp
        gpu.set_option("batch", max(64, min(pool, (samples + 255) // 256 * 256)))
        gpu.set_option("chunk_iters", 2)
        gpu.set_option("traversal", 1)
        gpu.set_option("trace_waves", 0)
        gpu.reset_film()
        gpu.render(ng.make_view(cam, spp, depth, sample_offset=1 << 22), seed=3, readback=False)
        gpu.reset_film()
        acc = {"samples": 0, "rays": 0}
        t1 = time.perf_counter()
        for k in range(steps):
            gpu.render(ng.make_view(cam, spp, depth, sample_offset=k * spp), seed=5, readback=False)
            s = gpu.stats()
            acc["samples"] += s.samples
            acc["rays"] += s.extension_rays + s.shadow_rays
        wall = time.perf_counter() - t1
        out[name] = {"workload": what, "triangles": int(len(sc.indices)), "msamples_per_s": round(acc["samples"] / wall
                                                                                                  / 1e6, 2),
                     "mrays_per_s": round(acc["rays"] / wall / 1e6, 2), "ms_per_frame": round(wall / steps * 1e3, 2),
                     "frames": steps, "path_pool": min(pool, (samples + 255) // 256 * 256)}
        log(f"[bench] {name}: {out[name]['msamples_per_s']} Msamples/s, {out[name]['mrays_per_s']} Mrays/s "
            f"({steps} frames of {samples} samples, {time.time() - t0:.1f}s with scene set-up)")
        del sc
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-autotune", action="store_true", help="keep --batch and --chunk-iters")
    ap.add_argument("--batch", type=int, default=1 << 24, help="path pool (--no-autotune)")
    ap.add_argument("--chunk-iters", type=int, default=2, help="pool iterations per done-flag check (--no-autotune)")
    ap.add_argument("--traversal", type=int, default=1, help="triangle schedule of the traversal (--no-autotune)")
    ap.add_argument("--trace-waves", type=int, default=0, help="waves-per-SIMD cap of traversal (--no-autotune)")
    ap.add_argument("--only", choices=["batched", "disney", "configs", "sun_room", "guide_frame"], default=None,
                    help="run one workload leg only (profiling); the JSON line then carries that leg as value")
    ap.add_argument("--disney-steps", type=int, default=4)
    ap.add_argument("--timing-steps", type=int, default=1,
                    help="steps of the kernel-timing pass (HIP events around every launch)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=20.0, help="wall-time budget of the CPU baseline sample")
    ap.add_argument("--no-time-to-mse", action="store_true", help="skip the time-to-relMSE leg")
    ap.add_argument("--no-config-legs", action="store_true",
                    help="skip the Cornell, hair and volume configurations' throughput")
    ap.add_argument("--config-frames", type=int, default=3, help="timed frames per configuration leg")
    ap.add_argument("--mse-ref-spp", type=int, default=1 << 14)
    ap.add_argument("--mse-start-spp", type=int, default=64)
    ap.add_argument("--mse-max-spp", type=int, default=1 << 12)
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

    gpu = ng.NoriGPU(0)
    if args.only == "sun_room":
        r = sun_room_legs(gpu, ng, args)
        print(json.dumps({"metric": "time to relMSE<=1e-3, guided vs plain, sun room", "unit": "ms",
                          "sun_room": r}), flush=True)
        gpu.close()
        return
    if args.only == "guide_frame":
        print(json.dumps({"metric": "GPU ms per frame, plain vs guided, sun room 128x128 x 1024 spp", "unit": "ms",
                          "guide_frame": guide_frame(gpu, ng)}), flush=True)
        gpu.close()
        return
    legs = {}
    cpu = None
    if args.only in (None, "batched"):
        t0 = time.time()
        sc = ng.generate_scene(5)
        views = batched_views(ng, sc.camera)
        log(f"[bench] batched scene: {len(sc.indices)} triangles, {len(sc.textures)} textures, "
            f"{sum(v.camera.width * v.camera.height * v.spp for v in views)} samples per step, "
            f"{time.time() - t0:.1f}s")
        if not args.no_cpu_baseline:
            cpu = cpu_baseline(sc, views, 12, args.cpu_seconds)
            log(f"[bench] cpu baseline {cpu['value']:.3f} Msamples/s on {cpu['cores']} threads "
                f"(nproc {cpu['nproc']}, {cpu['cpu_model']})")
        t0 = time.time()
        gpu.upload(sc)
        info = gpu.bvh_info()
        log(f"[bench] upload: BVH8 {info.nodes} nodes depth {info.max_depth}, build {info.build_ms:.0f} ms, "
            f"total {time.time() - t0:.1f}s")

        def bviews(off, count_pass):
            return [ng.make_view(v.camera, 1 if count_pass else v.spp, v.max_depth, sample_offset=off)
                    for v in views]

        tune_views = [ng.make_view(v.camera, 8, v.max_depth, sample_offset=0x40000000) for v in views]
        m = measure(gpu, ng, bviews, args.steps, args.warmup, not args.no_autotune, tune_views, args.batch,
                    args.chunk_iters, args.traversal, args.trace_waves, "batched",
                    max(0, min(args.timing_steps, args.steps)))
        m["roofline"] = roofline(m, "k_extend_batched_traffic.md", "k_extend<ALPHA> (BVH8 closest hit, alpha test)")
        m["workload"] = (f"batched textured scene, {len(sc.indices):,} triangles (2048^2 heightfield, 199 displaced "
                         f"icospheres, 20,000 alpha-tested cards), 128 1024^2 RGBA8 textures, thin lens r 0.05 "
                         f"focus 5, 16 views 256^2 .. 1920x1080 at 16 .. 256 spp in one render call, max depth 12, "
                         f"plain path_mis")
        legs["batched"] = m
        del sc

    if args.only in (None, "disney"):
        t0 = time.time()
        sc = ng.generate_scene(2)
        cam = sc.camera
        log(f"[bench] Disney grid: {len(sc.indices)} triangles, envmap {sc.env_rgb.shape[1]}x"
            f"{sc.env_rgb.shape[0]}, {time.time() - t0:.1f}s")
        gpu.upload(sc)
        if cpu is None and not args.no_cpu_baseline:
            cpu = cpu_baseline(sc, [ng.make_view(cam, 1, 12)] * 64, 12, args.cpu_seconds)

        def dviews(off, count_pass):
            return [ng.make_view(cam, 4 if count_pass else 256, 12, sample_offset=off)]

        steps = args.steps if args.only == "disney" else args.disney_steps
        m = measure(gpu, ng, dviews, steps, args.warmup, not args.no_autotune, [ng.make_view(cam, 256, 12)],
                    args.batch, args.chunk_iters, args.traversal, args.trace_waves, "disney",
                    max(0, min(args.timing_steps, steps)))
        m["roofline"] = roofline(m, "k_extend_disney_grid_traffic.md", "k_extend (BVH8 closest hit)")
        m["workload"] = (f"Disney icosphere grid 8x8, level 6, {len(sc.indices):,} triangles, 2048x1024 sky+sun "
                         f"envmap, 512x512, 256 spp per step, max depth 12, plain path_mis")
        legs["disney"] = m
        if not args.no_time_to_mse and args.only is None:
            legs["time_to_rel_mse"] = time_to_rel_mse_legs(gpu, ng, sc, args)
        del sc
        if not args.no_time_to_mse and args.only is None:
            legs["time_to_rel_mse"]["sun_room"] = sun_room_legs(gpu, ng, args)

    if args.only in (None, "configs") and not args.no_config_legs:
        legs["configs"] = config_legs(gpu, ng, max(1, args.config_frames), args.batch)
        if args.only == "configs":
            print(json.dumps({"metric": METRIC, "unit": "Msamples/s", "other_configs": legs["configs"]}), flush=True)
            gpu.close()
            return

    head = legs["batched"] if "batched" in legs else legs["disney"]
    result = {
        "metric": METRIC,
        "value": round(head["msamples_per_s"], 3),
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": head["workload"], "samples_per_step": head["samples_per_step"],
                   "path_pool": head["autotune"]["pool"] if head["autotune"] else args.batch,
                   "autotune": head["autotune"]},
        "mrays_per_s": round(head["mrays_per_s"], 2),
        "gpu_ms": head["gpu_ms"],
        "traversal": head["traversal"],
        "roofline": head["roofline"],
        "cpu_baseline": cpu,
    }
    if cpu:
        result["gpu_over_cpu"] = round(head["msamples_per_s"] / cpu["value"], 1)
    if "batched" in legs and "disney" in legs:
        d = legs["disney"]
        result["disney_grid"] = {"workload": d["workload"], "msamples_per_s": round(d["msamples_per_s"], 3),
                                 "ms_per_step": round(d["ms_per_step"], 3), "mrays_per_s": round(d["mrays_per_s"], 2),
                                 "gpu_ms": d["gpu_ms"], "traversal": d["traversal"], "roofline": d["roofline"],
                                 "autotune": d["autotune"]}
    if "configs" in legs:
        result["other_configs"] = legs["configs"]
    if "time_to_rel_mse" in legs:
        result["time_to_rel_mse"] = legs["time_to_rel_mse"]
    print(json.dumps(result), flush=True)
    gpu.close()


if __name__ == "__main__":
    main()
