#!/usr/bin/env python3
"""Throughput of the integrator hot path on one MI355X: the product-sphere workload (a seeded 2,045,952-triangle
displaced sphere in 16 Disney parts under a 2048 x 1024 sky with a 0.5-degree sun at 1e4 x; 1920 x 1080,
256 spp, max depth 12), rendered by the gfx950 wavefront path tracer through the C-ABI.

A step is one wavefront launch of --spl samples for every pixel of the 1080p frame (default 128: 265 M paths,
about 60 GB of path state, in flight at once; deeper launches amortise the short tail bounces), so the default 2
timed steps are exactly the workload's 256 spp. Prints one JSON line:
  value            Mrays/s: extension + shadow rays the kernels traced, over the timed wall time
  spp_per_s        1080p frames' worth of samples per second;  msamples_per_s: the reference's "M samples/s"
  roofline         closest-hit kernel: algorithmic bytes per launch (ray in/out + node and triangle bytes of the
                   measured visits) over its HIP-event time, against 8 TB/s HBM3E
  roofline_shade   the dominant kernel (shade): rocprof's calibrated HBM bytes per launch over its HIP-event time
  cpu_baseline     the reference's own CPU renderer (oracle/_ref, its own compiler flags) on this host's cores,
                   on a bounded sample of the same workload
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "photon-engine_amd" / "python"))

HBM_PEAK_GBS = 8000.0
BF16_PEAK_TFLOPS = 2500.0  # MI355X dense bf16 MFMA
# Guiding MLP per query: 2 x (160*64 + 4*64*64) multiply-adds (the padded layer shapes the MFMAs compute).
GUIDE_FLOPS_PER_QUERY = 2 * (160 * 64 + 4 * 64 * 64)
NODE_BYTES, TRI_BYTES, RAY_IN_BYTES, HIT_OUT_BYTES = 128, 48, 32, 16


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def kernel_times(dev) -> dict:
    from photon_hip import capi

    out = (capi.KernelStats * 8)()
    n = C.c_int()
    capi.check("pe_profile_read", capi.load().pe_profile_read(dev.h, out, 8, C.byref(n)))
    return {out[k].name.decode(): {"launches": int(out[k].launches), "ms": float(out[k].total_ms)} for k in range(n.value)}


def cpu_baseline(scene, view, rays_per_sample: float, target_s: float) -> dict | None:
    """The reference renderer (or, where it was not built, the oracle restatement) on this host's cores."""
    try:
        from oracle import bindings
    except Exception as e:  # noqa: BLE001
        log(f"cpu baseline unavailable: {e}")
        return None
    cores = len(os.sched_getaffinity(0))
    small = view
    if bindings.reference(strict=False) is not None:
        t = time.time()
        ref = bindings.RefScene(scene, strict=False)
        log(f"cpu baseline: reference scene + BVH built in {time.time() - t:.1f} s")
        _, _, t1 = ref.render(small, 0, 1, cores)
        count = max(1, min(64, int(target_s / max(t1, 1e-3))))
        _, _, secs = ref.render(small, 1, count, cores)
        samples = small.width * small.height * count
        msps = samples / secs / 1e6
        return {"value": msps * rays_per_sample, "unit": "Mrays/s", "cores": cores, "kind": "reference",
                "msamples_per_s": msps,
                "sample": (f"Renderer::renderSamples of the same 1920x1080 product-sphere scene, samples 1..{count} "
                           f"({secs:.1f} s on {cores} threads, reference flags -O3 -ffast-math -mavx2 -mfma); "
                           f"M samples/s x {rays_per_sample:.3f} rays/sample counted by the GPU kernels on the "
                           f"same sample sequence")}
    from photon_hip import camera_perspective, settings

    o = bindings.OracleScene(scene)
    cam = camera_perspective(view.position, view.target, view.up, view.vfov, view.aspect)
    st = settings(max_bounces=view.max_bounces, seed=12345)
    rows = max(1, view.height // 16)
    _, secs, rays = o.render(cam, st, view.width, view.height, 0, 1, threads=cores, rows=(0, rows))
    return {"value": (rays[0] + rays[1]) / secs / 1e6, "unit": "Mrays/s", "cores": cores, "kind": "port",
            "sample": f"oracle restatement, 1 spp on rows 0..{rows} of the 1920x1080 frame, {cores} threads"}


def serving_run(dev, main_scene, seconds: float, rate: float, seed: int = 3, drain_s: float = 150.0) -> dict:
    """The serving stream: Poisson arrivals at `rate` jobs/s for `seconds`, each job a scene drawn from a resident
    pool (C0 box + icosphere, C1 displaced spheres of 0.1 M / 2 M / 5 M triangles, the C2 studio with its LED
    strip, dispersive prism, absorbing liquid and subsurface block), a film of 256^2, 512^2, 1920x1080 or 2048^2, and log-uniform 16..1024 spp with a 16-spp preview.
    Returns p50/p99 time to preview and to completion, offered and served load."""
    import numpy as np

    from photon_hip import camera_perspective, configs, settings
    from photon_hip import capi
    from photon_hip.serve import Server

    t = time.time()
    pool = [("c0_box_icosphere_20k", *configs.box_icosphere(level=5)),
            ("c1_sphere_100k", *configs.product_sphere(lon=256, lat=200, env_width=1024, env_height=512)),
            ("c1_sphere_5m", *configs.product_sphere(lon=1600, lat=1600)),
            ("c2_lights_media_204k", *configs.lights_media_studio())]
    resident = [(name, dev.upload(sc), view) for name, sc, view in pool]
    resident.append(("c1_sphere_2m", main_scene[0], main_scene[1]))
    log(f"serving: {len(resident)} resident scenes ready in {time.time() - t:.1f} s")
    rng = np.random.default_rng(seed)
    sizes = [(256, 256), (512, 512), (1920, 1080), (2048, 2048)]
    n_jobs = int(rng.poisson(rate * seconds))
    arrivals = np.sort(rng.uniform(0.0, seconds, n_jobs))
    plan = []
    for a in arrivals:
        name, gs, view = resident[int(rng.integers(len(resident)))]
        w, h = sizes[int(rng.integers(len(sizes)))]
        spp = int(round(math.exp(rng.uniform(math.log(16), math.log(1024)))))
        plan.append((float(a), name, gs, view, w, h, spp))
    srv = Server(dev, slots=4, chunk_spp=16, max_admitted=4096)
    offered = sum(w * h * spp for _, _, _, _, w, h, spp in plan)
    ids = []
    t0 = srv.clock()
    k = 0
    next_log = 5.0
    while k < len(plan):
        now = srv.clock() - t0
        while k < len(plan) and plan[k][0] <= now:
            _, name, gs, view, w, h, spp = plan[k]
            cam = camera_perspective(view.position, view.target, view.up, view.vfov, w / h)
            st = settings(max_bounces=view.max_bounces, seed=12345)
            ids.append((srv.submit(gs, cam, st, w, h, spp, 16), w * h * spp))
            k += 1
        if now >= next_log:
            log(f"serving: {k}/{len(plan)} jobs submitted at t={now:.0f} s")
            next_log += 5.0
        time.sleep(0.002)
    deadline = time.time() + drain_s
    stats = []
    for i, (jid, work) in enumerate(ids):
        st = srv.wait(jid, max(0, int((deadline - time.time()) * 1000)))
        if not st.ended:
            srv.cancel(jid)
            st = srv.wait(jid, -1)
        stats.append((st, work))
        if (i + 1) % 50 == 0:
            log(f"serving: {i + 1}/{len(ids)} jobs ended")
    wall = srv.clock() - t0
    for jid, _ in ids:
        srv.release(jid)
    srv.close()
    done = [s for s, _ in stats if s.state == capi.PE_JOB_DONE]
    prev = np.array([s.t_preview - s.t_submit for s, _ in stats if s.t_preview >= 0])
    comp = np.array([s.t_done - s.t_submit for s in done])
    served = sum(s.samples_done * (w * h * spp) / spp for (s, _), (_, _, _, _, w, h, spp) in zip(stats, plan))
    rays = sum(s.extension_rays + s.shadow_rays for s, _ in stats)
    pct = lambda a, q: round(float(np.percentile(a, q)), 3) if len(a) else None  # noqa: E731
    return {"workload": f"Poisson {rate:g} jobs/s for {seconds:g} s (seed {seed}); scenes {[r[0] for r in resident]}; "
                        f"films 256^2/512^2/1080p/2048^2; spp log-uniform 16-1024; preview 16 spp",
            "slots": 4, "chunk_spp": 16, "jobs": len(ids), "completed": len(done),
            "preview_p50_s": pct(prev, 50), "preview_p99_s": pct(prev, 99),
            "done_p50_s": pct(comp, 50), "done_p99_s": pct(comp, 99),
            "offered_msamples_per_s": round(offered / seconds / 1e6, 1),
            "served_msamples_per_s": round(served / wall / 1e6, 1),
            "mrays_per_s": round(rays / wall / 1e6, 1), "wall_s": round(wall, 1)}


def guiding_run(dev, budget_spp: int, ref_spp: int, spl: int = 16, seed: int = 13) -> dict:
    """Equal-time comparison on the learned-guiding workload (the 493k-triangle room lit through a 5 cm door gap,
    1920x1080, depth 10): the wall time of a plain render of `budget_spp` spp is the budget; the guided render
    (guide created, trained online during its first launches, then guiding) adds launches of `spl` spp into
    its film until it has used the same time. Both are compared with a plain `ref_spp`-spp image of independent
    samples by relMSE."""
    import numpy as np

    from photon_hip import camera_perspective, configs, settings, render
    from photon_hip.guide import Guide, render_guided

    sc, view = configs.slit_room()
    t = time.time()
    gs = dev.upload(sc)
    log(f"guiding: {sc.num_triangles} triangles uploaded in {time.time() - t:.1f} s")
    cam = camera_perspective(view.position, view.target, view.up, view.vfov, view.aspect)
    st = settings(max_bounces=view.max_bounces, seed=12345, samples_per_launch=spl)

    def mean(film):
        s, c, _, _ = film.download()
        return s / np.maximum(c, 1)[..., None]

    def relmse(x, r):
        return float(np.mean((x - r) ** 2 / (r ** 2 + 0.01)))

    ref_film = dev.film(view.width, view.height)
    t = time.time()
    render(dev, gs, cam, settings(max_bounces=view.max_bounces, seed=12345), ref_film, 1 << 20, ref_spp)
    dev.synchronize()
    log(f"guiding: reference {ref_spp} spp in {time.time() - t:.1f} s")
    ref = mean(ref_film)
    plain_film = dev.film(view.width, view.height)
    render(dev, gs, cam, st, plain_film, 1 << 22, spl)  # warm-up (kernels, buffers), other samples
    plain_film = dev.film(view.width, view.height)
    dev.synchronize()
    t0 = time.perf_counter()
    render(dev, gs, cam, st, plain_film, 0, budget_spp)
    dev.synchronize()
    budget = time.perf_counter() - t0
    plain = mean(plain_film)
    guided_film = dev.film(view.width, view.height)
    dev.synchronize()
    t0 = time.perf_counter()
    guide = Guide(dev, gs, seed=seed, learning_rate=2e-3, steps_per_launch=16, train_launches=3, guided_bounces=1)
    spp = 0
    while time.perf_counter() - t0 < budget:
        render_guided(dev, gs, guide, cam, st, guided_film, spp, spl)
        dev.synchronize()
        spp += spl
    used = time.perf_counter() - t0
    stats = guide.stats()
    guided = mean(guided_film)
    # MFMA roofline of the fused inference kernel: a separately profiled guided render with the trained guide.
    dev.synchronize()
    dev._lib.pe_profile_enable(dev.h, 1)
    before = guide.stats()["inferred"]
    render_guided(dev, gs, guide, cam, st, dev.film(view.width, view.height), 1 << 21, 2 * spl)
    dev.synchronize()
    kt = kernel_times(dev)
    dev._lib.pe_profile_enable(dev.h, 0)
    queries = guide.stats()["inferred"] - before
    guide.close()
    flops = queries * GUIDE_FLOPS_PER_QUERY
    infer = kt.get("guide_infer", {"launches": 0, "ms": 0.0})
    tflops = flops / (infer["ms"] * 1e-3) / 1e12 if infer["ms"] > 0 else 0.0
    e_p, e_g = relmse(plain, ref), relmse(guided, ref)
    return {"workload": f"slit room {sc.num_triangles} tris, {view.width}x{view.height}, depth {view.max_bounces}, "
                        f"5 cm door gap; reference {ref_spp} spp plain",
            "budget_s": round(budget, 3), "guided_s": round(used, 3), "plain_spp": budget_spp, "guided_spp": spp,
            "plain_relmse": round(e_p, 5), "guided_relmse": round(e_g, 5), "ratio": round(e_g / e_p, 3),
            "train_steps": stats["steps"], "records": stats["records"], "final_loss": round(stats["last_loss"], 3),
            "kernels_ms": {k: round(v["ms"], 2) for k, v in kt.items()},
            "roofline": {"kernel": "guide_infer", "bound": "mfma", "achieved": round(tflops, 1),
                         "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": round(tflops / BF16_PEAK_TFLOPS, 4),
                         "traffic": None, "flops_per_query": GUIDE_FLOPS_PER_QUERY, "queries": queries,
                         "launches": infer["launches"],
                         "avg_launch_ms": round(infer["ms"] / max(1, infer["launches"]), 4)}}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--spl", type=int, default=128, help="samples per pixel per step (one wavefront launch)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="target duration of the CPU-baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--small", action="store_true", help="reduced scene (smoke runs only; not the workload)")
    ap.add_argument("--serving-seconds", type=float, default=5.0,
                    help="length of the serving stream's arrivals (0: no serving run)")
    ap.add_argument("--serving-rate", type=float, default=20.0, help="serving arrivals per second")
    ap.add_argument("--guiding-spp", type=int, default=256,
                    help="plain spp whose wall time is the equal-time budget of the guiding comparison (0: skip)")
    ap.add_argument("--guiding-ref-spp", type=int, default=2048)
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="only run a plain device-to-device copy of this many MiB (HBM counter calibration)")
    args = ap.parse_args()
    if args.copy_mib > 0:
        import torch

        n = args.copy_mib * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_mib": args.copy_mib, "bytes_read": n * 4, "bytes_written": n * 4}), flush=True)
        return 0
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus}); "
              f"a frame shards by sample range across GPUs with one film reduction, which is not implemented.",
              file=sys.stderr)
        return 2

    from photon_hip import Device, camera_perspective, configs, render, settings

    t = time.time()
    if args.small:
        scene, view = configs.product_sphere(width=480, height=270, lon=256, lat=200, env_width=512, env_height=256)
        workload = "product_sphere_small (smoke only)"
    else:
        scene, view = configs.product_sphere()
        workload = "product_sphere_2M_1080p"
    log(f"scene generated in {time.time() - t:.1f} s: {scene.num_triangles} triangles")
    dev = Device(0)
    t = time.time()
    gs = dev.upload(scene)
    log(f"uploaded (BVH build + copies) in {time.time() - t:.1f} s: {gs.stats()}")
    cam = camera_perspective(view.position, view.target, view.up, view.vfov, view.aspect)
    st = settings(max_bounces=view.max_bounces, seed=12345, samples_per_launch=args.spl)
    film = dev.film(view.width, view.height)
    pixels = view.width * view.height

    first = 0
    for _ in range(args.warmup):
        render(dev, gs, cam, st, film, first, args.spl)
        first += args.spl
    # Traversal visit counts of one step (untimed, counting kernels are slower).
    st_count = settings(max_bounces=view.max_bounces, seed=12345, samples_per_launch=args.spl)
    st_count.count_visits = 1
    scratch = dev.film(view.width, view.height)
    rc = render(dev, gs, cam, st_count, scratch, first, args.spl)
    ext_nodes = rc.node_visits / max(1, rc.extension_rays)
    ext_tris = rc.tri_visits / max(1, rc.extension_rays)
    sh_nodes = rc.shadow_node_visits / max(1, rc.shadow_rays)
    sh_tris = rc.shadow_tri_visits / max(1, rc.shadow_rays)
    del scratch
    log(f"visits per closest-hit ray: {ext_nodes:.1f} nodes, {ext_tris:.1f} triangles; "
        f"per shadow ray: {sh_nodes:.1f} nodes, {sh_tris:.1f} triangles")

    lib_prof = dev._lib.pe_profile_enable
    dev.synchronize()
    lib_prof(dev.h, 1)
    ext = shadow = 0
    t0 = time.perf_counter()
    for k in range(args.steps):
        r = render(dev, gs, cam, st, film, first, args.spl)
        first += args.spl
        ext += r.extension_rays
        shadow += r.shadow_rays
        if (k + 1) % 16 == 0:
            log(f"step {k + 1}/{args.steps}")
    dev.synchronize()
    wall = time.perf_counter() - t0
    kt = kernel_times(dev)
    lib_prof(dev.h, 0)

    samples = pixels * args.steps * args.spl
    mrays = (ext + shadow) / wall / 1e6
    isect = kt.get("intersect", {"launches": 0, "ms": 0.0})
    bytes_per_ray = RAY_IN_BYTES + HIT_OUT_BYTES + ext_nodes * NODE_BYTES + ext_tris * TRI_BYTES
    isect_gbs = (ext * bytes_per_ray) / (isect["ms"] * 1e-3) / 1e9 if isect["ms"] > 0 else 0.0
    dominant = max(kt.items(), key=lambda kv: kv[1]["ms"])[0] if kt else "intersect"
    traffic = None
    tfile = ROOT / "profiles" / "intersect_traffic.json"
    if tfile.exists():
        try:
            traffic = json.loads(tfile.read_text()).get("bytes_per_launch")
        except Exception:  # noqa: BLE001
            traffic = None
    result = {
        "metric": "Mrays/s",
        "value": round(mrays, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded generator: displaced sphere, sky + sun environment)",
        "config": {"workload": workload, "triangles": scene.num_triangles, "width": view.width,
                   "height": view.height, "spp": args.steps * args.spl, "samples_per_step": args.spl,
                   "max_depth": view.max_bounces, "rr_depth": 3, "seed": 12345, "filter": "box 1px"},
        "spp_per_s": round(args.steps * args.spl / wall, 3),
        "msamples_per_s": round(samples / wall / 1e6, 2),
        "extension_rays": ext,
        "shadow_rays": shadow,
        "rays_per_sample": round((ext + shadow) / samples, 4),
        "kernels_ms": {k: round(v["ms"], 2) for k, v in kt.items()},
        "dominant_kernel": dominant,
        "roofline": {"kernel": "intersect", "bound": "hbm", "achieved": round(isect_gbs, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(isect_gbs / HBM_PEAK_GBS, 4),
                     "traffic": traffic,
                     "bytes_per_ray": round(bytes_per_ray, 1), "node_visits_per_ray": round(ext_nodes, 2),
                     "tri_visits_per_ray": round(ext_tris, 2),
                     "launches": isect["launches"],
                     "avg_launch_ms": round(isect["ms"] / max(1, isect["launches"]), 4)},
    }
    # The dominant kernel, shade, has no single algorithmic byte count (its gathers depend on the branch each
    # path takes); its line gives the HBM traffic rocprof measured per launch over the HIP-event time per launch.
    shade = kt.get("shade", {"launches": 0, "ms": 0.0})
    sfile = ROOT / "profiles" / "shade_traffic.json"
    if shade["launches"] > 0 and shade["ms"] > 0 and sfile.exists():
        sbytes = json.loads(sfile.read_text()).get("bytes_per_launch")
        shade_ms = shade["ms"] / shade["launches"]
        gbs = sbytes / (shade_ms * 1e-3) / 1e9
        result["roofline_shade"] = {"kernel": "shade", "bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS,
                                    "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": sbytes,
                                    "basis": "measured HBM bytes per launch (rocprof, calibrated) / HIP-event time",
                                    "launches": shade["launches"], "avg_launch_ms": round(shade_ms, 4)}
    if args.serving_seconds > 0:
        log("serving stream ...")
        result["serving"] = serving_run(dev, (gs, view), args.serving_seconds, args.serving_rate)
    if args.guiding_spp > 0:
        log("guiding (equal time) ...")
        result["guiding"] = guiding_run(dev, args.guiding_spp, args.guiding_ref_spp)
    if not args.no_cpu:
        log("cpu baseline ...")
        cb = cpu_baseline(scene, view, (ext + shadow) / samples, args.cpu_seconds)
        result["cpu_baseline"] = cb
        if cb:
            result["gpu_vs_cpu"] = round(mrays / cb["value"], 1) if cb["value"] > 0 else None
    else:
        result["cpu_baseline"] = None
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
