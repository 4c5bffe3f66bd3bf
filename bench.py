#!/usr/bin/env python3
"""Throughput of the integrator hot path on one MI355X: the product-sphere workload (a seeded 2,045,952-triangle
displaced sphere in 16 Disney parts under a 2048 x 1024 sky with a 0.5-degree sun at 1e4 x; 1920 x 1080,
256 spp, max depth 12), rendered by the gfx950 wavefront path tracer through the C-ABI.

A step is one pe_render call of --spl samples for every pixel of the 1080p frame (default 128: 265 M samples
through a pool of 2^26 path ids regenerated as paths end, about 16 GB of path state plus 4 GB of per-sample
results), so two steps are the workload's 256 spp. Prints one JSON line:
  value            Mrays/s: extension + shadow rays the kernels traced, over the timed wall time
  spp_per_s        1080p frames' worth of samples per second;  msamples_per_s: the reference's "M samples/s"
  spp_per_s_by_call  the same workload issued as calls of 1, 8 and 32 spp (RenderSession's batch sizes)
  roofline         the dominant kernel: its algorithmic bytes per launch (per-branch model, counts from the
                   kernels' own counters) over its HIP-event time, against 8 TB/s HBM3E
  rooflines        the same for every wavefront kernel; closest hit also against its measured HBM traffic
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
DEFAULT_POOL = 1 << 26            # pe_settings.paths_in_flight's default
PATH_BYTES = 4 * 16 + 16 + 6 * 16 + 8 + 3 * 16   # device bytes per path id (one shadow sample)
RESULT_BYTES = 16                 # per window sample
BF16_PEAK_TFLOPS = 2500.0  # MI355X dense bf16 MFMA
# Guiding MLP per query: 2 x (160*64 + 4*64*64) multiply-adds (the padded layer shapes the MFMAs compute).
GUIDE_FLOPS_PER_QUERY = 2 * (160 * 64 + 4 * 64 * 64)
NODE_BYTES, TRI_BYTES, RAY_IN_BYTES, HIT_OUT_BYTES = 128, 48, 32, 16


# Algorithmic bytes per unit of each wavefront kernel (the layouts of csrc/integrator.hip and pe_traverse.h):
QUEUE_ENTRY = 48            # ray queue record read by shade: origin | path id, direction | tag, hit record
PATH_STATE_READ = 80        # continuation entry: throughput, radiance, ids, previous point and normal
PEND_RW = 32                # held-back NEE slot read and cleared, per shadow sample
HIT_VERTEX = 32 + 3 * (16 + 16 + 8) + 336   # TriShade, three vertices (position, normal, uv), GpuMaterial
ENV_LOOKUP = 48             # bilinear environment texel fetch
NEE_SAMPLE = 100 + 32 + 16  # light / environment sample tables, shadow-queue record, held-back slot
BSDF_TABLES = 32            # energy-table lookups of a BSDF sample
CONTINUE_WRITE = 112        # next ray, throughput, radiance, ids, previous point and normal
END_WRITE = 20              # per-sample result and the free-list entry
REGEN_PER_SAMPLE = 4 + 32 + 16   # free-list entry, camera ray record, per-sample result of a deferred path
FILM_PER_PIXEL, FILM_PER_SAMPLE = 20 + 20, 16


def algorithmic_bytes(c: dict) -> dict:
    """Bytes per kernel for the counts c of one or more renders: extension rays E, camera entries (samples) N,
    hits H, shadow rays R, node / triangle visits of closest and any hit, film pixels P."""
    cont = c["E"] - c["N"]
    miss = c["E"] - c["H"]
    return {
        "intersect": c["E"] * (RAY_IN_BYTES + HIT_OUT_BYTES) + c["nodes"] * NODE_BYTES + c["tris"] * TRI_BYTES,
        "shade": (c["E"] * QUEUE_ENTRY + cont * (PATH_STATE_READ + PEND_RW) + c["H"] * (HIT_VERTEX + BSDF_TABLES)
                  + miss * ENV_LOOKUP + c["R"] * NEE_SAMPLE + cont * CONTINUE_WRITE + c["N"] * END_WRITE),
        "shadow": c["R"] * RAY_IN_BYTES + c["snodes"] * NODE_BYTES + c["stris"] * TRI_BYTES,
        "regen": c["N"] * REGEN_PER_SAMPLE,
        "film": c["P"] * FILM_PER_PIXEL + c["N"] * FILM_PER_SAMPLE,
    }


def host_cpu() -> dict:
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0))}


def traffic_for(kernel: str, config: dict) -> tuple[int | None, str]:
    """Calibrated HBM bytes per launch from profiles/<kernel>_traffic.json, when it was measured on this same
    configuration (workload, samples per call, path pool); otherwise None."""
    f = ROOT / "profiles" / f"{kernel}_traffic.json"
    if not f.exists():
        return None, "no measurement"
    try:
        d = json.loads(f.read_text())
    except Exception:  # noqa: BLE001
        return None, "unreadable"
    if d.get("config") != config:
        return None, f"measured on {d.get('config')}, not this configuration"
    return int(d["bytes_per_launch"]), f"rocprof FETCH_SIZE + WRITE_SIZE, {f.name}"


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
    host = host_cpu()
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
                "msamples_per_s": msps, **host,
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
    return {"value": (rays[0] + rays[1]) / secs / 1e6, "unit": "Mrays/s", "cores": cores, "kind": "port", **host,
            "sample": f"oracle restatement, 1 spp on rows 0..{rows} of the 1920x1080 frame, {cores} threads"}


def serving_run(dev, main_scene, seconds: float, rate: float, seed: int = 3, drain_s: float = 30.0) -> dict:
    """The serving stream: Poisson arrivals at `rate` jobs/s for `seconds`, each job a scene drawn from a resident
    pool (C0 box + icosphere, C1 displaced spheres of 0.1 M / 2 M / 5 M triangles, the C2 studio with its LED
    strip, dispersive prism, absorbing liquid and subsurface block), a film of 256^2, 512^2, 1920x1080 or 2048^2,
    and log-uniform 16..1024 spp with a 16-spp preview. The stream offers several times what one GPU renders, so
    completions queue by design; previews go first. After the last arrival the server gets `drain_s` for the
    outstanding previews, then the unfinished jobs are cancelled. Returns p50/p99 time to preview (overall and
    per 10-s window of arrivals) and to completion, offered and served load, and the server's memory counters."""
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
    next_log = 10.0
    while k < len(plan):
        now = srv.clock() - t0
        while k < len(plan) and plan[k][0] <= now:
            _, name, gs, view, w, h, spp = plan[k]
            cam = camera_perspective(view.position, view.target, view.up, view.vfov, w / h)
            st = settings(max_bounces=view.max_bounces, seed=12345)
            ids.append(srv.submit(gs, cam, st, w, h, spp, 16))
            k += 1
        if now >= next_log:
            log(f"serving: {k}/{len(plan)} jobs submitted at t={now:.0f} s; {srv.stats()['admitted']} admitted")
            next_log += 10.0
        time.sleep(0.002)
    # Drain: until every preview has landed (or drain_s), then cancel what is left.
    deadline = time.time() + drain_s
    while time.time() < deadline:
        if all(srv.poll(j).t_preview >= 0 or srv.poll(j).ended for j in ids):
            break
        time.sleep(0.05)
    wall = srv.clock() - t0
    mem = srv.stats()
    for j in ids:
        if not srv.poll(j).ended:
            srv.cancel(j)
    stats = [srv.wait(j, -1) for j in ids]
    for j in ids:
        srv.release(j)
    srv.close()
    done = [s for s in stats if s.state == capi.PE_JOB_DONE]
    prev = np.array([s.t_preview - s.t_submit for s in stats if s.t_preview >= 0])
    comp = np.array([s.t_done - s.t_submit for s in done])
    served = sum(s.samples_done * w * h for s, (_, _, _, _, w, h, _) in zip(stats, plan))
    rays = sum(s.extension_rays + s.shadow_rays for s in stats)
    pct = lambda a, q: round(float(np.percentile(a, q)), 3) if len(a) else None  # noqa: E731
    windows = []
    for w0 in np.arange(0.0, seconds, 10.0):
        sel = [s.t_preview - s.t_submit for s, p in zip(stats, plan) if w0 <= p[0] < w0 + 10.0 and s.t_preview >= 0]
        missing = sum(1 for s, p in zip(stats, plan) if w0 <= p[0] < w0 + 10.0 and s.t_preview < 0)
        windows.append({"t": [float(w0), float(w0 + 10.0)], "jobs": len(sel) + missing,
                        "preview_p50_s": pct(sel, 50), "preview_p99_s": pct(sel, 99), "no_preview": missing})
    return {"workload": f"Poisson {rate:g} jobs/s for {seconds:g} s (seed {seed}); scenes {[r[0] for r in resident]}; "
                        f"films 256^2/512^2/1080p/2048^2; spp log-uniform 16-1024; preview 16 spp",
            "slots": 4, "chunk_spp": 16, "jobs": len(ids), "completed": len(done),
            "previews": int(len(prev)), "preview_p50_s": pct(prev, 50), "preview_p99_s": pct(prev, 99),
            "preview_p99_by_10s_window": windows,
            "done_p50_s": pct(comp, 50), "done_p99_s": pct(comp, 99),
            "offered_msamples_per_s": round(offered / seconds / 1e6, 1),
            "served_msamples_per_s": round(served / wall / 1e6, 1),
            "mrays_per_s": round(rays / wall / 1e6, 1), "wall_s": round(wall, 1),
            "memory": {"budget_gb": round(mem["memory_budget"] / 1e9, 1), "used_gb": round(mem["memory_used"] / 1e9, 1),
                       "lanes_gb": round(mem["lane_bytes"] / 1e9, 1), "scenes_parked": mem["scenes_parked"],
                       "scenes_restored": mem["scenes_restored"]}}


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
    guide = Guide(dev, gs, seed=seed, learning_rate=1e-3, steps_per_launch=16, train_launches=3, guided_bounces=1)
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
    ap.add_argument("--spl", type=int, default=128, help="samples per pixel per step (one pe_render call)")
    ap.add_argument("--pool", type=int, default=0, help="path ids in flight (0: the library default, 2^26)")
    ap.add_argument("--no-call-sweep", dest="call_sweep", action="store_false",
                    help="skip the spp/s of 1-, 8- and 32-spp calls")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="target duration of the CPU-baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--small", action="store_true", help="reduced scene (smoke runs only; not the workload)")
    ap.add_argument("--serving-seconds", type=float, default=60.0,
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
    pool = args.pool if args.pool > 0 else DEFAULT_POOL
    st = settings(max_bounces=view.max_bounces, seed=12345, samples_per_launch=args.spl, paths_in_flight=pool)
    film = dev.film(view.width, view.height)
    pixels = view.width * view.height

    first = 0
    for _ in range(args.warmup):
        render(dev, gs, cam, st, film, first, args.spl)
        first += args.spl
    # Counts of one step (untimed; the counting kernels are slower): traversal visits, hits, deepest stack.
    st_count = settings(max_bounces=view.max_bounces, seed=12345, samples_per_launch=args.spl, paths_in_flight=pool)
    st_count.count_visits = 1
    scratch = dev.film(view.width, view.height)
    rc = render(dev, gs, cam, st_count, scratch, first, args.spl)
    del scratch
    per_ext = {"nodes": rc.node_visits / max(1, rc.extension_rays), "tris": rc.tri_visits / max(1, rc.extension_rays),
               "hits": rc.closest_hits / max(1, rc.extension_rays)}
    per_shadow = {"nodes": rc.shadow_node_visits / max(1, rc.shadow_rays),
                  "tris": rc.shadow_tri_visits / max(1, rc.shadow_rays)}
    log(f"per closest-hit ray: {per_ext['nodes']:.1f} nodes, {per_ext['tris']:.1f} triangles, {per_ext['hits']:.3f} "
        f"hit; per shadow ray: {per_shadow['nodes']:.1f} nodes, {per_shadow['tris']:.1f} triangles; deepest "
        f"stack {rc.max_stack_depth}")

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
    counts = {"E": ext, "N": samples, "H": ext * per_ext["hits"], "R": shadow, "nodes": ext * per_ext["nodes"],
              "tris": ext * per_ext["tris"], "snodes": shadow * per_shadow["nodes"],
              "stris": shadow * per_shadow["tris"], "P": pixels * args.steps}
    algo = algorithmic_bytes(counts)
    bench_config = {"workload": workload, "samples_per_call": args.spl, "paths_in_flight": pool}
    rooflines = {}
    for kernel, total_bytes in algo.items():
        k = kt.get(kernel, {"launches": 0, "ms": 0.0})
        if k["launches"] == 0 or k["ms"] <= 0:
            continue
        per_launch = total_bytes / k["launches"]
        avg_s = k["ms"] * 1e-3 / k["launches"]
        gbs = per_launch / avg_s / 1e9
        traffic, basis = traffic_for(kernel, bench_config)
        entry = {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                 "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": traffic, "algorithmic_bytes_per_launch": round(per_launch),
                 "launches": k["launches"], "avg_launch_ms": round(avg_s * 1e3, 4), "traffic_basis": basis}
        if traffic is not None:
            entry["measured_frac"] = round(traffic / avg_s / 1e9 / HBM_PEAK_GBS, 4)
        if gbs > HBM_PEAK_GBS:
            # Node and triangle re-reads are served by L2 and the 256 MB Infinity Cache: the algorithmic rate is
            # a cache-served rate, not an HBM fraction; the fraction is the measured one, or none.
            entry["cache_served_gbs"] = entry.pop("achieved")
            entry["achieved"] = round(traffic / avg_s / 1e9, 1) if traffic is not None else None
            entry["frac"] = entry.get("measured_frac")
            entry["basis"] = "measured HBM traffic (algorithmic bytes exceed the HBM peak: cache-served)"
        else:
            entry["basis"] = "algorithmic bytes per launch / HIP-event time"
        rooflines[kernel] = entry
    timed = {k: v for k, v in kt.items() if k in rooflines}
    dominant = max(timed.items(), key=lambda kv: kv[1]["ms"])[0] if timed else None
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
                   "paths_in_flight": pool, "max_depth": view.max_bounces, "rr_depth": 3, "seed": 12345,
                   "filter": "box 1px"},
        "spp_per_s": round(args.steps * args.spl / wall, 3),
        "msamples_per_s": round(samples / wall / 1e6, 2),
        "extension_rays": ext,
        "shadow_rays": shadow,
        "rays_per_sample": round((ext + shadow) / samples, 4),
        "path_state_gb": round((pool * PATH_BYTES + pixels * args.spl * RESULT_BYTES) / 1e9, 2),
        "kernels_ms": {k: round(v["ms"], 2) for k, v in kt.items()},
        "dominant_kernel": dominant,
        "roofline": dict(kernel=dominant, **rooflines[dominant]) if dominant else None,
        "rooflines": rooflines,
        "deepest_traversal_stack": rc.max_stack_depth,
    }
    if args.call_sweep:
        result["spp_per_s_by_call"] = call_sweep(dev, gs, cam, view, pool, first)
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


def call_sweep(dev, gs, cam, view, pool, first) -> dict:
    """spp/s of the same workload issued as pe_render calls of 1, 8 and 32 spp (RenderSession's Interactive passes,
    HipDevice's cancellable chunks, a mid-size Final batch), 64 spp of each."""
    from photon_hip import render, settings

    out = {}
    film = dev.film(view.width, view.height)
    for spl in (1, 8, 32):
        st = settings(max_bounces=view.max_bounces, seed=12345, samples_per_launch=spl, paths_in_flight=pool)
        render(dev, gs, cam, st, film, first, spl)
        first += spl
        dev.synchronize()
        t0 = time.perf_counter()
        for _ in range(64 // spl):
            render(dev, gs, cam, st, film, first, spl)
            first += spl
        dev.synchronize()
        out[str(spl)] = round(64 / (time.perf_counter() - t0), 1)
        log(f"calls of {spl} spp: {out[str(spl)]} spp/s")
    return out


if __name__ == "__main__":
    sys.exit(main())
