"""Benchmark: plain MIS path tracing of the large-mesh scene on one MI355X.

Workload (one step = one full render): level-8 displaced icosphere (1,310,720 triangles,
rough GGX conductor patches) on a Lambertian ground under a 512x256 sky map, 1024x1024,
64 spp stratified, max depth 6, Russian roulette from depth 3. Synthetic procedural input,
built once; the scene (BVH + tables) and the film stay resident in HBM.

Prints one JSON line with samples/s, the HBM roofline of the dominant kernel (BVH traversal
of extension rays; algorithmic bytes from the library's traversal work counters, duration
from HIP events recorded around every launch on the library's stream) and the CPU oracle
timed on this host on a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# algorithmic bytes of the extension-traversal kernel
RAY_IO_BYTES = 4 + 4 + 32 + 16 + 4   # queue entry, flags, ray (2 x float4), hit record, shade-queue entry
NODE_BYTES = 64                      # one 4-wide node visit reads one compressed node: a 16-byte header
                                     # (origin, exponents), 32 bytes of 8-bit child bounds, 4 child refs
TRI_BYTES = 36                       # three float3 vertices
SEGMENT_IO_BYTES = 292               # per path segment: ray read 32, hit write + read 32, path state read +
                                     # write 128, shadow ray write + read 64, occlusion 4, next ray 32


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--spp", type=int, default=64)
    ap.add_argument("--size", type=int, default=1024)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--copy-mb", type=int, default=0,
                    help="only run a plain device copy of this many MiB (scale reference for PMC byte counters)")
    ap.add_argument("--tuning-from", default="",
                    help="take the launch configuration from the \"tuning\" of an earlier bench line (a JSON "
                         "file) instead of autotuning: a profiled run then holds only the timed renders' launches")
    ap.add_argument("--pool", type=int, default=0,
                    help="path-pool slots for the benchmark render, overriding the tuning (0: keep it)")
    ap.add_argument("--no-guiding", action="store_true", help="skip the guided-vs-plain leg")
    ap.add_argument("--guide-size", type=int, default=512, help="guided-vs-plain leg: image size")
    ap.add_argument("--guide-ref-spp", type=int, default=32768,
                    help="guided-vs-plain leg: reference spp (eight independent pieces)")
    ap.add_argument("--only-guiding", action="store_true",
                    help="run only the guided-vs-plain legs and print their JSON object")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving-mix leg")
    ap.add_argument("--only-serving", action="store_true",
                    help="run only the serving-mix leg and print its JSON object")
    ap.add_argument("--serve-requests", type=int, default=48,
                    help="serving leg: requests of the Poisson stream (lambda 40/s, seed 4)")
    ap.add_argument("--no-volume", action="store_true", help="skip the volumetric leg")
    ap.add_argument("--volume-stride", type=int, default=1,
                    help="volumetric leg's CPU sample: every k-th 32x32 block of the 512x512 render")
    ap.add_argument("--only-volume", action="store_true",
                    help="run only the volumetric leg and print its JSON object (for profiling)")
    ap.add_argument("--cpu-stride", type=int, default=1,
                    help="CPU sample: every k-th 32x32 block in creation order (1: the whole render)")
    return ap.parse_args()


def host_cpu():
    """Threads of the CPU baseline and what the host reports. One GPU's share of the host is 16
    hardware threads (OMP_NUM_THREADS is set to it), whatever nproc shows for the whole machine."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        affinity = os.cpu_count() or 1
    share = int(os.environ.get("OMP_NUM_THREADS", "16") or 16)
    threads = max(1, min(share, affinity))
    return threads, {"host_nproc": os.cpu_count(), "host_affinity": affinity, "cpu_model": model,
                     "threads_rule": "one GPU's share of the host: min(OMP_NUM_THREADS (16 on the GPU host), "
                                     "CPUs in this process's affinity mask)"}


def cpu_baseline(sd, spp, stride):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O

    threads, host = host_cpu()
    o = O.OracleScene(sd)
    t0 = time.perf_counter()
    rgb, w, st, nblocks = o.render(spp=spp, threads=threads, blocks=(0, 0, stride))
    dt = time.perf_counter() - t0
    samples = nblocks * 32 * 32 * spp
    return {
        "value": samples / dt,
        "unit": "samples/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"the whole render" if stride == 1 else f"every {stride}th 32x32 block of the same render")
                  + f" ({nblocks} blocks, {samples} samples, {dt:.1f} s), reference schedule (blocks seeded (x, y), "
                  "LIFO pop)",
        "rays_per_s": float(st[0]) / dt,
        **host,
    }


def serving(ctx, big, n_req, lam=40.0, seed=4):
    """Serving mix: a Poisson stream (lambda req/s, seed 4) of requests, each drawing a scene
    uniformly from the Cornell box (plain MIS, depth 8), the large displaced sphere (plain MIS,
    depth 6) and the heterogeneous-medium Cornell box (VolPathTracing, 256^3 grid, depth 16), a
    resolution from {128,256,512,1024}^2 and spp from {4,16,64,256}. The three scenes stay
    resident; the server's own thread runs the loop (rtg_server_start); a producer thread submits
    at the arrival times and this thread waits for each request (rtg_wait); latency = submit to
    finish as the library measures it."""
    import threading

    import numpy as np
    from raytracer_amd import rtg, scenes

    sds = [scenes.cornell(), big[0], scenes.vol_cornell()]
    resident = [ctx.upload(sds[0]), big[1], ctx.upload(sds[2])]
    names = ["cornell", "displaced_sphere", "vol_cornell"]
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / lam, n_req)
    kinds = rng.integers(0, 3, n_req)
    sizes = rng.choice([128, 256, 512, 1024], n_req)
    spps = rng.choice([4, 16, 64, 256], n_req)
    reqs = []
    for g, k, sz, spp in zip(gaps, kinds, sizes, spps):
        sd = sds[int(k)]
        reqs.append((float(g), int(k), dict(sd.camera, width=int(sz), height=int(sz)),
                     dict(sd.render, spp=int(spp), seed=int(rng.integers(1, 1 << 30)))))
    offered = sum(c["width"] * c["height"] * r["spp"] for _, _, c, r in reqs)
    arrival_s = float(gaps.sum())
    with rtg.Server(ctx) as srv:
        srv.start()
        ids = []

        def producer():
            t = time.perf_counter()
            for g, k, cam, rp in reqs:
                t += g
                dt = t - time.perf_counter()
                if dt > 0:
                    time.sleep(dt)
                ids.append(srv.submit(resident[k], cam, rp))

        t0 = time.perf_counter()
        th = threading.Thread(target=producer)
        th.start()
        lat = []
        done = 0
        while done < n_req:
            if done >= len(ids):
                time.sleep(0.001)
                continue
            status, _, ms = srv.wait(ids[done], want_image=False)
            assert status == "done", status
            lat.append(ms)
            done += 1
        th.join()
        makespan = time.perf_counter() - t0
        st = srv.stats()
    lat = np.array(lat)
    for r in (resident[0], resident[2]):
        r.release()
    per_scene = {}
    for k in range(3):
        sel = kinds == k
        if sel.any():
            per_scene[names[k]] = {"requests": int(sel.sum()), "p50_ms": float(np.percentile(lat[sel], 50)),
                                   "p99_ms": float(np.percentile(lat[sel], 99))}
    return {
        "requests": n_req, "lambda_per_s": lam, "seed": seed, "scenes": names,
        "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
        "mean_ms": float(lat.mean()), "max_ms": float(lat.max()),
        "samples_per_s": st["samples"] / makespan, "makespan_s": makespan, "arrival_span_s": arrival_s,
        "offered_samples_per_s": offered / arrival_s,
        "max_running": st["max_running"], "film_bytes_peak": st["film_bytes_peak"],
        "completed": st["completed"], "iterations": st["iterations"], "graph_launches": st["graph_launches"],
        "per_scene": per_scene,
    }


def relmse(x, r, noise=None):
    """mean (x - r)^2 / (r^2 + 1e-2); `noise` (per-pixel variance of the reference, estimated from
    two independent halves) is subtracted so the reference's own noise does not count."""
    import numpy as np
    d = (x - r) ** 2
    if noise is not None:
        d = d - noise
    return float(np.mean(d / (r * r + 1e-2)))


GUIDE_FLOPS_PER_QUERY = 2 * (144 * 64 + 64 * 64 + 64 * 64 + 64 * 40)  # network multiply-adds x 2
MFMA_BF16_PEAK_TFLOPS = 2500.0


GLOSSY_REFERENCE = os.path.join(ROOT, "tests", "golden", "glossy_cornell_ref.npz")


def guiding_leg(ctx, scene, size, ref_spp, target=1e-3, ladder=(256, 1024, 4096), piece_spp=4096,
                max_spp=262144, methods=(("plain", 0), ("guided", 1), ("exit_only", 2))):
    """Plain vs guided on one scene, both on the GPU: error at a short ladder of spp counts and the
    time to relMSE <= target.

    The judge of the error is the plain tracer's converged image: for the glossy-indirect scene at
    512x512 the committed reference (tests/golden/glossy_cornell_ref.npz: 2^20 spp as 64 independent
    pieces, FP64 accumulation), otherwise ref_spp samples rendered here as eight pieces; the spread of
    the pieces estimates the reference's own per-pixel noise, which is subtracted from every squared
    error so that it does not count against either method. Every render starts from scratch (a
    guided render initialises and trains its network online inside the timed call).
    Methods: rtg_render_desc.guiding 0 (plain, the reference's estimator), 1 (learned guiding with
    the rough-dielectric exit proposals), 2 (the exit proposals alone).
    Ladder: one render per point; the guided/plain error ratio and error x time ratio at 1024 spp
    (the configuration's sample budget). Time to target: independent renders of piece_spp samples
    (seeds 5000 + k) are averaged progressively, the error measured after each, until the average
    reaches the target or max_spp; the time is the sum of the pieces' wall times, interpolated
    log-log between the two pieces around the crossing."""
    import numpy as np
    from raytracer_amd import rtg, scenes

    sd = getattr(scenes, scene)(width=size, height=size)
    g = ctx.upload(sd)
    cam = sd.camera
    t0 = time.perf_counter()
    if scene == "glossy_cornell" and size == 512 and os.path.exists(GLOSSY_REFERENCE):
        z = np.load(GLOSSY_REFERENCE)
        ref = z["mean"].astype(np.float64)
        noise = z["var_mean"].astype(np.float64)
        ref_desc = f"committed, {int(z['pieces'])} x {int(z['spp'])} spp"
    else:
        pieces = []
        for s in range(8):
            h, _ = g.render(cam, dict(sd.render, spp=ref_spp // 8, seed=100 + s))
            pieces.append(rtg.resolve(h).astype(np.float64))
        pieces = np.array(pieces)
        ref = pieces.mean(0)
        noise = pieces.var(0, ddof=1) / len(pieces)  # variance of the reference mean
        ref_desc = f"rendered here, 8 x {ref_spp // 8} spp"
    t_ref = time.perf_counter() - t0
    res = {"scene": f"{scene} {size}x{size}", "target_relmse": target, "reference": ref_desc,
           "reference_own_relmse": float(np.mean(noise / (ref * ref + 1e-2))), "reference_s": t_ref,
           "methods": {name: gd for name, gd in methods}, "piece_spp": piece_spp, "max_spp": max_spp}
    extra = {}
    for name, gd in methods:
        ctx.set_profiling(rtg.PROFILE_EVENTS)
        pts = []
        for i, spp in enumerate(ladder):
            t1 = time.perf_counter()
            out, st = g.render(cam, dict(sd.render, spp=spp, seed=1000 + i, guiding=gd))
            dt = time.perf_counter() - t1
            e = relmse(rtg.resolve(out).astype(np.float64), ref, noise)
            pts.append((spp, dt, e))
            print(f"[bench]   {scene} {name} {spp} spp: {dt:.2f} s, relMSE {e:.3e}", file=sys.stderr, flush=True)
            if gd == 1 and spp == 1024:
                infer_s = st.stage_ms[2] / 1e3
                tflops = st.guide_queries * GUIDE_FLOPS_PER_QUERY / infer_s / 1e12 if infer_s > 0 else 0.0
                extra = {"train_steps": int(st.guide_train_steps), "train_ms": st.guide_train_ms,
                         "infer_ms": st.stage_ms[2], "render_ms": st.milliseconds, "queries": int(st.guide_queries),
                         "infer_tflops": tflops, "mfma_util": tflops / MFMA_BF16_PEAK_TFLOPS,
                         "training_samples": int(st.guide_samples), "final_loss": st.guide_loss}
        ctx.set_profiling(0)
        # progressive average of independent pieces until the target
        acc = np.zeros_like(ref)
        prog = []
        t_sum = 0.0
        k = 0
        while (k + 1) * piece_spp <= max_spp:
            t1 = time.perf_counter()
            out, _ = g.render(cam, dict(sd.render, spp=piece_spp, seed=5000 + k, guiding=gd))
            img = rtg.resolve(out).astype(np.float64)
            t_sum += time.perf_counter() - t1
            acc += img
            k += 1
            e = relmse(acc / k, ref, noise)
            prog.append((k * piece_spp, t_sum, e))
            if e <= target:
                break
        print(f"[bench]   {scene} {name} progressive: {k} pieces, {k * piece_spp} spp, {t_sum:.1f} s, relMSE {e:.3e}",
              file=sys.stderr, flush=True)
        t_to, spp_to = None, None
        if prog[-1][2] <= target:
            if len(prog) == 1:
                t_to, spp_to = prog[0][1], float(prog[0][0])
            else:
                (s0, ta, e0), (s1, tb, e1) = prog[-2], prog[-1]
                f = (np.log(target) - np.log(e0)) / (np.log(e1) - np.log(e0)) if e0 > e1 else 1.0
                t_to = float(np.exp(np.log(ta) + f * (np.log(tb) - np.log(ta))))
                spp_to = float(np.exp(np.log(s0) + f * (np.log(s1) - np.log(s0))))
        res[name] = {"ladder": [{"spp": p[0], "s": round(p[1], 4), "relmse": p[2]} for p in pts],
                     "progressive": [{"spp": p[0], "s": round(p[1], 3), "relmse": p[2]} for p in prog],
                     "s_to_target": t_to, "spp_to_target": spp_to,
                     "samples_per_s_to_target": (spp_to * size * size / t_to) if t_to else None}
    at = {name: next(p for p in res[name]["ladder"] if p["spp"] == 1024) for name, _ in methods}
    for name, _ in methods[1:]:
        a, b = at[name], at["plain"]
        res[f"{name}_error_ratio_1024spp"] = a["relmse"] / b["relmse"]
        res[f"{name}_error_x_time_ratio_1024spp"] = (a["relmse"] * a["s"]) / (b["relmse"] * b["s"])
        res[f"{name}_over_plain_time"] = None
        if res["plain"]["s_to_target"] and res[name]["s_to_target"]:
            res[f"{name}_over_plain_time"] = res[name]["s_to_target"] / res["plain"]["s_to_target"]
            res[f"{name}_meets_half_time_target"] = res[f"{name}_over_plain_time"] <= 0.5
    # the headline pair: learned guiding (the network and the exit proposals) against plain; the
    # fastest guided mode beside it
    res["error_ratio_1024spp"] = res.get("guided_error_ratio_1024spp")
    res["error_x_time_ratio_1024spp"] = res.get("guided_error_x_time_ratio_1024spp")
    res["guided_over_plain_time"] = res.get("guided_over_plain_time")
    if res["guided_over_plain_time"] is not None:
        res["meets_half_time_target"] = res["guided_over_plain_time"] <= 0.5
    timed = [(res[f"{n}_over_plain_time"], n) for n, _ in methods[1:] if res.get(f"{n}_over_plain_time")]
    if timed:
        res["fastest_guided_mode"] = {"method": min(timed)[1], "over_plain_time": min(timed)[0],
                                      "meets_half_time_target": min(timed)[0] <= 0.5}
    res["guide_network"] = extra
    return res


def volumetric_leg(ctx, run_cpu, stride):
    """Heterogeneous volumetric path tracing: the Cornell box over [0,1]^3 filled by a 256^3 fBm
    density (5 octaves, seed 3, clipped to [0,4]), albedo 0.9, Henyey-Greenstein g 0.3, 512x512,
    128 spp, depth 16, delta tracking with ratio-tracking transmittance. One warm-up render, one
    timed render; the CPU oracle on every stride-th 32x32 block of the same render."""
    import torch
    from raytracer_amd import scenes

    sd = scenes.vol_cornell()
    scene = ctx.upload(sd)
    cam, rp = sd.camera, sd.render
    W, H = cam["width"], cam["height"]
    film = torch.zeros(H * W * 4, dtype=torch.float32, device="cuda")
    scene.render_device(cam, rp, film.data_ptr())
    torch.cuda.synchronize()
    film.zero_()
    t0 = time.perf_counter()
    st = scene.render_device(cam, rp, film.data_ptr())
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    out = {
        "workload": f"vol_cornell heterogeneous medium 256^3 VolPathTracing {W}x{H} {rp['spp']}spp depth{rp['max_depth']}",
        "value": st.samples / wall,
        "unit": "samples/s",
        "ms": wall * 1e3,
        "rays_per_s": (st.extension_rays + st.shadow_rays) / wall,
        "iterations": st.iterations,
        "nan_samples": st.nan_samples,
    }
    if run_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib as O

        threads, host = host_cpu()
        o = O.OracleScene(sd)
        t0 = time.perf_counter()
        _, _, cst, nblocks = o.render(spp=rp["spp"], threads=threads, blocks=(0, 0, stride))
        dt = time.perf_counter() - t0
        samples = nblocks * 32 * 32 * rp["spp"]
        out["cpu_baseline"] = {
            "value": samples / dt, "unit": "samples/s", "cores": threads, "kind": "port", **host,
            "sample": (f"the whole {W}x{H} render" if stride == 1 else f"every {stride}th 32x32 block of the same "
                       f"{W}x{H} render") + f" ({nblocks} blocks, {samples} samples, {dt:.1f} s)",
        }
        out["speedup_vs_cpu"] = out["value"] / out["cpu_baseline"]["value"]
    del scene
    return out


def copy_reference(mb):
    import torch

    n = mb * (1 << 20) // 4
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_reference_mib": mb, "bytes_read_per_copy": n * 4, "bytes_written_per_copy": n * 4}))


def main():
    a = parse()
    if a.copy_mb:
        copy_reference(a.copy_mb)
        return
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {a.gpus}); the multi-GPU design is described, not built.",
              file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    from raytracer_amd import rtg, scenes

    def progress(msg):
        print(f"[bench] {msg}", file=sys.stderr, flush=True)

    def leg_context():
        # a leg run on its own takes the launch configuration of an earlier full run when given
        # (the full run's legs share the benchmark render's tuned context)
        c = rtg.Context(0)
        if a.tuning_from:
            with open(a.tuning_from) as f:
                prev = json.loads(f.read().strip().splitlines()[-1])["tuning"]
            prev.pop("trial_samples_per_s", None)
            c.set_tuning(**prev)
        return c

    if a.only_volume:
        progress("volumetric leg only")
        print(json.dumps(volumetric_leg(leg_context(), not a.no_cpu, a.volume_stride)), flush=True)
        return
    if a.only_guiding:
        gctx = leg_context()
        progress("guided vs plain only")
        print(json.dumps({"guiding": guiding_leg(gctx, "glossy_cornell", a.guide_size, a.guide_ref_spp)}), flush=True)
        return

    progress("building the displaced-sphere scene")
    sd = scenes.displaced_sphere(width=a.size, height=a.size, spp=a.spp)
    ctx = rtg.Context(0)
    scene = ctx.upload(sd)
    if a.only_serving:
        progress(f"serving mix only ({a.serve_requests} requests)")
        print(json.dumps(serving(ctx, (sd, scene), a.serve_requests)), flush=True)
        return
    cam, rp = sd.camera, sd.render
    W, H = cam["width"], cam["height"]
    film = torch.zeros(H * W * 4, dtype=torch.float32, device="cuda")

    # launch configuration for this scene (pool size, traversal waves and refill) from the autotuner
    if a.tuning_from:
        with open(a.tuning_from) as f:
            prev = json.loads(f.read().strip().splitlines()[-1])["tuning"]
        tune_rate = prev.pop("trial_samples_per_s", None)
        ctx.set_tuning(**prev)
        tuning = ctx.tuning()
    else:
        progress("autotune")
        tuning, tune_rate = ctx.autotune(scene, cam, rp, trial_samples=W * H * rp["spp"])
    if a.pool:
        ctx.set_tuning(pool_slots=a.pool)
        tuning = ctx.tuning()

    # one counted render at the chosen configuration (the pool size sets the rays per launch):
    # traversal work per launch (not timed)
    progress("counted render")
    ctx.set_profiling(rtg.PROFILE_COUNT)
    film.zero_()
    st = scene.render_device(cam, rp, film.data_ptr())
    counted = dict(ext_rays=st.extension_rays, nodes=st.nodes_visited, prims=st.prims_tested,
                   sh_rays=st.shadow_rays, sh_nodes=st.shadow_nodes_visited, sh_prims=st.shadow_prims_tested,
                   iters=st.iterations)

    ctx.set_profiling(rtg.PROFILE_EVENTS)
    for _ in range(a.warmup):
        film.zero_()
        scene.render_device(cam, rp, film.data_ptr())
    torch.cuda.synchronize()
    progress(f"timing {a.steps} steps")
    stage_ms = np.zeros(len(rtg.STAGES))
    launches = np.zeros(len(rtg.STAGES))
    dev_ms = 0.0
    samples = 0
    rays = 0
    t0 = time.perf_counter()
    for _ in range(a.steps):
        film.zero_()
        st = scene.render_device(cam, rp, film.data_ptr())
        stage_ms += np.array(st.stage_ms[:])
        launches += np.array(st.stage_launches[:])
        dev_ms += st.milliseconds
        samples += st.samples
        rays += st.extension_rays + st.shadow_rays
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1e3 / a.steps
    value = samples / wall

    # HBM bytes per extension-traversal dispatch from the committed rocprofv3 PMC summary of this
    # workload (FETCH_SIZE and WRITE_SIZE, each scaled by what the counter reads for a copy of known
    # size); null when the summary is absent or was taken on another workload
    workload = f"displaced_sphere plain MIS path tracing {W}x{H} {rp['spp']}spp depth{rp['max_depth']}"
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "k_extend_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            tj = json.load(f)
        if tj.get("workload") == workload:
            per_pool = tj.get("hbm_bytes_per_dispatch_by_pool", {})
            if str(tuning["pool_slots"]) in per_pool:
                traffic = float(per_pool[str(tuning["pool_slots"])])

    # roofline of the dominant kernel (extension traversal)
    ext_bytes = (counted["ext_rays"] * RAY_IO_BYTES + counted["nodes"] * NODE_BYTES
                 + counted["prims"] * TRI_BYTES)
    bytes_per_launch = ext_bytes / max(1, counted["iters"])
    path_bytes = (counted["ext_rays"] * SEGMENT_IO_BYTES + (counted["nodes"] + counted["sh_nodes"]) * NODE_BYTES
                  + (counted["prims"] + counted["sh_prims"]) * TRI_BYTES)
    avg_launch_s = stage_ms[1] / max(1, launches[1]) / 1e3
    achieved = bytes_per_launch / avg_launch_s / 1e9 if avg_launch_s > 0 else 0.0
    out = {
        "metric": "samples/s to relMSE≤1e-3 (guided vs plain GPU); p50/p99 request latency",
        "value": value,
        "unit": "samples/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (procedural scene generated in-process)",
        "config": {
            "workload": workload,
            "triangles": sd.num_triangles(), "width": W, "height": H, "spp": rp["spp"],
            "max_depth": rp["max_depth"], "min_depth": rp["min_depth"], "sampler": "stratified",
        },
        "value_measures": "plain MIS samples/s of the large-mesh render named in config.workload (the "
                          "guided-vs-plain time to relMSE 1e-3 is the guiding object, request latency the "
                          "serving object)",
        "rays_per_s": rays / wall,
        "device_ms_per_step": dev_ms / a.steps,
        "tuning": dict(tuning, trial_samples_per_s=tune_rate),
        "stage_ms_per_step": {k: float(v) / a.steps for k, v in zip(rtg.STAGES, stage_ms)},
        "roofline": {
            "bound": "hbm",
            "kernel": "k_extend (closest-hit BVH traversal + material classification)",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "bytes_per_launch": bytes_per_launch,
            "avg_launch_ms": avg_launch_s * 1e3,
            "per_ray": {"nodes": counted["nodes"] / max(1, counted["ext_rays"]),
                        "prims": counted["prims"] / max(1, counted["ext_rays"])},
        },
        # the whole step under the per-segment byte model of the path (292 B of ray, hit, path-state
        # and shadow-ray I/O per extension segment, plus node and triangle bytes of both
        # traversals; node bytes at this BVH's 64-byte nodes) over the step's wall time
        "path_model": {
            "bytes_per_step": path_bytes,
            "achieved": path_bytes / (ms_per_step / 1e3) / 1e9,
            "unit": "GB/s",
            "frac": path_bytes / (ms_per_step / 1e3) / 1e9 / HBM_PEAK_GBS,
            "segment_io_bytes": SEGMENT_IO_BYTES, "node_bytes": NODE_BYTES, "tri_bytes": TRI_BYTES,
        },
    }
    if not a.no_guiding:
        progress("guided vs plain (glossy-indirect scene)")
        out["guiding"] = guiding_leg(ctx, "glossy_cornell", a.guide_size, a.guide_ref_spp)
        progress("guided vs plain (indirect-only Lambertian variant)")
        out["guiding_indirect"] = guiding_leg(ctx, "indirect_cornell", a.guide_size, a.guide_ref_spp,
                                              piece_spp=1024, max_spp=32768, methods=(("plain", 0), ("guided", 1)))
    if not a.no_volume:
        progress("volumetric leg (heterogeneous medium, 512x512 128spp)")
        out["volumetric"] = volumetric_leg(ctx, not a.no_cpu, a.volume_stride)
    if not a.no_serving:
        progress(f"serving mix ({a.serve_requests} requests)")
        # latency over throughput: the serving loop runs at the default pool of 2^21 path slots
        # (the benchmark render's tuned pool batches more work per iteration and raises p50)
        ctx.set_tuning(pool_slots=1 << 21)
        out["serving"] = serving(ctx, (sd, scene), a.serve_requests)
        out["serving"]["tuning"] = ctx.tuning()
    if not a.no_cpu:
        progress("cpu baseline (oracle)")
        out["cpu_baseline"] = cpu_baseline(sd, rp["spp"], a.cpu_stride)
        out["speedup_vs_cpu"] = value / out["cpu_baseline"]["value"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
