"""Throughput of the path-tracing hot path on one MI355X.

Workload (one "step"): the 1920x1080 heightfield render -- 1024^2 cells
(2,097,152 triangles) displaced by seeded 6-octave fBm (amplitude 0.1),
per-triangle GGX-over-Lambert materials, sky 1.0 plus a 0.5 deg sun of
radiance 5e4, 64 spp, max depth 16, StandardPathTracer (NEE + BSDF MIS) with
the sun/sky sampled as a light.  Scene build and upload happen before the
timed region; the film lives in HBM.

Prints one JSON line: Mpaths/s (the metric), Mrays/s, the roofline of the
dominant kernel (closest-hit traversal) and the CPU baseline (the C++ CPU
checker in oracle/, OpenMP over 16x16 tiles as RenderTarget.h:110-127, on a
bounded sample of the same workload).
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E
TRI_BYTES = 48                 # one triangle record; a node's size comes from the scene (bytes_nodes / nodes)
EXT_RAY_IO_BYTES = 32 + 16     # ray in (o, d) + hit out
SHADOW_RAY_IO_BYTES = 48 + 16  # shadow record in + radiance update
QUERY_IO_BYTES = {"closest": 32 + 16, "any": 32 + 1}   # xt_ray in + xt_hit / occlusion byte out
SHADE_NMAP_BYTES_PER_PATH = 64 + 16 + 48 + 48 + 96 + 32 + 64 + 64 + 48 + 32


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--spp", type=int, default=64)
    ap.add_argument("--cells", type=int, default=1024)
    ap.add_argument("--max-depth", type=int, default=16)
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="target wall time of the CPU-baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--workload", choices=["heightfield", "cornell", "guiding", "batch", "nmap", "rays"],
                    default="heightfield",
                    help="cornell: the 256x256, 16 spp, depth 8, seed 1 Cornell box; "
                         "guiding: equal-time relMSE of guided vs plain path tracing on the occluded room; "
                         "batch: 64 small scenes rendered one by one and as one graph-replayed batch; "
                         "nmap: the displaced plane with microfacet-based normal mapping, 16 camera variants")
    ap.add_argument("--plane-grid", type=int, default=4096, help="nmap: vertices per side of the displaced plane")
    ap.add_argument("--plane-width", type=int, default=3840)
    ap.add_argument("--plane-height", type=int, default=2160)
    ap.add_argument("--plane-spp", type=int, default=1, help="nmap: samples per pixel per variant per step")
    ap.add_argument("--variants", type=int, default=16)
    ap.add_argument("--guided", action="store_true",
                    help="nmap: train the guiding network online, then render every variant guided in one "
                         "xt_render_batch call per step, and compare guided with plain quality")
    ap.add_argument("--scenes", type=int, default=64, help="batch: number of scenes")
    ap.add_argument("--autotune", action="store_true", help="batch: autotune each scene's launch shapes first")
    ap.add_argument("--batch-side", type=int, default=0, help="batch: override every scene's image side")
    ap.add_argument("--batch-spp", type=int, default=0, help="batch: override every scene's spp")
    ap.add_argument("--budget", type=float, default=6.0, help="guiding: seconds of rendering per method")
    ap.add_argument("--ref-spp", type=int, default=16384,
                    help="samples per pixel of the FP64 reference each image's relMSE is stated against")
    ap.add_argument("--no-ref", action="store_true", help="skip the converged reference (no relMSE in the line)")
    ap.add_argument("--guided-runs", type=int, default=3,
                    help="guiding: guided runs from fresh networks (seeds 5..4+n); the median ratio is reported")
    ap.add_argument("--room-tris", type=int, default=500_000)
    ap.add_argument("--calib-copy", action="store_true",
                    help="first run a plain 1 GiB streaming kernel (out = in * 1) to calibrate rocprof byte counters")
    return ap.parse_args()


def pmc_from_profiles(name):
    """Per-launch counters of the dominant kernel measured with rocprofv3
    --pmc, if a summary for this workload is committed under profiles/:
    hbm_bytes_per_launch (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel
    calibration) and valu_insts_per_launch (SQ_INSTS_VALU, wave-level)."""
    path = os.path.join(ROOT, "profiles", name)
    out = {}
    if not os.path.exists(path):
        return out
    text = open(path).read()
    for key in ("hbm_bytes_per_launch", "valu_insts_per_launch"):
        m = re.search(rf"^{key}:\s*([0-9.eE+]+)", text, re.M)
        if m:
            out[key] = float(m.group(1))
    return out


def traffic_from_profiles(kernel_tag, name=None):
    """HBM bytes per launch (rocprofv3 --pmc FETCH_SIZE + WRITE_SIZE, scaled by
    the copy-kernel calibration) from the committed summary, or None."""
    return pmc_from_profiles(name or f"{kernel_tag}_pmc.txt").get("hbm_bytes_per_launch")


# VALU issue peak (wave-level instructions per second): the published fp32
# vector peak, 157.3 TFLOP/s = 256 CUs x 2.4 GHz x 128 FMA lanes, is two wave64
# instructions per CU per cycle.  xt_device_valu_rate's probe (16 independent
# FMA chains per lane) is reported beside it: it issues about 0.75 of this,
# well above one instruction per CU per cycle.
VALU_ISSUE_PEAK = 2 * 256 * 2.4e9
_valu_probe = None


def valu_probe_rate():
    global _valu_probe
    if _valu_probe is None:
        from xitils_amd import api
        _valu_probe = api.Device(0).valu_rate()
    return _valu_probe


def kernel_roofline(kernel, algo_bytes_per_launch, avg_launch_ms, pmc, extra=None):
    """Roofline object of a memory-side kernel.  achieved / frac: algorithmic
    bytes per launch (traversal: ray records plus every node and triangle
    fetch at full size, the fetch counts pinned by the host walk of the same
    BVH; shade: the path-state records it reads and writes) over the measured
    launch time, against the 8 TB/s HBM peak.  The bytes that really
    cross HBM (rocprof PMC) and the VALU issue rate (SQ_INSTS_VALU) are
    reported beside it, and `bound` names what limits the kernel: "hbm" or
    "valu-issue" when that measured rate is at 70% of its peak or more, else
    "latency" (neither unit saturated: the kernel waits on its dependent
    loads and tests)."""
    sec = avg_launch_ms * 1e-3
    achieved = algo_bytes_per_launch / sec / 1e9
    r = {
        "bound": "hbm",
        "achieved": round(achieved, 2),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": pmc.get("hbm_bytes_per_launch"),
        "kernel": kernel,
        "bytes_per_launch": round(algo_bytes_per_launch),
        "avg_launch_ms": round(avg_launch_ms, 4),
    }
    if "hbm_bytes_per_launch" in pmc:
        r["hbm_measured_frac"] = round(pmc["hbm_bytes_per_launch"] / sec / 1e9 / HBM_PEAK_GBS, 4)
    if "valu_insts_per_launch" in pmc:
        r["valu_issue_frac"] = round(pmc["valu_insts_per_launch"] / sec / VALU_ISSUE_PEAK, 4)
        r["valu_issue_peak"] = round(VALU_ISSUE_PEAK / 1e9, 1)       # G wave-instructions/s
        r["valu_probe_rate"] = round(valu_probe_rate() / 1e9, 1)     # measured, same unit
    vf, hf = r.get("valu_issue_frac"), r.get("hbm_measured_frac")
    if vf is not None and hf is not None:
        if vf >= 0.7 and vf > hf:
            r["bound"] = "valu-issue"
        elif hf < 0.7:
            r["bound"] = "latency"
        r["note"] = ("frac counts the kernel's algorithmic bytes at full size; most of them hit in L2/MALL, so the "
                     "bytes crossing HBM (hbm_measured_frac) are fewer; with VALU issue (valu_issue_frac) below "
                     "its peak too, the kernel is bound by the latency of its dependent fetches and tests"
                     if r["bound"] == "latency" else
                     "frac counts the kernel's algorithmic bytes at full size; the measured rates name the bound")
    if extra:
        r.update(extra)
    return r


def cpu_baseline(desc, args, max_depth, flags):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    o = pyoracle.OracleScene(desc)
    W, H = args.width, args.height
    t0 = time.perf_counter()
    o.render(W, H, 1, seed=1, max_depth=max_depth, flags=flags, nthreads=threads)
    t1 = time.perf_counter()
    one = t1 - t0
    spp = max(1, min(args.spp, int(args.cpu_seconds / max(one, 1e-3))))
    t0 = time.perf_counter()
    o.render(W, H, spp, seed=1, max_depth=max_depth, flags=flags, nthreads=threads, spp_offset=1)
    dt = time.perf_counter() - t0
    paths = W * H * spp
    ext, shd = (int(x) for x in o.last_stats)
    cpu_model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                cpu_model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {
        "value": round(paths / dt / 1e6, 4),
        "unit": "Mpaths/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{W}x{H} at {spp} spp of the same scene/seed (oracle C++ restatement, -O3 -march=x86-64-v3, "
                  f"OpenMP over 16x16 tiles, reference-shaped binary BVH), {dt:.1f} s wall; "
                  f"{(ext + shd) / dt / 1e6:.2f} Mrays/s; {cpu_model}",
    }


def guiding_main(args):
    """Equal-time relMSE, guided vs plain, on the 1280x720 occluded room
    (8x4x3 m room lit through a 5 cm door slit by an exterior light of
    radiance 500; 500k triangles of seeded boxes and spheres, 50/50 Lambert /
    GGX alpha U[0.1, 0.4]) against a 16,384-spp reference accumulated in FP64
    on the GPU.  Two guide configurations run on the same reference: the
    stated one (Adam lr 1e-2, training during the first 25% of the budget;
    the headline value) and the library's defaults (lr and training share
    tuned on this scene).  Guide networks start from held-out seeds (5, 6, 7,
    not the seeds the defaults were tuned with); plain and guided each render
    with two render seeds."""
    import torch  # (load torch's HIP runtime before the library opens the device)
    from xitils_amd import api, equal_time, quality, scenes
    W, H = 1280, 720
    max_depth = args.max_depth
    log(f"[bench] building occluded room ({args.room_tris} triangles)")
    desc = scenes.occluded_room(W, H, n_tris=args.room_tris)
    dev = api.Device(0)
    scene = api.Scene(dev, desc)
    info = scene.info()
    log(f"[bench] scene {info}")
    scene.guide_init(seed=1)
    t0 = time.perf_counter()
    ref = quality.reference_image(scene, W, H, spp=args.ref_spp, max_depth=max_depth)
    ref_s = time.perf_counter() - t0
    log(f"[bench] reference {args.ref_spp} spp in {ref_s:.1f} s ({W * H * args.ref_spp / ref_s / 1e6:.1f} Mpaths/s)")
    guide_seeds = list(range(5, 5 + args.guided_runs))
    default_lr = float(api.Scene.guide_config().learning_rate)
    configs = {
        "stated": {"learning_rate": 1e-2, "train_fraction": 0.25},
        "tuned": {"learning_rate": default_lr, "train_fraction": equal_time.TRAIN_FRACTION},
    }
    results = {}
    for name, c in configs.items():
        log(f"[bench] guide configuration {name}: {c}")
        res = equal_time.compare(scene, W, H, args.budget, ref, log=log, max_depth=max_depth,
                                 guide_seeds=guide_seeds, render_seeds=[1, 2],
                                 guide_config={"learning_rate": c["learning_rate"]},
                                 train_fraction=c["train_fraction"])
        # the values the runs used, read back from the library
        lr_used = float(api.Scene.guide_config(learning_rate=c["learning_rate"]).learning_rate)
        results[name] = {"config": {"learning_rate": lr_used, "train_fraction": c["train_fraction"],
                                    "guide_prob": float(api.Scene.guide_config().guide_prob)}, "result": res}
    head = results["stated"]["result"]
    out = {
        "metric": "equal-time relMSE guided vs plain GPU PT (16k-spp GPU FP64 reference)",
        "value": round(head["relmse_ratio_equal_time"], 4),
        "unit": "relMSE(plain) / relMSE(guided) at equal wall time",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 0,
        "ms_per_step": round(args.budget * 1e3, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 render, bf16/fp32 MLP",
        "data": "synthetic",
        "config": {"workload": f"occluded_room_{W}x{H}_{info['num_prims']}tris_depth{max_depth}",
                   "budget_s": args.budget, "reference_spp": args.ref_spp, "guide_seeds": guide_seeds,
                   "render_seeds": [1, 2], **results["stated"]["config"], "parallelism": "1 GPU"},
        "time_speedup": head["time_speedup"],
        "guided_seconds_to_plain_relmse": head["guided_seconds_to_plain_relmse"],
        "plain": head["plain"],
        "plain_runs": head["plain_runs"],
        "guided": head["guided"],
        "guided_runs": head["runs"],
        "tuned_defaults": {"config": results["tuned"]["config"],
                           "relmse_ratio_equal_time": round(results["tuned"]["result"]["relmse_ratio_equal_time"], 4),
                           "time_speedup": results["tuned"]["result"]["time_speedup"],
                           "guided_runs": results["tuned"]["result"]["runs"],
                           "note": "learning rate and training share tuned on this scene (network seeds 1-4)"},
        "reference_seconds": round(ref_s, 2),
    }
    # the network's arithmetic: 33,792 FLOP per query (2 x (96x64 + 64x64 + 64x64 + 64x40)); a training
    # sample adds the weight gradients (33,792) and the activation gradients of layers 1-3
    # (2 x (64x40 + 64x64 + 64x64) = 21,504): 89,088 FLOP.  Peak: 2.5 PFLOP/s dense bf16 MFMA.
    g = head["guided"]
    mlp = {"peak_tflops_bf16": 2500.0}
    if g.get("guide_queries") and g.get("ms_guide_infer"):
        tf = g["guide_queries"] * 33792 / (g["ms_guide_infer"] * 1e-3) / 1e12
        mlp["infer"] = {"queries": g["guide_queries"], "flop_per_query": 33792, "tflops": round(tf, 2),
                        "frac": round(tf / 2500.0, 4)}
    # training alone: 1-spp training renders with the device idle between them, so
    # that each burst of Adam steps runs by itself (the equal-time runs overlap the
    # bursts with the next render, which stretches them)
    film = torch.zeros((H, W, 3), dtype=torch.float32, device="cuda:0")
    scene.guide_init(seed=5, learning_rate=1e-2)
    tr_flags = api.XT_FLAG_GUIDE | api.XT_FLAG_GUIDE_TRAIN
    n = steps = 0
    ms = 0.0
    for k in range(9):
        scene.render_device(film.data_ptr(), W, H, 1, spp_offset=k, seed=3, flags=tr_flags, max_depth=max_depth)
        torch.cuda.synchronize()
        st = scene.stats()
        n += st["guide_grad_samples"]
        steps += st["guide_steps"]
        ms += st["ms_guide_train"]
    scene.render_device(film.data_ptr(), W, H, 1, spp_offset=9, seed=3, flags=api.XT_FLAG_GUIDE, max_depth=max_depth)
    ms += scene.stats()["ms_guide_train"]   # the last burst
    if n and ms > 0:
        tf = n * 89088 / (ms * 1e-3) / 1e12
        mlp["train"] = {"minibatch_samples": n, "adam_steps": steps, "ms": round(ms, 3),
                        "us_per_step": round(ms * 1e3 / max(steps, 1), 2), "flop_per_sample": 89088,
                        "tflops": round(tf, 2), "frac": round(tf / 2500.0, 4),
                        "note": "9 one-spp training renders, the device idle between them"}
    if g.get("grad_samples") and g.get("ms_guide_train"):
        tf = g["grad_samples"] * 89088 / (g["ms_guide_train"] * 1e-3) / 1e12
        mlp["train_overlapped"] = {"minibatch_samples": g["grad_samples"], "ms": round(g["ms_guide_train"], 1),
                                   "tflops": round(tf, 2),
                                   "note": "the equal-time run's bursts, sharing the GPU with the next render"}
    out["mlp"] = mlp
    out["curves"] = head["curves"]
    print(json.dumps(out), flush=True)


def image_quality(scene, W, H, spp, ref_spp, **kw):
    """relMSE of one `spp`-sample image (the bench's seed) against a
    `ref_spp`-sample reference accumulated in FP64 on the GPU with another
    seed; both outside the timed region."""
    from xitils_amd import quality
    img = scene.render(W, H, spp, **kw) / spp
    t0 = time.perf_counter()
    kw = {k: v for k, v in kw.items() if k != "seed"}
    ref = quality.reference_image(scene, W, H, spp=ref_spp, **kw)
    return {"relmse": quality.rel_mse(img, ref), "image_spp": spp, "reference_spp": ref_spp,
            "reference": "same renderer, FP64 film, seed 0x5EED", "reference_seconds": round(time.perf_counter() - t0, 2)}


def cornell_main(args):
    """The Cornell box as configured: unit room of 5 Lambert walls (10
    triangles; 0.73 grey, red (0.63, 0.065, 0.05), green (0.14, 0.45, 0.09)),
    a 0.3 short box and a 0.6 tall box (12 triangles each, rotated -18/+18
    degrees), a 0.25 x 0.2 ceiling light of radiance (17, 12, 4); 256 x 256,
    16 spp, max depth 8, seed 1, StandardPathTracer.  One step renders the
    image once (1,048,576 paths) through xt_render into a device film."""
    import torch
    from xitils_amd import api, scenes
    W = H = 256
    spp, depth = 16, 8
    desc = scenes.cornell(W, H)
    dev = api.Device(0)
    scene = api.Scene(dev, desc)
    info = scene.info()
    film = torch.zeros((H, W, 3), dtype=torch.float32, device="cuda:0")
    kw = dict(seed=1, max_depth=depth)
    scene.set_counting(True)
    scene.render_device(film.data_ptr(), W, H, spp, **kw)
    cst = scene.stats()
    scene.set_counting(False)
    steps = max(args.steps, 50)
    for _ in range(max(args.warmup, 5)):
        scene.render_device(film.data_ptr(), W, H, spp, **kw)
    agg = {"ms_extend": 0.0, "ms_shade": 0.0, "ms_shadow": 0.0, "ms_other": 0.0, "extension_rays": 0,
           "shadow_rays": 0, "extend_launches": 0}
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        scene.render_device(film.data_ptr(), W, H, spp, **kw)
        st = scene.stats()
        for key in agg:
            agg[key] += st[key]
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    paths = W * H * spp * steps
    n_ext = cst["extension_rays"]
    node_bytes = info["bytes_nodes"] // max(1, info["bvh_nodes"])
    bytes_ext = n_ext * EXT_RAY_IO_BYTES + cst["ext_node_fetches"] * node_bytes + cst["ext_tri_fetches"] * TRI_BYTES
    launches = agg["extend_launches"] / steps
    out = {
        "metric": "Mpaths/s (Cornell box 256x256, 16 spp)",
        "value": round(paths / wall / 1e6, 3),
        "unit": "Mpaths/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": max(args.warmup, 5),
        "ms_per_step": round(wall * 1e3 / steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "cornell_256x256_16spp_depth8_seed1", "triangles": info["num_prims"], "width": W,
                   "height": H, "spp": spp, "max_depth": depth, "integrator": "StandardPathTracer (NEE+BSDF MIS)",
                   "parallelism": "1 GPU"},
        "mrays_per_s": round((agg["extension_rays"] + agg["shadow_rays"]) / wall / 1e6, 2),
        "stage_ms_per_step": {k: round(agg[k] / steps, 3) for k in ("ms_extend", "ms_shade", "ms_shadow", "ms_other")},
        "roofline": kernel_roofline("k_trace<closest> (extend)", bytes_ext / launches,
                                    agg["ms_extend"] / max(1, agg["extend_launches"]), {},
                                    {"note": "36 triangles: the tree and triangles sit in L1/LDS-sized caches; "
                                             "launch and latency bound, Mrays/s is the figure of merit"}),
    }
    if not args.no_ref:
        out["quality"] = image_quality(scene, W, H, spp, args.ref_spp, **kw)
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        o = pyoracle.OracleScene(desc)
        reps, t0 = 0, time.perf_counter()
        while reps == 0 or time.perf_counter() - t0 < min(args.cpu_seconds, 10.0):
            o.render(W, H, spp, nthreads=threads, **kw)
            reps += 1
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(W * H * spp * reps / dt / 1e6, 4), "unit": "Mpaths/s", "cores": threads,
                               "kind": "port",
                               "sample": f"the full 256x256 x 16 spp image rendered {reps} times by the oracle C++ "
                                         f"restatement (OpenMP over 16x16 tiles), {dt:.1f} s wall"}
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


def batch_main(args):
    """Aggregate Mpaths/s and per-scene latency over the 64-scene batch
    (master seed 3: unit rooms with tessellated spheres and boxes, 1e4..2e6
    triangles, 128^2..1024^2 pixels, 4..256 spp, depth 4..32, one area light
    each), rendered (a) one xt_render call after another and as one
    xt_render_batch call in its three forms: (b) the default -- items of at
    most one graph replay share one wavefront (the items' paths walked in
    windows, each window one captured graph replay), larger ones run per item
    concurrently -- (c) every item in the wavefront and (d) a graph per item
    on per-scene streams.  value is (b)."""
    import numpy as np
    import torch
    from xitils_amd import api, scenes
    dev = api.Device(0)
    t0 = time.perf_counter()
    jobs = []
    tris = 0
    for k in range(args.scenes):
        desc, w, h, spp, depth = scenes.batch_scene(k)
        if args.batch_side:
            w = h = args.batch_side
        if args.batch_spp:
            spp = args.batch_spp
        sc = api.Scene(dev, desc)
        tris += sc.info()["num_prims"]
        film = torch.zeros((h, w, 3), dtype=torch.float32, device="cuda:0")
        jobs.append((sc, w, h, spp, depth, film))
    build_s = time.perf_counter() - t0
    paths = sum(w * h * spp for _, w, h, spp, _, _ in jobs)
    log(f"[bench] {len(jobs)} scenes, {tris} triangles, {paths / 1e6:.1f} M paths, built in {build_s:.1f} s")
    tuned = []
    if args.autotune:
        for sc, w, h, spp, depth, _ in jobs:
            tuned.append(sc.autotune(w, h, spp, seed=1, max_depth=depth))
        log(f"[bench] autotuned: {tuned}")

    def sequential():
        done = []
        torch.cuda.synchronize()
        t = time.perf_counter()
        for sc, w, h, spp, depth, film in jobs:
            sc.render_device(film.data_ptr(), w, h, spp, seed=1, max_depth=depth)   # synchronous
            done.append((time.perf_counter() - t) * 1e3)
        torch.cuda.synchronize()
        return time.perf_counter() - t, done

    def batched(mode):
        if mode != "hybrid":
            os.environ["XITILS_AMD_BATCH"] = mode
        try:
            torch.cuda.synchronize()
            t = time.perf_counter()
            _, ms = api.render_batch([(sc, w, h, spp, (film.data_ptr(), False), dict(seed=1, max_depth=depth))
                                      for sc, w, h, spp, depth, film in jobs])
            torch.cuda.synchronize()
            return time.perf_counter() - t, ms
        finally:
            os.environ.pop("XITILS_AMD_BATCH", None)

    sequential()                       # warm-up (first-touch allocations)
    seq_s, seq_done = sequential()
    log(f"[bench] sequential: {seq_s:.3f} s")
    timed = {}
    for mode in ("items", "unified", "hybrid"):
        batched(mode)                  # warm-up: graph captures, the wavefront's buffers
        timed[mode] = batched(mode)
        log(f"[bench] xt_render_batch ({mode}): {timed[mode][0]:.3f} s")
    bat_s, bat_lat = timed["hybrid"]
    item_s, item_ms = timed["items"]
    uni_s, uni_lat = timed["unified"]
    pct = lambda v, q: float(np.percentile(np.asarray(v), q))  # noqa: E731
    out = {
        "metric": "aggregate Mpaths/s over a 64-scene batch",
        "value": round(paths / bat_s / 1e6, 2),
        "unit": "Mpaths/s",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 1,
        "ms_per_step": round(bat_s * 1e3, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"batch{len(jobs)}_seed3" + (f"_side{args.batch_side}" if args.batch_side else "") +
                               (f"_spp{args.batch_spp}" if args.batch_spp else ""), "scenes": len(jobs), "triangles": tris,
                   "paths": paths, "autotune": bool(args.autotune), "parallelism": "1 GPU"},
        "latency_ms": {"p50": round(pct(bat_lat, 50), 3), "p99": round(pct(bat_lat, 99), 3),
                       "note": "xt_render_batch's item times: a wavefront item's is the end of the window holding its "
                               "last path, a graph item's its graph on its stream, both from HIP events"},
        "sequential": {"mpaths_per_s": round(paths / seq_s / 1e6, 2), "seconds": round(seq_s, 3),
                       "latency_ms": {"p50": round(pct(seq_done, 50), 3), "p99": round(pct(seq_done, 99), 3)},
                       "note": "one synchronous xt_render per scene; latency = host completion time from the start"},
        "per_item_graphs": {"mpaths_per_s": round(paths / item_s / 1e6, 2), "seconds": round(item_s, 3),
                            "item_gpu_ms": {"p50": round(pct(item_ms, 50), 3), "p99": round(pct(item_ms, 99), 3)},
                            "note": "XITILS_AMD_BATCH=items: a graph per item on per-scene streams"},
        "unified_all": {"mpaths_per_s": round(paths / uni_s / 1e6, 2), "seconds": round(uni_s, 3),
                        "latency_ms": {"p50": round(pct(uni_lat, 50), 3), "p99": round(pct(uni_lat, 99), 3)},
                        "note": "XITILS_AMD_BATCH=unified: every item in the one wavefront"},
        "speedup_vs_sequential": round(seq_s / bat_s, 3),
        "build_s": round(build_s, 1),
    }
    if tuned:
        out["tuned_blocks_per_cu"] = tuned
    print(json.dumps(out), flush=True)


def nmap_main(args):
    """Mpaths/s of the displaced-plane workload: a plane of plane_grid^2
    vertices (2 (grid - 1)^2 triangles) displaced by a low-frequency sine sum,
    shaded with MicrofacetNormalMapping over GGX (alpha 0.2) and the 32-sine
    normal map, lit by a 2048x1024 lat-long HDR sky with a 2e4 sun sampled
    as a light; one step renders every camera variant at plane_width x
    plane_height, plane_spp spp, depth 12, as one xt_render_batch call (the
    variants share the scene, so they share one wavefront).  The same steps
    rendered one xt_render call per variant give the per-stage times and the
    shade kernel's roofline, and their rate is reported beside the value."""
    import numpy as np
    import torch
    from xitils_amd import api, scenes
    W, H, spp, depth, nv = args.plane_width, args.plane_height, args.plane_spp, 12, args.variants
    flags = scenes.FLAG_ENV_NEE
    t0 = time.perf_counter()
    desc = scenes.displaced_plane(grid=args.plane_grid, width=W, height=H)
    gen_s = time.perf_counter() - t0
    log(f"[bench] displaced plane {args.plane_grid}^2: {desc.num_triangles} triangles, generated in {gen_s:.1f} s")
    dev = api.Device(0)
    t0 = time.perf_counter()
    scene = api.Scene(dev, desc)
    upload_s = time.perf_counter() - t0
    info = scene.info()
    log(f"[bench] scene {info}, build+upload {upload_s:.1f} s")
    cams = [scenes.displaced_plane_camera(v, W, H) for v in range(nv)]
    films = [torch.zeros((H, W, 3), dtype=torch.float32, device="cuda:0") for _ in range(nv)]
    if args.calib_copy:
        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda:0")   # 1 GiB read + 1 GiB written
        dst = torch.empty_like(src)
        torch.mul(src, 1.0, out=dst)
        torch.cuda.synchronize()
        del src, dst
        log("[bench] calibration copy done (1 GiB in, 1 GiB out)")
    kw = dict(seed=1, max_depth=depth, flags=flags)
    for k in range(args.warmup):
        scene.render_device(films[0].data_ptr(), W, H, spp, camera=cams[0], **kw)
        log(f"[bench] warmup {k} done")
    for f in films:
        f.zero_()
    agg = {"ms_extend": 0.0, "ms_shade": 0.0, "ms_shadow": 0.0, "ms_other": 0.0, "extension_rays": 0,
           "shadow_rays": 0, "ms_total": 0.0, "extend_launches": 0}
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        for v in range(nv):
            scene.render_device(films[v].data_ptr(), W, H, spp, camera=cams[v], spp_offset=k * spp, **kw)
            st = scene.stats()
            for key in agg:
                agg[key] += st[key]
        log(f"[bench] step {k} done")
    torch.cuda.synchronize()
    seq_wall = time.perf_counter() - t0
    paths = W * H * spp * nv * args.steps
    means = [float(f.mean()) / (spp * args.steps) for f in films]
    finite = all(bool(torch.isfinite(f).all()) for f in films)
    # the value: every variant of a step in one xt_render_batch call
    bfilms = [torch.zeros((H, W, 3), dtype=torch.float32, device="cuda:0") for _ in range(nv)]

    def batch_step(k):
        api.render_batch([(scene, W, H, spp, (bfilms[v].data_ptr(), False),
                           dict(camera=cams[v], spp_offset=k * spp, **kw)) for v in range(nv)])
    for k in range(args.warmup):
        batch_step(args.steps + k)
    for f in bfilms:
        f.zero_()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        batch_step(k)
        log(f"[bench] batched step {k} done")
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    same_films = all(bool(torch.equal(a, b)) for a, b in zip(films, bfilms))
    del bfilms
    stages = {k: round(agg[k] / args.steps, 2) for k in ("ms_extend", "ms_shade", "ms_shadow", "ms_other")}
    # roofline of the dominant kernel, the shade step with the normal-map lookup and the random walk:
    # per queued path 64 B queue in + 16 hit + 48 triangle + 48 normals + 96 normal-map attributes +
    # 32 material + 64 (four bilinear texels) + 64 queue out + 48 shadow record + 32 radiance/pending
    launches = max(1, agg["extend_launches"])
    pmc = pmc_from_profiles("shade_displaced_plane4096_nmap.txt") if (args.plane_grid, W, H) == (4096, 3840, 2160) \
        else {}
    roofline = kernel_roofline("k_shade<false, true> (shade + normal map + two-facet random walk)",
                               agg["extension_rays"] * SHADE_NMAP_BYTES_PER_PATH / launches,
                               agg["ms_shade"] / launches, pmc, {"bytes_per_path": SHADE_NMAP_BYTES_PER_PATH})
    out = {
        "metric": "Mpaths/s (displaced plane, microfacet normal mapping)",
        "value": round(paths / wall / 1e6, 3),
        "unit": "Mpaths/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / args.steps, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"displaced_plane{args.plane_grid}_{W}x{H}_{spp}spp_x{nv}variants_depth{depth}_latlong2048",
                   "triangles": info["num_prims"], "width": W, "height": H, "spp": spp, "variants": nv,
                   "max_depth": depth, "material": "MicrofacetNormalMapping over GGX alpha 0.2, 1024^2 32-sine normal map",
                   "parallelism": "1 GPU"},
        "mrays_per_s": round((agg["extension_rays"] + agg["shadow_rays"]) / wall / 1e6, 1),
        "rays_per_path": round((agg["extension_rays"] + agg["shadow_rays"]) / paths, 3),
        "per_variant_renders": {"mpaths_per_s": round(paths / seq_wall / 1e6, 3),
                                "ms_per_step": round(seq_wall * 1e3 / args.steps, 2),
                                "films_equal_batched": same_films,
                                "note": "one xt_render per variant: stage times and the roofline come from these"},
        "stage_ms_per_step": stages,
        "roofline": roofline,
        "film_means": [round(m, 4) for m in means[:4]],
        "finite": finite,
        "build": {"scene_gen_s": round(gen_s, 1), "bvh_build_upload_s": round(upload_s, 1),
                  "bvh_nodes": info["bvh_nodes"], "bvh_depth": info["bvh_depth"]},
    }
    ref = None
    if not args.no_ref or args.guided:
        # converged reference of variant 0 at 960x540 (at 3840x2160 the 16,384-spp
        # FP64 reference would be 136 G paths)
        from xitils_amd import quality
        t0 = time.perf_counter()
        ref = quality.reference_image(scene, NMAP_QW, NMAP_QH, spp=args.ref_spp, chunk=512, seed=0x5EED, camera=cams[0],
                                      max_depth=depth, flags=flags)
        ref_s = time.perf_counter() - t0
        log(f"[bench] reference of variant 0 at {NMAP_QW}x{NMAP_QH}, {args.ref_spp} spp, FP64: {ref_s:.1f} s")
        if not args.no_ref:
            img = scene.render(NMAP_QW, NMAP_QH, NMAP_QSPP, camera=cams[0], seed=1, max_depth=depth, flags=flags) / NMAP_QSPP
            out["quality"] = {"relmse": quality.rel_mse(img, ref), "image_spp": NMAP_QSPP, "reference_spp": args.ref_spp,
                              "resolution": f"{NMAP_QW}x{NMAP_QH}", "camera": "variant 0",
                              "reference": "same renderer, FP64 film, seed 0x5EED",
                              "reference_seconds": round(ref_s, 2)}
    if args.guided:
        out.update(nmap_guided(args, scene, cams, W, H, spp, depth, flags, paths / wall / 1e6, ref))
    if not args.no_cpu_baseline:
        log("[bench] CPU baseline (oracle) ...")
        cdesc = desc
        cdesc.camera = cams[0]
        cargs = argparse.Namespace(**vars(args))
        cargs.width, cargs.height, cargs.spp = W, H, spp
        out["cpu_baseline"] = cpu_baseline(cdesc, cargs, depth, flags)
    print(json.dumps(out), flush=True)


NMAP_QW, NMAP_QH, NMAP_QSPP = 960, 540, 64   # the displaced plane's relMSE images


def nmap_guided(args, scene, cams, W, H, spp, depth, flags, plain_mpaths, ref):
    """The displaced plane with the guiding network: online training (one
    1-spp training render per camera variant), then timed steps that render
    every variant guided in one xt_render_batch call, then plain and guided
    at equal spp on variant 0 against the converged reference `ref`
    (NMAP_QW x NMAP_QH, args.ref_spp spp, FP64)."""
    import numpy as np
    import torch
    from xitils_amd import api, quality
    nv = len(cams)
    scratch = torch.zeros((H, W, 3), dtype=torch.float32, device="cuda:0")
    scene.guide_init(seed=1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    steps = samples = 0
    for v in range(nv):
        scene.render_device(scratch.data_ptr(), W, H, 1, camera=cams[v], seed=2, spp_offset=v, max_depth=depth,
                            flags=flags | api.XT_FLAG_GUIDE | api.XT_FLAG_GUIDE_TRAIN)
        st = scene.stats()
        steps += st["guide_steps"]
        samples += st["guide_samples"]
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t0
    log(f"[bench] guide trained: {nv} training renders, {steps} Adam steps, {samples} samples, {train_s:.2f} s")
    films = [torch.zeros((H, W, 3), dtype=torch.float32, device="cuda:0") for _ in range(nv)]
    gflags = flags | api.XT_FLAG_GUIDE

    def batch_step(k):
        api.render_batch([(scene, W, H, spp, (films[v].data_ptr(), False),
                           dict(camera=cams[v], seed=1, spp_offset=k * spp, max_depth=depth, flags=gflags))
                          for v in range(nv)])
    batch_step(0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        batch_step(k + 1)
        log(f"[bench] guided batch step {k} done")
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    guided_mpaths = W * H * spp * nv * args.steps / wall / 1e6
    finite = all(bool(torch.isfinite(f).all()) for f in films)
    # quality on variant 0 at reduced resolution, equal spp
    qw, qh, qspp = NMAP_QW, NMAP_QH, NMAP_QSPP
    res = {}
    for name, f in (("plain", flags), ("guided", gflags)):
        img = torch.zeros((qh, qw, 3), dtype=torch.float32, device="cuda:0")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        scene.render_device(img.data_ptr(), qw, qh, qspp, camera=cams[0], seed=3, max_depth=depth, flags=f)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        res[name] = {"relmse": quality.rel_mse(img.cpu().numpy().astype(np.float64) / qspp, ref), "seconds": dt}
    eff = (res["plain"]["relmse"] * res["plain"]["seconds"]) / (res["guided"]["relmse"] * res["guided"]["seconds"])
    log(f"[bench] guided {guided_mpaths:.1f} Mpaths/s (batched), quality {res}, efficiency ratio {eff:.3f}")
    return {
        "guided": {
            "mpaths_per_s": round(guided_mpaths, 2), "plain_mpaths_per_s": round(plain_mpaths, 2),
            "calls": "one xt_render_batch call of %d variants per step" % nv, "finite": finite,
            "training": {"renders": nv, "adam_steps": steps, "samples": samples, "seconds": round(train_s, 2)},
            "quality_variant0": {"resolution": f"{qw}x{qh}", "spp": qspp, "reference_spp": args.ref_spp,
                                 "plain": {k: round(v, 5) for k, v in res["plain"].items()},
                                 "guided": {k: round(v, 5) for k, v in res["guided"].items()},
                                 "efficiency_ratio": round(eff, 3),
                                 "note": "relMSE x render time, plain over guided (>1: guiding pays)"},
        }
    }


def rays_main(args):
    """Ray-query microbenchmark on the 1024^2-cell height field: one batch
    each of 1080p primary rays, cosine-weighted diffuse bounce rays from
    random surface points and shadow rays into the sun's cone (seed 7,
    2,073,600 rays each), resident in HBM, through xt_trace_closest
    (primary, diffuse) and xt_trace_any (shadow).  A step traces all three
    batches; value = rays / summed kernel time (HIP events on the library's
    stream)."""
    import numpy as np
    import torch
    from xitils_amd import api, scenes
    W, H = args.width, args.height
    log(f"[bench] building heightfield {args.cells}^2 cells")
    desc = scenes.heightfield(cells=args.cells, width=W, height=H)
    dev = api.Device(0)
    scene = api.Scene(dev, desc)
    info = scene.info()
    node_bytes = info["bytes_nodes"] // max(1, info["bvh_nodes"])
    log(f"[bench] scene {info}")
    if args.calib_copy:
        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda:0")   # 1 GiB read + 1 GiB written
        dst = torch.empty_like(src)
        torch.mul(src, 1.0, out=dst)
        torch.cuda.synchronize()
        del src, dst
        log("[bench] calibration copy done (1 GiB in, 1 GiB out)")
    batches, per_kind = {}, {}
    for kind in scenes.RAY_BATCH_KINDS:
        rays = scenes.ray_batch(desc, kind, W, H)
        batches[kind] = rays
        n = rays.shape[0]
        q = "any" if kind == "shadow" else "closest"
        d_rays = torch.from_numpy(rays).to("cuda:0")
        out = torch.empty(n, dtype=torch.uint8, device="cuda:0") if q == "any" else \
            torch.empty((n, 4), dtype=torch.float32, device="cuda:0")
        call = scene.trace_any_dev if q == "any" else scene.trace_closest_dev
        torch.cuda.synchronize()
        scene.set_counting(True)
        call(d_rays.data_ptr(), n, out.data_ptr())
        cst = scene.stats()
        scene.set_counting(False)
        for _ in range(max(1, args.warmup)):
            call(d_rays.data_ptr(), n, out.data_ptr())
        key = "ms_shadow" if q == "any" else "ms_extend"
        ms = []
        for _ in range(args.steps):
            call(d_rays.data_ptr(), n, out.data_ptr())
            ms.append(scene.stats()[key])
        avg = float(np.mean(ms))
        nodes = cst["shadow_node_fetches" if q == "any" else "ext_node_fetches"]
        tris = cst["shadow_tri_fetches" if q == "any" else "ext_tri_fetches"]
        algo = n * QUERY_IO_BYTES[q] + nodes * node_bytes + tris * TRI_BYTES
        res = out.cpu().numpy()
        hit = float(res.mean()) if q == "any" else float(np.mean(res[:, 1].view(np.uint32) != 0xFFFFFFFF))
        per_kind[kind] = {
            "query": q, "rays": n, "avg_kernel_ms": round(avg, 4), "mrays_per_s": round(n / avg / 1e3, 1),
            "node_fetches_per_ray": round(nodes / n, 3), "tri_fetches_per_ray": round(tris / n, 3),
            "hit_fraction": round(hit, 4), "algorithmic_bytes_per_ray": round(algo / n, 1),
            "achieved_gbs": round(algo / (avg * 1e-3) / 1e9, 1),
        }
        log(f"[bench] {kind}: {per_kind[kind]}")
        del d_rays, out
    total_ms = sum(v["avg_kernel_ms"] for v in per_kind.values())
    total_rays = sum(v["rays"] for v in per_kind.values())
    dom_kind = max(per_kind, key=lambda k: per_kind[k]["avg_kernel_ms"])
    dom = per_kind[dom_kind]
    pmc = pmc_from_profiles(f"rays_{dom_kind}_heightfield1080p_pmc.txt") if (args.cells, W, H) == (1024, 1920, 1080) \
        else {}
    roofline = kernel_roofline(f"k_trace<{'any' if dom['query'] == 'any' else 'closest'}> ({dom_kind} rays)",
                                  dom["algorithmic_bytes_per_ray"] * dom["rays"], dom["avg_kernel_ms"], pmc)
    out = {
        "metric": "Mrays/s (ray-query microbenchmark: primary, diffuse, shadow)",
        "value": round(total_rays / total_ms / 1e3, 1),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(total_ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"raybatch_heightfield{args.cells}_{W}x{H}_seed7", "triangles": info["num_prims"],
                   "rays_per_kind": W * H, "parallelism": "1 GPU"},
        "per_kind": per_kind,
        "roofline": roofline,
    }
    if not args.no_cpu_baseline:
        log("[bench] CPU baseline (oracle) ...")
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        o = pyoracle.OracleScene(desc)
        budget = args.cpu_seconds / len(batches)
        done = secs = 0.0
        agree = []
        for kind, rays in batches.items():
            fn = (lambda r: o.trace_any(r, mode=0, nthreads=threads)) if kind == "shadow" else \
                (lambda r: o.trace_closest(r, mode=0, nthreads=threads))
            t0 = time.perf_counter()
            fn(rays[:20000])
            one = (time.perf_counter() - t0) / 20000
            m = int(min(rays.shape[0], max(20000, budget / max(one, 1e-9))))
            t0 = time.perf_counter()
            ref = fn(rays[:m])
            secs += time.perf_counter() - t0
            done += m
            got = scene.trace_any(rays[:m]) if kind == "shadow" else scene.trace_closest(rays[:m])
            same = np.array_equal(ref, got) if kind == "shadow" else \
                np.array_equal(ref[:, 1].view(np.uint32), got[:, 1].view(np.uint32))
            agree.append(f"{kind} {m} rays{'' if same else ' (MISMATCH)'}")
        out["cpu_baseline"] = {
            "value": round(done / secs / 1e6, 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": f"first rays of each batch ({', '.join(agree)}; GPU results identical where not marked) "
                      f"through the oracle C++ restatement's reference-shaped binary BVH, {secs:.1f} s wall",
        }
    print(json.dumps(out), flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}): the renderer shards by sample range "
              f"with the scene replicated per GPU and one film reduction, which is not built", file=sys.stderr)
        sys.exit(2)
    if args.workload == "guiding":
        return guiding_main(args)
    if args.workload == "batch":
        return batch_main(args)
    if args.workload == "cornell":
        return cornell_main(args)
    if args.workload == "nmap":
        return nmap_main(args)
    if args.workload == "rays":
        return rays_main(args)
    import numpy as np
    import torch
    from xitils_amd import api, scenes

    W, H, spp, max_depth = args.width, args.height, args.spp, args.max_depth
    flags = scenes.FLAG_ENV_NEE
    log(f"[bench] building heightfield {args.cells}^2 cells")
    t0 = time.perf_counter()
    desc = scenes.heightfield(cells=args.cells, width=W, height=H)
    gen_s = time.perf_counter() - t0
    dev = api.Device(0)
    t0 = time.perf_counter()
    scene = api.Scene(dev, desc)
    upload_s = time.perf_counter() - t0
    info = scene.info()
    log(f"[bench] scene {info}")
    film = torch.zeros((H, W, 3), dtype=torch.float32, device="cuda:0")
    if args.calib_copy:
        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda:0")   # 1 GiB read + 1 GiB written
        dst = torch.empty_like(src)
        torch.mul(src, 1.0, out=dst)
        torch.cuda.synchronize()
        del src, dst
        log("[bench] calibration copy done (1 GiB in, 1 GiB out)")
    kw = dict(seed=1, max_depth=max_depth, flags=flags)

    # fetch counts of the identical work (same seed, same rays): one untimed counting render
    scene.set_counting(True)
    scene.render_device(film.data_ptr(), W, H, spp, **kw)
    torch.cuda.synchronize()
    cst = scene.stats()
    scene.set_counting(False)
    film.zero_()
    log(f"[bench] counting pass: {cst}")

    for k in range(args.warmup):
        scene.render_device(film.data_ptr(), W, H, spp, spp_offset=0, **kw)
        log(f"[bench] warmup {k} done")
    torch.cuda.synchronize()
    agg = {"ms_extend": 0.0, "ms_shade": 0.0, "ms_shadow": 0.0, "ms_other": 0.0, "extension_rays": 0,
           "shadow_rays": 0, "extend_launches": 0, "ms_total": 0.0}
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        scene.render_device(film.data_ptr(), W, H, spp, **kw)
        st = scene.stats()
        for key in agg:
            agg[key] += st[key]
        log(f"[bench] step {k}: {st['ms_total']:.1f} ms")
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1e3 / args.steps
    paths = W * H * spp * args.steps
    mpaths = paths / wall / 1e6
    rays = agg["extension_rays"] + agg["shadow_rays"]
    # roofline of the closest-hit traversal (dominant kernel)
    n_ext = cst["extension_rays"]
    node_bytes = info["bytes_nodes"] // max(1, info["bvh_nodes"])
    bytes_ext = n_ext * EXT_RAY_IO_BYTES + cst["ext_node_fetches"] * node_bytes + cst["ext_tri_fetches"] * TRI_BYTES
    per_step_ext_bytes = bytes_ext
    launches = agg["extend_launches"] / args.steps
    pmc = pmc_from_profiles("extend_heightfield1080p_pmc.txt") if (args.cells, W, H, spp) == (1024, 1920, 1080, 64) \
        else {}
    roofline = kernel_roofline(
        "k_trace<closest> (extend)", per_step_ext_bytes / launches, agg["ms_extend"] / agg["extend_launches"], pmc,
        {"per_ray": {"node_fetches": round(cst["ext_node_fetches"] / n_ext, 3),
                     "tri_fetches": round(cst["ext_tri_fetches"] / n_ext, 3),
                     "bytes": round(bytes_ext / n_ext, 1)}})
    out = {
        "metric": "Mpaths/s at 1080p",
        "value": round(mpaths, 3),
        "unit": "Mpaths/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"heightfield{args.cells}_{W}x{H}_{spp}spp_depth{max_depth}_sky+sun_envNEE",
                   "triangles": info["num_prims"], "width": W, "height": H, "spp": spp, "max_depth": max_depth,
                   "integrator": "StandardPathTracer (NEE+BSDF MIS)", "parallelism": "1 GPU"},
        "mrays_per_s": round(rays / wall / 1e6, 2),
        "stage_ms_per_step": {k: round(agg[k] / args.steps, 2) for k in ("ms_extend", "ms_shade", "ms_shadow", "ms_other")},
        "build": {"scene_gen_s": round(gen_s, 2), "bvh_build_upload_s": round(upload_s, 2),
                  "bvh_nodes": info["bvh_nodes"], "bvh_depth": info["bvh_depth"]},
        "roofline": roofline,
    }
    if not args.no_ref:
        log(f"[bench] relMSE of one {spp}-spp image against a {args.ref_spp}-spp FP64 reference ...")
        out["quality"] = image_quality(scene, W, H, spp, args.ref_spp, **kw)
        log(f"[bench] quality {out['quality']}")
    if not args.no_cpu_baseline:
        log("[bench] CPU baseline (oracle) ...")
        out["cpu_baseline"] = cpu_baseline(desc, args, max_depth, flags)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
