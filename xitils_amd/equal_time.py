"""Equal-time comparison of guided and plain path tracing.

Both renderers accumulate progressively into a device film (each call adds
`spp` samples at the next sample index, the way the reference's
RenderTarget::render is called once per frame, RenderTarget.h:110-127) until a
wall-clock budget is spent.  After each call the running mean is compared
with a converged reference (relMSE as quality.rel_mse, computed on the device
in float64 and read back after the run).  The clock runs from the first call
to the last without pausing: the guided renderer's Adam steps run on a stream
of their own while the next render runs, and time the GPU spends on them
after a call returns counts against the calls that follow.  The error
computation is inside the clock for both renderers (tens of microseconds a
call).  The guided renderer trains its network online during the first
`train_fraction` of its budget (one sample per pixel per call, Adam steps
after each call) and samples with the trained network after that; its
training samples stay in the film.

The film is a torch tensor: import torch before the library opens its device,
so that both share one HIP runtime.
"""
from __future__ import annotations

import math
import time

import numpy as np

from . import api

# Share of the budget spent training online (one sample per pixel per call,
# Adam steps after each call); chosen on the occluded room (see compare).
TRAIN_FRACTION = 0.15


def progressive(scene, width, height, budget_s, ref, *, guided=False, train_fraction=TRAIN_FRACTION, spp_per_call=4,
                train_spp_per_call=1, seed=1, flags=0, max_depth=16, max_spp=1 << 20, keep_training_samples=True,
                log=None):
    """Returns the error curve [(seconds, spp, relMSE)] of one progressive run.
    keep_training_samples=False restarts the film when training ends (the
    samples rendered with the untrained network are dropped; the sample
    index keeps counting, so no sample is rendered twice)."""
    import torch
    film = torch.zeros((height, width, 3), dtype=torch.float32, device="cuda:0")
    ref_d = torch.as_tensor(np.asarray(ref, np.float64), device="cuda:0")
    inv_den = 1.0 / (ref_d * ref_d + 0.01)   # quality.rel_mse's denominator
    points, errs = [], []
    t = 0.0
    spp = 0          # sample index of the next call
    film_spp = 0     # samples in the film
    was_training = False
    steps = samples = queries = grad_samples = 0
    infer_ms = train_ms = 0.0
    torch.cuda.synchronize()
    start = time.perf_counter()
    while t < budget_s and spp < max_spp:
        training = guided and t < train_fraction * budget_s
        if was_training and not training and not keep_training_samples:
            film.zero_()
            film_spp = 0
        was_training = training
        f = flags | (api.XT_FLAG_GUIDE if guided else 0) | (api.XT_FLAG_GUIDE_TRAIN if training else 0)
        n = train_spp_per_call if training else spp_per_call
        # the film's readers and writers on torch's stream have finished (the
        # render writes it on the library's stream); the library's training
        # stream is not waited for
        torch.cuda.current_stream().synchronize()
        scene.render_device(film.data_ptr(), width, height, n, spp_offset=spp, seed=seed, flags=f, max_depth=max_depth)
        spp += n
        film_spp += n
        d = film.double() * (1.0 / film_spp) - ref_d
        errs.append((d * d * inv_den).mean())
        t = time.perf_counter() - start
        points.append((t, spp))
        st = scene.stats()
        steps += st["guide_steps"]
        samples += st["guide_samples"]
        infer_ms += st["ms_guide_infer"]
        queries += st["guide_queries"]
        grad_samples += st["guide_grad_samples"]
        train_ms += st["ms_guide_train"]
    torch.cuda.synchronize()
    e_host = torch.stack(errs).cpu().numpy() if errs else np.zeros(0)
    curve = [(a, b, float(e)) for (a, b), e in zip(points, e_host)]
    if log:
        log(f"[equal-time] guided={guided}: {spp} spp in {t:.3f} s, relMSE {curve[-1][2]:.4e}, "
            f"adam steps {steps}, samples {samples}, infer {infer_ms:.1f} ms, train {train_ms:.1f} ms")
    return curve, {"adam_steps": steps, "train_samples": samples, "ms_guide_infer": infer_ms,
                   "ms_guide_train": train_ms, "guide_queries": queries, "grad_samples": grad_samples}


def time_to_reach(curve, target):
    """First time the curve's relMSE drops to `target` (log-linear between
    checkpoints), or None."""
    prev = None
    for t, _, e in curve:
        if e <= target:
            if prev is None or prev[1] <= e:
                return t
            t0, e0 = prev
            a = (math.log(e0) - math.log(target)) / (math.log(e0) - math.log(e))
            return t0 + a * (t - t0)
        prev = (t, e)
    return None


def compare(scene, width, height, budget_s, ref, *, log=None, guide_seeds=None, render_seeds=None,
            guide_config=None, **kw):
    """Runs plain, then guided for the same budget; returns a summary dict.
    Plain runs once per render seed (render_seeds, default [kw's seed or 1]);
    their mean final relMSE is the plain error.  With guide_seeds, the guided
    run is repeated from a fresh network per seed (scene.guide_init(seed=...,
    **guide_config)), each with a render seed of its own (100 + the plain
    seeds, cycled), and the run with the median ratio is reported, all runs
    listed (training is stochastic: single runs spread by about 10%).
    Without, one guided run uses the scene's current network."""
    seeds = list(render_seeds) if render_seeds else [kw.pop("seed", 1)]
    kw.pop("seed", None)
    plains = []
    for rs in seeds:
        curve, _ = progressive(scene, width, height, budget_s, ref, guided=False, log=log, seed=rs, **kw)
        plains.append({"seed": rs, "curve": curve})
    e_plain = float(np.mean([p["curve"][-1][2] for p in plains]))
    t_plain = float(np.mean([p["curve"][-1][0] for p in plains]))
    runs = []
    for i, gseed in enumerate(guide_seeds or [None]):
        if gseed is not None:
            scene.guide_init(seed=gseed, **(guide_config or {}))
        rs = 100 + seeds[i % len(seeds)]
        guided, gstats = progressive(scene, width, height, budget_s, ref, guided=True, log=log, seed=rs, **kw)
        t_reach = time_to_reach(guided, e_plain)
        runs.append({"seed": gseed, "render_seed": rs, "curve": guided, "stats": gstats, "relmse": guided[-1][2],
                     "ratio": e_plain / guided[-1][2], "t_reach": t_reach})
    med = sorted(runs, key=lambda r: r["ratio"])[len(runs) // 2]
    guided, gstats, t_reach = med["curve"], med["stats"], med["t_reach"]
    plain = plains[0]["curve"]
    return {
        "budget_s": budget_s,
        "plain": {"spp": int(np.mean([p["curve"][-1][1] for p in plains])), "seconds": round(t_plain, 4),
                  "relmse": e_plain},
        "plain_runs": [{"seed": p["seed"], "spp": p["curve"][-1][1], "relmse": p["curve"][-1][2]} for p in plains],
        "guided": {"spp": guided[-1][1], "seconds": round(guided[-1][0], 4), "relmse": guided[-1][2],
                   "seed": med["seed"], **gstats},
        "relmse_ratio_equal_time": med["ratio"],
        "guided_seconds_to_plain_relmse": t_reach,
        "time_speedup": (t_plain / t_reach) if t_reach else None,
        "runs": [{"seed": r["seed"], "render_seed": r["render_seed"], "spp": r["curve"][-1][1], "relmse": r["relmse"],
                  "ratio": round(r["ratio"], 4), "time_speedup": (t_plain / r["t_reach"]) if r["t_reach"] else None}
                 for r in runs],
        "curves": {"plain": [(round(a, 4), b, c) for a, b, c in plain[:: max(1, len(plain) // 40)]],
                   "guided": [(round(a, 4), b, c) for a, b, c in guided[:: max(1, len(guided) // 40)]]},
    }
