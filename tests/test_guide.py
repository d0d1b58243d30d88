"""Guiding network: the numpy reference (tests/guide_ref.py) is pinned by
finite differences and closed forms on the CPU; the MFMA kernels are checked
against it on the GPU, and guided rendering against plain rendering.

Tolerances: bf16 rounding happens at the same points on both sides, but the
fp32 accumulation order differs (MFMA vs numpy), which can move a hidden
activation by one bf16 ulp (2^-8 relative); outputs and gradients are
compared to 2e-2 of their scale, means to 1e-3."""
import numpy as np
import pytest

import guide_ref as G
from xitils_amd import api, scenes

LO = np.array([-1.0, -0.5, -2.0])
HI = np.array([3.0, 1.5, 1.0])


# ---------------------------------------------------------------- CPU pins

def test_bf16_rounding():
    x = np.array([1.0, 1.00390625, 1.005859375, 1.001953125, -3.14159265, 0.0, 65504.0], np.float32)
    r = G.bf16(x)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7 -> even (1.0); 1 + 3*2^-9 rounds up
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == np.float32(1.0078125) and r[3] == 1.0
    assert r[4] == np.float32(-3.140625) and r[5] == 0.0 and r[6] == 65536.0


def test_encoding_closed_form():
    q = G.random_queries(64, LO, HI, 1)
    x = G.encode(q, LO, HI, np.float64)
    u = (q[:, :3] - LO) / (HI - LO)
    for a in range(3):
        for k in range(G.G_BANDS):
            np.testing.assert_allclose(x[:, 24 * a + 2 * k], np.sin(np.pi * 2 ** k * u[:, a]), atol=1e-9)
            np.testing.assert_allclose(x[:, 24 * a + 2 * k + 1], np.cos(np.pi * 2 ** k * u[:, a]), atol=1e-9)
    np.testing.assert_array_equal(x[:, 72:78], q[:, 3:9])
    assert not x[:, 78:].any()


def test_time_to_reach_is_log_linear():
    """equal_time.time_to_reach: the first crossing of the error curve, with
    log(error) linear in time between checkpoints (closed form here)."""
    from xitils_amd import equal_time
    curve = [(1.0, 4, 1.0), (2.0, 8, 0.5), (3.0, 12, 0.25)]
    assert equal_time.time_to_reach(curve, 2.0) == 1.0        # reached at the first checkpoint
    assert equal_time.time_to_reach(curve, 0.5) == 2.0        # exactly at a checkpoint
    # between 2 s (0.5) and 3 s (0.25): 0.35 = 0.5 * 2^-a at a = log2(0.5 / 0.35)
    assert equal_time.time_to_reach(curve, 0.35) == pytest.approx(2.0 + np.log2(0.5 / 0.35), rel=1e-12)
    assert equal_time.time_to_reach(curve, 0.1) is None       # never reached


def test_vmf_pdf_integrates_to_one():
    # midpoint rule over (cos theta, phi) about a tilted mean direction
    mu = np.array([0.3, -0.5, 0.81])
    mu /= np.linalg.norm(mu)
    n = 2000
    c = -1 + (np.arange(n) + 0.5) * 2 / n
    ph = (np.arange(n) + 0.5) * 2 * np.pi / n
    C, Ph = np.meshgrid(c, ph, indexing="ij")
    S = np.sqrt(1 - C * C)
    w = np.stack([S * np.cos(Ph), S * np.sin(Ph), C], -1).reshape(-1, 3)
    for k in (0.05, 1.0, 10.0, 200.0):
        integral = G.vmf_pdf(mu, k, w).sum() * (2 / n) * (2 * np.pi / n)
        assert abs(integral - 1) < 2e-3, (k, integral)


def test_backward_matches_finite_differences():
    P = G.init_params(3).astype(np.float64)
    q = G.random_queries(8, LO, HI, 5)
    d = G.random_dirs(8, 6)
    w = np.random.default_rng(7).uniform(0.1, 2, 8)
    g = G.backward(P, q, d, w, LO, HI, emulate=False)
    rng = np.random.default_rng(0)
    idx = np.concatenate([rng.choice(np.arange(a, b), k) for a, b, k in
                          [(G.OFF_W0, G.OFF_B0, 6), (G.OFF_B0, G.OFF_W1, 3), (G.OFF_W1, G.OFF_B1, 4),
                           (G.OFF_W2, G.OFF_B2, 4), (G.OFF_B2, G.OFF_W3, 3), (G.OFF_W3, G.OFF_W3 + 40 * 64, 8),
                           (G.OFF_B3, G.OFF_B3 + 40, 8)]])
    for i in idx:
        h = 1e-6
        Pp, Pm = P.copy(), P.copy()
        Pp[i] += h
        Pm[i] -= h
        fd = (G.loss(Pp, q, d, w, LO, HI) - G.loss(Pm, q, d, w, LO, HI)) / (2 * h)
        assert abs(g[i] - fd) <= 1e-6 + 1e-5 * abs(fd), (i, g[i], fd)
    # the padded output rows 40..47 carry no gradient
    assert not g[G.OFF_W3 + 40 * 64:G.OFF_B3].any() and not g[G.OFF_B3 + 40:].any()


def test_adam_first_step_is_signed_lr():
    g = np.array([1e-3, -2.0, 5.0, 0.0])
    P, m, v = G.adam(np.zeros(4), np.zeros(4), np.zeros(4), g, 1, lr=0.01)
    np.testing.assert_allclose(P[:3], -0.01 * np.sign(g[:3]), rtol=1e-4)
    assert P[3] == 0.0


# ---------------------------------------------------------------- GPU

@pytest.fixture(scope="module")
def gscene():
    print("[gpu] guide: opening device 0, cornell scene", flush=True)
    dev = api.Device(0)
    sc = api.Scene(dev, scenes.cornell(32, 32))
    sc.guide_init(seed=3)
    return dev, sc


def _bounds(sc):
    lo, hi = sc.guide_bounds()
    return lo.astype(np.float64), hi.astype(np.float64)


@pytest.mark.gpu
def test_guide_init_and_bounds(gscene):
    _, sc = gscene
    lo, hi = _bounds(sc)
    np.testing.assert_allclose(lo, [0, 0, 0], atol=1e-6)
    np.testing.assert_allclose(hi, [1, 1, 1], atol=1e-6)
    P = sc.guide_params()
    assert P.shape == (G.NPARAMS,) and np.isfinite(P).all()
    assert not P[G.OFF_W3 + 40 * 64:G.OFF_B3].any()     # padded rows stay zero
    sc.guide_set_params(P * 0.5)
    np.testing.assert_array_equal(sc.guide_params(), (P * 0.5).astype(np.float32))


@pytest.mark.gpu
def test_guide_forward_matches_reference(gscene):
    _, sc = gscene
    lo, hi = _bounds(sc)
    P = G.init_params(11)
    sc.guide_set_params(P)
    for n in (1, 63, 64, 1000, 5000):
        print(f"[gpu] guide forward n={n}", flush=True)
        q = G.random_queries(n, lo - 0.2, hi + 0.2, n)          # includes clamped positions
        y = sc.guide_forward(q)
        yr, _ = G.forward(P, G.encode(q, lo, hi), emulate=True)
        yr = yr[:, :40]
        scale = np.abs(yr).max()
        err = np.abs(y - yr)
        print(f"[gpu]   max|y|={scale:.3f} max err={err.max():.2e} mean err={err.mean():.2e}", flush=True)
        assert err.max() <= 2e-2 * scale and err.mean() <= 1e-3 * scale


@pytest.mark.gpu
def test_guide_gradient_matches_reference(gscene):
    _, sc = gscene
    lo, hi = _bounds(sc)
    P = G.init_params(12)
    sc.guide_set_params(P)
    for n, scale, clip in ((1, 1.0, np.inf), (1000, 0.5, np.inf), (4096, 2.0, 3.0)):
        print(f"[gpu] guide gradient n={n}", flush=True)
        q = G.random_queries(n, lo, hi, 100 + n)
        d = G.random_dirs(n, 200 + n)
        w = np.random.default_rng(n).exponential(1.0, n).astype(np.float32)
        g = sc.guide_gradient(q, d, w, scale, clip)
        gr = G.backward(P, q, d, w, lo, hi, emulate=True, weight_scale=scale, weight_clip=clip)
        g64 = G.backward(P, q, d, w, lo, hi, emulate=False, weight_scale=scale, weight_clip=clip)
        rel = np.linalg.norm(g - gr) / np.linalg.norm(gr)
        rel64 = np.linalg.norm(g - g64) / np.linalg.norm(g64)
        print(f"[gpu]   rel err vs bf16 reference {rel:.2e}, vs fp64 {rel64:.2e}", flush=True)
        assert rel <= 2e-2 and rel64 <= 5e-2
        for a, b in ((G.OFF_W0, G.OFF_B0), (G.OFF_W1, G.OFF_B1), (G.OFF_W2, G.OFF_B2), (G.OFF_W3, G.OFF_B3),
                     (G.OFF_B3, G.NPARAMS)):
            blk = np.linalg.norm(g[a:b] - gr[a:b]) / max(np.linalg.norm(gr[a:b]), 1e-30)
            assert blk <= 5e-2, (a, b, blk)


@pytest.mark.gpu
def test_guide_gradient_over_several_block_passes(gscene):
    # more samples than one pass of the gradient kernel's blocks holds (one
    # 128-sample group per CU): the blocks add later groups into their rows
    _, sc = gscene
    lo, hi = _bounds(sc)
    P = G.init_params(15)
    sc.guide_set_params(P)
    n = 70001
    print(f"[gpu] guide gradient n={n}", flush=True)
    q = G.random_queries(n, lo, hi, 7)
    d = G.random_dirs(n, 8)
    w = np.random.default_rng(9).exponential(1.0, n).astype(np.float32)
    g = sc.guide_gradient(q, d, w, 1.0, 4.0)
    gr = G.backward(P, q, d, w, lo, hi, emulate=True, weight_scale=1.0, weight_clip=4.0)
    rel = np.linalg.norm(g - gr) / np.linalg.norm(gr)
    print(f"[gpu]   rel err vs bf16 reference {rel:.2e}", flush=True)
    assert rel <= 2e-2


@pytest.mark.gpu
def test_adam_repack_equals_prepare(gscene):
    """Adam writes each updated parameter into the bf16 inference and backward
    copies itself; those must equal the copies prepared from the parameters."""
    _, sc = gscene
    lo, hi = _bounds(sc)
    sc.guide_set_params(G.init_params(14))
    n = 3000
    q = G.random_queries(n, lo, hi, 4)
    d = G.random_dirs(n, 5)
    w = np.random.default_rng(6).exponential(1.0, n).astype(np.float32)
    for _ in range(3):
        sc.guide_train_step(q, d, w)
    y1 = sc.guide_forward(q)
    g1 = sc.guide_gradient(q, d, w, 1.0, 5.0)
    sc.guide_set_params(sc.guide_params())       # re-prepares both copies from the fp32 parameters
    y2 = sc.guide_forward(q)
    g2 = sc.guide_gradient(q, d, w, 1.0, 5.0)
    assert np.array_equal(y1, y2)
    assert np.array_equal(g1, g2)


@pytest.mark.gpu
def test_guide_train_step_is_adam(gscene):
    _, sc = gscene
    lo, hi = _bounds(sc)
    P = G.init_params(13)
    sc.guide_set_params(P)
    n = 2048
    q = G.random_queries(n, lo, hi, 1)
    d = G.random_dirs(n, 2)
    w = np.random.default_rng(3).exponential(1.0, n).astype(np.float32)
    sc.guide_train_step(q, d, w)
    P1 = sc.guide_params().astype(np.float64)
    g = G.backward(P, q, d, w, lo, hi, emulate=True, weight_scale=n / w.astype(np.float64).sum(), weight_clip=20.0) / n
    sig = np.abs(g) > 1e-3 * np.abs(g).max()
    dP = P1 - P
    # first Adam step: -lr * g / (|g| + eps) = -lr * sign(g) where |g| >> eps
    # (lr: the library default the fixture's guide_init used)
    lr = float(np.float32(api.Scene.guide_config().learning_rate))
    assert lr > 0.0
    np.testing.assert_allclose(dP[sig], -lr * np.sign(g[sig]), rtol=1e-3, atol=1e-6)
    assert np.abs(dP).max() <= lr * (1 + 1e-5)


@pytest.mark.gpu
def test_guide_learns_a_concentrated_target(gscene):
    """Training on directions drawn from a vMF about each query's normal
    (kappa = 40) concentrates the predicted mixture there."""
    _, sc = gscene
    lo, hi = _bounds(sc)
    sc.guide_init(seed=5)
    rng = np.random.default_rng(9)

    def batch(n, seed):
        q = G.random_queries(n, lo, hi, seed)
        nrm = q[:, 3:6].astype(np.float64)
        # Jakob's sampler about nrm with kappa = 40
        k = 40.0
        u1, u2 = rng.random(n), rng.random(n)
        W = 1 + np.log(u1 + (1 - u1) * np.exp(-2 * k)) / k
        t = np.stack([-nrm[:, 1], nrm[:, 0], np.zeros(n)], 1) + 1e-9
        t /= np.linalg.norm(t, axis=1, keepdims=True)
        b = np.cross(nrm, t)
        phi = 2 * np.pi * u2
        s = np.sqrt(np.maximum(0, 1 - W * W))
        d = W[:, None] * nrm + (s * np.cos(phi))[:, None] * t + (s * np.sin(phi))[:, None] * b
        return q, d.astype(np.float32)

    qv, dv = batch(2048, 999)
    before = np.log(G.mixture_pdf(sc.guide_forward(qv).astype(np.float64), dv)).mean()
    for it in range(150):
        q, d = batch(4096, it)
        sc.guide_train_step(q, d, np.ones(len(q), np.float32))
    after = np.log(G.mixture_pdf(sc.guide_forward(qv).astype(np.float64), dv)).mean()
    print(f"[gpu] mean log q: before {before:.3f} after {after:.3f}", flush=True)
    # uniform sphere: log(1/4pi) = -2.53; the target's own entropy bound: log(40/2pi) - 1 = 0.85
    assert before < -1.5 and after > 0.0


def _mean_err(a, b):
    return abs(a.mean() - b.mean()) / b.mean()


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["cornell", "material_zoo"])
def test_guided_render_matches_plain(name):
    """Guided sampling (random and trained networks) is an unbiased estimator
    of the same image as plain NEE + BSDF MIS."""
    print(f"[gpu] guided vs plain: {name}", flush=True)
    dev = api.Device(0)
    desc = scenes.cornell(32, 32) if name == "cornell" else scenes.material_zoo(32, 32)
    sc = api.Scene(dev, desc)
    sc.guide_init(seed=7)
    W = H = 32
    spp = 16384   # a render's image mean has sd ~0.1% here (heavy-tailed specular paths)
    plain = sc.render(W, H, spp, seed=11, max_depth=8) / spp
    guided = sc.render(W, H, spp, seed=12, max_depth=8, flags=api.XT_FLAG_GUIDE) / spp
    st = sc.stats()
    assert st["ms_guide_infer"] > 0 and st["guide_steps"] == 0
    e_rand = _mean_err(guided, plain)
    # train for a few progressive calls, then render with the trained network
    for k in range(8):
        sc.render(W, H, 16, seed=100 + k, max_depth=8, flags=api.XT_FLAG_GUIDE | api.XT_FLAG_GUIDE_TRAIN)
        st = sc.stats()
        assert st["guide_steps"] > 0 and st["guide_samples"] > 0
    trained = sc.render(W, H, spp, seed=13, max_depth=8, flags=api.XT_FLAG_GUIDE) / spp
    e_tr = _mean_err(trained, plain)
    from helpers import rel_mse
    r1, r2 = rel_mse(guided, plain), rel_mse(trained, plain)
    print(f"[gpu]   mean rel err: random {e_rand:.2e} trained {e_tr:.2e}; relMSE {r1:.2e} {r2:.2e}", flush=True)
    assert np.isfinite(guided).all() and np.isfinite(trained).all()
    assert e_rand < 5e-3 and e_tr < 5e-3
    assert r1 < 5e-3 and r2 < 5e-3


@pytest.mark.gpu
def test_guided_furnace_closed_form():
    """Convex Lambert box (albedo a) under a uniform sky of 1: every hit pixel
    converges to a whatever the guide samples (random and trained nets)."""
    print("[gpu] guided furnace", flush=True)
    a = 0.6
    dev = api.Device(0)
    sc = api.Scene(dev, scenes.furnace_object(a))
    sc.guide_init(seed=2)
    spp = 2048
    for trained in (False, True):
        if trained:
            for k in range(6):
                sc.render(32, 32, 8, seed=50 + k, flags=api.XT_FLAG_GUIDE | api.XT_FLAG_GUIDE_TRAIN)
        v = sc.render(32, 32, spp, seed=9 if trained else 8, flags=api.XT_FLAG_GUIDE) / spp
        # pixels whose film footprint (pixel +- half a pixel of jitter) lies on
        # the box: fully covered in a plain render, eroded by one pixel
        plain = sc.render(32, 32, 256, seed=3) / 256
        full = np.all(np.abs(plain - a) <= 1e-5, axis=2)
        empty = np.all(plain == 1.0, axis=2)
        hit = full.copy()
        hit[1:-1, 1:-1] = full[:-2, :-2] & full[:-2, 1:-1] & full[:-2, 2:] & full[1:-1, :-2] & full[1:-1, 1:-1] & \
            full[1:-1, 2:] & full[2:, :-2] & full[2:, 1:-1] & full[2:, 2:]
        hit[0, :] = hit[-1, :] = hit[:, 0] = hit[:, -1] = False
        px = v[hit].mean(axis=1)
        print(f"[gpu]   trained={trained}: {hit.sum()} interior pixels, mean {px.mean():.5f} (a = {a}), "
              f"worst pixel {np.abs(px - a).max():.4f}", flush=True)
        assert hit.sum() > 50
        assert abs(px.mean() - a) < 2e-3 * a
        assert np.abs(px - a).max() < 0.05 * a


@pytest.mark.gpu
def test_equal_time_compare_reports_the_median_run():
    """equal_time.compare with several guide seeds: one plain run, one guided
    run per fresh network, and the reported ratio is the median of theirs."""
    import torch  # noqa: F401  (the films are torch tensors on the library's HIP runtime)
    from xitils_amd import equal_time, quality
    desc = scenes.cornell(48, 48)
    print("[gpu] equal-time compare, cornell 48x48, 3 seeds", flush=True)
    sc = api.Scene(api.Device(0), desc)
    ref = quality.reference_image(sc, 48, 48, spp=256, chunk=64, max_depth=8)
    res = equal_time.compare(sc, 48, 48, 0.15, ref, max_depth=8, guide_seeds=[1, 2, 3])
    ratios = sorted(r["ratio"] for r in res["runs"])
    assert [r["seed"] for r in res["runs"]] == [1, 2, 3]
    assert res["relmse_ratio_equal_time"] == pytest.approx(ratios[1], rel=1e-3)
    assert res["guided"]["seed"] in (1, 2, 3) and res["plain"]["spp"] > 0
    assert all(np.isfinite(r["relmse"]) and r["spp"] > 0 for r in res["runs"])
    # the curve's device-side relMSE is quality.rel_mse of the film
    curve, _ = equal_time.progressive(sc, 48, 48, 0.05, ref, max_depth=8, seed=7)
    spp = curve[-1][1]
    film = sc.render(48, 48, spp, seed=7, max_depth=8)
    assert curve[-1][2] == pytest.approx(quality.rel_mse(film / spp, ref), rel=1e-4)
    assert all(t1 > t0 for (t0, _, _), (t1, _, _) in zip(curve, curve[1:]))


@pytest.mark.gpu
def test_training_bursts_read_the_burst_before_last():
    """A training render's Adam steps (a burst) run on a stream of their own
    while the next render runs.  A training render samples with the network
    published by the burst before last, any other render with the last one.
    (Training itself is not bit-reproducible: samples are appended through
    atomics, so gradient sums run in varying order; films are compared only
    where the network they read is fixed.)"""
    print("[gpu] asynchronous training bursts, cornell 32x32", flush=True)
    dev = api.Device(0)
    desc = scenes.cornell(32, 32)
    tr = dict(max_depth=8, flags=api.XT_FLAG_GUIDE | api.XT_FLAG_GUIDE_TRAIN)
    a, b, c, d = (api.Scene(dev, desc) for _ in range(4))
    for sc in (a, b, c, d):
        sc.guide_init(seed=5)
    fa = [a.render(32, 32, 8, seed=60 + k, **tr) for k in range(4)]          # back to back
    fb = [b.render(32, 32, 8, seed=60, **tr)]
    p1 = b.guide_params()                                                    # waits for the first burst
    assert b.stats()["guide_steps"] > 0
    fb += [b.render(32, 32, 8, seed=60 + k, **tr) for k in (1, 2)]
    # the first two training renders read the initial network, whenever the host waits
    assert np.array_equal(fa[0], fb[0]) and np.array_equal(fa[1], fb[1])
    assert np.array_equal(c.render(32, 32, 8, seed=61, **tr), fa[1])
    # the third reads the first burst's network
    d.guide_set_params(p1)
    assert np.array_equal(d.render(32, 32, 8, seed=62, **tr), fb[2])
    assert not np.array_equal(fb[2], c.render(32, 32, 8, seed=62, **tr))
    # a render that does not train reads the last burst's network: the same
    # film as from the final parameters set again (which re-derive every copy)
    kw = dict(seed=70, max_depth=8, flags=api.XT_FLAG_GUIDE)
    g1 = a.render(32, 32, 8, **kw)
    a.guide_set_params(a.guide_params())
    assert np.array_equal(a.render(32, 32, 8, **kw), g1)
