"""numpy reference of the guiding network (xitils_amd/csrc/xt_guide.h,
guide.hip): frequency encoding, 3x64 ReLU MLP, 8-lobe vMF mixture head,
negative-log-likelihood gradient and Adam.

With emulate=True the forward/backward round to bf16 at the same points as
the MFMA kernels (inputs, weights, hidden activations, output and hidden
gradients) and accumulate in fp32; with emulate=False everything runs in
float64 (used against finite differences).
"""
import numpy as np

G_IN, G_HID, G_OUT, G_OUT_PAD, G_LOBES, G_BANDS = 96, 64, 40, 48, 8, 12
OFF_W0 = 0
OFF_B0 = OFF_W0 + G_HID * G_IN
OFF_W1 = OFF_B0 + G_HID
OFF_B1 = OFF_W1 + G_HID * G_HID
OFF_W2 = OFF_B1 + G_HID
OFF_B2 = OFF_W2 + G_HID * G_HID
OFF_W3 = OFF_B2 + G_HID
OFF_B3 = OFF_W3 + G_OUT_PAD * G_HID
NPARAMS = OFF_B3 + G_OUT_PAD
KLOG_MIN, KLOG_MAX = -3.0, 9.0


def bf16(x):
    """Round float32 values to the nearest bf16 (ties to even), as float32."""
    a = np.ascontiguousarray(x, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).reshape(a.shape)


def unpack(P):
    P = np.asarray(P)
    return (P[OFF_W0:OFF_B0].reshape(G_HID, G_IN), P[OFF_B0:OFF_W1], P[OFF_W1:OFF_B1].reshape(G_HID, G_HID),
            P[OFF_B1:OFF_W2], P[OFF_W2:OFF_B2].reshape(G_HID, G_HID), P[OFF_B2:OFF_W3],
            P[OFF_W3:OFF_B3].reshape(G_OUT_PAD, G_HID), P[OFF_B3:NPARAMS])


def encode(q, lo, hi, dtype=np.float32):
    """96 features per query (p, n, wo as 9 floats): 12 frequency bands of the
    position scaled to [0, 1] (sin, cos by the doubling recurrence), n, wo."""
    q = np.asarray(q, dtype).reshape(-1, 9)
    lo = np.asarray(lo, dtype)
    inv = (dtype(1.0) / (np.asarray(hi, dtype) - lo)).astype(dtype)
    x = np.zeros((q.shape[0], G_IN), dtype)
    u = np.clip((q[:, 0:3] - lo) * inv, 0, 1).astype(dtype)
    pi = dtype(np.pi)
    for a in range(3):
        s = np.sin(pi * u[:, a]).astype(dtype)
        c = np.cos(pi * u[:, a]).astype(dtype)
        for k in range(G_BANDS):
            x[:, a * 2 * G_BANDS + 2 * k] = s
            x[:, a * 2 * G_BANDS + 2 * k + 1] = c
            # sin 2x = (s + s) c, cos 2x = 1 - 2 s^2 with one rounding (an fma)
            s, c = ((s + s) * c).astype(dtype), (1.0 - 2.0 * s.astype(np.float64) * s).astype(dtype)
    x[:, 72:75] = q[:, 3:6]
    x[:, 75:78] = q[:, 6:9]
    return x


def forward(P, X, emulate=True):
    """Returns (Y [n, 48], activations) with activations = (X, H1, H2, H3)."""
    W0, b0, W1, b1, W2, b2, W3, b3 = unpack(P)
    if emulate:
        rd = bf16
        acc = np.float32
        W0, W1, W2, W3 = (bf16(w) for w in (W0, W1, W2, W3))
        X = bf16(X)
    else:
        rd = lambda v: v  # noqa: E731
        acc = np.float64
        X = X.astype(np.float64)
    H1 = rd(np.maximum(X.astype(acc) @ W0.T.astype(acc) + b0, 0).astype(acc))
    H2 = rd(np.maximum(H1.astype(acc) @ W1.T.astype(acc) + b1, 0).astype(acc))
    H3 = rd(np.maximum(H2.astype(acc) @ W2.T.astype(acc) + b2, 0).astype(acc))
    Y = (H3.astype(acc) @ W3.T.astype(acc) + b3).astype(acc)
    return Y, (X, H1, H2, H3)


def decode(Y):
    """Raw outputs [n, >=40] -> mu [n, 8, 3], kappa [n, 8], alpha [n, 8]."""
    Y = np.asarray(Y)[:, :G_OUT].reshape(-1, G_LOBES, 5)
    m = Y[:, :, 0:3]
    ln = np.linalg.norm(m, axis=2, keepdims=True)
    mu = np.where(ln > 1e-12, m / np.maximum(ln, 1e-300), np.array([0, 0, 1.0]))
    kappa = np.exp(np.clip(Y[:, :, 3], KLOG_MIN, KLOG_MAX))
    a = Y[:, :, 4] - Y[:, :, 4].max(axis=1, keepdims=True)
    alpha = np.exp(a) / np.exp(a).sum(axis=1, keepdims=True)
    return mu, kappa, alpha


def vmf_pdf(mu, kappa, w):
    """vMF(w; mu, kappa) = kappa / (2 pi (1 - e^{-2 kappa})) e^{kappa (mu.w - 1)}."""
    return kappa / (2 * np.pi * (1 - np.exp(-2 * kappa))) * np.exp(kappa * ((mu * w).sum(-1) - 1))


def mixture_pdf(Y, w):
    mu, kappa, alpha = decode(Y)
    return (alpha * vmf_pdf(mu, kappa, np.asarray(w)[:, None, :])).sum(1)


def nll_grad_y(Y, w, weight):
    """d(-weight log q(w)) / dY [n, 40] (float64)."""
    Y = np.asarray(Y, np.float64)[:, :G_OUT]
    n = Y.shape[0]
    w = np.asarray(w, np.float64).reshape(n, 1, 3)
    weight = np.asarray(weight, np.float64).reshape(n, 1)
    mu, kappa, alpha = decode(Y)
    v = vmf_pdf(mu, kappa, w)
    q = np.maximum((alpha * v).sum(1, keepdims=True), 1e-30)
    r = alpha * v / q
    m = Y.reshape(n, G_LOBES, 5)[:, :, 0:3]
    ln = np.maximum(np.linalg.norm(m, axis=2, keepdims=True), 1e-12)
    mw = (mu * w).sum(-1, keepdims=True)
    gm = (w - mw * mu) * (r * kappa)[:, :, None] / ln
    dk = 1 / kappa - 2 / np.expm1(2 * kappa) + (mw[:, :, 0] - 1)
    kl = Y.reshape(n, G_LOBES, 5)[:, :, 3]
    gk = np.where((kl > KLOG_MIN) & (kl < KLOG_MAX), r * dk * kappa, 0.0)
    ga = r - alpha
    g = np.concatenate([gm, gk[:, :, None], ga[:, :, None]], axis=2).reshape(n, G_OUT)
    return -weight * g


def loss(P, q, dirs, weights, lo, hi):
    """sum_i -w_i log q(dir_i | query_i), float64, no rounding."""
    X = encode(q, lo, hi, np.float64)
    Y, _ = forward(np.asarray(P, np.float64), X, emulate=False)
    return float(-(np.asarray(weights, np.float64) * np.log(mixture_pdf(Y, dirs))).sum())


def backward(P, q, dirs, weights, lo, hi, emulate=True, weight_scale=1.0, weight_clip=np.inf):
    """Flat gradient of sum_i -min(w_i * scale, clip) log q(dir_i | query_i)."""
    dt = np.float32 if emulate else np.float64
    rd = bf16 if emulate else (lambda v: v)
    X = encode(q, lo, hi, np.float32 if emulate else np.float64)
    Y, (X, H1, H2, H3) = forward(np.asarray(P, dt), X, emulate)
    W0, b0, W1, b1, W2, b2, W3, b3 = (np.asarray(a, dt) for a in unpack(P))
    if emulate:
        W0, W1, W2, W3 = (bf16(w) for w in (W0, W1, W2, W3))
    w = np.minimum(np.asarray(weights, np.float64) * weight_scale, weight_clip)
    dY = np.zeros((Y.shape[0], G_OUT_PAD), dt)
    dY[:, :G_OUT] = rd(nll_grad_y(Y, dirs, w).astype(dt))
    g = np.zeros(NPARAMS, np.float64)
    g[OFF_W3:OFF_B3] = (dY.T.astype(np.float64) @ H3.astype(np.float64)).ravel()
    g[OFF_B3:NPARAMS] = dY.sum(0)
    dZ3 = rd(((dY @ W3) * (H3 > 0)).astype(dt))
    g[OFF_W2:OFF_B2] = (dZ3.T.astype(np.float64) @ H2.astype(np.float64)).ravel()
    g[OFF_B2:OFF_W3] = dZ3.sum(0)
    dZ2 = rd(((dZ3 @ W2) * (H2 > 0)).astype(dt))
    g[OFF_W1:OFF_B1] = (dZ2.T.astype(np.float64) @ H1.astype(np.float64)).ravel()
    g[OFF_B1:OFF_W2] = dZ2.sum(0)
    dZ1 = rd(((dZ2 @ W1) * (H1 > 0)).astype(dt))
    g[OFF_W0:OFF_B0] = (dZ1.T.astype(np.float64) @ X.astype(np.float64)).ravel()
    g[OFF_B0:OFF_W1] = dZ1.sum(0)
    return g


def adam(P, m, v, g, t, lr=1e-2, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (Kingma & Ba) at step count t (1-based)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return P - lr * mh / (np.sqrt(vh) + eps), m, v


def init_params(seed=0, out_scale=0.1):
    """He-uniform hidden layers, small output layer (the shape of the
    library's initialisation, not its exact values)."""
    rng = np.random.default_rng(seed)
    P = np.zeros(NPARAMS, np.float32)
    P[OFF_W0:OFF_B0] = rng.uniform(-1, 1, G_HID * G_IN) * np.sqrt(6 / G_IN)
    P[OFF_W1:OFF_B1] = rng.uniform(-1, 1, G_HID * G_HID) * np.sqrt(6 / G_HID)
    P[OFF_W2:OFF_B2] = rng.uniform(-1, 1, G_HID * G_HID) * np.sqrt(6 / G_HID)
    P[OFF_W3:OFF_W3 + G_OUT * G_HID] = rng.uniform(-1, 1, G_OUT * G_HID) * out_scale * np.sqrt(6 / G_HID)
    P[OFF_B0:OFF_W1] = rng.uniform(-0.1, 0.1, G_HID)
    P[OFF_B3:OFF_B3 + G_OUT] = rng.uniform(-0.5, 0.5, G_OUT)
    return P


def random_queries(n, lo, hi, seed=0):
    rng = np.random.default_rng(seed)
    p = rng.uniform(lo, hi, (n, 3))
    nn = rng.normal(size=(n, 3))
    nn /= np.linalg.norm(nn, axis=1, keepdims=True)
    wo = rng.normal(size=(n, 3))
    wo /= np.linalg.norm(wo, axis=1, keepdims=True)
    return np.concatenate([p, nn, wo], 1).astype(np.float32)


def random_dirs(n, seed=1):
    rng = np.random.default_rng(seed)
    d = rng.normal(size=(n, 3))
    return (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
