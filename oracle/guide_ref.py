"""CPU restatement of the path-guiding network — TEST INFRASTRUCTURE.

Only tests/ import this module, as the checker of the HIP guiding kernels
(jarabotransientrendering_amd/csrc/tr_guide.h, tr_guide.hip). The network has no counterpart in
the reference renderer; what pins this restatement is mathematics (tests/test_guide_oracle.py):
the vMF normalisation integrates to one, the analytic backward pass agrees with central finite
differences of the float64 loss, and Adam matches the published update (Kingma & Ba 2015, Alg. 1).

Network: features(48) -> 64 -> 64 -> 64 -> 64 -> 40, ReLU; outputs per lobe i (8 lobes):
direction r = o[5i:5i+3] (mu = r/|r|), concentration kappa = min(softplus(o[5i+3]) + 0.01, 2000),
weight logit o[5i+4] (softmax over lobes). Parameters: per layer W[in][out] (in-major) then b[out].
"""
from __future__ import annotations

import numpy as np

LAYERS = [(48, 64), (64, 64), (64, 64), (64, 64), (64, 40)]
N_LOBES = 8
KAPPA_MIN, KAPPA_MAX = 1e-2, 2e3
N_PARAMS = sum(i * o + o for i, o in LAYERS)  # 18,216
N_WEIGHTS = sum(i * o for i, o in LAYERS)     # 17,920


def unpack(params):
    """[(W (in,out), b (out,)), ...] views of a flat parameter vector."""
    out, k = [], 0
    p = np.asarray(params)
    for i, o in LAYERS:
        W = p[k:k + i * o].reshape(i, o)
        k += i * o
        b = p[k:k + o]
        k += o
        out.append((W, b))
    return out


def bf16(x):
    """Round to the nearest bfloat16 (ties to even), returned as float32."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def encode(raw):
    """Raw vertex inputs (n, 12): pos xyz (scene-normalised), wo xyz, normal xyz, t, medium, 0
    -> features (n, 48): pos, sin/cos(2^f pi pos) f=0..3, wo, n, t, sin/cos(2^f pi t) f=0..4,
    medium, 3 zeros."""
    raw = np.asarray(raw, dtype=np.float64)
    n = raw.shape[0]
    f = np.zeros((n, 48))
    k = 0
    f[:, 0:3] = raw[:, 0:3]
    k = 3
    for c in range(3):
        for q in range(4):
            a = np.pi * (2.0 ** q) * raw[:, c]
            f[:, k] = np.sin(a)
            f[:, k + 1] = np.cos(a)
            k += 2
    f[:, k:k + 6] = raw[:, 3:9]
    k += 6
    f[:, k] = raw[:, 9]
    k += 1
    for q in range(5):
        a = np.pi * (2.0 ** q) * raw[:, 9]
        f[:, k] = np.sin(a)
        f[:, k + 1] = np.cos(a)
        k += 2
    f[:, k] = raw[:, 10]
    return f


def forward(params, raw, emulate_bf16=True):
    """Network outputs (n, 40). With emulate_bf16 the features, weights and hidden activations are
    rounded to bfloat16 as on the matrix cores (accumulation stays in float64 here)."""
    x = encode(raw)
    if emulate_bf16:
        x = bf16(x).astype(np.float64)
    layers = unpack(params)
    for li, (W, b) in enumerate(layers):
        Wd = bf16(W).astype(np.float64) if emulate_bf16 else np.asarray(W, np.float64)
        x = x @ Wd + np.asarray(b, np.float64)
        if li < len(layers) - 1:
            x = np.maximum(x, 0.0)
            if emulate_bf16:
                x = bf16(x).astype(np.float64)
    return x


def softplus(x):
    x = np.asarray(x, dtype=np.float64)
    return np.where(x > 15.0, x, np.log1p(np.exp(np.minimum(x, 15.0))))


def mixture(o):
    """(mu (n,8,3), kappa (n,8), weights (n,8)) from raw outputs (n,40)."""
    o = np.asarray(o, dtype=np.float64).reshape(-1, N_LOBES, 5)
    r = o[:, :, 0:3]
    rn = np.linalg.norm(r, axis=2, keepdims=True)
    mu = np.where(rn > 1e-20, r / np.maximum(rn, 1e-300), np.array([0.0, 0.0, 1.0]))
    kappa = np.minimum(softplus(o[:, :, 3]) + KAPPA_MIN, KAPPA_MAX)
    lg = o[:, :, 4] - o[:, :, 4].max(axis=1, keepdims=True)
    w = np.exp(lg)
    w /= w.sum(axis=1, keepdims=True)
    return mu, kappa, w


def vmf_log_pdf(mu, kappa, omega):
    """log of kappa / (2 pi (1 - e^{-2 kappa})) * exp(kappa (mu.omega - 1))."""
    c = np.einsum("...k,...k->...", mu, omega)
    return np.log(kappa) - np.log(2 * np.pi) - np.log(-np.expm1(-2 * kappa)) + kappa * (c - 1)


def mixture_log_pdf(o, omega):
    mu, kappa, w = mixture(o)
    om = np.asarray(omega, dtype=np.float64)[:, None, :]
    lt = np.log(w) + vmf_log_pdf(mu, kappa, om)
    m = lt.max(axis=1, keepdims=True)
    return (m + np.log(np.exp(lt - m).sum(axis=1, keepdims=True)))[:, 0]


def loss(params, records, emulate_bf16=False):
    """Mean of -w_norm log q(omega) over records (n,16); w_norm = w / mean(w)."""
    rec = np.asarray(records, dtype=np.float64).reshape(-1, 16)
    raw = record_raw(rec)
    w = rec[:, 11]
    wn = w / (w.sum() / len(w))
    o = forward(params, raw, emulate_bf16)
    lq = mixture_log_pdf(o, rec[:, 12:15])
    return float(np.mean(np.where(wn > 0, -wn * lq, 0.0)))


def record_raw(rec):
    rec = np.asarray(rec, dtype=np.float64).reshape(-1, 16)
    raw = np.zeros((len(rec), 12))
    raw[:, 0:3] = rec[:, 0:3]
    raw[:, 3:6] = rec[:, 4:7]
    raw[:, 6:9] = rec[:, 8:11]
    raw[:, 9] = rec[:, 3]
    raw[:, 10] = rec[:, 7]
    return raw


def output_grad(o, omega, w):
    """d(-w log q)/d(raw outputs), float64 (n,40); the analytic form the kernel implements."""
    o = np.asarray(o, dtype=np.float64).reshape(-1, N_LOBES, 5)
    n = o.shape[0]
    om = np.asarray(omega, dtype=np.float64)
    w = np.asarray(w, dtype=np.float64)
    r = o[:, :, 0:3]
    rn = np.linalg.norm(r, axis=2)
    mu = np.where(rn[..., None] > 1e-20, r / np.maximum(rn[..., None], 1e-300), np.array([0.0, 0.0, 1.0]))
    raw_k = o[:, :, 3]
    kappa = np.minimum(softplus(raw_k) + KAPPA_MIN, KAPPA_MAX)
    lg = o[:, :, 4] - o[:, :, 4].max(axis=1, keepdims=True)
    pi = np.exp(lg) / np.exp(lg).sum(axis=1, keepdims=True)
    c = np.einsum("nik,nk->ni", mu, om)
    lt = np.log(pi) + np.log(kappa) - np.log(2 * np.pi) - np.log(-np.expm1(-2 * kappa)) + kappa * (c - 1)
    m = lt.max(axis=1, keepdims=True)
    log_q = m + np.log(np.exp(lt - m).sum(axis=1, keepdims=True))
    gamma = np.exp(lt - log_q)
    d = np.zeros((n, N_LOBES, 5))
    dlogf = -w[:, None] * gamma
    d[:, :, 4] = -w[:, None] * (gamma - pi)
    gk = np.where(kappa < 1e-3, 1 - kappa / 3, 1 / kappa - 2 / np.expm1(2 * kappa))
    sig = 1 / (1 + np.exp(-raw_k))
    clamped = softplus(raw_k) + KAPPA_MIN >= KAPPA_MAX
    d[:, :, 3] = np.where(clamped, 0.0, dlogf * (gk + c - 1) * sig)
    s = np.where(rn > 1e-20, dlogf * kappa / np.maximum(rn, 1e-300), 0.0)
    d[:, :, 0:3] = s[..., None] * (om[:, None, :] - c[..., None] * mu)
    return d.reshape(n, 40), (-w * log_q[:, 0])


def gradient(params, records, emulate_bf16=False):
    """Mean loss gradient over records in float64 — the reference for the kernel's BF16
    forward/backward. With emulate_bf16, features, weights, hidden activations and the
    back-propagated output gradients are rounded to bfloat16 where the kernel rounds them."""
    rd = (lambda a: bf16(a).astype(np.float64)) if emulate_bf16 else (lambda a: a)
    rec = np.asarray(records, dtype=np.float64).reshape(-1, 16)
    raw = record_raw(rec)
    w = rec[:, 11]
    wn = np.where(w > 0, w / (w.sum() / len(w)), 0.0)
    x = rd(encode(raw))
    layers = unpack(np.asarray(params, np.float64))
    Ws = [rd(W) for W, _ in layers]
    acts = [x]
    pre = []
    for li, (W, b) in enumerate(layers):
        z = acts[-1] @ Ws[li] + b
        pre.append(z)
        acts.append(rd(np.maximum(z, 0.0)) if li < len(layers) - 1 else z)
    dY, _ = output_grad(acts[-1], rec[:, 12:15], wn)
    dY = rd(dY)
    grads = [None] * len(layers)
    for li in range(len(layers) - 1, -1, -1):
        X = acts[li]
        gW = X.T @ dY
        gb = dY.sum(axis=0)
        grads[li] = (gW, gb)
        if li > 0:
            dX = dY @ Ws[li].T
            dY = rd(dX * (pre[li - 1] > 0))
    flat = np.concatenate([np.concatenate([gW.ravel(), gb]) for gW, gb in grads])
    return flat / len(rec)


def adam(params, m, v, g, t, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (Kingma & Ba 2015, Algorithm 1); returns (params, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return params - lr * mh / (np.sqrt(vh) + eps), m, v


def vmf_sample(mu, kappa, u1, u2):
    """Directions from one lobe (W. Jakob 2012 inversion), for distribution checks."""
    ct = 1 + np.log(u1 + (1 - u1) * np.exp(-2 * kappa)) / kappa
    ct = np.clip(ct, -1, 1)
    st = np.sqrt(np.maximum(0, 1 - ct * ct))
    phi = 2 * np.pi * u2
    sign = np.copysign(1.0, mu[2])
    a = -1 / (sign + mu[2])
    b = mu[0] * mu[1] * a
    t = np.array([1 + sign * mu[0] ** 2 * a, sign * b, -sign * mu[0]])
    bb = np.array([b, sign + mu[1] ** 2 * a, -mu[1]])
    return (np.outer(st * np.cos(phi), t) + np.outer(st * np.sin(phi), bb) + np.outer(ct, mu))


def initial_like(seed=0, out_scale=0.1):
    """Random parameters of the same distribution family as the library's initialisation
    (Glorot-uniform, output layer scaled); for tests only."""
    rng = np.random.default_rng(seed)
    p = []
    for li, (i, o) in enumerate(LAYERS):
        bnd = np.sqrt(6.0 / (i + o)) * (out_scale if li == len(LAYERS) - 1 else 1.0)
        p.append(rng.uniform(-bnd, bnd, i * o))
        p.append(rng.normal(0, 0.05, o))
    return np.concatenate(p).astype(np.float32)
