"""FP64 oracle of the path-guiding network (numpy).

TEST INFRASTRUCTURE ONLY: loaded by tests/ (and bench.py's cpu_baseline leg)
as the checker of the HIP guiding kernels.

The network has no counterpart in the reference (it is one of the three
additions on top of the reference path tracer): a frequency encoding of the
position (12 bands, sin and cos, 3 axes = 72), the shading normal (3) and
the outgoing direction (3) -> 78 inputs, 4 hidden layers of width 64 with
ReLU, 40 outputs = 8 von Mises-Fisher lobes x (mean direction 3, log kappa
1, weight logit 1).  Training minimises the weighted negative
log-likelihood  sum_j w_j * -log q(wi_j | x_j), whose gradient is the
KL-divergence gradient of Mueller et al. 2019 (Neural Importance Sampling)
for samples drawn from a known pdf with w_j = F(wi_j) / p_sample(wi_j).
Parameters are updated with Adam (Kingma & Ba 2015).

Parameter layout (one flat float vector, 20,136 values):
  W1 [64 x 78], b1 [64], W2 [64 x 64], b2, W3, b3, W4, b4, W5 [40 x 64], b5 [40]
Weights are stored output-major (row n holds the weights of output n).
"""
from __future__ import annotations

import numpy as np

N_IN = 78
N_HID = 64
N_OUT = 40
N_LOBES = 8
BANDS = 12
LAYERS = [(N_IN, N_HID), (N_HID, N_HID), (N_HID, N_HID), (N_HID, N_HID), (N_HID, N_OUT)]
KAPPA_LO, KAPPA_HI = -5.0, 7.0  # log-kappa clamp


def param_count() -> int:
    return sum(i * o + o for i, o in LAYERS)


def unpack(theta: np.ndarray):
    out, off = [], 0
    for i, o in LAYERS:
        W = theta[off: off + i * o].reshape(o, i)
        off += i * o
        b = theta[off: off + o]
        off += o
        out.append((W, b))
    return out


def encode(pos01: np.ndarray, nrm: np.ndarray, wo: np.ndarray) -> np.ndarray:
    """pos01: positions normalised to the scene box [0,1]^3 (N,3); nrm, wo unit (N,3).
    Feature order: for axis a in (x,y,z), band b in 0..11: sin(2^b pi p_a), cos(2^b pi p_a);
    then n (3), then wo (3)."""
    p = np.asarray(pos01, np.float64)
    freq = (2.0 ** np.arange(BANDS)) * np.pi
    ang = p[:, :, None] * freq[None, None, :]  # N,3,12
    sc = np.stack([np.sin(ang), np.cos(ang)], axis=-1).reshape(p.shape[0], 3 * BANDS * 2)
    return np.concatenate([sc, nrm, wo], axis=1)


def bf16(a: np.ndarray) -> np.ndarray:
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


# Taylor coefficients of sin(pi a) (degree 11 .. 1) and cos(pi a) (12 .. 2),
# rounded to float32: pi^k / k! with alternating signs.
_SIN_PI = [np.float32(float.fromhex(h)) for h in
           ("-0x1.e3075p-8", "0x1.507834p-4", "-0x1.32d2ccp-1", "0x1.466bc6p+1", "-0x1.4abbcep+2", "0x1.921fb6p+1")]
_COS_PI = [np.float32(float.fromhex(h)) for h in
           ("0x1.f9d38ap-10", "-0x1.a6d1f2p-6", "0x1.e1f506p-3", "-0x1.55d3c8p+0", "0x1.03c1fp+2", "-0x1.3bd3ccp+2")]


def _sincospi_series(t: np.ndarray):
    """sin(pi t), cos(pi t) for float32 |t| <= 1, operation for operation as
    the device evaluates them (guide.hip sinCosPiSeries): reflection to
    a <= 1/2, Horner on a^2, every product and sum rounded to float32."""
    f = np.float32
    a0 = np.abs(t)
    refl = a0 > f(0.5)
    a = np.where(refl, f(1.0) - a0, a0).astype(np.float32)
    z = a * a
    ps = np.full_like(z, _SIN_PI[0])
    for k in _SIN_PI[1:]:
        ps = ps * z + k
    pc = np.full_like(z, _COS_PI[0])
    for k in _COS_PI[1:]:
        pc = pc * z + k
    sn = ps * a
    cs = pc * z + f(1.0)
    cs = np.where(refl, -cs, cs)
    sn = np.where(t < f(0.0), -sn, sn)
    return sn.astype(np.float32), cs.astype(np.float32)


def encode_f32(pos01, nrm, wo) -> np.ndarray:
    """encode() as the device forms it in float32: per axis, p reduced exactly
    to t = p - 2 rint(p / 2), (sin, cos)(pi t) from a float32 Taylor series,
    then band b + 1 from band b by the double-angle step
    (s, c) -> (2 s c, (c - s)(c + s)).  Pinned against encode() (float64
    sin/cos of pi 2^b p) to the recurrence's error growth (< 1e-3)."""
    f = np.float32
    p = np.asarray(pos01, np.float32)
    t = (p - f(2.0) * np.rint(f(0.5) * p)).astype(np.float32)
    sn, cs = _sincospi_series(t)
    bands = []
    for b in range(BANDS):
        if b > 0:
            sn, cs = (f(2.0) * sn) * cs, (cs - sn) * (cs + sn)
        bands.append(np.stack([sn, cs], axis=-1))
    sc = np.stack(bands, axis=2).astype(np.float64).reshape(p.shape[0], 3 * BANDS * 2)
    return np.concatenate([sc, np.asarray(nrm, np.float64), np.asarray(wo, np.float64)], axis=1)


def forward_bf16(theta, x):
    """Forward with the device's rounding points: inputs, weights and hidden
    activations rounded to bfloat16, products accumulated exactly (float64),
    biases in float32.  Returns the raw outputs (N, 40)."""
    h = bf16(x)
    layers = unpack(np.asarray(theta, np.float32))
    for li, (W, b) in enumerate(layers):
        z = h @ bf16(W).T + b.astype(np.float64)
        if li < len(layers) - 1:
            h = bf16(np.maximum(z, 0.0))
        else:
            h = z
    return h


def forward(theta, x):
    """Returns (raw outputs (N,40), activations list [x, a1..a4])."""
    acts = [x]
    h = x
    layers = unpack(theta)
    for li, (W, b) in enumerate(layers):
        z = h @ W.T + b
        if li < len(layers) - 1:
            h = np.maximum(z, 0.0)
            acts.append(h)
        else:
            h = z
    return h, acts


def decode(raw):
    """raw (N,40) -> mu (N,8,3), kappa (N,8), pi (N,8), plus pieces for gradients."""
    r = raw.reshape(-1, N_LOBES, 5)
    v = r[..., 0:3]
    vn = np.linalg.norm(v, axis=-1, keepdims=True)
    vn_safe = np.maximum(vn, 1e-12)
    mu = v / vn_safe
    lk = r[..., 3]
    lkc = np.clip(lk, KAPPA_LO, KAPPA_HI)
    kappa = np.exp(lkc)
    logits = r[..., 4]
    m = logits.max(-1, keepdims=True)
    e = np.exp(logits - m)
    pi = e / e.sum(-1, keepdims=True)
    return mu, kappa, pi, dict(v=v, vn=vn_safe, lk=lk, lkc=lkc)


def vmf_log_pdf(mu, kappa, w):
    """log of kappa / (2 pi (1 - exp(-2 kappa))) * exp(kappa (mu.w - 1)); w (N,3)."""
    c = np.einsum("nkd,nd->nk", mu, w)
    return np.log(kappa) - np.log(2 * np.pi) - np.log(-np.expm1(-2 * kappa)) + kappa * (c - 1.0)


def mixture_pdf(raw, w):
    mu, kappa, pi, _ = decode(raw)
    return (pi * np.exp(vmf_log_pdf(mu, kappa, w))).sum(-1)


def loss_and_grad_raw(raw, w, weight):
    """L = sum_j weight_j * -log q(w_j); returns (L, dL/draw (N,40))."""
    mu, kappa, pi, aux = decode(raw)
    lf = vmf_log_pdf(mu, kappa, w)  # N,8
    lq_parts = np.log(pi) + lf
    m = lq_parts.max(-1, keepdims=True)
    logq = (m + np.log(np.exp(lq_parts - m).sum(-1, keepdims=True)))[:, 0]
    L = float(-(weight * logq).sum())
    resp = np.exp(lq_parts - logq[:, None])  # responsibilities r_k
    W = weight[:, None]
    g = np.zeros((raw.shape[0], N_LOBES, 5))
    # logits
    g[..., 4] = -W * (resp - pi)
    # log kappa (through the clamp)
    with np.errstate(over="ignore"):  # expm1(2 kappa) = inf for large kappa: the term is 0
        dlogf_dk = 1.0 / kappa - 2.0 / np.expm1(2 * kappa) + (np.einsum("nkd,nd->nk", mu, w) - 1.0)
    inside = (aux["lk"] > KAPPA_LO) & (aux["lk"] < KAPPA_HI)
    g[..., 3] = -W * resp * dlogf_dk * kappa * inside
    # mean direction through the normalisation
    c = np.einsum("nkd,nd->nk", mu, w)
    tang = w[:, None, :] - c[..., None] * mu
    g[..., 0:3] = -(W * resp)[..., None] * kappa[..., None] * tang / aux["vn"]
    return L, g.reshape(-1, N_OUT)


def backward(theta, x, w, weight):
    """Returns (loss, gradient wrt theta) for a batch (analytic backprop)."""
    raw, acts = forward(theta, x)
    L, d = loss_and_grad_raw(raw, w, weight)
    layers = unpack(theta)
    grads = []
    for li in range(len(layers) - 1, -1, -1):
        W, b = layers[li]
        a_prev = acts[li]
        gW = d.T @ a_prev
        gb = d.sum(0)
        grads.append((gW, gb))
        if li > 0:
            d = (d @ W) * (acts[li] > 0)
    grads.reverse()
    flat = np.concatenate([np.concatenate([gW.reshape(-1), gb]) for gW, gb in grads])
    return L, flat


def _onb(n):
    """Orthonormal basis around unit vectors n (N,3): the branchless
    construction of Duff et al. 2017 (the device's onb())."""
    s = np.where(n[:, 2] >= 0.0, 1.0, -1.0)
    a = -1.0 / (s + n[:, 2])
    c = n[:, 0] * n[:, 1] * a
    t = np.stack([1.0 + s * n[:, 0] ** 2 * a, s * c, -s * n[:, 0]], axis=1)
    b = np.stack([c, s + n[:, 1] ** 2 * a, -n[:, 1]], axis=1)
    return t, b


def vmf_sample(mu, kappa, u1, u2):
    """vMF directions by Jakob 2012 ("Numerically stable sampling of the von
    Mises-Fisher distribution on S^2"): cos = 1 + log1p((1-u1) expm1(-2k)) / k."""
    wz = 1.0 + np.log1p((1.0 - u1) * np.expm1(-2.0 * kappa)) / kappa
    wz = np.clip(wz, -1.0, 1.0)
    sz = np.sqrt(np.maximum(1.0 - wz * wz, 0.0))
    phi = 2.0 * np.pi * u2
    t, b = _onb(mu)
    d = t * (sz * np.cos(phi))[:, None] + b * (sz * np.sin(phi))[:, None] + mu * wz[:, None]
    return d / np.linalg.norm(d, axis=1, keepdims=True)


def mixture_sample(raw, u):
    """Lobe k = first index with u0 < cumulative weight (last lobe otherwise),
    then vmf_sample with (u1, u2).  Returns (directions (N,3), lobe (N,))."""
    mu, kappa, pi, _ = decode(raw)
    cum = np.cumsum(pi, axis=1)
    k = np.minimum((u[:, :1] >= cum[:, :-1]).sum(axis=1), N_LOBES - 1)
    rows = np.arange(raw.shape[0])
    return vmf_sample(mu[rows, k], kappa[rows, k], u[:, 1], u[:, 2]), k


def backward_bf16(theta, x, w, weight):
    """backward() with the device's rounding points: the forward of
    forward_bf16 (bf16 inputs, weights and hidden activations), the raw-output
    gradient of the unnormalised sum_j weight_j (-log q_j) rounded to bf16,
    and every back-propagated delta rounded to bf16 after the ReLU mask;
    products and sums exact, the summed gradient divided by the batch's total
    weight at the end (guide.hip kGuideTrainRows / kGuideReduceStep).
    Returns (loss, grad) of the loss normalised by the total weight."""
    layers = unpack(np.asarray(theta, np.float32))
    acts = [bf16(x)]
    h = acts[0]
    for li, (W, b) in enumerate(layers):
        z = h @ bf16(W).T + b.astype(np.float64)
        if li < len(layers) - 1:
            h = bf16(np.maximum(z, 0.0))
            acts.append(h)
        else:
            h = z
    raw = h.astype(np.float32).astype(np.float64)
    wsum = weight.sum()
    L, d = loss_and_grad_raw(raw, w, weight)
    d = bf16(d)
    grads = []
    for li in range(len(layers) - 1, -1, -1):
        W, _ = layers[li]
        grads.append((d.T @ acts[li], d.sum(0)))
        if li > 0:
            d = bf16((d @ bf16(W)) * (acts[li] > 0))
    grads.reverse()
    return L / wsum, np.concatenate([np.concatenate([gW.reshape(-1), gb]) for gW, gb in grads]) / wsum


def adam_step(theta, m, v, g, t, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (t = 1-based step); returns new (theta, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return theta - lr * mh / (np.sqrt(vh) + eps), m, v


def init_params(seed: int = 1) -> np.ndarray:
    """He-uniform weights, zero biases (seeded numpy generator)."""
    rng = np.random.default_rng(seed)
    parts = []
    for i, o in LAYERS:
        lim = np.sqrt(6.0 / i)
        parts.append(rng.uniform(-lim, lim, size=o * i))
        parts.append(np.zeros(o))
    return np.concatenate(parts)
