"""Pins the FP64 guiding-network oracle (oracle/guide.py) against references
that do not share its code: torch's Linear/ReLU stack and autograd for the
network and the loss gradient, the closed-form vMF marginal in cos(theta),
quadrature of the pdf, Adam's first step, and special values of the encoding.
"""
import numpy as np
import pytest
import torch

from oracle import guide as G


def _records(n, seed):
    rng = np.random.default_rng(seed)
    pos = rng.random((n, 3))
    nrm = rng.normal(size=(n, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    wo = rng.normal(size=(n, 3))
    wo /= np.linalg.norm(wo, axis=1, keepdims=True)
    wi = rng.normal(size=(n, 3))
    wi /= np.linalg.norm(wi, axis=1, keepdims=True)
    w = rng.exponential(size=n)
    return pos, nrm, wo, wi, w


def _torch_mlp(theta):
    layers = []
    for li, (W, b) in enumerate(G.unpack(theta)):
        lin = torch.nn.Linear(W.shape[1], W.shape[0]).double()
        with torch.no_grad():
            lin.weight.copy_(torch.from_numpy(np.array(W)))
            lin.bias.copy_(torch.from_numpy(np.array(b)))
        layers.append(lin)
        if li < 4:
            layers.append(torch.nn.ReLU())
    return torch.nn.Sequential(*layers)


def _torch_loss(raw, wi, w):
    """sum_j w_j * -log sum_k pi_k vMF(wi_j; mu_k, kappa_k), written with torch ops."""
    r = raw.reshape(-1, 8, 5)
    mu = torch.nn.functional.normalize(r[..., :3], dim=-1)
    kappa = torch.exp(torch.clamp(r[..., 3], G.KAPPA_LO, G.KAPPA_HI))
    logpi = torch.log_softmax(r[..., 4], dim=-1)
    cos = (mu * wi[:, None, :]).sum(-1)
    logf = torch.log(kappa / (2 * np.pi * (1 - torch.exp(-2 * kappa)))) + kappa * (cos - 1)
    return -(w * torch.logsumexp(logpi + logf, dim=-1)).sum()


def test_param_layout():
    assert G.param_count() == 20136
    layers = G.unpack(np.arange(G.param_count(), dtype=np.float64))
    assert [W.shape for W, _ in layers] == [(64, 78), (64, 64), (64, 64), (64, 64), (40, 64)]
    assert layers[0][1][0] == 78 * 64  # first bias follows W1


def test_encoding_special_values():
    pos = np.array([[0.0, 0.5, 0.25]])
    e = G.encode(pos, [[0, 0, 1]], [[1, 0, 0]])[0]
    sx, cx = e[0:24:2], e[1:24:2]
    np.testing.assert_allclose(sx, 0.0, atol=1e-12)
    np.testing.assert_allclose(cx, 1.0, atol=1e-12)
    sy, cy = e[24:48:2], e[25:48:2]
    np.testing.assert_allclose([sy[0], cy[0], sy[1], cy[1]], [1.0, 0.0, 0.0, -1.0], atol=1e-12)
    np.testing.assert_allclose(cy[2:], 1.0, atol=1e-9)
    sz = e[48:72:2]
    np.testing.assert_allclose(sz[:2], [np.sin(np.pi / 4), 1.0], atol=1e-12)
    np.testing.assert_array_equal(e[72:], [0, 0, 1, 1, 0, 0])
    # float32 device form: the double-angle steps double the error per band,
    # to below 1e-3 at band 11 (under the bf16 rounding of the features)
    pos, nrm, wo, _, _ = _records(256, 3)
    assert np.abs(G.encode(pos, nrm, wo) - G.encode_f32(pos, nrm, wo)).max() < 1e-3


def test_sincospi_series():
    """The device's series for sin / cos(pi t) against float64 on a dense grid
    of [-1, 1] (a wrong coefficient or reflection shows up as > 1e-5), and the
    exact values at the quarter points."""
    t = np.linspace(-1.0, 1.0, 400001).astype(np.float32)
    sn, cs = G._sincospi_series(t)
    ref = np.pi * t.astype(np.float64)
    assert np.abs(sn - np.sin(ref)).max() < 3e-7
    assert np.abs(cs - np.cos(ref)).max() < 3e-7
    q = np.array([-1.0, -0.5, 0.0, 0.5, 1.0], np.float32)
    sq, cq = G._sincospi_series(q)
    np.testing.assert_allclose(sq, [0, -1, 0, 1, 0], atol=1e-7)
    np.testing.assert_allclose(cq, [-1, 0, 1, 0, -1], atol=1e-7)


def test_bf16_rounding():
    x = np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, 1 + 2 ** -7 + 2 ** -9, -2.5, 0.0], np.float32)
    np.testing.assert_array_equal(G.bf16(x), [1.0, 1.0, 1 + 2 ** -6, 1 + 2 ** -7, -2.5, 0.0])


def test_forward_matches_torch():
    theta = G.init_params(3)
    pos, nrm, wo, _, _ = _records(500, 4)
    x = G.encode(pos, nrm, wo)
    raw, acts = G.forward(theta, x)
    ref = _torch_mlp(theta)(torch.from_numpy(x)).detach().numpy()
    np.testing.assert_allclose(raw, ref, rtol=1e-12, atol=1e-12)
    assert len(acts) == 5 and acts[1].shape == (500, 64)
    # the bf16 variant stays within bf16 rounding of the exact network
    rb = G.forward_bf16(theta, x)
    assert np.abs(rb - raw).max() < 0.05 * np.abs(raw).std() + 1e-3


def test_gradient_matches_torch_autograd():
    theta = G.init_params(5)
    theta[G.param_count() - 40:] = np.random.default_rng(1).normal(size=40)  # nonzero output biases
    pos, nrm, wo, wi, w = _records(300, 6)
    x = G.encode(pos, nrm, wo)
    L, g = G.backward(theta, x, wi, w)
    mlp = _torch_mlp(theta)
    loss = _torch_loss(mlp(torch.from_numpy(x)), torch.from_numpy(wi), torch.from_numpy(w))
    loss.backward()
    ref = np.concatenate([np.concatenate([lin.weight.grad.numpy().ravel(), lin.bias.grad.numpy()])
                          for lin in mlp if isinstance(lin, torch.nn.Linear)])
    assert L == pytest.approx(loss.item(), rel=1e-12)
    np.testing.assert_allclose(g, ref, rtol=1e-9, atol=1e-12 * np.abs(ref).max())


def test_loss_grad_raw_matches_finite_differences():
    rng = np.random.default_rng(7)
    raw = rng.normal(size=(20, 40))
    raw[:, 3::5] = rng.uniform(-6, 8, size=(20, 8))  # some log-kappas in the clamp
    _, _, _, wi, w = _records(20, 8)
    _, g = G.loss_and_grad_raw(raw, wi, w)
    h = 1e-6
    for j in range(40):
        d = np.zeros_like(raw)
        d[:, j] = h
        num = np.array([(G.loss_and_grad_raw(raw[i:i + 1] + d[i:i + 1], wi[i:i + 1], w[i:i + 1])[0] -
                         G.loss_and_grad_raw(raw[i:i + 1] - d[i:i + 1], wi[i:i + 1], w[i:i + 1])[0]) / (2 * h)
                        for i in range(20)])
        np.testing.assert_allclose(g[:, j], num, rtol=1e-5, atol=1e-6)


@pytest.mark.parametrize("log_kappa", [-5.0, -1.0, 0.0, 2.0, 5.0, 7.0])
def test_vmf_pdf_normalised(log_kappa):
    """2 pi integral over cos in [-1, 1] of the pdf = 1 (substitution
    s = 1 - cos, dense near the mode)."""
    kappa = np.exp(log_kappa)
    s = np.concatenate([np.linspace(0, 20 / kappa, 200001)[:-1], np.linspace(min(20 / kappa, 2.0), 2.0, 200001)])
    s = np.unique(np.clip(s, 0, 2))
    mu = np.tile([[0.0, 0.0, 1.0]], (len(s), 1))[:, None, :]
    wdir = np.stack([np.sqrt(np.maximum(0, 1 - (1 - s) ** 2)), 0 * s, 1 - s], axis=1)
    f = np.exp(G.vmf_log_pdf(mu, np.full((len(s), 1), kappa), wdir))[:, 0]
    assert 2 * np.pi * np.trapezoid(f, s) == pytest.approx(1.0, abs=2e-4)


@pytest.mark.parametrize("log_kappa", [-5.0, 0.0, 3.0, 7.0])
def test_vmf_sampler_matches_closed_form_cdf(log_kappa):
    """cos(theta) to the mean has CDF (e^{k(t-1)} - e^{-2k}) / (1 - e^{-2k});
    the azimuth is uniform."""
    kappa = np.exp(log_kappa)
    n = 200_000
    rng = np.random.default_rng(11)
    mu = rng.normal(size=3)
    mu /= np.linalg.norm(mu)
    u = rng.random((n, 2))
    d = G.vmf_sample(np.tile(mu, (n, 1)), np.full(n, kappa), u[:, 0], u[:, 1])
    np.testing.assert_allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-12)
    t = np.sort(d @ mu)
    cdf = (np.exp(kappa * (t - 1)) - np.exp(-2 * kappa)) / -np.expm1(-2 * kappa)
    ks = np.abs(cdf - np.arange(1, n + 1) / n).max()
    assert ks < 1.63 / np.sqrt(n) * 1.5  # ~1% level, padded
    # azimuth around mu uniform
    tb, bb = G._onb(mu[None, :])
    phi = np.arctan2(d @ bb[0], d @ tb[0])
    hist, _ = np.histogram(phi, bins=16, range=(-np.pi, np.pi))
    expected = n / 16
    assert ((hist - expected) ** 2 / expected).sum() < 50  # chi2, 15 dof


def test_mixture_sampler_matches_pdf():
    """Histogram of mixture samples over 8 x 16 equal-area sphere bins
    against the pdf integrated per bin (Monte Carlo with 4M uniform points)."""
    rng = np.random.default_rng(12)
    raw = rng.normal(size=(1, 40))
    raw[0, 3::5] = [-2, 0, 1, 2, 3, 0.5, 1.5, 2.5]
    n = 400_000
    d, _ = G.mixture_sample(np.repeat(raw, n, 0), rng.random((n, 3)))

    def bin_of(v):
        ct = np.clip(((v[:, 2] + 1) / 2 * 8).astype(int), 0, 7)
        ph = np.clip(((np.arctan2(v[:, 1], v[:, 0]) + np.pi) / (2 * np.pi) * 16).astype(int), 0, 15)
        return ct * 16 + ph

    counts = np.bincount(bin_of(d), minlength=128)
    m = 4_000_000
    z = rng.uniform(-1, 1, m)
    ph = rng.uniform(-np.pi, np.pi, m)
    q = np.stack([np.sqrt(1 - z * z) * np.cos(ph), np.sqrt(1 - z * z) * np.sin(ph), z], axis=1)
    pdf = np.concatenate([G.mixture_pdf(np.repeat(raw, 500_000, 0), q[i:i + 500_000])
                          for i in range(0, m, 500_000)])
    prob = np.bincount(bin_of(q), weights=pdf, minlength=128) * (4 * np.pi / m)
    assert prob.sum() == pytest.approx(1.0, abs=0.01)
    expected = prob / prob.sum() * n
    keep = expected > 20
    chi2 = ((counts[keep] - expected[keep]) ** 2 / expected[keep]).sum()
    assert chi2 < 2.0 * keep.sum() + 40


def test_adam_first_steps_closed_form():
    rng = np.random.default_rng(2)
    theta = rng.normal(size=50)
    g1, g2 = rng.normal(size=50), rng.normal(size=50)
    m0 = v0 = np.zeros(50)
    t1, m1, v1 = G.adam_step(theta, m0, v0, g1, 1, lr=0.01)
    # step 1: bias-corrected moments are g and g^2
    np.testing.assert_allclose(t1, theta - 0.01 * g1 / (np.abs(g1) + 1e-8), rtol=1e-12)
    t2, m2, v2 = G.adam_step(t1, m1, v1, g2, 2, lr=0.01)
    mh = (0.9 * 0.1 * g1 + 0.1 * g2) / (1 - 0.81)
    vh = (0.999 * 0.001 * g1 ** 2 + 0.001 * g2 ** 2) / (1 - 0.999 ** 2)
    np.testing.assert_allclose(t2, t1 - 0.01 * mh / (np.sqrt(vh) + 1e-8), rtol=1e-10)


def test_bf16_backward_structure(monkeypatch):
    """With rounding disabled, backward_bf16 is backward() normalised by the
    total weight (so only its rounding points differ from the pinned form)."""
    theta = G.init_params(9)
    pos, nrm, wo, wi, w = _records(200, 10)
    x = G.encode(pos, nrm, wo)
    L, g = G.backward(theta.astype(np.float32).astype(np.float64), x, wi, w)
    monkeypatch.setattr(G, "bf16", lambda a: np.asarray(a, np.float64))
    Lb, gb = G.backward_bf16(theta, x, wi, w)
    assert Lb == pytest.approx(L / w.sum(), rel=1e-6)
    # raw outputs stay rounded to float32 (as the device stores them)
    np.testing.assert_allclose(gb, g / w.sum(), rtol=1e-4, atol=1e-6 * np.abs(g).max() / w.sum())
