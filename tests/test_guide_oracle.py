"""CPU pins of the guiding-network restatement (oracle/guide_ref.py): closed forms of the von
Mises-Fisher mixture, central finite differences of the float64 loss, and the published Adam
update. These fix the checker that the GPU guiding kernels are compared with."""
import numpy as np
import pytest

from oracle import guide_ref as G


def _fib_sphere(n):
    i = np.arange(n) + 0.5
    z = 1 - 2 * i / n
    r = np.sqrt(1 - z * z)
    phi = np.pi * (1 + 5 ** 0.5) * i
    return np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=1)


def _random_outputs(rng, n, kappa_hi=3.0):
    o = rng.normal(size=(n, 40))
    o[:, 3::5] = rng.uniform(-2, kappa_hi, size=(n, 8))
    return o


def test_mixture_pdf_integrates_to_one():
    rng = np.random.default_rng(0)
    o = _random_outputs(rng, 4)
    pts = _fib_sphere(400_000)
    for k in range(len(o)):
        lq = G.mixture_log_pdf(np.repeat(o[k:k + 1], len(pts), axis=0), pts)
        integral = np.exp(lq).mean() * 4 * np.pi
        assert abs(integral - 1) < 2e-3, integral


def test_vmf_small_kappa_is_uniform_and_normalisation_is_stable():
    mu = np.array([0.0, 0.0, 1.0])
    for k in (1e-6, 1e-3):
        lp = G.vmf_log_pdf(mu, np.array(k), np.array([1.0, 0.0, 0.0]))
        assert abs(np.exp(lp) * 4 * np.pi - 1) < 2 * k
    # large kappa: peak density kappa / (2 pi)
    lp = G.vmf_log_pdf(mu, np.array(1500.0), mu)
    assert abs(np.exp(lp) - 1500 / (2 * np.pi)) < 1e-6 * 1500


@pytest.mark.parametrize("kappa", [0.5, 5.0, 80.0])
def test_vmf_sampling_mean_cosine(kappa):
    """E[mu.w] = coth(kappa) - 1/kappa for the vMF on S^2."""
    rng = np.random.default_rng(1)
    mu = np.array([0.3, -0.5, 0.8])
    mu /= np.linalg.norm(mu)
    w = G.vmf_sample(mu, kappa, rng.random(200_000), rng.random(200_000))
    assert np.allclose(np.linalg.norm(w, axis=1), 1, atol=1e-9)
    c = w @ mu
    expect = 1 / np.tanh(kappa) - 1 / kappa
    assert abs(c.mean() - expect) < 4 * c.std() / np.sqrt(len(c))


def test_output_gradient_matches_finite_differences():
    rng = np.random.default_rng(2)
    o = _random_outputs(rng, 6)
    om = rng.normal(size=(6, 3))
    om /= np.linalg.norm(om, axis=1, keepdims=True)
    w = rng.uniform(0.2, 3, 6)
    d, loss = G.output_grad(o, om, w)
    for n in range(6):
        for k in range(40):
            h = 1e-6
            op, omn = o.copy(), o.copy()
            op[n, k] += h
            omn[n, k] -= h
            fp = -w[n] * G.mixture_log_pdf(op[n:n + 1], om[n:n + 1])[0]
            fm = -w[n] * G.mixture_log_pdf(omn[n:n + 1], om[n:n + 1])[0]
            assert abs((fp - fm) / (2 * h) - d[n, k]) < 1e-6 * max(1, abs(d[n, k])), (n, k)
    assert np.allclose(loss, -w * G.mixture_log_pdf(o, om))


def _records(rng, n):
    rec = np.zeros((n, 16))
    rec[:, 0:3] = rng.random((n, 3))
    rec[:, 3] = rng.random(n)
    wo = rng.normal(size=(n, 3))
    rec[:, 4:7] = wo / np.linalg.norm(wo, axis=1, keepdims=True)
    rec[:, 7] = rng.random(n) < 0.3
    nn = rng.normal(size=(n, 3))
    rec[:, 8:11] = np.where(rec[:, 7:8] > 0, 0, nn / np.linalg.norm(nn, axis=1, keepdims=True))
    rec[:, 11] = rng.exponential(1.0, n)  # weight: w of the third float4
    om = rng.normal(size=(n, 3))
    rec[:, 12:15] = om / np.linalg.norm(om, axis=1, keepdims=True)
    return rec


def test_record_layout():
    """Record = 4 float4: (pos, t), (wo, medium), (normal, weight), (direction, 0)."""
    rec = np.arange(16, dtype=np.float64)[None, :]
    raw = G.record_raw(rec)
    assert raw[0].tolist() == [0, 1, 2, 4, 5, 6, 8, 9, 10, 3, 7, 0]


def test_parameter_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    p = G.initial_like(seed=4, out_scale=1.0).astype(np.float64)
    rec = _records(rng, 24)
    g = G.gradient(p, rec)
    assert g.shape == (G.N_PARAMS,)
    idx = np.concatenate([rng.choice(G.N_PARAMS, 40, replace=False),
                          np.arange(G.N_PARAMS - 40, G.N_PARAMS)])  # include output layer and biases
    for i in idx:
        h = 1e-5
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (G.loss(pp, rec) - G.loss(pm, rec)) / (2 * h)
        assert abs(fd - g[i]) < 1e-5 + 1e-4 * abs(g[i]), (i, fd, g[i])


def test_planted_gradient_mistake_is_caught():
    """Dropping the (mu.w - 1) term of d log vMF / d kappa must fail the finite-difference pin."""
    rng = np.random.default_rng(5)
    o = _random_outputs(rng, 3)
    om = rng.normal(size=(3, 3))
    om /= np.linalg.norm(om, axis=1, keepdims=True)
    w = np.ones(3)
    d, _ = G.output_grad(o, om, w)
    mu, kappa, pi = G.mixture(o)
    c = np.einsum("nik,nk->ni", mu, om)
    sig = 1 / (1 + np.exp(-o.reshape(3, 8, 5)[:, :, 3]))
    # the planted version differs from the analytic one exactly by the dropped term
    lt = np.log(pi) + G.vmf_log_pdf(mu, kappa, om[:, None, :])
    gamma = np.exp(lt - lt.max(1, keepdims=True))
    gamma /= gamma.sum(1, keepdims=True)
    planted = d.reshape(3, 8, 5)[:, :, 3] - (-w[:, None] * gamma) * (c - 1) * sig
    n, k = 0, 3
    h = 1e-6
    op, omn = o.copy(), o.copy()
    op[n, k] += h
    omn[n, k] -= h
    fd = (-G.mixture_log_pdf(op[:1], om[:1])[0] + G.mixture_log_pdf(omn[:1], om[:1])[0]) / (2 * h)
    assert abs(fd - d[0, 3]) < 1e-6 * max(1, abs(fd))
    assert abs(fd - planted[0, 0]) > 1e-4


def test_adam_first_step_closed_form():
    rng = np.random.default_rng(6)
    p = rng.normal(size=50)
    g = rng.normal(size=50)
    p1, m, v = G.adam(p, np.zeros(50), np.zeros(50), g, t=1, lr=1e-3)
    # first step: m_hat = g, v_hat = g^2 -> update lr * g / (|g| + eps)
    assert np.allclose(p1, p - 1e-3 * g / (np.abs(g) + 1e-8), rtol=0, atol=1e-12)
    p2, _, _ = G.adam(p1, m, v, g, t=2, lr=1e-3)
    assert np.allclose(p2, p1 - 1e-3 * g / (np.abs(g) + 1e-8), atol=1e-12)


def test_bf16_rounding_ties_to_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5], dtype=np.float32)
    r = G.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 2 ** -6 and r[3] == 1.0 and r[4] == -2.5
