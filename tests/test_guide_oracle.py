"""Pins of the float64 guiding oracle (oracle/guide_ref.py): analytic loss
gradients against central finite differences, the vMF mixture density
integrating to 1 by quadrature, and the one-blob encoding's partition of
unity."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import guide_ref as G  # noqa: E402


def random_batch(rng, n):
    feats = np.concatenate([rng.uniform(-1, 1, (n, 3)), rng.normal(size=(n, 3)), rng.normal(size=(n, 3))], 1)
    feats[:, 3:6] /= np.linalg.norm(feats[:, 3:6], axis=1, keepdims=True)
    feats[:, 6:9] /= np.linalg.norm(feats[:, 6:9], axis=1, keepdims=True)
    dirs = rng.normal(size=(n, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    w = rng.exponential(1.0, n)
    return feats, dirs, w


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = G.init_params(1).astype(np.float64)
    # larger output weights so every lobe term carries gradient
    wo, _ = G.OFFS[-1]
    p[wo:wo + 64 * 48] *= 10.0
    feats, dirs, w = random_batch(rng, 6)
    lo, hi = np.full(3, -1.0), np.full(3, 1.0)
    loss, g = G.loss_and_grad(p, feats, dirs, w, 1.0, lo, hi)
    assert np.isfinite(loss) and np.isfinite(g).all()
    idx = []
    for wo, bo in G.OFFS:
        idx += list(rng.choice(np.arange(wo, bo), 12, replace=False)) + list(range(bo, bo + 3))
    h = 1e-6
    for i in idx:
        q = p.copy()
        q[i] += h
        lp, _ = G.loss_and_grad(q, feats, dirs, w, 1.0, lo, hi)
        q[i] -= 2 * h
        lm, _ = G.loss_and_grad(q, feats, dirs, w, 1.0, lo, hi)
        num = (lp - lm) / (2 * h)
        assert abs(num - g[i]) <= 1e-6 + 1e-4 * abs(num), (i, num, g[i])


def test_weight_clamp_and_zero_weights():
    rng = np.random.default_rng(1)
    p = G.init_params(2)
    feats, dirs, w = random_batch(rng, 5)
    lo, hi = np.full(3, -1.0), np.full(3, 1.0)
    w2 = w.copy()
    w2[0] = 1e9  # clamped to 64 / wnorm
    l1, _ = G.loss_and_grad(p, feats, dirs, np.where(np.arange(5) == 0, 64.0, w), 1.0, lo, hi)
    l2, _ = G.loss_and_grad(p, feats, dirs, w2, 1.0, lo, hi)
    assert np.isclose(l1, l2)
    l0, g0 = G.loss_and_grad(p, feats, dirs, np.zeros(5), 1.0, lo, hi)
    assert l0 == 0.0 and not g0.any()


def sphere_quadrature(n_theta=400, n_phi=800):
    t = (np.arange(n_theta) + 0.5) * np.pi / n_theta
    f = (np.arange(n_phi) + 0.5) * 2 * np.pi / n_phi
    T, F = np.meshgrid(t, f, indexing="ij")
    d = np.stack([np.sin(T) * np.cos(F), np.cos(T), np.sin(T) * np.sin(F)], -1).reshape(-1, 3)
    dw = (np.sin(T) * (np.pi / n_theta) * (2 * np.pi / n_phi)).reshape(-1)
    return d, dw


def test_mixture_pdf_integrates_to_one():
    rng = np.random.default_rng(3)
    d, dw = sphere_quadrature()
    for _ in range(4):
        raw = rng.normal(size=48)
        raw[4::5][:8] = rng.uniform(-3, 4, 8)  # kappa from 0.05 to 55
        assert abs((G.pdf(raw, d) * dw).sum() - 1.0) < 2e-3


def test_vmf_closed_form():
    # a single dominant lobe: density at the mean is kappa / (2 pi (1 - e^{-2 kappa}))
    raw = np.zeros(48)
    raw[0::5][:8] = -50.0
    raw[0] = 50.0
    raw[1:4] = [0.0, 0.0, 2.0]
    raw[4] = np.log(10.0)
    k = 10.0
    assert np.isclose(G.pdf(raw, [[0, 0, 1]])[0], k / (2 * np.pi * (1 - np.exp(-2 * k))))
    assert np.isclose(G.pdf(raw, [[0, 0, -1]])[0], k / (2 * np.pi * (1 - np.exp(-2 * k))) * np.exp(-2 * k))


def test_one_blob_partition_of_unity():
    s = np.linspace(0.0, 1.0, 101)
    S = np.tile(s[:, None], (1, 7))
    X = G.one_blob(S)
    assert X.shape == (101, 80)
    # interior scalars: the blob weights of each dimension sum to 1
    for start, bins in [(0, 16), (48, 8)]:
        tot = X[:, start:start + bins].sum(axis=1)
        inner = (s * bins >= 0.5) & (s * bins <= bins - 0.5)
        assert np.allclose(tot[inner], 1.0)
        assert (tot <= 1.0 + 1e-12).all()


def test_lobe_means_point_at_their_target():
    """A lobe's mean is the unit vector from the vertex to its target point
    c + h t (c, h: centre and half extent of the bounds, t: the lobe's three
    outputs), so one output places a target that every vertex sees at its
    own direction; without a vertex (probes) the outputs are the direction."""
    lo, hi = np.array([-3.0, 0.0, -3.0]), np.array([3.0, 3.0, 3.0])
    target = np.array([2.0, 1.5, -1.0])
    c, h = 0.5 * (lo + hi), 0.5 * (hi - lo)
    raw = np.zeros(G.K_OUT)
    raw[1:4] = (target - c) / h
    verts = np.array([[0.0, 0.0, 0.0], [-2.5, 2.0, 2.0], [2.0, 0.2, -1.0]])
    _, mu, _, ln, _ = G.mixture(np.tile(raw, (3, 1)), verts, lo, hi)
    want = target - verts
    assert np.allclose(ln[:, 0], np.linalg.norm(want, axis=1))
    assert np.allclose(mu[:, 0], want / np.linalg.norm(want, axis=1, keepdims=True))
    _, mu_dir, _, _, _ = G.mixture(raw[None, :])
    assert np.allclose(mu_dir[0, 0], raw[1:4] / np.linalg.norm(raw[1:4]))
