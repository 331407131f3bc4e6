"""Float64 reference of the learned path-guiding network (test infrastructure
only: the checker of the device kernels, never part of the product).

The network has no counterpart in the reference (its `path` integrator samples
the BSDF only, module/rad.offline/src/renderer/path.cpp:192-210); this file
states the mathematics the device kernels implement:

* inputs: position (3 x 16 one-blob bins over the scene bounds), geometric
  normal and outgoing direction (polar (y + 1) / 2 and azimuth
  atan2(z, x) / 2pi + 1/2, 8 bins each) -> 80 features; one-blob value of a
  scalar s in bin i of k: max(0, 1 - |s k - (i + 1/2)|);
* 5 dense layers 80-64-64-64-64-48, ReLU between, parameters flat per layer
  as W[out][in] then b[out];
* outputs: 8 lobes x (logit, target xyz, log kappa) -> softmax weights,
  means pointing from the vertex x to the lobe's target point c + h t
  (c the bounds' centre, h their half extent, t the three outputs; a lobe
  probe without a vertex takes the outputs as the mean direction itself),
  kappa = exp(clamp(., -3, 6));
* vMF density kappa / (2 pi (1 - e^{-2 kappa})) e^{kappa (mu.w - 1)};
* loss (1/n) sum_j min(w_j * wnorm, 64) * -log q(dir_j | x_j).

Analytic gradients are pinned by central finite differences in the tests.
"""
import numpy as np

K_IN, K_HID, K_OUT, LOBES = 80, 64, 48, 8
DIMS = [(80, 64), (64, 64), (64, 64), (64, 64), (64, 48)]
MAX_WEIGHT = 64.0


def offsets():
    o, out = 0, []
    for nin, nout in DIMS:
        out.append((o, o + nin * nout))
        o += nin * nout + nout
    return out, o


OFFS, N_PARAMS = offsets()


def unpack(params):
    p = np.asarray(params, np.float64)
    layers = []
    for (wo, bo), (nin, nout) in zip(OFFS, DIMS):
        layers.append((p[wo:wo + nin * nout].reshape(nout, nin), p[bo:bo + nout]))
    return layers


def scalars(feats, lo, hi):
    f = np.asarray(feats, np.float64).reshape(-1, 9)
    lo, hi = np.asarray(lo, np.float64), np.asarray(hi, np.float64)
    ext = hi - lo
    inv = np.where(ext > 0, 1.0 / np.where(ext > 0, ext, 1.0), 0.0)
    s = np.empty((len(f), 7))
    s[:, 0:3] = np.clip((f[:, 0:3] - lo) * inv, 0.0, 1.0)
    for k, base in ((3, 3), (5, 6)):
        d = f[:, base:base + 3]
        s[:, k] = np.clip(0.5 * (d[:, 1] + 1.0), 0.0, 1.0)
        s[:, k + 1] = np.clip(np.arctan2(d[:, 2], d[:, 0]) / (2 * np.pi) + 0.5, 0.0, 1.0)
    return s


def one_blob(s):
    cols = []
    for dim, bins in [(0, 16), (1, 16), (2, 16), (3, 8), (4, 8), (5, 8), (6, 8)]:
        x = s[:, dim:dim + 1] * bins - (np.arange(bins) + 0.5)
        cols.append(np.maximum(0.0, 1.0 - np.abs(x)))
    return np.concatenate(cols, axis=1)


def forward(params, feats, lo, hi):
    """Returns (layer inputs A_0..A_4, pre-activations Z_0..Z_4); Z_4 = raw outputs."""
    A = [one_blob(scalars(feats, lo, hi))]
    Z = []
    for i, (W, b) in enumerate(unpack(params)):
        z = A[-1] @ W.T + b
        Z.append(z)
        if i < len(DIMS) - 1:
            A.append(np.maximum(z, 0.0))
    return A, Z


def lobe_frame(lo, hi):
    """Centre and half extent of the lobe-target frame."""
    lo, hi = np.asarray(lo, np.float64), np.asarray(hi, np.float64)
    return 0.5 * (lo + hi), np.maximum(0.5 * (hi - lo), 1e-6)


def mixture(raw, pos=None, lo=None, hi=None):
    raw = np.asarray(raw, np.float64).reshape(-1, K_OUT)[:, :5 * LOBES].reshape(-1, LOBES, 5)
    a = raw[:, :, 0]
    pi = np.exp(a - a.max(axis=1, keepdims=True))
    pi /= pi.sum(axis=1, keepdims=True)
    m = raw[:, :, 1:4]
    if pos is not None:
        c, h = lobe_frame(lo, hi)
        m = c + h * m - np.asarray(pos, np.float64).reshape(-1, 1, 3)
    ln = np.linalg.norm(m, axis=2, keepdims=True)
    mu = np.where(ln > 1e-6, m / np.where(ln > 1e-6, ln, 1.0), np.array([0.0, 1.0, 0.0]))
    kr = raw[:, :, 4]
    kappa = np.exp(np.clip(kr, -3.0, 6.0))
    return pi, mu, kappa, ln[:, :, 0], kr


def log_vmf(mu, kappa, w):
    c = np.einsum("nlk,nk->nl", mu, w)
    return np.log(kappa) - np.log(2 * np.pi) - np.log1p(-np.exp(-2 * kappa)) + kappa * (c - 1.0), c


def pdf(raw, dirs):
    """Mixture density of one raw row at each direction."""
    dirs = np.asarray(dirs, np.float64).reshape(-1, 3)
    pi, mu, kappa, _, _ = mixture(np.broadcast_to(np.asarray(raw, np.float64).reshape(1, -1)[:, :K_OUT], (1, K_OUT)))
    lv, _ = log_vmf(np.repeat(mu, len(dirs), 0), np.repeat(kappa, len(dirs), 0), dirs)
    return (np.repeat(pi, len(dirs), 0) * np.exp(lv)).sum(axis=1)


def loss_and_grad(params, feats, dirs, w, wnorm, lo, hi):
    """Mean loss and its gradient with respect to the flat parameters."""
    A, Z = forward(params, feats, lo, hi)
    raw = Z[-1]
    n = len(raw)
    dirs = np.asarray(dirs, np.float64).reshape(-1, 3)
    wn = np.minimum(np.asarray(w, np.float64) * wnorm, MAX_WEIGHT)
    wn = np.where(np.asarray(w) > 0, wn, 0.0)
    pos = np.asarray(feats, np.float64).reshape(-1, 9)[:, 0:3]
    pi, mu, kappa, ln, kr = mixture(raw, pos, lo, hi)
    lv, c = log_vmf(mu, kappa, dirs)
    t = np.log(pi) + lv
    tm = t.max(axis=1, keepdims=True)
    logq = tm[:, 0] + np.log(np.exp(t - tm).sum(axis=1))
    r = np.exp(t - logq[:, None])
    loss = (wn * -logq).mean()
    d = np.zeros((n, LOBES, 5))
    d[:, :, 0] = -wn[:, None] * (r - pi)
    e = np.exp(-2 * kappa)
    dk = 1.0 / kappa - 2 * e / (1 - e) + (c - 1.0)
    inside = (kr > -3.0) & (kr < 6.0)
    d[:, :, 4] = np.where(inside, -wn[:, None] * r * dk * kappa, 0.0)
    gm = (dirs[:, None, :] - mu * c[:, :, None]) * (kappa / np.where(ln > 1e-6, ln, 1.0))[:, :, None]
    gm = np.where((ln > 1e-6)[:, :, None], gm, 0.0) * lobe_frame(lo, hi)[1]
    d[:, :, 1:4] = -wn[:, None, None] * r[:, :, None] * gm
    dz = np.zeros((n, K_OUT))
    dz[:, :5 * LOBES] = d.reshape(n, -1)
    grad = np.zeros(N_PARAMS)
    layers = unpack(params)
    for i in range(len(DIMS) - 1, -1, -1):
        W, _ = layers[i]
        wo, bo = OFFS[i]
        grad[wo:wo + W.size] = (dz.T @ A[i]).reshape(-1) / n
        grad[bo:bo + W.shape[0]] = dz.sum(axis=0) / n
        if i > 0:
            dz = (dz @ W) * (A[i] > 0)
    return loss, grad


def init_params(seed=0):
    rng = np.random.default_rng(seed)
    p = np.zeros(N_PARAMS, np.float32)
    for i, ((wo, _), (nin, nout)) in enumerate(zip(OFFS, DIMS)):
        lim = np.sqrt(6.0 / nin) * (0.1 if i == len(DIMS) - 1 else 1.0)
        p[wo:wo + nin * nout] = rng.uniform(-lim, lim, nin * nout)
    return p
