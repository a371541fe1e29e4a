"""FP64 reference of the learned proposal -- TEST INFRASTRUCTURE ONLY.

The learned proposal has no counterpart in the reference (it is new work named
by the build's goal); this module is the CPU statement the GPU kernels are
checked against:

* features of a site from its 6 neighbours, in the frame of the local field
  h = sum_j S_j:  x = [beta, beta|h|, |h|/6, <q^2>, <q^3>, <q^4>],
  q_j = S_j . h/|h|  (rotation invariant, neighbour-order invariant);
* MLP: 6 -> H (ReLU) -> H (ReLU) -> H (ReLU) -> 3 outputs o;
  weights W1 [H,6], W2 [H,H], W3 [H,H], W4 [3,H], biases b1..b4;
* two-component mixture around h/|h|: kappa_k = softplus(o_k) + 1e-3,
  w_1 = sigmoid(o_2), w_2 = 1 - w_1; component densities in c = s . h/|h|:
    XY (von Mises, per radian):    f(c) = exp(kappa c) / (2 pi I0(kappa))
    Heisenberg (Fisher, per sr):   f(c) = kappa exp(kappa c) / (4 pi sinh kappa)
* loss = -log q(s | neighbours) (maximum likelihood of the single-site
  conditional), whose minimiser is the exact conditional (heat bath);
* Metropolis-Hastings ratio for s -> s':
    r = exp(beta J h.(s' - s)) * q(s) / q(s').
"""
import numpy as np
from scipy.special import i0e, i1e


def features(beta, nbrs):
    """nbrs: (n, 6, ncomp) neighbour spins -> (x (n,6), hhat (n,ncomp), |h| (n,))."""
    h = nbrs.sum(axis=1)
    hn = np.linalg.norm(h, axis=1)
    hhat = h / hn[:, None]
    q = np.einsum("njc,nc->nj", nbrs, hhat)
    x = np.stack([np.full(len(hn), beta), beta * hn, hn / 6.0, (q ** 2).mean(1), (q ** 3).mean(1), (q ** 4).mean(1)],
                 axis=1)
    return x, hhat, hn


def init_params(H, seed):
    rng = np.random.default_rng(seed)
    p = {
        "W1": rng.normal(0, np.sqrt(2 / 6), (H, 6)), "b1": np.zeros(H),
        "W2": rng.normal(0, np.sqrt(2 / H), (H, H)), "b2": np.zeros(H),
        "W3": rng.normal(0, np.sqrt(2 / H), (H, H)), "b3": np.zeros(H),
        "W4": rng.normal(0, 0.1 / np.sqrt(H), (3, H)), "b4": np.zeros(3),
    }
    return p


def relu(z):
    return np.maximum(z, 0.0)


def pre_activations(p, x):
    """Hidden pre-activations z1, z2, z3 (the ReLU kinks are where they vanish)."""
    z1 = x @ p["W1"].T + p["b1"]
    z2 = relu(z1) @ p["W2"].T + p["b2"]
    z3 = relu(z2) @ p["W3"].T + p["b3"]
    return z1, z2, z3


def kink_margin(p, x):
    """Per sample, the smallest |pre-activation| over every hidden unit: central
    differences of the loss with steps well below it never cross a ReLU kink."""
    return np.min(np.abs(np.concatenate(pre_activations(p, x), axis=1)), axis=1)


def forward(p, x):
    a1 = relu(x @ p["W1"].T + p["b1"])
    a2 = relu(a1 @ p["W2"].T + p["b2"])
    a3 = relu(a2 @ p["W3"].T + p["b3"])
    o = a3 @ p["W4"].T + p["b4"]
    return o, (a1, a2, a3)


def mixture(o):
    kap = np.logaddexp(0.0, o[:, :2]) + 1e-3
    w1 = 1.0 / (1.0 + np.exp(-o[:, 2]))
    return kap, np.stack([w1, 1.0 - w1], axis=1)


def log_component(ncomp, kap, c):
    """log f_k(c) for each component; kap (n,2), c (n,) -> (n,2)."""
    if ncomp == 2:
        return kap * (c[:, None] - 1.0) - np.log(2 * np.pi) - np.log(i0e(kap))
    return np.log(kap) + kap * (c[:, None] - 1.0) - np.log(2 * np.pi) - np.log(-np.expm1(-2 * kap))


def log_q(ncomp, o, c):
    kap, w = mixture(o)
    lf = log_component(ncomp, kap, c)
    m = lf.max(axis=1, keepdims=True)
    return (m[:, 0] + np.log((w * np.exp(lf - m)).sum(axis=1)))


def loss(p, ncomp, x, c):
    o, _ = forward(p, x)
    return -log_q(ncomp, o, c).mean()


def grad_fd(p, ncomp, x, c, eps=1e-6):
    """Central finite differences of the mean loss, every parameter (FP64)."""
    g = {}
    for k, v in p.items():
        gk = np.zeros_like(v)
        it = np.nditer(v, flags=["multi_index"])
        for _ in it:
            i = it.multi_index
            old = v[i]
            v[i] = old + eps
            lp = loss(p, ncomp, x, c)
            v[i] = old - eps
            lm = loss(p, ncomp, x, c)
            v[i] = old
            gk[i] = (lp - lm) / (2 * eps)
        g[k] = gk
    return g


def mean_cos(ncomp, kap):
    """E[c] under one component: I1/I0 (XY) or coth k - 1/k (Heisenberg)."""
    if ncomp == 2:
        return i1e(kap) / i0e(kap)
    return 1.0 / np.tanh(kap) - 1.0 / kap


def mh_log_ratio(ncomp, beta, J, hn, o, c_old, c_new):
    """log r for s -> s' with c = s . hhat; h.(s' - s) = |h| (c_new - c_old)."""
    return beta * J * hn * (c_new - c_old) + log_q(ncomp, o, c_old) - log_q(ncomp, o, c_new)


def pack_params(p, H):
    """Flat FP32 parameter vector in the device order: W1,b1,W2,b2,W3,b3,W4,b4 (row-major)."""
    return np.concatenate([p[k].reshape(-1) for k in ("W1", "b1", "W2", "b2", "W3", "b3", "W4", "b4")])


def unpack_params(v, H):
    shapes = (("W1", (H, 6)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("W3", (H, H)), ("b3", (H,)),
              ("W4", (3, H)), ("b4", (3,)))
    p, i = {}, 0
    for k, s in shapes:
        n = int(np.prod(s))
        p[k] = np.asarray(v[i:i + n], np.float64).reshape(s).copy()
        i += n
    return p


def n_params(H):
    return 6 * H + H + 2 * (H * H + H) + 3 * H + 3
