"""Learned proposal: gradients against FP64 finite differences of the oracle's
loss, exactness of the Metropolis-Hastings correction, equilibrium agreement
with the plain chain, and what training gains over the untrained network."""
import math
import os
import sys

import numpy as np
import pytest

import oracle_lib as o

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import learned as ref  # noqa: E402

pytestmark = pytest.mark.gpu
ising_hip = pytest.importorskip("ising_hip")


def bf16_round(x):
    """Round FP32 values to BF16 (nearest even), returned as float64."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def random_samples(nc, n, beta, seed):
    rng = np.random.default_rng(seed)
    axis = rng.normal(size=(n, 1, nc))
    nb = axis + rng.normal(size=(n, 6, nc)) * 0.9
    nb /= np.linalg.norm(nb, axis=2, keepdims=True)
    x, hhat, hn = ref.features(beta, nb)
    s = hhat + rng.normal(size=(n, nc)) * 0.7
    s /= np.linalg.norm(s, axis=1, keepdims=True)
    c = (s * hhat).sum(axis=1)
    return x.astype(np.float32), c.astype(np.float32)


def ensemble(model, L=4, R=1, temps=None, seed=3, **kw):
    if temps is None:
        temps = np.full(R, 1.0)
    return ising_hip.PTEnsemble(model, [L], R, np.asarray(temps, np.float64), seed, **kw)


@pytest.mark.parametrize("model,nc", [("xy", 2), ("heisenberg", 3)])
def test_gradients_match_fp64_finite_differences(model, nc):
    H = 8
    pt = ensemble(model)
    pt.learn_init(H, seed=11)
    rng = np.random.default_rng(5)
    params = (rng.normal(size=ref.n_params(H)) * 0.5).astype(np.float32)
    pt.learn_set_params(params)
    x, c = random_samples(nc, 256, 0.45, 9)
    p64 = ref.unpack_params(bf16_round(params), H)
    # central differences need a smooth loss: keep samples clear of the ReLU kinks
    keep = ref.kink_margin(p64, x.astype(np.float64)) > 1e-3
    assert keep.sum() > 200
    x, c = x[keep], c[keep]
    loss_gpu, g_gpu = pt.learn_eval(x, c)
    x64, c64 = x.astype(np.float64), c.astype(np.float64)
    loss_ref = ref.loss(p64, nc, x64, c64)
    assert abs(loss_gpu - loss_ref) < 2e-5 * max(1.0, abs(loss_ref)), (loss_gpu, loss_ref)
    g_fd = ref.pack_params(ref.grad_fd(p64, nc, x64, c64), H)
    scale = np.max(np.abs(g_fd))
    err = np.abs(g_gpu - g_fd)
    assert np.all(err <= 2e-3 * scale + 2e-3 * np.abs(g_fd)), (np.max(err), scale)
    cos = g_gpu @ g_fd / (np.linalg.norm(g_gpu) * np.linalg.norm(g_fd))
    assert cos > 0.9999


def forward_bf16(p, x):
    """FP64 forward with the fused sweep's roundings: BF16 inputs and
    activations (weights are BF16 already)."""
    a = bf16_round(x)
    for k in (1, 2, 3):
        a = bf16_round(np.maximum(a @ p["W%d" % k].T + p["b%d" % k], 0.0))
    return a @ p["W4"].T + p["b4"]


@pytest.mark.parametrize("H", [8, 16, 32, 64])
def test_matrix_core_inference_matches_the_bf16_forward(H):
    nc = 2
    pt = ensemble("xy")
    pt.learn_init(H, seed=3)
    rng = np.random.default_rng(H)
    params = (rng.normal(size=ref.n_params(H)) * (1.5 / np.sqrt(H))).astype(np.float32)
    pt.learn_set_params(params)
    x, _ = random_samples(nc, 1000, 0.45, 21)
    got = pt.learn_infer(x).astype(np.float64)
    p64 = ref.unpack_params(bf16_round(params), H)
    want = forward_bf16(p64, x.astype(np.float64))
    full, _ = ref.forward(p64, x.astype(np.float64))
    err = np.abs(got - want)
    scale = np.max(np.abs(want))
    # accumulation order can move an activation across a BF16 rounding
    # boundary now and then; anything systematic (a layout or
    # padding mistake) shows as errors of the order of the outputs themselves
    assert np.median(err) < 1e-3 * scale and np.max(err) < 3e-2 * scale, (np.median(err), np.max(err), scale)
    assert np.max(np.abs(got - full)) < 0.1 * scale


def test_inference_follows_the_parameters_after_training():
    """The fused pass reads the network as per-lane fragments built from the
    BF16 parameters; after Adam steps (which rewrite the parameters on the
    device) the inference path must use the updated network."""
    H, nc = 16, 2
    pt = ensemble("xy", L=8, R=4, temps=np.linspace(2.0, 2.4, 4))
    pt.warmup(50)
    pt.learn_init(H, seed=5)
    x, _ = random_samples(nc, 512, 0.45, 31)
    before = pt.learn_infer(x).astype(np.float64)
    pt.learn_train(20, batch=1024, lr=3e-2)
    after = pt.learn_infer(x).astype(np.float64)
    p64 = ref.unpack_params(bf16_round(pt.learn_params()), H)
    want = forward_bf16(p64, x.astype(np.float64))
    scale = np.max(np.abs(want))
    assert np.max(np.abs(after - before)) > 0.05 * scale  # training moved the outputs
    err = np.abs(after - want)
    assert np.median(err) < 1e-3 * scale and np.max(err) < 3e-2 * scale, (np.median(err), np.max(err), scale)


def test_hidden64_gradients_on_a_parameter_subset():
    H, nc = 64, 3
    pt = ensemble("heisenberg")
    pt.learn_init(H, seed=2)
    params = pt.learn_params()
    x, c = random_samples(nc, 512, 0.7, 4)
    p64 = ref.unpack_params(bf16_round(params), H)
    keep = ref.kink_margin(p64, x.astype(np.float64)) > 1e-3
    assert keep.sum() > 400
    x, c = x[keep], c[keep]
    _, g = pt.learn_eval(x, c)
    x64, c64 = x.astype(np.float64), c.astype(np.float64)
    flat = ref.pack_params(p64, H)
    rng = np.random.default_rng(1)
    for idx in rng.choice(len(flat), 24, replace=False):
        v = flat.copy()
        eps = 1e-5
        v[idx] += eps
        lp = ref.loss(ref.unpack_params(v, H), nc, x64, c64)
        v[idx] -= 2 * eps
        lm = ref.loss(ref.unpack_params(v, H), nc, x64, c64)
        fd = (lp - lm) / (2 * eps)
        assert abs(g[idx] - fd) < 2e-3 * np.max(np.abs(g)) + 2e-2 * abs(fd), (idx, g[idx], fd)


@pytest.mark.parametrize("model,nc,beta", [("heisenberg", 3, 0.6), ("xy", 2, 0.5)])
def test_learned_moves_sample_the_exact_conditional(model, nc, beta):
    """Colour 1 frozen; each colour-0 site must sample exp(beta s.h) whatever the
    (here untrained) network proposes: the MH ratio with q(s)/q(s') corrects it."""
    L = 4
    pt = ensemble(model, L=L, R=1, temps=[1.0 / beta], seed=8, delta=0.7)
    pt.learn_init(16, seed=4)
    pt.learn_enable(True)
    sp0 = pt.get_spins(0, 0).astype(np.float64).reshape(-1, nc)
    s = np.arange(L ** 3)
    colour0 = s[((s % L + (s // L) % L + s // (L * L)) & 1) == 0]
    hs = []
    for site in colour0:
        z, rem = divmod(site, L * L)
        y, x = divmod(rem, L)
        nbs = [((z * L + y) * L + (x + 1) % L), ((z * L + y) * L + (x - 1) % L), ((z * L + (y + 1) % L) * L + x),
               ((z * L + (y - 1) % L) * L + x), ((((z + 1) % L) * L + y) * L + x), ((((z - 1) % L) * L + y) * L + x)]
        hs.append(sp0[nbs].sum(axis=0))
    hs = np.array(hs)
    hn = np.linalg.norm(hs, axis=1)
    n_pass = 4000
    cs = np.zeros((n_pass, len(colour0)))
    for t in range(n_pass):
        pt.metropolis_pass(0, t)
        S = pt.get_spins(0, 0).astype(np.float64).reshape(-1, nc)[colour0]
        cs[t] = (S * hs).sum(axis=1) / hn
    tries, acc = pt.mh_stats()
    assert tries > 0 and acc < tries  # the untrained proposal is not the heat bath
    kap = beta * hn
    exact = ref.mean_cos(nc, kap)
    nb = 20
    bm = cs[: (n_pass // nb) * nb].reshape(nb, -1, len(colour0)).mean(axis=1)
    err = bm.std(axis=0, ddof=1) / math.sqrt(nb)
    z = (cs.mean(axis=0) - exact) / err
    assert np.all(np.abs(z) < 5.0), z
    assert abs(z.mean()) * math.sqrt(len(z)) < 5.0


@pytest.mark.parametrize("model,nc,beta", [("heisenberg", 3, 0.6), ("xy", 2, 0.5), ("heisenberg", 3, 2.5),
                                           ("xy", 2, 2.0)])
def test_heat_bath_samples_the_exact_conditional(model, nc, beta):
    """Colour 1 frozen; every heat-bath pass redraws each colour-0 site from
    exp(beta s.h): the mean and second moment of c = s.h^ per site against the
    closed forms (S^2: <c> = coth k - 1/k, <c^2> = 1 - 2 <c>/k; S^1:
    <c> = I1(k)/I0(k), <c^2> = 1 - <c>/k; k = beta |h|)."""
    L = 4
    pt = ensemble(model, L=L, R=1, temps=[1.0 / beta], seed=9, delta=0.7)
    pt.heat_bath(True)
    sp0 = pt.get_spins(0, 0).astype(np.float64).reshape(-1, nc)
    s = np.arange(L ** 3)
    colour0 = s[((s % L + (s // L) % L + s // (L * L)) & 1) == 0]
    hs = []
    for site in colour0:
        z, rem = divmod(site, L * L)
        y, x = divmod(rem, L)
        nbs = [((z * L + y) * L + (x + 1) % L), ((z * L + y) * L + (x - 1) % L), ((z * L + (y + 1) % L) * L + x),
               ((z * L + (y - 1) % L) * L + x), ((((z + 1) % L) * L + y) * L + x), ((((z - 1) % L) * L + y) * L + x)]
        hs.append(sp0[nbs].sum(axis=0))
    hs = np.array(hs)
    hn = np.linalg.norm(hs, axis=1)
    n_pass = 4000
    cs = np.zeros((n_pass, len(colour0)))
    for t in range(n_pass):
        pt.metropolis_pass(0, t)
        S = pt.get_spins(0, 0).astype(np.float64).reshape(-1, nc)[colour0]
        cs[t] = (S * hs).sum(axis=1) / hn
    assert np.array_equal(pt.get_spins(0, 0).reshape(-1, nc)[np.setdiff1d(s, colour0)],
                          sp0[np.setdiff1d(s, colour0)].astype(np.float32))
    kap = beta * hn
    m1 = ref.mean_cos(nc, kap)
    m2 = 1.0 - (2.0 if nc == 3 else 1.0) * m1 / kap
    nb = 20
    for series, exact in ((cs, m1), (cs ** 2, m2)):
        bm = series[: (n_pass // nb) * nb].reshape(nb, -1, len(colour0)).mean(axis=1)
        err = bm.std(axis=0, ddof=1) / math.sqrt(nb)
        z = (series.mean(axis=0) - exact) / err
        assert np.all(np.abs(z) < 5.0), z
        assert abs(z.mean()) * math.sqrt(len(z)) < 5.0
    # passes redraw independently: consecutive draws of a site are uncorrelated
    rho = np.mean([np.corrcoef(cs[:-1, i], cs[1:, i])[0, 1] for i in range(len(colour0))])
    assert abs(rho) < 0.05, rho


@pytest.mark.parametrize("model,temps", [("xy", np.linspace(2.18, 2.22, 4)), ("heisenberg", np.linspace(1.43, 1.46, 4))])
def test_heat_bath_chain_keeps_equilibrium(model, temps):
    """Heat-bath + overrelaxation sweeps against plain Metropolis +
    overrelaxation: energies per temperature within 4.5 combined sigma."""
    R, L, n = len(temps), 8, 20000
    kw = dict(delta=1.0, n_overrelax=2, exchange_every=10, measure_every=1, record_capacity=n + 2)
    res = {}
    for mode in ("plain", "heat"):
        pt = ising_hip.PTEnsemble(model, [L], R, np.asarray(temps), 3, **kw)
        pt.warmup(500)
        if mode == "heat":
            pt.heat_bath(True)
        pt.sample(n)
        tidx, e, mx, my, mz = pt.read_cont()
        res[mode] = (tidx[:, 0, :], e[:, 0, :] / L ** 3)
    for k in range(R):
        means = {}
        for mode in res:
            t, e = res[mode]
            ek = e[t == k]
            bm = ek[: len(ek) // 20 * 20].reshape(20, -1).mean(axis=1)
            means[mode] = (bm.mean(), bm.std(ddof=1) / math.sqrt(20))
        (a, ea), (b, eb) = means["plain"], means["heat"]
        assert abs(a - b) < 4.5 * math.hypot(ea, eb), (k, a, b, ea, eb)


def plain_vs_learned(model, temps, L, hidden, train_steps, sweeps, seed=1):
    R = len(temps)
    kw = dict(delta=1.0, n_overrelax=2, exchange_every=10, measure_every=1, record_capacity=sweeps + 2)
    out = {}
    for mode in ("plain", "untrained", "learned"):
        pt = ising_hip.PTEnsemble(model, [L], R, np.asarray(temps), seed, **kw)
        pt.warmup(500)
        if mode != "plain":
            pt.learn_init(hidden, seed=7)
            if mode == "learned":
                pt.learn_train(train_steps, batch=8192, lr=3e-3)
            pt.learn_enable(True)
        pt.sample(sweeps)
        tidx, e, mx, my, mz = pt.read_cont()
        out[mode] = (tidx[:, 0, :], e[:, 0, :] / L ** 3, pt.mh_stats())
    return out


@pytest.mark.parametrize("model,temps", [("xy", np.linspace(2.18, 2.22, 4)), ("heisenberg", np.linspace(1.43, 1.46, 4))])
def test_training_raises_acceptance_and_keeps_equilibrium(model, temps):
    """Trained on the plain chain's configurations, the network raises the
    learned moves' acceptance above the untrained network's, and the learned
    chain's energies agree with the plain chain's per temperature."""
    res = plain_vs_learned(model, temps, 8, 16, 400, 20000)
    _, _, (tu, au) = res["untrained"]
    _, _, (tl, al) = res["learned"]
    assert al / tl > 0.9 and al / tl > au / tu, (au / tu, al / tl)
    for k in range(4):
        means = {}
        for mode in ("plain", "learned"):
            t, e, _ = res[mode]
            ek = e[t == k]
            bm = ek[: len(ek) // 20 * 20].reshape(20, -1).mean(axis=1)
            means[mode] = (bm.mean(), bm.std(ddof=1) / math.sqrt(20))
        (a, ea), (b, eb) = means["plain"], means["learned"]
        assert abs(a - b) < 4.5 * math.hypot(ea, eb), (k, a, b, ea, eb)
