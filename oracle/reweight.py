"""FP64 restatement of the reference's WHAM reweighting -- TEST INFRASTRUCTURE ONLY.

Follows analysis/scripts/reweighting.py:45-108 (wham_reweight) step for step
(free-energy iteration, shift f_0 = 0, tolerance test, weights at the target
from the last iteration's denominator), and returns reweighted means of given
observables.
Pinned by closed-form cases (two-level system, Gaussian energies, single
histogram) in tests/test_oracle_reweight.py.
"""
import numpy as np


def wham(energy_lists, betas, observables=(), beta_targets=(), n_iter=100, tol=1e-8):
    K = len(betas)
    N_k = np.array([len(e) for e in energy_lists])
    E = np.concatenate(energy_lists)
    f = np.zeros(K)
    it = 0
    log_denom = np.logaddexp.reduce(np.log(N_k)[None, :] + f[None, :] - np.outer(E, betas), axis=1)
    for it in range(1, n_iter + 1):
        terms = np.log(N_k)[None, :] + f[None, :] - np.outer(E, betas)
        log_denom = np.logaddexp.reduce(terms, axis=1)
        f_new = np.array([-np.logaddexp.reduce(-betas[k] * E - log_denom) for k in range(K)])
        f_new -= f_new[0]
        done = np.max(np.abs(f_new - f)) < tol
        f = f_new
        if done:
            break
    # the weights use the last iteration's denominator (from f before its update),
    # as reweighting.py:101 does
    obs = [np.concatenate(o) for o in observables]
    out = np.zeros((len(beta_targets), len(obs)))
    for t, bt in enumerate(beta_targets):
        lw = -bt * E - log_denom
        w = np.exp(lw - lw.max())
        w /= w.sum()
        for j, o in enumerate(obs):
            out[t, j] = (w * o).sum()
    return f, it, out
