"""Fine-grid observables of the reference's analysis (analyze_gpu_fss.py:60-253)
by single-histogram reweighting, against closed forms: no reweighting gives
the plain means and the block standard error; Gaussian energies shift by
-delta_beta sigma^2 and keep C_v; a constant |m| gives U = 2/3.  Host code:
no GPU needed (the WHAM method is checked in the GPU tests)."""
import numpy as np
import pytest

import ising_hip


def test_no_reweighting_gives_plain_means_and_block_errors():
    rng = np.random.default_rng(5)
    betas = np.array([0.20, 0.22, 0.24])
    e_l = [rng.normal(-1000.0 * (k + 1), 30.0, 10000) for k in range(3)]
    m_l = [np.abs(rng.normal(0.3 + 0.1 * k, 0.05, 10000)) for k in range(3)]
    out = ising_hip.reweight_fine(e_l, m_l, betas, 1.0 / betas, n=4, method="single", n_blocks=20)
    for k in range(3):
        assert out["M"][k] == pytest.approx(m_l[k].mean(), rel=1e-12)
        assert out["M2"][k] == pytest.approx((m_l[k] ** 2).mean(), rel=1e-12)
        assert out["E"][k] == pytest.approx(e_l[k].mean() / 64.0, rel=1e-12)
        bm = m_l[k].reshape(20, 500).mean(axis=1)  # leave-one-block-out error of a mean = block standard error
        assert out["M_err"][k] == pytest.approx(np.sqrt(((bm - bm.mean()) ** 2).sum() / (20 * 19)), rel=1e-9)
        var_e = (e_l[k] ** 2).mean() - e_l[k].mean() ** 2
        assert out["Cv"][k] == pytest.approx(betas[k] ** 2 * var_e / 64.0, rel=1e-9)


def test_gaussian_energies_shift_by_minus_delta_beta_sigma_squared():
    rng = np.random.default_rng(11)
    mu, sigma, beta0, dbeta, N = -5000.0, 100.0, 0.5, 1e-3, 2_000_000
    e = rng.normal(mu, sigma, N)
    m = np.full(N, 0.4)
    out = ising_hip.reweight_fine([e], [m], [beta0], [1.0 / (beta0 + dbeta)], n=1, method="single")
    assert out["E"][0] == pytest.approx(mu - dbeta * sigma ** 2, abs=6 * sigma / np.sqrt(N))
    assert out["Cv"][0] == pytest.approx((beta0 + dbeta) ** 2 * sigma ** 2, rel=0.01)
    assert out["U"][0] == pytest.approx(2.0 / 3.0, abs=1e-12)  # |m| constant: <m^4> = <m^2>^2
    assert out["M_err"][0] == pytest.approx(0.0, abs=1e-12)


def test_series_by_temperature_and_errors():
    tidx = np.array([[0, 1], [1, 0], [0, 1]])
    bond = np.array([[10, -4], [6, 8], [12, -2]])
    mag = np.array([[-8, 4], [2, 6], [8, -4]])
    e_l, m_l = ising_hip.series_by_temperature(tidx, bond, mag, L=2, J=1.0)
    assert np.array_equal(e_l[0], [-10.0, -8.0, -12.0]) and np.array_equal(e_l[1], [4.0, -6.0, 2.0])
    assert np.array_equal(m_l[0], [1.0, 0.75, 1.0]) and np.array_equal(m_l[1], [0.5, 0.25, 0.5])
    with pytest.raises(ValueError):
        ising_hip.reweight_fine(e_l, m_l, [1.0, 0.9], [1.0], n=2, method="bogus")
