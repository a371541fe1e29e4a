"""Pins the oracle's Philox4x32-10, its deterministic exp and its uniforms."""
import math
import os

import numpy as np
import pytest

import oracle_lib as o

HERE = os.path.dirname(os.path.abspath(__file__))


def _kat():
    rows = []
    for line in open(os.path.join(HERE, "golden", "philox_kat.txt")):
        if line.startswith("#") or not line.strip():
            continue
        v = [int(w, 16) for w in line.split()]
        rows.append((v[:4], v[4:6], v[6:10]))
    return rows


@pytest.mark.parametrize("ctr,key,expect", _kat())
def test_philox_known_answers(ctr, key, expect):
    assert list(o.philox(ctr, key)) == expect


def test_philox_matches_rocrand_engine():
    rng = np.random.default_rng(7)
    rows = [(int(rng.integers(2 ** 32)), int(rng.integers(2 ** 30)), int(rng.integers(2 ** 32)),
             int(rng.integers(2 ** 32)), int(rng.integers(2 ** 32)), int(rng.integers(2 ** 32))) for _ in range(64)]
    ref = o.rocrand_philox(rows)
    if ref is None:
        pytest.skip("rocRAND check binary not built")
    for i, r in enumerate(rows):
        assert (o.philox(r[:4], r[4:]) == ref[i]).all(), r


def test_det_exp_against_libm():
    xs = np.concatenate([np.linspace(-700.0, 0.0, 20001), -np.logspace(-12, 2, 2000), [0.0, -1e-300]])
    for x in xs:
        ref = math.exp(x)
        got = o.lib.o_det_exp(float(x))
        assert abs(got - ref) <= 4.5e-16 * ref, x
    assert o.lib.o_det_exp(-800.0) == 0.0


def test_u01_range_and_resolution():
    assert o.lib.o_u01(0, 0) == 0.0
    top = o.lib.o_u01(0xFFFFFFFF, 0xFFFFFFFF)
    assert top < 1.0 and top == 1.0 - 2.0 ** -53
    assert o.lib.o_u01(0, 1 << 6) == 2.0 ** -53


def test_philox_uniformity_of_7bit_draws():
    # the 7-bit draws used for Ising acceptance (low 7 bits of each byte) are
    # uniform over one sweep's counters
    counts = np.zeros(16, np.int64)
    for g in range(4096):
        w = o.philox([g, 3, 7, 0x200], [42, 0])
        for word in w:
            for j in range(4):
                counts[((int(word) >> (8 * j)) & 0x7F) >> 3] += 1
    exp = counts.sum() / 16
    chi2 = float(((counts - exp) ** 2 / exp).sum())
    assert chi2 < 45.0  # 15 dof, p ~ 1e-4
