"""Pins the MLP oracle (oracle/mlp_oracle.py): backprop against central
finite differences of its own loss, and the bf16 rounding against hand
values."""
import numpy as np

import mlp_oracle as mo


def test_bf16_rounding_known_values():
    x = np.array([1.0, 1.00390625, 1.01171875, 1.0078125, -2.5, 3.0e-39, np.inf], np.float32)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: ties to even -> 1
    # 1 + 3*2^-8 lies between 1 + 2^-7 and 1 + 2^-6: tie -> even mantissa (1 + 2^-6)
    want = np.array([1.0, 1.0, 1.015625, 1.0078125, -2.5, np.float32(3.0e-39), np.inf], np.float32)
    got = mo.bf16_round(x)
    assert np.array_equal(got[:5], want[:5]) and np.isinf(got[6])
    # every output is exactly representable in bf16 (low 16 bits zero)
    assert (mo.bf16_round(np.random.default_rng(0).normal(size=1000).astype(np.float32)).view(np.uint32)
            & 0xFFFF == 0).all()


def test_backprop_matches_finite_differences():
    rng = np.random.default_rng(1)
    p = rng.normal(scale=0.3, size=mo.NP)
    x = rng.normal(size=(7, mo.NI))
    t = rng.normal(size=(7, mo.NO))
    L, g = mo.loss_grad(p, x, t, round_bf16=False)
    idx = rng.choice(mo.NP, 120, replace=False)
    eps = 1e-6
    for i in idx:
        d = np.zeros(mo.NP)
        d[i] = eps
        fd = (mo.loss_grad(p + d, x, t, False)[0] - mo.loss_grad(p - d, x, t, False)[0]) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_planted_mistake_in_backprop_is_caught():
    # a dropped ReLU mask in the second layer must break the FD check
    rng = np.random.default_rng(2)
    p = rng.normal(scale=0.5, size=mo.NP)
    x = rng.normal(size=(5, mo.NI))
    t = rng.normal(size=(5, mo.NO))
    W1, b1, W2, b2, W3, b3 = mo.unpack(p)
    y, (xx, z1, h1, z2, h2) = mo.forward(p, x, round_bf16=False)
    dy = (y - t) / len(x)
    wrong_dz2 = dy @ W3  # mask dropped
    right_dz2 = wrong_dz2 * (z2 > 0)
    assert not np.allclose(wrong_dz2.T @ h1, right_dz2.T @ h1)
