"""Pins of the FP64 closure oracle (oracle/mlp.py): the analytic backward
pass against central finite differences of the loss, the forward pass
against a layer-by-layer evaluation written out here, and the feature
definitions on a hand-built chain whose values are known in closed form."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import mlp  # noqa: E402


def test_param_count():
    assert mlp.N_PARAMS == 36998


def test_forward_matches_layerwise_evaluation():
    rng = np.random.default_rng(0)
    p = mlp.init_params(3)
    x = rng.normal(size=(7, 24))
    out, _ = mlp.forward(p, x)
    (W0, b0), (W1, b1), (W2, b2), (W3, b3) = mlp.unpack(p)
    h = x
    for W, b in ((W0, b0), (W1, b1), (W2, b2)):
        z = np.einsum("oi,mi->mo", W, h) + b
        h = z * (1.0 / (1.0 + np.exp(-z)))
    ref = np.einsum("oi,mi->mo", W3, h) + b3
    assert np.allclose(out, ref, rtol=1e-12, atol=1e-12)


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(1)
    p = mlp.init_params(5) * 0.5
    x = rng.normal(size=(16, 24))
    y = rng.normal(size=(16, 3)) * 2.0
    sdy = np.array([1.5, 0.7, 2.0])
    L, g = mlp.loss_and_grad(p, x, y, sdy)
    idx = np.concatenate([rng.choice(mlp.N_PARAMS, 60, replace=False),
                          [0, 3071, 3072 + 5, 36998 - 1, 36998 - 7, 36998 - 6 - 1]])
    eps = 1e-6
    for k in idx:
        pp, pm = p.copy(), p.copy()
        pp[k] += eps
        pm[k] -= eps
        fd = (mlp.loss_and_grad(pp, x, y, sdy)[0] - mlp.loss_and_grad(pm, x, y, sdy)[0]) / (2 * eps)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, g[k])


def test_loss_is_gaussian_nll():
    """with all weights zero except the output biases, the loss is the
    closed-form NLL of N(b_mu, exp(b_s)) averaged over the rows"""
    p = np.zeros(mlp.N_PARAMS)
    p[-6:] = [0.3, -0.2, 0.1, 0.5, -1.0, 0.0]
    x = np.random.default_rng(2).normal(size=(5, 24))
    y = np.random.default_rng(3).normal(size=(5, 3))
    L, _ = mlp.loss_and_grad(p, x, y)
    mu, s = np.array([0.3, -0.2, 0.1]), np.array([0.5, -1.0, 0.0])
    ref = np.mean(np.sum(0.5 * ((y - mu) ** 2 * np.exp(-s) + s), axis=1))
    assert abs(L - ref) < 1e-12


def test_features_on_a_known_chain():
    # core 0 with three arms of one bead each, plus a second bead on arm 1
    pos = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [2, 0, 0]], float)
    vel = np.array([[0, 0, 0], [1, 0, 0], [0, 0, 0], [0, 0, 0], [0, 2, 0]], float)
    bonds = [(0, 1), (0, 2), (0, 3), (1, 4)]
    f = mlp.features(pos, vel, np.full(5, 10.0), np.array([0, 0, 0, 0, 1]), bonds, shear=0.1)
    com = pos.mean(0)
    assert np.allclose(f[:, 0:3], pos - com)
    assert f[0, 17] == 1 and f[1:, 17].sum() == 0           # core flag from degree
    assert np.allclose(f[0, 9:15], 0)                        # the core's bond vectors are zero
    assert np.allclose(f[1, 9:12], [-1, 0, 0]) and np.allclose(f[1, 12:15], [1, 0, 0])
    assert np.allclose(f[4, 9:12], [-1, 0, 0]) and np.allclose(f[4, 12:15], [0, 0, 0])
    assert f[0, 15] == 3 and f[4, 15] == 1                   # neighbours within 1.5
    assert np.allclose(f[:, 6], vel[:, 0] - 0.1 * pos[:, 1])  # velocity relative to the flow
    assert np.allclose(f[0, 19:22], [1 / 3, 0, 0])           # neighbours 1, 2, 3 of the core
    assert np.allclose(f[4, 19:22], [1, -2, 0])              # its one neighbour is bead 1
    assert np.allclose(f[:, 22], 0.1) and np.allclose(f[:, 23], [3, 2, 1, 1, 1]) and np.allclose(f[:, 18], 0.1)
