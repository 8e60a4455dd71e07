"""Pins of the CPU network oracle (oracle/mlp_oracle.py) that checks the GPU training kernels:
autograd gradients against central finite differences, bf16 rounding against torch's own
float32 -> bfloat16 conversion, and Adam against its closed form after one step."""
import numpy as np
import pytest

from oracle import mlp_oracle as mo


def _net(rng, dims):
    W = [rng.normal(0, 1 / np.sqrt(i), (i, o)) for i, o in zip(dims[:-1], dims[1:])]
    b = [rng.normal(0, 0.1, o) for o in dims[1:]]
    return W, b


def test_gradients_against_finite_differences():
    rng = np.random.default_rng(0)
    dims = [5, 7, 6, 3]
    W, b = _net(rng, dims)
    X = rng.normal(size=(11, 5))
    T = rng.normal(size=(11, 3))
    loss, gW, gb = mo.loss_and_grads(W, b, X, T)
    h = 1e-6
    for l in range(len(W)):
        for (i, j) in [(0, 0), (W[l].shape[0] - 1, W[l].shape[1] - 1), (1, 2)]:
            Wp = [w.copy() for w in W]
            Wm = [w.copy() for w in W]
            Wp[l][i, j] += h
            Wm[l][i, j] -= h
            fd = (mo.loss_and_grads(Wp, b, X, T)[0] - mo.loss_and_grads(Wm, b, X, T)[0]) / (2 * h)
            assert abs(fd - gW[l][i, j]) <= 1e-7 * max(1.0, abs(fd))
        bp = [x.copy() for x in b]
        bm = [x.copy() for x in b]
        bp[l][0] += h
        bm[l][0] -= h
        fd = (mo.loss_and_grads(W, bp, X, T)[0] - mo.loss_and_grads(W, bm, X, T)[0]) / (2 * h)
        assert abs(fd - gb[l][0]) <= 1e-7 * max(1.0, abs(fd))
    # the loss is the mean squared error of the forward pass
    assert abs(loss - np.mean((mo.forward(W, b, X) - T) ** 2)) <= 1e-14


def test_bf16_round_matches_torch():
    import torch
    rng = np.random.default_rng(1)
    a = np.concatenate([rng.normal(size=1000) * 10.0 ** rng.integers(-8, 8, 1000), [0.0, -0.0, 1.0, 3.0 + 2 ** -8]])
    a = a.astype(np.float32)
    ref = torch.as_tensor(a).to(torch.bfloat16).to(torch.float64).numpy()
    np.testing.assert_array_equal(mo.bf16_round(a), ref)


def test_adam_first_step_is_signed_lr():
    g = np.array([0.3, -2.0, 1e-3])
    p, m, v = mo.adam(np.zeros(3), g, np.zeros(3), np.zeros(3), 1, 1e-3, eps=0.0)
    np.testing.assert_allclose(p, -1e-3 * np.sign(g), rtol=1e-12)
    np.testing.assert_allclose(m, 0.1 * g)
