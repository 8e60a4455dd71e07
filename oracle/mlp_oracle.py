"""CPU FP64 oracle of the surrogate network's forward pass, loss and gradients.

TEST INFRASTRUCTURE ONLY (imported by tests/): the checker for the library's BF16 matrix-core
training kernels. Network: dense layers with exact GELU (0.5 z (1 + erf(z / sqrt 2))) between
them, mean-squared-error loss; gradients by torch autograd in float64 on the CPU, and by central
finite differences for spot checks.
"""
from __future__ import annotations

import numpy as np


def bf16_round(a):
    """Rounds float32 values to the nearest bfloat16 (ties to even), returned as float64."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def forward(W, b, X, bf16_activations=False):
    """Outputs of the network (float64); W[l] [in x out], b[l] [out], X [rows x in]."""
    import torch
    a = torch.as_tensor(np.asarray(X, np.float64))
    for l, (w, bb) in enumerate(zip(W, b)):
        z = a @ torch.as_tensor(np.asarray(w, np.float64)) + torch.as_tensor(np.asarray(bb, np.float64))
        if l < len(W) - 1:
            a = torch.nn.functional.gelu(z)
            if bf16_activations:
                a = torch.as_tensor(bf16_round(a.numpy().astype(np.float32)))
        else:
            a = z
    return a.numpy()


def loss_and_grads(W, b, X, T):
    """MSE loss (mean over all outputs) and its gradients wrt every W and b (float64)."""
    import torch
    Wt = [torch.tensor(np.asarray(w, np.float64), requires_grad=True) for w in W]
    bt = [torch.tensor(np.asarray(bb, np.float64), requires_grad=True) for bb in b]
    a = torch.as_tensor(np.asarray(X, np.float64))
    for l in range(len(Wt)):
        z = a @ Wt[l] + bt[l]
        a = torch.nn.functional.gelu(z) if l < len(Wt) - 1 else z
    loss = ((a - torch.as_tensor(np.asarray(T, np.float64))) ** 2).mean()
    loss.backward()
    return float(loss), [w.grad.numpy() for w in Wt], [bb.grad.numpy() for bb in bt]


def adam(p, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (Kingma & Ba 2015) in float64; returns new (p, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v
