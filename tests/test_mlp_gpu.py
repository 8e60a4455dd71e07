"""GPU tests of the surrogate network kernels (pr_mlp_*) against the CPU FP64 oracle.

* the BF16 matrix-core product in all four operand layouts, ragged sizes, bias and GELU
  epilogue, against float64 products of the same bf16 values;
* forward, loss and backward of a 23-64-64-22 network against torch CPU float64 autograd
  (bf16 rounding of weights, inputs and activations bounds the agreement);
* Adam against its float64 closed form.
"""
import numpy as np
import pytest

from oracle import mlp_oracle as mo

pytestmark = pytest.mark.gpu


def _bf16(t):
    import torch
    return torch.as_tensor(np.asarray(t, np.float32)).to(torch.bfloat16)


@pytest.mark.parametrize("a_kfast,b_kfast", [(True, False), (True, True), (False, False), (False, True)])
@pytest.mark.parametrize("M,N,K", [(77, 45, 50), (128, 64, 256), (5, 130, 23)])
def test_gemm_layouts(gpu, a_kfast, b_kfast, M, N, K):
    import torch
    from pelelmex_amd import surrogate as sg
    rng = np.random.default_rng(M * 1000 + N + K)
    A = rng.normal(size=(M, K)).astype(np.float32)
    B = rng.normal(size=(K, N)).astype(np.float32)
    bias = rng.normal(size=N).astype(np.float32)
    Ab, Bb = mo.bf16_round(A), mo.bf16_round(B)
    ref = Ab @ Bb + bias
    # store A as [M, K] (k fastest) or [K, M] (m fastest); B as [K, N] or [N, K]
    Ad = (_bf16(A) if a_kfast else _bf16(A.T.copy())).to(gpu).contiguous()
    Bd = (_bf16(B.T.copy()) if b_kfast else _bf16(B)).to(gpu).contiguous()
    sam, sak = (K, 1) if a_kfast else (1, M)
    sbk, sbn = (1, K) if b_kfast else (N, 1)
    C = torch.full((M, N), float("nan"), dtype=torch.float32, device=gpu)
    Z = torch.full((M, N), float("nan"), dtype=torch.float32, device=gpu)
    Cb = torch.zeros((M, N), dtype=torch.bfloat16, device=gpu)
    bd = torch.as_tensor(bias).to(gpu)
    sg.gemm(M, N, K, Ad, sam, sak, Bd, sbk, sbn, bias=bd, gelu=True, Cf=C, ldc=N, Cb=Cb, ldcb=N, Z=Z, ldz=N)
    torch.cuda.synchronize()
    scale = np.abs(Ab) @ np.abs(Bb) + np.abs(bias)
    z = Z.cpu().numpy()
    assert np.all(np.abs(z - ref) <= 2e-6 * scale + 1e-6)
    import torch.nn.functional as F
    g = F.gelu(torch.as_tensor(ref)).numpy()
    assert np.all(np.abs(C.cpu().numpy() - g) <= 2e-6 * scale + 1e-5)
    assert np.all(np.abs(Cb.float().cpu().numpy() - g) <= 1e-2 * np.abs(g) + 1e-5 * scale + 1e-5)
    # accumulate mode adds the raw product
    C2 = torch.ones((M, N), dtype=torch.float32, device=gpu)
    sg.gemm(M, N, K, Ad, sam, sak, Bd, sbk, sbn, Cf=C2, ldc=N, accumulate=True)
    torch.cuda.synchronize()
    assert np.all(np.abs(C2.cpu().numpy() - (1.0 + Ab @ Bb)) <= 2e-6 * scale + 1e-6)


def test_forward_backward_against_autograd(gpu):
    import torch
    from pelelmex_amd import surrogate as sg
    dims = [23, 64, 64, 22]
    mlp = sg.MLP(dims, device=gpu, seed=3)
    rng = np.random.default_rng(4)
    for b in mlp.b:
        b.copy_(torch.as_tensor(rng.normal(0, 0.1, b.numel()).astype(np.float32)))
    B = 300
    X = rng.normal(size=(B, 23)).astype(np.float32)
    T = rng.normal(size=(B, 22)).astype(np.float32)
    Xd = _bf16(X).to(gpu).contiguous()
    Td = torch.as_tensor(T).to(gpu)
    out = mlp.forward(Xd, keep=True)
    loss, dY = mlp.mse(out, Td)
    mlp.backward(dY)
    torch.cuda.synchronize()
    Wr = [mo.bf16_round(w.cpu().numpy()) for w in mlp.W]
    br = [b.cpu().numpy().astype(np.float64) for b in mlp.b]
    Xr = mo.bf16_round(X)
    ref_out = mo.forward(Wr, br, Xr, bf16_activations=True)
    np.testing.assert_allclose(out.cpu().numpy(), ref_out, rtol=0, atol=1e-3 * np.abs(ref_out).max())
    l_ref, gW_ref, gb_ref = mo.loss_and_grads(Wr, br, Xr, T)
    assert abs(float(loss.item()) - l_ref) <= 1e-2 * l_ref
    for l in range(len(dims) - 1):
        for got, ref in ((mlp.gW[l].cpu().numpy(), gW_ref[l]), (mlp.gb[l].cpu().numpy(), gb_ref[l])):
            rel = np.linalg.norm(got - ref) / np.linalg.norm(ref)
            cos = np.sum(got * ref) / (np.linalg.norm(got) * np.linalg.norm(ref))
            assert rel <= 3e-2 and cos >= 0.999, (l, rel, cos)


def test_adam_against_closed_form(gpu):
    import torch
    from pelelmex_amd import surrogate as sg
    rng = np.random.default_rng(5)
    dims = [23, 40, 22]
    mlp = sg.MLP(dims, device=gpu, seed=1)
    p0 = [w.cpu().numpy().astype(np.float64) for w in mlp.W + mlp.b]
    grads = [rng.normal(size=w.shape).astype(np.float32) for w in mlp.W + mlp.b]
    for g, t in zip(grads, mlp.gW + mlp.gb):
        t.copy_(torch.as_tensor(g))
    mlp.adam(lr=1e-2)
    for g, t in zip(grads, mlp.gW + mlp.gb):
        t.copy_(torch.as_tensor(-0.5 * g))
    mlp.adam(lr=1e-2)
    torch.cuda.synchronize()
    for k, (p, g) in enumerate(zip(p0, grads)):
        m = np.zeros_like(p)
        v = np.zeros_like(p)
        p1, m, v = mo.adam(p, g.astype(np.float64), m, v, 1, 1e-2)
        p2, m, v = mo.adam(p1, -0.5 * g.astype(np.float64), m, v, 2, 1e-2)
        got = (mlp.W + mlp.b)[k].cpu().numpy()
        np.testing.assert_allclose(got, p2, rtol=1e-5, atol=1e-6)
    # the bf16 copies follow the fp32 parameters
    for w, wb in zip(mlp.W, mlp.Wb):
        np.testing.assert_array_equal(wb.float().cpu().numpy(), mo.bf16_round(w.cpu().numpy()).astype(np.float32))


def test_hidden_layer_forward_kernel_bits(gpu):
    """pr_surrogate_forward (the learned path's forward: the tiled 128 x 128 kernel with
    transposed weights for the hidden layers) gives the bits of the same network run layer by
    layer through pr_mlp_gemm, on a row count that is not a multiple of the tile."""
    import torch
    from pelelmex_amd import surrogate as sg
    from pelelmex_amd.reactor import Reactor
    model = sg.load(sg.WEIGHTS)
    dims = [int(d) for d in model["dims"]]
    mlp = sg.MLP(dims, device=gpu)
    mlp.set_flat_params(torch.as_tensor(model["params"], dtype=torch.float32))
    r = Reactor.create("ReactorCvode", "drm19", {"ode.rtol": 1e-10, "ode.atol": 1e-10})
    r.init(2, 1)
    sg.attach(r, model)
    g = torch.Generator(device="cpu").manual_seed(3)
    X = (torch.randn((1000, dims[0]), generator=g) * 1.5).to(torch.bfloat16).to(gpu)
    Y1 = r.surrogate_forward(X)
    Y2 = mlp.forward(X)
    torch.cuda.synchronize()
    assert torch.isfinite(Y1).all()
    assert torch.equal(Y1, Y2)
    r.close()
