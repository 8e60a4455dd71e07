"""The reactor at the n-dodecane size class: syn53, a synthetic 53-species / 268-reaction
mechanism (54-component state, 54 x 54 Jacobian and LU in the cell's global slot at 64 lanes),
on 60 atm states with 10 % of cells masked, against the CPU oracle bit for bit - for both builds
of the kernel (2 and 3 waves per SIMD)."""
import numpy as np
import pytest

from pelelmex_amd import workloads as W

from test_gpu_parity import KEYS, T_TOL, Y_TOL, _assert_bits, _compare, _gpu, _oracle, _sample

pytestmark = pytest.mark.gpu


def test_syn53_rhs_and_jacobian_bits(gpu):
    import torch
    from oracle import pyoracle as po
    from pelelmex_amd.reactor import Reactor
    box = _sample(W.syn53_chunk(0, nx=16, ny=16, nz=16), 96, seed=4)
    m = po.Mech("syn53")
    n = m.nsp + 1
    nc = box["T"].size
    y = np.vstack([box["rY"], box["T"][None, :]])
    t_eval = 0.41 * box["dt"]
    fref = np.zeros((n, nc))
    Jref = np.zeros((n, n, nc))
    for c in range(nc):
        cell = m.cell(box["rH"][c], box["rHsrc"][c], box["rYsrc"][:, c])
        fref[:, c] = m.rhs(cell, t_eval, y[:, c])
        Jref[:, :, c] = m.jac(cell, 0.0, y[:, c])
    r = Reactor.create("ReactorCvode", "syn53", KEYS)
    r.init(2, 1)
    dev = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64).to(gpu)
    f = r.rhs_batch(dev(y), dev(box["rYsrc"]), dev(box["rH"]), dev(box["rHsrc"]), t_eval).cpu().numpy()
    J = r.jac_batch(dev(y), dev(box["rYsrc"]), dev(box["rHsrc"])).cpu().numpy()
    r.close()
    _assert_bits("rhs", f, fref)
    _assert_bits("jac", J, Jref)


@pytest.mark.parametrize("waves", [2, 3])
def test_syn53_react_parity(gpu, waves):
    box = _sample(W.syn53_chunk(0, nx=32, ny=32, nz=32), 320, seed=6)
    assert np.any(box["mask"] == -1)
    ref = _oracle(box, max_order=4)
    keys = dict(KEYS, **{"cvode.max_order": 4, "hipbdf.waves_per_simd": waves})
    out = _gpu(box, keys, gpu)
    assert out["nfail"] == 0 and np.all(out["status"] == 0)
    assert out["stats"]["waves_per_simd"] == waves and out["stats"]["lanes_per_cell"] == 64
    dT, dY = _compare(out, ref, box["mask"])
    print(f"syn53 W={waves}: max dT/T {dT:.2e} max dY {dY:.2e} stats {out['stats']}")
    assert dT <= T_TOL and dY <= Y_TOL
    _assert_bits("rY", out["rY"], ref["rY"])
    _assert_bits("T", out["T"], ref["T"])
    _assert_bits("rH", out["rH"], ref["rH"])
    _assert_bits("FC", out["FC"], ref["FC"])
    cov = box["mask"] == -1
    assert np.all(out["FC"][cov] == 0) and np.all(out["FC"][~cov] > 0)
