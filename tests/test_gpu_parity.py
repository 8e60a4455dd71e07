"""GPU reactor (libpelelmex_amd.so, through its C-ABI) against the CPU oracle.

Tolerances are those the reactor is specified to meet against the CPU integrator at
rtol = atol = 1e-10: max |dT|/T <= 1e-6 and max |dY_k| <= 1e-8 (Y = rhoY / rho). rho h is
checked to round-off (it is rho h0 + dt Q_h on both sides), FC (RHS evaluations) must be
positive on active cells and zero on masked cells.
"""
import numpy as np
import pytest

from pelelmex_amd import workloads as W

pytestmark = pytest.mark.gpu

T_TOL = 1e-6
Y_TOL = 1e-8


def _oracle(box, **kw):
    from oracle import pyoracle as po
    m = po.Mech(box["mech"])
    return m.react(box["rY"], box["rYsrc"], box["T"], box["rH"], box["rHsrc"], box["mask"], box["dt"],
                   nthreads=8, **kw)


def _gpu(box, keys, gpu):
    import torch
    from pelelmex_amd.reactor import Reactor, box_tensors
    r = Reactor.create("ReactorCvode", box["mech"], keys)
    r.init(2, 1)
    t = box_tensors(box, gpu)
    status = torch.zeros(box["T"].size, dtype=torch.int32, device=gpu)
    nfail, _, tm = r.react(t["rY"], t["rYsrc"], t["T"], t["rH"], t["rHsrc"], t["FC"], t["mask"], box["dt"],
                           0.0, status=status)
    torch.cuda.synchronize()
    n = box["T"].size
    out = dict(rY=t["rY"].reshape(-1, n).cpu().numpy(), T=t["T"].reshape(n).cpu().numpy(),
               rH=t["rH"].reshape(n).cpu().numpy(), FC=t["FC"].reshape(n).cpu().numpy(),
               status=status.cpu().numpy(), nfail=nfail, time=tm, stats=r.stats())
    r.close()
    return out


def _compare(a, b, mask):
    act = mask != -1
    dT = np.abs(a["T"] - b["T"]) / b["T"]
    Ya = a["rY"] / a["rY"].sum(0)
    Yb = b["rY"] / b["rY"].sum(0)
    dY = np.abs(Ya - Yb)
    return float(dT[act].max()), float(dY[:, act].max())


def _sample(box, n, seed=0):
    rng = np.random.default_rng(seed)
    idx = np.sort(rng.choice(box["T"].size, n, replace=False))
    return W.subsample(box, idx)


KEYS = {"ode.rtol": 1e-10, "ode.atol": 1e-10, "cvode.max_order": 5, "cvode.max_substeps": 10000}


@pytest.mark.parametrize("lanes", [16, 8])
def test_h2_air_parity(gpu, lanes):
    box = _sample(W.h2_air_box(), 384, seed=1)
    ref = _oracle(box, max_order=5)
    out = _gpu(box, dict(KEYS, **{"hipbdf.lanes_per_cell": lanes}), gpu)
    assert out["nfail"] == 0 and np.all(out["status"] == 0)
    dT, dY = _compare(out, ref, box["mask"])
    print(f"LiDryer G={lanes}: max dT/T {dT:.2e} max dY {dY:.2e} stats {out['stats']}")
    assert dT <= T_TOL and dY <= Y_TOL
    np.testing.assert_allclose(out["rH"], box["rH"] + box["dt"] * box["rHsrc"], rtol=1e-15, atol=0)
    assert np.all(out["FC"] > 0)


@pytest.mark.parametrize("lanes", [32, 16, 64])
def test_ch4_flame_parity(gpu, lanes):
    full = W.ch4_air_flame_box()
    # every slab type: one full i-row plus a random sample
    idx = np.union1d(np.arange(64), np.random.default_rng(2).choice(full["T"].size, 448, replace=False))
    box = W.subsample(full, idx)
    ref = _oracle(box, max_order=5)
    out = _gpu(box, dict(KEYS, **{"hipbdf.lanes_per_cell": lanes}), gpu)
    assert out["nfail"] == 0 and np.all(out["status"] == 0)
    dT, dY = _compare(out, ref, box["mask"])
    print(f"drm19 G={lanes}: max dT/T {dT:.2e} max dY {dY:.2e} stats {out['stats']}")
    assert dT <= T_TOL and dY <= Y_TOL
    np.testing.assert_allclose(out["rH"], box["rH"] + box["dt"] * box["rHsrc"], rtol=1e-15, atol=0)


def test_mask_skips_cells(gpu):
    full = W.ch4_air_flame_box(nx=16, ny=4, nz=4)
    box = dict(full)
    mask = np.ones(full["T"].size, dtype=np.int32)
    mask[::3] = -1
    box["mask"] = mask
    out = _gpu(box, KEYS, gpu)
    assert out["nfail"] == 0
    cov = mask == -1
    np.testing.assert_array_equal(out["T"][cov], full["T"][cov])
    np.testing.assert_array_equal(out["rY"][:, cov], full["rY"][:, cov])
    assert np.all(out["FC"][cov] == 0) and np.all(out["FC"][~cov] > 0)
    ref = _oracle(box, max_order=5)
    dT, dY = _compare(out, ref, mask)
    assert dT <= T_TOL and dY <= Y_TOL
