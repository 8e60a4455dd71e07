"""pr_react_box_noalloc (no device allocation: the library's static scratch) against
pr_react_box, bit for bit, on two calls in a row through the C-ABI - a drm19 box and then a
smaller one of other cells - plus its size limit and refusals."""
import numpy as np
import pytest

from pelelmex_amd import workloads as W

pytestmark = pytest.mark.gpu

KEYS = {"ode.rtol": 1e-10, "ode.atol": 1e-10, "cvode.max_order": 4}


def _run(r, box, gpu, noalloc):
    import torch
    from pelelmex_amd.reactor import box_tensors
    t = box_tensors(box, gpu)
    # the previous advance's cost, as PeleLMeX passes it, so the cost order has work to do
    t["FC"].copy_(torch.as_tensor(np.arange(box["T"].size) % 97, dtype=torch.float64).reshape(t["FC"].shape))
    status = torch.full((box["T"].size,), 7, dtype=torch.int32, device=gpu)
    nfail, _, tm = r.react(t["rY"], t["rYsrc"], t["T"], t["rH"], t["rHsrc"], t["FC"], t["mask"], box["dt"], 0.0,
                           status=status, noalloc=noalloc)
    torch.cuda.synchronize()
    out = {k: t[k].cpu().numpy() for k in ("rY", "T", "rH", "FC")}
    out["status"] = status.cpu().numpy()
    out["nfail"] = nfail
    out["stats"] = r.stats()
    return out


def _same(a, b):
    for k in ("rY", "T", "rH", "FC", "status"):
        assert np.array_equal(a[k], b[k]), f"{k} differs between pr_react_box_noalloc and pr_react_box"
    assert a["nfail"] == b["nfail"]
    for k in ("cells", "nst", "nfe", "nje", "nsetups", "nni"):
        assert a["stats"][k] == b["stats"][k], k


def test_noalloc_matches_react_box_twice(gpu):
    from pelelmex_amd.reactor import Reactor
    full = W.ch4_air_flame_box(nx=32, ny=32, nz=32, seed=2002)
    big = full
    small = W.ch4_air_flame_box(nx=16, ny=8, nz=4, seed=77)
    small["mask"] = np.where(np.arange(small["T"].size) % 11 == 3, -1, 1).astype(np.int32)
    r = Reactor.create("ReactorCvode", "drm19", KEYS)
    r.init(2, 1)
    ref_big = _run(r, big, gpu, noalloc=False)
    ref_small = _run(r, small, gpu, noalloc=False)
    got_big = _run(r, big, gpu, noalloc=True)
    got_small = _run(r, small, gpu, noalloc=True)     # second call, smaller box: stale scratch must not leak in
    r.close()
    print(f"noalloc: {big['T'].size} cells then {small['T'].size}: blocks {got_big['stats']['blocks']} / "
          f"{got_small['stats']['blocks']}")
    _same(got_big, ref_big)
    _same(got_small, ref_small)
    assert np.all(got_small["FC"].reshape(-1)[small["mask"] == -1] == 0)


def test_noalloc_refusals(gpu):
    import torch
    from pelelmex_amd.reactor import Reactor, ReactorError, box_tensors
    box = W.ch4_air_flame_box(nx=4, ny=4, nz=4, seed=5)
    r = Reactor.create("ReactorCvode", "drm19", dict(KEYS, **{"hipbdf.autotune": 1}))
    r.init(2, 1)
    t = box_tensors(box, gpu)
    with pytest.raises(ReactorError, match="autotune"):
        r.react(t["rY"], t["rYsrc"], t["T"], t["rH"], t["rHsrc"], t["FC"], t["mask"], box["dt"], noalloc=True)
    r.close()
    # a box above the cost-order limit is refused before any launch (the arrays are never read)
    r = Reactor.create("ReactorCvode", "drm19", KEYS)
    r.init(2, 1)
    nz, ny, nx = 1, 2048, 2049
    rY = torch.empty((r.nsp, nz, ny, nx), dtype=torch.float64, device=gpu)
    T = torch.empty((nz, ny, nx), dtype=torch.float64, device=gpu)
    with pytest.raises(ReactorError, match="no-allocation limit"):
        r.react(rY, rY, T, T, T, T, None, 1e-6, noalloc=True)
    r.close()
