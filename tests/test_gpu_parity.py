"""GPU reactor (libpelelmex_amd.so, through its C-ABI) against the CPU oracle.

The GPU kernels and the oracle evaluate the same IEEE-754 expression trees (own exp / log on
both sides, no contraction, sums in component order), so the GPU results are required to be
bit-identical to the oracle's: right-hand sides, Jacobians, and whole react calls (rhoY, T,
rho h, FC). The specified tolerance (max |dT|/T <= 1e-6, max |dY_k| <= 1e-8 with Y = rhoY / rho)
is checked against integrators independent of the oracle: Fortran VODE run at the same
tolerances and the rtol-1e-13 Radau / VODE truth of the committed fixtures (tests/golden).
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


def _gpu(box, keys, gpu, typ=None):
    import torch
    from pelelmex_amd.reactor import Reactor, box_tensors
    r = Reactor.create("ReactorCvode", box["mech"], keys)
    r.init(2, 1)
    if typ is not None:
        r.set_typ_vals_ode(typ)
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


def _ulps(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    d = np.abs(a - b)
    sc = np.spacing(np.maximum(np.abs(a), np.abs(b)))
    return float(np.max(np.where(d == 0, 0.0, d / np.where(sc > 0, sc, 1e-300))))


def _assert_bits(name, got, ref):
    # NaN where both have NaN counts as equal: a cell whose integration failed can end in a
    # non-finite state (status < 0 says so), and the payload of a NaN carries no result
    if not np.array_equal(got, ref, equal_nan=True):
        raise AssertionError(f"{name}: not bit-identical to the oracle, max {_ulps(got, ref):.3g} ulp, "
                             f"{int(np.sum(got != ref))} of {got.size} entries differ")


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
    print(f"LiDryer G={lanes}: stats {out['stats']}")
    _assert_bits("rY", out["rY"], ref["rY"])
    _assert_bits("T", out["T"], ref["T"])
    _assert_bits("rH", out["rH"], ref["rH"])
    _assert_bits("FC", out["FC"], ref["FC"])
    assert np.all(out["FC"] > 0)


@pytest.mark.parametrize("lanes,waves", [(32, 0), (32, 3), (16, 0), (64, 0)])
def test_ch4_flame_parity(gpu, lanes, waves):
    full = W.ch4_air_flame_box()
    # every slab type: one full i-row plus a random sample
    idx = np.union1d(np.arange(64), np.random.default_rng(2).choice(full["T"].size, 448, replace=False))
    box = W.subsample(full, idx)
    ref = _oracle(box, max_order=5)
    out = _gpu(box, dict(KEYS, **{"hipbdf.lanes_per_cell": lanes, "hipbdf.waves_per_simd": waves}), gpu)
    assert out["nfail"] == 0 and np.all(out["status"] == 0)
    print(f"drm19 G={lanes} W={out['stats']['waves_per_simd']}: stats {out['stats']}")
    _assert_bits("rY", out["rY"], ref["rY"])
    _assert_bits("T", out["T"], ref["T"])
    _assert_bits("rH", out["rH"], ref["rH"])
    _assert_bits("FC", out["FC"], ref["FC"])


def test_ch4_flame_full_box_bits(gpu):
    """The whole 64^3 flame box (262144 cells, the benchmark's launch) at max_order 4 against the
    oracle on every cell, bit for bit, with the element moles of every cell moving only by the
    sources' (a linear invariant of the exact solution; 1e-12 of the cell's element total)."""
    from oracle import pyoracle as po
    from pelelmex_amd.mechanisms import load
    full = W.ch4_air_flame_box()
    out = _gpu(full, dict(KEYS, **{"cvode.max_order": 4}), gpu)
    assert out["nfail"] == 0 and np.all(out["status"] == 0)
    ref = po.Mech("drm19").react(full["rY"], full["rYsrc"], full["T"], full["rH"], full["rHsrc"], full["mask"],
                                 full["dt"], max_order=4, nthreads=16)
    assert np.all(ref["status"] == 0)
    for k in ("rY", "T", "rH", "FC"):
        _assert_bits(k, out[k], ref[k])
    mech = load("drm19")
    E = np.asarray(mech.elem_matrix, dtype=np.float64)
    Wk = np.asarray(mech.W, dtype=np.float64)[:, None]
    el0 = ((full["rY"] + full["dt"] * full["rYsrc"]) / Wk).T @ E
    el1 = (out["rY"] / Wk).T @ E
    drift = np.abs(el1 - el0) / np.abs(el0).sum(axis=1, keepdims=True)
    print(f"drm19 64^3: element drift max {drift.max():.2e}, kernel {out['stats']['kernel_ms']:.1f} ms")
    assert drift.max() <= 1e-12, drift.max()


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
    _assert_bits("rY", out["rY"], ref["rY"])
    _assert_bits("T", out["T"], ref["T"])


@pytest.mark.parametrize("lanes", [32, 64])
def test_single_cell_ragged_and_fully_masked_boxes(gpu, lanes):
    """Edge boxes of the persistent launch, bit for bit against the oracle: one cell (far fewer
    cells than resident groups), a ragged 3 x 5 x 7 box (an odd count: at 32 lanes the last
    wavefront's second cell slot finds the queue empty while the first integrates), and a fully
    masked box (no active cell; state unchanged, FC = 0, rho h still advanced by its source)."""
    full = W.ch4_air_flame_box(nx=16, ny=8, nz=4, seed=31)
    keys = dict(KEYS, **{"cvode.max_order": 4, "hipbdf.lanes_per_cell": lanes})
    one = W.subsample(full, np.array([295]))          # a front-band cell (i = 7, x = 0.47)
    rag = W.subsample(full, np.random.default_rng(4).choice(full["T"].size, 105, replace=False))
    rag["dims"] = (3, 5, 7)
    masked = dict(rag, mask=np.full(105, -1, dtype=np.int32))
    for name, box in (("one", one), ("ragged", rag), ("masked", masked)):
        out = _gpu(box, keys, gpu)
        ref = _oracle(box, max_order=4)
        assert out["nfail"] == 0 and np.all(out["status"] == 0), name
        for k in ("rY", "T", "rH", "FC"):
            _assert_bits(f"{name} {k}", out[k], ref[k])
        if name == "masked":
            np.testing.assert_array_equal(out["rY"], box["rY"])
            np.testing.assert_array_equal(out["T"], box["T"])
            assert np.all(out["FC"] == 0) and out["stats"]["cells"] == 0
        else:
            assert np.all(out["FC"] > 0) and out["stats"]["cells"] == box["T"].size


def test_argument_errors_and_empty_batch(gpu):
    """Host-side argument checks of pr_react_box / pr_react_batch (no launch happens): an empty box
    (hi < lo) and a dt_react that is zero, negative or not finite are errors with a message; an
    empty box list (a level with no box on this rank) integrates nothing and returns 0, and
    pr_collect after it reports 0 failed cells and empty statistics."""
    import ctypes as C
    import torch
    from pelelmex_amd import reactor as R
    from pelelmex_amd.reactor import Reactor, ReactorError, box_tensors
    box = W.ch4_air_flame_box(nx=4, ny=2, nz=2)
    r = Reactor.create("ReactorCvode", "drm19", KEYS)
    r.init(2, 1)
    t = box_tensors(box, gpu)
    for dt in (0.0, -1e-6, float("nan"), float("inf")):
        with pytest.raises(ReactorError, match="dt_react must be positive"):
            r.react(t["rY"], t["rYsrc"], t["T"], t["rH"], t["rHsrc"], t["FC"], t["mask"], dt)
    a = [R._array4(t["rY"], r.nsp), R._array4(t["rYsrc"], r.nsp)] + \
        [R._array4(t[k], 1) for k in ("T", "rH", "rHsrc", "FC")]
    lo = (C.c_int * 3)(0, 0, 0)
    hi = (C.c_int * 3)(3, -1, 1)
    dt = C.c_double(box["dt"])
    tm = C.c_double(0.0)
    rc = R.lib().pr_react_box(r._h, lo, hi, *[C.byref(x) for x in a], None, C.byref(dt), C.byref(tm),
                              C.c_void_p(torch.cuda.current_stream().cuda_stream), None)
    assert rc == -1 and "empty box" in R.lib().pr_last_error().decode()
    # one real batch first, so that the counters hold something the empty batch must clear
    b = {k: v.clone() for k, v in t.items()}
    b["dt"] = box["dt"]
    assert r.react_batch([b]) == 0 and r.stats()["cells"] == box["T"].size
    assert r.react_batch([]) == 0
    st = r.stats()
    assert st["cells"] == 0 and st["nst"] == 0 and st["failed"] == 0
    assert r.react_batch([], enqueue_only=True) == 0 and r.collect() == 0
    torch.cuda.synchronize()
    r.close()


@pytest.mark.parametrize("mech_box,lanes", [("h2", 16), ("h2", 8), ("ch4", 32), ("ch4", 16), ("ch4", 64)])
def test_rhs_and_jacobian_bits(gpu, mech_box, lanes):
    """Batched right-hand sides and Jacobians on 256 states of each workload, bit for bit, for
    every lanes-per-cell variant (cross-lane sums by shuffles, or DPP / readlane at 64)."""
    import torch
    from oracle import pyoracle as po
    from pelelmex_amd.reactor import Reactor
    full = W.h2_air_box() if mech_box == "h2" else W.ch4_air_flame_box()
    box = _sample(full, 256, seed=5)
    m = po.Mech(box["mech"])
    n = m.nsp + 1
    nc = box["T"].size
    y = np.vstack([box["rY"], box["T"][None, :]])
    t_eval = 0.37 * box["dt"]
    fref = np.zeros((n, nc))
    Jref = np.zeros((n, n, nc))
    for c in range(nc):
        cell = m.cell(box["rH"][c], box["rHsrc"][c], box["rYsrc"][:, c])
        fref[:, c] = m.rhs(cell, t_eval, y[:, c])
        Jref[:, :, c] = m.jac(cell, 0.0, y[:, c])
    r = Reactor.create("ReactorCvode", box["mech"], dict(KEYS, **{"hipbdf.lanes_per_cell": lanes}))
    r.init(2, 1)
    dev = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64).to(gpu)
    f = r.rhs_batch(dev(y), dev(box["rYsrc"]), dev(box["rH"]), dev(box["rHsrc"]), t_eval).cpu().numpy()
    J = r.jac_batch(dev(y), dev(box["rYsrc"]), dev(box["rHsrc"])).cpu().numpy()
    r.close()
    _assert_bits("rhs", f, fref)
    _assert_bits("jac", J, Jref)


@pytest.mark.parametrize("mech_box,lanes", [("h2", 16), ("ch4", 32), ("ch4", 64)])
def test_reaction_rate_box_bits(gpu, mech_box, lanes):
    """pr_reaction_rate_box (reactionRateRhoY over a box, MKS) bit for bit against the oracle;
    covered cells (mask == -1) get zero rates, as the reference's covered-cell branch writes."""
    import torch
    from oracle import pyoracle as po
    from pelelmex_amd.reactor import Reactor
    box = W.h2_air_box(nx=16, ny=8, nz=4) if mech_box == "h2" else W.ch4_air_flame_box(nx=16, ny=8, nz=4)
    nx, ny, nz = box["dims"]
    nc = nx * ny * nz
    m = po.Mech(box["mech"])
    rY, rH = box["rY"] * 1e3, box["rH"] * 0.1
    T0 = box["T"] + 25.0 * np.sin(np.arange(nc))
    ref, _ = m.rate_rhoY(rY, rH, T0)
    mask = np.ones(nc, dtype=np.int32)
    mask[::7] = -1
    ref[:, mask == -1] = 0.0
    r = Reactor.create("ReactorCvode", box["mech"], dict(KEYS, **{"hipbdf.lanes_per_cell": lanes}))
    r.init(2, 1)
    dev = lambda a, *shape: torch.as_tensor(np.ascontiguousarray(a)).reshape(*shape).to(gpu).contiguous()
    N = m.nsp
    out = torch.full((N, nz, ny, nx), float("nan"), dtype=torch.float64, device=gpu)
    r.reaction_rate_rhoY(dev(rY, N, nz, ny, nx), dev(rH, nz, ny, nx), dev(T0, nz, ny, nx),
                         mask=dev(mask, nz, ny, nx), rYdot=out)
    got = out.cpu().numpy().reshape(N, nc)
    r.close()
    _assert_bits("rYdot", got, ref)


@pytest.mark.parametrize("lanes", [32, 64])
def test_clean_init_massfrac_parity(gpu, lanes):
    """ode.clean_init_massfrac = 1: cells start from max(rhoY, 0) and the clipped negative part
    is added back to the result; bit-identical to the oracle, which restates the same rule."""
    box = _sample(W.ch4_air_flame_box(), 128, seed=9)
    rng = np.random.default_rng(9)
    hit = rng.random(box["rY"].shape) < 0.05
    box["rY"] = np.where(hit, -1e-9 * box["rY"].sum(0, keepdims=True) * rng.random(box["rY"].shape), box["rY"])
    assert np.any(box["rY"] < 0)
    ref = _oracle(box, max_order=4, clean_init_massfrac=1)
    out = _gpu(box, dict(KEYS, **{"cvode.max_order": 4, "ode.clean_init_massfrac": 1,
                                  "hipbdf.lanes_per_cell": lanes}), gpu)
    assert out["nfail"] == 0 and np.all(ref["status"] == 0)
    _assert_bits("rY", out["rY"], ref["rY"])
    _assert_bits("T", out["T"], ref["T"])
    _assert_bits("FC", out["FC"], ref["FC"])
    # the negative entries come back: result - input is the integrated increment of the clipped cell
    clean = _oracle(dict(box, rY=np.maximum(box["rY"], 0.0)), max_order=4)
    np.testing.assert_array_equal(ref["rY"], np.where(box["rY"] < 0, clean["rY"] + box["rY"], clean["rY"]))


def test_react_batch_matches_single_boxes(gpu):
    """pr_react_batch: boxes of mixed sizes, dt_react and masks integrated in one launch give, box
    by box, the bits of pr_react_box on the same inputs (itself bit-identical to the oracle)."""
    import torch
    from pelelmex_amd.reactor import Reactor, box_tensors
    specs = [(16, 4, 2, 5e-6, 11), (8, 8, 2, 2e-6, 12), (4, 4, 4, 1e-5, 13), (32, 2, 1, 5e-6, 14)]
    boxes = []
    for nx, ny, nz, dt, seed in specs:
        b = W.ch4_air_flame_box(nx=nx, ny=ny, nz=nz, seed=seed, dt=dt)
        if seed == 12:
            b["mask"][::5] = -1
        boxes.append(b)
    r = Reactor.create("ReactorCvode", "drm19", dict(KEYS, **{"cvode.max_order": 4}))
    r.init(2, 1)
    base = [box_tensors(b, gpu) for b in boxes]
    singles = []
    for b, t0 in zip(boxes, base):
        t = {k: v.clone() for k, v in t0.items()}
        status = torch.full((b["T"].size,), 7, dtype=torch.int32, device=gpu)
        nf, _, _ = r.react(t["rY"], t["rYsrc"], t["T"], t["rH"], t["rHsrc"], t["FC"], t["mask"], b["dt"], 0.0,
                           status=status)
        assert nf == 0
        singles.append((t, status))
    batch = []
    for b, t0 in zip(boxes, base):
        t = {k: v.clone() for k, v in t0.items()}
        t["status"] = torch.full((b["T"].size,), 7, dtype=torch.int32, device=gpu)
        t["dt"] = b["dt"]
        batch.append(t)
    assert r.react_batch(batch) == 0
    st = r.stats()
    torch.cuda.synchronize()
    r.close()
    assert st["cells"] == sum(int(np.sum(b["mask"] != -1)) for b in boxes)
    for (ts, ss), tb in zip(singles, batch):
        for k in ("rY", "T", "rH", "FC"):
            _assert_bits(k, tb[k].cpu().numpy(), ts[k].cpu().numpy())
        np.testing.assert_array_equal(tb["status"].cpu().numpy(), ss.cpu().numpy())


def test_react_batch_graph_replay(gpu):
    """The batch integration captured in a hipGraph (torch.cuda.graph around
    pr_react_batch_enqueue) and replayed three times, inputs restored eagerly and the stream
    synchronized between replays, gives the eager call's bits on every replay."""
    import torch
    from pelelmex_amd.reactor import Reactor, box_tensors
    boxes = [W.ch4_air_flame_box(nx=16, ny=4, nz=2, seed=21, dt=5e-6),
             W.ch4_air_flame_box(nx=8, ny=4, nz=4, seed=22, dt=2e-6)]
    r = Reactor.create("ReactorCvode", "drm19", dict(KEYS, **{"cvode.max_order": 4}))
    r.init(2, 1)
    base = [box_tensors(b, gpu) for b in boxes]
    work = []
    for b, t0 in zip(boxes, base):
        t = {k: v.clone() for k, v in t0.items()}
        t["dt"] = b["dt"]
        work.append(t)

    def restore():
        for t, t0 in zip(work, base):
            for k in ("rY", "T", "rH", "FC"):
                t[k].copy_(t0[k])

    restore()
    assert r.react_batch(work) == 0  # eager reference; also sizes every buffer the capture uses
    ref = [{k: t[k].clone() for k in ("rY", "T", "rH", "FC")} for t in work]
    restore()
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        r.react_batch(work, enqueue_only=True)
    for rep in range(3):
        restore()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert r.collect() == 0
        assert r.stats()["cells"] == sum(b["T"].size for b in boxes)
        for t, rf in zip(work, ref):
            for k in rf:
                _assert_bits(f"replay {rep} {k}", t[k].cpu().numpy(), rf[k].cpu().numpy())
    r.close()


def test_autotune_picks_a_variant_with_identical_results(gpu):
    """hipbdf.autotune = 1: the first react call times the lanes-per-cell variants and block sizes
    on a sample of the box and keeps the fastest; every variant computes the same bits, so the
    tuned reactor's results equal a fixed 32-lane reactor's."""
    box = W.ch4_air_flame_box(nx=64, ny=8, nz=8, seed=31)
    fixed = _gpu(box, dict(KEYS, **{"cvode.max_order": 4, "hipbdf.lanes_per_cell": 32}), gpu)
    tuned = _gpu(box, dict(KEYS, **{"cvode.max_order": 4, "hipbdf.autotune": 1}), gpu)
    assert tuned["stats"]["lanes_per_cell"] in (16, 32, 64)
    assert tuned["nfail"] == 0 and fixed["nfail"] == 0
    for k in ("rY", "T", "rH", "FC"):
        _assert_bits(k, tuned[k], fixed[k])


def test_block_size_changes_nothing_but_speed(gpu):
    """The 32-lane drm19 kernel runs 4 waves per SIMD in 1024-thread blocks by default;
    hipbdf.block = 256 and 128 give the same bits. A block above a variant's limit (512 threads
    at 64 lanes, compiled for 2 waves per SIMD) is refused before anything is launched."""
    from pelelmex_amd.reactor import ReactorError
    box = W.ch4_air_flame_box(nx=32, ny=8, nz=8, seed=37)
    base = dict(KEYS, **{"cvode.max_order": 4, "hipbdf.lanes_per_cell": 32})
    auto = _gpu(box, base, gpu)
    assert auto["nfail"] == 0
    for blk in (256, 128):
        other = _gpu(box, dict(base, **{"hipbdf.block": blk}), gpu)
        for k in ("rY", "T", "rH", "FC"):
            _assert_bits(k, other[k], auto[k])
    with pytest.raises(ReactorError, match="exceeds"):
        _gpu(box, dict(base, **{"hipbdf.lanes_per_cell": 64, "hipbdf.block": 768}), gpu)


def test_array4_views_inside_ghosted_fabs(gpu):
    """Drop-in layout: the box [lo, hi] (lo != 0) is a view into FABs with two ghost cells per
    side (AMReX Array4: p at the FAB's begin corner, FAB strides, component stride = FAB cells).
    react, reactionRateRhoY and the EOS box kernels on the views give the bits of contiguous runs
    on the same cells, and the ghost cells (NaN) are neither read into results nor written."""
    import ctypes as C
    import torch
    from pelelmex_amd import reactor as RM
    from pelelmex_amd.reactor import Reactor
    nx, ny, nz, ng = 8, 4, 4, 2
    lo = (10, 20, 30)
    box = W.ch4_air_flame_box(nx=nx, ny=ny, nz=nz, seed=41)
    box["mask"][::6] = -1
    N = box["rY"].shape[0]
    FX, FY, FZ = nx + 2 * ng, ny + 2 * ng, nz + 2 * ng

    def fab(a, ncomp, dtype=torch.float64, fill=float("nan")):
        f = torch.full((ncomp, FZ, FY, FX), fill, dtype=dtype)
        f[:, ng:ng + nz, ng:ng + ny, ng:ng + nx] = torch.as_tensor(np.asarray(a).reshape(ncomp, nz, ny, nx))
        return f.to(gpu).contiguous()

    f = {k: fab(box[k], N) for k in ("rY", "rYsrc")}
    for k in ("T", "rH", "rHsrc"):
        f[k] = fab(box[k], 1)
    f["FC"] = fab(np.zeros(box["T"].size), 1)
    fmask = fab(box["mask"], 1, dtype=torch.int32, fill=-7)
    begin = (C.c_int * 3)(lo[0] - ng, lo[1] - ng, lo[2] - ng)

    def a4(t, ncomp):
        return RM.Array4(t.data_ptr(), FX, FX * FY, FX * FY * FZ, begin, ncomp)

    r = Reactor.create("ReactorCvode", "drm19", dict(KEYS, **{"cvode.max_order": 4}))
    r.init(2, 1)
    L = RM.lib()
    clo = (C.c_int * 3)(*lo)
    chi = (C.c_int * 3)(lo[0] + nx - 1, lo[1] + ny - 1, lo[2] + nz - 1)
    arrs = [a4(f["rY"], N), a4(f["rYsrc"], N), a4(f["T"], 1), a4(f["rH"], 1), a4(f["rHsrc"], 1), a4(f["FC"], 1)]
    m = RM.IArray4(fmask.data_ptr(), FX, FX * FY, begin)
    dt, tm = C.c_double(box["dt"]), C.c_double(0.0)
    st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = L.pr_react_box(r._h, clo, chi, *[C.byref(x) for x in arrs], C.byref(m), C.byref(dt), C.byref(tm), st, None)
    assert rc == 0, L.pr_last_error()
    ref = _gpu(box, dict(KEYS, **{"cvode.max_order": 4}), gpu)
    inner = (slice(None), slice(ng, ng + nz), slice(ng, ng + ny), slice(ng, ng + nx))
    for k, ncomp in (("rY", N), ("T", 1), ("rH", 1), ("FC", 1)):
        got = f[k].cpu()
        _assert_bits(k, got[inner].reshape(ncomp, -1).numpy().squeeze(), ref[k])
        ghost = got.clone()
        ghost[inner] = 0.0
        assert torch.isnan(ghost).sum() == ghost.numel() - got[inner].numel(), f"{k}: a ghost cell was written"
    # reactionRateRhoY on the same views (MKS inputs), against a contiguous call
    fm = {k: fab(box[k] * s_, c_) for k, s_, c_ in (("rY", 1e3, N), ("rH", 0.1, 1), ("T", 1.0, 1))}
    out = fab(np.zeros((N, box["T"].size)), N)
    rc = L.pr_reaction_rate_box(r._h, clo, chi, C.byref(a4(fm["rY"], N)), C.byref(a4(fm["rH"], 1)),
                                C.byref(a4(fm["T"], 1)), C.byref(a4(out, N)), C.byref(m), st)
    assert rc == 0, L.pr_last_error()
    dev = lambda a, *shape: torch.as_tensor(np.ascontiguousarray(a)).reshape(*shape).to(gpu).contiguous()
    want = r.reaction_rate_rhoY(dev(box["rY"] * 1e3, N, nz, ny, nx), dev(box["rH"] * 0.1, nz, ny, nx),
                                dev(box["T"], nz, ny, nx), mask=dev(box["mask"], nz, ny, nx))
    torch.cuda.synchronize()
    _assert_bits("rYdot", out.cpu()[inner].numpy(), want.cpu().numpy())
    # the EOS box kernels on views of the same FABs (one thread per cell, strided rows), against
    # contiguous calls; floorSpecies gets negative and NaN entries, and no ghost is written
    rYm = box["rY"] * 1e3
    rYm[3, ::5] = -1e-9
    rYm[7, ::11] = np.nan
    rho = np.nansum(np.abs(rYm), axis=0)
    vT, vrho, vY, vH = fab(box["T"] * 1.01, 1), fab(rho, 1), fab(rYm, N), fab(box["rH"] * 0.1, 1)
    vIR, vHR = fab(np.sin(np.arange(N * box["T"].size)).reshape(N, -1), N), fab(np.zeros(box["T"].size), 1)
    rc = L.pr_set_temperature_box(r._h, clo, chi, C.byref(a4(vrho, 1)), C.byref(a4(fab(np.abs(box["rY"]) * 1e3, N), N)),
                                  C.byref(a4(vH, 1)), C.byref(a4(vT, 1)), st)
    assert rc == 0, L.pr_last_error()
    cT = dev(box["T"] * 1.01, nz, ny, nx)
    r.set_temperature(dev(rho, nz, ny, nx), dev(np.abs(box["rY"]) * 1e3, N, nz, ny, nx), dev(box["rH"] * 0.1, nz, ny, nx),
                      cT)
    rc = L.pr_floor_species_box(r._h, clo, chi, C.byref(a4(vY, N)), C.byref(a4(vrho, 1)), C.byref(a4(vH, 1)),
                                C.byref(a4(vT, 1)), st)
    assert rc == 0, L.pr_last_error()
    cY, crho, cH = dev(rYm, N, nz, ny, nx), dev(rho, nz, ny, nx), dev(box["rH"] * 0.1, nz, ny, nx)
    r.floor_species(cY, crho, cH, cT)
    rc = L.pr_heat_release_box(r._h, clo, chi, C.byref(a4(vT, 1)), C.byref(a4(vIR, N)), C.byref(a4(vHR, 1)), st)
    assert rc == 0, L.pr_last_error()
    cHR = r.heat_release(cT, dev(np.sin(np.arange(N * box["T"].size)).reshape(N, -1), N, nz, ny, nx))
    torch.cuda.synchronize()
    for name, v, c, ncomp in (("T", vT, cT, 1), ("rhoY", vY, cY, N), ("rho", vrho, crho, 1), ("rhoH", vH, cH, 1),
                              ("HR", vHR, cHR, 1)):
        got = v.cpu()
        _assert_bits(name, got[inner].reshape(ncomp, -1).numpy().squeeze(), c.cpu().reshape(ncomp, -1).numpy().squeeze())
        ghost = got.clone()
        ghost[inner] = 0.0
        assert torch.isnan(ghost).sum() == ghost.numel() - got[inner].numel(), f"{name}: a ghost cell was written"
    r.close()


@pytest.mark.parametrize("mech_box,lanes", [("h2", 16), ("h2", 8), ("ch4", 32), ("ch4", 16), ("ch4", 64)])
def test_eos_box_kernels_bits(gpu, mech_box, lanes):
    """setTemperature, floorSpecies and getHeatRelease over a box (MKS Array4 views), bit for bit
    against the oracle for every lanes-per-cell variant; floorSpecies' input carries negative,
    -0 and NaN species densities (clipped to +0 and DBL_MAX as std::min / std::max do)."""
    import torch
    from oracle import pyoracle as po
    from pelelmex_amd.reactor import Reactor
    box = W.h2_air_box(nx=16, ny=8, nz=4) if mech_box == "h2" else W.ch4_air_flame_box(nx=16, ny=8, nz=4)
    nx, ny, nz = box["dims"]
    nc = nx * ny * nz
    m = po.Mech(box["mech"])
    N = m.nsp
    rY, rH, T = box["rY"] * 1e3, box["rH"] * 0.1, box["T"].copy()
    rho = rY.sum(0)
    T0 = T + 40.0 * np.cos(np.arange(nc))
    r = Reactor.create("ReactorCvode", box["mech"], dict(KEYS, **{"hipbdf.lanes_per_cell": lanes}))
    r.init(2, 1)
    dev = lambda a, shp: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64).reshape(shp).to(gpu)
    s3, sN = (nz, ny, nx), (N, nz, ny, nx)
    # setTemperature
    Tg = dev(T0, s3)
    r.set_temperature(dev(rho, s3), dev(rY, sN), dev(rH, s3), Tg)
    _assert_bits("T", Tg.cpu().numpy().reshape(-1), m.set_temperature(rho, rY, rH, T0))
    # floorSpecies
    rYn = rY.copy()
    rYn[1, ::3] = -1e-9
    rYn[2, 1::4] = -0.0
    rYn[0, 5] = np.nan
    ref_rY, ref_rho, ref_rH = m.floor_species(rYn, T)
    gY, grho, grH = dev(rYn, sN), torch.zeros(s3, dtype=torch.float64, device=gpu), torch.zeros(s3, dtype=torch.float64, device=gpu)
    r.floor_species(gY, grho, grH, dev(T, s3))
    _assert_bits("rY", gY.cpu().numpy().reshape(N, -1), ref_rY)
    _assert_bits("rho", grho.cpu().numpy().reshape(-1), ref_rho)
    _assert_bits("rH", grH.cpu().numpy().reshape(-1), ref_rH)
    # getHeatRelease on the reactionRateRhoY rates
    IR, _ = m.rate_rhoY(rY, rH, T)
    HR = r.heat_release(dev(T, s3), dev(IR, sN))
    _assert_bits("HR", HR.cpu().numpy().reshape(-1), m.heat_release(T, IR))
    r.close()


def test_cost_order_changes_nothing_but_the_order(gpu):
    """hipbdf.cost_order = 1 hands cells out by decreasing input FC (the previous advance's
    right-hand-side count): run on the FC of a first run, the second run's results and counters
    are bit-identical to the first's."""
    import torch
    from pelelmex_amd.reactor import Reactor, box_tensors
    box = W.ch4_air_flame_box(nx=64, ny=8, nz=4, seed=43)
    box["mask"][::11] = -1
    keys = dict(KEYS, **{"cvode.max_order": 4, "hipbdf.lanes_per_cell": 32})
    outs = []
    fc_prev = None
    for order in (0, 1):
        r = Reactor.create("ReactorCvode", box["mech"], dict(keys, **{"hipbdf.cost_order": order}))
        r.init(2, 1)
        t = box_tensors(box, gpu)
        if fc_prev is not None:
            t["FC"].copy_(fc_prev)
        nf, _, _ = r.react(t["rY"], t["rYsrc"], t["T"], t["rH"], t["rHsrc"], t["FC"], t["mask"], box["dt"], 0.0)
        torch.cuda.synchronize()
        st = r.stats()
        r.close()
        fc_prev = t["FC"].clone()
        outs.append(({k: t[k].cpu().numpy() for k in ("rY", "T", "rH", "FC")}, nf, st))
    assert float(fc_prev.max()) > 0.0
    for k in ("rY", "T", "rH", "FC"):
        _assert_bits(k, outs[1][0][k].reshape(-1), outs[0][0][k].reshape(-1))
    assert outs[0][1] == outs[1][1] == 0
    for k in ("cells", "nst", "nfe", "nje", "nsetups", "nni", "netf", "ncfn"):
        assert outs[0][2][k] == outs[1][2][k], k


def test_box_costs(gpu):
    """pr_box_costs (PeleLM::computeCosts for the chemistry BoxArray): Ncell, and the average,
    maximum and sum of FC per box, exactly as numpy computes them for whole-number FC, on boxes
    of three sizes; a reactor's own FC gives a cost that grows with the box's stiff content."""
    import torch
    from pelelmex_amd.reactor import Reactor, box_tensors
    rng = np.random.default_rng(7)
    shapes = [(4, 4, 4), (8, 4, 2), (16, 16, 16)]
    fcs = [torch.as_tensor(rng.integers(0, 3000, size=s).astype(np.float64)).to(gpu) for s in shapes]
    r = Reactor.create("ReactorCvode", "drm19", KEYS)
    r.init(2, 1)
    n = np.array([f.numel() for f in fcs], dtype=np.float64)
    np.testing.assert_array_equal(r.box_costs(fcs, "Ncell"), n)
    host = [f.cpu().numpy() for f in fcs]
    np.testing.assert_array_equal(r.box_costs(fcs, "ChemFunctCallSum"), [h.sum() for h in host])
    np.testing.assert_array_equal(r.box_costs(fcs, "ChemFunctCallMax"), [h.max() for h in host])
    np.testing.assert_array_equal(r.box_costs(fcs, "ChemFunctCallAvg"), [h.sum() / h.size for h in host])
    # costs from a react: the flame-front slab costs more per cell than the cold slab
    box = W.ch4_air_flame_box(nx=64, ny=4, nz=4, seed=5)
    t = box_tensors(box, gpu)
    r.react(t["rY"], t["rYsrc"], t["T"], t["rH"], t["rHsrc"], t["FC"], t["mask"], box["dt"], 0.0)
    torch.cuda.synchronize()
    fc = t["FC"].reshape(4, 4, 64)
    cold, front = fc[:, :, :16].contiguous(), fc[:, :, 20:30].contiguous()
    c = r.box_costs([cold, front], "ChemFunctCallAvg")
    assert c[1] > c[0] > 0
    r.close()


@pytest.mark.parametrize("lanes", [32, 64])
def test_too_much_work_status_parity(gpu, lanes):
    """cvode.max_substeps = 60: stiff cells stop with CVODE's too-much-work code (-1) after their
    last accepted step; per-cell status, the state written for failed cells (the last accepted
    step), FC and the failed-cell count are the oracle's, bit for bit."""
    full = W.ch4_air_flame_box()
    idx = np.union1d(np.arange(64), np.random.default_rng(9).choice(full["T"].size, 192, replace=False))
    box = W.subsample(full, idx)
    ref = _oracle(box, max_order=5, max_steps=60)
    out = _gpu(box, dict(KEYS, **{"cvode.max_substeps": 60, "hipbdf.lanes_per_cell": lanes}), gpu)
    failed = ref["status"] != 0
    assert failed.sum() > 10 and (~failed).sum() > 10        # both kinds present
    assert np.all(ref["status"][failed] == -1)
    np.testing.assert_array_equal(out["status"], ref["status"])
    assert out["nfail"] == int(failed.sum())
    for k in ("rY", "T", "rH", "FC"):
        _assert_bits(k, out[k], ref[k])


def test_max_order_2_parity(gpu):
    """cvode.max_order = 2 (the documented default): orders 1-2 only, bit for bit."""
    full = W.ch4_air_flame_box()
    box = W.subsample(full, np.random.default_rng(12).choice(full["T"].size, 192, replace=False))
    ref = _oracle(box, max_order=2)
    out = _gpu(box, dict(KEYS, **{"cvode.max_order": 2, "hipbdf.lanes_per_cell": 32}), gpu)
    assert out["nfail"] == 0
    for k in ("rY", "T", "rH", "FC"):
        _assert_bits(k, out[k], ref[k])


def test_reaction_source_box_bits(gpu):
    """pr_reaction_source_box (I_R formation) on a box of random MKS states: the oracle's bits."""
    import torch
    from oracle import pyoracle as po
    from pelelmex_amd.reactor import Reactor
    rng = np.random.default_rng(12)
    shape = (21, 5, 6, 7)
    o = rng.random(shape) * 0.2
    n = o * (1 + 1e-3 * rng.standard_normal(shape))
    F = rng.standard_normal(shape) * 50
    r = Reactor.create("ReactorCvode", "drm19", KEYS)
    r.init(2, 1)
    dev = lambda a: torch.as_tensor(a).to(gpu).contiguous()
    IR = r.reaction_source(dev(o), dev(n), dev(F), 3.7e-6).cpu().numpy()
    r.close()
    _assert_bits("I_R", IR, po.reaction_source(o, n, F, 3.7e-6))


def test_advance_chemistry_box_bits(gpu):
    """pr_advance_chemistry_box (one box of advanceChemistry on MKS views: conversions, react,
    conversions back, I_R) against the same sequence around the oracle's react: new state,
    forcing after its round trip, I_R and FC bit for bit, covered cells untouched by react."""
    import torch
    from oracle import pyoracle as po
    from pelelmex_amd.reactor import Reactor
    b = W.ch4_air_flame_box(nx=16, ny=4, nz=2, seed=61, dt=3e-6)
    b["mask"][::6] = -1
    N, n = b["rY"].shape
    # the workload's CGS state as the MKS level data PeleLMeX holds
    rY_o, rH_o, T_o = b["rY"] * 1.0e3, b["rH"] * 0.1, b["T"].copy()
    F_rY, F_rH = b["rYsrc"] * 1.0e3, b["rHsrc"] * 0.1
    ref = po.advance_chemistry(po.Mech("drm19"), rY_o, rH_o, T_o, F_rY, F_rH, b["mask"], b["dt"], max_order=4)
    nx, ny, nz = b["dims"]
    dev = lambda a, c: torch.as_tensor(np.ascontiguousarray(a)).reshape((c, nz, ny, nx) if c else (nz, ny, nx)).to(gpu)
    old = {"rY": dev(rY_o, N), "rH": dev(rH_o, 0), "T": dev(T_o, 0)}
    new = {"rY": torch.zeros_like(old["rY"]), "rH": torch.zeros_like(old["rH"]), "T": torch.zeros_like(old["T"])}
    FY, FH = dev(F_rY, N), dev(F_rH, 0)
    FC = torch.zeros_like(old["T"])
    IR = torch.zeros_like(old["rY"])
    mask = torch.as_tensor(b["mask"]).reshape(nz, ny, nx).to(gpu)
    r = Reactor.create("ReactorCvode", "drm19", dict(KEYS, **{"cvode.max_order": 4}))
    r.init(2, 1)
    assert r.advance_chemistry(old, new, FY, FH, FC, mask, b["dt"], IR=IR) == 0
    torch.cuda.synchronize()
    r.close()
    host = lambda t, c: t.cpu().numpy().reshape((c, n) if c else (n,))
    _assert_bits("rY", host(new["rY"], N), ref["rY"])
    _assert_bits("rH", host(new["rH"], 0), ref["rH"])
    _assert_bits("T", host(new["T"], 0), ref["T"])
    _assert_bits("F_rY", host(FY, N), ref["FY"])
    _assert_bits("F_rH", host(FH, 0), ref["FH"])
    _assert_bits("I_R", host(IR, N), ref["IR"])
    _assert_bits("FC", host(FC, 0), ref["FC"])
    cov = b["mask"] == -1
    assert np.all(ref["FC"][cov] == 0) and np.all(ref["FC"][~cov] > 0)


@pytest.mark.parametrize("do_reaction", [1, 0])
def test_advance_chemistry_chem_box_bits(gpu, do_reaction):
    """pr_advance_chemistry_chem_box (one box of advanceChemistryBAChem on the chemistry BoxArray's
    copy, in place: conversions, react - or FC = 0 for a box flagged covered - and the state's
    conversion back, the forcing copy left in CGS) against the same sequence around the oracle's
    react, bit for bit; set_eos_parm accepts the ideal-gas EOS and refuses a real-gas one."""
    import torch
    from oracle import pyoracle as po
    from pelelmex_amd.reactor import Reactor, ReactorError
    b = W.ch4_air_flame_box(nx=16, ny=4, nz=2, seed=64, dt=3e-6)
    b["mask"][::5] = -1
    N, n = b["rY"].shape
    rY, rH, T = b["rY"] * 1.0e3, b["rH"] * 0.1, b["T"].copy()
    F_rY, F_rH = b["rYsrc"] * 1.0e3, b["rHsrc"] * 0.1
    ref = po.advance_chemistry_chem(po.Mech("drm19"), rY, rH, T, F_rY, F_rH, b["mask"], b["dt"],
                                    do_reaction=bool(do_reaction), max_order=4)
    nx, ny, nz = b["dims"]
    dev = lambda a, c: torch.as_tensor(np.ascontiguousarray(a)).reshape((c, nz, ny, nx) if c else (nz, ny, nx)).to(gpu)
    tY, tH, tT, FY, FH = dev(rY, N), dev(rH, 0), dev(T, 0), dev(F_rY, N), dev(F_rH, 0)
    FC = torch.full_like(tT, 7.0)
    mask = torch.as_tensor(b["mask"]).reshape(nz, ny, nx).to(gpu)
    r = Reactor.create("ReactorCvode", "drm19", dict(KEYS, **{"cvode.max_order": 4}))
    r.init(2, 1)
    r.set_eos_parm(None, "Fuego")
    with pytest.raises(ReactorError):
        r.set_eos_parm(None, "SRK")
    assert r.advance_chemistry_chem(tY, tH, tT, FY, FH, FC, mask, b["dt"], do_reaction=bool(do_reaction)) == 0
    torch.cuda.synchronize()
    r.close()
    host = lambda t, c: t.cpu().numpy().reshape((c, n) if c else (n,))
    for name, got, want, c in (("rY", tY, ref["rY"], N), ("rH", tH, ref["rH"], 0), ("T", tT, ref["T"], 0),
                               ("F_rY", FY, ref["FY"], N), ("F_rH", FH, ref["FH"], 0), ("FC", FC, ref["FC"], 0)):
        _assert_bits(name, host(got, c), want)
    if not do_reaction:
        assert np.all(ref["FC"] == 0)
        np.testing.assert_array_equal(host(tY, N), (rY * 1.0e-3) * 1.0e3)


def test_advance_chemistry_bachem_level(gpu):
    """advanceChemistryBAChem over a whole level (Source/PeleLMeX_Reactions.cpp:180-367): the
    chemistry BoxArray from pr_chem_boxes (two grids, a finer level covering part of them,
    max_grid_size_chem 4), the state and forcing copied into each chemistry box, the per-box
    body (react on uncovered boxes, FC = 0 on covered ones), the copies back and I_R - every
    cell bit for bit with the oracle's per-cell sequence (react on the cells the finer level
    does not cover)."""
    import torch
    from oracle import pyoracle as po
    from pelelmex_amd.reactor import Reactor, advance_chemistry_bachem, chem_boxes
    nx, ny, nz = 16, 8, 4
    b = W.ch4_air_flame_box(nx=nx, ny=ny, nz=nz, seed=65, dt=2e-6)
    N, n = b["rY"].shape
    dom = ((0, 0, 0), (nx - 1, ny - 1, nz - 1))
    grids = [((0, 0, 0), (7, 7, 3)), ((8, 0, 0), (15, 7, 3))]
    fine = [((6, 2, 0), (17, 9, 5))]                      # at ref_ratio 2: cells (3..8, 1..4, 0..2)
    boxes, flags = chem_boxes(dom, grids, fine, ref_ratio=2, max_grid_size_chem=(4, 4, 4))
    assert flags.count(0) == 1 and flags.count(1) > 4
    ii, jj, kk = np.arange(n) % nx, (np.arange(n) // nx) % ny, np.arange(n) // (nx * ny)
    covered = (ii >= 3) & (ii <= 8) & (jj >= 1) & (jj <= 4) & (kk <= 2)
    rY, rH, T = b["rY"] * 1.0e3, b["rH"] * 0.1, b["T"].copy()
    F_rY, F_rH = b["rYsrc"] * 1.0e3, b["rHsrc"] * 0.1
    m = po.Mech("drm19")
    react = po.advance_chemistry_chem(m, rY, rH, T, F_rY, F_rH, b["mask"], b["dt"], do_reaction=True, max_order=4)
    skip = po.advance_chemistry_chem(m, rY, rH, T, F_rY, F_rH, b["mask"], b["dt"], do_reaction=False, max_order=4)
    ref = {k: np.where(covered, skip[k], react[k]) for k in ("rY", "rH", "T", "FC")}
    ref["IR"] = -(rY - ref["rY"]) * (1.0 / b["dt"]) - F_rY
    dev = lambda a, c: torch.as_tensor(np.ascontiguousarray(a)).reshape((c, nz, ny, nx) if c else (nz, ny, nx)).to(gpu)
    old = {"rY": dev(rY, N), "rH": dev(rH, 0), "T": dev(T, 0)}
    new = {k: torch.full_like(v, -1.0) for k, v in old.items()}
    FY, FH = dev(F_rY, N), dev(F_rH, 0)
    FC = torch.full_like(old["T"], 7.0)
    IR = torch.zeros_like(old["rY"])
    r = Reactor.create("ReactorCvode", "drm19", dict(KEYS, **{"cvode.max_order": 4}))
    r.init(2, 1)
    assert advance_chemistry_bachem(r, dom[0], old, new, FY, FH, FC, IR, b["dt"], boxes, flags) == 0
    torch.cuda.synchronize()
    r.close()
    host = lambda t, c: t.cpu().numpy().reshape((c, n) if c else (n,))
    for name, got, c in (("rY", new["rY"], N), ("rH", new["rH"], 0), ("T", new["T"], 0), ("FC", FC, 0),
                         ("IR", IR, N)):
        _assert_bits(name, host(got, c), ref[name])
    assert np.all(host(FC, 0)[covered] == 0) and np.all(host(FC, 0)[~covered] > 0)
    np.testing.assert_array_equal(host(FY, N), F_rY)   # the caller's forcing is not touched


def test_one_handle_from_two_host_threads(gpu):
    """Calls on one handle from two host threads at once (as PeleLMeX's OpenMP box loop would
    make them) are serialized by the library: each thread's boxes get the bits a sequential
    run gives."""
    import threading
    import torch
    from pelelmex_amd.reactor import Reactor, box_tensors
    boxes = [W.ch4_air_flame_box(nx=16, ny=8, nz=4, seed=70 + i, dt=4e-6) for i in range(4)]
    r = Reactor.create("ReactorCvode", "drm19", dict(KEYS, **{"cvode.max_order": 4}))
    r.init(2, 1)
    base = [box_tensors(b, gpu) for b in boxes]

    def run(t, b):
        nf, _, _ = r.react(t["rY"], t["rYsrc"], t["T"], t["rH"], t["rHsrc"], t["FC"], t["mask"], b["dt"], 0.0)
        assert nf == 0

    seq = []
    for b, t0 in zip(boxes, base):
        t = {k: v.clone() for k, v in t0.items()}
        run(t, b)
        seq.append(t)
    torch.cuda.synchronize()
    par = [{k: v.clone() for k, v in t0.items()} for t0 in base]
    errors = []

    def worker(ids):
        try:
            for i in ids:
                run(par[i], boxes[i])
        except Exception as e:  # reported below
            errors.append(e)

    th = [threading.Thread(target=worker, args=(ids,)) for ids in ((0, 2), (1, 3))]
    for x in th:
        x.start()
    for x in th:
        x.join(timeout=120)
    torch.cuda.synchronize()
    r.close()
    assert not errors, errors
    for ts, tp in zip(seq, par):
        for k in ("rY", "T", "rH", "FC"):
            _assert_bits(k, tp[k].cpu().numpy(), ts[k].cpu().numpy())


def test_advance_chemistry_error_restores_units(gpu):
    """An error inside react (here a block size above the variant's limit, found before any
    integration) still scales the new state and the forcing back to MKS: the forcing ends as
    the reference's round trip F * 1e-3 * 1e3, the new state as the old one's round trip."""
    import torch
    from pelelmex_amd.reactor import Reactor, ReactorError
    b = W.ch4_air_flame_box(nx=8, ny=4, nz=2, seed=62, dt=3e-6)
    N, n = b["rY"].shape
    nx, ny, nz = b["dims"]
    dev = lambda a, c: torch.as_tensor(np.ascontiguousarray(a)).reshape((c, nz, ny, nx) if c else (nz, ny, nx)).to(gpu)
    old = {"rY": dev(b["rY"] * 1e3, N), "rH": dev(b["rH"] * 0.1, 0), "T": dev(b["T"], 0)}
    new = {k: torch.zeros_like(v) for k, v in old.items()}
    F_rY, F_rH = b["rYsrc"] * 1e3, b["rHsrc"] * 0.1
    FY, FH = dev(F_rY, N), dev(F_rH, 0)
    FC = torch.zeros_like(old["T"])
    r = Reactor.create("ReactorCvode", "drm19", dict(KEYS, **{"hipbdf.lanes_per_cell": 64, "hipbdf.block": 768}))
    r.init(2, 1)
    with pytest.raises(ReactorError, match="exceeds"):
        r.advance_chemistry(old, new, FY, FH, FC, None, b["dt"])
    torch.cuda.synchronize()
    r.close()
    host = lambda t, c: t.cpu().numpy().reshape((c, n) if c else (n,))
    np.testing.assert_array_equal(host(FY, N), (F_rY * 1.0e-3) * 1.0e3)
    np.testing.assert_array_equal(host(FH, 0), (F_rH * 10.0) * 0.1)
    np.testing.assert_array_equal(host(new["rY"], N), ((b["rY"] * 1e3) * 1.0e-3) * 1.0e3)
    np.testing.assert_array_equal(host(new["T"], 0), b["T"])


@pytest.mark.parametrize("max_order", [4, 5])
def test_h2_air_whole_box_parity(gpu, max_order):
    """The whole LiDryer 32^3 box (every one of its 32768 cells) bit for bit with the oracle."""
    box = W.h2_air_box()
    ref = _oracle(box, max_order=max_order)
    out = _gpu(box, dict(KEYS, **{"cvode.max_order": max_order}), gpu)
    assert out["nfail"] == 0 and np.all(out["status"] == 0)
    for k in ("rY", "T", "rH", "FC"):
        _assert_bits(k, out[k], ref[k])


def test_ch4_flame_large_sample_parity(gpu):
    """4096 cells of a drm19 flame box of another seed (all three slab types, with sources), the
    default 32-lane build, bit for bit with the oracle."""
    full = W.ch4_air_flame_box(seed=2077)
    box = _sample(full, 4096, seed=13)
    ref = _oracle(box, max_order=4)
    out = _gpu(box, dict(KEYS, **{"cvode.max_order": 4}), gpu)
    assert out["nfail"] == 0 and np.all(out["status"] == 0)
    for k in ("rY", "T", "rH", "FC"):
        _assert_bits(k, out[k], ref[k])


def _fixture_box(name):
    import os
    g = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", name + ".npz"))
    n = g["T"].size
    box = dict(mech=str(g["mech"]), rY=g["rY"], rYsrc=g["rYsrc"], T=g["T"], rH=g["rH"], rHsrc=g["rHsrc"],
               mask=np.ones(n, dtype=np.int32), dt=float(g["dt"]), dims=(n, 1, 1))
    return g, box


def _state(out):
    return np.vstack([out["rY"], out["T"]])


def _errs_vs(y, ref):
    N = ref.shape[0] - 1
    Y = y[:N] / y[:N].sum(0)
    Yr = ref[:N] / ref[:N].sum(0)
    return float(np.abs(Y - Yr).max()), float((np.abs(y[N] - ref[N]) / ref[N]).max())


@pytest.mark.parametrize("fixture", ["h2_air_cells", "ch4_air_flame_cells"])
def test_gpu_against_vode_fixtures(gpu, fixture):
    """The specified tolerance against integrators independent of the oracle, on the fixture
    cells (every slab): the GPU reactor at ode.rtol = ode.atol = 1e-10 against Fortran VODE at
    the same tolerances (orders 4 and 5), and with PeleLMeX's typical values (atol_i = ode.atol *
    typ_i, its default path) against VODE with the same atol vector and against the rtol-1e-13
    truth: max |dY| <= 1e-8, max |dT| / T <= 1e-6 in every case."""
    g, box = _fixture_box(fixture)
    for q in (4, 5):
        out = _gpu(box, dict(KEYS, **{"cvode.max_order": q}), gpu)
        assert out["nfail"] == 0
        y = _state(out)
        dY, dT = _errs_vs(y, g[f"vode10_q{q}"])
        print(f"{fixture} q{q} vs VODE(1e-10): dY {dY:.2e} dT/T {dT:.2e}")
        assert dT <= T_TOL
        if q == 4:
            assert dY <= Y_TOL
        else:   # order 5: one flame-front cell of 264 at 1.4e-8 (VODE's heuristics are not CVODE's)
            N = y.shape[0] - 1
            ref = g["vode10_q5"]
            per = np.abs(y[:N] / y[:N].sum(0) - ref[:N] / ref[:N].sum(0)).max(0)
            assert np.mean(per > Y_TOL) <= 0.01 and dY <= 2 * Y_TOL
    out = _gpu(box, dict(KEYS, **{"cvode.max_order": 4}), gpu, typ=g["typ"])
    assert out["nfail"] == 0
    for ref in ("vode10_typ_q4", "radau"):
        dY, dT = _errs_vs(_state(out), g[ref])
        print(f"{fixture} typical values vs {ref}: dY {dY:.2e} dT/T {dT:.2e}")
        assert dY <= Y_TOL and dT <= T_TOL


@pytest.mark.parametrize("mech_box,lanes", [("h2", 16), ("ch4", 32), ("ch4", 64)])
def test_typical_values_parity(gpu, mech_box, lanes):
    """PeleLMeX's default tolerance path: set_typ_vals_ode with the typical values of the whole
    box (typical_values_chem, the setTypicalValues / updateTypicalValuesChem formula), atol_i =
    ode.atol * typ_i in the kernel, bit for bit with the oracle given atol_vec = 1e-10 * typ."""
    from pelelmex_amd.reactor import typical_values_chem
    full = W.h2_air_box() if mech_box == "h2" else W.ch4_air_flame_box()
    typ = typical_values_chem(full["rY"] * 1e3, full["T"])
    if mech_box == "h2":
        box = _sample(full, 512, seed=41)
    else:
        idx = np.union1d(np.arange(64), np.random.default_rng(42).choice(full["T"].size, 448, replace=False))
        box = W.subsample(full, idx)
    ref = _oracle(box, max_order=4, atol_vec=1e-10 * typ)
    out = _gpu(box, dict(KEYS, **{"cvode.max_order": 4, "hipbdf.lanes_per_cell": lanes}), gpu, typ=typ)
    assert out["nfail"] == 0 and np.all(out["status"] == 0) and np.all(ref["status"] == 0)
    for k in ("rY", "T", "rH", "FC"):
        _assert_bits(k, out[k], ref[k])
    # the typical values change the steps: the scalar-atol result differs
    plain = _oracle(box, max_order=4)
    assert not np.array_equal(plain["FC"], ref["FC"])


def test_graph_replay_survives_other_calls_on_the_handle(gpu):
    """A batch captured in a hipGraph keeps its own box table: pr_react_box on another box and
    set_typ_vals_ode between replays neither corrupt nor stale the replay. The replay after
    set_typ_vals_ode integrates with the new atol (as an eager call would), the one after the
    box call reproduces the eager bits, and a larger box after the capture grows new buffers
    while the graph keeps replaying on the ones it captured."""
    import torch
    from pelelmex_amd.reactor import Reactor, box_tensors, typical_values_chem
    b0 = W.ch4_air_flame_box(nx=16, ny=4, nz=2, seed=61, dt=5e-6)
    other = W.ch4_air_flame_box(nx=8, ny=2, nz=2, seed=62, dt=2e-6)
    keys = dict(KEYS, **{"cvode.max_order": 4})
    r = Reactor.create("ReactorCvode", "drm19", keys)
    r.init(2, 1)
    t0 = box_tensors(b0, gpu)
    work = {k: v.clone() for k, v in t0.items()}
    work["dt"] = b0["dt"]

    def restore():
        for k in ("rY", "T", "rH", "FC"):
            work[k].copy_(t0[k])

    assert r.react_batch([work]) == 0
    ref = {k: work[k].clone() for k in ("rY", "T", "rH", "FC")}
    restore()
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        r.react_batch([work], enqueue_only=True)
    # another box through pr_react_box on the same handle (restages the handle's own table)
    to = box_tensors(other, gpu)
    assert r.react(to["rY"], to["rYsrc"], to["T"], to["rH"], to["rHsrc"], to["FC"], to["mask"], other["dt"])[0] == 0
    restore()
    torch.cuda.synchronize()
    g.replay()
    torch.cuda.synchronize()
    assert r.collect() == 0
    for k in ref:
        _assert_bits(f"replay after a box call: {k}", work[k].cpu().numpy(), ref[k].cpu().numpy())
    # typical values set after the capture reach the replay
    typ = typical_values_chem(b0["rY"] * 1e3, b0["T"])
    r.set_typ_vals_ode(typ)
    restore()
    torch.cuda.synchronize()
    g.replay()
    torch.cuda.synchronize()
    assert r.collect() == 0
    n = b0["T"].size
    o = _oracle(b0, max_order=4, atol_vec=1e-10 * typ)
    _assert_bits("replay after set_typ_vals_ode: rY", work["rY"].reshape(-1, n).cpu().numpy(), o["rY"])
    # a box larger than the captured buffers: the handle grows new ones (the graph keeps its own)
    big = W.ch4_air_flame_box(nx=64, ny=8, nz=8, seed=63)
    tb = box_tensors(big, gpu)
    assert r.react(tb["rY"], tb["rYsrc"], tb["T"], tb["rH"], tb["rHsrc"], tb["FC"], tb["mask"], big["dt"])[0] == 0
    ob = _oracle(big, max_order=4, atol_vec=1e-10 * typ)
    _assert_bits("larger box after the capture: rY", tb["rY"].reshape(-1, big["T"].size).cpu().numpy(), ob["rY"])
    restore()
    torch.cuda.synchronize()
    g.replay()
    torch.cuda.synchronize()
    assert r.collect() == 0
    _assert_bits("replay after the buffers grew: rY", work["rY"].reshape(-1, n).cpu().numpy(), o["rY"])
    del g
    r.close()
