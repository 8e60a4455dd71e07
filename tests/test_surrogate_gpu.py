"""GPU tests of the learned map + trust check + BDF fallback path of pr_react_box.

* cells the trust check rejects are integrated by the same BDF kernel from untouched inputs:
  bit-identical to a BDF-only call; accepted cells stay within max |dT| / T <= 1e-4 of it;
* the audit (the BDF re-solving sampled accepted cells) reports exactly the error a direct
  comparison gives on those cells;
* the trained network against the method without it: it takes more cells off the BDF than an
  untrained network of the same shape, and no more BDF work than the zero-increment predictor
  (the residual trust check certifies frozen cells only);
* bad networks are refused by pr_surrogate_load.
"""
import os

import numpy as np
import pytest

from pelelmex_amd import workloads as W

pytestmark = pytest.mark.gpu

KEYS = {"ode.rtol": 1e-10, "ode.atol": 1e-10, "cvode.max_order": 4, "cvode.max_substeps": 10000}


def _model():
    from pelelmex_amd import surrogate as sg
    if not os.path.exists(sg.WEIGHTS):
        pytest.skip("no trained network in the package")
    return sg.load(sg.WEIGHTS)


def _keys(model, **extra):
    k = dict(KEYS)
    if "trust_rtol" in model:
        k["surrogate.trust_rtol"] = float(model["trust_rtol"])
        k["surrogate.trust_atol"] = float(model["trust_atol"])
    k.update(extra)
    return k


def _run(keys, box, gpu, model=None):
    import torch
    from pelelmex_amd import surrogate as sg
    from pelelmex_amd.reactor import Reactor, box_tensors
    r = Reactor.create("ReactorCvode", "drm19", keys)
    r.init(2, 1)
    if model is not None:
        sg.attach(r, model)
    t = box_tensors(box, gpu)
    rH0 = t["rH"].clone()
    nf, _, _ = r.react(t["rY"], t["rYsrc"], t["T"], t["rH"], t["rHsrc"], t["FC"], t["mask"], box["dt"], 0.0)
    torch.cuda.synchronize()
    st = r.stats()
    r.close()
    out = {k: t[k].cpu().numpy().reshape(t[k].shape[0], -1) if k == "rY" else t[k].cpu().numpy().reshape(-1)
           for k in ("rY", "T", "rH", "FC")}
    out["rH0"] = rH0.cpu().numpy().reshape(-1)
    return nf, st, out


def _box():
    return W.ch4_air_flame_box(nx=64, ny=8, nz=4, seed=2002)


def test_fallback_exact_and_accepted_within_tolerance(gpu):
    model = _model()
    box = _box()
    nc = box["T"].size
    nf0, _, ref = _run(KEYS, box, gpu)
    nf1, st, sur = _run(_keys(model), box, gpu, model)
    assert nf0 == 0 and nf1 == 0
    acc = sur["FC"] == 1.0
    assert st["sur_accepted"] == int(acc.sum()) and st["sur_accepted"] + st["sur_rejected"] == nc
    rej = ~acc
    for k in ("rY", "T", "FC"):
        np.testing.assert_array_equal(sur[k][..., rej], ref[k][..., rej])
    np.testing.assert_array_equal(sur["rH"], ref["rH"])
    if acc.any():
        dT = np.abs(sur["T"][acc] - ref["T"][acc]) / ref["T"][acc]
        assert dT.max() <= 1e-4, dT.max()


def test_audit_reports_the_direct_error(gpu):
    model = _model()
    box = _box()
    _, _, ref = _run(KEYS, box, gpu)
    _, st, sur = _run(_keys(model, **{"surrogate.audit_frac": 1.0}), box, gpu, model)
    acc = sur["FC"] == 1.0
    if not acc.any():
        pytest.skip("network accepted no cell of this box")
    assert st["sur_audited"] == int(acc.sum())
    direct = np.max(np.abs(ref["T"][acc] - sur["T"][acc]) / ref["T"][acc])
    assert st["audit_max_dT_rel"] == pytest.approx(direct, rel=1e-12, abs=1e-300)
    rho_r = ref["rY"][:, acc].sum(0)
    rho_s = sur["rY"][:, acc].sum(0)
    dY = np.max(np.abs(ref["rY"][:, acc] / rho_r - sur["rY"][:, acc] / rho_s))
    assert st["audit_max_dY"] == pytest.approx(dY, rel=1e-12, abs=1e-300)


def test_network_against_the_zero_increment_predictor(gpu):
    """What the trained network takes off the BDF, against two baselines on the same box with the
    same trust tolerances: an untrained network of the same shape, and the zero-increment
    predictor (surrogate.zero_increment = 1: the exact source step alone, no network pass). The
    trained network beats the untrained one; against the zero increment it removes no more BDF
    work (within 3 % of the box's right-hand-side evaluations): the one-step residual check only
    certifies chemically frozen cells, whatever the prediction (tests/test_trust_check_limits.py
    shows the exact solution itself failing it on the reacting cells). The zero-increment path
    keeps the same guarantees: accepted cells within 1e-4 of the BDF, the rest bit-identical."""
    from pelelmex_amd import surrogate as sg
    model = _model()
    box = _box()
    _, st0, ref = _run(KEYS, box, gpu)
    _, st1, sur = _run(_keys(model), box, gpu, model)
    untrained = sg.MLP([int(d) for d in model["dims"]], device=gpu, seed=11)
    _, st2, _ = _run(_keys(model), box, gpu, (untrained, model["xnorm"], model["ynorm"]))
    _, st3, zer = _run(_keys(model, **{"surrogate.zero_increment": 1}), box, gpu, model)
    bdf_rhs = ref["FC"].sum()
    removed = lambda out: ref["FC"][out["FC"] == 1.0].sum() / bdf_rhs
    print(f"accepted: trained {st1['sur_accepted']}, untrained {st2['sur_accepted']}, zero increment "
          f"{st3['sur_accepted']}; BDF work removed: trained {removed(sur):.4f}, zero {removed(zer):.4f}")
    assert st1["sur_accepted"] >= 0.1 * box["T"].size
    assert st2["sur_accepted"] < st1["sur_accepted"]
    assert abs(removed(sur) - removed(zer)) <= 0.03
    acc = zer["FC"] == 1.0
    assert acc.sum() == st3["sur_accepted"] and acc.any()
    np.testing.assert_array_equal(zer["T"][~acc], ref["T"][~acc])
    np.testing.assert_array_equal(zer["rY"][:, ~acc], ref["rY"][:, ~acc])
    assert np.max(np.abs(zer["T"][acc] - ref["T"][acc]) / ref["T"][acc]) <= 1e-4


def test_load_rejects_bad_networks(gpu):
    from pelelmex_amd.reactor import Reactor, ReactorError
    r = Reactor.create("ReactorCvode", "drm19", KEYS)
    r.init(2, 1)
    N = r.nsp
    with pytest.raises(ReactorError):       # wrong input width
        r.load_surrogate([N + 1, 8, N + 1], np.zeros((N + 1) * 8 + 8 + 8 * (N + 1) + N + 1), np.ones((2, N + 1)),
                         np.ones((3, N + 1)))
    with pytest.raises(ReactorError):       # zero standard deviation
        xn = np.ones((2, N + 2))
        xn[1, 3] = 0.0
        r.load_surrogate([N + 2, 8, N + 1], np.zeros((N + 2) * 8 + 8 + 8 * (N + 1) + N + 1), xn, np.ones((3, N + 1)))
    r.close()


def _batch_boxes():
    # mixed sizes, dt_react (the network saw 5e-6 only) and a mask
    specs = [(32, 8, 4, 5e-6, 51), (16, 16, 2, 2e-6, 52), (8, 8, 8, 5e-6, 53), (64, 2, 2, 4e-6, 54)]
    boxes = []
    for nx, ny, nz, dt, seed in specs:
        b = W.ch4_air_flame_box(nx=nx, ny=ny, nz=nz, seed=seed, dt=dt)
        if seed == 52:
            b["mask"][::7] = -1
        boxes.append(b)
    return boxes


def test_batch_with_learned_map_matches_single_boxes(gpu):
    """pr_react_batch with a learned map: every box's cells go through the map (one network pass
    over all boxes), the rejected cells of all boxes through one BDF launch; box by box the bits
    equal pr_react_box with the same map, and the accepted / rejected / audited counts add up."""
    import torch
    from pelelmex_amd import surrogate as sg
    from pelelmex_amd.reactor import Reactor, box_tensors
    model = _model()
    boxes = _batch_boxes()
    keys = _keys(model, **{"surrogate.audit_frac": 0.05})
    r = Reactor.create("ReactorCvode", "drm19", keys)
    r.init(2, 1)
    sg.attach(r, model)
    base = [box_tensors(b, gpu) for b in boxes]
    singles, acc, rej, aud = [], 0, 0, 0
    for b, t0 in zip(boxes, base):
        t = {k: v.clone() for k, v in t0.items()}
        status = torch.full((b["T"].size,), 7, dtype=torch.int32, device=gpu)
        nf, _, _ = r.react(t["rY"], t["rYsrc"], t["T"], t["rH"], t["rHsrc"], t["FC"], t["mask"], b["dt"], 0.0,
                           status=status)
        assert nf == 0
        st = r.stats()
        acc, rej, aud = acc + st["sur_accepted"], rej + st["sur_rejected"], aud + st["sur_audited"]
        singles.append((t, status))
    assert acc > 0 and rej > 0
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
    assert (st["sur_accepted"], st["sur_rejected"], st["sur_audited"]) == (acc, rej, aud)
    assert st["audit_max_dT_rel"] <= 1e-4
    for (ts, ss), tb in zip(singles, batch):
        for k in ("rY", "T", "rH", "FC"):
            np.testing.assert_array_equal(tb[k].cpu().numpy(), ts[k].cpu().numpy())
        np.testing.assert_array_equal(tb["status"].cpu().numpy(), ss.cpu().numpy())


def test_batch_with_learned_map_graph_replay(gpu):
    """The learned map + batched fallback captured in one hipGraph (torch.cuda.graph around
    pr_react_batch_enqueue) and replayed three times, inputs restored eagerly and the stream
    synchronized between replays: every replay gives the eager call's bits and counts."""
    import torch
    from pelelmex_amd import surrogate as sg
    from pelelmex_amd.reactor import Reactor, box_tensors
    model = _model()
    boxes = _batch_boxes()
    r = Reactor.create("ReactorCvode", "drm19", _keys(model, **{"surrogate.audit_frac": 0.05}))
    r.init(2, 1)
    sg.attach(r, model)
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
    assert r.react_batch(work) == 0  # eager reference; sizes every buffer the capture uses
    st0 = r.stats()
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
        st = r.stats()
        assert (st["sur_accepted"], st["sur_rejected"], st["sur_audited"]) == \
            (st0["sur_accepted"], st0["sur_rejected"], st0["sur_audited"])
        for t, rf in zip(work, ref):
            for k in rf:
                np.testing.assert_array_equal(t[k].cpu().numpy(), rf[k].cpu().numpy(), err_msg=f"replay {rep} {k}")
    r.close()
