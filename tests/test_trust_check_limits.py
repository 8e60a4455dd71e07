"""What the learned map's FP64 trust check can certify on the drm19 flame-like box (CPU, oracle).

The trust check accepts a predicted state y1 when the one-step backward-Euler residual
r = y1 - y0 - dt f(dt, y1) is small in the weighted RMS norm, weights 1 / (trust_rtol |y1| +
trust_atol). Here y1 is the oracle's own rtol = atol = 1e-10 solution - the best prediction a
network could make. On the flame-front and burnt cells of the box the exact solution fails the
check: their chemistry has transients much faster than dt_react = 5e-6 s, so r is of the size of
the increment itself. Only chemically frozen cells (the 300 K slab) pass, and those the
zero-increment predictor (the source step alone) passes too. A residual-certified network
therefore cannot take more BDF work off this box than the zero-increment predictor, however
accurate it is; the trapezoidal residual y1 - y0 - dt/2 (f(0, y0) + f(dt, y1)) fails the same
cells."""
import numpy as np
import pytest

from pelelmex_amd import workloads as W


@pytest.fixture(scope="module")
def cells():
    from oracle import pyoracle as po
    full = W.ch4_air_flame_box()
    nx = full["dims"][0]
    i = np.arange(full["T"].size) % nx
    xc = (i + 0.5) / nx
    slab = np.where(xc < 0.3, 0, np.where(xc < 0.5, 1, 2))
    rng = np.random.default_rng(11)
    idx = np.sort(np.concatenate([rng.choice(np.where(slab == s)[0], 60, replace=False) for s in range(3)]))
    box = W.subsample(full, idx)
    m = po.Mech("drm19")
    ref = m.react(box["rY"], box["rYsrc"], box["T"], box["rH"], box["rHsrc"], box["mask"], box["dt"], max_order=4,
                  nthreads=4)
    assert np.all(ref["status"] == 0)
    return m, box, ref, slab[idx]


def _wrms(r, y1, rtol, atol):
    return float(np.sqrt(np.mean((r / (rtol * np.abs(y1) + atol)) ** 2)))


@pytest.mark.parametrize("rtol,atol", [(3e-3, 3e-11), (1e-2, 1e-9)])
def test_exact_solution_fails_the_residual_check_on_reacting_cells(cells, rtol, atol):
    m, box, ref, slab = cells
    dt = box["dt"]
    N = m.nsp
    passed = {"be": np.zeros(slab.size, bool), "trap": np.zeros(slab.size, bool), "zero": np.zeros(slab.size, bool)}
    for c in range(slab.size):
        cell = m.cell(box["rH"][c], box["rHsrc"][c], box["rYsrc"][:, c])
        y0 = np.concatenate([box["rY"][:, c], [box["T"][c]]])
        y1 = np.concatenate([ref["rY"][:, c], [ref["T"][c]]])
        f1 = m.rhs(cell, dt, y1)
        f0 = m.rhs(cell, 0.0, y0)
        passed["be"][c] = _wrms(y1 - y0 - dt * f1, y1, rtol, atol) <= 1.0
        passed["trap"][c] = _wrms(y1 - y0 - 0.5 * dt * (f0 + f1), y1, rtol, atol) <= 1.0
        # the zero-increment prediction: the source step alone, T from the new rho h
        yz = y0.copy()
        yz[:N] = y0[:N] + dt * box["rYsrc"][:, c]
        rho = yz[:N].sum()
        yz[N] = m.hy2t((box["rH"][c] + dt * box["rHsrc"][c]) / rho, yz[:N] / rho, y0[N])[0]
        passed["zero"][c] = _wrms(yz - y0 - dt * m.rhs(cell, dt, yz), yz, rtol, atol) <= 1.0
    cold, front, burnt = (slab == 0), (slab == 1), (slab == 2)
    for kind in ("be", "trap"):
        assert passed[kind][front].mean() == 0.0, kind            # no flame-front cell
        assert passed[kind][burnt].mean() <= 0.1, kind            # at most a few burnt cells
        assert passed[kind][cold].mean() >= 0.9, kind             # the frozen slab
    assert passed["zero"][cold].mean() >= 0.9                     # which the zero increment passes too
    fc = ref["FC"]
    work = lambda ok: fc[ok].sum() / fc.sum()
    assert work(passed["be"]) <= work(passed["zero"]) + 0.02      # so the exact map removes ~no more BDF work


def test_bdf_work_at_the_metric_accuracy(cells):
    """How much cheaper the box gets when 1e-4 in T is all that is asked: the same BDF at
    rtol 1e-6 .. 1e-4 (atol 1e-10) needs 1/3.5 .. 1/4 of the right-hand-side evaluations of
    rtol = atol = 1e-10, with max |dT|/T of 1e-4 .. 5e-4 against it. Resolving the fast
    transients of the front and burnt cells costs most of the work at any tolerance, which bounds
    what a certified shortcut of the dt_react map can save on this box."""
    m, box, ref, slab = cells
    for rtol, ratio_max in ((1e-6, 5.0), (1e-4, 5.0)):
        out = m.react(box["rY"], box["rYsrc"], box["T"], box["rH"], box["rHsrc"], box["mask"], box["dt"], rtol=rtol,
                      atol=1e-10, max_order=4, nthreads=4)
        assert np.all(out["status"] == 0)
        ratio = ref["FC"].sum() / out["FC"].sum()
        dT = np.max(np.abs(out["T"] - ref["T"]) / ref["T"])
        assert 2.0 < ratio < ratio_max, (rtol, ratio)
        assert dT < 1e-3
