"""Generate the golden reactor fixtures in tests/golden/ (run in the dev container, commit the .npz).

For a few hundred cells of each workload the reactor ODE (the oracle's right-hand side, i.e. the
constant-pressure rhoY/T system with rho h(t) = rho h0 + t Q_h) is integrated over dt_react with
independent integrators from scipy:
  * truth: solve_ivp(method="Radau"), rtol = 1e-13, atol = 1e-20, analytic Jacobian, and
    scipy.integrate.ode("vode", method="bdf") (Fortran VODE, CVODE's predecessor: the same
    Nordsieck BDF, WRMS error test and Newton iteration) at rtol = 1e-13, atol = 1e-20;
  * the reference's tolerances: VODE at rtol = atol = 1e-10 with maximum order 4 and 5 (what
    CVODE run at PeleLMeX's ode.rtol = ode.atol = 1e-10 is compared with), and VODE at
    rtol = 1e-10 with the per-component absolute tolerances atol_i = 1e-10 * typ_i that
    PeleLMeX's typical values produce (Source/PeleLMeX_Utils.cpp:1363-1475, use_typ_vals_chem = 1),
    typ computed over the whole box the cells are drawn from.
Cells: every slab of each box (drm19: one full i-row plus stratified draws from the cold, front
and burnt slabs; LiDryer: a random draw of the 32^3 box).

Usage: python tests/golden/make_golden.py
"""
import os
import sys
from multiprocessing import Pool

import numpy as np
from scipy.integrate import ode, solve_ivp

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import pyoracle as po  # noqa: E402
from pelelmex_amd import workloads as W  # noqa: E402
from pelelmex_amd.reactor import typical_values_chem  # noqa: E402

_JOB = {}


def _vode(f, jac, y0, dt, rtol, atol, order):
    v = ode(f, jac).set_integrator("vode", method="bdf", rtol=rtol, atol=atol, nsteps=500000,
                                   with_jacobian=True, order=order)
    v.set_initial_value(y0, 0.0)
    y = v.integrate(dt)
    assert v.successful()
    return y


def integrate(c):
    sub, atol_typ = _JOB["box"], _JOB["atol_typ"]
    m = po.Mech(sub["mech"])
    y0 = np.append(sub["rY"][:, c], sub["T"][c])
    cell = m.cell(sub["rH"][c], sub["rHsrc"][c], sub["rYsrc"][:, c])
    f = lambda t, y: m.rhs(cell, t, y)
    jac = lambda t, y: m.jac(cell, t, y)
    dt = sub["dt"]
    r = solve_ivp(f, (0.0, dt), y0, method="Radau", rtol=1e-13, atol=1e-20, jac=jac)
    assert r.success, r.message
    return dict(radau=r.y[:, -1], vode=_vode(f, jac, y0, dt, 1e-13, 1e-20, 5),
                vode10_q4=_vode(f, jac, y0, dt, 1e-10, 1e-10, 4), vode10_q5=_vode(f, jac, y0, dt, 1e-10, 1e-10, 5),
                vode10_typ_q4=_vode(f, jac, y0, dt, 1e-10, atol_typ, 4))


def make(name, box, idx):
    typ = typical_values_chem(box["rY"] * 1e3, box["T"])      # the box's state in MKS, as PeleLMeX holds it
    atol_typ = 1e-10 * typ
    sub = W.subsample(box, idx)
    _JOB.update(box=sub, atol_typ=atol_typ)
    with Pool(8) as pool:
        res = pool.map(integrate, range(len(idx)))
    arrays = {k: np.array([r[k] for r in res]).T for k in res[0]}
    path = os.path.join(HERE, f"{name}.npz")
    np.savez_compressed(path, mech=box["mech"], idx=np.asarray(idx), rY=sub["rY"], rYsrc=sub["rYsrc"],
                        T=sub["T"], rH=sub["rH"], rHsrc=sub["rHsrc"], dt=sub["dt"], typ=typ, **arrays)
    print(path, len(idx))


def drm19_cells(box, seed=12):
    """One full i-row (all three slabs) plus 40 cold, 80 front and 80 burnt cells at random."""
    rng = np.random.default_rng(seed)
    nx = box["dims"][0]
    x = (np.arange(box["T"].size) % nx + 0.5) / nx
    pick = [np.arange(nx)]
    for lo, hi, k in ((0.0, 0.3, 40), (0.3, 0.5, 80), (0.5, 1.0, 80)):
        cand = np.flatnonzero((x >= lo) & (x < hi) & (np.arange(x.size) >= nx))
        pick.append(rng.choice(cand, k, replace=False))
    return np.concatenate(pick)


if __name__ == "__main__":
    b1 = W.h2_air_box()
    make("h2_air_cells", b1, np.random.default_rng(11).choice(b1["T"].size, 240, replace=False))
    b2 = W.ch4_air_flame_box()
    make("ch4_air_flame_cells", b2, drm19_cells(b2))
