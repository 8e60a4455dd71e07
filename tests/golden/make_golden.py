"""Generate the golden reactor fixtures in tests/golden/ (run in the dev container, commit the .npz).

For a few cells of each workload the reactor ODE (the oracle's right-hand side, i.e. the
constant-pressure rhoY/T system with rho h(t) = rho h0 + t Q_h) is integrated over dt_react with
two independent integrators from scipy at tight tolerance:
  * solve_ivp(method="Radau"), rtol = 1e-13, atol = 1e-20, with the analytic Jacobian;
  * scipy.integrate.ode("vode", method="bdf") (Fortran VODE, CVODE's predecessor), rtol = 1e-13.
The fixtures store inputs and both results; tests compare the oracle's BDF at rtol = atol = 1e-10
(and, on the GPU, the HIP reactor) against them.

Usage: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np
from scipy.integrate import ode, solve_ivp

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import pyoracle as po  # noqa: E402
from pelelmex_amd import workloads as W  # noqa: E402


def integrate(m, box, c):
    N = m.nsp
    y0 = np.append(box["rY"][:, c], box["T"][c])
    cell = m.cell(box["rH"][c], box["rHsrc"][c], box["rYsrc"][:, c])
    f = lambda t, y: m.rhs(cell, t, y)
    jac = lambda t, y: m.jac(cell, t, y)
    dt = box["dt"]
    r = solve_ivp(f, (0.0, dt), y0, method="Radau", rtol=1e-13, atol=1e-20, jac=jac)
    assert r.success, r.message
    yr = r.y[:, -1]
    v = ode(f, jac).set_integrator("vode", method="bdf", rtol=1e-13, atol=1e-20, nsteps=200000,
                                   with_jacobian=True, order=5)
    v.set_initial_value(y0, 0.0)
    yv = v.integrate(dt)
    assert v.successful()
    return yr, yv


def make(name, box, idx):
    m = po.Mech(box["mech"])
    sub = W.subsample(box, idx)
    out_r, out_v = [], []
    for c in range(len(idx)):
        yr, yv = integrate(m, sub, c)
        out_r.append(yr)
        out_v.append(yv)
    path = os.path.join(HERE, f"{name}.npz")
    np.savez_compressed(path, mech=box["mech"], idx=np.asarray(idx), rY=sub["rY"], rYsrc=sub["rYsrc"],
                        T=sub["T"], rH=sub["rH"], rHsrc=sub["rHsrc"], dt=sub["dt"],
                        radau=np.array(out_r).T, vode=np.array(out_v).T)
    print(path, len(idx))


if __name__ == "__main__":
    b1 = W.h2_air_box()
    make("h2_air_32_cells", b1, np.random.default_rng(11).choice(b1["T"].size, 24, replace=False))
    b2 = W.ch4_air_flame_box()
    idx = np.concatenate([np.arange(0, 64, 3), np.random.default_rng(12).choice(b2["T"].size, 8, replace=False)])
    make("ch4_air_flame_64_cells", b2, idx)
