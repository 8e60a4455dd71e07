"""Algorithmic FLOP and byte model of the reactor, for roofline reporting.

Per cell-advance the integrator performs, in the kernel's own counters,
  nfe right-hand sides, nje Jacobians, nsetups LU factorizations of M = I - gamma J,
  nni linear solves (Newton iterations) and nst accepted steps,
so the algorithmic work of a react call is

  F = nfe F_rhs + nje F_jac + nsetups F_lu + nni F_newton + nst F_step.

The static counts below follow the kernels' expression trees with add / sub / mul / div /
compare-free ops = 1 flop, fma = 2 flops and exp / log = 1 flop each (a conservative count:
the transcendental expansions are not charged). The enthalpy inversion in the right-hand
side is charged for 2 Newton iterations. Memory: the algorithmic HBM traffic of a react call
is the state in and out, (3 N + 6) 8 + 4 bytes per cell (rhoY, sources, T, rho h and its
source, FC, mask).
"""
from __future__ import annotations

from .mechanisms import load

NASA_HCP = 18          # cp/R (4 fma) + h/RT (5 fma)
GATHER_BLOCK = 8       # PR_GATHER_BLOCK of csrc/mech_common.h: block sums of the production rates
NASA_FULL = 35         # + s/R (4 fma, mul, add) + g (sub) + dcp (3 fma)
HY2T_ITERS = 2


def _reaction_flops(r, jac):
    nre, npr = len(r.reac), min(len(r.prod), 3)   # the kernel reads 3 product slots at most
    f = 5                                   # k = exp(fma(fma)))
    if r.third is not None:
        f += 1 + 2 * len([e for e in r.eff.values() if e != 1.0])  # [M]
    f += (nre - 1) + (npr - 1)              # concentration products
    if r.rev:
        f += (nre + npr - 1) + 3            # dG, fma, exp
    if r.third is None or r.third == "M":
        f += 3 + (4 if r.third == "M" else 0)
        if jac:
            f += 3 + (nre + npr - 1) + 2 + 4 + (2 if r.third == "M" else 0)
    else:
        f += 5 + 2                          # k0, Pr
        if r.troe is not None:
            f += 25 + (3 if len(r.troe) == 4 else 0)
        f += 6
        if jac:
            f += 3 + 3 + (nre + npr - 1) + 2 + 40
    f += 1                                  # q = fwd - rev
    return f


def static_counts(mech_name):
    m = load(mech_name)
    N, R = m.nsp, m.nr
    n = N + 1
    ninc = 0
    row_work = 0
    combine = 0                             # additions of a species' block sums
    for k, s in enumerate(m.species):
        nk = 0
        for r in m.reactions:
            nu = r.prod.count(s) - r.reac.count(s)
            if nu != 0:
                ninc += 1
                nk += 1
                row_work += 8 + 3 * (len(r.reac) + len(r.prod))
                if r.third is not None:
                    row_work += 3 + 2 * len([e for e in r.eff.values() if e != 1.0])
        combine += max(-(-nk // GATHER_BLOCK) - 1, 0)
    rx_rhs = sum(_reaction_flops(r, False) for r in m.reactions)
    rx_jac = sum(_reaction_flops(r, True) for r in m.reactions)
    F_rhs = (N + 3) + HY2T_ITERS * ((NASA_HCP + 8) * N + 5) + 5 + (NASA_FULL + 5) * N + rx_rhs \
        + 2 * ninc + combine + 2 * N + 4 * N + 2
    F_jac = 5 + (NASA_FULL + 6) * N + N + rx_jac + row_work + (3 * N + 1) * N + 2 * N + 12 * N + 4 \
        + (2 * N + 4) * n
    lu = 2 * n * n
    for k in range(n):
        rows = n - k - 1
        lu += 1 + rows + 2 * rows * rows
    F_solve = 2 * n * (n - 1) + n + n
    F_newton = F_solve + 4 * n + 2 * n + 3 * n + 6
    F_step = n * (4 * 5 // 2 + 2 * 5 + 3 + 9) + 40
    bytes_cell = (3 * N + 6) * 8 + 4
    return dict(F_rhs=F_rhs, F_jac=F_jac, F_lu=lu, F_newton=F_newton, F_step=F_step, bytes_cell=bytes_cell,
                n=n, N=N, R=R, ninc=ninc)


def flops(mech_name, stats):
    c = static_counts(mech_name)
    return (stats["nfe"] * c["F_rhs"] + stats["nje"] * c["F_jac"] + stats["nsetups"] * c["F_lu"]
            + stats["nni"] * c["F_newton"] + stats["nst"] * c["F_step"])


MI355X_FP64_TFLOPS = 78.6
MI355X_HBM_GBS = 8000.0
