"""chemTest harness: the (state, forcing) pairs PeleLMeX's `evaluate` writes for the chemistry
(`pltEvaluate` with the "chemTest" variable, Source/PeleLMeX_Evaluate.cpp:188-207 and
Source/PeleLMeX_Setup.cpp:1232-1247: components rhoY(<species>) x N, rhoH, Temp,
F_rhoY(<species>) x N, F_rhoH, MKS; the header time holds dt, Source/PeleLMeX_Evaluate.cpp:86-93)
advanced over dt by this reactor, every FAB of the chosen levels in one batched launch, with the
unit conversions of advanceChemistry (Source/PeleLMeX_Reactions.cpp:68-79 in, 110-120 out).

    python -m pelelmex_amd.chemtest <pltEvaluate> --mech drm19 [--out <plt>]

The result is a plotfile of the same boxes with rhoY(<species>), rhoH, Temp after the advance
and FC (right-hand-side evaluations per cell): reference pairs for regression fixtures and
surrogate training from production states.
"""
from __future__ import annotations

import argparse

from .plotfile import Fab, PlotFile, read_plotfile, write_plotfile


def chemtest_names(species):
    """The component names of the chemTest evaluate variable (Source/PeleLMeX_Setup.cpp:1232-1247)."""
    return ([f"rhoY({s})" for s in species] + ["rhoH", "Temp"] + [f"F_rhoY({s})" for s in species] +
            ["F_rhoH"])


def react_plotfile(pf: PlotFile, reactor, species, dt=None, levels=None, device="cuda"):
    """Advances every FAB of `pf` (a chemTest plotfile) by dt (default: the header time) with
    `reactor` (a pelelmex_amd.reactor.Reactor); returns (PlotFile of rhoY, rhoH, Temp, FC,
    number of failed cells)."""
    import torch
    idx = {n: i for i, n in enumerate(pf.names)}
    try:
        cY = [idx[f"rhoY({s})"] for s in species]
        cH, cT = idx["rhoH"], idx["Temp"]
        cFY = [idx[f"F_rhoY({s})"] for s in species]
        cFH = idx["F_rhoH"]
    except KeyError as e:
        raise ValueError(f"not a chemTest plotfile for this mechanism: missing {e}") from None
    dt = float(pf.time if dt is None else dt)
    levs = range(len(pf.levels)) if levels is None else levels
    boxes, where = [], []
    for lev in levs:
        for b, fb in enumerate(pf.levels[lev]):
            d = torch.as_tensor(fb.data, dtype=torch.float64)
            # MKS -> CGS as advanceChemistry does before react
            t = dict(rY=(d[cY] * 1.0e-3).to(device).contiguous(), rYsrc=(d[cFY] * 1.0e-3).to(device).contiguous(),
                     T=d[cT].to(device).contiguous(), rH=(d[cH] * 10.0).to(device).contiguous(),
                     rHsrc=(d[cFH] * 10.0).to(device).contiguous())
            t["FC"] = torch.zeros_like(t["T"])
            t["dt"] = dt
            boxes.append(t)
            where.append((lev, b))
    nfail = reactor.react_batch(boxes) if boxes else 0
    names = [f"rhoY({s})" for s in species] + ["rhoH", "Temp", "FC"]
    out = PlotFile(names, dt, pf.prob_lo, pf.prob_hi, pf.ref_ratio, pf.domains, pf.dx,
                   [[] for _ in pf.levels])
    for (lev, b), t in zip(where, boxes):
        fb = pf.levels[lev][b]
        # CGS -> MKS as advanceChemistry does after react
        data = torch.cat([t["rY"] * 1.0e3, (t["rH"] * 0.1)[None], t["T"][None], t["FC"][None]]).cpu().numpy()
        out.levels[lev].append(Fab(fb.lo, fb.hi, data))
    return out, nfail


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("plotfile")
    ap.add_argument("--mech", default="drm19")
    ap.add_argument("--out", default=None)
    ap.add_argument("--dt", type=float, default=None)
    ap.add_argument("--rtol", type=float, default=1e-10)
    ap.add_argument("--atol", type=float, default=1e-10)
    ap.add_argument("--max-order", type=int, default=4)
    a = ap.parse_args(argv)
    from .mechanisms import load
    from .reactor import Reactor
    species = load(a.mech).species
    r = Reactor.create("ReactorCvode", a.mech, {"ode.rtol": a.rtol, "ode.atol": a.atol, "cvode.max_order": a.max_order})
    r.init(2, 1)
    pf = read_plotfile(a.plotfile)
    out, nfail = react_plotfile(pf, r, species, dt=a.dt)
    r.close()
    ncells = sum(f.data[0].size for lev in out.levels for f in lev)
    print(f"{ncells} cells advanced over dt = {out.time:g} s, {nfail} failed")
    if a.out:
        write_plotfile(a.out, out)
    return 0 if nfail == 0 else 1


if __name__ == "__main__":
    raise SystemExit(main())
