"""Mechanism generator: CHEMKIN-style spec (a Python module) -> C header of constant tables.

The emitted header (``pelelmex_amd/csrc/mech/<name>.h``) is plain C99 so that the CPU oracle
(gcc) and the HIP library (hipcc) read the same data. It holds:

* species names, molecular weights (g/mol) from standard atomic weights, element counts;
* NASA-7 fits per species, packed as ``[T_mid, high a1..a7, low a1..a7]``;
* per reaction: type (0 elementary, 1 third body, 2 Lindemann falloff, 3 Troe falloff),
  reversibility, reactant / product molecule slots (a stoichiometric coefficient of 2 fills two
  slots), ``ln A``, ``beta``, activation temperature ``Ea/R`` (K), the low-pressure triple for
  falloff reactions, Troe parameters, and collision efficiencies stored as a default plus
  sparse ``(species, eff - default)`` corrections;
* the species -> (reaction, net stoichiometry) incidence lists in CSR form, used by the
  species-parallel gather on the GPU.

This mirrors what PelePhysics' CEPTR writes into ``mechanism.H`` (named in
CMake/BuildPelePhysicsLib.cmake:8,126-131 of the reference) in content, not in form: the
reference header is generated straight-line code, this one is data that kernels interpret.

Run ``python -m pelelmex_amd.mechanisms.gen`` to regenerate every header.
"""
from __future__ import annotations

import importlib
import math
import os
import re
import sys

ATOMIC_WEIGHTS = {"H": 1.008, "C": 12.011, "O": 15.999, "N": 14.007, "Ar": 39.95}
R_CAL = 8.314462618 / 4.184  # cal / (mol K)
MECHANISMS = ["lidryer", "drm19", "syn53", "dodecane_lu"]

_HERE = os.path.dirname(os.path.abspath(__file__))
OUT_DIR = os.path.join(os.path.dirname(_HERE), "csrc", "mech")


class Reaction:
    def __init__(self, eq, A, b, E):
        self.eq = eq
        self.A, self.b, self.E = A, b, E
        self.reac, self.prod = [], []  # molecule slots (species names, repeated)
        self.rev = True
        self.third = None  # None | "M" | "(+M)"
        self.low = None
        self.troe = None
        self.eff = {}
        self.dup = False


def _side(text):
    slots = []
    for tok in text.split("+"):
        tok = tok.strip()
        if not tok:
            continue
        m = re.match(r"^(\d+)(.*)$", tok)
        n, name = (int(m.group(1)), m.group(2)) if m else (1, tok)
        slots += [name] * n
    return slots


def parse_reactions(text):
    out = []
    for raw in text.strip().splitlines():
        line = raw.split("!")[0].strip()
        if not line:
            continue
        up = line.upper()
        if up.startswith("LOW/"):
            out[-1].low = [float(x) for x in line.split("/")[1].split()]
            continue
        if up.startswith("TROE/"):
            out[-1].troe = [float(x) for x in line.split("/")[1].split()]
            continue
        if up.startswith("DUPLICATE"):
            out[-1].dup = True
            continue
        if "/" in line and "=" not in line:
            toks = line.split("/")
            for i in range(0, len(toks) - 1, 2):
                name = toks[i].strip()
                if name:
                    out[-1].eff[name] = float(toks[i + 1])
            continue
        parts = line.split()
        eq, A, b, E = parts[0], float(parts[1]), float(parts[2]), float(parts[3])
        r = Reaction(eq, A, b, E)
        if "<=>" in eq:
            lhs, rhs = eq.split("<=>")
        elif "=>" in eq:
            lhs, rhs = eq.split("=>")
            r.rev = False
        else:
            lhs, rhs = eq.split("=")
        if "(+M)" in lhs:
            r.third = "(+M)"
            lhs, rhs = lhs.replace("(+M)", ""), rhs.replace("(+M)", "")
        sl, sr = _side(lhs), _side(rhs)
        if "M" in sl:
            r.third = "M"
            sl = [s for s in sl if s != "M"]
            sr = [s for s in sr if s != "M"]
        r.reac, r.prod = sl, sr
        out.append(r)
    return out


def nasa_eval(coef, T):
    """cp/R, h/RT, s/R of one NASA-7 range."""
    a = coef
    cp = a[0] + T * (a[1] + T * (a[2] + T * (a[3] + T * a[4])))
    h = a[0] + T * (a[1] / 2 + T * (a[2] / 3 + T * (a[3] / 4 + T * a[4] / 5))) + a[5] / T
    s = a[0] * math.log(T) + T * (a[1] + T * (a[2] / 2 + T * (a[3] / 3 + T * a[4] / 4))) + a[6]
    return cp, h, s


def load(name):
    mod = importlib.import_module(f"pelelmex_amd.mechanisms.{name}")
    rx = parse_reactions(mod.REACTIONS)
    return mod, rx


def validate(mod, rx):
    """Element balance of every reaction and NASA-7 continuity at T_mid; returns a report."""
    sp = mod.SPECIES
    rep = {"continuity": {}, "nreact": len(rx)}
    for r in rx:
        for s in r.reac + r.prod + list(r.eff):
            if s not in sp:
                raise ValueError(f"{mod.NAME}: unknown species {s} in {r.eq}")
        for el in ATOMIC_WEIGHTS:
            lhs = sum(mod.COMPOSITION[s].get(el, 0) for s in r.reac)
            rhs = sum(mod.COMPOSITION[s].get(el, 0) for s in r.prod)
            if lhs != rhs:
                raise ValueError(f"{mod.NAME}: element {el} unbalanced in {r.eq}")
        if r.third == "(+M)" and r.low is None:
            raise ValueError(f"{mod.NAME}: falloff without LOW in {r.eq}")
    for s in sp:
        tl, tm, th, hi, lo = mod.THERMO[s]
        a = nasa_eval(hi, tm)
        b = nasa_eval(lo, tm)
        rep["continuity"][s] = max(abs(x - y) / max(abs(x), 1e-30) for x, y in zip(a, b))
    return rep


def _fmt(x):
    return repr(float(x)) if x == x else "0.0"


def emit(name):
    mod, rx = load(name)
    validate(mod, rx)
    sp = mod.SPECIES
    idx = {s: i for i, s in enumerate(sp)}
    nsp, nr = len(sp), len(rx)
    elements = [e for e in ATOMIC_WEIGHTS if any(mod.COMPOSITION[s].get(e, 0) for s in sp)]
    W = [sum(ATOMIC_WEIGHTS[e] * n for e, n in mod.COMPOSITION[s].items()) for s in sp]
    P = f"pr_{mod.NAME.lower()}"
    L = []
    w = L.append
    w(f"/* Generated by pelelmex_amd/mechanisms/gen.py from mechanisms/{name}.py: do not edit. */")
    w(f"#ifndef PR_MECH_{mod.NAME.upper()}_H")
    w(f"#define PR_MECH_{mod.NAME.upper()}_H")
    w('#include "../mech_common.h"')
    w(f"#define PR_{mod.NAME.upper()}_NSP {nsp}")
    w(f"#define PR_{mod.NAME.upper()}_NR {nr}")
    w(f"static const char* const {P}_species[{nsp}] = {{" + ", ".join(f'"{s}"' for s in sp) + "};")
    w(f"static const char* const {P}_elements[{len(elements)}] = {{"
      + ", ".join(f'"{e}"' for e in elements) + "};")
    w(f"static const double {P}_W[{nsp}] = {{" + ", ".join(_fmt(x) for x in W) + "};")
    w(f"static const int {P}_elem[{nsp * len(elements)}] = {{"
      + ", ".join(str(mod.COMPOSITION[s].get(e, 0)) for s in sp for e in elements) + "};")
    rows = []
    for s in sp:
        tl, tm, th, hi, lo = mod.THERMO[s]
        rows.append("  " + ", ".join(_fmt(x) for x in [tm] + list(hi) + list(lo)))
    w(f"static const double {P}_nasa[{nsp * 15}] = {{\n" + ",\n".join(rows) + "};")

    rtype, rev, nreac, reac, nprod, prod, arr, low, troe, ntroe = ([] for _ in range(10))
    eff_dflt, eff_start, eff_sp, eff_val = [], [0], [], []
    for r in rx:
        if r.third is None:
            t = 0
        elif r.third == "M":
            t = 1
        else:
            t = 3 if r.troe is not None else 2
        rtype.append(t)
        rev.append(1 if r.rev else 0)
        # up to 3 molecules per side are stored (the slots the reverse rate reads); an
        # irreversible reaction may list up to 6 products, which enter the production rates
        # through the incidence lists only
        if len(r.reac) > 3 or len(r.prod) > (3 if r.rev else 6):
            raise ValueError(f"too many molecules on a side of {r.eq}")
        nreac.append(len(r.reac))
        reac += [idx[s] for s in r.reac] + [0] * (3 - len(r.reac))
        ps = r.prod[:3]
        nprod.append(len(ps))
        prod += [idx[s] for s in ps] + [0] * (3 - len(ps))
        arr += [math.log(r.A), r.b, r.E / R_CAL]
        if r.low is not None:
            low += [math.log(r.low[0]), r.low[1], r.low[2] / R_CAL]
        else:
            low += [0.0, 0.0, 0.0]
        if r.troe is not None:
            tr = list(r.troe) + [0.0] * (4 - len(r.troe))
            troe += tr
            ntroe.append(len(r.troe))
        else:
            troe += [0.0] * 4
            ntroe.append(0)
        if t > 0:
            eff_dflt.append(1.0)
            for s, e in r.eff.items():
                if e != 1.0:
                    eff_sp.append(idx[s])
                    eff_val.append(e - 1.0)
        else:
            eff_dflt.append(0.0)
        eff_start.append(len(eff_sp))

    def arr_int(nm, v):
        w(f"static const int {P}_{nm}[{max(len(v), 1)}] = {{" + (", ".join(map(str, v)) or "0") + "};")

    def arr_dbl(nm, v):
        w(f"static const double {P}_{nm}[{max(len(v), 1)}] = {{"
          + (", ".join(_fmt(x) for x in v) or "0.0") + "};")

    arr_int("rtype", rtype)
    arr_int("rev", rev)
    arr_int("nreac", nreac)
    arr_int("reac", reac)
    arr_int("nprod", nprod)
    arr_int("prod", prod)
    arr_dbl("arr", arr)
    arr_dbl("low", low)
    arr_dbl("troe", troe)
    arr_int("ntroe", ntroe)
    arr_dbl("eff_dflt", eff_dflt)
    arr_int("eff_start", eff_start)
    arr_int("eff_sp", eff_sp)
    arr_dbl("eff_val", eff_val)

    # species -> reaction incidence with net stoichiometric coefficient
    inc_start, inc_r, inc_nu = [0], [], []
    for k, s in enumerate(sp):
        for j, r in enumerate(rx):
            nu = r.prod.count(s) - r.reac.count(s)
            if nu != 0:
                inc_r.append(j)
                inc_nu.append(nu)
        inc_start.append(len(inc_r))
    arr_int("inc_start", inc_start)
    arr_int("inc_r", inc_r)
    arr_int("inc_nu", inc_nu)
    w(f"static const pr_mech_t {P} = {{")
    w(f'  "{mod.NAME}", {nsp}, {nr}, {len(elements)}, {len(eff_sp)}, {len(inc_r)},')
    w(f"  {P}_species, {P}_elements, {P}_W, {P}_nasa, {P}_elem,")
    w(f"  {P}_rtype, {P}_rev, {P}_nreac, {P}_reac, {P}_nprod, {P}_prod,")
    w(f"  {P}_arr, {P}_low, {P}_troe, {P}_ntroe,")
    w(f"  {P}_eff_dflt, {P}_eff_start, {P}_eff_sp, {P}_eff_val,")
    w(f"  {P}_inc_start, {P}_inc_r, {P}_inc_nu}};")
    w("#endif")
    os.makedirs(OUT_DIR, exist_ok=True)
    path = os.path.join(OUT_DIR, f"{name}.h")
    with open(path, "w") as f:
        f.write("\n".join(L) + "\n")
    return path


def main(argv=None):
    for name in (argv or MECHANISMS):
        mod, rx = load(name)
        rep = validate(mod, rx)
        worst = max(rep["continuity"].items(), key=lambda kv: kv[1])
        print(f"{name}: {len(mod.SPECIES)} species, {len(rx)} reactions, "
              f"worst T_mid mismatch {worst[0]} {worst[1]:.2e} -> {emit(name)}")


if __name__ == "__main__":
    main(sys.argv[1:])
