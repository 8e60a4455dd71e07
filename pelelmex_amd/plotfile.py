"""AMReX native plotfiles (HyperCLaw-V1.1 header, VisMF v1 `Cell_H`, FAB data files), the
format PeleLMeX's `evaluate` writes (`amrex::WriteMultiLevelPlotfile`,
Source/PeleLMeX_Evaluate.cpp:86-93) - read into numpy, and written for fixtures.

Layout restated from the format itself (AMReX is not part of the reference tree):
  <plt>/Header       text: "HyperCLaw-V1.1", ncomp, the ncomp names, dim, time, finest level,
                     prob_lo, prob_hi, refinement ratios, level domains, level steps, cell sizes
                     per level, coordinate system, boundary width, then per level
                     "lev nboxes time", its step, the boxes' physical bounds and "Level_<l>/Cell"
  <plt>/Level_<l>/Cell_H   "1" (VisMF version), how, ncomp, nghost, the BoxArray
                     "(n 0 ((lo) (hi) (type)) ... )", n, "FabOnDisk: <file> <offset>" per box,
                     then per-box component minima and maxima
  <plt>/Level_<l>/Cell_D_xxxxx   per FAB: one header line
                     "FAB ((8, (64 11 52 0 1 12 0 1023)),(8, <byte order>))((lo) (hi) (type)) ncomp"
                     then ncomp x cells float64 values, x fastest, component slowest; byte order
                     (8 7 6 5 4 3 2 1) = little-endian, (1 2 3 4 5 6 7 8) = big-endian.
"""
from __future__ import annotations

import os
import re
from dataclasses import dataclass, field

import numpy as np

_BOX = re.compile(r"\(\((-?\d+),(-?\d+),(-?\d+)\) \((-?\d+),(-?\d+),(-?\d+)\) \((\d+),(\d+),(\d+)\)\)")
_FAB = re.compile(r"FAB \(\(8, \(([^)]*)\)\),\(8, \(([^)]*)\)\)\)" + _BOX.pattern + r" (\d+)")
_LITTLE = "8 7 6 5 4 3 2 1"
_BIG = "1 2 3 4 5 6 7 8"
_FORMAT = "64 11 52 0 1 12 0 1023"


@dataclass
class Fab:
    lo: tuple
    hi: tuple
    data: np.ndarray          # [ncomp, nz, ny, nx]


@dataclass
class PlotFile:
    names: list
    time: float
    prob_lo: tuple = (0.0, 0.0, 0.0)
    prob_hi: tuple = (1.0, 1.0, 1.0)
    ref_ratio: list = field(default_factory=list)
    domains: list = field(default_factory=list)      # per level (lo, hi)
    dx: list = field(default_factory=list)           # per level (dx, dy, dz)
    levels: list = field(default_factory=list)       # per level: list of Fab

    def comp(self, name):
        return self.names.index(name)


def _box_str(lo, hi):
    return f"(({lo[0]},{lo[1]},{lo[2]}) ({hi[0]},{hi[1]},{hi[2]}) (0,0,0))"


def _parse_box(m, off=0):
    g = [int(m.group(off + i)) for i in range(1, 10)]
    if any(g[6:9]):
        raise ValueError("only cell-centred boxes are supported")
    return tuple(g[0:3]), tuple(g[3:6])


def read_plotfile(path: str) -> PlotFile:
    with open(os.path.join(path, "Header")) as f:
        lines = f.read().split("\n")
    it = iter(lines)
    ver = next(it).strip()
    if ver != "HyperCLaw-V1.1":
        raise ValueError(f"unsupported plotfile version {ver!r}")
    ncomp = int(next(it))
    names = [next(it).strip() for _ in range(ncomp)]
    if int(next(it)) != 3:
        raise ValueError("only 3-D plotfiles are supported")
    time = float(next(it))
    finest = int(next(it))
    prob_lo = tuple(float(x) for x in next(it).split())
    prob_hi = tuple(float(x) for x in next(it).split())
    ref_ratio = [int(x) for x in next(it).split()]
    dom_line = next(it)
    domains = [_parse_box(m) for m in _BOX.finditer(dom_line)]
    next(it)  # level steps
    dx = [tuple(float(x) for x in next(it).split()) for _ in range(finest + 1)]
    next(it)  # coordinate system
    next(it)  # boundary width
    pf = PlotFile(names, time, prob_lo, prob_hi, ref_ratio, domains, dx)
    for lev in range(finest + 1):
        lv, nbox, _t = next(it).split()
        if int(lv) != lev:
            raise ValueError("level records out of order")
        next(it)  # level step
        for _ in range(int(nbox) * 3):
            next(it)  # physical box bounds
        rel = next(it).strip()
        pf.levels.append(_read_level(os.path.join(path, rel + "_H"), os.path.dirname(os.path.join(path, rel)), ncomp))
    return pf


def _read_level(cell_h: str, ldir: str, ncomp: int):
    with open(cell_h) as f:
        text = f.read()
    entries = re.findall(r"FabOnDisk: (\S+) (\d+)", text)
    fabs = []
    for fname, off in entries:
        with open(os.path.join(ldir, fname), "rb") as g:
            g.seek(int(off))
            head = b""
            while not head.endswith(b"\n"):
                ch = g.read(1)
                if not ch:
                    raise ValueError("truncated FAB header")
                head += ch
            m = _FAB.match(head.decode().strip())
            if not m:
                raise ValueError(f"unrecognised FAB header {head!r}")
            if " ".join(m.group(1).split()) != _FORMAT:
                raise ValueError("only IEEE float64 FAB data is supported")
            order = " ".join(m.group(2).split())
            if order not in (_LITTLE, _BIG):
                raise ValueError(f"unsupported byte order {order}")
            lo, hi = _parse_box(m, 2)
            nc = int(m.group(12))
            if nc != ncomp:
                raise ValueError("FAB component count differs from the header")
            shape = (nc, hi[2] - lo[2] + 1, hi[1] - lo[1] + 1, hi[0] - lo[0] + 1)
            n = int(np.prod(shape))
            raw = g.read(8 * n)
            if len(raw) != 8 * n:
                raise ValueError("truncated FAB data")
            data = np.frombuffer(raw, dtype="<f8" if order == _LITTLE else ">f8").astype(np.float64).reshape(shape)
        fabs.append(Fab(lo, hi, data))
    return fabs


def write_plotfile(path: str, pf: PlotFile) -> None:
    """Writes `pf` (every level's FABs in one Cell_D_00000 file, little-endian float64)."""
    os.makedirs(path, exist_ok=True)
    finest = len(pf.levels) - 1
    ncomp = len(pf.names)
    L = ["HyperCLaw-V1.1", str(ncomp), *pf.names, "3", repr(float(pf.time)), str(finest),
         " ".join(repr(float(x)) for x in pf.prob_lo) + " ", " ".join(repr(float(x)) for x in pf.prob_hi) + " ",
         " ".join(str(r) for r in pf.ref_ratio) + " ",
         " ".join(_box_str(lo, hi) for lo, hi in pf.domains) + " ",
         " ".join("0" for _ in pf.levels) + " "]
    L += [" ".join(repr(float(x)) for x in d) + " " for d in pf.dx]
    L += ["0", "0"]
    for lev, fabs in enumerate(pf.levels):
        L.append(f"{lev} {len(fabs)} {float(pf.time)!r}")
        L.append("0")
        dxl = pf.dx[lev]
        for fb in fabs:
            for d in range(3):
                a = pf.prob_lo[d] + fb.lo[d] * dxl[d]
                b = pf.prob_lo[d] + (fb.hi[d] + 1) * dxl[d]
                L.append(f"{a!r} {b!r}")
        L.append(f"Level_{lev}/Cell")
        ldir = os.path.join(path, f"Level_{lev}")
        os.makedirs(ldir, exist_ok=True)
        offsets = []
        with open(os.path.join(ldir, "Cell_D_00000"), "wb") as g:
            for fb in fabs:
                offsets.append(g.tell())
                d = np.ascontiguousarray(fb.data, dtype="<f8")
                if d.shape[0] != ncomp:
                    raise ValueError("FAB component count differs from the names")
                g.write(f"FAB ((8, ({_FORMAT})),(8, ({_LITTLE}))){_box_str(fb.lo, fb.hi)} {ncomp}\n".encode())
                g.write(d.tobytes())
        H = ["1", "0", str(ncomp), "0", f"({len(fabs)} 0"]
        H += [_box_str(fb.lo, fb.hi) for fb in fabs]
        H += [")", str(len(fabs))]
        H += [f"FabOnDisk: Cell_D_00000 {o}" for o in offsets]
        H += ["", f"{len(fabs)},{ncomp}"]
        H += ["".join(f"{v!r}," for v in fb.data.reshape(ncomp, -1).min(1)) for fb in fabs]
        H += ["", f"{len(fabs)},{ncomp}"]
        H += ["".join(f"{v!r}," for v in fb.data.reshape(ncomp, -1).max(1)) for fb in fabs]
        with open(os.path.join(ldir, "Cell_H"), "w") as f:
            f.write("\n".join(H) + "\n")
    with open(os.path.join(path, "Header"), "w") as f:
        f.write("\n".join(L) + "\n")
