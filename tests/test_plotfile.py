"""AMReX native plotfile reader / writer (pelelmex_amd.plotfile) and the chemTest harness's
file side (CPU). The header and FAB records follow the plotfile format AMReX writes for
PeleLMeX's `evaluate` (Source/PeleLMeX_Evaluate.cpp:86-93); AMReX itself is not in the reference
tree, so the layout is pinned by a hand-written record of that format (below) and by round
trips, not by a file written by AMReX (parity of the reader with AMReX files: unpinned)."""
import os

import numpy as np
import pytest

from pelelmex_amd.chemtest import chemtest_names
from pelelmex_amd.plotfile import Fab, PlotFile, read_plotfile, write_plotfile


def _pf(rng, ncomp=5, names=None):
    names = names or [f"c{i}" for i in range(ncomp)]
    lv0 = [Fab((0, 0, 0), (7, 3, 1), rng.standard_normal((ncomp, 2, 4, 8))),
           Fab((8, 0, 0), (15, 3, 1), rng.standard_normal((ncomp, 2, 4, 8)))]
    lv1 = [Fab((4, 2, 0), (11, 5, 3), rng.standard_normal((ncomp, 4, 4, 8)))]
    return PlotFile(names, 2.5e-6, (0.0, 0.0, 0.0), (0.016, 0.004, 0.002), [2],
                    [((0, 0, 0), (15, 3, 1)), ((0, 0, 0), (31, 7, 3))],
                    [(1e-3, 1e-3, 1e-3), (5e-4, 5e-4, 5e-4)], [lv0, lv1])


def test_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    pf = _pf(rng)
    write_plotfile(str(tmp_path / "plt"), pf)
    back = read_plotfile(str(tmp_path / "plt"))
    assert back.names == pf.names and back.time == pf.time and back.ref_ratio == [2]
    assert back.domains == pf.domains and back.dx == pf.dx
    for la, lb in zip(pf.levels, back.levels):
        assert len(la) == len(lb)
        for a, b in zip(la, lb):
            assert a.lo == b.lo and a.hi == b.hi
            np.testing.assert_array_equal(a.data, b.data)


def test_hand_written_records(tmp_path):
    """A one-box plotfile written out by hand from the format (header text, Cell_H, a FAB record
    in big-endian order): the reader recovers names, time, box and values (x fastest)."""
    d = tmp_path / "plt"
    (d / "Level_0").mkdir(parents=True)
    (d / "Header").write_text("HyperCLaw-V1.1\n2\nrhoH\nTemp\n3\n1e-06\n0\n0 0 0 \n1 1 1 \n \n"
                              "((0,0,0) (1,0,0) (0,0,0)) \n0 \n0.5 1 1 \n0\n0\n0 1 1e-06\n0\n"
                              "0 1\n0 1\n0 1\nLevel_0/Cell\n")
    vals = np.array([1.0, 2.0, 300.0, 301.0])            # comp 0 (x = 0, 1), comp 1 (x = 0, 1)
    head = b"FAB ((8, (64 11 52 0 1 12 0 1023)),(8, (1 2 3 4 5 6 7 8)))((0,0,0) (1,0,0) (0,0,0)) 2\n"
    pad = b"xyz"                                         # the FAB starts at a nonzero offset
    (d / "Level_0" / "Cell_D_00000").write_bytes(pad + head + vals.astype(">f8").tobytes())
    (d / "Level_0" / "Cell_H").write_text(f"1\n0\n2\n0\n(1 0\n((0,0,0) (1,0,0) (0,0,0))\n)\n1\n"
                                          f"FabOnDisk: Cell_D_00000 {len(pad)}\n\n1,2\n1,300,\n\n1,2\n2,301,\n")
    pf = read_plotfile(str(d))
    assert pf.names == ["rhoH", "Temp"] and pf.time == 1e-6 and pf.dx == [(0.5, 1.0, 1.0)]
    fb = pf.levels[0][0]
    assert fb.lo == (0, 0, 0) and fb.hi == (1, 0, 0)
    np.testing.assert_array_equal(fb.data.reshape(2, -1), [[1.0, 2.0], [300.0, 301.0]])


def test_rejects_other_formats(tmp_path):
    rng = np.random.default_rng(1)
    write_plotfile(str(tmp_path / "plt"), _pf(rng))
    f = tmp_path / "plt" / "Level_0" / "Cell_D_00000"
    raw = f.read_bytes().replace(b"64 11 52 0 1 12 0 1023", b"32 8 23 0 1 9 0 127", 1)
    f.write_bytes(raw)
    with pytest.raises(ValueError):
        read_plotfile(str(tmp_path / "plt"))


def test_chemtest_names():
    n = chemtest_names(["H2", "O2"])
    assert n == ["rhoY(H2)", "rhoY(O2)", "rhoH", "Temp", "F_rhoY(H2)", "F_rhoY(O2)", "F_rhoH"]
    assert len(chemtest_names(["a"] * 21)) == 2 * (21 + 1) + 1
