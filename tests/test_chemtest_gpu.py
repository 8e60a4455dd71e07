"""chemTest harness on the GPU: a two-level chemTest plotfile (MKS state + forcing, dt in the
header time) advanced by pelelmex_amd.chemtest gives, FAB by FAB, the CPU oracle's bits on the
same cells after advanceChemistry's unit conversions, and round-trips through the writer."""
import numpy as np
import pytest

from pelelmex_amd import workloads as W

pytestmark = pytest.mark.gpu

KEYS = {"ode.rtol": 1e-10, "ode.atol": 1e-10, "cvode.max_order": 4, "cvode.max_substeps": 10000}


def _fab_from_box(box, species):
    from pelelmex_amd.chemtest import chemtest_names
    nx, ny, nz = box["dims"]
    N = len(species)
    comps = np.concatenate([box["rY"] * 1e3, (box["rH"] * 0.1)[None], box["T"][None], box["rYsrc"] * 1e3,
                            (box["rHsrc"] * 0.1)[None]])
    assert comps.shape[0] == len(chemtest_names(species)) == 2 * N + 3
    return comps.reshape(2 * N + 3, nz, ny, nx)


def test_chemtest_plotfile_matches_react_box(gpu, tmp_path):
    import torch
    from pelelmex_amd.chemtest import chemtest_names, react_plotfile
    from pelelmex_amd.mechanisms import load
    from pelelmex_amd.plotfile import Fab, PlotFile, read_plotfile, write_plotfile
    from pelelmex_amd.reactor import Reactor
    species = load("drm19").species
    dt = 5e-6
    b0 = W.ch4_air_flame_box(nx=16, ny=4, nz=2, seed=61)
    b1 = W.ch4_air_flame_box(nx=8, ny=4, nz=4, seed=62)
    b2 = W.ch4_air_flame_box(nx=8, ny=8, nz=2, seed=63)
    pf = PlotFile(chemtest_names(species), dt, (0, 0, 0), (0.016, 0.004, 0.004), [2],
                  [((0, 0, 0), (15, 3, 3)), ((0, 0, 0), (31, 7, 7))], [(1e-3,) * 3, (5e-4,) * 3],
                  [[Fab((0, 0, 0), (15, 3, 1), _fab_from_box(b0, species)),
                    Fab((0, 0, 2), (7, 3, 5), _fab_from_box(b1, species))],
                   [Fab((4, 0, 0), (11, 7, 1), _fab_from_box(b2, species))]])
    write_plotfile(str(tmp_path / "pltEvaluate"), pf)
    back = read_plotfile(str(tmp_path / "pltEvaluate"))
    r = Reactor.create("ReactorCvode", "drm19", KEYS)
    r.init(2, 1)
    out, nfail = react_plotfile(back, r, species, device=gpu)
    assert nfail == 0
    got = [f.data for lev in out.levels for f in lev]
    from oracle import pyoracle as po
    om = po.Mech("drm19")
    for box, g in zip((b0, b1, b2), got):
        # the same cells through the CPU oracle, MKS -> CGS -> MKS as advanceChemistry converts
        N = box["rY"].shape[0]
        ref = po.advance_chemistry_chem(om, box["rY"] * 1e3, box["rH"] * 0.1, box["T"], box["rYsrc"] * 1e3,
                                        box["rHsrc"] * 0.1, None, dt, max_order=4)
        nx, ny, nz = box["dims"]
        want = np.concatenate([ref["rY"], ref["rH"][None], ref["T"][None], ref["FC"][None]]).reshape(N + 3, nz, ny, nx)
        assert np.array_equal(g, want), "chemTest output differs from the oracle"
        assert g[-1].min() > 0                       # every cell integrated
    r.close()
    write_plotfile(str(tmp_path / "pltOut"), out)
    again = read_plotfile(str(tmp_path / "pltOut"))
    assert again.names[-1] == "FC" and again.time == dt
