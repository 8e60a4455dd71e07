"""Extract premixed-flame profile data from the PeleLMeX reference tree into tests/golden/ (run in the
dev container, where the reference is mounted; commit the outputs). The profiles are data files of
the reference (BSD-3-Clause, LICENSE:1-3), read by its PMF initial conditions
(pmf.datafile in Exec/RegTests/FlameSheet/flamesheet-drm19-3d.inp:30 and
Exec/RegTests/EnclosedFlame/enclosedflame-3d.inp; Exec/UnitTests/DodecaneLu/pelelmex_prob.H:84-87
reads the columns positionally). Columns: X (cm), temp (K), u (cm/s), rho (g/cm^3), then the
mole fractions in the mechanism's species order.

Outputs:
  pmf_ends.json              first and last rows of both drm19 profiles
  pmf_drm19_10atm.npz        the whole 10 atm drm19 profile (Exec/RegTests/EnclosedFlame/
                             drm19_H2CH4_10Atm.dat, 705 points, P = 10 atm per
                             enclosedflame-3d.inp:27) on its own adaptive grid, so the species
                             balance can be checked on the points it was solved on
  pmf_dodecane_lu.npz        the whole n-dodecane profile (Exec/UnitTests/DodecaneLu/
                             dodecaneLu_pmf.dat, 971 points; P = 1 atm)

Usage: python tests/golden/make_pmf_fixtures.py
"""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference"


def read_pmf(rel):
    lines = open(os.path.join(REF, rel)).read().splitlines()
    cols = [c.strip('"') for c in lines[0].split("=", 1)[1].split()]
    rows = [[float(v) for v in ln.split()] for ln in lines[2:] if ln.strip()]
    return cols, np.array(rows)


def main():
    src = "PeleLMeX reference tree, BSD-3-Clause (LICENSE:1-3)"
    profiles = {}
    for key, rel, P in (("drm19_pmf", "Exec/RegTests/FlameSheet/drm19_pmf.dat", 1.0),
                        ("drm19_H2CH4_10Atm", "Exec/RegTests/EnclosedFlame/drm19_H2CH4_10Atm.dat", 10.0)):
        cols, d = read_pmf(rel)
        profiles[key] = {"file": rel, "P_atm": P, "columns": cols, "first": d[0].tolist(), "last": d[-1].tolist(),
                         "rows": int(d.shape[0])}
    json.dump({"source": src + "; first and last rows of the premixed-flame profiles", "profiles": profiles},
              open(os.path.join(HERE, "pmf_ends.json"), "w"), indent=1)

    cols, d = read_pmf("Exec/RegTests/EnclosedFlame/drm19_H2CH4_10Atm.dat")
    np.savez_compressed(os.path.join(HERE, "pmf_drm19_10atm.npz"), columns=np.array(cols), data=d, P_atm=10.0,
                        source=src + ", Exec/RegTests/EnclosedFlame/drm19_H2CH4_10Atm.dat (all rows)")
    cols, d = read_pmf("Exec/UnitTests/DodecaneLu/dodecaneLu_pmf.dat")
    np.savez_compressed(os.path.join(HERE, "pmf_dodecane_lu.npz"), columns=np.array(cols), data=d, P_atm=1.0,
                        source=src + ", Exec/UnitTests/DodecaneLu/dodecaneLu_pmf.dat (all rows)")
    print("written")


if __name__ == "__main__":
    main()
