"""syn53: a SYNTHETIC 53-species, 268-reaction test mechanism at the size of n-dodecane_lu.

It is not n-dodecane_lu and computes no real fuel's chemistry. That mechanism's reactions are not in
the reference tree (its PelePhysics submodule is empty): only its species order is, in
Exec/UnitTests/DodecaneLu/dodecaneLu_pmf.dat:1. This one is used only to exercise and measure the
reactor at that size: 54-component state, 54 x 54 Jacobian and LU (23.3 KB per cell each), about
five reactions per lane of a 64-lane group, every reaction type the kernels interpret.

Construction (deterministic):
  * the 21 drm19 species with their data, followed by 32 isomer species X01..X32. Each isomer has
    the composition of a drm19 parent species and the parent's NASA-7 fit with the enthalpy and
    entropy constants (a6, a7) raised by a fixed dH / R and dS / R in both ranges. Continuity at
    T_mid is kept, and each isomer sits dH above its parent;
  * the 84 drm19 reactions;
  * per isomer X_i of parent P_i, four bimolecular isomerizations X_i + B <=> P_i + B with
    B in {H, OH, O2, HO2}, and one unimolecular channel:
      - Troe falloff X_i (+M) <=> P_i (+M) for i = 1..16,
      - Lindemann falloff for i = 17..24,
      - a third-body reaction X_i + M <=> P_i + M for i = 25..32;
  * 24 cross reactions X_i + X_(i+1) <=> P_i + P_(i+1), i = 1..24.
Every reaction balances its elements, because isomers share their parent's composition. Reverse
rates follow from the NASA-7 fits, as for every reversible reaction. The unimolecular channels
are fast at flame temperatures (k_inf ~ 1e13 exp(-20000 K / T) s^-1) and frozen below ~800 K.
That mixes stiff and frozen cells, the regime the reactor is built for.
Units: A in mol-cm-s, Ea in cal/mol.
"""

from . import drm19 as _d

NAME = "syn53"

_PARENTS = ["CH4", "CH3", "CH2O", "CH3O", "C2H4", "C2H5", "C2H6", "HCO", "HO2", "H2O", "OH",
            "O2", "H2", "CO", "CO2", "CH2",
            "CH4", "C2H6", "C2H4", "CH2O", "CH3", "C2H5", "CH3O", "H2O",
            "CH4", "C2H6", "CO2", "O2", "H2", "C2H4", "CH2O", "HO2"]
_ISO = [f"X{i + 1:02d}" for i in range(32)]
PARENT = dict(zip(_ISO, _PARENTS))

SPECIES = list(_d.SPECIES) + _ISO

COMPOSITION = dict(_d.COMPOSITION)
for _x, _p in PARENT.items():
    COMPOSITION[_x] = dict(_d.COMPOSITION[_p])


def _shift(fit, dh_R, ds_R):
    c = list(fit)
    c[5] += dh_R
    c[6] += ds_R
    return c


THERMO = dict(_d.THERMO)
for _i, (_x, _p) in enumerate(PARENT.items()):
    _tl, _tm, _th, _hi, _lo = _d.THERMO[_p]
    _dh = 2500.0 + 250.0 * (_i % 8)          # K: dH = 5.0 .. 8.5 kcal/mol above the parent
    _ds = 0.5 - 0.25 * (_i % 3)              # dS / R in {0.5, 0.25, 0}
    THERMO[_x] = (_tl, _tm, _th, _shift(_hi, _dh, _ds), _shift(_lo, _dh, _ds))

_EFF = "H2/2.0/ H2O/6.0/ CH4/2.0/ CO/1.5/ CO2/2.0/ C2H6/3.0/ AR/0.7/"


def _extra():
    L = []
    for i, x in enumerate(_ISO):
        p = PARENT[x]
        L.append(f"{x}+H<=>{p}+H     {1.0e13 * (1 + i % 3):.3E}   0.000   {4000.0 + 200.0 * (i % 5):.2f}")
        L.append(f"{x}+OH<=>{p}+OH   {5.0e12:.3E}   0.000   {1500.0 + 100.0 * (i % 4):.2f}")
        L.append(f"{x}+O2<=>{p}+O2   {1.0e12:.3E}   0.000   {28000.0 + 500.0 * (i % 4):.2f}")
        L.append(f"{x}+HO2<=>{p}+HO2 {2.0e12:.3E}   0.000   {9000.0 + 300.0 * (i % 6):.2f}")
        ea = 38000.0 + 1000.0 * (i % 5)
        if i < 16:
            L.append(f"{x}(+M)<=>{p}(+M)  {1.0e13:.3E}   0.000   {ea:.2f}")
            L.append(f"    LOW/{1.0e17:.3E} 0.000 {ea - 8000.0:.2f}/")
            L.append(f"    TROE/{0.55 + 0.02 * (i % 5):.4f} {90.0 + 10.0 * (i % 3):.2f} {2000.0 + 100.0 * (i % 4):.2f} {6000.0:.2f}/")
            L.append(f"    {_EFF}")
        elif i < 24:
            L.append(f"{x}(+M)<=>{p}(+M)  {8.0e12:.3E}   0.000   {ea:.2f}")
            L.append(f"    LOW/{5.0e16:.3E} 0.000 {ea - 6000.0:.2f}/")
            L.append(f"    {_EFF}")
        else:
            L.append(f"{x}+M<=>{p}+M      {2.0e15:.3E}   0.000   {ea - 4000.0:.2f}")
            L.append(f"    {_EFF}")
    for i in range(24):
        x, y = _ISO[i], _ISO[i + 1]
        L.append(f"{x}+{y}<=>{PARENT[x]}+{PARENT[y]}  {1.0e12:.3E}   0.000   {12000.0 + 500.0 * (i % 4):.2f}")
    return "\n".join(L)


REACTIONS = _d.REACTIONS + "\n" + _extra() + "\n"
