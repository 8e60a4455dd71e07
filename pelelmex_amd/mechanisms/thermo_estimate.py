"""NASA-7 fits for the dodecane_lu species that have no GRI-Mech / Li et al. thermo entry.

Method (published, named): Benson group additivity (S.W. Benson, Thermochemical Kinetics, 2nd ed.,
1976; group values as tabulated by N. Cohen and S.W. Benson, Chem. Rev. 93 (1993) 2419) for the
closed-shell molecules, and the hydrogen-bond-increment method for radicals (T.H. Lay,
J.W. Bozzelli, A.M. Dean, E.R. Ritter, J. Phys. Chem. 99 (1995) 14514): a radical is its parent
molecule minus one H atom, with dHf = dHf(parent) + D(R-H) - dHf(H), the entropy changed by the
symmetry and spin terms, and the heat capacity reduced by the lost C-H modes. Entropies carry
-R ln(sigma) for the symmetry number and +R ln(n) for n lumped structural isomers. The C2 / C3
species that are not group-additive (C2H2, C2H3, CH2CHO, aC3H5, C2H3CHO) take literature
dHf / S298 / Cp(T) values (JANAF and the compilations above, rounded).

From (dHf298, S298, Cp at 300..1500 K) each species gets the two-range NASA-7 form used by the
reactor (T_mid = 1000 K): a least-squares fit of Cp/R in each range with the two ranges' Cp, dCp/dT
equal at T_mid, Cp extended beyond 1500 K by the fit itself (bounded heat capacity growth of
the 1500 K slope), and a6 / a7 fixed by H and S at 298.15 K in the low range and by continuity of
H and S at T_mid in the high range. The coefficients are printed rounded to 9 significant digits
and pasted into dodecane_lu.py, so that the mechanism data is a plain table (run this module to
regenerate them: python -m pelelmex_amd.mechanisms.thermo_estimate).

The lumped hydroperoxyalkyl C12OOH takes its enthalpy and entropy relative to C12H25O2 from the
partial equilibrium the reference flame holds between them (see species_data). The formation
enthalpy of NC12H26 is not taken from the groups but from the reference's own
adiabatic premixed n-dodecane flame (Exec/UnitTests/DodecaneLu/dodecaneLu_pmf.dat): the value that
makes the unburnt (298 K) and burnt ends of that profile isenthalpic with the GRI thermo of the
products (-69.975 kcal/mol; the group value -69.70 kcal/mol is 0.28 kcal/mol above it).
"""
from __future__ import annotations

import math

import numpy as np

R = 1.987204  # cal / (mol K)
TS = np.array([300.0, 400.0, 500.0, 600.0, 800.0, 1000.0, 1500.0])

# group: (dHf298 kcal/mol, S298 cal/mol/K, Cp at TS cal/mol/K)
GROUPS = {
    "C-(C)(H)3": (-10.20, 30.41, [6.19, 7.84, 9.40, 10.79, 13.02, 14.77, 17.58]),
    "C-(C)2(H)2": (-4.93, 9.42, [5.50, 6.95, 8.25, 9.35, 11.07, 12.34, 14.25]),
    "Cd-(H)2": (6.26, 27.61, [5.10, 6.36, 7.51, 8.50, 10.07, 11.27, 13.19]),
    "Cd-(C)(H)": (8.59, 7.97, [4.16, 5.03, 5.81, 6.50, 7.65, 8.45, 9.62]),
    "C-(Cd)(C)(H)2": (-4.76, 9.80, [5.12, 6.86, 8.32, 9.49, 11.22, 12.48, 14.36]),
    "C-(Cd)(H)3": (-10.20, 30.41, [6.19, 7.84, 9.40, 10.79, 13.02, 14.77, 17.58]),
    "C-(O)(C)2(H)": (-7.20, -11.00, [4.80, 6.64, 8.10, 8.73, 9.81, 10.40, 11.51]),
    "C-(O)(C)(H)2": (-8.10, 9.80, [4.99, 6.85, 8.30, 9.43, 11.11, 12.33, 14.10]),
    "O-(O)(C)": (-4.50, 9.40, [3.90, 4.31, 4.60, 4.84, 5.32, 5.80, 6.50]),
    "O-(O)(H)": (-16.27, 27.83, [5.21, 5.72, 6.17, 6.66, 7.15, 7.61, 8.43]),
    "CO-(C)2": (-31.40, 15.00, [5.59, 6.32, 7.09, 7.76, 8.89, 9.61, 10.75]),
    "C-(CO)(C)(H)2": (-5.20, 9.60, [6.20, 7.70, 8.80, 9.60, 11.10, 12.20, 13.80]),
}

# hydrogen-bond increments (kcal/mol for D(R-H); Cp change of the lost C-H modes)
D_PRIMARY, D_SECONDARY, D_ALLYL_PRIMARY, D_ROO_H = 101.1, 98.5, 88.8, 85.3
HF_H = 52.10
DCP_CH = np.array([-0.50, -1.00, -1.50, -1.80, -2.20, -2.40, -2.70])


def additive(groups, sigma=1, isomers=1):
    h = sum(GROUPS[g][0] * n for g, n in groups.items())
    s = sum(GROUPS[g][1] * n for g, n in groups.items()) - R * math.log(sigma) + R * math.log(isomers)
    cp = sum(np.array(GROUPS[g][2]) * n for g, n in groups.items())
    return h, s, cp


def radical(parent, D, sigma_parent, sigma_rad, isomers=1):
    """HBI radical of a parent (h, s, cp): one H removed from a C-H (or O-H) bond of strength D."""
    h, s, cp = parent
    s_r = s + R * math.log(sigma_parent / sigma_rad) + R * math.log(2.0) + R * math.log(isomers)
    return h + D - HF_H, s_r, cp + DCP_CH


def alkane(n):
    return additive({"C-(C)(H)3": 2, "C-(C)2(H)2": n - 2}, sigma=18)


def alkene(n):
    """1-alkene CnH2n (n >= 4)."""
    g = {"Cd-(H)2": 1, "Cd-(C)(H)": 1, "C-(Cd)(C)(H)2": 1, "C-(C)(H)3": 1}
    if n > 4:
        g["C-(C)2(H)2"] = n - 4
    return additive(g, sigma=3)


def species_data():
    out = {}
    # literature values (kcal/mol, cal/mol/K)
    out["C2H2"] = (54.19, 48.00, np.array([10.55, 11.97, 12.97, 13.77, 15.00, 15.99, 17.73]))
    out["C2H3"] = (71.50, 56.30, np.array([10.00, 11.60, 13.00, 14.20, 16.10, 17.60, 19.90]))
    out["CH2CHO"] = (3.50, 63.00, np.array([13.10, 15.60, 17.70, 19.40, 21.90, 23.70, 26.20]))
    out["aC3H5"] = (40.90, 61.50, np.array([15.10, 18.40, 21.50, 24.10, 28.20, 31.30, 36.20]))
    out["C2H3CHO"] = (-17.80, 67.40, np.array([17.00, 20.60, 23.70, 26.20, 29.90, 32.60, 36.70]))
    c3h6 = additive({"Cd-(H)2": 1, "Cd-(C)(H)": 1, "C-(Cd)(H)3": 1}, sigma=3)
    out["C3H6"] = c3h6
    propane = additive({"C-(C)(H)3": 2, "C-(C)2(H)2": 1}, sigma=18)
    out["nC3H7"] = radical(propane, D_PRIMARY, 18, 3)
    out["C4H81"] = alkene(4)
    out["pC4H9"] = radical(alkane(4), D_PRIMARY, 18, 3)
    # 3-buten-1-yl and 4-penten-1-yl: primary radicals of the 1-alkenes (their beta-scissions give
    # C2H4 + C2H3 and C2H4 + aC3H5)
    out["C4H7"] = radical(alkene(4), D_PRIMARY, 3, 1)
    out["C5H9"] = radical(alkene(5), D_PRIMARY, 3, 1)
    for n in (5, 6, 7, 8, 9, 10, 12):
        out[f"C{n}H{2 * n}"] = alkene(n)
    for n in (5, 6, 7, 8, 9, 10, 12):
        out[f"PXC{n}H{2 * n + 1}"] = radical(alkane(n), D_PRIMARY, 18, 3)
    out["NC12H26"] = alkane(12)
    dod = alkane(12)
    out["SXC12H25"] = radical(dod, D_SECONDARY, 18, 9)
    out["S3XC12H25"] = radical(dod, D_SECONDARY, 18, 9, isomers=4)   # 3- to 6-dodecyl, lumped
    # low-temperature species (secondary positions; 2 optical forms per stereocentre)
    rooh = additive({"C-(O)(C)2(H)": 1, "O-(O)(C)": 1, "O-(O)(H)": 1, "C-(C)(H)3": 2, "C-(C)2(H)2": 9},
                    sigma=9, isomers=2)
    h, s, cp = rooh
    out["C12H25O2"] = (h + D_ROO_H - HF_H, s + R * math.log(2.0) - R * math.log(3.0), cp + DCP_CH * 0.8)
    # the lumped QOOH: relative to RO2, dH = +7.75 kcal/mol and dS = +5.75 cal/mol/K, the partial
    # equilibrium RO2 <=> QOOH that the reference's n-dodecane flame holds from 520 to 980 K
    # ([C12OOH] / [C12H25O2] = exp(5.75 / R - 7750 / (R T)) within 5 % at every point of
    # dodecaneLu_pmf.dat there; a plain HBI estimate gives +13.2 kcal/mol, 300x less QOOH at 700 K)
    h_ro2, s_ro2, cp_ro2 = out["C12H25O2"]
    out["C12OOH"] = (h_ro2 + 7.75, s_ro2 + 5.75, cp_ro2)
    bis = additive({"C-(O)(C)2(H)": 2, "O-(O)(C)": 2, "O-(O)(H)": 2, "C-(C)(H)3": 2, "C-(C)2(H)2": 8},
                   sigma=9, isomers=4)
    h, s, cp = bis
    out["O2C12H24OOH"] = (h + D_ROO_H - HF_H, s + R * math.log(2.0) - R * math.log(3.0), cp + DCP_CH * 0.8)
    out["OC12H23OOH"] = additive({"CO-(C)2": 1, "C-(CO)(C)(H)2": 1, "C-(O)(C)2(H)": 1, "O-(O)(C)": 1,
                                  "O-(O)(H)": 1, "C-(C)(H)3": 2, "C-(C)2(H)2": 7}, sigma=9, isomers=2)
    return out


def fit_nasa(h298, s298, cp):
    """Two-range NASA-7 (T_mid 1000 K, 300..3500 K) from dHf298 (kcal/mol), S298 (cal/mol/K) and Cp
    at TS (cal/mol/K): returns (low[7], high[7])."""
    cpR = np.asarray(cp) / R
    # Cp beyond 1500 K: continued with a slope falling to zero at 3500 K (bounded growth)
    slope = (cpR[-1] - cpR[-2]) / (TS[-1] - TS[-2])
    ext_T = np.array([2000.0, 2500.0, 3000.0, 3500.0])
    ext = cpR[-1] + slope * (ext_T - 1500.0) * (1.0 - (ext_T - 1500.0) / 4000.0)
    T_all = np.concatenate([TS, ext_T])
    c_all = np.concatenate([cpR, ext])
    # cubic-spline-free smooth interpolant: fit each range with a quartic in T, continuity of Cp and
    # dCp/dT at 1000 K imposed through Lagrange multipliers
    lo_m = T_all <= 1000.0 + 1e-9
    hi_m = T_all >= 1000.0 - 1e-9
    Tl, cl = T_all[lo_m], c_all[lo_m]
    Th, ch = T_all[hi_m], c_all[hi_m]
    # dense resampling by monotone piecewise-linear interpolation (numerically benign LSQ)
    Tl_d = np.linspace(Tl.min(), 1000.0, 60)
    Th_d = np.linspace(1000.0, 3500.0, 120)
    cl_d = np.interp(Tl_d, T_all, c_all)
    ch_d = np.interp(Th_d, T_all, c_all)
    V = lambda T: np.vstack([T ** k for k in range(5)]).T
    dV = lambda T: np.vstack([k * T ** (k - 1) if k else 0 * T for k in range(5)]).T
    sc = np.array([1.0, 1e-3, 1e-6, 1e-9, 1e-12])      # coefficient scaling for conditioning
    A = np.zeros((Tl_d.size + Th_d.size, 10))
    A[:Tl_d.size, :5] = V(Tl_d) * sc
    A[Tl_d.size:, 5:] = V(Th_d) * sc
    b = np.concatenate([cl_d, ch_d])
    Cc = np.zeros((2, 10))
    Cc[0, :5] = V(np.array([1000.0]))[0] * sc
    Cc[0, 5:] = -V(np.array([1000.0]))[0] * sc
    Cc[1, :5] = dV(np.array([1000.0]))[0] * sc
    Cc[1, 5:] = -dV(np.array([1000.0]))[0] * sc
    K = np.block([[2 * A.T @ A, Cc.T], [Cc, np.zeros((2, 2))]])
    rhs = np.concatenate([2 * A.T @ b, np.zeros(2)])
    sol = np.linalg.solve(K, rhs)[:10]
    lo = list(sol[:5] * sc)
    hi = list(sol[5:] * sc)

    def H_RT(a, T):   # without a6
        return a[0] + a[1] * T / 2 + a[2] * T ** 2 / 3 + a[3] * T ** 3 / 4 + a[4] * T ** 4 / 5

    def S_R(a, T):    # without a7
        return a[0] * math.log(T) + a[1] * T + a[2] * T ** 2 / 2 + a[3] * T ** 3 / 3 + a[4] * T ** 4 / 4

    T0 = 298.15
    a6l = h298 * 1000.0 / R - H_RT(lo, T0) * T0
    a7l = s298 / R - S_R(lo, T0)
    Tm = 1000.0
    h_m = (H_RT(lo, Tm) * Tm + a6l)
    s_m = S_R(lo, Tm) + a7l
    a6h = h_m - H_RT(hi, Tm) * Tm
    a7h = s_m - S_R(hi, Tm)
    return lo + [a6l, a7l], hi + [a6h, a7h]


def _g(x):
    return float(f"{x:.9g}")


def table(hf_overrides=None):
    """{species: (300.0, 1000.0, 3500.0, high[7], low[7])} rounded to 9 significant digits."""
    out = {}
    for sp, (h, s, cp) in species_data().items():
        if hf_overrides and sp in hf_overrides:
            h = hf_overrides[sp]
        lo, hi = fit_nasa(h, s, cp)
        out[sp] = (300.0, 1000.0, 3500.0, [_g(x) for x in hi], [_g(x) for x in lo])
    return out


if __name__ == "__main__":
    for sp, (tl, tm, th, hi, lo) in table().items():
        print(f'    "{sp}": ({tl}, {tm}, {th},\n        {hi},\n        {lo}),')
