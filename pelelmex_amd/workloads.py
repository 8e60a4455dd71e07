"""Seeded synthetic reactor inputs (box states in CGS) for the benchmark configurations.

Every generator returns component-major arrays in AMReX Array4 order (i fastest, then j, k;
species as the slowest index): ``rY[N, ncells]``, ``rYsrc[N, ncells]``, ``T``, ``rH``, ``rHsrc``,
``mask`` (int32, 1 = integrate, -1 = covered) with ``cell = i + nx * (j + ny * k)``.
Random numbers come from numpy's Philox4x32-10 generator keyed by the seed, drawn in a fixed
order, so the same seed gives the same box on every machine.

Recipes (documented choices where the configuration text leaves them open):
* H2/air (LiDryer, 32^3, seed 1001): phi ~ U[0.5, 2] sets the unburnt H2:O2:N2 = 2 phi : 1 : 3.76
  mole ratios; complete-combustion products are H2O plus excess O2 (lean) or H2 (rich) with the
  same N2; burn fraction c ~ U[0, 1] mixes unburnt and product mass fractions; radicals H, O,
  OH, HO2, H2O2 each get Y ~ 10^U[-8, -3], then Y is renormalised to sum 1; T ~ U[1000, 1800] K;
  P = 1 atm; no sources; dt = 1e-6 s.
* CH4/air (drm19, 64^3, seed 2002): slabs along i with x = (i + 0.5) / 64: x < 0.3 cold unburnt
  CH4:O2:N2 = phi : 2 : 7.52, phi ~ U[0.6, 1.4], T = 300 K; x < 0.5 flame front: unburnt/product
  mix with c ~ U[0, 1], radicals H, O, OH, HO2, CH2, CH2(S), CH3, HCO, CH3O, C2H5 each
  Y ~ 10^U[-7, -2], T ~ U[1200, 2200] K; otherwise burnt products (carbon to CO, hydrogen to
  H2O as far as oxygen allows, remaining oxygen to CO2 then O2, remaining hydrogen to H2) with
  dissociation radicals H, O, OH each Y ~ 10^U[-6, -3], T ~ U[1800, 2300] K; renormalised.
  External sources rhoY_src_k = 0.05 |rhoY_k| / dt * xi_k and rhoH_src = 0.05 |rhoH| / dt * xi_h
  with xi ~ N(0, 1); P = 1 atm; dt = 5e-6 s.
* n-dodecane (dodecane_lu, 128^3 in eight 64^3 chunks, seed 3003 + chunk): see dodecane_box.
* syn53 at the n-dodecane size (synthetic 53-species mechanism, 128^3 in eight 64^3 chunks,
  seed 3003 + chunk): mixture fraction Z ~ Beta(0.5, 2) between a CH4 fuel stream and air
  (O2:N2 = 1 : 3.76), phi = (Z / (1 - Z)) / (Z_st / (1 - Z_st)) with Z_st = 0.0552; burn
  fraction c ~ U[0, 1] toward the CH4 products at min(phi, 2); radicals H, O, OH, HO2, CH3,
  HCO each Y ~ 10^U[-8, -3] and every isomer X01..X32 Y ~ 10^U[-8, -4]; renormalised;
  T ~ U[700, 2400] K; P = 60 atm; no sources; dt = 2e-6 s; 10 % of cells masked off (mask -1)
  at random.
"""
from __future__ import annotations

import numpy as np

from .mechanisms import PATM, load


def _rng(seed):
    return np.random.Generator(np.random.Philox(int(seed)))


def _finish(mech, Y, T, P, dt, src_frac, rng, dims, mask=None):
    Y = Y / Y.sum(axis=1, keepdims=True)
    rho = mech.density(P, T, Y)
    rY = (rho[:, None] * Y).T.copy()
    rH = rho * mech.h_mass(T, Y)
    n = T.size
    if src_frac > 0.0:
        xi = rng.standard_normal((mech.nsp + 1, n))
        rYsrc = src_frac * np.abs(rY) / dt * xi[:-1]
        rHsrc = src_frac * np.abs(rH) / dt * xi[-1]
    else:
        rYsrc = np.zeros_like(rY)
        rHsrc = np.zeros(n)
    if mask is None:
        mask = np.ones(n, dtype=np.int32)
    return dict(mech=mech.name, rY=rY, rYsrc=rYsrc, T=T.copy(), rH=rH, rHsrc=rHsrc,
                mask=mask.astype(np.int32), dt=float(dt), P=float(P), dims=tuple(dims))


def h2_air_box(nx=32, ny=32, nz=32, seed=1001, dt=1e-6):
    """LiDryer H2/air box (32^3 = 32768 cells by default)."""
    m = load("LiDryer")
    rng = _rng(seed)
    n = nx * ny * nz
    phi = rng.uniform(0.5, 2.0, n)
    c = rng.uniform(0.0, 1.0, n)
    T = rng.uniform(1000.0, 1800.0, n)
    X_u = np.zeros((n, m.nsp))
    X_u[:, m.index["H2"]] = 2 * phi
    X_u[:, m.index["O2"]] = 1.0
    X_u[:, m.index["N2"]] = 3.76
    X_b = np.zeros((n, m.nsp))
    X_b[:, m.index["H2O"]] = 2 * np.minimum(phi, 1.0)
    X_b[:, m.index["O2"]] = np.maximum(1.0 - phi, 0.0)
    X_b[:, m.index["H2"]] = 2 * np.maximum(phi - 1.0, 0.0)
    X_b[:, m.index["N2"]] = 3.76
    Y = (1 - c)[:, None] * m.X_to_Y(X_u) + c[:, None] * m.X_to_Y(X_b)
    for s in ("H", "O", "OH", "HO2", "H2O2"):
        Y[:, m.index[s]] += 10.0 ** rng.uniform(-8.0, -3.0, n)
    return _finish(m, Y, T, PATM, dt, 0.0, rng, (nx, ny, nz))


def _ch4_products(m, phi):
    n = phi.size
    C, H, O = phi.copy(), 4 * phi, np.full(n, 4.0)
    CO = C.copy()
    O = O - C
    H2O = np.minimum(H / 2, O)
    O = O - H2O
    H = H - 2 * H2O
    CO2 = np.minimum(CO, O)
    CO = CO - CO2
    O = O - CO2
    X = np.zeros((n, m.nsp))
    X[:, m.index["CO"]] = CO
    X[:, m.index["CO2"]] = CO2
    X[:, m.index["H2O"]] = H2O
    X[:, m.index["O2"]] = O / 2
    X[:, m.index["H2"]] = H / 2
    X[:, m.index["N2"]] = 7.52
    return X


def ch4_air_flame_box(nx=64, ny=64, nz=64, seed=2002, dt=5e-6, src_frac=0.05):
    """drm19 CH4/air flame-like box (64^3 = 262144 cells by default)."""
    m = load("drm19")
    rng = _rng(seed)
    n = nx * ny * nz
    i = np.arange(n) % nx
    x = (i + 0.5) / nx
    cold = x < 0.3
    front = (~cold) & (x < 0.5)
    burnt = ~(cold | front)
    phi = rng.uniform(0.6, 1.4, n)
    c = rng.uniform(0.0, 1.0, n)
    T_front = rng.uniform(1200.0, 2200.0, n)
    T_burnt = rng.uniform(1800.0, 2300.0, n)
    X_u = np.zeros((n, m.nsp))
    X_u[:, m.index["CH4"]] = phi
    X_u[:, m.index["O2"]] = 2.0
    X_u[:, m.index["N2"]] = 7.52
    Y_u = m.X_to_Y(X_u)
    Y_b = m.X_to_Y(_ch4_products(m, phi))
    Y = np.where(cold[:, None], Y_u, np.where(front[:, None], (1 - c)[:, None] * Y_u + c[:, None] * Y_b, Y_b))
    front_rad = ("H", "O", "OH", "HO2", "CH2", "CH2(S)", "CH3", "HCO", "CH3O", "C2H5")
    for s in front_rad:
        draw = 10.0 ** rng.uniform(-7.0, -2.0, n)
        Y[:, m.index[s]] += np.where(front, draw, 0.0)
    for s in ("H", "O", "OH"):
        draw = 10.0 ** rng.uniform(-6.0, -3.0, n)
        Y[:, m.index[s]] += np.where(burnt, draw, 0.0)
    T = np.where(cold, 300.0, np.where(front, T_front, T_burnt))
    return _finish(m, Y, T, PATM, dt, src_frac, rng, (nx, ny, nz))


def syn53_chunk(chunk=0, nx=64, ny=64, nz=64, seed=3003, dt=2e-6, masked=0.10):
    """One 64^3 chunk (262144 cells) of the 128^3 syn53 box at 60 atm."""
    m = load("syn53")
    rng = _rng(seed + int(chunk))
    n = nx * ny * nz
    Z = rng.beta(0.5, 2.0, n)
    zst = 0.0552
    phi = (Z / np.maximum(1.0 - Z, 1e-12)) / (zst / (1.0 - zst))
    c = rng.uniform(0.0, 1.0, n)
    T = rng.uniform(700.0, 2400.0, n)
    Y_air = m.X_to_Y(np.array([[1.0 if s == "O2" else (3.76 if s == "N2" else 0.0) for s in m.species]]))[0]
    Y_u = (1.0 - Z)[:, None] * Y_air[None, :]
    Y_u[:, m.index["CH4"]] += Z
    Y_b = m.X_to_Y(_ch4_products(m, np.minimum(phi, 2.0)))
    Y = (1 - c)[:, None] * Y_u + c[:, None] * Y_b
    for s in ("H", "O", "OH", "HO2", "CH3", "HCO"):
        Y[:, m.index[s]] += 10.0 ** rng.uniform(-8.0, -3.0, n)
    for s in m.species[21:]:
        Y[:, m.index[s]] += 10.0 ** rng.uniform(-8.0, -4.0, n)
    mask = np.where(rng.uniform(0.0, 1.0, n) < masked, -1, 1).astype(np.int32)
    return _finish(m, Y, T, 60.0 * PATM, dt, 0.0, rng, (nx, ny, nz), mask=mask)


def _fuel_air_products(m, X_u, fuel, nC, nH):
    """Complete-combustion products of unburnt mole fractions X_u (fuel CnCHnH + O2 + N2): lean or
    stoichiometric -> CO2, H2O and leftover O2; rich -> carbon to CO, hydrogen to H2O, the remaining
    oxygen turning CO into CO2; too rich to oxidize every carbon to CO -> the fuel burns to CO + H2O
    as far as the oxygen goes and the rest stays fuel. Element-conserving, per unburnt mole."""
    n = X_u.shape[0]
    F = X_u[:, m.index[fuel]]
    O = 2.0 * X_u[:, m.index["O2"]]
    C, H = nC * F, nH * F
    X = np.zeros((n, m.nsp))
    X[:, m.index["N2"]] = X_u[:, m.index["N2"]]
    lean = O >= 2 * C + H / 2
    mid = (~lean) & (O >= C + H / 2)
    rich = ~(lean | mid)
    X[:, m.index["CO2"]] = np.where(lean, C, np.where(mid, O - C - H / 2, 0.0))
    X[:, m.index["H2O"]] = np.where(rich, nH / 2 * O / (nC + nH / 2), H / 2)
    X[:, m.index["O2"]] = np.where(lean, (O - 2 * C - H / 2) / 2, 0.0)
    X[:, m.index["CO"]] = np.where(mid, C - (O - C - H / 2), np.where(rich, nC * O / (nC + nH / 2), 0.0))
    X[:, m.index[fuel]] = np.where(rich, F - O / (nC + nH / 2), 0.0)
    return X


def dodecane_box(nx=64, ny=64, nz=64, seed=3003, dt=2e-6, masked=0.10, P_atm=60.0):
    """n-dodecane box (dodecane_lu; one 64^3 chunk = 262144 cells by default) at 60 atm: mixture
    fraction Z ~ Beta(0.5, 2) between pure NC12H26 and air (O2 : N2 = 1 : 3.76), burn fraction
    c ~ U[0, 1] toward the element-conserving complete-combustion products of the same Z
    (_fuel_air_products), radicals H, O, OH, HO2, H2O2 each Y ~ 10^U[-8, -3], renormalised,
    T ~ U[700, 2400] K, no sources, dt 2e-6 s, a fraction `masked` of cells covered (mask -1)."""
    m = load("dodecane_lu")
    rng = _rng(seed)
    n = nx * ny * nz
    Z = rng.beta(0.5, 2.0, n)
    c = rng.uniform(0.0, 1.0, n)
    T = rng.uniform(700.0, 2400.0, n)
    W_F = m.W[m.index["NC12H26"]]
    W_air = (m.W[m.index["O2"]] + 3.76 * m.W[m.index["N2"]]) / 4.76
    nF = Z / W_F
    nA = (1.0 - Z) / W_air
    X_u = np.zeros((n, m.nsp))
    X_u[:, m.index["NC12H26"]] = nF / (nF + nA)
    X_u[:, m.index["O2"]] = nA / (nF + nA) / 4.76
    X_u[:, m.index["N2"]] = nA / (nF + nA) * 3.76 / 4.76
    Y_u = m.X_to_Y(X_u)
    Y_b = m.X_to_Y(_fuel_air_products(m, X_u, "NC12H26", 12, 26))
    Y = (1 - c)[:, None] * Y_u + c[:, None] * Y_b
    for s in ("H", "O", "OH", "HO2", "H2O2"):
        Y[:, m.index[s]] += 10.0 ** rng.uniform(-8.0, -3.0, n)
    mask = np.where(rng.uniform(0.0, 1.0, n) < masked, -1, 1).astype(np.int32)
    return _finish(m, Y, T, P_atm * PATM, dt, 0.0, rng, (nx, ny, nz), mask=mask)


def dodecane_chunk(chunk=0, nx=64, ny=64, nz=64, seed=3003, dt=2e-6, masked=0.10):
    """One 64^3 chunk (262144 cells) of the 128^3 n-dodecane box at 60 atm (seed 3003 + chunk)."""
    return dodecane_box(nx, ny, nz, seed=seed + int(chunk), dt=dt, masked=masked)


CONFIGS = {
    "h2_air_32": h2_air_box,
    "ch4_air_flame_64": ch4_air_flame_box,
    "syn53_60atm_64": syn53_chunk,
    "dodecane_60atm_64": dodecane_chunk,
}


def subsample(box, idx):
    """A workload restricted to the cells idx (kept in order), as a 1D box."""
    idx = np.asarray(idx)
    out = dict(box)
    for k in ("rY", "rYsrc"):
        out[k] = np.ascontiguousarray(box[k][:, idx])
    for k in ("T", "rH", "rHsrc", "mask"):
        out[k] = np.ascontiguousarray(box[k][idx])
    out["dims"] = (idx.size, 1, 1)
    return out
