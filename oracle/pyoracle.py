"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, as the checker or the timed CPU baseline; never by the pelelmex_amd package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

RHS_FN = C.CFUNCTYPE(C.c_int, C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_void_p)
JAC_FN = C.CFUNCTYPE(C.c_int, C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double),
                     C.POINTER(C.c_double), C.c_void_p)


class BdfOpts(C.Structure):
    _fields_ = [("rtol", C.c_double), ("atol", C.c_double), ("atol_vec", C.POINTER(C.c_double)),
                ("max_order", C.c_int), ("max_steps", C.c_int), ("clean_init_massfrac", C.c_int)]


class BdfStats(C.Structure):
    _fields_ = [("nst", C.c_long), ("nfe", C.c_long), ("nje", C.c_long), ("nsetups", C.c_long),
                ("nni", C.c_long), ("netf", C.c_long), ("ncfn", C.c_long), ("q_last", C.c_int),
                ("h_last", C.c_double)]


class Cell(C.Structure):
    _fields_ = [("mech", C.c_int), ("rhoh0", C.c_double), ("rhohsrc", C.c_double),
                ("rYsrc", C.POINTER(C.c_double))]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        dp, ip = C.POINTER(C.c_double), C.POINTER(C.c_int)
        L.orc_mech_find.argtypes = [C.c_char_p]
        L.orc_thermo.argtypes = [C.c_int, C.c_double, dp, dp, dp]
        L.orc_rates.argtypes = [C.c_int, dp, C.c_double, dp, dp]
        L.orc_hy2t.argtypes = [C.c_int, C.c_double, dp, dp]
        L.orc_reactor_rhs.argtypes = [C.c_double, dp, dp, C.c_void_p]
        L.orc_reactor_rhs_frozen.argtypes = [C.c_double, dp, dp, C.c_void_p]
        L.orc_reactor_jac.argtypes = [C.c_double, dp, dp, dp, C.c_void_p]
        L.orc_lu_factor.argtypes = [C.c_int, dp, ip]
        L.orc_lu_solve.argtypes = [C.c_int, dp, ip, dp]
        L.orc_bdf_solve.argtypes = [C.c_int, C.c_double, C.c_double, dp, RHS_FN, JAC_FN, C.c_void_p,
                                    C.POINTER(BdfOpts), C.POINTER(BdfStats)]
        L.orc_react.argtypes = [C.c_int, C.c_long, dp, dp, dp, dp, dp, dp, ip, C.c_double,
                                C.POINTER(BdfOpts), C.c_int, ip, C.POINTER(C.c_long)]
        L.orc_mech_weights.argtypes = [C.c_int, dp]
        L.orc_rates_fr.argtypes = [C.c_int, dp, C.c_double, dp, dp]
        L.orc_rate_rhoY.argtypes = [C.c_int, C.c_long, dp, dp, dp, dp, dp]
        L.orc_set_temperature.argtypes = [C.c_int, C.c_long, dp, dp, dp, dp]
        L.orc_floor_species.argtypes = [C.c_int, C.c_long, dp, dp, dp, dp]
        L.orc_heat_release.argtypes = [C.c_int, C.c_long, dp, dp, dp]
        L.orc_exp.argtypes = [C.c_double]
        L.orc_exp.restype = C.c_double
        L.orc_log.argtypes = [C.c_double]
        L.orc_log.restype = C.c_double
        _LIB = L
    return _LIB


def _p(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


class Mech:
    """Mechanism tables as seen by the oracle."""

    def __init__(self, name):
        L = lib()
        self.id = L.orc_mech_find(name.encode())
        if self.id < 0:
            raise KeyError(name)
        self.name = name
        self.nsp = L.orc_mech_nsp(self.id)
        self.nr = L.orc_mech_nr(self.id)
        self.W = np.zeros(self.nsp)
        L.orc_mech_weights(self.id, _p(self.W))

    def thermo(self, T):
        cp, h, s = (np.zeros(self.nsp) for _ in range(3))
        lib().orc_thermo(self.id, float(T), _p(cp), _p(h), _p(s))
        return cp, h, s

    def rates(self, Cmol, T):
        Cmol = np.ascontiguousarray(Cmol, dtype=np.float64)
        q, w = np.zeros(self.nr), np.zeros(self.nsp)
        lib().orc_rates(self.id, _p(Cmol), float(T), _p(q), _p(w))
        return q, w

    def rates_fr(self, Cmol, T):
        Cmol = np.ascontiguousarray(Cmol, dtype=np.float64)
        f, r = np.zeros(self.nr), np.zeros(self.nr)
        lib().orc_rates_fr(self.id, _p(Cmol), float(T), _p(f), _p(r))
        return f, r

    def rate_rhoY(self, rY, rH, T0):
        """reactionRateRhoY on cells stored component-major (rY shape (N, ncells)); MKS.
        Returns (rhoYdot (N, ncells), T (ncells))."""
        rY = np.ascontiguousarray(rY, dtype=np.float64)
        rH = np.ascontiguousarray(rH, dtype=np.float64)
        T0 = np.ascontiguousarray(T0, dtype=np.float64)
        nc = rH.size
        out, T = np.zeros((self.nsp, nc)), np.zeros(nc)
        lib().orc_rate_rhoY(self.id, nc, _p(rY), _p(rH), _p(T0), _p(out), _p(T))
        return out, T

    def set_temperature(self, rho, rY, rH, T):
        """setTemperature on cells (MKS; rY (N, ncells)); returns the new T."""
        rho, rY, rH = (np.ascontiguousarray(x, dtype=np.float64) for x in (rho, rY, rH))
        T = np.array(T, dtype=np.float64, copy=True)
        lib().orc_set_temperature(self.id, T.size, _p(rho), _p(rY), _p(rH), _p(T))
        return T

    def floor_species(self, rY, T):
        """floorSpecies on cells (MKS); returns (clipped rY, rho, rH)."""
        rY = np.array(rY, dtype=np.float64, order="C", copy=True)
        T = np.ascontiguousarray(T, dtype=np.float64)
        rho, rH = np.zeros(T.size), np.zeros(T.size)
        lib().orc_floor_species(self.id, T.size, _p(rY), _p(rho), _p(rH), _p(T))
        return rY, rho, rH

    def heat_release(self, T, IR):
        """getHeatRelease on cells (MKS; IR (N, ncells)); returns HR (W m^-3)."""
        T, IR = (np.ascontiguousarray(x, dtype=np.float64) for x in (T, IR))
        HR = np.zeros(T.size)
        lib().orc_heat_release(self.id, T.size, _p(T), _p(IR), _p(HR))
        return HR

    def hy2t(self, h, Y, Tguess):
        Y = np.ascontiguousarray(Y, dtype=np.float64)
        T = np.array([Tguess], dtype=np.float64)
        it = lib().orc_hy2t(self.id, float(h), _p(Y), _p(T))
        return T[0], it

    def enthalpy_mass(self, T, Y):
        _, hRT, _ = self.thermo(T)
        return float(np.sum(Y * hRT * 8.31446261815324e7 * T / self.W))

    def cell(self, rhoh0, rhohsrc, rYsrc):
        src = np.ascontiguousarray(rYsrc, dtype=np.float64)
        c = Cell(self.id, float(rhoh0), float(rhohsrc), _p(src))
        c._keep = src
        return c

    def rhs(self, cell, t, y, frozen=False):
        y = np.ascontiguousarray(y, dtype=np.float64)
        f = np.zeros_like(y)
        fn = lib().orc_reactor_rhs_frozen if frozen else lib().orc_reactor_rhs
        fn(float(t), _p(y), _p(f), C.cast(C.pointer(cell), C.c_void_p))
        return f

    def jac(self, cell, t, y):
        y = np.ascontiguousarray(y, dtype=np.float64)
        n = y.size
        J = np.zeros((n, n))
        lib().orc_reactor_jac(float(t), _p(y), None, _p(J), C.cast(C.pointer(cell), C.c_void_p))
        return J

    def react(self, rY, rYsrc, T, rH, rHsrc, mask, dt, rtol=1e-10, atol=1e-10, max_order=5,
              max_steps=10000, nthreads=1, atol_vec=None, clean_init_massfrac=0):
        """Box react on component-major arrays (shape [N, ncells]); returns new arrays."""
        rY = np.array(rY, dtype=np.float64, order="C", copy=True)
        T = np.array(T, dtype=np.float64, copy=True)
        rH = np.array(rH, dtype=np.float64, copy=True)
        rYsrc = np.ascontiguousarray(rYsrc, dtype=np.float64)
        rHsrc = np.ascontiguousarray(rHsrc, dtype=np.float64)
        ncell = T.size
        FC = np.zeros(ncell)
        status = np.zeros(ncell, dtype=np.int32)
        mask = np.ascontiguousarray(mask if mask is not None else np.ones(ncell), dtype=np.int32)
        av = None
        if atol_vec is not None:
            av = np.ascontiguousarray(atol_vec, dtype=np.float64)
        o = BdfOpts(rtol, atol, _p(av) if av is not None else None, max_order, max_steps, int(clean_init_massfrac))
        nst = C.c_long(0)
        lib().orc_react(self.id, ncell, _p(rY), _p(rYsrc), _p(T), _p(rH), _p(rHsrc), _p(FC),
                        _ip(mask), float(dt), C.byref(o), int(nthreads), _ip(status), C.byref(nst))
        return dict(rY=rY, T=T, rH=rH, FC=FC, status=status, nst=nst.value)


def reaction_source(rY_old, rY_new, F, dt):
    """I_R formation after the advance (Source/PeleLMeX_Reactions.cpp:141-160):
    I_R = -(rhoY_old - rhoY_new) * dt_inv - F_rhoY with dt_inv = 1 / dt formed once, entrywise in
    IEEE float64 in that order (numpy does not contract)."""
    dt_inv = 1.0 / float(dt)
    return -(np.asarray(rY_old, dtype=np.float64) - np.asarray(rY_new, dtype=np.float64)) * dt_inv \
        - np.asarray(F, dtype=np.float64)


def advance_chemistry(mech, rY_o, rH_o, T_o, F_rY, F_rH, mask, dt, **react_kw):
    """One box of PeleLM::advanceChemistry(lev, dt, Forcing) on MKS arrays (component-major
    [N, ncells] / [ncells]): the CGS conversions of Source/PeleLMeX_Reactions.cpp:68-79, the
    oracle's react, the conversions back of :110-120 (the forcing scaled to CGS and back, as the
    reference scales it in place) and I_R of :141-160. Returns the new state, the forcing after
    its round trip, I_R, FC and status."""
    rYn = np.asarray(rY_o, dtype=np.float64) * 1.0e-3
    FY = np.asarray(F_rY, dtype=np.float64) * 1.0e-3
    rHn = np.asarray(rH_o, dtype=np.float64) * 10.0
    FH = np.asarray(F_rH, dtype=np.float64) * 10.0
    out = mech.react(rYn, FY, np.array(T_o, dtype=np.float64), rHn, FH, mask, dt, **react_kw)
    rYn = out["rY"] * 1.0e3
    FY = FY * 1.0e3
    return dict(rY=rYn, rH=out["rH"] * 0.1, T=out["T"], FY=FY, FH=FH * 0.1, FC=out["FC"], status=out["status"],
                IR=reaction_source(rY_o, rYn, FY, dt))


def advance_chemistry_chem(mech, rY, rH, T, F_rY, F_rH, mask, dt, do_reaction=True, **react_kw):
    """One box of PeleLM::advanceChemistryBAChem on the chemistry BoxArray's copy (MKS,
    component-major [N, ncells] / [ncells]; Source/PeleLMeX_Reactions.cpp:220-305): rhoY and its
    forcing times 1e-3, rho h and its forcing times 10 (:228-238); the oracle's react when
    do_reaction, else FC = 0 and the state untouched (:260-280); the state (only) back to MKS
    (:284-291). Returns the new state, the forcing copy (in CGS), FC and status."""
    rYc = np.asarray(rY, dtype=np.float64) * 1.0e-3
    FY = np.asarray(F_rY, dtype=np.float64) * 1.0e-3
    rHc = np.asarray(rH, dtype=np.float64) * 10.0
    FH = np.asarray(F_rH, dtype=np.float64) * 10.0
    T = np.array(T, dtype=np.float64)
    if do_reaction:
        out = mech.react(rYc, FY, T, rHc, FH, mask, dt, **react_kw)
        rYc, rHc, T, FC, status = out["rY"], out["rH"], out["T"], out["FC"], out["status"]
    else:
        FC, status = np.zeros(T.size), np.zeros(T.size, dtype=np.int32)
    return dict(rY=rYc * 1.0e3, rH=rHc * 0.1, T=T, FY=FY, FH=FH, FC=FC, status=status)


def exp(x):
    return lib().orc_exp(float(x))


def log(x):
    return lib().orc_log(float(x))


def lu_factor(A):
    A = np.array(A, dtype=np.float64, order="C", copy=True)
    n = A.shape[0]
    piv = np.zeros(n, dtype=np.int32)
    ier = lib().orc_lu_factor(n, _p(A), _ip(piv))
    return A, piv, ier


def lu_solve(LU, piv, b):
    b = np.array(b, dtype=np.float64, copy=True)
    lib().orc_lu_solve(LU.shape[0], _p(np.ascontiguousarray(LU)), _ip(piv), _p(b))
    return b


def bdf_solve(f, jac, y0, t0, tout, rtol, atol, max_order=5, max_steps=100000):
    """Integrate a Python-defined system with the oracle's BDF (small test problems only)."""
    y = np.array(y0, dtype=np.float64, copy=True)
    n = y.size

    def _f(t, yp, fp, ud):
        yy = np.ctypeslib.as_array(yp, shape=(n,))
        out = np.ctypeslib.as_array(fp, shape=(n,))
        out[:] = f(t, yy.copy())
        return 0

    def _j(t, yp, fyp, Jp, ud):
        yy = np.ctypeslib.as_array(yp, shape=(n,))
        out = np.ctypeslib.as_array(Jp, shape=(n * n,))
        out[:] = np.asarray(jac(t, yy.copy()), dtype=np.float64).reshape(-1)
        return 0

    cf, cj = RHS_FN(_f), JAC_FN(_j)
    av = None
    if np.ndim(atol) > 0:
        av = np.ascontiguousarray(atol, dtype=np.float64)
        o = BdfOpts(rtol, 0.0, _p(av), max_order, max_steps)
    else:
        o = BdfOpts(rtol, float(atol), None, max_order, max_steps)
    st = BdfStats()
    ier = lib().orc_bdf_solve(n, float(t0), float(tout), _p(y), cf, cj, None, C.byref(o), C.byref(st))
    return y, ier, st
