"""Python mirror of PelePhysics' ReactorBase for the MI355X reactor (ctypes over the C-ABI).

Reference interface (call sites in PeleLMeX):
  ReactorBase::create(name)           Source/PeleLMeX_Setup.cpp:152-153
  init(reactor_type, ncells)          Source/PeleLMeX_Setup.cpp:154
  set_typ_vals_ode(typ)               Source/PeleLMeX_Utils.cpp:1472
  react(box, rY, rYsrc, T, rE, rEsrc, FC, mask, dt_react, time, stream)
                                      Source/PeleLMeX_Reactions.cpp:101-108
  close()                             Source/PeleLMeX.cpp:26

Device arrays are torch tensors on the GPU (plumbing only: the integration runs in
libpelelmex_amd.so). Box arrays use the AMReX Array4 order, i.e. a C-contiguous tensor of
shape [ncomp, nz, ny, nx] (or [nz, ny, nx] for one component). The library must be built;
nothing here falls back to another implementation.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpelelmex_amd.so")
_LIB = None

EXPORTS = ["pr_create", "pr_set_typ_vals", "pr_react_box", "pr_react_box_noalloc", "pr_react_batch", "pr_react_batch_enqueue", "pr_collect", "pr_rhs_batch", "pr_jac_batch",
           "pr_reaction_rate_box", "pr_box_costs", "pr_chem_boxes", "pr_set_temperature_box", "pr_floor_species_box",
           "pr_heat_release_box", "pr_reaction_source_box", "pr_advance_chemistry_box",
           "pr_advance_chemistry_chem_box", "pr_set_eos_parm", "pr_surrogate_load", "pr_surrogate_features", "pr_surrogate_forward", "pr_mlp_gemm",
           "pr_mlp_gelu_backward", "pr_mlp_column_sum", "pr_mlp_mse", "pr_mlp_adam", "pr_mlp_to_bf16",
           "pr_get_stats", "pr_num_species", "pr_destroy", "pr_last_error"]


class Array4(C.Structure):
    _fields_ = [("p", C.c_void_p), ("jstride", C.c_int64), ("kstride", C.c_int64),
                ("nstride", C.c_int64), ("begin", C.c_int * 3), ("ncomp", C.c_int)]


class IArray4(C.Structure):
    _fields_ = [("p", C.c_void_p), ("jstride", C.c_int64), ("kstride", C.c_int64),
                ("begin", C.c_int * 3)]


class Box(C.Structure):
    _fields_ = [("lo", C.c_int * 3), ("hi", C.c_int * 3), ("rY", Array4), ("rYsrc", Array4), ("T", Array4),
                ("rH", Array4), ("rHsrc", Array4), ("FC", Array4), ("mask", IArray4), ("dt_react", C.c_double),
                ("status", C.c_void_p)]


class Stats(C.Structure):
    _fields_ = [("cells", C.c_int64), ("failed", C.c_int64), ("nst", C.c_int64), ("nfe", C.c_int64),
                ("nje", C.c_int64), ("nsetups", C.c_int64), ("nni", C.c_int64), ("netf", C.c_int64),
                ("ncfn", C.c_int64), ("kernel_ms", C.c_double), ("lanes_per_cell", C.c_int),
                ("blocks", C.c_int), ("sur_accepted", C.c_int64), ("sur_rejected", C.c_int64),
                ("sur_audited", C.c_int64), ("audit_failed", C.c_int64), ("total_ms", C.c_double),
                ("surrogate_ms", C.c_double), ("audit_ms", C.c_double), ("audit_max_dT_rel", C.c_double),
                ("audit_max_dY", C.c_double), ("block_threads", C.c_int), ("waves_per_simd", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def lib():
    """Load libpelelmex_amd.so; raises if it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m pelelmex_amd.build` "
                               "(or __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        L.pr_create.argtypes = [C.c_char_p, C.c_char_p, C.c_int, C.c_int, C.c_char_p, C.POINTER(C.c_void_p)]
        L.pr_create.restype = C.c_int
        L.pr_set_typ_vals.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.c_int]
        L.pr_react_box.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(Array4),
                                   C.POINTER(Array4), C.POINTER(Array4), C.POINTER(Array4), C.POINTER(Array4),
                                   C.POINTER(Array4), C.POINTER(IArray4), C.POINTER(C.c_double),
                                   C.POINTER(C.c_double), C.c_void_p, C.c_void_p]
        L.pr_react_box.restype = C.c_int
        L.pr_react_box_noalloc.argtypes = L.pr_react_box.argtypes
        L.pr_react_box_noalloc.restype = C.c_int
        L.pr_react_batch.argtypes = [C.c_void_p, C.c_int, C.POINTER(Box), C.c_void_p]
        L.pr_react_batch.restype = C.c_int
        L.pr_react_batch_enqueue.argtypes = [C.c_void_p, C.c_int, C.POINTER(Box), C.c_void_p]
        L.pr_collect.argtypes = [C.c_void_p, C.c_void_p]
        L.pr_rhs_batch.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                   C.c_double, C.c_void_p, C.c_void_p]
        L.pr_jac_batch.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                   C.c_void_p]
        L.pr_reaction_rate_box.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(Array4),
                                           C.POINTER(Array4), C.POINTER(Array4), C.POINTER(Array4),
                                           C.POINTER(IArray4), C.c_void_p]
        L.pr_reaction_rate_box.restype = C.c_int
        a4 = C.POINTER(Array4)
        L.pr_box_costs.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), a4, C.c_int,
                                   C.POINTER(C.c_double), C.c_void_p]
        L.pr_set_temperature_box.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int), a4, a4, a4, a4,
                                             C.c_void_p]
        L.pr_floor_species_box.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int), a4, a4, a4, a4,
                                           C.c_void_p]
        L.pr_heat_release_box.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int), a4, a4, a4, C.c_void_p]
        L.pr_reaction_source_box.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int), a4, a4, a4, a4,
                                             C.c_double, C.c_void_p]
        L.pr_advance_chemistry_box.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int)] + [a4] * 9 + \
            [C.POINTER(IArray4), a4, C.c_double, C.c_void_p, C.c_void_p]
        L.pr_advance_chemistry_chem_box.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int)] + \
            [a4] * 6 + [C.POINTER(IArray4), C.c_int, C.c_double, C.c_void_p, C.c_void_p]
        L.pr_set_eos_parm.argtypes = [C.c_void_p, C.c_void_p, C.c_char_p]
        vp, i64, fp = C.c_void_p, C.c_int64, C.POINTER(C.c_float)
        L.pr_surrogate_load.argtypes = [vp, C.c_int, C.POINTER(C.c_int), fp, fp, fp]
        L.pr_surrogate_features.argtypes = [vp, i64, vp, vp, vp, vp, vp, vp]
        L.pr_surrogate_forward.argtypes = [vp, i64, vp, vp, vp]
        L.pr_mlp_gemm.argtypes = [C.c_int, C.c_int, C.c_int, vp, i64, i64, vp, i64, i64, vp, C.c_int, vp, i64, vp,
                                  i64, vp, i64, C.c_int, vp]
        L.pr_mlp_gelu_backward.argtypes = [i64, C.c_int, vp, i64, vp, i64, vp, i64, vp, i64, vp]
        L.pr_mlp_column_sum.argtypes = [i64, C.c_int, vp, i64, vp, C.c_int, vp]
        L.pr_mlp_mse.argtypes = [i64, C.c_int, vp, i64, vp, i64, vp, vp, vp]
        L.pr_mlp_adam.argtypes = [i64, vp, vp, vp, vp, vp, C.c_float, C.c_float, C.c_float, C.c_float, C.c_int, vp]
        L.pr_mlp_to_bf16.argtypes = [i64, vp, vp, vp]
        ip3 = C.POINTER(C.c_int)
        L.pr_chem_boxes.argtypes = [ip3, ip3, C.c_int, ip3, ip3, C.c_int, ip3, ip3, C.c_int, ip3, C.c_int, C.c_int,
                                    ip3, ip3, ip3]
        L.pr_chem_boxes.restype = C.c_int
        L.pr_get_stats.argtypes = [C.c_void_p, C.POINTER(Stats)]
        L.pr_num_species.argtypes = [C.c_void_p]
        L.pr_destroy.argtypes = [C.c_void_p]
        L.pr_destroy.restype = None
        L.pr_last_error.restype = C.c_char_p
        _LIB = L
    return _LIB


def _keys_text(keys):
    if keys is None:
        return None
    if isinstance(keys, str):
        return keys
    return "\n".join(f"{k} = {v}" for k, v in keys.items())


def _array4(t, ncomp):
    if not t.is_contiguous():
        raise ValueError("box arrays must be contiguous [ncomp, nz, ny, nx] / [nz, ny, nx]")
    shp = t.shape[-3:]
    nz, ny, nx = shp
    return Array4(t.data_ptr(), nx, nx * ny, nx * ny * nz, (C.c_int * 3)(0, 0, 0), ncomp)


class ReactorError(RuntimeError):
    pass


class Reactor:
    """ReactorBase mirror: create(name) -> init(reactor_type, ncells) -> react(...) -> close()."""

    def __init__(self, name, mechanism, keys=None):
        self.name = name
        self.mechanism = mechanism
        self.keys = keys
        self._h = None

    @staticmethod
    def create(name="ReactorCvode", mechanism="drm19", keys=None):
        return Reactor(name, mechanism, keys)

    def init(self, reactor_type=2, ncells=1):
        h = C.c_void_p()
        txt = _keys_text(self.keys)
        rc = lib().pr_create(self.name.encode(), self.mechanism.encode(), int(reactor_type), int(ncells),
                             txt.encode() if txt else None, C.byref(h))
        if rc != 0:
            raise ReactorError(f"pr_create failed ({rc}): {lib().pr_last_error().decode()}")
        self._h = h
        self.nsp = lib().pr_num_species(h)
        return 0

    def set_typ_vals_ode(self, typ):
        import numpy as np
        a = np.ascontiguousarray(typ, dtype=np.float64)
        rc = lib().pr_set_typ_vals(self._h, a.ctypes.data_as(C.POINTER(C.c_double)), a.size)
        if rc != 0:
            raise ReactorError(lib().pr_last_error().decode())

    def react(self, rY, rYsrc, T, rH, rHsrc, FC, mask, dt_react, time=0.0, stream=None, status=None,
              noalloc=False):
        """Box react on device tensors; returns (n_failed, dt_react, time). noalloc=True calls
        pr_react_box_noalloc (no device allocation; the library's static scratch)."""
        import torch
        if self._h is None:
            raise ReactorError("react() before init()")
        nz, ny, nx = T.shape
        lo = (C.c_int * 3)(0, 0, 0)
        hi = (C.c_int * 3)(nx - 1, ny - 1, nz - 1)
        a = [_array4(rY, self.nsp), _array4(rYsrc, self.nsp), _array4(T, 1), _array4(rH, 1),
             _array4(rHsrc, 1), _array4(FC, 1)]
        m = None
        if mask is not None:
            m = IArray4(mask.data_ptr(), nx, nx * ny, (C.c_int * 3)(0, 0, 0))
        dt = C.c_double(dt_react)
        tm = C.c_double(time)
        st = stream if stream is not None else torch.cuda.current_stream()
        fn = lib().pr_react_box_noalloc if noalloc else lib().pr_react_box
        rc = fn(self._h, lo, hi, *[C.byref(x) for x in a], C.byref(m) if m is not None else None,
                C.byref(dt), C.byref(tm), C.c_void_p(st.cuda_stream),
                C.c_void_p(status.data_ptr()) if status is not None else None)
        if rc < 0:
            raise ReactorError(f"{fn.__name__} failed ({rc}): {lib().pr_last_error().decode()}")
        return rc, dt.value, tm.value

    def react_batch(self, boxes, stream=None, enqueue_only=False):
        """Several boxes in one launch. boxes: list of dicts with device tensors rY, rYsrc, T, rH,
        rHsrc, FC (shapes as for react), optional mask and status, and dt. Returns n_failed, or
        0 after only enqueueing the work when enqueue_only (then collect() waits for it)."""
        import torch
        arr = (Box * len(boxes))()
        for b, x in enumerate(boxes):
            nz, ny, nx = x["T"].shape
            m = x.get("mask")
            arr[b] = Box((C.c_int * 3)(0, 0, 0), (C.c_int * 3)(nx - 1, ny - 1, nz - 1), _array4(x["rY"], self.nsp),
                         _array4(x["rYsrc"], self.nsp), _array4(x["T"], 1), _array4(x["rH"], 1),
                         _array4(x["rHsrc"], 1), _array4(x["FC"], 1),
                         IArray4(m.data_ptr() if m is not None else None, nx, nx * ny, (C.c_int * 3)(0, 0, 0)),
                         float(x["dt"]), x["status"].data_ptr() if x.get("status") is not None else None)
        st = stream if stream is not None else torch.cuda.current_stream()
        fn = lib().pr_react_batch_enqueue if enqueue_only else lib().pr_react_batch
        rc = fn(self._h, len(boxes), arr, C.c_void_p(st.cuda_stream))
        if rc < 0:
            raise ReactorError(f"pr_react_batch failed ({rc}): {lib().pr_last_error().decode()}")
        return rc

    def collect(self, stream=None):
        """pr_collect: waits for the stream; failed cells of the last enqueued integration."""
        import torch
        st = stream if stream is not None else torch.cuda.current_stream()
        rc = lib().pr_collect(self._h, C.c_void_p(st.cuda_stream))
        if rc < 0:
            raise ReactorError(f"pr_collect failed ({rc}): {lib().pr_last_error().decode()}")
        return rc

    def rhs_batch(self, y, rYsrc, rhoh0, rhohsrc, t=0.0, stream=None):
        """Right-hand side for a batch: y [n, ncells], rYsrc [N, ncells] (device tensors)."""
        import torch
        n, nc = y.shape
        f = torch.empty_like(y)
        st = stream if stream is not None else torch.cuda.current_stream()
        rc = lib().pr_rhs_batch(self._h, nc, y.data_ptr(), rYsrc.data_ptr(), rhoh0.data_ptr(), rhohsrc.data_ptr(),
                                float(t), f.data_ptr(), C.c_void_p(st.cuda_stream))
        if rc != 0:
            raise ReactorError(lib().pr_last_error().decode())
        return f

    def jac_batch(self, y, rYsrc, rhohsrc, stream=None):
        """Analytic Jacobians for a batch; returns [n, n, ncells]."""
        import torch
        n, nc = y.shape
        J = torch.empty((n, n, nc), dtype=y.dtype, device=y.device)
        st = stream if stream is not None else torch.cuda.current_stream()
        rc = lib().pr_jac_batch(self._h, nc, y.data_ptr(), rYsrc.data_ptr(), rhohsrc.data_ptr(), J.data_ptr(),
                                C.c_void_p(st.cuda_stream))
        if rc != 0:
            raise ReactorError(lib().pr_last_error().decode())
        return J

    def reaction_rate_rhoY(self, rY, rH, T, mask=None, rYdot=None, stream=None):
        """reactionRateRhoY over a box of device tensors (MKS: rY [N, nz, ny, nx], rH and T
        [nz, ny, nx]); returns rYdot [N, nz, ny, nx] (zero where mask == -1)."""
        import torch
        nz, ny, nx = T.shape
        if rYdot is None:
            rYdot = torch.empty_like(rY)
        lo = (C.c_int * 3)(0, 0, 0)
        hi = (C.c_int * 3)(nx - 1, ny - 1, nz - 1)
        a = [_array4(rY, self.nsp), _array4(rH, 1), _array4(T, 1), _array4(rYdot, self.nsp)]
        m = IArray4(mask.data_ptr(), nx, nx * ny, (C.c_int * 3)(0, 0, 0)) if mask is not None else None
        st = stream if stream is not None else torch.cuda.current_stream()
        rc = lib().pr_reaction_rate_box(self._h, lo, hi, *[C.byref(x) for x in a],
                                        C.byref(m) if m is not None else None, C.c_void_p(st.cuda_stream))
        if rc != 0:
            raise ReactorError(f"pr_reaction_rate_box failed ({rc}): {lib().pr_last_error().decode()}")
        return rYdot

    COST_METHODS = {"Ncell": 0, "ChemFunctCallAvg": 1, "ChemFunctCallMax": 2, "ChemFunctCallSum": 3}

    def box_costs(self, fcs, method="ChemFunctCallAvg", stream=None):
        """Chemistry cost per box (PeleLM::computeCosts) from FC tensors [nz, ny, nx] (one per box);
        returns a numpy array."""
        import numpy as np
        import torch
        n = len(fcs)
        lo = (C.c_int * (3 * n))(*([0] * 3 * n))
        hi = (C.c_int * (3 * n))(*[v for f in fcs for v in (f.shape[2] - 1, f.shape[1] - 1, f.shape[0] - 1)])
        arr = (Array4 * n)(*[_array4(f, 1) for f in fcs])
        out = (C.c_double * n)()
        st = stream if stream is not None else torch.cuda.current_stream()
        rc = lib().pr_box_costs(self._h, n, lo, hi, arr, self.COST_METHODS[method], out, C.c_void_p(st.cuda_stream))
        if rc != 0:
            raise ReactorError(f"pr_box_costs failed ({rc}): {lib().pr_last_error().decode()}")
        return np.array(out[:], dtype=np.float64)

    def _eos(self, fn, arrays, shape, stream):
        import torch
        nz, ny, nx = shape
        lo = (C.c_int * 3)(0, 0, 0)
        hi = (C.c_int * 3)(nx - 1, ny - 1, nz - 1)
        st = stream if stream is not None else torch.cuda.current_stream()
        rc = getattr(lib(), fn)(self._h, lo, hi, *[C.byref(x) for x in arrays], C.c_void_p(st.cuda_stream))
        if rc != 0:
            raise ReactorError(f"{fn} failed ({rc}): {lib().pr_last_error().decode()}")

    def set_temperature(self, rho, rY, rH, T, stream=None):
        """setTemperature over a box of device tensors (MKS: rY [N, nz, ny, nx], rho, rH and T
        [nz, ny, nx]); T is updated in place (HY2T from its stored value)."""
        self._eos("pr_set_temperature_box", [_array4(rho, 1), _array4(rY, self.nsp), _array4(rH, 1), _array4(T, 1)],
                  T.shape, stream)
        return T

    def floor_species(self, rY, rho, rH, T, stream=None):
        """floorSpecies over a box: clips rY to [0, DBL_MAX] in place and rewrites rho and rH."""
        self._eos("pr_floor_species_box", [_array4(rY, self.nsp), _array4(rho, 1), _array4(rH, 1), _array4(T, 1)],
                  T.shape, stream)
        return rY, rho, rH

    def heat_release(self, T, IR, HR=None, stream=None):
        """getHeatRelease over a box: HR = -sum_k h_k(T) IR_k (W m^-3) from T [nz, ny, nx] and
        IR [N, nz, ny, nx] (kg m^-3 s^-1)."""
        import torch
        if HR is None:
            HR = torch.empty_like(T)
        self._eos("pr_heat_release_box", [_array4(T, 1), _array4(IR, self.nsp), _array4(HR, 1)], T.shape, stream)
        return HR

    def reaction_source(self, rY_old, rY_new, F, dt, IR=None, stream=None):
        """I_R = -(rY_old - rY_new) / dt - F over a box of device tensors [N, nz, ny, nx] (MKS)."""
        import torch
        if IR is None:
            IR = torch.empty_like(rY_new)
        nz, ny, nx = rY_new.shape[1:]
        lo = (C.c_int * 3)(0, 0, 0)
        hi = (C.c_int * 3)(nx - 1, ny - 1, nz - 1)
        st = stream if stream is not None else torch.cuda.current_stream()
        arrs = [_array4(x, self.nsp) for x in (rY_old, rY_new, F, IR)]
        rc = lib().pr_reaction_source_box(self._h, lo, hi, *[C.byref(x) for x in arrs], C.c_double(float(dt)),
                                          C.c_void_p(st.cuda_stream))
        if rc != 0:
            raise ReactorError(f"pr_reaction_source_box failed ({rc}): {lib().pr_last_error().decode()}")
        return IR

    def advance_chemistry(self, old, new, F_rY, F_rH, FC, mask, dt, IR=None, stream=None, status=None):
        """One box of advanceChemistry on MKS device tensors: old / new are dicts of rY [N, nz, ny, nx],
        rH and T [nz, ny, nx] (new is written), the forcing F_rY / F_rH is scaled to CGS and back in
        place, FC, mask (int32 or None), I_R written when given. Returns the failed-cell count."""
        import torch
        nz, ny, nx = new["T"].shape
        lo = (C.c_int * 3)(0, 0, 0)
        hi = (C.c_int * 3)(nx - 1, ny - 1, nz - 1)
        a = [_array4(old["rY"], self.nsp), _array4(old["rH"], 1), _array4(old["T"], 1), _array4(new["rY"], self.nsp),
             _array4(new["rH"], 1), _array4(new["T"], 1), _array4(F_rY, self.nsp), _array4(F_rH, 1), _array4(FC, 1)]
        m = IArray4(mask.data_ptr(), nx, nx * ny, (C.c_int * 3)(0, 0, 0)) if mask is not None else None
        ir = _array4(IR, self.nsp) if IR is not None else Array4(None, 0, 0, 0, (C.c_int * 3)(0, 0, 0), 0)
        st = stream if stream is not None else torch.cuda.current_stream()
        rc = lib().pr_advance_chemistry_box(self._h, lo, hi, *[C.byref(x) for x in a],
                                            C.byref(m) if m is not None else None, C.byref(ir), C.c_double(float(dt)),
                                            C.c_void_p(st.cuda_stream),
                                            C.c_void_p(status.data_ptr()) if status is not None else None)
        if rc < 0:
            raise ReactorError(f"pr_advance_chemistry_box failed ({rc}): {lib().pr_last_error().decode()}")
        return rc

    def advance_chemistry_chem(self, rY, rH, T, F_rY, F_rH, FC, mask, dt, do_reaction=True, stream=None,
                               status=None):
        """One box of advanceChemistryBAChem on the chemistry BoxArray's copy (device tensors, MKS,
        in place): rY [N, nz, ny, nx], rH, T, FC [nz, ny, nx], the forcing copy F_rY / F_rH (left in
        CGS, as the reference leaves it); do_reaction False = a covered box (FC = 0, no
        integration). Returns the failed-cell count."""
        import torch
        nz, ny, nx = T.shape
        lo = (C.c_int * 3)(0, 0, 0)
        hi = (C.c_int * 3)(nx - 1, ny - 1, nz - 1)
        a = [_array4(rY, self.nsp), _array4(rH, 1), _array4(T, 1), _array4(F_rY, self.nsp), _array4(F_rH, 1),
             _array4(FC, 1)]
        m = IArray4(mask.data_ptr(), nx, nx * ny, (C.c_int * 3)(0, 0, 0)) if mask is not None else None
        st = stream if stream is not None else torch.cuda.current_stream()
        rc = lib().pr_advance_chemistry_chem_box(self._h, lo, hi, *[C.byref(x) for x in a],
                                                 C.byref(m) if m is not None else None, int(bool(do_reaction)),
                                                 C.c_double(float(dt)), C.c_void_p(st.cuda_stream),
                                                 C.c_void_p(status.data_ptr()) if status is not None else None)
        if rc < 0:
            raise ReactorError(f"pr_advance_chemistry_chem_box failed ({rc}): {lib().pr_last_error().decode()}")
        return rc

    def set_eos_parm(self, eos_parm=None, eos_model="Fuego"):
        """ReactorBase::set_eos_parm: records the EOS parameter block (device pointer or None)."""
        rc = lib().pr_set_eos_parm(self._h, eos_parm, eos_model.encode() if eos_model else None)
        if rc != 0:
            raise ReactorError(f"pr_set_eos_parm failed ({rc}): {lib().pr_last_error().decode()}")

    def load_surrogate(self, dims, params, xnorm, ynorm):
        """pr_surrogate_load: dims (layer widths), params (fp32 W0, b0, W1, b1, ... flattened),
        xnorm / ynorm ((mean, sd) stacked); host arrays."""
        import numpy as np
        d = np.ascontiguousarray(dims, dtype=np.int32)
        p = np.ascontiguousarray(params, dtype=np.float32)
        xn = np.ascontiguousarray(xnorm, dtype=np.float32)
        yn = np.ascontiguousarray(ynorm, dtype=np.float32)
        fp = C.POINTER(C.c_float)
        rc = lib().pr_surrogate_load(self._h, d.size - 1, d.ctypes.data_as(C.POINTER(C.c_int)), p.ctypes.data_as(fp),
                                     xn.ctypes.data_as(fp), yn.ctypes.data_as(fp))
        if rc != 0:
            raise ReactorError(f"pr_surrogate_load failed ({rc}): {lib().pr_last_error().decode()}")
        self.surrogate_dims = [int(x) for x in d]

    def surrogate_forward(self, X, stream=None):
        """The loaded network on normalised bf16 features X [nc, dims[0]] (device); fp32 [nc, dims[-1]]."""
        import torch
        nc = X.shape[0]
        nout = self.surrogate_dims[-1]
        Y = torch.empty((nc, nout), dtype=torch.float32, device=X.device)
        st = stream if stream is not None else torch.cuda.current_stream()
        rc = lib().pr_surrogate_forward(self._h, nc, X.data_ptr(), Y.data_ptr(), C.c_void_p(st.cuda_stream))
        if rc != 0:
            raise ReactorError(f"pr_surrogate_forward failed ({rc}): {lib().pr_last_error().decode()}")
        return Y

    def surrogate_features(self, rY, T, dt, normalised=False, stream=None):
        """Network inputs for a batch (device tensors rY [N, nc], T [nc], dt [nc]): raw fp32
        [nc, N + 2], or normalised bf16 with a loaded network."""
        import torch
        nc = T.numel()
        st = stream if stream is not None else torch.cuda.current_stream()
        if normalised:
            out = torch.empty((nc, self.nsp + 2), dtype=torch.bfloat16, device=T.device)
            args = (None, out.data_ptr())
        else:
            out = torch.empty((nc, self.nsp + 2), dtype=torch.float32, device=T.device)
            args = (out.data_ptr(), None)
        rc = lib().pr_surrogate_features(self._h, nc, rY.data_ptr(), T.data_ptr(), dt.data_ptr(), *args,
                                         C.c_void_p(st.cuda_stream))
        if rc != 0:
            raise ReactorError(f"pr_surrogate_features failed ({rc}): {lib().pr_last_error().decode()}")
        return out

    def stats(self):
        s = Stats()
        lib().pr_get_stats(self._h, C.byref(s))
        return s.as_dict()

    def close(self):
        if self._h is not None:
            lib().pr_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def chem_boxes(domain, grids, fine=(), ref_ratio=2, max_grid_size_chem=(0, 0, 0), finest=False):
    """The chemistry BoxArray of a level (pr_chem_boxes; PeleLM::resetCoveredMask,
    Source/PeleLMeX_Utils.cpp:728-788). domain: (lo, hi); grids / fine: lists of (lo, hi) with
    inclusive cell-index triples, fine at the finer level's resolution. Returns (boxes, flags):
    the uncovered (reacted, flag 1) boxes chopped to max_grid_size_chem, then the coarsened
    finer-level boxes (flag 0); on the finest level its grids chopped, all flag 1."""
    import numpy as np
    def arr(bs, k):
        a = np.ascontiguousarray(np.array([b[k] for b in bs], dtype=np.int32).reshape(-1, 3))
        return a
    dlo = (C.c_int * 3)(*domain[0])
    dhi = (C.c_int * 3)(*domain[1])
    glo, ghi = arr(grids, 0), arr(grids, 1)
    flo, fhi = arr(fine, 0), arr(fine, 1)
    ms = (C.c_int * 3)(*max_grid_size_chem)
    ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int))
    cap = 0
    while True:
        blo = np.zeros((max(cap, 1), 3), dtype=np.int32)
        bhi = np.zeros((max(cap, 1), 3), dtype=np.int32)
        fl = np.zeros(max(cap, 1), dtype=np.int32)
        n = lib().pr_chem_boxes(dlo, dhi, len(grids), ip(glo), ip(ghi), len(fine), ip(flo), ip(fhi), int(ref_ratio), ms,
                                int(bool(finest)), cap, ip(blo), ip(bhi), ip(fl))
        if n < 0:
            raise ReactorError(f"pr_chem_boxes failed ({n}): {lib().pr_last_error().decode()}")
        if n <= cap:
            return [(tuple(int(x) for x in blo[b]), tuple(int(x) for x in bhi[b])) for b in range(n)], \
                [int(x) for x in fl[:n]]
        cap = n


def advance_chemistry_bachem(reactor, domain_lo, old, new, F_rY, F_rH, FC, IR, dt, boxes, flags, mask=None,
                             stream=None):
    """PeleLM::advanceChemistryBAChem for one level (Source/PeleLMeX_Reactions.cpp:180-367) on
    device tensors laid out over the level's index space from domain_lo (MKS): old / new dicts of
    rY [N, nz, ny, nx], rH, T [nz, ny, nx]; the forcing F_rY, F_rH; FC and I_R written. Each
    chemistry box (boxes, flags from chem_boxes) gets a contiguous copy of the old state and of
    the forcing (ParallelCopy, :207-208), the per-box body pr_advance_chemistry_chem_box (react
    when flagged, else FC = 0), and its state and FC copied back into new / FC (:311-312); I_R =
    -(rY_old - rY_new) dt_inv - F_rY over the boxes (:322-347). Returns the failed-cell count."""
    import torch
    nfail = 0
    dt_inv = 1.0 / float(dt)
    for (lo, hi), flag in zip(boxes, flags):
        sl = tuple(slice(lo[d] - domain_lo[d], hi[d] - domain_lo[d] + 1) for d in (2, 1, 0))
        cY = old["rY"][(slice(None),) + sl].contiguous()
        cH = old["rH"][sl].contiguous()
        cT = old["T"][sl].contiguous()
        fY = F_rY[(slice(None),) + sl].contiguous()
        fH = F_rH[sl].contiguous()
        cF = torch.empty_like(cT)
        m = mask[sl].contiguous() if mask is not None else None
        nfail += reactor.advance_chemistry_chem(cY, cH, cT, fY, fH, cF, m, dt, do_reaction=bool(flag), stream=stream)
        new["rY"][(slice(None),) + sl] = cY
        new["rH"][sl] = cH
        new["T"][sl] = cT
        FC[sl] = cF
        oY = old["rY"][(slice(None),) + sl]
        IR[(slice(None),) + sl] = -(oY - cY) * dt_inv - F_rY[(slice(None),) + sl]
    return nfail


def typical_values_chem(rY_mks, T, Ymin=1.0e-10):
    """The typical values PeleLMeX hands the reactor (set_typ_vals_ode), from a state in MKS as
    the level MultiFabs hold it: rY_mks [N, ncells] (kg m^-3), T [ncells] (K).

    PeleLM::setTypicalValues (Source/PeleLMeX_Utils.cpp:1363-1392): rho_typ = (max rho + min rho)/2,
    Y_typ,k = (max rhoY_k + min rhoY_k) / 2 / rho_typ, T_typ = (max T + min T) / 2 (rho = sum_k rhoY_k
    here; PeleLMeX keeps it as its own component). PeleLM::updateTypicalValuesChem (:1444-1475):
    typ_k = max(Ymin rho_typ 1e-3, Y_typ,k rho_typ 1e-3) (g cm^-3), typ_N = T_typ, with Ymin =
    peleLM.typical_values_Ymin (default 1e-10, Source/PeleLMeX.H:2003). The reactor then integrates
    with atol_i = ode.atol * typ_i (ode.atol is the pre-factor of the typical values,
    Docs/sphinx/manual/LMeXControls.rst:304)."""
    import numpy as np
    rY_mks = np.asarray(rY_mks, dtype=np.float64)
    T = np.asarray(T, dtype=np.float64)
    rho = rY_mks.sum(0)
    rho_typ = 0.5 * (rho.max() + rho.min())
    Y_typ = 0.5 * (rY_mks.max(1) + rY_mks.min(1)) / rho_typ
    typ = np.empty(rY_mks.shape[0] + 1)
    typ[:-1] = np.maximum(Ymin * rho_typ * 1.0e-3, Y_typ * rho_typ * 1.0e-3)
    typ[-1] = 0.5 * (T.max() + T.min())
    return typ


def box_tensors(box, device="cuda"):
    """Device tensors in Array4 order for a workload dict from pelelmex_amd.workloads."""
    import torch
    nx, ny, nz = box["dims"]
    N = box["rY"].shape[0]

    def t(a, nc):
        x = torch.as_tensor(a, dtype=torch.float64)
        x = x.reshape(nc, nz, ny, nx) if nc > 1 else x.reshape(nz, ny, nx)
        return x.to(device).contiguous()

    out = dict(rY=t(box["rY"], N), rYsrc=t(box["rYsrc"], N), T=t(box["T"], 1), rH=t(box["rH"], 1),
               rHsrc=t(box["rHsrc"], 1))
    out["FC"] = torch.zeros_like(out["T"])
    out["mask"] = torch.as_tensor(box["mask"], dtype=torch.int32).reshape(nz, ny, nx).to(device).contiguous()
    return out
