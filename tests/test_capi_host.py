"""Host-side checks of the C-ABI library and the Python mirror (CPU only: no GPU compute).

* libpelelmex_amd.so loads and exports every function include/pele_reactor_capi.h declares.
* pr_create consumes the reference's reactor keys and rejects malformed or unknown ones, an
  unsupported reactor_type, unknown mechanisms and integrators, before touching the device.
* the seeded workload generators are deterministic and produce physical states.
"""
import ctypes as C
import os
import re

import numpy as np
import pytest

from pelelmex_amd import reactor as R
from pelelmex_amd import workloads as W
from pelelmex_amd.mechanisms import PATM, load

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pele_reactor_capi.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pr_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    if not os.path.exists(R.LIB_PATH):
        pytest.skip("library not built")
    L = R.lib()
    names = declared_functions()
    assert names, "no declarations found"
    for nm in names:
        assert hasattr(L, nm), nm
    assert set(names) == set(R.EXPORTS)


def _create(integrator=b"ReactorCvode", mech=b"drm19", rtype=2, keys=None):
    h = C.c_void_p()
    rc = R.lib().pr_create(integrator, mech, rtype, 1, keys, C.byref(h))
    return rc, h


@pytest.mark.parametrize("keys,code", [
    (b"ode.rtol = -1", -4),
    (b"ode.atol=abc", -4),
    (b"cvode.max_order = 9", -4),
    (b"ode.unknown_key = 1", -4),
    (b"hipbdf.block = 100", -4),
    (b"justtext", -4),
    (b"ode.clean_init_massfrac = 2", -4),
    (b"surrogate.trust_rtol = 0", -4),
    (b"surrogate.audit_frac = 1.5", -4),
    (b"surrogate.unknown = 1", -4),
    (b"hipbdf.lanes_per_cell = 3", -2),
])
def test_pr_create_key_errors(keys, code):
    if not os.path.exists(R.LIB_PATH):
        pytest.skip("library not built")
    rc, h = _create(keys=keys)
    assert rc == code, (rc, R.lib().pr_last_error())
    assert not h.value
    assert R.lib().pr_last_error()


def test_pr_create_rejects_unsupported():
    if not os.path.exists(R.LIB_PATH):
        pytest.skip("library not built")
    assert _create(rtype=1)[0] == -2
    assert _create(mech=b"gri30")[0] == -2
    assert _create(mech=b"dodecane_lu")[0] in (0, -3)      # supported; -3 without a HIP device
    assert _create(integrator=b"ReactorRK64")[0] == -2
    # keys of other namespaces pass through untouched; a key error is reported before any HIP call
    rc, _ = _create(keys=b"peleLM.chem_integrator = ReactorCvode\node.rtol = 1e-8 cvode.max_order=4")
    assert rc in (0, -3)
    rc, _ = _create(keys=b"ode.clean_init_massfrac = 1 surrogate.audit_frac = 0.02 surrogate.enable = 0")
    assert rc in (0, -3)


def test_workloads_deterministic_and_physical():
    a = W.ch4_air_flame_box(nx=16, ny=2, nz=2)
    b = W.ch4_air_flame_box(nx=16, ny=2, nz=2)
    for k in ("rY", "rYsrc", "T", "rH", "rHsrc"):
        np.testing.assert_array_equal(a[k], b[k])
    m = load("drm19")
    rho = a["rY"].sum(0)
    Y = a["rY"] / rho
    assert np.all(a["rY"] >= 0) and np.allclose(Y.sum(0), 1.0, rtol=1e-14)
    np.testing.assert_allclose(rho, m.density(PATM, a["T"], Y.T), rtol=1e-12)
    np.testing.assert_allclose(a["rH"] / rho, m.h_mass(a["T"], Y.T), rtol=1e-12)
    i = np.arange(a["T"].size) % 16
    x = (i + 0.5) / 16
    assert np.all(a["T"][x < 0.3] == 300.0)
    assert np.all((a["T"][(x >= 0.3) & (x < 0.5)] >= 1200) & (a["T"][(x >= 0.3) & (x < 0.5)] <= 2200))
    assert np.all(a["T"][x >= 0.5] >= 1800)
    h = W.h2_air_box(nx=8, ny=2, nz=2)
    assert np.all((h["T"] >= 1000) & (h["T"] <= 1800)) and np.all(h["rYsrc"] == 0)


def test_cpp_adapter_compiles_and_links(tmp_path):
    """include/pele_reactor_hip.H against stand-ins of amrex::Box / Array4, linked to the library."""
    if not os.path.exists(R.LIB_PATH):
        pytest.skip("library not built")
    src = tmp_path / "adapter.cpp"
    src.write_text(r'''
#include "pele_reactor_hip.H"
#include <cstdio>
struct Dim3 { int x, y, z; };
struct IntVect { int v[3]; int operator[](int i) const { return v[i]; } };
struct Box { IntVect lo, hi; int smallEnd(int d) const { return lo[d]; } int bigEnd(int d) const { return hi[d]; } };
template <class T> struct Array4 { T* p; long long jstride, kstride, nstride; Dim3 begin, end; int ncomp; };
int main() {
  const int dlo[3] = {0, 0, 0}, dhi[3] = {15, 15, 15}, mgs[3] = {8, 8, 8};
  pele_hip::ChemBoxes cb = pele_hip::chem_boxes(dlo, dhi, {0, 0, 0}, {15, 15, 15}, {8, 8, 8}, {15, 15, 15}, 2, mgs,
                                                false);
  if (cb.flag.size() < 2 || cb.flag.back() != 0 || cb.flag.front() != 1) return 2;
  pele_hip::ReactorHipBDF r("ReactorCvode", "drm19", "ode.rtol = -5");
  try { r.init(2, 1); } catch (const pele_hip::ReactorError& e) { std::printf("%s\n", e.what()); return 0; }
  return 1;
}
template int pele_hip::ReactorHipBDF::react<Box, Array4<double>, Array4<int>>(
    const Box&, Array4<double> const&, Array4<double> const&, Array4<double> const&, Array4<double> const&,
    Array4<double> const&, Array4<double> const&, Array4<int> const&, double&, double&, void*);
template int pele_hip::ReactorHipBDF::react_noalloc<Box, Array4<double>, Array4<int>>(
    const Box&, Array4<double> const&, Array4<double> const&, Array4<double> const&, Array4<double> const&,
    Array4<double> const&, Array4<double> const&, Array4<int> const&, double&, double&, void*);
struct EosParmStandIn { int dummy; };
template void pele_hip::ReactorHipBDF::set_eos_parm<EosParmStandIn>(EosParmStandIn const*, const char*);
template int pele_hip::ReactorHipBDF::advance_chemistry_chem<Box, Array4<double>, Array4<int>>(
    const Box&, Array4<double> const&, Array4<double> const&, Array4<double> const&, Array4<double> const&,
    Array4<double> const&, Array4<double> const&, Array4<int> const&, int, double, void*);
template int pele_hip::ReactorHipBDF::advance_chemistry<Box, Array4<double>, Array4<int>>(
    const Box&, Array4<double> const&, Array4<double> const&, Array4<double> const&, Array4<double> const&,
    Array4<double> const&, Array4<double> const&, Array4<double> const&, Array4<double> const&,
    Array4<double> const&, Array4<int> const&, Array4<double> const&, double, void*);
''')
    import subprocess
    exe = tmp_path / "adapter"
    libdir = os.path.dirname(R.LIB_PATH)
    subprocess.run(["g++", "-std=c++17", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe),
                    "-L", libdir, "-lpelelmex_amd", f"-Wl,-rpath,{libdir}"], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True)
    assert "ode.rtol" in out.stdout


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus 2 exits with status 2 and says so, before touching a GPU."""
    import subprocess
    import sys
    out = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                         text=True, timeout=120)
    assert out.returncode == 2
    assert "one MI355X" in out.stderr
