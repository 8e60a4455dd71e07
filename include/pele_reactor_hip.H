// C++ adapter with the ReactorBase member functions PeleLMeX calls, over the C-ABI of
// pele_reactor_capi.h. Header-only; templated on the caller's Box / Array4 types so it binds to
// AMReX's amrex::Box, amrex::Array4<Real> and amrex::Array4<int> without including AMReX.
//
// Reference call sites this adapter serves (PeleLMeX):
//   m_reactor = ReactorBase::create(m_chem_integrator);   Source/PeleLMeX_Setup.cpp:152-153
//   m_reactor->init(reactor_type = 2, ncells_chem = 1);    Source/PeleLMeX_Setup.cpp:154
//   m_reactor->set_eos_parm(eos_parms.device_parm());      Source/PeleLMeX_Setup.cpp:155
//   m_reactor->set_typ_vals_ode(typical_values);           Source/PeleLMeX_Utils.cpp:1472
//   m_reactor->react(bx, rhoY, frcExt, T, rhoH, frcEExt, fcl, mask, dt, time, stream)
//                                                          Source/PeleLMeX_Reactions.cpp:101-108
//   m_reactor->close();                                    Source/PeleLMeX.cpp:26
// A PelePhysics ReactorBase subclass registered under a factory name (ReactorBase::Register)
// forwards these calls to this class, which holds the C-ABI handle.
#ifndef PELE_REACTOR_HIP_H
#define PELE_REACTOR_HIP_H

#include <stdexcept>
#include <string>
#include <vector>

#include "pele_reactor_capi.h"

namespace pele_hip {

class ReactorError : public std::runtime_error {
 public:
  explicit ReactorError(const std::string& m) : std::runtime_error(m) {}
};

// Array4 view conversion: p points at (begin.x, begin.y, begin.z, 0).
template <class A4>
inline pr_array4 to_pr(const A4& a) {
  pr_array4 r;
  r.p = const_cast<double*>(a.p);
  r.jstride = a.jstride;
  r.kstride = a.kstride;
  r.nstride = a.nstride;
  r.begin[0] = a.begin.x;
  r.begin[1] = a.begin.y;
  r.begin[2] = a.begin.z;
  r.ncomp = a.ncomp;
  return r;
}

template <class I4>
inline pr_iarray4 to_pr_mask(const I4& a) {
  pr_iarray4 r;
  r.p = a.p;
  r.jstride = a.jstride;
  r.kstride = a.kstride;
  r.begin[0] = a.begin.x;
  r.begin[1] = a.begin.y;
  r.begin[2] = a.begin.z;
  return r;
}

class ReactorHipBDF {
 public:
  // integrator: the factory name PeleLMeX passes ("ReactorCvode" keeps existing inputs valid);
  // mechanism: the compile-time Chemistry_Model of the build ("drm19", "LiDryer");
  // keys: the ode.* / cvode.* / hipbdf.* ParmParse entries as "key = value" lines.
  ReactorHipBDF(std::string integrator, std::string mechanism, std::string keys)
      : integrator_(std::move(integrator)), mechanism_(std::move(mechanism)), keys_(std::move(keys)) {}
  ReactorHipBDF(const ReactorHipBDF&) = delete;
  ReactorHipBDF& operator=(const ReactorHipBDF&) = delete;
  ~ReactorHipBDF() { close(); }

  int init(int reactor_type, int ncells) {
    close();
    int rc = pr_create(integrator_.c_str(), mechanism_.c_str(), reactor_type, ncells,
                       keys_.empty() ? nullptr : keys_.c_str(), &h_);
    if (rc != 0) throw ReactorError(std::string("pr_create: ") + pr_last_error());
    return 0;
  }

  // ReactorBase::set_eos_parm(eos_parms.device_parm()) (Source/PeleLMeX_Setup.cpp:155). The
  // reactor's EOS is the ideal-gas mixture (Fuego) of the mechanism: the block is recorded, not read.
  template <class EosParm>
  void set_eos_parm(EosParm const* parm, const char* eos_model = "Fuego") {
    int rc = pr_set_eos_parm(h_, static_cast<const void*>(parm), eos_model);
    if (rc != 0) throw ReactorError(std::string("pr_set_eos_parm: ") + pr_last_error());
  }

  void set_typ_vals_ode(const std::vector<double>& typ) {
    int rc = pr_set_typ_vals(h_, typ.data(), static_cast<int>(typ.size()));
    if (rc != 0) throw ReactorError(std::string("pr_set_typ_vals: ") + pr_last_error());
  }

  // Box form of react (device arrays, CGS units). Returns the number of cells that failed to
  // integrate; like the reference, the caller may ignore it. Blocks until the stream is done.
  template <class Box, class A4, class I4>
  int react(const Box& box, A4 const& rY, A4 const& rYsrc, A4 const& T, A4 const& rE, A4 const& rEsrc,
            A4 const& FC, I4 const& mask, double& dt_react, double& time, void* stream = nullptr) {
    const int lo[3] = {box.smallEnd(0), box.smallEnd(1), box.smallEnd(2)};
    const int hi[3] = {box.bigEnd(0), box.bigEnd(1), box.bigEnd(2)};
    pr_array4 a[6] = {to_pr(rY), to_pr(rYsrc), to_pr(T), to_pr(rE), to_pr(rEsrc), to_pr(FC)};
    pr_iarray4 m = to_pr_mask(mask);
    int rc = pr_react_box(h_, lo, hi, &a[0], &a[1], &a[2], &a[3], &a[4], &a[5], &m, &dt_react, &time, stream,
                          nullptr);
    if (rc < 0) throw ReactorError(std::string("pr_react_box: ") + pr_last_error());
    return rc;
  }

  // The per-box body of PeleLM::advanceChemistry(lev, dt, Forcing)
  // (Source/PeleLMeX_Reactions.cpp:54-160) on MKS arrays: new = old in CGS, react, back to MKS,
  // with the forcing scaled in place as the reference scales it, and I_R when I_R.p != nullptr.
  template <class Box, class A4, class I4>
  int advance_chemistry(const Box& box, A4 const& rhoY_o, A4 const& rhoH_o, A4 const& temp_o, A4 const& rhoY_n,
                        A4 const& rhoH_n, A4 const& temp_n, A4 const& extF_rhoY, A4 const& extF_rhoH, A4 const& fcl,
                        I4 const& mask, A4 const& I_R, double dt, void* stream = nullptr) {
    const int lo[3] = {box.smallEnd(0), box.smallEnd(1), box.smallEnd(2)};
    const int hi[3] = {box.bigEnd(0), box.bigEnd(1), box.bigEnd(2)};
    pr_array4 a[10] = {to_pr(rhoY_o), to_pr(rhoH_o), to_pr(temp_o), to_pr(rhoY_n),    to_pr(rhoH_n),
                       to_pr(temp_n), to_pr(extF_rhoY), to_pr(extF_rhoH), to_pr(fcl), to_pr(I_R)};
    pr_iarray4 m = to_pr_mask(mask);
    int rc = pr_advance_chemistry_box(h_, lo, hi, &a[0], &a[1], &a[2], &a[3], &a[4], &a[5], &a[6], &a[7], &a[8], &m,
                                      &a[9], dt, stream, nullptr);
    if (rc < 0) throw ReactorError(std::string("pr_advance_chemistry_box: ") + pr_last_error());
    return rc;
  }

  // The per-box body of PeleLM::advanceChemistryBAChem (Source/PeleLMeX_Reactions.cpp:220-305) on
  // the chemistry BoxArray's copy (MKS, in place): do_reaction is m_baChemFlag[lev][mfi.index()].
  template <class Box, class A4, class I4>
  int advance_chemistry_chem(const Box& box, A4 const& rhoY, A4 const& rhoH, A4 const& temp, A4 const& extF_rhoY,
                             A4 const& extF_rhoH, A4 const& fcl, I4 const& mask, int do_reaction, double dt,
                             void* stream = nullptr) {
    const int lo[3] = {box.smallEnd(0), box.smallEnd(1), box.smallEnd(2)};
    const int hi[3] = {box.bigEnd(0), box.bigEnd(1), box.bigEnd(2)};
    pr_array4 a[6] = {to_pr(rhoY), to_pr(rhoH), to_pr(temp), to_pr(extF_rhoY), to_pr(extF_rhoH), to_pr(fcl)};
    pr_iarray4 m = to_pr_mask(mask);
    int rc = pr_advance_chemistry_chem_box(h_, lo, hi, &a[0], &a[1], &a[2], &a[3], &a[4], &a[5], &m, do_reaction,
                                           dt, stream, nullptr);
    if (rc < 0) throw ReactorError(std::string("pr_advance_chemistry_chem_box: ") + pr_last_error());
    return rc;
  }

  // react for callers that cannot allocate device memory (no reference counterpart;
  // pr_react_box_noalloc: the library's static per-GPU scratch, its limits in the C header)
  template <class Box, class A4, class I4>
  int react_noalloc(const Box& box, A4 const& rY, A4 const& rYsrc, A4 const& T, A4 const& rE, A4 const& rEsrc,
                    A4 const& FC, I4 const& mask, double& dt_react, double& time, void* stream = nullptr) {
    const int lo[3] = {box.smallEnd(0), box.smallEnd(1), box.smallEnd(2)};
    const int hi[3] = {box.bigEnd(0), box.bigEnd(1), box.bigEnd(2)};
    pr_array4 a[6] = {to_pr(rY), to_pr(rYsrc), to_pr(T), to_pr(rE), to_pr(rEsrc), to_pr(FC)};
    pr_iarray4 m = to_pr_mask(mask);
    int rc = pr_react_box_noalloc(h_, lo, hi, &a[0], &a[1], &a[2], &a[3], &a[4], &a[5], &m, &dt_react, &time,
                                  stream, nullptr);
    if (rc < 0) throw ReactorError(std::string("pr_react_box_noalloc: ") + pr_last_error());
    return rc;
  }

  void close() {
    if (h_) pr_destroy(h_);
    h_ = nullptr;
  }

  pr_stats stats() const {
    pr_stats s{};
    pr_get_stats(h_, &s);
    return s;
  }

 private:
  std::string integrator_, mechanism_, keys_;
  pr_handle* h_ = nullptr;
};

// The chemistry BoxArray of a level and its flags (PeleLM::resetCoveredMask,
// Source/PeleLMeX_Utils.cpp:728-788) through pr_chem_boxes: lo / hi triples of the level's grids
// and of the finer level's grids (at their resolution; empty on the finest level). Returns the
// chemistry boxes' lo / hi triples and flags (1: react, 0: covered by the finer level).
struct ChemBoxes {
  std::vector<int> lo, hi, flag;
};
inline ChemBoxes chem_boxes(const int domain_lo[3], const int domain_hi[3], const std::vector<int>& grid_lo,
                            const std::vector<int>& grid_hi, const std::vector<int>& fine_lo,
                            const std::vector<int>& fine_hi, int ref_ratio, const int max_grid_size_chem[3],
                            bool finest) {
  const int ng = static_cast<int>(grid_lo.size() / 3), nf = static_cast<int>(fine_lo.size() / 3);
  ChemBoxes out;
  int n = pr_chem_boxes(domain_lo, domain_hi, ng, grid_lo.data(), grid_hi.data(), nf, fine_lo.data(),
                        fine_hi.data(), ref_ratio, max_grid_size_chem, finest ? 1 : 0, 0, nullptr, nullptr, nullptr);
  if (n < 0) throw ReactorError(std::string("pr_chem_boxes: ") + pr_last_error());
  out.lo.resize(3 * static_cast<size_t>(n));
  out.hi.resize(3 * static_cast<size_t>(n));
  out.flag.resize(static_cast<size_t>(n));
  n = pr_chem_boxes(domain_lo, domain_hi, ng, grid_lo.data(), grid_hi.data(), nf, fine_lo.data(), fine_hi.data(),
                    ref_ratio, max_grid_size_chem, finest ? 1 : 0, n, out.lo.data(), out.hi.data(), out.flag.data());
  if (n < 0) throw ReactorError(std::string("pr_chem_boxes: ") + pr_last_error());
  return out;
}

}  // namespace pele_hip
#endif
