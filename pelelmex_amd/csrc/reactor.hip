// Host side of the MI355X reactor: the C-ABI of include/pele_reactor_capi.h.
//
// pr_create      <- ReactorBase::create + init + the ode.* / cvode.* ParmParse keys
//                   (Source/PeleLMeX_Setup.cpp:146-155, Docs/sphinx/manual/LMeXControls.rst:300-308)
// pr_set_typ_vals<- ReactorBase::set_typ_vals_ode (Source/PeleLMeX_Utils.cpp:1444-1475)
// pr_react_box   <- ReactorBase::react(box, ...) (Source/PeleLMeX_Reactions.cpp:101-108)
// pr_destroy     <- ReactorBase::close (Source/PeleLMeX.cpp:26)
#include <hip/hip_runtime.h>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/pele_reactor_capi.h"
#include "surrogate_kernel.h"
#include "chem_boxes.h"
#include "eos_kernel.h"
#include "mech/drm19.h"
#include "mech/lidryer.h"
#include "mech/syn53.h"
#include "mech/dodecane_lu.h"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(x)                                                                        \
  do {                                                                                   \
    hipError_t e_ = (x);                                                                 \
    if (e_ != hipSuccess) return fail(PR_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

const pr_mech_t* const kMechs[] = {&pr_lidryer, &pr_drm19, &pr_syn53, &pr_dodecane_lu};

const pr_mech_t* find_mech(const char* name) {
  if (!name) return nullptr;
  for (const pr_mech_t* m : kMechs) {
    const char* a = m->name;
    const char* b = name;
    while (*a && *b && std::tolower((unsigned char)*a) == std::tolower((unsigned char)*b)) ++a, ++b;
    if (*a == 0 && *b == 0) return m;
  }
  return nullptr;
}

using Launch = hipError_t (*)(const prk::KParams&, int blocks, int block, size_t shmem, hipStream_t);
using EvalLaunch = hipError_t (*)(int mode, const unsigned char* blob, int blob_bytes, int neff, int ninc, long ncells,
                                  const double* y, const double* rYsrc, const double* rhoh0, const double* rhohsrc,
                                  double t, double* out, double* gscr, int blocks, int block, size_t shmem,
                                  hipStream_t st);
using Occ = hipError_t (*)(int* nb, int block, size_t shmem);
using RateLaunch = hipError_t (*)(const prk::RateParams&, int blocks, int block, size_t shmem, hipStream_t);
using TrustLaunch = hipError_t (*)(const prk::TrustParams&, int blocks, int block, size_t shmem, hipStream_t);
using FeatLaunch = hipError_t (*)(const prk::FeatParams&, hipStream_t);
using PrepLaunch = hipError_t (*)(const prk::PrepParams&, int blocks, int block, size_t shmem, hipStream_t);
using EosLaunch = hipError_t (*)(int mode, const prk::EosParams&, int blocks, int block, size_t shmem, hipStream_t);

template <int N, int R, int G, int W>
hipError_t launch_t(const prk::KParams& P, int blocks, int block, size_t shmem, hipStream_t st) {
  hipLaunchKernelGGL((prk::react_kernel<N, R, G, W>), dim3(blocks), dim3(block), shmem, st, P);
  return hipGetLastError();
}

template <int N, int R, int G>
hipError_t eval_t(int mode, const unsigned char* blob, int blob_bytes, int neff, int ninc, long ncells, const double* y,
                  const double* rYsrc, const double* rhoh0, const double* rhohsrc, double t, double* out, double* gscr,
                  int blocks, int block, size_t shmem, hipStream_t st) {
  if (mode == 0)
    hipLaunchKernelGGL((prk::eval_kernel<N, R, G, 0>), dim3(blocks), dim3(block), shmem, st, blob, blob_bytes, neff,
                       ninc, ncells, y, rYsrc, rhoh0, rhohsrc, t, out, gscr);
  else
    hipLaunchKernelGGL((prk::eval_kernel<N, R, G, 1>), dim3(blocks), dim3(block), shmem, st, blob, blob_bytes, neff,
                       ninc, ncells, y, rYsrc, rhoh0, rhohsrc, t, out, gscr);
  return hipGetLastError();
}

template <int N, int R, int G>
hipError_t rate_t(const prk::RateParams& P, int blocks, int block, size_t shmem, hipStream_t st) {
  hipLaunchKernelGGL((prk::rate_kernel<N, R, G>), dim3(blocks), dim3(block), shmem, st, P);
  return hipGetLastError();
}

template <int N, int R, int G>
hipError_t trust_t(const prk::TrustParams& P, int blocks, int block, size_t shmem, hipStream_t st) {
  hipLaunchKernelGGL((prk::trust_kernel<N, R, G>), dim3(blocks), dim3(block), shmem, st, P);
  return hipGetLastError();
}

template <int N, int R, int G>
hipError_t prep_t(const prk::PrepParams& P, int blocks, int block, size_t shmem, hipStream_t st) {
  hipLaunchKernelGGL((prk::prep_kernel<N, R, G>), dim3(blocks), dim3(block), shmem, st, P);
  return hipGetLastError();
}

template <int N, int R, int G>
hipError_t eos_t(int mode, const prk::EosParams& P, int blocks, int block, size_t shmem, hipStream_t st) {
  if (mode == prk::EOS_SET_T)
    hipLaunchKernelGGL((prk::eos_cell_kernel<N, R, prk::EOS_SET_T>), dim3(blocks), dim3(block), shmem, st, P);
  else if (mode == prk::EOS_FLOOR)
    hipLaunchKernelGGL((prk::eos_cell_kernel<N, R, prk::EOS_FLOOR>), dim3(blocks), dim3(block), shmem, st, P);
  else
    hipLaunchKernelGGL((prk::eos_cell_kernel<N, R, prk::EOS_HEAT_RELEASE>), dim3(blocks), dim3(block), shmem, st, P);
  return hipGetLastError();
}

template <int N>
hipError_t feat_t(const prk::FeatParams& P, hipStream_t st) {
  if (P.ncells <= 0) return hipSuccess;
  hipLaunchKernelGGL((prk::featurize_kernel<N>), dim3((unsigned)((P.ncells + 127) / 128)), dim3(128), 0, st, P);
  return hipGetLastError();
}

template <int N, int R, int G, int W>
hipError_t occ_t(int* nb, int block, size_t shmem) {
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(
      nb, reinterpret_cast<const void*>(&prk::react_kernel<N, R, G, W>), block, shmem);
}

template <int N, int R, int G>
constexpr int scratch_bytes() {
  return prk::group_doubles<N, R, G>() * 8;
}

struct Variant {
  const char* mech;
  int G;
  int waves;    // waves per SIMD react_kernel is compiled for; blocks of up to 256 * waves threads
  Launch launch;
  EvalLaunch eval;
  Occ occ;
  int scratch;  // LDS bytes per cell group
  int slot;     // global scratch bytes per cell group
  RateLaunch rate;
  TrustLaunch trust;
  FeatLaunch feat;
  PrepLaunch prep;
  EosLaunch eos;
};

#define PR_VARIANT(NAME, NS, NR, G, W) \
  {NAME, G, W, &launch_t<NS, NR, G, W>, &eval_t<NS, NR, G>, &occ_t<NS, NR, G, W>, scratch_bytes<NS, NR, G>(), \
   prk::Cell<NS, NR, G>::gslot_doubles * 8, &rate_t<NS, NR, G>, &trust_t<NS, NR, G>, &feat_t<NS>, &prep_t<NS, NR, G>, &eos_t<NS, NR, G>}

// drm19 at 32 lanes keeps its LU factors in the global slot (the group's LDS then holds 2 KB
// instead of 6 KB), which lets 4 waves per SIMD be resident: 32 cells in a 1024-thread block, 44
// VGPRs spilled, 7 % faster than 3 waves with the factors in LDS (the integrator is latency
// bound). At 64 lanes the group's LDS does not fit 3 waves; at 16 lanes 4 cells share a wave and
// registers already spill at 2. LiDryer at 16 lanes fits 3 waves without spills (21% faster than
// 2 on a 32^3 box); at 8 lanes 4 waves spill and are slower. A 54-component state (syn53) keeps
// its LU factors in its global slot too: 223 VGPRs at 2 waves (no spills), 45 spilled at 3.
const Variant kVariants[] = {
    PR_VARIANT("LiDryer", PR_LIDRYER_NSP, PR_LIDRYER_NR, 16, 3),
    PR_VARIANT("LiDryer", PR_LIDRYER_NSP, PR_LIDRYER_NR, 8, 2),
    PR_VARIANT("drm19", PR_DRM19_NSP, PR_DRM19_NR, 32, 4),
    PR_VARIANT("drm19", PR_DRM19_NSP, PR_DRM19_NR, 32, 3),
    PR_VARIANT("drm19", PR_DRM19_NSP, PR_DRM19_NR, 16, 2),
    PR_VARIANT("drm19", PR_DRM19_NSP, PR_DRM19_NR, 64, 2),
    PR_VARIANT("syn53", PR_SYN53_NSP, PR_SYN53_NR, 64, 2),
    PR_VARIANT("syn53", PR_SYN53_NSP, PR_SYN53_NR, 64, 3),
    PR_VARIANT("dodecane_lu", PR_DODECANE_LU_NSP, PR_DODECANE_LU_NR, 64, 2),
    PR_VARIANT("dodecane_lu", PR_DODECANE_LU_NSP, PR_DODECANE_LU_NR, 64, 3),
};

const Variant* find_variant(const char* mech, int G, int W = 0) {
  for (const Variant& v : kVariants)
    if (std::strcmp(v.mech, mech) == 0 && (G == 0 || v.G == G) && (W == 0 || v.waves == W)) return &v;
  return nullptr;
}

std::vector<unsigned char> pack_blob(const pr_mech_t* m) {
  prk::BlobLayout L = prk::blob_layout(m->nsp, m->nr, m->neff, m->ninc);
  std::vector<unsigned char> b(L.bytes, 0);
  auto* sp = reinterpret_cast<prk::SpRec*>(b.data() + L.off_sp);
  for (int k = 0; k < m->nsp; ++k) {
    prk::SpRec& o = sp[k];
    for (int rg = 0; rg < 2; ++rg) {
      const double* a = m->nasa + 15 * k + (rg == 0 ? 8 : 1);
      for (int c = 0; c < 5; ++c) o.cp[rg][c] = a[c];
      o.h[rg][0] = a[0];
      o.h[rg][1] = a[1] / 2.0;
      o.h[rg][2] = a[2] / 3.0;
      o.h[rg][3] = a[3] / 4.0;
      o.h[rg][4] = a[4] / 5.0;
      o.h[rg][5] = a[5];
      o.s[rg][0] = a[0];
      o.s[rg][1] = a[1];
      o.s[rg][2] = a[2] / 2.0;
      o.s[rg][3] = a[3] / 3.0;
      o.s[rg][4] = a[4] / 4.0;
      o.s[rg][5] = a[6];
      o.dcp[rg][0] = a[1];
      o.dcp[rg][1] = 2.0 * a[2];
      o.dcp[rg][2] = 3.0 * a[3];
      o.dcp[rg][3] = 4.0 * a[4];
    }
    o.W = m->W[k];
    o.invW = 1.0 / m->W[k];
    o.RW = PR_RU / m->W[k];
    o.tmid = m->nasa[15 * k];
  }
  // Reactions are stored grouped by kind (Troe, Lindemann, third body, elementary; stable within
  // a kind) so that the lanes of one rate round mostly take the same code path. Incidence lists
  // keep the mechanism's reaction order and point at the stored slots, so production rates and
  // Jacobian rows are still summed in mechanism order.
  std::vector<int> slot_of(m->nr);
  {
    int next = 0;
    for (int kind = PR_RX_TROE; kind >= PR_RX_ELEM; --kind)
      for (int r = 0; r < m->nr; ++r)
        if (m->rtype[r] == kind) slot_of[r] = next++;
  }
  auto* rx = reinterpret_cast<prk::RxRec*>(b.data() + L.off_rx);
  for (int r = 0; r < m->nr; ++r) {
    prk::RxRec& o = rx[slot_of[r]];
    o.lnA = m->arr[3 * r];
    o.beta = m->arr[3 * r + 1];
    o.Ta = m->arr[3 * r + 2];
    o.lnA0 = m->low[3 * r];
    o.beta0 = m->low[3 * r + 1];
    o.Ta0 = m->low[3 * r + 2];
    for (int c = 0; c < 4; ++c) o.troe[c] = m->troe[4 * r + c];
    o.effd = m->eff_dflt[r];
    o.pad = 0.0;
    o.info = m->rtype[r] | (m->rev[r] << 4) | (m->nreac[r] << 8) | (m->nprod[r] << 12) | (m->ntroe[r] << 16);
    int ra = 0, pa = 0;
    for (int s = 0; s < m->nreac[r]; ++s) ra |= m->reac[3 * r + s] << (8 * s);
    for (int s = 0; s < m->nprod[r]; ++s) pa |= m->prod[3 * r + s] << (8 * s);
    o.reac = ra;
    o.prod = pa;
    o.eff = m->eff_start[r] | (m->eff_start[r + 1] << 16);
  }
  auto* esp = reinterpret_cast<int*>(b.data() + L.off_effsp);
  auto* eval = reinterpret_cast<double*>(b.data() + L.off_effval);
  for (int e = 0; e < m->neff; ++e) {
    esp[e] = m->eff_sp[e];
    eval[e] = m->eff_val[e];
  }
  auto* incs = reinterpret_cast<int*>(b.data() + L.off_incs);
  auto* inc = reinterpret_cast<int*>(b.data() + L.off_inc);
  for (int k = 0; k <= m->nsp; ++k) incs[k] = m->inc_start[k];
  for (int e = 0; e < m->ninc; ++e) inc[e] = (slot_of[m->inc_r[e]] & 0xFFFF) | (m->inc_nu[e] * 65536);
  // the production-rate gather: per entry the byte offset of its rate slot and nu as a double,
  // and the incidence lists cut into PR_GATHER_BLOCK-entry blocks
  auto* incq = reinterpret_cast<int*>(b.data() + L.off_incq);
  auto* incnu = reinterpret_cast<double*>(b.data() + L.off_incnu);
  auto* kblk = reinterpret_cast<int*>(b.data() + L.off_kblk);
  auto* blk = reinterpret_cast<int*>(b.data() + L.off_blk);
  for (int e = 0; e < m->ninc; ++e) {
    incq[e] = slot_of[m->inc_r[e]] * (int)sizeof(double);
    incnu[e] = (double)m->inc_nu[e];
  }
  int nb = 0;
  for (int k = 0; k < m->nsp; ++k) {
    kblk[k] = nb;
    for (int e = m->inc_start[k]; e < m->inc_start[k + 1]; e += PR_GATHER_BLOCK) blk[nb++] = e;
  }
  kblk[m->nsp] = nb;
  blk[nb] = m->ninc;
  return b;
}

struct Keys {
  double rtol = 1e-10, atol = 1e-10;
  int max_order = 2, max_substeps = 10000;
  std::string solve_type = "GMRES";
  int clean_init_massfrac = 0, analytical_jacobian = 1, verbose = 0;
  int lanes_per_cell = 0, block = 0, autotune = 0, waves = 0;
  int cost_order = 1;  // hand cells out by decreasing input FC (PeleLMeX passes the previous advance's)
  // learned map in front of the BDF (used once weights are loaded with pr_surrogate_load)
  int surr_enable = 1;
  double surr_trust_rtol = 1e-3, surr_trust_atol = 1e-9, surr_elem_tol = 0.05, surr_audit_frac = 0.01;
  int surr_audit_seed = 12345;
  // 1: the learned path with every increment zero (the source step alone, no network pass): the
  // predictor the network has to beat, under the same trust check, fallback and audit
  int surr_zero = 0;
};

int parse_keys(const char* text, Keys& K) {
  if (!text) return 0;
  std::string s(text);
  for (char& c : s)
    if (c == '\n' || c == ';' || c == '\t') c = ' ';
  // normalise "a = b" to "a=b"
  std::string t;
  for (size_t i = 0; i < s.size(); ++i) {
    if (s[i] == ' ') {
      size_t j = i;
      while (j < s.size() && s[j] == ' ') ++j;
      bool before = !t.empty() && t.back() == '=';
      bool after = j < s.size() && s[j] == '=';
      if (!before && !after) t.push_back(' ');
      i = j - 1;
    } else {
      t.push_back(s[i]);
    }
  }
  std::istringstream in(t);
  std::string tok;
  while (in >> tok) {
    size_t eq = tok.find('=');
    if (eq == std::string::npos) return fail(PR_ERR_KEY, "malformed key/value: " + tok);
    std::string k = tok.substr(0, eq), v = tok.substr(eq + 1);
    if (!v.empty() && v.front() == '"') v = v.substr(1);
    if (!v.empty() && v.back() == '"') v.pop_back();
    char* end = nullptr;
    auto num = [&](double& d) {
      d = std::strtod(v.c_str(), &end);
      return end && *end == 0 && !v.empty();
    };
    double d = 0;
    if (k == "ode.rtol") {
      if (!num(d) || !(d > 0)) return fail(PR_ERR_KEY, "bad ode.rtol");
      K.rtol = d;
    } else if (k == "ode.atol") {
      if (!num(d) || !(d > 0)) return fail(PR_ERR_KEY, "bad ode.atol");
      K.atol = d;
    } else if (k == "cvode.max_order") {
      if (!num(d) || d < 1 || d > 5) return fail(PR_ERR_KEY, "cvode.max_order must be 1..5");
      K.max_order = (int)d;
    } else if (k == "cvode.max_substeps") {
      if (!num(d) || d < 1) return fail(PR_ERR_KEY, "bad cvode.max_substeps");
      K.max_substeps = (int)d;
    } else if (k == "cvode.solve_type") {
      K.solve_type = v;
    } else if (k == "ode.clean_init_massfrac") {
      if (!num(d)) return fail(PR_ERR_KEY, "bad ode.clean_init_massfrac");
      K.clean_init_massfrac = (int)d;
    } else if (k == "ode.analytical_jacobian") {
      if (!num(d)) return fail(PR_ERR_KEY, "bad ode.analytical_jacobian");
      K.analytical_jacobian = (int)d;
    } else if (k == "ode.verbose") {
      if (!num(d)) return fail(PR_ERR_KEY, "bad ode.verbose");
      K.verbose = (int)d;
    } else if (k == "hipbdf.lanes_per_cell") {
      if (!num(d)) return fail(PR_ERR_KEY, "bad hipbdf.lanes_per_cell");
      K.lanes_per_cell = (int)d;
    } else if (k == "hipbdf.waves_per_simd") {
      if (!num(d) || d < 0 || d > 4 || d != (int)d) return fail(PR_ERR_KEY, "hipbdf.waves_per_simd must be 0 to 4");
      K.waves = (int)d;
    } else if (k == "hipbdf.cost_order") {
      if (!num(d) || (d != 0 && d != 1)) return fail(PR_ERR_KEY, "hipbdf.cost_order must be 0 or 1");
      K.cost_order = (int)d;
    } else if (k == "hipbdf.autotune") {
      if (!num(d) || (d != 0 && d != 1)) return fail(PR_ERR_KEY, "hipbdf.autotune must be 0 or 1");
      K.autotune = (int)d;
    } else if (k == "hipbdf.block") {
      if (!num(d) || ((int)d % 64) != 0 || d < 64 || d > 1024)
        return fail(PR_ERR_KEY, "hipbdf.block must be a multiple of 64 from 64 to 1024");
      K.block = (int)d;
    } else if (k == "surrogate.enable") {
      if (!num(d)) return fail(PR_ERR_KEY, "bad surrogate.enable");
      K.surr_enable = (int)d;
    } else if (k == "surrogate.trust_rtol") {
      if (!num(d) || !(d > 0)) return fail(PR_ERR_KEY, "bad surrogate.trust_rtol");
      K.surr_trust_rtol = d;
    } else if (k == "surrogate.trust_atol") {
      if (!num(d) || !(d > 0)) return fail(PR_ERR_KEY, "bad surrogate.trust_atol");
      K.surr_trust_atol = d;
    } else if (k == "surrogate.elem_tol") {
      if (!num(d) || !(d >= 0)) return fail(PR_ERR_KEY, "bad surrogate.elem_tol");
      K.surr_elem_tol = d;
    } else if (k == "surrogate.audit_frac") {
      if (!num(d) || d < 0 || d > 1) return fail(PR_ERR_KEY, "surrogate.audit_frac must be in [0, 1]");
      K.surr_audit_frac = d;
    } else if (k == "surrogate.zero_increment") {
      if (!num(d) || (d != 0 && d != 1)) return fail(PR_ERR_KEY, "surrogate.zero_increment must be 0 or 1");
      K.surr_zero = (int)d;
    } else if (k == "surrogate.audit_seed") {
      if (!num(d)) return fail(PR_ERR_KEY, "bad surrogate.audit_seed");
      K.surr_audit_seed = (int)d;
    } else if (k.rfind("ode.", 0) == 0 || k.rfind("cvode.", 0) == 0 || k.rfind("hipbdf.", 0) == 0 ||
               k.rfind("surrogate.", 0) == 0) {
      return fail(PR_ERR_KEY, "unknown reactor key " + k);
    }
    // keys of other namespaces are not the reactor's and are left alone
  }
  return 0;
}

}  // namespace

// Loaded network and the per-box buffers of the surrogate path.
struct Surrogate {
  std::vector<int> dims;            // layer widths, dims[0] = N + 2, back = N + 1
  std::vector<__bf16*> W;           // [dims[l] x dims[l+1]] bf16
  std::vector<__bf16*> Wt;          // W^T [dims[l+1] x dims[l]] bf16 (hidden-layer forward kernel)
  std::vector<float*> b;            // [dims[l+1]]
  float *xmu = nullptr, *xsd = nullptr, *ymu = nullptr, *ysd = nullptr, *yeps = nullptr;
  double *Mel = nullptr, *Minv = nullptr;
  int E = 0;
  long cap = 0;                     // cells the buffers hold
  __bf16 *X = nullptr, *H0 = nullptr, *H1 = nullptr;
  float* Yout = nullptr;
  double* Tstar = nullptr;
  double* scale = nullptr;
  int* rej = nullptr;
  unsigned int* cnt = nullptr;      // [0] rejected [1] audited [2] accepted [3] fallback queue [4] audit queue
  unsigned int aud_cap = 0;
  double *aud_in = nullptr, *aud_out = nullptr, *aud_box = nullptr, *aud_err = nullptr;
  int* aud_cell = nullptr;
  double* aud_dt = nullptr;         // [aud_cap] dt_react of each audited cell
  std::vector<void*> owned;
};

struct pr_handle {
  // Calls on one handle are serialized (its buffers, counters and statistics are shared state);
  // recursive because entry points call each other (pr_advance_chemistry_box -> pr_react_box).
  mutable std::recursive_mutex mu;
  int batch_surr = 0;            // the last pr_react_batch_enqueue ran the learned map (pr_collect reads its counters)
  unsigned int batch_acap = 0;   // and audited up to this many cells
  const pr_mech_t* mech = nullptr;
  const Variant* var = nullptr;
  Keys keys;
  int device = 0;
  int block = 128;
  unsigned char* d_blob = nullptr;
  int blob_bytes = 0;
  double* d_atolv = nullptr;
  void* d_counters = nullptr;  // two regions of 128 B: queue (4 B), nfail (4 B), stats (8 x 8 B at +64)
  double* d_gscr = nullptr;    // per-group global scratch slots
  size_t gscr_bytes = 0;
  bool use_typ = false;
  pr_stats last{};
  hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr, ev3 = nullptr;
  Surrogate* surr = nullptr;
  bool tuned = false;        // hipbdf.autotune has chosen var / block
  int* d_order = nullptr;     // hipbdf.cost_order: cell order of the box (cells, capacity order_cap)
  unsigned int* d_hist = nullptr;  // its bucket counts and cursors (2 x COST_BUCKETS)
  long order_cap = 0;
  void* d_boxtab = nullptr;  // device box table (descriptors, then cell prefix sums)
  size_t boxtab_bytes = 0;
  // Set once pr_react_batch_enqueue has been captured into a graph: the graph's nodes hold this
  // handle's device buffers, so buffers a later call outgrows are retired (freed at pr_destroy)
  // instead of freed, and the graph keeps replaying on the ones it captured.
  bool captured = false;
  std::vector<void*> retired;
  const void* eos_parm = nullptr;  // set_eos_parm's block: recorded, not read (ideal-gas EOS)
};

static int resolve_noalloc_scratch(int device);

// Frees a device buffer of the handle, or keeps it until pr_destroy once a captured graph may
// reference it.
static void release(pr_handle* h, void* p) {
  if (!p) return;
  if (h->captured)
    h->retired.push_back(p);
  else
    (void)hipFree(p);
}

extern "C" {

const char* pr_last_error(void) { return g_err.c_str(); }

int pr_create(const char* integrator, const char* mechanism, int reactor_type, int ncells,
              const char* keyvals, pr_handle** out) {
  (void)ncells;
  if (!out) return fail(PR_ERR_ARG, "out is NULL");
  *out = nullptr;
  if (!integrator || (std::strcmp(integrator, "ReactorCvode") != 0 && std::strcmp(integrator, "ReactorHipBDF") != 0))
    return fail(PR_ERR_UNSUPPORTED, std::string("unsupported integrator: ") + (integrator ? integrator : "(null)"));
  if (reactor_type != 2)
    return fail(PR_ERR_UNSUPPORTED, "only reactor_type 2 (constant pressure, rho h) is implemented");
  const pr_mech_t* m = find_mech(mechanism);
  if (!m) return fail(PR_ERR_UNSUPPORTED, std::string("unknown mechanism: ") + (mechanism ? mechanism : "(null)"));
  Keys K;
  int rc = parse_keys(keyvals, K);
  if (rc) return rc;
  if (K.clean_init_massfrac != 0 && K.clean_init_massfrac != 1)
    return fail(PR_ERR_KEY, "ode.clean_init_massfrac must be 0 or 1");
  int G = K.lanes_per_cell;
  if (G == 0) G = (m->nsp + 1 <= 16) ? 16 : (m->nsp + 1 <= 32 ? 32 : 64);
  const Variant* v = find_variant(m->name, G, K.waves);
  if (!v)
    return fail(PR_ERR_UNSUPPORTED,
                "no kernel variant for this mechanism, hipbdf.lanes_per_cell and hipbdf.waves_per_simd");
  auto* h = new pr_handle();
  h->mech = m;
  h->var = v;
  h->keys = K;
  h->block = K.block;
  hipError_t e = hipGetDevice(&h->device);
  if (e != hipSuccess) {
    delete h;
    return fail(PR_ERR_HIP, std::string("hipGetDevice: ") + hipGetErrorString(e));
  }
  if (int rc = resolve_noalloc_scratch(h->device)) {
    delete h;
    return rc;
  }
  std::vector<unsigned char> blob = pack_blob(m);
  h->blob_bytes = (int)blob.size();
  {
    // the gather's block sums are held in the group's 5 N species doubles of LDS
    const prk::BlobLayout L = prk::blob_layout(m->nsp, m->nr, m->neff, m->ninc);
    const int nb = reinterpret_cast<const int*>(blob.data() + L.off_kblk)[m->nsp];
    if (nb > 5 * m->nsp) {
      delete h;
      return fail(PR_ERR_UNSUPPORTED, "mechanism has more production-rate gather blocks than 5 per species");
    }
  }
  // per-component absolute tolerances: ode.atol everywhere until set_typ_vals_ode scales them
  // (the kernels always read this array, so a captured graph sees later typical values)
  const std::vector<double> atol0((size_t)m->nsp + 1, K.atol);
  if (hipMalloc(&h->d_blob, blob.size()) != hipSuccess || hipMalloc(&h->d_counters, 256) != hipSuccess ||
      hipMalloc(&h->d_atolv, sizeof(double) * (m->nsp + 1)) != hipSuccess ||
      hipMemcpy(h->d_atolv, atol0.data(), sizeof(double) * atol0.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(h->d_blob, blob.data(), blob.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipEventCreate(&h->ev0) != hipSuccess || hipEventCreate(&h->ev1) != hipSuccess ||
      hipEventCreate(&h->ev2) != hipSuccess || hipEventCreate(&h->ev3) != hipSuccess) {
    pr_destroy(h);
    return fail(PR_ERR_HIP, "device allocation failed");
  }
  *out = h;
  return 0;
}

int pr_set_typ_vals(pr_handle* h, const double* typ, int n) {
  if (!h || !typ) return fail(PR_ERR_ARG, "NULL argument");
  std::lock_guard<std::recursive_mutex> lock(h->mu);
  if (n != h->mech->nsp + 1) return fail(PR_ERR_ARG, "typical values need NUM_SPECIES + 1 entries");
  std::vector<double> at(n);
  for (int i = 0; i < n; ++i) {
    if (!(typ[i] > 0)) return fail(PR_ERR_ARG, "typical values must be positive");
    at[i] = h->keys.atol * typ[i];
  }
  // no launch of this handle may still be reading the old values (graph replays included)
  HIPCHK(hipSetDevice(h->device));
  HIPCHK(hipDeviceSynchronize());
  HIPCHK(hipMemcpy(h->d_atolv, at.data(), sizeof(double) * n, hipMemcpyHostToDevice));
  h->use_typ = true;
  return 0;
}

int pr_num_species(const pr_handle* h) { return h ? h->mech->nsp : PR_ERR_ARG; }

// Largest block (64..256 threads) whose LDS (tables + per-group scratch) stays within 80 KiB,
// so that two blocks share a CU's 160 KiB.
// LDS per CU on gfx950
constexpr size_t kLdsPerCU = 163840;

// Block size of the one-cell-per-group helper kernels (launch bound prk::BLOCK): the largest
// that lets two blocks share a CU's LDS.
static int pick_block(const pr_handle* h) {
  if (h->block && h->block <= prk::BLOCK) return h->block;
  const int G = h->var->G;
  for (int b = prk::BLOCK; b >= 64; b -= 64)
    if ((size_t)h->blob_bytes + (size_t)(b / G) * h->var->scratch <= kLdsPerCU / 2) return b;
  return 64;
}

// Block size of the BDF kernel: hipbdf.block when set, else the largest block (up to
// 256 * waves threads) whose LDS (mechanism tables + per-group areas) lets the variant's waves
// per SIMD be resident, i.e. ceil(256 * waves / b) blocks per CU.
static int bdf_block(const pr_handle* h) {
  if (h->block) return h->block;
  const int G = h->var->G, W = h->var->waves;
  for (int b = 256 * W; b >= 64; b -= 64) {
    const size_t per_cu = (size_t)((256 * W + b - 1) / b);
    if ((size_t)h->blob_bytes + (size_t)(b / G) * h->var->scratch <= kLdsPerCU / per_cu) return b;
  }
  return 64;
}

static int ensure_gscr(pr_handle* h, long slots) {
  const size_t need = (size_t)slots * (size_t)h->var->slot;
  if (need <= h->gscr_bytes) return 0;
  release(h, h->d_gscr);
  h->d_gscr = nullptr;
  h->gscr_bytes = 0;
  HIPCHK(hipMalloc(&h->d_gscr, need));
  h->gscr_bytes = need;
  return 0;
}

static bool arr_ok(const pr_array4* a, int ncomp) { return a && a->p && (ncomp <= 1 || a->nstride != 0); }

static prk::Arr to_arr(const pr_array4* a) {
  prk::Arr r;
  r.p = a->p;
  r.js = a->jstride;
  r.ks = a->kstride;
  r.ns = a->nstride;
  r.bx = a->begin[0];
  r.by = a->begin[1];
  r.bz = a->begin[2];
  return r;
}

// Launches the persistent BDF kernel for up to work_max queue entries; zeroes P.queue, P.stats
// and P.nfail first (all on the stream).
// With gscr given (the no-allocation path), the group slots come from that buffer of gscr_cap
// bytes and the grid is capped to the groups it holds; otherwise the handle's slots grow to fit.
static int bdf_launch(pr_handle* h, prk::KParams& P, long work_max, hipStream_t st, double* gscr = nullptr,
                      size_t gscr_cap = 0) {
  const int G = h->var->G;
  const int block = bdf_block(h);
  if (block > 256 * h->var->waves)
    return fail(PR_ERR_KEY, "hipbdf.block " + std::to_string(block) + " exceeds " + std::to_string(256 * h->var->waves) +
                                " threads for " + std::to_string(G) + " lanes per cell");
  const size_t shmem = (size_t)h->blob_bytes + (size_t)(block / G) * h->var->scratch;
  int dev_cus = 0, nb = 0;
  HIPCHK(hipDeviceGetAttribute(&dev_cus, hipDeviceAttributeMultiprocessorCount, h->device));
  HIPCHK(h->var->occ(&nb, block, shmem));
  if (nb < 1) return fail(PR_ERR_HIP, "kernel does not fit on a CU with this block / LDS size");
  long cap = (long)nb * dev_cus;
  if (gscr) {
    const long fit = (long)(gscr_cap / ((size_t)(block / G) * (size_t)h->var->slot));
    if (fit < 1) return fail(PR_ERR_ARG, "scratch too small for one block of this kernel variant");
    cap = fit < cap ? fit : cap;
  }
  long want = (work_max + (block / G) - 1) / (block / G);
  if (want < 1) want = 1;
  const int blocks = (int)(want < cap ? want : cap);
  if (gscr) {
    P.gscr = gscr;
  } else {
    int rc = ensure_gscr(h, (long)blocks * (block / G));
    if (rc) return rc;
    P.gscr = h->d_gscr;
  }
  hipLaunchKernelGGL(prk::zero_counters, dim3(1), dim3(64), 0, st, P.queue, P.stats, P.nfail);
  HIPCHK(hipGetLastError());
  HIPCHK(h->var->launch(P, blocks, block, shmem, st));
  h->last.lanes_per_cell = G;
  h->last.blocks = blocks;
  h->last.block_threads = block;
  h->last.waves_per_simd = h->var->waves;
  return 0;
}

// Copies box descriptors and their cell prefix sums to the device table. The bytes travel in
// the arguments of store_words launches (no host staging buffer), so a captured launch sequence
// replays the table of its capture even after other calls restage the handle's table.
// With fixed (the no-allocation path), the table goes to that buffer of fixed_cap bytes.
static int stage_boxes(pr_handle* h, const std::vector<prk::BoxDesc>& d, const std::vector<long>& start,
                       hipStream_t st, const prk::BoxDesc** dd, const long** ds, void* fixed = nullptr,
                       size_t fixed_cap = 0) {
  const size_t nd = d.size() * sizeof(prk::BoxDesc), ns = start.size() * sizeof(long);
  const size_t off = (nd + 255) / 256 * 256, bytes = off + ns;
  if (fixed && bytes > fixed_cap) return fail(PR_ERR_ARG, "box table exceeds its scratch area");
  void* table = fixed ? fixed : h->d_boxtab;
  if (!fixed && bytes > h->boxtab_bytes) {
    release(h, h->d_boxtab);
    h->d_boxtab = nullptr;
    h->boxtab_bytes = 0;
    HIPCHK(hipMalloc(&h->d_boxtab, bytes));
    h->boxtab_bytes = bytes;
    table = h->d_boxtab;
  }
  std::vector<unsigned long long> words((bytes + 7) / 8, 0ull);
  std::memcpy(words.data(), d.data(), nd);
  std::memcpy(reinterpret_cast<unsigned char*>(words.data()) + off, start.data(), ns);
  auto* dst = static_cast<unsigned long long*>(table);
  for (size_t w0 = 0; w0 < words.size(); w0 += prk::kStoreWords) {
    const int n = (int)std::min<size_t>(prk::kStoreWords, words.size() - w0);
    prk::StoreWords chunk;
    std::memcpy(chunk.w, words.data() + w0, sizeof(unsigned long long) * (size_t)n);
    hipLaunchKernelGGL(prk::store_words, dim3(1), dim3(256), 0, st, dst + w0, n, chunk);
  }
  HIPCHK(hipGetLastError());
  *dd = static_cast<const prk::BoxDesc*>(table);
  *ds = reinterpret_cast<const long*>(static_cast<unsigned char*>(table) + off);
  return 0;
}

static int surr_alloc(Surrogate* S, void** p, size_t bytes) {
  HIPCHK(hipMalloc(p, bytes));
  S->owned.push_back(*p);
  return 0;
}

// Buffers of the surrogate path for boxes of up to ncells cells.
static int ensure_surr_buffers(pr_handle* h, long ncells) {
  Surrogate* S = h->surr;
  if (ncells <= S->cap) return 0;
  const int N = h->mech->nsp;
  void** bufs[] = {(void**)&S->X, (void**)&S->H0, (void**)&S->H1, (void**)&S->Yout, (void**)&S->Tstar,
                   (void**)&S->scale, (void**)&S->rej, (void**)&S->aud_in, (void**)&S->aud_out, (void**)&S->aud_box,
                   (void**)&S->aud_err, (void**)&S->aud_cell, (void**)&S->aud_dt};
  for (void** q : bufs) {
    if (*q) {
      release(h, *q);
      for (auto it = S->owned.begin(); it != S->owned.end(); ++it)
        if (*it == *q) {
          S->owned.erase(it);
          break;
        }
    }
    *q = nullptr;  // no freed address survives a failed reallocation below
  }
  S->cap = 0;
  int hid = 1;
  for (size_t l = 1; l + 1 < S->dims.size(); ++l) hid = S->dims[l] > hid ? S->dims[l] : hid;
  const double frac = h->keys.surr_audit_frac;
  const long acap = frac > 0 ? (long)(ncells * frac * 1.5) + 64 : 0;
  S->aud_cap = (unsigned int)(acap < ncells ? acap : ncells);
  const long ac = S->aud_cap > 0 ? S->aud_cap : 1;
  int rc = 0;
  if ((rc = surr_alloc(S, (void**)&S->X, sizeof(__bf16) * ncells * (N + 2))) ||
      (rc = surr_alloc(S, (void**)&S->H0, sizeof(__bf16) * ncells * hid)) ||
      (rc = surr_alloc(S, (void**)&S->H1, sizeof(__bf16) * ncells * hid)) ||
      (rc = surr_alloc(S, (void**)&S->Yout, sizeof(float) * ncells * (N + 1))) ||
      (rc = surr_alloc(S, (void**)&S->rej, sizeof(int) * ncells)) ||
      (rc = surr_alloc(S, (void**)&S->Tstar, sizeof(double) * ncells)) ||
      (rc = surr_alloc(S, (void**)&S->scale, sizeof(double) * ncells * (N + 1))) ||
      (rc = surr_alloc(S, (void**)&S->aud_in, sizeof(double) * ac * (2 * N + 3))) ||
      (rc = surr_alloc(S, (void**)&S->aud_out, sizeof(double) * ac * (N + 1))) ||
      (rc = surr_alloc(S, (void**)&S->aud_box, sizeof(double) * ac * (2 * N + 4))) ||
      (rc = surr_alloc(S, (void**)&S->aud_err, sizeof(double) * 2 * ac)) ||
      (rc = surr_alloc(S, (void**)&S->aud_cell, sizeof(int) * ac)) ||
      (rc = surr_alloc(S, (void**)&S->aud_dt, sizeof(double) * ac)))
    return rc;
  S->cap = ncells;
  return 0;
}

static hipError_t gemm_launch(const prk::GemmParams& G, hipStream_t st) {
  if (G.M <= 0 || G.N <= 0) return hipSuccess;
  dim3 grid((unsigned)((G.N + prk::GT_N - 1) / prk::GT_N), (unsigned)((G.M + prk::GT_M - 1) / prk::GT_M));
  hipLaunchKernelGGL(prk::gemm_bf16_kernel, grid, dim3(256), 0, st, G);
  return hipGetLastError();
}

// Network forward over feature rows [r0, r0 + m) of S->X: hidden layers GELU into bf16, last
// layer fp32 into the same rows of S->Yout.
static int mlp_forward_rows(pr_handle* h, long r0, long m, hipStream_t st) {
  Surrogate* S = h->surr;
  const int nl = (int)S->dims.size() - 1;
  const __bf16* in = S->X + r0 * S->dims[0];
  for (int l = 0; l < nl; ++l) {
    prk::GemmParams G{};
    G.M = (int)m;
    G.N = S->dims[l + 1];
    G.K = S->dims[l];
    G.A = in;
    G.sam = G.K;
    G.sak = 1;
    G.B = S->W[l];
    G.sbk = G.N;
    G.sbn = 1;
    G.bias = S->b[l];
    const bool last = (l == nl - 1);
    G.epi = last ? prk::EPI_NONE : prk::EPI_GELU;
    __bf16* out = ((l % 2 == 0) ? S->H0 : S->H1) + r0 * G.N;
    if (last) {
      G.C = S->Yout + r0 * G.N;
      G.ldc = G.N;
    } else {
      G.Cb = out;
      G.ldcb = G.N;
    }
    if (!last && G.K % prk::GT_K == 0 && G.N % prk::FT_N == 0 && G.K % 8 == 0) {
      // hidden layer: the tiled forward kernel (same bits as gemm_bf16_kernel)
      dim3 grid((unsigned)(G.N / prk::FT_N), (unsigned)((G.M + prk::FT_M - 1) / prk::FT_M));
      hipLaunchKernelGGL(prk::gemm_bf16_fwd_kernel, grid, dim3(256), 0, st, G.M, G.N, G.K, in, S->Wt[l], G.bias,
                         G.epi, out, (long)G.N);
      HIPCHK(hipGetLastError());
    } else {
      HIPCHK(gemm_launch(G, st));
    }
    in = out;
  }
  return 0;
}

// Network forward over ncells feature rows S->X into S->Yout.
// Rows go in chunks whose row tiles fit the grid's y dimension (at most 65535 blocks).
static int mlp_forward(pr_handle* h, long ncells, hipStream_t st) {
  constexpr long kRows = 32767L * 128;  // a multiple of both kernels' row tiles, 65535 * 64 at most
  for (long r0 = 0; r0 < ncells; r0 += kRows) {
    const long m = ncells - r0 < kRows ? ncells - r0 : kRows;
    if (int rc = mlp_forward_rows(h, r0, m, st)) return rc;
  }
  return 0;
}

// hipbdf.autotune: on the first react call, every lanes-per-cell variant of the mechanism, with
// its own block size and with 128 and 256 threads, integrates the same evenly spaced sample (up to 65536
// cells) of the box (a copy:
// the box itself is untouched); the fastest per cell is kept for the handle. All variants give
// identical results, so the choice only affects speed.
static int autotune(pr_handle* h, const prk::KParams& Pbase, const prk::BoxDesc& B, hipStream_t st) {
  if (!h->keys.autotune || h->tuned) return 0;
  h->tuned = true;
  const int N = h->mech->nsp;
  // the sample must be many times the cells resident on the GPU at once, or the launch tail
  // (not the steady state) decides the ranking
  const long S = B.ncells < 65536 ? B.ncells : 65536;
  if (S < 4096) return 0;  // too small to tell variants apart
  double* buf = nullptr;
  const size_t per = (size_t)(2 * N + 4);
  HIPCHK(hipMalloc(&buf, sizeof(double) * S * per * 2));
  double* pristine = buf;
  double* work = buf + S * per;
  hipLaunchKernelGGL(prk::gather_sample_kernel, dim3((unsigned)((S + 127) / 128)), dim3(128), 0, st, B, N, S,
                     pristine);
  hipError_t e = hipGetLastError();
  const Variant* best_v = h->var;
  int best_b = h->block;
  float best_ms = 1e30f;
  prk::BoxDesc sd{};
  sd.nx = (int)S;
  sd.ny = 1;
  sd.ncells = S;
  auto arr1 = [&](double* p, long ns) {
    prk::Arr r;
    r.p = p;
    r.js = S;
    r.ks = S;
    r.ns = ns;
    r.bx = r.by = r.bz = 0;
    return r;
  };
  sd.rY = arr1(work, S);
  sd.rYsrc = arr1(work + (long)N * S, S);
  sd.T = arr1(work + 2L * N * S, 0);
  sd.rH = arr1(work + (2L * N + 1) * S, 0);
  sd.rHsrc = arr1(work + (2L * N + 2) * S, 0);
  sd.FC = arr1(work + (2L * N + 3) * S, 0);
  sd.dt = B.dt;
  std::vector<prk::BoxDesc> descs(1, sd);
  std::vector<long> starts = {0L, S};
  const Variant* v0 = h->var;
  const int b0 = h->block;
  for (const Variant& v : kVariants) {
    if (e != hipSuccess) break;
    if (std::strcmp(v.mech, h->mech->name) != 0) continue;
    for (int blk : {0, 256, 128}) {  // 0: the variant's own choice (bdf_block)
      h->var = &v;
      h->block = 0;
      if (blk != 0 && blk == bdf_block(h)) continue;
      h->block = blk;
      prk::KParams P = Pbase;
      const prk::BoxDesc* db = nullptr;
      const long* dsx = nullptr;
      if ((e = hipMemcpyAsync(work, pristine, sizeof(double) * S * per, hipMemcpyDeviceToDevice, st)) != hipSuccess) break;
      int rc = stage_boxes(h, descs, starts, st, &db, &dsx);
      if (rc) {
        e = hipErrorUnknown;
        break;
      }
      P.boxes = db;
      P.box_start = dsx;
      P.nbox = 1;
      P.ncells = S;
      P.cell_list = nullptr;
      P.ncells_dev = nullptr;
      if ((e = hipEventRecord(h->ev2, st)) != hipSuccess) break;
      rc = bdf_launch(h, P, S, st);
      if (rc) {
        e = hipErrorUnknown;
        break;
      }
      if ((e = hipEventRecord(h->ev3, st)) != hipSuccess || (e = hipStreamSynchronize(st)) != hipSuccess) break;
      float ms = 0.f;
      if ((e = hipEventElapsedTime(&ms, h->ev2, h->ev3)) != hipSuccess) break;
      if (ms < best_ms) {
        best_ms = ms;
        best_v = &v;
        best_b = blk;
      }
    }
  }
  (void)hipStreamSynchronize(st);
  (void)hipFree(buf);
  h->var = (e == hipSuccess) ? best_v : v0;
  h->block = (e == hipSuccess) ? best_b : b0;
  if (e != hipSuccess) return fail(PR_ERR_HIP, std::string("autotune: ") + hipGetErrorString(e));
  return 0;
}

// hipbdf.cost_order: the launch's queue entries (all, or the listed ones) handed out by
// decreasing input FC, the previous advance's cost; P.cell_list becomes the ordered list (the
// listed count P.ncells_dev is unchanged). No memset and no allocation once sized for ncells.
static int cost_order_launch(pr_handle* h, prk::KParams& P, long ncells, hipStream_t st, int* order = nullptr,
                             unsigned int* hist = nullptr) {
  if (order && hist) {  // the no-allocation path: caller-sized buffers
    hipLaunchKernelGGL(prk::zero_u32, dim3(1), dim3(256), 0, st, hist, 2 * prk::COST_BUCKETS);
  } else if (h->order_cap < ncells) {
    release(h, h->d_order);
    h->d_order = nullptr;
    h->order_cap = 0;
    HIPCHK(hipMalloc(&h->d_order, sizeof(int) * (size_t)ncells));
    h->order_cap = ncells;
  }
  if (!(order && hist)) {
    if (!h->d_hist) HIPCHK(hipMalloc(&h->d_hist, sizeof(unsigned int) * 2 * prk::COST_BUCKETS));
    hipLaunchKernelGGL(prk::zero_u32, dim3(1), dim3(256), 0, st, h->d_hist, 2 * prk::COST_BUCKETS);
    order = h->d_order;
    hist = h->d_hist;
  }
  const int* list = P.cell_list;
  const unsigned int* nlist = P.ncells_dev;
  const unsigned hgrid = (unsigned)((ncells + 255) / 256 < 1024 ? (ncells + 255) / 256 : 1024);
  const unsigned sgrid = (unsigned)((ncells + 255) / 256);  // one entry per thread
  hipLaunchKernelGGL(prk::cost_hist_kernel, dim3(hgrid), dim3(256), 0, st, P.boxes, P.box_start, P.nbox, list, nlist,
                     ncells, hist);
  hipLaunchKernelGGL(prk::cost_scan_kernel, dim3(1), dim3(64), 0, st, hist);
  hipLaunchKernelGGL(prk::cost_scatter_kernel, dim3(sgrid), dim3(256), 0, st, P.boxes, P.box_start, P.nbox, list,
                     nlist, ncells, hist, order);
  HIPCHK(hipGetLastError());
  P.cell_list = order;
  return 0;
}

// Blocks for a grid-stride launch of the surrogate's per-cell kernels over ncells cells.
static int surr_blocks(pr_handle* h, long ncells, int* block, size_t* shmem, int* gblocks) {
  const int G = h->var->G;
  *block = pick_block(h);
  *shmem = (size_t)h->blob_bytes + (size_t)(*block / G) * h->var->scratch;
  int dev_cus = 0;
  HIPCHK(hipDeviceGetAttribute(&dev_cus, hipDeviceAttributeMultiprocessorCount, h->device));
  const long want = (ncells + (*block / G) - 1) / (*block / G);
  *gblocks = (int)(want < 8L * dev_cus ? want : 8L * dev_cus);
  if (*gblocks < 1) *gblocks = 1;
  return 0;
}

// The learned map's first half for one box: features and rate scales of its cells, written to
// rows [base, base + cells) of the surrogate buffers.
static int surr_prep(pr_handle* h, const prk::BoxDesc& d, long base, hipStream_t st) {
  Surrogate* S = h->surr;
  const int N = h->mech->nsp;
  int block = 0, gblocks = 0;
  size_t shmem = 0;
  if (int rc = surr_blocks(h, d.ncells, &block, &shmem, &gblocks)) return rc;
  prk::PrepParams F{};
  F.lo0 = d.lo0;
  F.lo1 = d.lo1;
  F.lo2 = d.lo2;
  F.nx = d.nx;
  F.ny = d.ny;
  F.ncells = d.ncells;
  F.rY = d.rY;
  F.rYsrc = d.rYsrc;
  F.T = d.T;
  F.rH = d.rH;
  F.rHsrc = d.rHsrc;
  F.mask = d.mask;
  F.has_mask = d.has_mask;
  F.clean = h->keys.clean_init_massfrac;
  F.dt = d.dt;
  F.xmu = S->xmu;
  F.xsd = S->xsd;
  F.X = S->X + base * (N + 2);
  F.Tstar = S->Tstar + base;
  F.scale = S->scale + base * (N + 1);
  F.blob = h->d_blob;
  F.blob_bytes = h->blob_bytes;
  F.neff = h->mech->neff;
  F.ninc = h->mech->ninc;
  HIPCHK(h->var->prep(F, gblocks, block, shmem, st));
  return 0;
}

// The second half, after the network has run on rows [base, base + cells): the FP64 trust check
// writes accepted cells in place and lists the others as base + box-local index.
static int surr_trust(pr_handle* h, const prk::BoxDesc& d, long base, hipStream_t st) {
  Surrogate* S = h->surr;
  const int N = h->mech->nsp;
  int block = 0, gblocks = 0;
  size_t shmem = 0;
  if (int rc = surr_blocks(h, d.ncells, &block, &shmem, &gblocks)) return rc;
  prk::TrustParams Q{};
  Q.lo0 = d.lo0;
  Q.lo1 = d.lo1;
  Q.lo2 = d.lo2;
  Q.nx = d.nx;
  Q.ny = d.ny;
  Q.ncells = d.ncells;
  Q.rY = d.rY;
  Q.rYsrc = d.rYsrc;
  Q.T = d.T;
  Q.rH = d.rH;
  Q.rHsrc = d.rHsrc;
  Q.FC = d.FC;
  Q.mask = d.mask;
  Q.has_mask = d.has_mask;
  Q.clean = h->keys.clean_init_massfrac;
  Q.status = d.status;
  Q.dt = d.dt;
  Q.Yout = S->Yout + base * (N + 1);
  Q.ymu = S->ymu;
  Q.ysd = S->ysd;
  Q.yeps = S->yeps;
  Q.Mel = S->Mel;
  Q.Minv = S->Minv;
  Q.E = S->E;
  Q.trust_rtol = h->keys.surr_trust_rtol;
  Q.trust_atol = h->keys.surr_trust_atol;
  Q.elem_tol = h->keys.surr_elem_tol;
  Q.blob = h->d_blob;
  Q.blob_bytes = h->blob_bytes;
  Q.neff = h->mech->neff;
  Q.ninc = h->mech->ninc;
  Q.rej_list = S->rej;
  Q.cell_base = (int)base;
  Q.counters = S->cnt;
  Q.audit_cap = S->aud_cap;
  const double thr = h->keys.surr_audit_frac * 4294967296.0;
  Q.audit_thresh = thr >= 4294967295.0 ? 0xFFFFFFFFu : (unsigned int)thr;
  Q.audit_seed = (unsigned int)h->keys.surr_audit_seed;
  Q.aud_in = S->aud_in;
  Q.aud_out = S->aud_out;
  Q.aud_cell = S->aud_cell;
  Q.aud_dt = S->aud_dt;
  Q.Tstar = S->Tstar + base;
  Q.scale = S->scale + base * (N + 1);
  Q.zero = h->keys.surr_zero;
  HIPCHK(h->var->trust(Q, gblocks, block, shmem, st));
  return 0;
}

// The audit batch: a 1D box over the audit buffers (acap cells, component-major), each cell
// with the dt_react of the cell it re-solves.
static prk::BoxDesc audit_desc(pr_handle* h) {
  Surrogate* S = h->surr;
  const int N = h->mech->nsp;
  const long c = S->aud_cap;
  double* a_rY = S->aud_box;
  double* a_src = a_rY + (long)N * c;
  double* a_T = a_src + (long)N * c;
  double* a_rH = a_T + c;
  double* a_rHsrc = a_rH + c;
  double* a_FC = a_rHsrc + c;
  auto arr1 = [&](double* p, long ns) {
    prk::Arr r;
    r.p = p;
    r.js = c;
    r.ks = c;
    r.ns = ns;
    r.bx = r.by = r.bz = 0;
    return r;
  };
  prk::BoxDesc a{};
  a.nx = (int)c;
  a.ny = 1;
  a.ncells = c;
  a.rY = arr1(a_rY, c);
  a.rYsrc = arr1(a_src, c);
  a.T = arr1(a_T, 0);
  a.rH = arr1(a_rH, 0);
  a.rHsrc = arr1(a_rHsrc, 0);
  a.FC = arr1(a_FC, 0);
  a.dt = 0.0;
  a.dtv = S->aud_dt;
  return a;
}

// Re-solves the audited cells with the BDF (queue and counters in the second half of
// d_counters) and compares them with the learned result.
static int audit_launch(pr_handle* h, const prk::KParams& P, const prk::BoxDesc& adesc, const prk::BoxDesc* dbox,
                        const long* dstart, hipStream_t st) {
  Surrogate* S = h->surr;
  const int N = h->mech->nsp;
  const unsigned int acap = S->aud_cap;
  auto* base = static_cast<unsigned char*>(h->d_counters);
  const unsigned grid = (unsigned)((acap + 127) / 128);
  hipLaunchKernelGGL(prk::audit_unpack_kernel, dim3(grid), dim3(128), 0, st, N, &S->cnt[1], acap, S->aud_in,
                     adesc.rY.p, adesc.rYsrc.p, adesc.T.p, adesc.rH.p, adesc.rHsrc.p, adesc.FC.p);
  HIPCHK(hipGetLastError());
  prk::KParams A = P;
  A.one = adesc;
  A.ncells = acap;
  A.boxes = dbox;
  A.box_start = dstart;
  A.nbox = 1;
  A.cell_list = nullptr;
  A.ncells_dev = &S->cnt[1];
  A.queue = reinterpret_cast<unsigned int*>(base + 128);
  A.nfail = reinterpret_cast<unsigned int*>(base + 132);
  A.stats = reinterpret_cast<unsigned long long*>(base + 192);
  if (int rc = bdf_launch(h, A, acap, st)) return rc;
  hipLaunchKernelGGL(prk::audit_err_kernel, dim3(grid), dim3(128), 0, st, N, &S->cnt[1], acap, adesc.rY.p,
                     adesc.T.p, S->aud_out, S->aud_err);
  HIPCHK(hipGetLastError());
  return 0;
}

// Accepted / rejected / audited counts and the worst audit error of the last learned-map pass.
static int surr_collect(pr_handle* h, unsigned int acap, hipStream_t st) {
  Surrogate* S = h->surr;
  unsigned int sc[3] = {0u, 0u, 0u};
  HIPCHK(hipMemcpyAsync(sc, S->cnt, sizeof(sc), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  h->last.sur_accepted = sc[2];
  h->last.sur_rejected = sc[0];
  h->last.sur_audited = sc[1] < acap ? sc[1] : acap;
  if (h->last.sur_audited > 0) {
    std::vector<double> err(2 * (size_t)h->last.sur_audited);
    unsigned int afail = 0;
    auto* base = static_cast<unsigned char*>(h->d_counters);
    HIPCHK(hipMemcpy(err.data(), S->aud_err, sizeof(double) * err.size(), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(&afail, base + 132, sizeof(afail), hipMemcpyDeviceToHost));
    double mT = 0.0, mY = 0.0;
    for (size_t a = 0; a < err.size(); a += 2) {
      mT = err[a] > mT || err[a] != err[a] ? err[a] : mT;
      mY = err[a + 1] > mY || err[a + 1] != err[a + 1] ? err[a + 1] : mY;
    }
    h->last.audit_max_dT_rel = mT;
    h->last.audit_max_dY = mY;
    h->last.audit_failed = afail;
  }
  return 0;
}

}  // extern "C"

// The no-allocation path's layout of the device scratch array (byte offsets).
constexpr size_t kFxCounters = 0;       // queue, nfail, stats: 256 B
constexpr size_t kFxTable = 256;        // box table: one descriptor and its cell prefix sums
constexpr size_t kFxTableBytes = 3840;
constexpr size_t kFxHist = 4096;        // cost-order bucket counts and cursors
constexpr size_t kFxOrder = 5120;       // cost order: one int per cell
constexpr long kFxMaxOrderCells = 1L << 22;

static size_t align256(size_t x) { return (x + 255) / 256 * 256; }

// The no-allocation react's device scratch (pr_react_box_noalloc): the counters, box table,
// cost order and BDF group slots of that entry point are carved from this array. Every device
// the library runs on holds its own copy until the process ends. Host code never passes the
// array's name to a kernel: its device address is resolved once per device by
// resolve_noalloc_scratch (from pr_create, with the handle's device current) and kept below.
constexpr size_t kNoAllocScratchBytes = (size_t)PR_NOALLOC_SCRATCH_BYTES;
static_assert(kNoAllocScratchBytes > kFxOrder + sizeof(int) * (size_t)kFxMaxOrderCells,
              "the no-allocation scratch must hold the cost order of its largest box");
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

static std::mutex g_scratch_mu;
static std::vector<unsigned char*> g_scratch_ptr;  // device address of scratch, per device (null: unresolved)

static int resolve_noalloc_scratch(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (device < 0) return fail(PR_ERR_ARG, "bad device");
  if ((int)g_scratch_ptr.size() <= device) g_scratch_ptr.resize((size_t)device + 1, nullptr);
  if (g_scratch_ptr[(size_t)device]) return 0;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return fail(PR_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_scratch_ptr[(size_t)device] = static_cast<unsigned char*>(p);
  return 0;
}

static unsigned char* noalloc_scratch(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return (device >= 0 && device < (int)g_scratch_ptr.size()) ? g_scratch_ptr[(size_t)device] : nullptr;
}

// One react call over a box (pr_react_box). fx (nullable) is the no-allocation path's device
// scratch of fx_bytes bytes: the counters, box table, cost order and group slots are carved from
// it and nothing is allocated.
static int react_box_core(pr_handle* h, const int lo[3], const int hi[3], const pr_array4* rY,
                          const pr_array4* rYsrc, const pr_array4* T, const pr_array4* rH,
                          const pr_array4* rHsrc, const pr_array4* FC, const pr_iarray4* mask,
                          double* dt_react, double* time, void* stream, int* status_out, unsigned char* fx,
                          size_t fx_bytes) {
  if (!h || !lo || !hi || !dt_react) return fail(PR_ERR_ARG, "NULL argument");
  std::lock_guard<std::recursive_mutex> lock(h->mu);
  const int N = h->mech->nsp;
  if (!arr_ok(rY, N) || !arr_ok(rYsrc, N) || !arr_ok(T, 1) || !arr_ok(rH, 1) || !arr_ok(rHsrc, 1) || !arr_ok(FC, 1))
    return fail(PR_ERR_ARG, "missing array");
  if ((rY->ncomp && rY->ncomp < N) || (rYsrc->ncomp && rYsrc->ncomp < N))
    return fail(PR_ERR_ARG, "rY / rYsrc need NUM_SPECIES components");
  const long nx = (long)hi[0] - lo[0] + 1, ny = (long)hi[1] - lo[1] + 1, nz = (long)hi[2] - lo[2] + 1;
  if (nx <= 0 || ny <= 0 || nz <= 0) return fail(PR_ERR_ARG, "empty box");
  const long ncells = nx * ny * nz;
  if (ncells > 0x7fffffffL) return fail(PR_ERR_ARG, "box too large");
  const double dt = *dt_react;
  if (!(dt > 0.0) || !std::isfinite(dt)) return fail(PR_ERR_ARG, "dt_react must be positive and finite");
  HIPCHK(hipSetDevice(h->device));
  hipStream_t st = (hipStream_t)stream;
  prk::KParams P{};
  P.one.lo0 = lo[0];
  P.one.lo1 = lo[1];
  P.one.lo2 = lo[2];
  P.one.nx = (int)nx;
  P.one.ny = (int)ny;
  P.ncells = ncells;
  P.one.ncells = ncells;
  P.one.rY = to_arr(rY);
  P.one.rYsrc = to_arr(rYsrc);
  P.one.T = to_arr(T);
  P.one.rH = to_arr(rH);
  P.one.rHsrc = to_arr(rHsrc);
  P.one.FC = to_arr(FC);
  P.one.has_mask = mask && mask->p;
  if (P.one.has_mask) {
    P.one.mask.p = mask->p;
    P.one.mask.js = mask->jstride;
    P.one.mask.ks = mask->kstride;
    P.one.mask.bx = mask->begin[0];
    P.one.mask.by = mask->begin[1];
    P.one.mask.bz = mask->begin[2];
  }
  P.one.dt = dt;
  P.rtol = h->keys.rtol;
  P.atol = h->keys.atol;
  P.atolv = h->d_atolv;
  P.max_order = h->keys.max_order;
  P.max_steps = h->keys.max_substeps;
  P.clean = h->keys.clean_init_massfrac;
  P.blob = h->d_blob;
  P.blob_bytes = h->blob_bytes;
  P.neff = h->mech->neff;
  P.ninc = h->mech->ninc;
  if (fx) {
    if (h->surr && h->keys.surr_enable)
      return fail(PR_ERR_UNSUPPORTED, "the no-allocation react does not run the learned map (surrogate.enable = 0)");
    if (h->keys.autotune && !h->tuned)
      return fail(PR_ERR_UNSUPPORTED, "hipbdf.autotune allocates its sample: tune with a first pr_react_box call");
    if (h->keys.cost_order && ncells > kFxMaxOrderCells)
      return fail(PR_ERR_ARG, "box of " + std::to_string(ncells) + " cells exceeds the no-allocation limit of " +
                                  std::to_string(kFxMaxOrderCells) + " cells with hipbdf.cost_order = 1");
  }
  auto* base = fx ? fx + kFxCounters : static_cast<unsigned char*>(h->d_counters);
  P.queue = reinterpret_cast<unsigned int*>(base);
  P.nfail = reinterpret_cast<unsigned int*>(base + 4);
  P.stats = reinterpret_cast<unsigned long long*>(base + 64);
  P.one.status = status_out;
  h->last = pr_stats{};

  {
    int rc = autotune(h, P, P.one, st);
    if (rc) return rc;
  }
  Surrogate* S = (!fx && h->surr && h->keys.surr_enable) ? h->surr : nullptr;
  HIPCHK(hipEventRecord(h->ev0, st));
  if (S) {
    int rc = ensure_surr_buffers(h, ncells);
    if (rc) return rc;
    if ((rc = surr_prep(h, P.one, 0, st))) return rc;
    if (!h->keys.surr_zero && (rc = mlp_forward(h, ncells, st))) return rc;
    HIPCHK(hipMemsetAsync(S->cnt, 0, 8 * sizeof(unsigned int), st));
    if ((rc = surr_trust(h, P.one, 0, st))) return rc;
    P.cell_list = S->rej;
    P.ncells_dev = &S->cnt[0];
  }
  HIPCHK(hipEventRecord(h->ev1, st));
  // box table: [0] this box; [1] the audit batch (a 1D box of acap cells) when there is one
  const unsigned int acap = S ? S->aud_cap : 0u;
  std::vector<prk::BoxDesc> descs(1, P.one);
  std::vector<long> starts = {0L, ncells};
  prk::BoxDesc adesc{};
  if (S && acap > 0) {
    adesc = audit_desc(h);
    descs.push_back(adesc);
    starts.push_back(0L);
    starts.push_back((long)acap);
  }
  const prk::BoxDesc* dboxes = nullptr;
  const long* dstart = nullptr;
  {
    int rc = fx ? stage_boxes(h, descs, starts, st, &dboxes, &dstart, fx + kFxTable, kFxTableBytes)
                : stage_boxes(h, descs, starts, st, &dboxes, &dstart);
    if (rc) return rc;
  }
  P.boxes = dboxes;
  P.box_start = dstart;
  P.nbox = 1;
  if (h->keys.cost_order) {
    int rc = fx ? cost_order_launch(h, P, ncells, st, reinterpret_cast<int*>(fx + kFxOrder),
                                    reinterpret_cast<unsigned int*>(fx + kFxHist))
                : cost_order_launch(h, P, ncells, st);
    if (rc) return rc;
  }
  {
    int rc;
    if (fx) {
      const size_t g0 = align256(kFxOrder + (h->keys.cost_order ? sizeof(int) * (size_t)ncells : 0));
      rc = bdf_launch(h, P, ncells, st, reinterpret_cast<double*>(fx + g0), fx_bytes - g0);
    } else {
      rc = bdf_launch(h, P, ncells, st);
    }
    if (rc) return rc;
  }
  HIPCHK(hipEventRecord(h->ev2, st));
  if (S && acap > 0) {
    // audit: the BDF re-solves the sampled accepted cells from their inputs, as the 1D box [1]
    // of the table
    int rc = audit_launch(h, P, adesc, dboxes + 1, dstart + 2, st);
    if (rc) return rc;
  }
  HIPCHK(hipEventRecord(h->ev3, st));
  unsigned long long hs[8];
  unsigned int nfail = 0;
  HIPCHK(hipMemcpyAsync(hs, P.stats, sizeof(hs), hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(&nfail, P.nfail, sizeof(nfail), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  float ms_all = 0.f, ms_sur = 0.f, ms_bdf = 0.f;
  HIPCHK(hipEventElapsedTime(&ms_all, h->ev0, h->ev3));
  HIPCHK(hipEventElapsedTime(&ms_sur, h->ev0, h->ev1));
  HIPCHK(hipEventElapsedTime(&ms_bdf, h->ev1, h->ev2));
  h->last.cells = (int64_t)hs[0];
  h->last.nst = (int64_t)hs[1];
  h->last.nfe = (int64_t)hs[2];
  h->last.nje = (int64_t)hs[3];
  h->last.nsetups = (int64_t)hs[4];
  h->last.nni = (int64_t)hs[5];
  h->last.netf = (int64_t)hs[6];
  h->last.ncfn = (int64_t)hs[7];
  h->last.failed = nfail;
  h->last.kernel_ms = ms_bdf;
  h->last.total_ms = ms_all;
  h->last.surrogate_ms = ms_sur;
  if (S) {
    int rc = surr_collect(h, acap, st);
    if (rc) return rc;
    if (h->last.sur_audited > 0) {
      float ms_aud = 0.f;
      HIPCHK(hipEventElapsedTime(&ms_aud, h->ev2, h->ev3));
      h->last.audit_ms = ms_aud;
    }
  }
  if (time) *time += dt;
  return (int)nfail;
}

extern "C" {

int pr_react_box(pr_handle* h, const int lo[3], const int hi[3], const pr_array4* rY,
                 const pr_array4* rYsrc, const pr_array4* T, const pr_array4* rH,
                 const pr_array4* rHsrc, const pr_array4* FC, const pr_iarray4* mask,
                 double* dt_react, double* time, void* stream, int* status_out) {
  return react_box_core(h, lo, hi, rY, rYsrc, T, rH, rHsrc, FC, mask, dt_react, time, stream, status_out, nullptr,
                        0);
}

int pr_react_box_noalloc(pr_handle* h, const int lo[3], const int hi[3], const pr_array4* rY,
                         const pr_array4* rYsrc, const pr_array4* T, const pr_array4* rH,
                         const pr_array4* rHsrc, const pr_array4* FC, const pr_iarray4* mask,
                         double* dt_react, double* time, void* stream, int* status_out) {
  if (!h) return fail(PR_ERR_ARG, "NULL handle");
  unsigned char* fx = noalloc_scratch(h->device);
  if (!fx) return fail(PR_ERR_HIP, "the no-allocation scratch of this device was not resolved by pr_create");
  return react_box_core(h, lo, hi, rY, rYsrc, T, rH, rHsrc, FC, mask, dt_react, time, stream, status_out, fx,
                        kNoAllocScratchBytes);
}

int pr_react_batch_enqueue(pr_handle* h, int nbox, const pr_box* boxes, void* stream) {
  if (!h || nbox < 0 || (nbox > 0 && !boxes)) return fail(PR_ERR_ARG, "bad argument");
  std::lock_guard<std::recursive_mutex> lock(h->mu);
  if (nbox == 0) {
    // A level with no box on this rank (PeleLMeX's MFIter loop then runs zero times): nothing to
    // integrate. The counters are cleared on the stream so that pr_collect reports 0 failed
    // cells and empty statistics; a kernel, not a memset, so the call can be captured.
    HIPCHK(hipSetDevice(h->device));
    auto* base = static_cast<unsigned char*>(h->d_counters);
    hipLaunchKernelGGL(prk::zero_counters, dim3(1), dim3(64), 0, (hipStream_t)stream,
                       reinterpret_cast<unsigned int*>(base), reinterpret_cast<unsigned long long*>(base + 64),
                       reinterpret_cast<unsigned int*>(base + 4));
    HIPCHK(hipGetLastError());
    h->batch_surr = 0;
    h->batch_acap = 0;
    return 0;
  }
  const int N = h->mech->nsp;
  std::vector<prk::BoxDesc> descs((size_t)nbox);
  std::vector<long> starts((size_t)nbox + 1, 0L);
  for (int b = 0; b < nbox; ++b) {
    const pr_box& x = boxes[b];
    if (!arr_ok(&x.rY, N) || !arr_ok(&x.rYsrc, N) || !arr_ok(&x.T, 1) || !arr_ok(&x.rH, 1) ||
        !arr_ok(&x.rHsrc, 1) || !arr_ok(&x.FC, 1))
      return fail(PR_ERR_ARG, "box " + std::to_string(b) + ": missing array");
    if ((x.rY.ncomp && x.rY.ncomp < N) || (x.rYsrc.ncomp && x.rYsrc.ncomp < N))
      return fail(PR_ERR_ARG, "box " + std::to_string(b) + ": rY / rYsrc need NUM_SPECIES components");
    const long nx = (long)x.hi[0] - x.lo[0] + 1, ny = (long)x.hi[1] - x.lo[1] + 1, nz = (long)x.hi[2] - x.lo[2] + 1;
    if (nx <= 0 || ny <= 0 || nz <= 0) return fail(PR_ERR_ARG, "box " + std::to_string(b) + ": empty box");
    if (!(x.dt_react > 0.0) || !std::isfinite(x.dt_react))
      return fail(PR_ERR_ARG, "box " + std::to_string(b) + ": dt_react must be positive and finite");
    prk::BoxDesc& d = descs[(size_t)b];
    d.lo0 = x.lo[0];
    d.lo1 = x.lo[1];
    d.lo2 = x.lo[2];
    d.nx = (int)nx;
    d.ny = (int)ny;
    d.ncells = nx * ny * nz;
    d.rY = to_arr(&x.rY);
    d.rYsrc = to_arr(&x.rYsrc);
    d.T = to_arr(&x.T);
    d.rH = to_arr(&x.rH);
    d.rHsrc = to_arr(&x.rHsrc);
    d.FC = to_arr(&x.FC);
    d.has_mask = x.mask.p != nullptr;
    if (d.has_mask) {
      d.mask.p = x.mask.p;
      d.mask.js = x.mask.jstride;
      d.mask.ks = x.mask.kstride;
      d.mask.bx = x.mask.begin[0];
      d.mask.by = x.mask.begin[1];
      d.mask.bz = x.mask.begin[2];
    }
    d.status = x.status;
    d.dt = x.dt_react;
    starts[(size_t)b + 1] = starts[(size_t)b] + d.ncells;
  }
  const long total = starts.back();
  if (total > 0x7fffffffL) return fail(PR_ERR_ARG, "too many cells");
  HIPCHK(hipSetDevice(h->device));
  hipStream_t st = (hipStream_t)stream;
  {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    HIPCHK(hipStreamIsCapturing(st, &cs));
    if (cs != hipStreamCaptureStatusNone) h->captured = true;
  }
  prk::KParams P{};
  P.one = descs[0];
  P.ncells = total;
  P.rtol = h->keys.rtol;
  P.atol = h->keys.atol;
  P.atolv = h->d_atolv;
  P.max_order = h->keys.max_order;
  P.max_steps = h->keys.max_substeps;
  P.clean = h->keys.clean_init_massfrac;
  P.blob = h->d_blob;
  P.blob_bytes = h->blob_bytes;
  P.neff = h->mech->neff;
  P.ninc = h->mech->ninc;
  auto* base = static_cast<unsigned char*>(h->d_counters);
  P.queue = reinterpret_cast<unsigned int*>(base);
  P.nfail = reinterpret_cast<unsigned int*>(base + 4);
  P.stats = reinterpret_cast<unsigned long long*>(base + 64);
  {
    size_t big = 0;  // tune on the largest box
    for (size_t b = 1; b < descs.size(); ++b)
      if (descs[b].ncells > descs[big].ncells) big = b;
    int rc = autotune(h, P, descs[big], st);
    if (rc) return rc;
  }
  // With a learned map loaded: every box's cells through the map (features per box, one network
  // pass over all boxes, trust check per box with its own dt), then one persistent BDF launch
  // over the rejected cells of all boxes. No memset and no allocation once buffers are sized, so
  // the sequence can be captured in a hipGraph.
  Surrogate* S = (h->surr && h->keys.surr_enable) ? h->surr : nullptr;
  h->batch_surr = 0;
  h->batch_acap = 0;
  if (S) {
    int rc = ensure_surr_buffers(h, total);
    if (rc) return rc;
    for (int b = 0; b < nbox; ++b)
      if ((rc = surr_prep(h, descs[(size_t)b], starts[(size_t)b], st))) return rc;
    if (!h->keys.surr_zero && (rc = mlp_forward(h, total, st))) return rc;
    hipLaunchKernelGGL(prk::zero_u32, dim3(1), dim3(64), 0, st, S->cnt, 8);
    HIPCHK(hipGetLastError());
    for (int b = 0; b < nbox; ++b)
      if ((rc = surr_trust(h, descs[(size_t)b], starts[(size_t)b], st))) return rc;
    P.cell_list = S->rej;
    P.ncells_dev = &S->cnt[0];
  }
  prk::BoxDesc adesc{};
  const bool audit = S && S->aud_cap > 0;
  if (audit) {  // the audit batch: box [nbox] of the table, with its own start pair
    adesc = audit_desc(h);
    descs.push_back(adesc);
    starts.push_back(0L);
    starts.push_back((long)S->aud_cap);
  }
  const prk::BoxDesc* dboxes = nullptr;
  const long* dstart = nullptr;
  {
    int rc = stage_boxes(h, descs, starts, st, &dboxes, &dstart);
    if (rc) return rc;
    P.boxes = dboxes;
    P.box_start = dstart;
    P.nbox = nbox;
  }
  if (h->keys.cost_order) {  // cells of every box by decreasing input FC
    int rc = cost_order_launch(h, P, total, st);
    if (rc) return rc;
  }
  {
    int rc = bdf_launch(h, P, total, st);
    if (rc) return rc;
  }
  if (audit) {
    int rc = audit_launch(h, P, adesc, dboxes + nbox, dstart + nbox + 1, st);
    if (rc) return rc;
  }
  h->batch_surr = S ? 1 : 0;
  h->batch_acap = S ? S->aud_cap : 0u;  // read after ensure_surr_buffers has sized the audit
  return 0;
}

int pr_collect(pr_handle* h, void* stream) {
  if (!h) return fail(PR_ERR_ARG, "NULL argument");
  std::lock_guard<std::recursive_mutex> lock(h->mu);
  HIPCHK(hipSetDevice(h->device));
  hipStream_t st = (hipStream_t)stream;
  auto* base = static_cast<unsigned char*>(h->d_counters);
  unsigned long long hs[8];
  unsigned int nfail = 0;
  HIPCHK(hipMemcpyAsync(hs, base + 64, sizeof(hs), hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(&nfail, base + 4, sizeof(nfail), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  const int G = h->last.lanes_per_cell, blocks = h->last.blocks;
  const int bt = h->last.block_threads, wv = h->last.waves_per_simd;
  h->last = pr_stats{};
  h->last.lanes_per_cell = G;
  h->last.blocks = blocks;
  h->last.block_threads = bt;
  h->last.waves_per_simd = wv;
  h->last.cells = (int64_t)hs[0];
  h->last.nst = (int64_t)hs[1];
  h->last.nfe = (int64_t)hs[2];
  h->last.nje = (int64_t)hs[3];
  h->last.nsetups = (int64_t)hs[4];
  h->last.nni = (int64_t)hs[5];
  h->last.netf = (int64_t)hs[6];
  h->last.ncfn = (int64_t)hs[7];
  h->last.failed = nfail;
  if (h->batch_surr && h->surr) {
    int rc = surr_collect(h, h->batch_acap, st);
    if (rc) return rc;
  }
  return (int)nfail;
}

int pr_react_batch(pr_handle* h, int nbox, const pr_box* boxes, void* stream) {
  if (!h) return fail(PR_ERR_ARG, "NULL argument");
  std::lock_guard<std::recursive_mutex> lock(h->mu);
  hipStream_t st = (hipStream_t)stream;
  HIPCHK(hipEventRecord(h->ev0, st));
  int rc = pr_react_batch_enqueue(h, nbox, boxes, stream);
  if (rc) return rc;
  HIPCHK(hipEventRecord(h->ev1, st));
  rc = pr_collect(h, stream);
  if (rc < 0) return rc;
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, h->ev0, h->ev1));
  h->last.kernel_ms = ms;
  h->last.total_ms = ms;
  return rc;
}

static int eval_batch(pr_handle* h, int mode, int64_t ncells, const double* y, const double* rYsrc,
                      const double* rhoh0, const double* rhohsrc, double t, double* out, void* stream) {
  if (!h || !y || !rYsrc || !rhohsrc || !out || (mode == 0 && !rhoh0) || ncells <= 0)
    return fail(PR_ERR_ARG, "bad argument");
  std::lock_guard<std::recursive_mutex> lock(h->mu);
  HIPCHK(hipSetDevice(h->device));
  const int G = h->var->G;
  const int block = pick_block(h);
  const size_t shmem = (size_t)h->blob_bytes + (size_t)(block / G) * h->var->scratch;
  int dev_cus = 0, nb = 0;
  HIPCHK(hipDeviceGetAttribute(&dev_cus, hipDeviceAttributeMultiprocessorCount, h->device));
  HIPCHK(h->var->occ(&nb, block, shmem));
  if (nb < 1) nb = 1;
  const long want = (ncells + (block / G) - 1) / (block / G);
  const long cap = (long)nb * dev_cus;
  const int blocks = (int)(want < cap ? want : cap);
  {
    int rc = ensure_gscr(h, (long)blocks * (block / G));
    if (rc) return rc;
  }
  hipStream_t st = (hipStream_t)stream;
  HIPCHK(h->var->eval(mode, h->d_blob, h->blob_bytes, h->mech->neff, h->mech->ninc, (long)ncells, y, rYsrc, rhoh0,
                      rhohsrc, t, out, h->d_gscr, blocks, block, shmem, st));
  HIPCHK(hipStreamSynchronize(st));
  return 0;
}

int pr_rhs_batch(pr_handle* h, int64_t ncells, const double* y, const double* rYsrc, const double* rhoh0,
                 const double* rhohsrc, double t, double* f, void* stream) {
  return eval_batch(h, 0, ncells, y, rYsrc, rhoh0, rhohsrc, t, f, stream);
}

int pr_jac_batch(pr_handle* h, int64_t ncells, const double* y, const double* rYsrc, const double* rhohsrc,
                 double* J, void* stream) {
  return eval_batch(h, 1, ncells, y, rYsrc, nullptr, rhohsrc, 0.0, J, stream);
}

int pr_reaction_rate_box(pr_handle* h, const int lo[3], const int hi[3], const pr_array4* rY,
                         const pr_array4* rH, const pr_array4* T, const pr_array4* rYdot,
                         const pr_iarray4* mask, void* stream) {
  if (!h || !lo || !hi) return fail(PR_ERR_ARG, "NULL argument");
  std::lock_guard<std::recursive_mutex> lock(h->mu);
  const int N = h->mech->nsp;
  if (!arr_ok(rY, N) || !arr_ok(rH, 1) || !arr_ok(T, 1) || !arr_ok(rYdot, N)) return fail(PR_ERR_ARG, "missing array");
  if ((rY->ncomp && rY->ncomp < N) || (rYdot->ncomp && rYdot->ncomp < N))
    return fail(PR_ERR_ARG, "rY / rYdot need NUM_SPECIES components");
  const long nx = (long)hi[0] - lo[0] + 1, ny = (long)hi[1] - lo[1] + 1, nz = (long)hi[2] - lo[2] + 1;
  if (nx <= 0 || ny <= 0 || nz <= 0) return fail(PR_ERR_ARG, "empty box");
  const long ncells = nx * ny * nz;
  if (ncells > 0x7fffffffL) return fail(PR_ERR_ARG, "box too large");
  HIPCHK(hipSetDevice(h->device));
  prk::RateParams P{};
  P.lo0 = lo[0];
  P.lo1 = lo[1];
  P.lo2 = lo[2];
  P.nx = (int)nx;
  P.ny = (int)ny;
  P.ncells = ncells;
  P.rY = to_arr(rY);
  P.rH = to_arr(rH);
  P.T = to_arr(T);
  P.out = to_arr(rYdot);
  P.has_mask = mask && mask->p;
  if (P.has_mask) {
    P.mask.p = mask->p;
    P.mask.js = mask->jstride;
    P.mask.ks = mask->kstride;
    P.mask.bx = mask->begin[0];
    P.mask.by = mask->begin[1];
    P.mask.bz = mask->begin[2];
  }
  P.blob = h->d_blob;
  P.blob_bytes = h->blob_bytes;
  P.neff = h->mech->neff;
  P.ninc = h->mech->ninc;
  const int G = h->var->G;
  const int block = pick_block(h);
  const size_t shmem = (size_t)h->blob_bytes + (size_t)(block / G) * h->var->scratch;
  int dev_cus = 0;
  HIPCHK(hipDeviceGetAttribute(&dev_cus, hipDeviceAttributeMultiprocessorCount, h->device));
  const long want = (ncells + (block / G) - 1) / (block / G);
  const long cap = 8L * dev_cus;
  const int blocks = (int)(want < cap ? want : cap);
  hipStream_t st = (hipStream_t)stream;
  HIPCHK(h->var->rate(P, blocks, block, shmem, st));
  HIPCHK(hipStreamSynchronize(st));
  return 0;
}

// Launches one of the EOS box kernels (one thread per cell, grid-stride) and synchronizes.
static int eos_launch(pr_handle* h, int mode, const int lo[3], const int hi[3], prk::EosParams& P, void* stream) {
  std::lock_guard<std::recursive_mutex> lock(h->mu);
  const long nx = (long)hi[0] - lo[0] + 1, ny = (long)hi[1] - lo[1] + 1, nz = (long)hi[2] - lo[2] + 1;
  if (nx <= 0 || ny <= 0 || nz <= 0) return fail(PR_ERR_ARG, "empty box");
  const long ncells = nx * ny * nz;
  if (ncells > 0x7fffffffL) return fail(PR_ERR_ARG, "box too large");
  HIPCHK(hipSetDevice(h->device));
  P.lo0 = lo[0];
  P.lo1 = lo[1];
  P.lo2 = lo[2];
  P.nx = (int)nx;
  P.ny = (int)ny;
  P.ncells = ncells;
  P.blob = h->d_blob;
  P.blob_bytes = h->blob_bytes;
  P.neff = h->mech->neff;
  P.ninc = h->mech->ninc;
  const int block = 256;
  const size_t shmem = (size_t)h->mech->nsp * sizeof(prk::SpRec);  // the species records
  int dev_cus = 0;
  HIPCHK(hipDeviceGetAttribute(&dev_cus, hipDeviceAttributeMultiprocessorCount, h->device));
  const long want = (ncells + block - 1) / block;
  const long cap = 8L * dev_cus;
  const int blocks = (int)(want < cap ? want : cap);
  hipStream_t st = (hipStream_t)stream;
  HIPCHK(h->var->eos(mode, P, blocks, block, shmem, st));
  HIPCHK(hipStreamSynchronize(st));
  return 0;
}

int pr_chem_boxes(const int domain_lo[3], const int domain_hi[3], int ngrid, const int* grid_lo, const int* grid_hi,
                  int nfine, const int* fine_lo, const int* fine_hi, int ref_ratio, const int max_grid_size_chem[3],
                  int finest, int cap, int* box_lo, int* box_hi, int* flags) {
  if (!domain_lo || !domain_hi || !max_grid_size_chem || ngrid < 0 || nfine < 0 || cap < 0)
    return fail(PR_ERR_ARG, "bad argument");
  if ((ngrid > 0 && (!grid_lo || !grid_hi)) || (nfine > 0 && (!fine_lo || !fine_hi)) ||
      (cap > 0 && (!box_lo || !box_hi || !flags)))
    return fail(PR_ERR_ARG, "NULL array");
  if (!finest && ref_ratio < 1) return fail(PR_ERR_ARG, "ref_ratio must be >= 1 below the finest level");
  auto box = [](const int* lo, const int* hi) {
    prc::IBox b;
    for (int d = 0; d < 3; ++d) {
      b.lo[d] = lo[d];
      b.hi[d] = hi[d];
    }
    return b;
  };
  const prc::IBox dom = box(domain_lo, domain_hi);
  if (!dom.ok()) return fail(PR_ERR_ARG, "empty domain");
  std::vector<prc::IBox> grids, fine, out;
  std::vector<int> fl;
  for (int g = 0; g < ngrid; ++g) {
    grids.push_back(box(grid_lo + 3 * g, grid_hi + 3 * g));
    if (!grids.back().ok()) return fail(PR_ERR_ARG, "empty grid box");
  }
  for (int g = 0; g < nfine; ++g) {
    fine.push_back(box(fine_lo + 3 * g, fine_hi + 3 * g));
    if (!fine.back().ok()) return fail(PR_ERR_ARG, "empty finer-level box");
  }
  prc::chem_box_array(dom, grids, fine, ref_ratio, max_grid_size_chem, finest != 0, out, fl);
  if (out.size() > 0x7fffffffu) return fail(PR_ERR_ARG, "too many boxes");
  for (size_t b = 0; b < out.size() && (long)b < cap; ++b) {
    for (int d = 0; d < 3; ++d) {
      box_lo[3 * b + d] = out[b].lo[d];
      box_hi[3 * b + d] = out[b].hi[d];
    }
    flags[b] = fl[b];
  }
  return (int)out.size();
}

int pr_box_costs(pr_handle* h, int nbox, const int* lo, const int* hi, const pr_array4* FC, int method,
                 double* costs, void* stream) {
  if (!h || nbox < 0 || (nbox > 0 && (!lo || !hi || !FC || !costs))) return fail(PR_ERR_ARG, "NULL argument");
  std::lock_guard<std::recursive_mutex> lock(h->mu);
  if (method < prk::COST_NCELL || method > prk::COST_FC_SUM) return fail(PR_ERR_ARG, "unknown cost method");
  if (nbox == 0) return 0;
  std::vector<prk::CostBox> bx((size_t)nbox);
  for (int b = 0; b < nbox; ++b) {
    const long nx = (long)hi[3 * b] - lo[3 * b] + 1, ny = (long)hi[3 * b + 1] - lo[3 * b + 1] + 1,
               nz = (long)hi[3 * b + 2] - lo[3 * b + 2] + 1;
    if (nx <= 0 || ny <= 0 || nz <= 0) return fail(PR_ERR_ARG, "empty box");
    if (nx * ny * nz > 0x7fffffffL) return fail(PR_ERR_ARG, "box too large");
    if (method != prk::COST_NCELL && !arr_ok(&FC[b], 1)) return fail(PR_ERR_ARG, "missing FC array");
    bx[b].lo0 = lo[3 * b];
    bx[b].lo1 = lo[3 * b + 1];
    bx[b].lo2 = lo[3 * b + 2];
    bx[b].nx = (int)nx;
    bx[b].ny = (int)ny;
    bx[b].ncells = nx * ny * nz;
    bx[b].FC = to_arr(&FC[b]);
  }
  if (method == prk::COST_NCELL) {
    for (int b = 0; b < nbox; ++b) costs[b] = (double)bx[b].ncells;
    return 0;
  }
  HIPCHK(hipSetDevice(h->device));
  hipStream_t st = (hipStream_t)stream;
  void* d = nullptr;
  const size_t bytes = sizeof(prk::CostBox) * (size_t)nbox + sizeof(double) * (size_t)nbox;
  HIPCHK(hipMalloc(&d, bytes));
  prk::CostBox* dbx = static_cast<prk::CostBox*>(d);
  double* dout = reinterpret_cast<double*>(static_cast<char*>(d) + sizeof(prk::CostBox) * (size_t)nbox);
  hipError_t e = hipMemcpyAsync(dbx, bx.data(), sizeof(prk::CostBox) * (size_t)nbox, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(prk::box_cost_kernel, dim3((unsigned)nbox), dim3(256), 0, st, dbx, method, dout);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(costs, dout, sizeof(double) * (size_t)nbox, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  (void)hipFree(d);
  if (e != hipSuccess) return fail(PR_ERR_HIP, std::string("pr_box_costs: ") + hipGetErrorString(e));
  return 0;
}

int pr_set_temperature_box(pr_handle* h, const int lo[3], const int hi[3], const pr_array4* rho,
                           const pr_array4* rY, const pr_array4* rH, const pr_array4* T, void* stream) {
  if (!h || !lo || !hi) return fail(PR_ERR_ARG, "NULL argument");
  const int N = h->mech->nsp;
  if (!arr_ok(rho, 1) || !arr_ok(rY, N) || !arr_ok(rH, 1) || !arr_ok(T, 1)) return fail(PR_ERR_ARG, "missing array");
  if (rY->ncomp && rY->ncomp < N) return fail(PR_ERR_ARG, "rY needs NUM_SPECIES components");
  prk::EosParams P{};
  P.rho = to_arr(rho);
  P.rY = to_arr(rY);
  P.rH = to_arr(rH);
  P.T = to_arr(T);
  return eos_launch(h, prk::EOS_SET_T, lo, hi, P, stream);
}

int pr_floor_species_box(pr_handle* h, const int lo[3], const int hi[3], const pr_array4* rY,
                         const pr_array4* rho, const pr_array4* rH, const pr_array4* T, void* stream) {
  if (!h || !lo || !hi) return fail(PR_ERR_ARG, "NULL argument");
  const int N = h->mech->nsp;
  if (!arr_ok(rY, N) || !arr_ok(rho, 1) || !arr_ok(rH, 1) || !arr_ok(T, 1)) return fail(PR_ERR_ARG, "missing array");
  if (rY->ncomp && rY->ncomp < N) return fail(PR_ERR_ARG, "rY needs NUM_SPECIES components");
  prk::EosParams P{};
  P.rY = to_arr(rY);
  P.rho = to_arr(rho);
  P.rH = to_arr(rH);
  P.T = to_arr(T);
  return eos_launch(h, prk::EOS_FLOOR, lo, hi, P, stream);
}

int pr_heat_release_box(pr_handle* h, const int lo[3], const int hi[3], const pr_array4* T,
                        const pr_array4* IR, const pr_array4* HR, void* stream) {
  if (!h || !lo || !hi) return fail(PR_ERR_ARG, "NULL argument");
  const int N = h->mech->nsp;
  if (!arr_ok(T, 1) || !arr_ok(IR, N) || !arr_ok(HR, 1)) return fail(PR_ERR_ARG, "missing array");
  if (IR->ncomp && IR->ncomp < N) return fail(PR_ERR_ARG, "I_R needs NUM_SPECIES components");
  prk::EosParams P{};
  P.T = to_arr(T);
  P.IR = to_arr(IR);
  P.HR = to_arr(HR);
  return eos_launch(h, prk::EOS_HEAT_RELEASE, lo, hi, P, stream);
}

int pr_reaction_source_box(pr_handle* h, const int lo[3], const int hi[3], const pr_array4* rY_old,
                           const pr_array4* rY_new, const pr_array4* extF_rY, const pr_array4* I_R, double dt,
                           void* stream) {
  if (!h || !lo || !hi) return fail(PR_ERR_ARG, "NULL argument");
  std::lock_guard<std::recursive_mutex> lock(h->mu);
  const int N = h->mech->nsp;
  if (!arr_ok(rY_old, N) || !arr_ok(rY_new, N) || !arr_ok(extF_rY, N) || !arr_ok(I_R, N))
    return fail(PR_ERR_ARG, "missing array");
  for (const pr_array4* a : {rY_old, rY_new, extF_rY, I_R})
    if (a->ncomp && a->ncomp < N) return fail(PR_ERR_ARG, "arrays need NUM_SPECIES components");
  if (!(dt > 0.0) || !std::isfinite(dt)) return fail(PR_ERR_ARG, "dt must be positive and finite");
  const long nx = (long)hi[0] - lo[0] + 1, ny = (long)hi[1] - lo[1] + 1, nz = (long)hi[2] - lo[2] + 1;
  if (nx <= 0 || ny <= 0 || nz <= 0) return fail(PR_ERR_ARG, "empty box");
  const long ncells = nx * ny * nz;
  if (ncells > 0x7fffffffL) return fail(PR_ERR_ARG, "box too large");
  HIPCHK(hipSetDevice(h->device));
  prk::IrParams P{};
  P.lo0 = lo[0];
  P.lo1 = lo[1];
  P.lo2 = lo[2];
  P.nx = (int)nx;
  P.ny = (int)ny;
  P.ncells = ncells;
  P.N = N;
  P.rYo = to_arr(rY_old);
  P.rYn = to_arr(rY_new);
  P.F = to_arr(extF_rY);
  P.IR = to_arr(I_R);
  P.dt_inv = 1.0 / dt;
  int dev_cus = 0;
  HIPCHK(hipDeviceGetAttribute(&dev_cus, hipDeviceAttributeMultiprocessorCount, h->device));
  const long want = (ncells * N + 255) / 256, cap = 16L * dev_cus;
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(prk::ir_kernel, dim3((unsigned)(want < cap ? want : cap)), dim3(256), 0, st, P);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(st));
  return 0;
}

int pr_advance_chemistry_box(pr_handle* h, const int lo[3], const int hi[3], const pr_array4* rY_old,
                             const pr_array4* rH_old, const pr_array4* T_old, const pr_array4* rY_new,
                             const pr_array4* rH_new, const pr_array4* T_new, const pr_array4* extF_rY,
                             const pr_array4* extF_rH, const pr_array4* FC, const pr_iarray4* mask,
                             const pr_array4* I_R, double dt, void* stream, int* status_out) {
  if (!h || !lo || !hi) return fail(PR_ERR_ARG, "NULL argument");
  std::lock_guard<std::recursive_mutex> lock(h->mu);
  const int N = h->mech->nsp;
  if (!arr_ok(rY_old, N) || !arr_ok(rH_old, 1) || !arr_ok(T_old, 1) || !arr_ok(rY_new, N) ||
      !arr_ok(rH_new, 1) || !arr_ok(T_new, 1) || !arr_ok(extF_rY, N) || !arr_ok(extF_rH, 1) || !arr_ok(FC, 1))
    return fail(PR_ERR_ARG, "missing array");
  for (const pr_array4* a : {rY_old, rY_new, extF_rY})
    if (a->ncomp && a->ncomp < N) return fail(PR_ERR_ARG, "species arrays need NUM_SPECIES components");
  if (I_R && I_R->p && I_R->ncomp && I_R->ncomp < N) return fail(PR_ERR_ARG, "I_R needs NUM_SPECIES components");
  if (!(dt > 0.0) || !std::isfinite(dt)) return fail(PR_ERR_ARG, "dt must be positive and finite");
  const long nx = (long)hi[0] - lo[0] + 1, ny = (long)hi[1] - lo[1] + 1, nz = (long)hi[2] - lo[2] + 1;
  if (nx <= 0 || ny <= 0 || nz <= 0) return fail(PR_ERR_ARG, "empty box");
  const long ncells = nx * ny * nz;
  if (ncells > 0x7fffffffL) return fail(PR_ERR_ARG, "box too large");
  HIPCHK(hipSetDevice(h->device));
  prk::ChemBoxParams P{};
  P.lo0 = lo[0];
  P.lo1 = lo[1];
  P.lo2 = lo[2];
  P.nx = (int)nx;
  P.ny = (int)ny;
  P.ncells = ncells;
  P.N = N;
  P.rYo = to_arr(rY_old);
  P.rHo = to_arr(rH_old);
  P.To = to_arr(T_old);
  P.rYn = to_arr(rY_new);
  P.rHn = to_arr(rH_new);
  P.Tn = to_arr(T_new);
  P.FY = to_arr(extF_rY);
  P.FH = to_arr(extF_rH);
  if (I_R && I_R->p) P.IR = to_arr(I_R);
  P.dt_inv = 1.0 / dt;
  int dev_cus = 0;
  HIPCHK(hipDeviceGetAttribute(&dev_cus, hipDeviceAttributeMultiprocessorCount, h->device));
  const long want = (ncells + 255) / 256, cap = 16L * dev_cus;
  const unsigned grid = (unsigned)(want < cap ? want : cap);
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(prk::chem_units_kernel<false>, dim3(grid), dim3(256), 0, st, P);
  HIPCHK(hipGetLastError());
  double dt_react = dt, time = 0.0;
  const int nfail = pr_react_box(h, lo, hi, rY_new, extF_rY, T_new, rH_new, extF_rH, FC, mask, &dt_react, &time,
                                 stream, status_out);
  // the scaling back runs after a failed react too, so the caller's forcing is restored
  const std::string react_err = nfail < 0 ? g_err : std::string();
  hipLaunchKernelGGL(prk::chem_units_kernel<true>, dim3(grid), dim3(256), 0, st, P);
  const hipError_t e1 = hipGetLastError();
  const hipError_t e2 = hipStreamSynchronize(st);
  if (nfail < 0) {
    g_err = react_err;
    return nfail;
  }
  if (e1 != hipSuccess) return fail(PR_ERR_HIP, std::string("chem_units_kernel: ") + hipGetErrorString(e1));
  if (e2 != hipSuccess) return fail(PR_ERR_HIP, std::string("hipStreamSynchronize: ") + hipGetErrorString(e2));
  return nfail;
}

int pr_advance_chemistry_chem_box(pr_handle* h, const int lo[3], const int hi[3], const pr_array4* rY,
                                  const pr_array4* rH, const pr_array4* T, const pr_array4* extF_rY,
                                  const pr_array4* extF_rH, const pr_array4* FC, const pr_iarray4* mask,
                                  int do_reaction, double dt, void* stream, int* status_out) {
  if (!h || !lo || !hi) return fail(PR_ERR_ARG, "NULL argument");
  std::lock_guard<std::recursive_mutex> lock(h->mu);
  const int N = h->mech->nsp;
  if (!arr_ok(rY, N) || !arr_ok(rH, 1) || !arr_ok(T, 1) || !arr_ok(extF_rY, N) || !arr_ok(extF_rH, 1) ||
      !arr_ok(FC, 1))
    return fail(PR_ERR_ARG, "missing array");
  for (const pr_array4* a : {rY, extF_rY})
    if (a->ncomp && a->ncomp < N) return fail(PR_ERR_ARG, "species arrays need NUM_SPECIES components");
  if (!(dt > 0.0) || !std::isfinite(dt)) return fail(PR_ERR_ARG, "dt must be positive and finite");
  const long nx = (long)hi[0] - lo[0] + 1, ny = (long)hi[1] - lo[1] + 1, nz = (long)hi[2] - lo[2] + 1;
  if (nx <= 0 || ny <= 0 || nz <= 0) return fail(PR_ERR_ARG, "empty box");
  const long ncells = nx * ny * nz;
  if (ncells > 0x7fffffffL) return fail(PR_ERR_ARG, "box too large");
  HIPCHK(hipSetDevice(h->device));
  prk::ChemBoxParams P{};
  P.lo0 = lo[0];
  P.lo1 = lo[1];
  P.lo2 = lo[2];
  P.nx = (int)nx;
  P.ny = (int)ny;
  P.ncells = ncells;
  P.N = N;
  P.rYn = to_arr(rY);
  P.rHn = to_arr(rH);
  P.Tn = to_arr(T);
  P.FY = to_arr(extF_rY);
  P.FH = to_arr(extF_rH);
  const prk::Arr fc = to_arr(FC);
  int dev_cus = 0;
  HIPCHK(hipDeviceGetAttribute(&dev_cus, hipDeviceAttributeMultiprocessorCount, h->device));
  const long want = (ncells + 255) / 256, cap = 16L * dev_cus;
  const unsigned grid = (unsigned)(want < cap ? want : cap);
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(prk::chem_units_inplace_kernel<0>, dim3(grid), dim3(256), 0, st, P, fc);
  HIPCHK(hipGetLastError());
  int nfail = 0;
  std::string react_err;
  if (do_reaction) {
    double dt_react = dt, time = 0.0;
    nfail = pr_react_box(h, lo, hi, rY, extF_rY, T, rH, extF_rH, FC, mask, &dt_react, &time, stream, status_out);
    if (nfail < 0) react_err = g_err;   // the state is still scaled back to MKS below
    hipLaunchKernelGGL(prk::chem_units_inplace_kernel<1>, dim3(grid), dim3(256), 0, st, P, fc);
  } else {
    h->last = pr_stats{};
    hipLaunchKernelGGL(prk::chem_units_inplace_kernel<2>, dim3(grid), dim3(256), 0, st, P, fc);
  }
  const hipError_t e1 = hipGetLastError();
  const hipError_t e2 = hipStreamSynchronize(st);
  if (nfail < 0) {
    g_err = react_err;
    return nfail;
  }
  if (e1 != hipSuccess) return fail(PR_ERR_HIP, std::string("chem_units_inplace_kernel: ") + hipGetErrorString(e1));
  if (e2 != hipSuccess) return fail(PR_ERR_HIP, std::string("hipStreamSynchronize: ") + hipGetErrorString(e2));
  return nfail;
}

int pr_set_eos_parm(pr_handle* h, const void* eos_parm, const char* eos_model) {
  if (!h) return fail(PR_ERR_ARG, "NULL argument");
  std::lock_guard<std::recursive_mutex> lock(h->mu);
  if (eos_model && std::strcmp(eos_model, "Fuego") != 0 && std::strcmp(eos_model, "GammaLaw") != 0)
    return fail(PR_ERR_UNSUPPORTED, std::string("EOS model ") + eos_model +
                                        ": the reactor implements the ideal-gas mixture (Fuego) EOS only");
  h->eos_parm = eos_parm;
  return 0;
}

int pr_surrogate_load(pr_handle* h, int nlayers, const int* dims, const float* params, const float* xnorm,
                      const float* ynorm) {
  if (!h || !dims || !params || !xnorm || !ynorm || nlayers < 1 || nlayers > 64) return fail(PR_ERR_ARG, "bad argument");
  std::lock_guard<std::recursive_mutex> lock(h->mu);
  const int N = h->mech->nsp;
  if (dims[0] != N + 2 || dims[nlayers] != N + 1)
    return fail(PR_ERR_ARG, "surrogate needs NUM_SPECIES + 2 inputs and NUM_SPECIES + 1 outputs");
  for (int l = 0; l <= nlayers; ++l)
    if (dims[l] < 1 || dims[l] > 8192) return fail(PR_ERR_ARG, "layer width out of range");
  for (int l = 0; l < 2 * dims[0]; ++l)
    if (!std::isfinite(xnorm[l]) || (l >= dims[0] && !(xnorm[l] > 0))) return fail(PR_ERR_ARG, "bad xnorm");
  for (int l = 0; l < 3 * dims[nlayers]; ++l)
    if (!std::isfinite(ynorm[l]) || (l >= dims[nlayers] && l < 2 * dims[nlayers] && !(ynorm[l] > 0)) ||
        (l >= 2 * dims[nlayers] && ynorm[l] < 0))
      return fail(PR_ERR_ARG, "bad ynorm");
  // element matrix M_ek = a_ek / W_k and (M M^T)^-1 by Gauss-Jordan
  const int E = h->mech->nelem;
  if (E < 1 || E > prk::SURR_MAXE) return fail(PR_ERR_UNSUPPORTED, "too many elements for the trust check");
  std::vector<double> Mel((size_t)E * N), MM((size_t)E * E, 0.0), Minv((size_t)E * E, 0.0);
  for (int e = 0; e < E; ++e)
    for (int k = 0; k < N; ++k) Mel[(size_t)e * N + k] = h->mech->elem[k * E + e] / h->mech->W[k];
  for (int e = 0; e < E; ++e)
    for (int f = 0; f < E; ++f) {
      double a = 0.0;
      for (int k = 0; k < N; ++k) a += Mel[(size_t)e * N + k] * Mel[(size_t)f * N + k];
      MM[(size_t)e * E + f] = a;
    }
  for (int e = 0; e < E; ++e) Minv[(size_t)e * E + e] = 1.0;
  for (int col = 0; col < E; ++col) {
    int pr = col;
    for (int r = col + 1; r < E; ++r)
      if (std::fabs(MM[(size_t)r * E + col]) > std::fabs(MM[(size_t)pr * E + col])) pr = r;
    if (!(std::fabs(MM[(size_t)pr * E + col]) > 0)) return fail(PR_ERR_UNSUPPORTED, "singular element matrix");
    for (int j = 0; j < E; ++j) {
      std::swap(MM[(size_t)col * E + j], MM[(size_t)pr * E + j]);
      std::swap(Minv[(size_t)col * E + j], Minv[(size_t)pr * E + j]);
    }
    const double d = MM[(size_t)col * E + col];
    for (int j = 0; j < E; ++j) {
      MM[(size_t)col * E + j] /= d;
      Minv[(size_t)col * E + j] /= d;
    }
    for (int r = 0; r < E; ++r)
      if (r != col) {
        const double f = MM[(size_t)r * E + col];
        for (int j = 0; j < E; ++j) {
          MM[(size_t)r * E + j] -= f * MM[(size_t)col * E + j];
          Minv[(size_t)r * E + j] -= f * Minv[(size_t)col * E + j];
        }
      }
  }
  HIPCHK(hipSetDevice(h->device));
  auto* S = new Surrogate();
  auto cleanup = [&](int code, const char* msg) {
    for (void* q : S->owned) (void)hipFree(q);
    delete S;
    return fail(code, msg);
  };
  S->dims.assign(dims, dims + nlayers + 1);
  S->E = E;
  size_t off = 0;
  for (int l = 0; l < nlayers; ++l) {
    const size_t nw = (size_t)dims[l] * dims[l + 1], nb = (size_t)dims[l + 1];
    float* tmp = nullptr;
    __bf16* w = nullptr;
    float* b = nullptr;
    if (surr_alloc(S, (void**)&w, nw * sizeof(__bf16)) || surr_alloc(S, (void**)&b, nb * sizeof(float)) ||
        hipMalloc(&tmp, nw * sizeof(float)) != hipSuccess)
      return cleanup(PR_ERR_HIP, "device allocation failed");
    bool ok = hipMemcpy(tmp, params + off, nw * sizeof(float), hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(b, params + off + nw, nb * sizeof(float), hipMemcpyHostToDevice) == hipSuccess;
    if (ok) {
      hipLaunchKernelGGL(prk::f32_to_bf16_kernel, dim3((unsigned)((nw + 255) / 256)), dim3(256), 0, 0, (long)nw, tmp, w);
      ok = hipGetLastError() == hipSuccess && hipDeviceSynchronize() == hipSuccess;
    }
    // the transposed copy for the hidden-layer forward kernel (k fastest in both operands)
    __bf16* wt = nullptr;
    if (ok) {
      std::vector<float> tr(nw);
      for (int i = 0; i < dims[l]; ++i)
        for (int o = 0; o < dims[l + 1]; ++o) tr[(size_t)o * dims[l] + i] = params[off + (size_t)i * dims[l + 1] + o];
      ok = surr_alloc(S, (void**)&wt, nw * sizeof(__bf16)) == 0 &&
           hipMemcpy(tmp, tr.data(), nw * sizeof(float), hipMemcpyHostToDevice) == hipSuccess;
      if (ok) {
        hipLaunchKernelGGL(prk::f32_to_bf16_kernel, dim3((unsigned)((nw + 255) / 256)), dim3(256), 0, 0, (long)nw, tmp, wt);
        ok = hipGetLastError() == hipSuccess && hipDeviceSynchronize() == hipSuccess;
      }
    }
    (void)hipFree(tmp);
    if (!ok) return cleanup(PR_ERR_HIP, "weight upload failed");
    S->W.push_back(w);
    S->Wt.push_back(wt);
    S->b.push_back(b);
    off += nw + nb;
  }
  const int nin = dims[0], nout = dims[nlayers];
  if (surr_alloc(S, (void**)&S->xmu, 2 * nin * sizeof(float)) || surr_alloc(S, (void**)&S->ymu, 3 * nout * sizeof(float)) ||
      surr_alloc(S, (void**)&S->Mel, Mel.size() * sizeof(double)) ||
      surr_alloc(S, (void**)&S->Minv, Minv.size() * sizeof(double)) ||
      surr_alloc(S, (void**)&S->cnt, 8 * sizeof(unsigned int)))
    return cleanup(PR_ERR_HIP, "device allocation failed");
  S->xsd = S->xmu + nin;
  S->ysd = S->ymu + nout;
  S->yeps = S->ymu + 2 * nout;
  if (hipMemcpy(S->xmu, xnorm, 2 * nin * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(S->ymu, ynorm, 3 * nout * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(S->Mel, Mel.data(), Mel.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(S->Minv, Minv.data(), Minv.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
    return cleanup(PR_ERR_HIP, "upload failed");
  if (h->surr) {
    for (void* q : h->surr->owned) release(h, q);
    delete h->surr;
  }
  h->surr = S;
  return 0;
}

int pr_surrogate_forward(pr_handle* h, int64_t ncells, const void* X, float* Y, void* stream) {
  if (!h || !X || !Y || ncells < 0) return fail(PR_ERR_ARG, "bad argument");
  std::lock_guard<std::recursive_mutex> lock(h->mu);
  if (!h->surr) return fail(PR_ERR_ARG, "no network loaded");
  if (ncells == 0) return 0;
  if (ncells > 0x7fffffffL) return fail(PR_ERR_ARG, "too many rows");
  HIPCHK(hipSetDevice(h->device));
  int rc = ensure_surr_buffers(h, (long)ncells);
  if (rc) return rc;
  Surrogate* S = h->surr;
  hipStream_t st = (hipStream_t)stream;
  HIPCHK(hipMemcpyAsync(S->X, X, sizeof(__bf16) * (size_t)ncells * S->dims[0], hipMemcpyDeviceToDevice, st));
  if ((rc = mlp_forward(h, (long)ncells, st))) return rc;
  const int nout = S->dims.back();
  HIPCHK(hipMemcpyAsync(Y, S->Yout, sizeof(float) * (size_t)ncells * nout, hipMemcpyDeviceToDevice, st));
  HIPCHK(hipStreamSynchronize(st));
  return 0;
}

int pr_surrogate_features(pr_handle* h, int64_t ncells, const double* rY, const double* T, const double* dt,
                          float* raw, void* X, void* stream) {
  if (!h || !rY || !T || !dt || ncells <= 0 || (!raw && !X)) return fail(PR_ERR_ARG, "bad argument");
  std::lock_guard<std::recursive_mutex> lock(h->mu);
  if (X && !h->surr) return fail(PR_ERR_ARG, "normalised features need a loaded surrogate");
  HIPCHK(hipSetDevice(h->device));
  prk::FeatParams F{};
  F.ncells = ncells;
  F.batch = 1;
  F.rY.p = const_cast<double*>(rY);
  F.T.p = const_cast<double*>(T);
  F.dtv = dt;
  F.xmu = X ? h->surr->xmu : nullptr;
  F.xsd = X ? h->surr->xsd : nullptr;
  F.raw = raw;
  F.X = static_cast<__bf16*>(X);
  hipStream_t st = (hipStream_t)stream;
  HIPCHK(h->var->feat(F, st));
  return 0;
}

int pr_mlp_gemm(int M, int N, int K, const void* A, int64_t sam, int64_t sak, const void* B, int64_t sbk,
                int64_t sbn, const float* bias, int gelu, float* C, int64_t ldc, void* Cb, int64_t ldcb, float* Z,
                int64_t ldz, int accumulate, void* stream) {
  if (M < 0 || N < 0 || K < 1 || !A || !B || (accumulate && !C) || (!C && !Cb && !Z))
    return fail(PR_ERR_ARG, "bad gemm arguments");
  if ((sak != 1 && sam != 1) || (sbk != 1 && sbn != 1)) return fail(PR_ERR_ARG, "each operand needs a unit stride");
  prk::GemmParams G{};
  G.M = M;
  G.N = N;
  G.K = K;
  G.A = static_cast<const __bf16*>(A);
  G.sam = sam;
  G.sak = sak;
  G.B = static_cast<const __bf16*>(B);
  G.sbk = sbk;
  G.sbn = sbn;
  G.bias = bias;
  G.epi = gelu ? prk::EPI_GELU : prk::EPI_NONE;
  G.C = C;
  G.ldc = ldc;
  G.Cb = static_cast<__bf16*>(Cb);
  G.ldcb = ldcb;
  G.Z = Z;
  G.ldz = ldz;
  G.accumulate = accumulate;
  HIPCHK(gemm_launch(G, (hipStream_t)stream));
  return 0;
}

int pr_mlp_gelu_backward(int64_t rows, int cols, const float* dA, int64_t lda, const float* Z, int64_t ldz,
                         void* dZb, int64_t ldb, float* dZ, int64_t ldd, void* stream) {
  if (rows < 0 || cols < 1 || !dA || !Z || !dZb) return fail(PR_ERR_ARG, "bad argument");
  const long n = (long)rows * cols;
  if (n == 0) return 0;
  hipLaunchKernelGGL(prk::gelu_bwd_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                     (long)rows, cols, dA, (long)lda, Z, (long)ldz, static_cast<__bf16*>(dZb), (long)ldb, dZ, (long)ldd);
  HIPCHK(hipGetLastError());
  return 0;
}

int pr_mlp_column_sum(int64_t rows, int cols, const float* d, int64_t ld, float* out, int accumulate, void* stream) {
  if (rows < 0 || cols < 1 || !d || !out) return fail(PR_ERR_ARG, "bad argument");
  hipLaunchKernelGGL(prk::col_sum_kernel, dim3((unsigned)((cols + 63) / 64)), dim3(256), 0, (hipStream_t)stream,
                     (long)rows, cols, d, (long)ld, out, accumulate);
  HIPCHK(hipGetLastError());
  return 0;
}

int pr_mlp_mse(int64_t rows, int cols, const float* Y, int64_t ldy, const float* T, int64_t ldt, float* dY,
               float* loss, void* stream) {
  if (rows < 1 || cols < 1 || !Y || !T || !dY || !loss) return fail(PR_ERR_ARG, "bad argument");
  const long n = (long)rows * cols;
  hipLaunchKernelGGL(prk::mse_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, (long)rows,
                     cols, Y, (long)ldy, T, (long)ldt, dY, loss);
  HIPCHK(hipGetLastError());
  return 0;
}

int pr_mlp_adam(int64_t n, float* p, const float* g, float* m, float* v, void* pb, float lr, float beta1, float beta2,
                float eps, int step, void* stream) {
  if (n < 0 || !p || !g || !m || !v || step < 1) return fail(PR_ERR_ARG, "bad argument");
  if (n == 0) return 0;
  const float bc1 = 1.0f - std::pow(beta1, (float)step), bc2 = 1.0f - std::pow(beta2, (float)step);
  hipLaunchKernelGGL(prk::adam_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, (long)n,
                     p, g, m, v, static_cast<__bf16*>(pb), lr, beta1, beta2, eps, bc1, bc2);
  HIPCHK(hipGetLastError());
  return 0;
}

int pr_mlp_to_bf16(int64_t n, const float* a, void* b, void* stream) {
  if (n < 0 || !a || !b) return fail(PR_ERR_ARG, "bad argument");
  if (n == 0) return 0;
  hipLaunchKernelGGL(prk::f32_to_bf16_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                     (long)n, a, static_cast<__bf16*>(b));
  HIPCHK(hipGetLastError());
  return 0;
}

int pr_get_stats(const pr_handle* h, pr_stats* out) {
  if (!h || !out) return fail(PR_ERR_ARG, "NULL argument");
  std::lock_guard<std::recursive_mutex> lock(h->mu);
  *out = h->last;
  return 0;
}

void pr_destroy(pr_handle* h) {
  if (!h) return;
  if (h->d_blob) (void)hipFree(h->d_blob);
  if (h->d_counters) (void)hipFree(h->d_counters);
  if (h->d_atolv) (void)hipFree(h->d_atolv);
  if (h->d_gscr) (void)hipFree(h->d_gscr);
  if (h->ev0) (void)hipEventDestroy(h->ev0);
  if (h->ev1) (void)hipEventDestroy(h->ev1);
  if (h->ev2) (void)hipEventDestroy(h->ev2);
  if (h->ev3) (void)hipEventDestroy(h->ev3);
  if (h->d_boxtab) (void)hipFree(h->d_boxtab);
  if (h->d_order) (void)hipFree(h->d_order);
  for (void* q : h->retired) (void)hipFree(q);
  if (h->d_hist) (void)hipFree(h->d_hist);
  if (h->surr) {
    for (void* q : h->surr->owned) (void)hipFree(q);
    delete h->surr;
  }
  delete h;
}

}  // extern "C"
