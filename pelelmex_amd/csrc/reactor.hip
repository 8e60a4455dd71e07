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
#include <sstream>
#include <string>
#include <vector>

#include "../../include/pele_reactor_capi.h"
#include "react_kernel.h"
#include "mech/drm19.h"
#include "mech/lidryer.h"

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

const pr_mech_t* const kMechs[] = {&pr_lidryer, &pr_drm19};

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

template <int N, int R, int G>
hipError_t launch_t(const prk::KParams& P, int blocks, int block, size_t shmem, hipStream_t st) {
  hipLaunchKernelGGL((prk::react_kernel<N, R, G>), dim3(blocks), dim3(block), shmem, st, P);
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
hipError_t occ_t(int* nb, int block, size_t shmem) {
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(
      nb, reinterpret_cast<const void*>(&prk::react_kernel<N, R, G>), block, shmem);
}

template <int N, int R, int G>
constexpr int scratch_bytes() {
  return prk::group_doubles<N, R, G>() * 8;
}

struct Variant {
  const char* mech;
  int G;
  Launch launch;
  EvalLaunch eval;
  Occ occ;
  int scratch;  // LDS bytes per cell group
  int slot;     // global scratch bytes per cell group
  RateLaunch rate;
};

#define PR_VARIANT(NAME, NS, NR, G) \
  {NAME, G, &launch_t<NS, NR, G>, &eval_t<NS, NR, G>, &occ_t<NS, NR, G>, scratch_bytes<NS, NR, G>(), \
   prk::Cell<NS, NR, G>::gslot_doubles * 8, &rate_t<NS, NR, G>}

const Variant kVariants[] = {
    PR_VARIANT("LiDryer", PR_LIDRYER_NSP, PR_LIDRYER_NR, 16),
    PR_VARIANT("LiDryer", PR_LIDRYER_NSP, PR_LIDRYER_NR, 8),
    PR_VARIANT("drm19", PR_DRM19_NSP, PR_DRM19_NR, 32),
    PR_VARIANT("drm19", PR_DRM19_NSP, PR_DRM19_NR, 16),
    PR_VARIANT("drm19", PR_DRM19_NSP, PR_DRM19_NR, 64),
};

const Variant* find_variant(const char* mech, int G) {
  for (const Variant& v : kVariants)
    if (std::strcmp(v.mech, mech) == 0 && (G == 0 || v.G == G)) return &v;
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
  return b;
}

struct Keys {
  double rtol = 1e-10, atol = 1e-10;
  int max_order = 2, max_substeps = 10000;
  std::string solve_type = "GMRES";
  int clean_init_massfrac = 0, analytical_jacobian = 1, verbose = 0;
  int lanes_per_cell = 0, block = 0;
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
    } else if (k == "hipbdf.block") {
      if (!num(d) || ((int)d % 64) != 0 || d < 64 || d > 256) return fail(PR_ERR_KEY, "hipbdf.block must be 64, 128, 192 or 256");
      K.block = (int)d;
    } else if (k.rfind("ode.", 0) == 0 || k.rfind("cvode.", 0) == 0 || k.rfind("hipbdf.", 0) == 0) {
      return fail(PR_ERR_KEY, "unknown reactor key " + k);
    }
    // keys of other namespaces are not the reactor's and are left alone
  }
  return 0;
}

}  // namespace

struct pr_handle {
  const pr_mech_t* mech = nullptr;
  const Variant* var = nullptr;
  Keys keys;
  int device = 0;
  int block = 128;
  unsigned char* d_blob = nullptr;
  int blob_bytes = 0;
  double* d_atolv = nullptr;
  void* d_counters = nullptr;  // queue (4 B), nfail (4 B), stats (8 x 8 B)
  double* d_gscr = nullptr;    // per-group global scratch slots
  size_t gscr_bytes = 0;
  bool use_typ = false;
  pr_stats last{};
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

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
  if (K.clean_init_massfrac != 0)
    return fail(PR_ERR_UNSUPPORTED, "ode.clean_init_massfrac = 1 is not implemented");
  int G = K.lanes_per_cell;
  if (G == 0) G = (m->nsp + 1 <= 16) ? 16 : (m->nsp + 1 <= 32 ? 32 : 64);
  const Variant* v = find_variant(m->name, G);
  if (!v) return fail(PR_ERR_UNSUPPORTED, "no kernel variant for this mechanism and hipbdf.lanes_per_cell");
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
  std::vector<unsigned char> blob = pack_blob(m);
  h->blob_bytes = (int)blob.size();
  if (hipMalloc(&h->d_blob, blob.size()) != hipSuccess || hipMalloc(&h->d_counters, 128) != hipSuccess ||
      hipMalloc(&h->d_atolv, sizeof(double) * (m->nsp + 1)) != hipSuccess ||
      hipMemcpy(h->d_blob, blob.data(), blob.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipEventCreate(&h->ev0) != hipSuccess || hipEventCreate(&h->ev1) != hipSuccess) {
    pr_destroy(h);
    return fail(PR_ERR_HIP, "device allocation failed");
  }
  *out = h;
  return 0;
}

int pr_set_typ_vals(pr_handle* h, const double* typ, int n) {
  if (!h || !typ) return fail(PR_ERR_ARG, "NULL argument");
  if (n != h->mech->nsp + 1) return fail(PR_ERR_ARG, "typical values need NUM_SPECIES + 1 entries");
  std::vector<double> at(n);
  for (int i = 0; i < n; ++i) {
    if (!(typ[i] > 0)) return fail(PR_ERR_ARG, "typical values must be positive");
    at[i] = h->keys.atol * typ[i];
  }
  HIPCHK(hipMemcpy(h->d_atolv, at.data(), sizeof(double) * n, hipMemcpyHostToDevice));
  h->use_typ = true;
  return 0;
}

int pr_num_species(const pr_handle* h) { return h ? h->mech->nsp : PR_ERR_ARG; }

// Largest block (64..256 threads) whose LDS (tables + per-group scratch) stays within 64 KiB.
static int pick_block(const pr_handle* h) {
  if (h->block) return h->block;
  const int G = h->var->G;
  for (int b = 256; b >= 64; b -= 64)
    if ((size_t)h->blob_bytes + (size_t)(b / G) * h->var->scratch <= 65536) return b;
  return 64;
}

static int ensure_gscr(pr_handle* h, long slots) {
  const size_t need = (size_t)slots * (size_t)h->var->slot;
  if (need <= h->gscr_bytes) return 0;
  if (h->d_gscr) (void)hipFree(h->d_gscr);
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

int pr_react_box(pr_handle* h, const int lo[3], const int hi[3], const pr_array4* rY,
                 const pr_array4* rYsrc, const pr_array4* T, const pr_array4* rH,
                 const pr_array4* rHsrc, const pr_array4* FC, const pr_iarray4* mask,
                 double* dt_react, double* time, void* stream, int* status_out) {
  if (!h || !lo || !hi || !dt_react) return fail(PR_ERR_ARG, "NULL argument");
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
  P.lo0 = lo[0];
  P.lo1 = lo[1];
  P.lo2 = lo[2];
  P.nx = (int)nx;
  P.ny = (int)ny;
  P.ncells = ncells;
  P.rY = to_arr(rY);
  P.rYsrc = to_arr(rYsrc);
  P.T = to_arr(T);
  P.rH = to_arr(rH);
  P.rHsrc = to_arr(rHsrc);
  P.FC = to_arr(FC);
  P.has_mask = mask && mask->p;
  if (P.has_mask) {
    P.mask.p = mask->p;
    P.mask.js = mask->jstride;
    P.mask.ks = mask->kstride;
    P.mask.bx = mask->begin[0];
    P.mask.by = mask->begin[1];
    P.mask.bz = mask->begin[2];
  }
  P.dt = dt;
  P.rtol = h->keys.rtol;
  P.atol = h->keys.atol;
  P.atolv = h->use_typ ? h->d_atolv : nullptr;
  P.max_order = h->keys.max_order;
  P.max_steps = h->keys.max_substeps;
  P.blob = h->d_blob;
  P.blob_bytes = h->blob_bytes;
  P.neff = h->mech->neff;
  P.ninc = h->mech->ninc;
  auto* base = static_cast<unsigned char*>(h->d_counters);
  P.queue = reinterpret_cast<unsigned int*>(base);
  P.nfail = reinterpret_cast<unsigned int*>(base + 4);
  P.stats = reinterpret_cast<unsigned long long*>(base + 64);
  P.status = status_out;

  const int G = h->var->G;
  const int block = pick_block(h);
  const size_t shmem = (size_t)h->blob_bytes + (size_t)(block / G) * h->var->scratch;
  int dev_cus = 0, nb = 0;
  HIPCHK(hipDeviceGetAttribute(&dev_cus, hipDeviceAttributeMultiprocessorCount, h->device));
  HIPCHK(h->var->occ(&nb, block, shmem));
  if (nb < 1) return fail(PR_ERR_HIP, "kernel does not fit on a CU with this block / LDS size");
  const long groups_needed = ncells;
  const long cap = (long)nb * dev_cus;
  const long want = (groups_needed + (block / G) - 1) / (block / G);
  const int blocks = (int)(want < cap ? want : cap);
  {
    int rc = ensure_gscr(h, (long)blocks * (block / G));
    if (rc) return rc;
  }
  P.gscr = h->d_gscr;

  hipLaunchKernelGGL(prk::zero_counters, dim3(1), dim3(64), 0, st, P.queue, P.stats, P.nfail);
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(h->ev0, st));
  HIPCHK(h->var->launch(P, blocks, block, shmem, st));
  HIPCHK(hipEventRecord(h->ev1, st));
  unsigned long long hs[8];
  unsigned int nfail = 0;
  HIPCHK(hipMemcpyAsync(hs, P.stats, sizeof(hs), hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(&nfail, P.nfail, sizeof(nfail), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, h->ev0, h->ev1));
  h->last.cells = (int64_t)hs[0];
  h->last.nst = (int64_t)hs[1];
  h->last.nfe = (int64_t)hs[2];
  h->last.nje = (int64_t)hs[3];
  h->last.nsetups = (int64_t)hs[4];
  h->last.nni = (int64_t)hs[5];
  h->last.netf = (int64_t)hs[6];
  h->last.ncfn = (int64_t)hs[7];
  h->last.failed = nfail;
  h->last.kernel_ms = ms;
  h->last.lanes_per_cell = G;
  h->last.blocks = blocks;
  if (time) *time += dt;
  return (int)nfail;
}

static int eval_batch(pr_handle* h, int mode, int64_t ncells, const double* y, const double* rYsrc,
                      const double* rhoh0, const double* rhohsrc, double t, double* out, void* stream) {
  if (!h || !y || !rYsrc || !rhohsrc || !out || (mode == 0 && !rhoh0) || ncells <= 0)
    return fail(PR_ERR_ARG, "bad argument");
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

int pr_get_stats(const pr_handle* h, pr_stats* out) {
  if (!h || !out) return fail(PR_ERR_ARG, "NULL argument");
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
  delete h;
}

}  // extern "C"
