// smc_api.cpp -- host side of libsmc: parameter semantics, derived
// quantities, exact footprint tables, device buffers and the C-ABI entry
// points declared in include/smc.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <stdexcept>
#include <string>
#include <vector>

#include "smc_internal.h"

using namespace smc;

namespace {

struct SmcError : std::runtime_error {
  int code;
  SmcError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIPCHK(x)                                                                        \
  do {                                                                                   \
    hipError_t err_ = (x);                                                               \
    if (err_ != hipSuccess)                                                              \
      throw SmcError(SMC_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(err_));     \
  } while (0)

template <class T>
T* dalloc(size_t n) {
  void* p = nullptr;
  if (n == 0) n = 1;
  HIPCHK(hipMalloc(&p, n * sizeof(T)));
  return (T*)p;
}

}  // namespace

// Regge96.cpp:27-50 (pp total, baryon-baryon elastic); MCnucl.cpp:58-64
double smc::regge_sigma_inel(double ecm) {
  const double s = ecm * ecm;
  const double sig = 22.0 * std::pow(s, 0.079) + 56.1 * std::pow(s, -0.46);
  const double bel = 2.0 * 2.3 + 2.0 * 2.3 + 4.0 * std::pow(s, 0.0808) - 4.2;
  const double sigel = 0.0511 * sig * sig / bel;
  return sig - sigel;
}

namespace {

// Woods-Saxon radius / diffuseness per mass number (Nucleus.cpp:65-111)
void ws_table(int A, double& R, double& a) {
  const double Ad = (double)A;
  R = 1.12 * std::pow(Ad, 0.333333) - 0.86 / std::pow(Ad, 0.333333);
  a = 0.54;
  if (A == 197) { R = 6.42; a = 0.45; }
  else if (A == 63) { R = 4.28; a = 0.5; }
  else if (A == 238) { R = 6.86; a = 0.44; }
  else if (A == 208) { R = 6.67; a = 0.44; }
  else if (A == 129) { R = 5.36; a = 0.590; }
}

// Deformation per mass number (Nucleus.cpp:122-148)
void deform_table(int A, double& b2, double& b4) {
  b2 = 0.0;
  b4 = 0.0;
  if (A == 197) { b2 = -0.13; b4 = -0.03; }
  else if (A == 63) { b2 = 0.162; b4 = 0.006; }
  else if (A == 129) { b2 = 0.162; b4 = -0.003; }
  else if (A == 238) { b2 = 0.28; b4 = 0.093; }
}

struct HostTable {
  std::vector<int> thr;
  int vlo = 0;
  float a = 0.f, b = 0.f;
};

// Thresholds of a monotone non-decreasing step function on [pmin, pmax].
HostTable build_steps(const std::function<long long(long long)>& f, long long pmin, long long pmax) {
  HostTable T;
  const long long vlo = f(pmin), vhi = f(pmax);
  if (vhi < vlo) throw SmcError(SMC_ERR_ARG, "footprint table is not monotone");
  T.vlo = (int)vlo;
  const long long n = vhi - vlo;
  T.thr.resize((size_t)n);
  for (long long k = 0; k < n; ++k) {
    const long long target = vlo + k + 1;
    long long lo = (k > 0) ? T.thr[(size_t)k - 1] : pmin, hi = pmax;  // f(hi) >= target
    while (lo < hi) {
      const long long mid = lo + (hi - lo) / 2;
      if (f(mid) >= target) hi = mid; else lo = mid + 1;
    }
    T.thr[(size_t)k] = (int)lo;
  }
  if (n >= 2) {
    const double a = (double)(n - 1) / ((double)T.thr[(size_t)n - 1] - (double)T.thr[0]);
    T.a = (float)a;
    T.b = (float)(1.0 - a * (double)T.thr[0]);
  } else {
    T.a = 0.f;
    T.b = 0.f;
  }
  return T;
}

}  // namespace

// Scales of the tile smearing's flush-to-zero disk cut.  With the cut the
// product (K gx) gy reaches the smallest normal float 2^-126 exactly at
// gx gy = exp2(-cut2 c2), c2 = inv2w2 log2 e as the kernel forms it;
// weights carry 2^64 so that w p stays normal.  Without a cut (cut2
// beyond ~100 decades of the Gaussian) K = 2^64 and nothing flushes.
void smc::set_tile_scales(DevCfg& g) {
  const float c2 = g.inv2w2 * 1.4426950408889634f;
  const double arg = (double)g.cut2 * (double)c2;
  if (arg < 100.0) {
    g.tile_k = (float)std::ldexp(std::exp2(arg), -126);
    g.tile_w = 0x1p64f;
  } else {
    g.tile_k = 0x1p64f;
    g.tile_w = 1.f;
  }
  g.tile_out = (float)(1.0 / ((double)g.tile_k * (double)g.tile_w));
}

struct smc_ctx {
  smc_params p{};
  smc_derived d{};
  DevCfg cfg{};
  EvBufs buf{};
  hipStream_t stream = nullptr;
  std::string err;
  // smc_run_graph: the captured batch of graph_nev events (k_event's node
  // takes new key / event0 values per replay)
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  hipGraphNode_t graph_event_node = nullptr;
  hipKernelNodeParams graph_event_params{};
  std::vector<void*> graph_args;
  uint2 graph_key{0, 0};
  unsigned long long graph_event0 = 0;
  int graph_nev = 0, graph_warm = 0, graph_cm_nu = 0;
  // mixed collision systems (smc_set_systems)
  std::vector<NucSys> mix_sys;  // 2 per system
  std::vector<int> mix_of;      // system of event g: mix_of[g % size]
  NucSys* d_mix_sys = nullptr;
  int* d_mix_of = nullptr;
  unsigned long long batch_event0 = 0;
  unsigned long long batch_seed = 0;  // seed of the current batch's sampled events
  int batch_mode = 0;
  int cap = 0;        // max events
  int nev = 0;        // events currently held
  int have_events = 0, have_density = 0, have_thickness = 0, have_moments = 0;
  int grid_sparse = 0;      // grid holds only the tiles inside each event's source rectangle
  // smc_run_stream: second record buffer, pinned staging, the copy stream and
  // its events (created on first use)
  double* evout_alt = nullptr;
  double* stage[2] = {nullptr, nullptr};
  hipStream_t copy_stream = nullptr;
  hipEvent_t ev_done[2] = {nullptr, nullptr}, ev_copied[2] = {nullptr, nullptr};
  int* dtab = nullptr;
  float* grid = nullptr;
  float* ta = nullptr;
  float* tb = nullptr;
  float* tmp = nullptr;
  double* evout = nullptr;
  double* cm = nullptr;     // centre-of-mass partials [cap][max(tile units, band units)][4]
  double* band = nullptr;   // moment partials [cap][moments_band_doubles()]
  int cm_nu = 0;            // units per event in cm for the current grid (0: none)
  void* mrects = nullptr;   // hot-spot mask rectangles [cap][2 kAMax] (8 bytes each)
  float* aux0 = nullptr;              // rho_binary / spectator A grids [cap][nx][ny]
  float* aux1 = nullptr;              // spectator B grids
  int4* spec_src = nullptr;           // spectator source lists [cap][2 kAMax]
  int4* spec_swin = nullptr;
  int* spec_counts = nullptr;         // [cap][8] (counts of the spectator lists)
  double* avg_rot = nullptr;          // averaged profiles: per event {xc, yc, angle, total}
  unsigned char* avg_acc = nullptr;   // per event: included in the average
  double* avg_sum = nullptr;          // [2][avg_nord][nx][ny] sums (rotated, recentred)
  double* avg_part = nullptr;         // [kAccumSegments][nx][ny] partial sums over event segments
  double* avg_ext = nullptr;          // [10][avg_nord][nx][ny] sums of the optional averages
  double* grid64 = nullptr;           // FP64 density grids [grid64_cap][nx][ny] (smc_density_f64)
  int grid64_cap = 0;
  int avg_nord = 0;
  int* nmrect = nullptr;    // [cap]
  double* kln_table = nullptr;
  // MC-KLN quadrature nodes (FP32 for the per-cell kernel, FP64 for points)
  void* kln_ptf = nullptr;
  void* kln_nodes_f = nullptr;
  void* kln_ptd = nullptr;
  void* kln_nodes_d = nullptr;
  unsigned* kln_list = nullptr;   // active cells of the per-cell quadrature (an event chunk), then its count
  int kln_npt = 0, kln_nn = 0;
  KlnConst kln_kc{};
  // differentiable batches
  DiffBufs dif{};
  void* dif_alloc[16] = {};
  int dif_nev = 0, dif_fwd = 0;
  uint64_t dif_seed = 0, dif_event0 = 0;
  double dif_theta0[5] = {}, dif_theta[5] = {};
  double* dif_out = nullptr;
  double* dif_cot = nullptr;
  // timing
  int timing = 0;
  struct Pending { int k; hipEvent_t a, b; };
  std::vector<Pending> pending;
  double tms[SMC_K_COUNT] = {0};
  int tcount[SMC_K_COUNT] = {0};
};

namespace {

int fail(smc_ctx* c, const SmcError& e) {
  if (c) c->err = e.what();
  return e.code;
}

struct TimerScope {
  smc_ctx* c;
  int k;
  hipEvent_t a = nullptr, b = nullptr;
  TimerScope(smc_ctx* c_, int k_) : c(c_), k(k_) {
    if (c->timing) {
      HIPCHK(hipEventCreate(&a));
      HIPCHK(hipEventCreate(&b));
      HIPCHK(hipEventRecord(a, c->stream));
    }
  }
  void done() {
    if (c->timing) {
      HIPCHK(hipEventRecord(b, c->stream));
      c->pending.push_back({k, a, b});
      a = b = nullptr;
    }
  }
};

void check_launch(int rc, const char* what) {
  if (rc != 0) throw SmcError(SMC_ERR_HIP, std::string(what) + ": " + hipGetErrorString((hipError_t)rc));
}

void gauss_legendre01(int n, std::vector<double>& x, std::vector<double>& w);

// sigma_gg of the Gaussian collision test: Newton on
//   sigma_in = int_0^{5w} 2 pi b db (1 - exp(-sigma_gg T_pp(b))),
//   T_pp(b) = e^{-b^2/4w^2} / (4 pi w^2),
// with a 38-node Gauss-Legendre rule, started at 10 fm^2 and stopped once a
// step is at most 1e-4 -- the iteration of GaussianNucleonsCal.cpp:130-163.
double sig_eff(double sigin_mb, double w) {
  std::vector<double> x, wt;
  gauss_legendre01(38, x, wt);
  const double sigin = sigin_mb * 0.1, bmax = 5.0 * w, w4 = 4.0 * w * w;
  double s = 10.0, s0;
  int it = 0;
  do {
    s0 = s;
    double sum = 0.0, dn = 0.0;
    for (int i = 0; i < 38; ++i) {
      const double b = x[i] * bmax, db = wt[i] * bmax;
      const double tpp = std::exp(-b * b / w4) / (M_PI * w4);
      sum += 2 * M_PI * b * db * (1.0 - std::exp(-s * tpp));
      dn += 2 * M_PI * b * db * tpp * std::exp(-s * tpp);
    }
    s -= (sum - sigin) / dn;
  } while (std::fabs(s - s0) > 1e-4 && ++it < 1000);
  return s;
}

void validate_and_derive(smc_ctx* c) {
  smc_params& p = c->p;
  smc_derived& d = c->d;
  if (!(p.which_mc_model == 5 || p.which_mc_model == 7 || p.which_mc_model == 1))
    throw SmcError(SMC_ERR_UNSUPPORTED, "which_mc_model must be 1, 5 or 7");
  if (p.which_mc_model == 5 && p.sub_model != 1 && p.sub_model != 2)
    throw SmcError(SMC_ERR_ARG, "MC-Glauber sub_model must be 1 or 2 (MCnucl.cpp:704-720)");
  if (p.which_mc_model == 1 && p.sub_model != 7)
    throw SmcError(SMC_ERR_UNSUPPORTED, "MC-KLN requires sub_model 7 (KLN uGD)");
  if (!(p.shape_of_nucleons >= 2 && p.shape_of_nucleons <= 4))
    throw SmcError(SMC_ERR_UNSUPPORTED, "shape_of_nucleons must be 2, 3 or 4 (Gaussian nucleons)");
  if (p.shape_of_nucleons == 3 && !(p.gaussian_lambda > 0))
    throw SmcError(SMC_ERR_ARG, "shape_of_nucleons 3 needs gaussian_lambda > 0");
  // Gaussian entropy: 2..9 share the Gaussian density (MCnucl.cpp:752-756,
  // 853-859) and the fluctuation law (:1278-1281, :314); 1 (disk) and 3
  // (quark fluctuations) are not on this path
  if (!(p.shape_of_entropy == 2 || (p.shape_of_entropy >= 4 && p.shape_of_entropy <= 9)))
    throw SmcError(SMC_ERR_UNSUPPORTED, "shape_of_entropy must be 2 or 4..9 (Gaussian entropy)");
  // MCnucl.cpp:363-383: 1 disk, 2 Gaussian; other values follow
  // shape_of_entropy: 2 the Gaussian test, 4..9 the disk
  d.criterion = (p.collision_criterion >= 1 && p.collision_criterion <= 4) ? p.collision_criterion
                                                                           : (p.shape_of_entropy == 2 ? 2 : 1);
  if (d.criterion > 2)
    throw SmcError(SMC_ERR_UNSUPPORTED, "collision_criterion 3 and 4 (quark / density collisions) are not on this path");
  if (!(p.cc_fluctuation_model == 0 || p.cc_fluctuation_model == 6))
    throw SmcError(SMC_ERR_UNSUPPORTED, "cc_fluctuation_model must be 0 or 6");
  const int As[2] = {p.Aproj, p.Atarg};
  for (int i = 0; i < 2; ++i) {
    if (As[i] < 1 || As[i] > kAMax) throw SmcError(SMC_ERR_ARG, "A must be in [1, 256]");
    if (As[i] == 3 || As[i] == 4 || As[i] == 12 || As[i] == 16)
      throw SmcError(SMC_ERR_UNSUPPORTED, "light nuclei read configuration tables (not on this path)");
  }
  if (!(p.dx > 0 && p.dy > 0 && p.maxx > 0 && p.maxy > 0)) throw SmcError(SMC_ERR_ARG, "bad grid");
  if (p.ecc_from_order < 1 || p.ecc_to_order > 9 || p.ecc_from_order > p.ecc_to_order)
    throw SmcError(SMC_ERR_ARG, "ecc orders must satisfy 1 <= from <= to <= 9");
  if (p.bmax < p.bmin || p.bmin < 0) throw SmcError(SMC_ERR_ARG, "bad impact parameter range");
  if (p.smear_path < 0 || p.smear_path > 4) throw SmcError(SMC_ERR_ARG, "smear_path must be 0, 1, 2, 3 or 4");

  // sigma_inel and width (MCnucl.cpp:58-64, 93; GaussianNucleonsCal.cpp:34-54)
  d.siginNN = (p.sigma_nn_mb > 0) ? p.sigma_nn_mb : regge_sigma_inel(p.ecm);
  double ratio3 = 0.0;  // shape 3: sigma_in / sigma_gg (GaussianNucleonsCal.cpp:41-46)
  if (p.shape_of_nucleons == 3) {
    // (gamma + int_lambda^{lambda+100} e^-t / t dt + ln lambda) / lambda, the
    // integral by 16-node Gauss-Legendre on each of 100 unit intervals
    std::vector<double> gx, gw;
    gauss_legendre01(16, gx, gw);
    const double lam = p.gaussian_lambda;
    double e1 = 0.0;
    for (int k = 0; k < 100; ++k)
      for (int i = 0; i < 16; ++i) {
        const double t = lam + k + gx[i];
        e1 += gw[i] * std::exp(-t) / t;
      }
    ratio3 = (0.57721566490153286 + e1 + std::log(lam)) / lam;
  }
  if (p.shape_of_nucleons == 2) d.width = std::sqrt(0.1 * d.siginNN / M_PI) / std::sqrt(8);
  else if (p.shape_of_nucleons == 3)
    d.width = std::sqrt(d.siginNN * 0.1 / (4 * M_PI * p.gaussian_lambda * ratio3));
  else d.width = p.gauss_nucl_width;
  d.dsq = 0.1 * d.siginNN / M_PI / 1;
  const double w = d.width;
  if (!(w > 0.05 && w <= 1.0)) throw SmcError(SMC_ERR_ARG, "nucleon width must be in (0.05, 1] fm");
  if (d.criterion == 1 && std::sqrt(d.dsq) > 8 * w)
    throw SmcError(SMC_ERR_UNSUPPORTED, "collision disk wider than the nucleon box (sweep pruning differs)");
  d.sigma_gg = (p.shape_of_nucleons == 3) ? d.siginNN * 0.1 / ratio3 : sig_eff(d.siginNN, w);
  d.box_lim = (int64_t)std::floor(8.0 * w / kQ);
  // grid (MCnucl.cpp:33-40)
  d.xmin = -p.maxx;
  d.ymin = -p.maxy;
  d.nx = (int)((p.maxx - d.xmin) / p.dx + 0.1) + 1;
  d.ny = (int)((p.maxy - d.ymin) / p.dy + 0.1) + 1;
  if (d.nx > 2048 || d.ny > 2048) throw SmcError(SMC_ERR_ARG, "grid too large");
  if (p.maxx + 4 * std::max(w, p.quark_width) > 56 || p.maxy > 56)
    throw SmcError(SMC_ERR_ARG, "grid extent beyond the 56 fm position range");

  // nuclear geometry (Nucleus.cpp:65-148), extension keys override
  const double Rext[2] = {p.proj_ws_R, p.targ_ws_R}, aext[2] = {p.proj_ws_a, p.targ_ws_a};
  const double b2ext[2] = {p.proj_beta2, p.targ_beta2}, b4ext[2] = {p.proj_beta4, p.targ_beta4};
  const int defo[2] = {p.proj_deformed, p.targ_deformed};
  for (int i = 0; i < 2; ++i) {
    double R, a, b2, b4;
    ws_table(As[i], R, a);
    deform_table(As[i], b2, b4);
    if (Rext[i] > 0) R = Rext[i];
    if (aext[i] > 0) a = aext[i];
    if (!std::isnan(b2ext[i])) b2 = b2ext[i];
    if (!std::isnan(b4ext[i])) b4 = b4ext[i];
    if (!defo[i]) b2 = b4 = 0.0;
    d.ws_R[i] = R;
    d.ws_a[i] = a;
    d.beta2[i] = b2;
    d.beta4[i] = b4;
    NucSys& ns = c->cfg.nuc[i];
    ns.A = As[i];
    ns.kind = As[i] == 1 ? kProton : (As[i] == 2 ? kDeuteron : kWoodsSaxon);
    ns.deformed = defo[i] ? 1 : 0;
    ns.R = (float)R;
    ns.a = (float)a;
    ns.beta2 = (float)b2;
    ns.beta4 = (float)b4;
    ns.rmax = (float)(R + 2.5);
    ns.rwmax = (float)(1.0 / (1.0 + std::exp(-R / a)));
    ns.inv_a = (float)(1.0 / a);
  }
  d.dmin = (p.nucleon_min_dist >= 0) ? p.nucleon_min_dist : 0.9;

  // exact collision predicate: largest n = dxq^2 + dyq^2 with
  // fl(fl(sqrt(n q^2))^2) <= dsq (MCnucl.cpp:359-366)
  {
    auto hit = [&](long long n) {
      const double d2 = (double)n * (kQ * kQ);
      const double b = std::sqrt(d2);
      return b * b <= d.dsq;
    };
    long long lo = 0, hi = (long long)1 << 52;
    while (lo < hi) {
      const long long mid = lo + (hi - lo + 1) / 2;
      if (hit(mid)) lo = mid; else hi = mid - 1;
    }
    d.hit_n_max = lo;
  }
  // KLN table step (MCnucl.cpp:915-919)
  d.kln_dT = 10.0 / d.siginNN / 1 / std::max(1, p.tmax_subdivision);
  d.kln_tmax = std::max(1, p.tmax_subdivision) * (p.tmax - 1) + 1;
  // sixPoint2dInterp reads rows and columns i .. i + 2 (arsenal.cpp:33-54)
  if (p.which_mc_model == 1 && !p.kln_direct && d.kln_tmax < 3)
    throw SmcError(SMC_ERR_ARG, "the dN/dy table needs tmax_subdivision * (tmax - 1) + 1 >= 3");

  DevCfg& g = c->cfg;
  g.dmin2 = (float)(d.dmin * d.dmin);
  g.bmin2 = p.bmin * p.bmin;
  g.bmax2 = p.bmax * p.bmax;
  g.npmin = p.Npmin;
  g.npmax = p.Npmax;
  g.max_attempts = p.max_attempts > 0 ? p.max_attempts : 10000;
  g.max_pairs = p.max_pairs > 0 ? p.max_pairs : std::min(p.Aproj * p.Atarg, 8192);
  g.hit_n = d.hit_n_max;
  g.hit_lim = (int)std::floor(std::sqrt((double)d.hit_n_max)) + 1;
  g.crit_gauss = d.criterion == 2;
  g.spec_attempts = 1;
  g.box_lim = (int)d.box_lim;
  g.gauss_k = (float)(1.4426950408889634 * kQ * kQ / (4.0 * w * w));
  g.gauss_A = (float)(d.sigma_gg / (4.0 * M_PI * w * w));
  g.fluct = p.cc_fluctuation_model == 6;
  {
    // MCnucl.cpp:1271-1301 (shape_of_entropy != 3)
    const double eps = 1e-8, th = p.cc_fluctuation_Gamma_theta, gk = 1. / th;
    g.gk_part = (float)((1 - p.alpha + eps) / 2. * gk);
    g.gs_part = (float)(2. / (1 - p.alpha + eps) * th);
    g.gk_bin = (float)((p.alpha + eps) * gk);
    g.gs_bin = (float)(1. / (p.alpha + eps) * th);
    if (g.fluct && !(th > 0)) throw SmcError(SMC_ERR_ARG, "cc_fluctuation_Gamma_theta must be > 0");
  }
  g.nx = d.nx;
  g.ny = d.ny;
  g.xmin = d.xmin;
  g.ymin = d.ymin;
  g.dx = p.dx;
  g.dy = p.dy;
  g.w = (float)w;
  g.inv2w2 = (float)(1.0 / (2 * w * w));
  g.norm = (float)(1 / (2 * M_PI * w * w));
  g.cut2 = (float)(25. * w * w);
  set_tile_scales(g);
  const double hb = 4 * std::max(w, p.quark_width);
  g.hb_h = (int)std::llround(hb / kH);
  g.alpha = (float)p.alpha;
  g.pref_part = (float)((1.0 - p.alpha) / 2.);
  g.use_binary = p.alpha > 1e-8;
  g.uli = p.which_mc_model == 5 && p.sub_model == 2;
  g.final_factor = p.finalFactor;
  g.from_order = p.ecc_from_order;
  g.to_order = p.ecc_to_order;
}

// Footprint tables reproducing the reference's FP64 index expressions.
void build_tables(smc_ctx* c) {
  const smc_params& p = c->p;
  const double w = c->d.width;
  const double d_max = 5. * w;  // MCnucl.cpp:455-456, 648-649
  const double Xmin = c->d.xmin, Ymin = c->d.ymin, dx = p.dx, dy = p.dy;
  const int nx = c->d.nx, ny = c->d.ny;
  const double qw = p.quark_width;
  const long long PH = (long long)(56.0 / kH), PQ = (long long)(120.0 / kQ), PO = (long long)(60.0 / kQ);
  auto clampi = [](long long v, long long lo, long long hi) { return std::max(lo, std::min(hi, v)); };
  // participant box edge: base box (Particle.cpp:19-20) united with the
  // quark boxes at the centre (Particle.cpp:94-99, Quark.cpp:5-10)
  auto boxL = [&](double x) { return std::min(x - 8 * w / 2, x - 8 * qw / 2); };
  auto boxR = [&](double x) { return std::max(x + 8 * w / 2, x + 8 * qw / 2); };
  std::vector<HostTable> T(kNumTables);
  T[kWinLoX] = build_steps([&](long long ph) { double x = ph * kH; return clampi((int)((x - d_max - Xmin) / dx), 0, nx); }, -PH, PH);
  T[kWinHiX] = build_steps([&](long long ph) { double x = ph * kH; return clampi((int)((x + d_max - Xmin) / dx), 0, nx); }, -PH, PH);
  T[kWinLoY] = build_steps([&](long long ph) { double y = ph * kH; return clampi((int)((y - d_max - Ymin) / dy), 0, ny); }, -PH, PH);
  T[kWinHiY] = build_steps([&](long long ph) { double y = ph * kH; return clampi((int)((y + d_max - Ymin) / dy), 0, ny); }, -PH, PH);
  // MCnucl.cpp:831-844
  T[kBoxLoX] = build_steps([&](long long ph) { double x = ph * kH; return clampi((int)((boxL(x) - Xmin) / dx), 0, nx); }, -PH, PH);
  T[kBoxHiX] = build_steps([&](long long ph) { double x = ph * kH; return clampi((int)((boxR(x) - Xmin) / dx), 0, nx); }, -PH, PH);
  T[kBoxLoY] = build_steps([&](long long ph) { double y = ph * kH; return clampi((int)((boxL(y) - Ymin) / dy), 0, ny); }, -PH, PH);
  T[kBoxHiY] = build_steps([&](long long ph) { double y = ph * kH; return clampi((int)((boxR(y) - Ymin) / dy), 0, ny); }, -PH, PH);
  // MakeDensity.cpp:2355-2357 and :2399-2400
  T[kMaskX] = build_steps([&](long long dq) { return (long long)(int)((dq * kQ) / dx); }, 0, PQ);
  T[kMaskY] = build_steps([&](long long dq) { return (long long)(int)((dq * kQ) / dy); }, 0, PQ);
  T[kOffX] = build_steps([&](long long rq) { return (long long)(int)((rq * kQ - Xmin) / dx); }, -PO, 0);
  T[kOffY] = build_steps([&](long long rq) { return (long long)(int)((rq * kQ - Ymin) / dy); }, -PO, 0);
  size_t total = 0;
  for (auto& t : T) total += t.thr.size();
  c->dtab = dalloc<int>(total);
  size_t off = 0;
  for (int i = 0; i < kNumTables; ++i) {
    StepTable& st = c->cfg.tab[i];
    st.thr = c->dtab + off;
    st.n = (int)T[i].thr.size();
    st.vlo = T[i].vlo;
    st.a = T[i].a;
    st.b = T[i].b;
    if (st.n) HIPCHK(hipMemcpy(c->dtab + off, T[i].thr.data(), st.n * sizeof(int), hipMemcpyHostToDevice));
    off += T[i].thr.size();
  }
}

// Gauss-Legendre nodes and weights on [0, 1] (Newton iteration on P_n).
void gauss_legendre01(int n, std::vector<double>& x, std::vector<double>& w) {
  x.assign(n, 0.0);
  w.assign(n, 0.0);
  for (int i = 0; i < (n + 1) / 2; ++i) {
    double z = std::cos(M_PI * (i + 0.75) / (n + 0.5)), pp = 1.0;
    for (int it = 0; it < 100; ++it) {
      double p1 = 1.0, p2 = 0.0;
      for (int j = 1; j <= n; ++j) {
        const double p3 = p2;
        p2 = p1;
        p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
      }
      pp = n * (z * p1 - p2) / (z * z - 1.0);
      const double z1 = z;
      z = z1 - p1 / pp;
      if (std::fabs(z - z1) < 1e-16) break;
    }
    x[i] = 0.5 * (1.0 - z);
    x[n - 1 - i] = 0.5 * (1.0 + z);
    w[i] = w[n - 1 - i] = 1.0 / ((1.0 - z * z) * pp * pp);
  }
}

// Precomputes the cell-independent parts of KLNModel::func (KLNModel.cpp:
// 219-277) on the (pT, kT, phi) product rule over the unit cube, with the
// normalisation of getdNdy (KLNModel.cpp:97-110, g2hfac = 2 x 1.0).
void build_kln(smc_ctx* c) {
  const smc_params& p = c->p;
  const int npt = std::max(1, p.kln_n_pt), nkt = std::max(1, p.kln_n_kt), nphi = std::max(1, p.kln_n_phi);
  const double ecm = p.ecm, lam = p.lambda, ptmin = p.kln_pt_min, ptmax = p.kln_pt_max;
  const double CF = (3.0 * 3.0 - 1.0) / (2 * 3.0), hbarc = 0.197327053;
  const double Norm = 2. / CF / (hbarc * hbarc), g2hfac = 2.0 * 1.0;
  const double fac = CF * 2. / (3. * M_PI * M_PI);
  const double lq2 = 0.2 * 0.2, amax = 0.5, beta0 = (33.0 - 2.0 * 3.0) / (12 * M_PI);
  auto alpha = [&](double q2) { return (q2 <= lq2) ? amax : std::min(amax, 1.0 / (beta0 * std::log(q2 / lq2))); };
  std::vector<double> xp, wp, xk, wk, xf, wf;
  gauss_legendre01(npt, xp, wp);
  gauss_legendre01(nkt, xk, wk);
  gauss_legendre01(nphi, xf, wf);
  // phi and 2 pi - phi give the same k1^2, k2^2: merge the symmetric pairs
  std::vector<double> uf, ww;
  for (int j = 0; j < nphi / 2; ++j) { uf.push_back(xf[j]); ww.push_back(2.0 * wf[j]); }
  if (nphi % 2) { uf.push_back(xf[nphi / 2]); ww.push_back(wf[nphi / 2]); }
  int nn = nkt * (int)uf.size();
  const int nn_pad = (nn + 15) & ~15;  // the FP32 kernel walks nodes 16 at a time (zero nodes add 0)
  std::vector<double> ptd(2 * npt), nd(4 * (size_t)npt * nn_pad, 0.0);
  std::vector<float> ptf(2 * npt), nf(4 * (size_t)npt * nn_pad, 0.f);
  for (int a = 0; a < npt; ++a) {
    const double pt = ptmin + xp[a] * (ptmax - ptmin);
    const double x = pt / ecm;
    double P = 0.0, g = 0.0;
    if (x <= 1.0) {
      g = 2. / 1.53 * std::pow(0.01 / x, lam);
      P = wp[a] * fac * fac * std::pow(1.0 - x, 8.) * 2.0 * M_PI * (ptmax - ptmin) * 2.0 * M_PI * pt / 4.0 *
          (1.0 * g2hfac * Norm * 9. / 32.);
    }
    ptd[2 * a] = P;
    ptd[2 * a + 1] = g;
    ptf[2 * a] = (float)P;
    ptf[2 * a + 1] = (float)g;
    int m = 0;
    for (int b = 0; b < nkt; ++b)
      for (size_t f = 0; f < uf.size(); ++f, ++m) {
        const double kt = pt * xk[b], phi = 2 * M_PI * uf[f];
        const double k1 = 0.25 * (pt * pt + kt * kt + 2 * kt * pt * std::cos(phi));
        const double k2 = 0.25 * (pt * pt + kt * kt - 2 * pt * kt * std::cos(phi));
        const double cn = wk[b] * ww[f] * xk[b] * alpha(std::max(std::max(k1, k2), pt * pt));
        const size_t o = 4 * ((size_t)a * nn_pad + m);
        nd[o] = cn;
        nd[o + 1] = 1.0 / k1;
        nd[o + 2] = 1.0 / k2;
        nf[o] = (float)cn;
        nf[o + 1] = (float)(1.0 / k1);
        nf[o + 2] = (float)(1.0 / k2);
      }
  }
  c->kln_npt = npt;
  c->kln_nn = nn_pad;
  c->kln_kc = KlnConst{lq2, amax, 1.0 / beta0};
  c->kln_ptf = dalloc<float>(ptf.size());
  c->kln_nodes_f = dalloc<float>(nf.size());
  c->kln_ptd = dalloc<double>(ptd.size());
  c->kln_nodes_d = dalloc<double>(nd.size());
  HIPCHK(hipMemcpy(c->kln_ptf, ptf.data(), ptf.size() * sizeof(float), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->kln_nodes_f, nf.data(), nf.size() * sizeof(float), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->kln_ptd, ptd.data(), ptd.size() * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->kln_nodes_d, nd.data(), nd.size() * sizeof(double), hipMemcpyHostToDevice));
}

void alloc_buffers(smc_ctx* c) {
  const size_t E = (size_t)c->cap;
  const size_t MP = (size_t)c->cfg.max_pairs;
  EvBufs& b = c->buf;
  b.pos = dalloc<int>(E * 2 * kAMax * 2);
  b.fluct = dalloc<float>(E * 2 * kAMax);
  b.wound = dalloc<int>(E * 2 * kAMax);
  b.pairs = dalloc<unsigned>(E * MP);
  b.pfluct = dalloc<float>(E * MP);
  b.paw = c->cfg.uli ? dalloc<unsigned char>(E * MP) : nullptr;
  b.counts = dalloc<int>(E * 8);
  b.bval = dalloc<double>(E);
  b.raw = dalloc<int>(E * 2 * kAMax * 2);
  b.src_cap = 2 * kAMax + (int)MP;
  b.src = dalloc<int4>(E * (size_t)b.src_cap);
  b.swin = dalloc<int4>(E * (size_t)b.src_cap);
  c->evout = dalloc<double>(E * SMC_EV_STRIDE);
  c->cm = dalloc<double>(E * (size_t)std::max(tile_units_max(c->cfg), moments_cm_units_min()) * 4);
  c->band = dalloc<double>(E * (size_t)moments_band_doubles());
  c->mrects = dalloc<unsigned long long>(E * 2 * kAMax);
  c->nmrect = dalloc<int>(E);
  HIPCHK(hipMemset(b.counts, 0, E * 8 * sizeof(int)));
}

float* grid_buffer(smc_ctx* c, float** slot) {
  if (!*slot) *slot = dalloc<float>((size_t)c->cap * c->d.nx * c->d.ny);
  return *slot;
}

void sync(smc_ctx* c) { HIPCHK(hipStreamSynchronize(c->stream)); }

SmearSpec density_spec(const smc_ctx* c) {
  const DevCfg& g = c->cfg;
  SmearSpec s{};
  if (c->p.which_mc_model == 5 && g.uli) {
    // MCnucl.cpp:713-717, 724-760 ("Ulrich" Glauber, sub_model 2): no
    // participant term; pair k adds f_k (alpha + (1 - alpha) aw_k) G, the
    // product held in the pair's source weight (k_event)
    s.enabled[kSrcPair] = g.use_binary;
    s.win[kSrcPair] = kWin5;
    s.scale[kSrcPair] = g.norm;
    s.use_fluct[kSrcPair] = 1;
    s.combine = 0;
  } else if (c->p.which_mc_model == 5) {
    // MCnucl.cpp:689-760: (1-alpha)/2 sum_part f T_n + alpha sum_coll f G
    s.enabled[kSrcPartA] = s.enabled[kSrcPartB] = 1;
    s.win[kSrcPartA] = s.win[kSrcPartB] = kWinBox;
    s.scale[kSrcPartA] = s.scale[kSrcPartB] = g.pref_part * g.norm;
    s.use_fluct[kSrcPartA] = s.use_fluct[kSrcPartB] = g.fluct;
    s.enabled[kSrcPair] = g.use_binary;
    s.win[kSrcPair] = kWin5;
    s.scale[kSrcPair] = g.alpha * g.norm;
    s.use_fluct[kSrcPair] = g.fluct;
    s.combine = 0;
  } else {
    // MCnucl.cpp:762-811: sqrt(rho_A rho_B), rho from addDensity
    s.enabled[kSrcPartA] = s.enabled[kSrcPartB] = 1;
    s.win[kSrcPartA] = s.win[kSrcPartB] = kWinBox;
    s.scale[kSrcPartA] = s.scale[kSrcPartB] = g.norm;
    s.use_fluct[kSrcPartA] = s.use_fluct[kSrcPartB] = g.fluct;
    s.acc[kSrcPartB] = 1;
    s.combine = 1;
  }
  return s;
}

SmearSpec thickness_spec(const smc_ctx* c) {
  // MCnucl.cpp:432-478: participants, +-5w window, getSmoothTn (no fluct)
  SmearSpec s{};
  s.enabled[kSrcPartA] = s.enabled[kSrcPartB] = 1;
  s.win[kSrcPartA] = s.win[kSrcPartB] = kWin5;
  s.scale[kSrcPartA] = s.scale[kSrcPartB] = c->cfg.norm;
  s.acc[kSrcPartB] = 1;
  s.combine = 2;
  return s;
}

// smear_path: 0 exact footprint on VALU tiles (k_tile: 32 x 64 by default,
// one wave, disk cut by flushing denormal products); 1 the same footprint as
// an FP32 MFMA contraction over the square window with the out-of-disk
// products subtracted (k_mtile); 2 exact footprint, per-cell exponentials
// (k_smear); 3 and 4 as 1 with BF16x3 and BF16x6 split-precision MFMA.
// Every path takes every model and grid size.
static int mfma_prec(const smc_ctx* c) {
  switch (c->p.smear_path) {
    case 1: return 0;
    case 3: return 1;
    case 4: return 2;
    default: return -1;
  }
}

// centre-of-mass units per event written by the density launch of the
// current path (0: none)
static int density_cm_units(const smc_ctx* c) {
  if (c->p.smear_path == 2) return 0;
  return mfma_prec(c) >= 0 ? mtile_units(c->cfg) : tile_units(c->cfg);
}

void run_smear(smc_ctx* c, const SmearSpec& spec, float* out0, float* out1) {
  TimerScope ts(c, SMC_K_SMEAR);
  const int prec = mfma_prec(c);
  if (prec >= 0)
    check_launch(launch_mtile(c->cfg, c->buf, spec, c->nev, out0, out1, prec, nullptr, 0, 0.0, c->stream), "k_mtile");
  else if (c->p.smear_path == 2)
    check_launch(launch_smear(c->cfg, c->buf, spec, c->nev, out0, out1, c->stream), "k_smear");
  else
    check_launch(launch_tile(c->cfg, c->buf, spec, c->nev, out0, out1, c->stream), "k_tile");
  ts.done();
}

// Completes the density grid of a sparse launch (tiles outside an event's
// source rectangle not stored) before whole grids are read.
void densify_grid(smc_ctx* c) {
  if (!c->grid_sparse) return;
  check_launch(launch_zero_outside(c->grid, c->buf.counts, c->d.nx, c->d.ny, c->nev, c->stream), "k_zero_outside");
  c->grid_sparse = 0;
}

void kln_points(smc_ctx* c, int n, const double* ta_dev, const double* tb_dev, double* out_dev) {
  TimerScope ts(c, SMC_K_KLN);
  check_launch(launch_kln_points(ta_dev, tb_dev, out_dev, n, c->kln_ptd, c->kln_nodes_d, c->kln_npt, c->kln_nn,
                                 c->kln_kc, c->stream),
               "k_kln_points");
  ts.done();
}

// MCnucl::makeTable (MCnucl.cpp:911-960): table[i][j] = getdNdy(0, i dT, j dT),
// row and column 0 zero.
void ensure_kln_table(smc_ctx* c) {
  if (c->kln_table) return;
  const int tmax = c->d.kln_tmax;
  const double dT = c->d.kln_dT;
  std::vector<double> a((size_t)tmax * tmax), b((size_t)tmax * tmax);
  for (int i = 0; i < tmax; ++i)
    for (int j = 0; j < tmax; ++j) {
      const bool in = i > 0 && j > 0;
      a[(size_t)i * tmax + j] = in ? dT * i : 0.0;
      b[(size_t)i * tmax + j] = in ? dT * j : 0.0;
    }
  double* da = dalloc<double>(a.size());
  double* db = dalloc<double>(b.size());
  c->kln_table = dalloc<double>(a.size());
  HIPCHK(hipMemcpyAsync(da, a.data(), a.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(db, b.data(), b.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
  kln_points(c, (int)a.size(), da, db, c->kln_table);
  sync(c);
  (void)hipFree(da);
  (void)hipFree(db);
}

// MCnucl::setDensity for which_mc_model 1 (MCnucl.cpp:654-687): T_A, T_B
// (MCnucl.cpp:398-478), then per-cell dN/dy from the table (reference) or
// from the quadrature in every cell (kln_direct).  With the table on the
// default smearing path both steps run in one tile launch, which also
// writes the centre-of-mass partials of the dN/dy grid.
// The smearing launch of the density step on the current path and tile
// shape -- what run_density and run_kln_density time as SMC_K_SMEAR -- with
// no timing and no state change.  MC-KLN through the table (not on the
// per-cell-exponential path): T_A, T_B and the table step in one launch into
// the grid; other MC-KLN paths: T_A, T_B into their grids.
void density_smear_launch(smc_ctx* c) {
  const int prec = mfma_prec(c);
  const int nu = density_cm_units(c);
  if (c->p.which_mc_model == 1) {
    if (!c->p.kln_direct && c->p.smear_path != 2) {
      SmearSpec spec = thickness_spec(c);
      spec.cm_out = c->cm;
      spec.sparse = 1;
      float* s = grid_buffer(c, &c->grid);
      if (prec >= 0)
        check_launch(launch_mtile(c->cfg, c->buf, spec, c->nev, s, nullptr, prec, c->kln_table, c->d.kln_tmax,
                                  c->d.kln_dT, c->stream),
                     "k_mtile (KLN)");
      else
        check_launch(launch_tile_kln(c->cfg, c->buf, spec, c->nev, s, c->kln_table, c->d.kln_tmax, c->d.kln_dT,
                                     c->stream),
                     "k_tile (KLN)");
      return;
    }
    const SmearSpec spec = thickness_spec(c);
    float* a = grid_buffer(c, &c->ta);
    float* b = grid_buffer(c, &c->tb);
    if (prec >= 0) check_launch(launch_mtile(c->cfg, c->buf, spec, c->nev, a, b, prec, nullptr, 0, 0.0, c->stream), "k_mtile");
    else if (c->p.smear_path == 2) check_launch(launch_smear(c->cfg, c->buf, spec, c->nev, a, b, c->stream), "k_smear");
    else check_launch(launch_tile(c->cfg, c->buf, spec, c->nev, a, b, c->stream), "k_tile");
    return;
  }
  SmearSpec spec = density_spec(c);
  // the tile kernels add the centre-of-mass partials of the stored grid (one
  // accumulator, or sqrt(T_A T_B) combined in the same launch) and skip the
  // stores of tiles outside each event's source rectangle
  if (nu) {
    spec.cm_out = c->cm;
    spec.sparse = 1;
  }
  float* g = grid_buffer(c, &c->grid);
  float* t = (spec.combine && !nu) ? grid_buffer(c, &c->tmp) : nullptr;
  if (prec >= 0) check_launch(launch_mtile(c->cfg, c->buf, spec, c->nev, g, t, prec, nullptr, 0, 0.0, c->stream), "k_mtile");
  else if (c->p.smear_path == 2) check_launch(launch_smear(c->cfg, c->buf, spec, c->nev, g, t, c->stream), "k_smear");
  else check_launch(launch_tile(c->cfg, c->buf, spec, c->nev, g, t, c->stream), "k_tile");
}

// MCnucl::setDensity for which_mc_model 1 (MCnucl.cpp:654-687): T_A, T_B
// (MCnucl.cpp:398-478), then per-cell dN/dy from the table (reference) or
// from the quadrature in every cell (kln_direct).  Through the table both
// steps run in one tile launch, which also writes the centre-of-mass
// partials of the dN/dy grid.
void run_kln_density(smc_ctx* c) {
  c->cm_nu = 0;
  if (!c->p.kln_direct) ensure_kln_table(c);
  {
    TimerScope ts(c, SMC_K_SMEAR);
    density_smear_launch(c);
    ts.done();
  }
  if (!c->p.kln_direct && c->p.smear_path != 2) {
    c->cm_nu = density_cm_units(c);
    c->grid_sparse = 1;
    return;
  }
  c->have_thickness = 1;
  float* a = grid_buffer(c, &c->ta);
  float* b = grid_buffer(c, &c->tb);
  float* s = grid_buffer(c, &c->grid);
  const size_t ncell_ev = (size_t)c->d.nx * c->d.ny;
  TimerScope ts(c, SMC_K_KLN);
  if (c->p.kln_direct) {
    // events in chunks of at most 2^28 cells: the active-cell list (32-bit
    // cell indices) stays within 1 GiB
    const int chunk = (int)std::max<size_t>(1, std::min<size_t>((size_t)c->cap, ((size_t)1 << 28) / ncell_ev));
    const size_t cells = (size_t)chunk * ncell_ev;
    if (!c->kln_list) c->kln_list = dalloc<unsigned>(cells + 1);
    for (int e0 = 0; e0 < c->nev; e0 += chunk) {
      const size_t o = (size_t)e0 * ncell_ev, n = (size_t)std::min(chunk, c->nev - e0) * ncell_ev;
      check_launch(launch_kln_direct(a + o, b + o, s + o, n, c->kln_ptf, c->kln_nodes_f, c->kln_npt, c->kln_nn,
                                     c->kln_kc, c->kln_list, c->kln_list + cells, c->stream),
                   "k_kln_direct");
    }
  } else {
    check_launch(launch_kln_table(a, b, s, (int)ncell_ev, c->nev, c->d.ny, c->kln_table, c->d.kln_tmax, c->d.kln_dT,
                                  c->buf.counts, c->stream),
                 "k_kln_table");
  }
  ts.done();
  c->grid_sparse = 0;
}

void run_density(smc_ctx* c) {
  if (!c->have_events) throw SmcError(SMC_ERR_STATE, "no events: call smc_sample or smc_set_positions first");
  if (c->p.which_mc_model == 1) {
    run_kln_density(c);
    c->have_density = 1;
    return;
  }
  TimerScope ts(c, SMC_K_SMEAR);
  density_smear_launch(c);
  ts.done();
  c->cm_nu = density_cm_units(c);
  c->grid_sparse = c->cm_nu ? 1 : 0;
  c->have_density = 1;
}

void run_moments(smc_ctx* c, int mode) {
  if (!c->have_density) run_density(c);
  TimerScope ts(c, SMC_K_MOMENTS);
  DevCfg g = c->cfg;
  if (mode == 1) g.final_factor = 1.0;
  check_launch(launch_moments(g, c->buf, c->grid, c->nev, c->evout, mode, c->cm_nu ? c->cm : nullptr, c->cm_nu,
                              c->cm, c->band, c->mrects, c->nmrect, c->stream),
               "k_moments");
  ts.done();
}

// system of event ev of the current batch, -1 when the batch is unmixed
static int event_system(const smc_ctx* c, int ev) {
  if (c->mix_of.empty() || c->batch_mode == 1) return -1;
  return c->mix_of[(c->batch_event0 + (unsigned long long)ev) % c->mix_of.size()];
}

// record slot SMC_EV_SYSTEM of copied-out records
static void tag_systems(const smc_ctx* c, double* ev_out, int nev) {
  if (!ev_out || event_system(c, 0) < 0) return;
  for (int e = 0; e < nev; ++e) ev_out[(size_t)e * SMC_EV_STRIDE + SMC_EV_SYSTEM] = event_system(c, e);
}

void run_event(smc_ctx* c, int nev, uint64_t seed, uint64_t event0, int fed) {
  c->batch_event0 = event0;
  c->batch_seed = seed;
  c->batch_mode = fed;
  TimerScope ts(c, SMC_K_EVENT);
  check_launch(launch_event(c->cfg, c->buf, nev, seed, event0, fed, c->stream), "k_event");
  ts.done();
  c->nev = nev;
  c->have_events = 1;
  c->have_density = c->have_thickness = c->have_moments = 0;
}

// ---------------------------------------------------------------- operation 2 grids
// MCnucl::calculate_rho_binary (MCnucl.cpp:481-527): the normalised Gaussian
// of every collision pair, +-5w window, d^2 <= 25 w^2, no fluctuation factor.
void run_rho_binary(smc_ctx* c, float* out) {
  SmearSpec s{};
  s.enabled[kSrcPair] = 1;
  s.win[kSrcPair] = kWin5;
  s.scale[kSrcPair] = c->cfg.norm;
  run_smear(c, s, out, nullptr);
}

// MCnucl::calculate_spectator_density (MCnucl.cpp:530-614): the same
// Gaussians on the nucleons without a collision, projectile (rapidity > 0)
// and target separately.
void run_spectator_density(smc_ctx* c, float* out_a, float* out_b) {
  const size_t E = (size_t)c->cap, cap = 2 * kAMax;
  if (!c->spec_src) {
    c->spec_src = dalloc<int4>(E * cap);
    c->spec_swin = dalloc<int4>(E * cap);
    c->spec_counts = dalloc<int>(E * 8);
  }
  EvBufs sb = c->buf;
  sb.src = c->spec_src;
  sb.swin = c->spec_swin;
  sb.counts = c->spec_counts;
  sb.src_cap = (int)cap;
  TimerScope ts(c, SMC_K_SMEAR);
  check_launch(launch_spectator_src(c->cfg, c->buf, sb, c->nev, c->stream), "k_spectator_src");
  SmearSpec s{};
  for (int k : {kSrcPartA, kSrcPartB}) {
    s.enabled[k] = 1;
    s.win[k] = kWin5;
    s.scale[k] = c->cfg.norm;
  }
  s.acc[kSrcPartB] = 1;
  s.combine = 2;
  check_launch(launch_tile(c->cfg, sb, s, c->nev, out_a, out_b, c->stream), "k_tile");
  ts.done();
}

// ---------------------------------------------------------------- averaged profiles
// Operation 3 (MakeDensity.cpp:736-1650, pT-integrated): for each order
// from..to in turn, optionally the recentred profile (reaction plane,
// generate_reaction_plane_avg_profile), then the profile recentred and
// turned to the order-n angle; the sources stay moved from one order to the
// next as in the reference.  Sums of finalFactor * rho over accepted events.
void ensure_cm(smc_ctx* c) {
  if (c->cm_nu) return;
  check_launch(launch_cm_partials(c->cfg, c->buf, c->grid, c->nev, c->cm, c->stream), "k_mom_cm");
  c->cm_nu = moments_cm_units_min();
}

// the optional averages in the order of smc_avg_extra's members
enum : int { kExTatb = 0, kExRhoBin, kExTa, kExTb, kExSpecA, kExSpecB, kExTatbRp, kExRhoBinRp, kExTaRp, kExTbRp, kExN };

void average_profiles(smc_ctx* c, int from, int to, const unsigned char* accept, double* sums, double* rp_sums,
                      const smc_avg_extra* ex) {
  if (!c->have_events) throw SmcError(SMC_ERR_STATE, "no events: call smc_sample or smc_run first");
  if (from < 1 || to < from || to > 9) throw SmcError(SMC_ERR_ARG, "orders must satisfy 1 <= from <= to <= 9");
  if (!sums) throw SmcError(SMC_ERR_ARG, "sums is NULL");
  double* want[kExN] = {};
  if (ex) {
    double* const m[kExN] = {ex->tatb,   ex->rho_binary, ex->ta,         ex->tb,    ex->spec_a,
                             ex->spec_b, ex->tatb_rp,    ex->rho_binary_rp, ex->ta_rp, ex->tb_rp};
    std::copy(m, m + kExN, want);
  }
  if (!rp_sums && (want[kExTatbRp] || want[kExRhoBinRp] || want[kExTaRp] || want[kExTbRp]))
    throw SmcError(SMC_ERR_ARG, "the reaction-plane averages need rp_sums (the recentring pass)");
  const bool need_t[2] = {want[kExTatbRp] || want[kExTaRp] || want[kExTbRp], want[kExTatb] || want[kExTa] || want[kExTb]};
  const bool need_rb[2] = {want[kExRhoBinRp] != nullptr, want[kExRhoBin] != nullptr};
  const bool need_sp = want[kExSpecA] || want[kExSpecB];
  bool any_ex = false;
  for (double* w : want) any_ex |= w != nullptr;
  const int nev = c->nev, nord = to - from + 1;
  const size_t N = (size_t)c->d.nx * c->d.ny;
  if (!c->avg_rot) {
    c->avg_rot = dalloc<double>((size_t)c->cap * 4);
    c->avg_acc = dalloc<unsigned char>((size_t)c->cap);
    c->avg_part = dalloc<double>((size_t)kAccumSegments * N);
  }
  if (c->avg_nord < nord || (any_ex && !c->avg_ext)) {
    if (c->avg_sum) HIPCHK(hipFree(c->avg_sum));
    if (c->avg_ext) HIPCHK(hipFree(c->avg_ext));
    c->avg_nord = std::max(nord, c->avg_nord);
    c->avg_sum = dalloc<double>(2 * (size_t)c->avg_nord * N);
    c->avg_ext = any_ex ? dalloc<double>((size_t)kExN * c->avg_nord * N) : nullptr;
  }
  std::vector<unsigned char> acc(nev);
  if (accept) {
    for (int e = 0; e < nev; ++e) acc[e] = accept[e] ? 1 : 0;
  } else {
    std::vector<int> cnt((size_t)nev * 8);
    HIPCHK(hipMemcpyAsync(cnt.data(), c->buf.counts, cnt.size() * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    sync(c);
    for (int e = 0; e < nev; ++e) acc[e] = cnt[(size_t)e * 8 + 3] == 0;
  }
  HIPCHK(hipMemcpyAsync(c->avg_acc, acc.data(), nev, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemsetAsync(c->avg_sum, 0, 2 * (size_t)nord * N * sizeof(double), c->stream));
  if (any_ex) HIPCHK(hipMemsetAsync(c->avg_ext, 0, (size_t)kExN * nord * N * sizeof(double), c->stream));
  if (!c->have_density) run_density(c);
  // the spectators as a source list of their own, from the unmoved nucleons;
  // they move with the sources below
  EvBufs sb = c->buf;
  if (need_sp) {
    const size_t E = (size_t)c->cap, cap = 2 * kAMax;
    if (!c->spec_src) {
      c->spec_src = dalloc<int4>(E * cap);
      c->spec_swin = dalloc<int4>(E * cap);
      c->spec_counts = dalloc<int>(E * 8);
    }
    sb.src = c->spec_src;
    sb.swin = c->spec_swin;
    sb.counts = c->spec_counts;
    sb.src_cap = (int)cap;
    check_launch(launch_spectator_src(c->cfg, c->buf, sb, nev, c->stream), "k_spectator_src");
  }
  const int nx = c->d.nx, ny = c->d.ny;
  auto accum = [&](const float* g, const int* counts, double ff, int slot, int o, const float* g2 = nullptr) {
    check_launch(launch_accum_avg(g, counts, nx, ny, nev, c->avg_acc, ff, c->avg_part,
                                  c->avg_ext + ((size_t)slot * nord + o) * N, c->stream, g2),
                 "k_accum_avg");
  };
  const double ff = c->p.finalFactor;
  for (int o = 0; o < nord; ++o) {
    const int n = from + o;
    for (int pass = rp_sums ? 0 : 1; pass < 2; ++pass) {  // 0: recentre only, 1: recentre and turn
      ensure_cm(c);
      check_launch(launch_cm_angle(c->cfg, c->buf, c->grid, c->cm, c->cm_nu, n, pass, nev, c->avg_rot, c->stream),
                   "k_cm_angle");
      check_launch(launch_rotate_src(c->cfg, c->buf, c->avg_rot, nev, c->stream), "k_rotate_src");
      if (need_sp) check_launch(launch_rotate_src(c->cfg, sb, c->avg_rot, nev, c->stream), "k_rotate_src");
      c->have_density = c->have_thickness = c->have_moments = 0;
      run_density(c);
      double* dst = c->avg_sum + ((size_t)(pass == 0 ? nord : 0) + o) * N;
      check_launch(launch_accum_avg(c->grid, c->buf.counts, c->d.nx, c->d.ny, nev, c->avg_acc, ff, c->avg_part, dst,
                                    c->stream),
                   "k_accum_avg");
      const bool rp = pass == 0;
      if (need_t[pass]) {  // MCnucl::calculateThickness of the moved participants
        SmearSpec ts = thickness_spec(c);
        ts.sparse = 1;
        float* ta = grid_buffer(c, &c->ta);
        float* tb = grid_buffer(c, &c->tb);
        run_smear(c, ts, ta, tb);
        if (want[rp ? kExTaRp : kExTa]) accum(ta, c->buf.counts, 1.0, rp ? kExTaRp : kExTa, o);
        if (want[rp ? kExTbRp : kExTb]) accum(tb, c->buf.counts, 1.0, rp ? kExTbRp : kExTb, o);
        if (want[rp ? kExTatbRp : kExTatb]) accum(ta, c->buf.counts, 1.0, rp ? kExTatbRp : kExTatb, o, tb);
      }
      if (need_rb[pass]) {  // MCnucl::calculate_rho_binary of the moved pairs
        SmearSpec rs{};
        rs.enabled[kSrcPair] = 1;
        rs.win[kSrcPair] = kWin5;
        rs.scale[kSrcPair] = c->cfg.norm;
        rs.sparse = 1;
        float* g = grid_buffer(c, &c->aux0);
        run_smear(c, rs, g, nullptr);
        accum(g, c->buf.counts, 1.0, rp ? kExRhoBinRp : kExRhoBin, o);
      }
      if (!rp && need_sp) {  // MCnucl::calculate_spectator_density of the moved spectators
        SmearSpec ss{};
        for (int k : {kSrcPartA, kSrcPartB}) {
          ss.enabled[k] = 1;
          ss.win[k] = kWin5;
          ss.scale[k] = c->cfg.norm;
        }
        ss.acc[kSrcPartB] = 1;
        ss.combine = 2;
        ss.sparse = 1;
        float* ga = grid_buffer(c, &c->aux0);
        float* gb = grid_buffer(c, &c->aux1);
        check_launch(launch_tile(c->cfg, sb, ss, nev, ga, gb, c->stream), "k_tile");
        if (want[kExSpecA]) accum(ga, sb.counts, 1.0, kExSpecA, o);
        if (want[kExSpecB]) accum(gb, sb.counts, 1.0, kExSpecB, o);
      }
    }
  }
  std::vector<double> h(2 * (size_t)nord * N);
  HIPCHK(hipMemcpyAsync(h.data(), c->avg_sum, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  std::vector<double> hx(any_ex ? (size_t)kExN * nord * N : 0);
  if (any_ex)
    HIPCHK(hipMemcpyAsync(hx.data(), c->avg_ext, hx.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  sync(c);
  for (size_t i = 0; i < (size_t)nord * N; ++i) sums[i] += h[i];
  if (rp_sums)
    for (size_t i = 0; i < (size_t)nord * N; ++i) rp_sums[i] += h[(size_t)nord * N + i];
  for (int k = 0; k < kExN; ++k)
    if (want[k])
      for (size_t i = 0; i < (size_t)nord * N; ++i) want[k][i] += hx[(size_t)k * nord * N + i];
}

// ---------------------------------------------------------------- differentiable batches
void diff_check_params(const smc_ctx* c) {
  const smc_params& p = c->p;
  const smc_derived& d = c->d;
  if (p.which_mc_model != 5 || p.sub_model != 1)
    throw SmcError(SMC_ERR_UNSUPPORTED, "differentiable batches need which_mc_model 5, sub_model 1");
  if (d.criterion != 1) throw SmcError(SMC_ERR_UNSUPPORTED, "differentiable batches need the disk criterion (1)");
  for (int nu = 0; nu < 2; ++nu)
    if (c->cfg.nuc[nu].kind != kWoodsSaxon || c->cfg.nuc[nu].deformed)
      throw SmcError(SMC_ERR_UNSUPPORTED, "differentiable batches need spherical Woods-Saxon nuclei");
  if (d.ws_R[0] != d.ws_R[1] || d.ws_a[0] != d.ws_a[1])
    throw SmcError(SMC_ERR_UNSUPPORTED, "differentiable batches share R and a between the nuclei");
  if (!(p.cc_fluctuation_model == 0 || p.cc_fluctuation_model == 6))
    throw SmcError(SMC_ERR_UNSUPPORTED, "cc_fluctuation_model must be 0 or 6");
  // k_diff_moments holds the whole hot-spot mask of an event in LDS
  if ((size_t)d.nx * ((d.ny + 31) / 32) * sizeof(unsigned) > 65536)
    throw SmcError(SMC_ERR_UNSUPPORTED, "differentiable batches need nx * ceil(ny / 32) <= 16384 (the LDS mask)");
}

template <class T>
T* dif_alloc(smc_ctx* c, int slot, size_t n) {
  if (c->dif_alloc[slot]) (void)hipFree(c->dif_alloc[slot]);
  T* ptr = dalloc<T>(n);
  c->dif_alloc[slot] = ptr;
  return ptr;
}

void diff_setup(smc_ctx* c, uint64_t seed, uint64_t event0, int nev) {
  diff_check_params(c);
  const size_t E = (size_t)nev, SC = (size_t)c->buf.src_cap;
  const int nwy = (c->d.ny + 31) / 32;
  DiffBufs& d = c->dif;
  d.pending = dif_alloc<int>(c, 0, E);
  d.attempt = dif_alloc<int>(c, 1, E);
  d.topo_ok = dif_alloc<int>(c, 2, E);
  d.bval = dif_alloc<double>(c, 3, E);
  d.du = dif_alloc<double>(c, 4, E * 2 * kAMax * 4);
  d.dpos = dif_alloc<double>(c, 5, E * 2 * kAMax * 6);
  d.dsrc = dif_alloc<double2>(c, 6, E * SC);
  d.gam_v = dif_alloc<double>(c, 7, E * SC);
  d.dsc = dif_alloc<double>(c, 8, E * kDsStride);
  d.mask = dif_alloc<unsigned>(c, 9, E * c->d.nx * nwy);
  d.gsrc = dif_alloc<double4>(c, 10, E * SC);
  d.gev = dif_alloc<double>(c, 11, E * 5);
  d.wstab = dif_alloc<double>(c, 13, kWsTabLen);
  d.cotc = dif_alloc<double>(c, 14, E * kCotStride);
  std::vector<double> x, w;
  gauss_legendre01(64, x, w);
  std::vector<double> glh(128);
  for (int i = 0; i < 64; ++i) {
    glh[2 * i] = x[i];
    glh[2 * i + 1] = w[i];
  }
  double* gl = dif_alloc<double>(c, 12, 128);
  HIPCHK(hipMemcpy(gl, glh.data(), glh.size() * sizeof(double), hipMemcpyHostToDevice));
  d.gl = reinterpret_cast<const double2*>(gl);
  d.ngl = 64;
  if (c->dif_out) (void)hipFree(c->dif_out);
  if (c->dif_cot) (void)hipFree(c->dif_cot);
  c->dif_out = dalloc<double>(E * kDiffOut);
  c->dif_cot = dalloc<double>(E * 3);
  HIPCHK(hipMemsetAsync(d.gam_v, 0, E * SC * sizeof(double), c->stream));
  // theta0 from the context's parameters
  c->dif_theta0[0] = c->d.width;
  c->dif_theta0[1] = c->d.ws_R[0];
  c->dif_theta0[2] = c->d.ws_a[0];
  c->dif_theta0[3] = c->p.alpha;
  c->dif_theta0[4] = 1.0 / c->p.cc_fluctuation_Gamma_theta;
  // draw topologies until every event passes MakeDensity.cpp:2147-2162
  std::vector<int> pend(E, 1), att(E, 0), cnt(E * 8), tok(E);
  int left = nev;
  for (int round = 0; left > 0; ++round) {
    if (round >= 200) throw SmcError(SMC_ERR_NO_EVENT, "differentiable batch: events not accepted after 200 rounds");
    HIPCHK(hipMemcpyAsync(d.pending, pend.data(), E * sizeof(int), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(d.attempt, att.data(), E * sizeof(int), hipMemcpyHostToDevice, c->stream));
    check_launch(launch_diff_topo(c->cfg, c->buf, d, nev, seed, event0, c->dif_theta0[1], c->dif_theta0[2],
                                  c->stream),
                 "k_diff_topo");
    check_launch(launch_event(c->cfg, c->buf, nev, seed, event0, 1, c->stream), "k_event");
    HIPCHK(hipMemcpyAsync(cnt.data(), c->buf.counts, cnt.size() * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(tok.data(), d.topo_ok, E * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    sync(c);
    left = 0;
    for (size_t e = 0; e < E; ++e) {
      if (!pend[e]) continue;
      const int np = cnt[e * 8] + cnt[e * 8 + 1], nc = cnt[e * 8 + 2], st = cnt[e * 8 + 3];
      const bool ok = tok[e] && st == 0 && nc > 0 && np >= c->p.Npmin && np <= c->p.Npmax;
      if (ok) {
        pend[e] = 0;
      } else {
        ++att[e];
        ++left;
      }
    }
  }
  // the Gamma uniforms (and the theta0 sources) are part of the topology
  check_launch(launch_diff_forward_pre(c->cfg, c->buf, d, nev, seed, event0, c->dif_theta0,
                                       c->p.cc_fluctuation_model == 6, c->stream),
               "k_diff_pos/src");
  sync(c);
  c->nev = nev;
  c->have_events = 1;
  c->have_density = c->have_thickness = c->have_moments = 0;
  c->dif_nev = nev;
  c->dif_seed = seed;
  c->dif_event0 = event0;
  c->dif_fwd = 0;
}

// The grid geometry of the context with the Gaussian at width th[0]; the
// footprint is the frozen integer window alone (no disk cut), so the
// density is smooth in theta.
DevCfg diff_cfg(const smc_ctx* c, const double* th) {
  DevCfg cw = c->cfg;
  cw.inv2w2 = (float)(1.0 / (2.0 * th[0] * th[0]));
  cw.norm = (float)(1.0 / (2.0 * M_PI * th[0] * th[0]));
  cw.cut2 = 1e30f;
  set_tile_scales(cw);
  return cw;
}

SmearSpec diff_spec(const smc_ctx* c, const double* th) {
  SmearSpec s = density_spec(c);
  const double norm = 1.0 / (2.0 * M_PI * th[0] * th[0]);
  s.scale[kSrcPartA] = s.scale[kSrcPartB] = (float)(0.5 * (1.0 - th[3]) * norm);
  s.scale[kSrcPair] = (float)(th[3] * norm);
  s.enabled[kSrcPair] = th[3] > 1e-8;
  s.use_fluct[kSrcPartA] = s.use_fluct[kSrcPartB] = s.use_fluct[kSrcPair] = 1;
  return s;
}

void diff_forward(smc_ctx* c, const double* th) {
  if (!c->dif_nev) throw SmcError(SMC_ERR_STATE, "call smc_diff_setup first");
  if (!(th[0] > 0.0 && th[1] > 0.0 && th[2] > 0.0 && th[3] >= 0.0 && th[3] < 1.0 && th[4] > 0.0))
    throw SmcError(SMC_ERR_ARG, "theta outside w, R, a, k > 0, 0 <= alpha < 1");
  const int nev = c->dif_nev;
  std::copy(th, th + 5, c->dif_theta);
  const int fluct = c->p.cc_fluctuation_model == 6;
  TimerScope t0(c, SMC_K_EVENT);
  check_launch(launch_diff_forward_pre(c->cfg, c->buf, c->dif, nev, c->dif_seed, c->dif_event0, th, fluct, c->stream),
               "k_diff_pos/src");
  t0.done();
  float* grid = grid_buffer(c, &c->grid);
  TimerScope t1(c, SMC_K_SMEAR);
  SmearSpec spec = diff_spec(c, th);
  spec.sparse = 1;  // k_diff_moments reads each event's source rectangle only
  check_launch(launch_tile(diff_cfg(c, th), c->buf, spec, nev, grid, nullptr, c->stream), "k_tile");
  c->cm_nu = 0;
  c->grid_sparse = 1;
  t1.done();
  TimerScope t2(c, SMC_K_MOMENTS);
  const double ffdxdy = c->p.finalFactor * c->p.dx * c->p.dy;
  check_launch(launch_diff_moments(c->cfg, c->buf, c->dif, grid, nev, ffdxdy, c->dif_out, c->stream),
               "k_diff_moments");
  t2.done();
  c->have_density = 1;
  c->dif_fwd = 1;
}

void diff_backward(smc_ctx* c, const double* cot, double* grad) {
  if (!c->dif_fwd) throw SmcError(SMC_ERR_STATE, "call smc_diff_forward first");
  const int nev = c->dif_nev;
  HIPCHK(hipMemcpyAsync(c->dif_cot, cot, (size_t)nev * 3 * sizeof(double), hipMemcpyHostToDevice, c->stream));
  float* G = grid_buffer(c, &c->tmp);
  const double ffdxdy = c->p.finalFactor * c->p.dx * c->p.dy;
  TimerScope ts(c, SMC_K_ADJOINT);
  check_launch(launch_diff_backward(diff_cfg(c, c->dif_theta), c->buf, diff_spec(c, c->dif_theta), c->dif, nev,
                                    c->dif_cot, ffdxdy, G,
                                    c->dif_theta, c->stream),
               "k_diff_cot/gather/chain");
  ts.done();
  std::vector<double> gev((size_t)nev * 5);
  HIPCHK(hipMemcpyAsync(gev.data(), c->dif.gev, gev.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  sync(c);
  for (int q = 0; q < 5; ++q) grad[q] = 0.0;
  for (int e = 0; e < nev; ++e)
    for (int q = 0; q < 5; ++q) grad[q] += gev[(size_t)e * 5 + q];
}

}  // namespace

extern "C" {

smc_status smc_params_default(smc_params* p) {
  if (!p) return SMC_ERR_ARG;
  std::memset(p, 0, sizeof(*p));
  // parameters.dat of the reference
  p->which_mc_model = 7;
  p->sub_model = 1;
  p->lambda = 0.138;
  p->tmax = 71;
  p->tmax_subdivision = 3;
  p->alpha = 0.118;
  p->Aproj = 208;
  p->Atarg = 208;
  p->proj_deformed = 0;
  p->targ_deformed = 0;
  p->shape_of_nucleons = 2;
  p->gaussian_lambda = 4.14;
  p->collision_criterion = 2;
  p->shape_of_entropy = 2;
  p->quark_width = 0.3;
  p->gauss_nucl_width = 0.812;
  p->ecm = 5020.;
  p->bmin = 0.;
  p->bmax = 20.;
  p->Npmin = 2;
  p->Npmax = 500;
  p->cc_fluctuation_model = 6;
  p->cc_fluctuation_Gamma_theta = 0.75;
  p->maxx = 15.;
  p->maxy = 15.;
  p->dx = 0.1;
  p->dy = 0.1;
  p->finalFactor = 40.0;
  p->ecc_from_order = 1;
  p->ecc_to_order = 9;
  // extensions: reference behaviour
  p->sigma_nn_mb = -1;
  p->proj_ws_R = p->proj_ws_a = p->targ_ws_R = p->targ_ws_a = -1;
  p->proj_beta2 = p->proj_beta4 = p->targ_beta2 = p->targ_beta4 = NAN;
  p->nucleon_min_dist = -1;
  p->kln_n_pt = 32;
  p->kln_n_kt = 32;
  p->kln_n_phi = 48;
  p->kln_pt_min = 0.1;
  p->kln_pt_max = 12.0;
  p->kln_direct = 0;
  p->max_events = 1024;
  p->max_pairs = 0;
  p->max_attempts = 10000;
  p->smear_path = 0;
  return SMC_OK;
}

size_t smc_params_sizeof(void) { return sizeof(smc_params); }
size_t smc_derived_sizeof(void) { return sizeof(smc_derived); }

// the event kernel instance: the second sampling slot only when it is on
// and some system of the context has a proton or deuteron projectile on a
// Woods-Saxon target
static void update_spec(smc_ctx* c) {
  bool light = c->cfg.nuc[0].kind != kWoodsSaxon && c->cfg.nuc[1].kind == kWoodsSaxon;
  for (size_t i = 0; i + 1 < c->mix_sys.size(); i += 2)
    light = light || (c->mix_sys[i].kind != kWoodsSaxon && c->mix_sys[i + 1].kind == kWoodsSaxon);
  c->cfg.spec_live = c->cfg.spec_attempts && light;
}

smc_status smc_create(smc_ctx** out, const smc_params* p) {
  if (!out || !p) return SMC_ERR_ARG;
  *out = nullptr;
  smc_ctx* c = new smc_ctx();
  try {
    c->p = *p;
    c->cap = p->max_events > 0 ? p->max_events : 1024;
    validate_and_derive(c);
    update_spec(c);
    HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    build_tables(c);
    build_kln(c);
    alloc_buffers(c);
    *out = c;
    return SMC_OK;
  } catch (const SmcError& e) {
    static thread_local std::string last;
    last = e.what();
    std::fprintf(stderr, "smc_create: %s\n", e.what());
    smc_destroy(c);
    return e.code;
  }
}

void smc_destroy(smc_ctx* c) {
  if (!c) return;
  if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  if (c->d_mix_sys) (void)hipFree(c->d_mix_sys);
  if (c->d_mix_of) (void)hipFree(c->d_mix_of);
  for (auto& pe : c->pending) { (void)hipEventDestroy(pe.a); (void)hipEventDestroy(pe.b); }
  void* ptrs[] = {c->buf.pos, c->buf.fluct, c->buf.wound, c->buf.pairs, c->buf.pfluct, c->buf.paw, c->buf.counts,
                  c->buf.bval, c->buf.raw, c->buf.src, c->buf.swin, c->dtab, c->grid, c->ta, c->tb, c->tmp, c->evout,
                  c->cm, c->band, c->mrects, c->nmrect, c->avg_rot, c->avg_acc, c->avg_sum, c->avg_part, c->avg_ext, c->aux0,
                  c->aux1, c->grid64, c->spec_src, c->spec_swin, c->spec_counts, c->kln_table, c->kln_ptf, c->kln_nodes_f, c->kln_ptd, c->kln_nodes_d, c->kln_list};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  for (void* q : c->dif_alloc)
    if (q) (void)hipFree(q);
  if (c->dif_out) (void)hipFree(c->dif_out);
  if (c->dif_cot) (void)hipFree(c->dif_cot);
  if (c->evout_alt) (void)hipFree(c->evout_alt);
  for (int i = 0; i < 2; ++i) {
    if (c->stage[i]) (void)hipHostFree(c->stage[i]);
    if (c->ev_done[i]) (void)hipEventDestroy(c->ev_done[i]);
    if (c->ev_copied[i]) (void)hipEventDestroy(c->ev_copied[i]);
  }
  if (c->copy_stream) (void)hipStreamDestroy(c->copy_stream);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

const char* smc_last_error(const smc_ctx* c) { return c ? c->err.c_str() : "null ctx"; }

smc_status smc_get_derived(const smc_ctx* c, smc_derived* out) {
  if (!c || !out) return SMC_ERR_ARG;
  *out = c->d;
  return SMC_OK;
}

smc_status smc_sample(smc_ctx* c, uint64_t seed, uint64_t event0, int nev) {
  if (!c) return SMC_ERR_ARG;
  try {
    if (nev <= 0 || nev > c->cap) throw SmcError(SMC_ERR_CAPACITY, "nev outside [1, max_events]");
    run_event(c, nev, seed, event0, 0);
    return SMC_OK;
  } catch (const SmcError& e) { return fail(c, e); }
}

smc_status smc_sample_b(smc_ctx* c, uint64_t seed, uint64_t event0, int nev, const double* b) {
  if (!c) return SMC_ERR_ARG;
  try {
    if (nev <= 0 || nev > c->cap) throw SmcError(SMC_ERR_CAPACITY, "nev outside [1, max_events]");
    if (!b) throw SmcError(SMC_ERR_ARG, "null impact parameters");
    for (int e = 0; e < nev; ++e)
      if (!(b[e] >= 0.0 && b[e] < 2 * (56.0 - 12.0))) throw SmcError(SMC_ERR_ARG, "impact parameter out of range");
    HIPCHK(hipMemcpyAsync(c->buf.bval, b, (size_t)nev * sizeof(double), hipMemcpyHostToDevice, c->stream));
    run_event(c, nev, seed, event0, 2);
    return SMC_OK;
  } catch (const SmcError& e) { return fail(c, e); }
}

smc_status smc_set_positions(smc_ctx* c, int nev, const double* xy_proj, const double* xy_targ,
                             const double* b, uint64_t fluct_seed) {
  if (!c) return SMC_ERR_ARG;
  try {
    if (nev <= 0 || nev > c->cap) throw SmcError(SMC_ERR_CAPACITY, "nev outside [1, max_events]");
    if (!xy_proj || !xy_targ) throw SmcError(SMC_ERR_ARG, "null positions");
    const int A = c->p.Aproj, B = c->p.Atarg;
    std::vector<int> raw((size_t)nev * 2 * kAMax * 2, 0);
    std::vector<double> bv((size_t)nev, 0.0);
    for (int e = 0; e < nev; ++e) {
      for (int nu = 0; nu < 2; ++nu) {
        const int n = nu ? B : A;
        const double* src = (nu ? xy_targ : xy_proj) + (size_t)e * n * 2;
        for (int i = 0; i < 2 * n; ++i) {
          const double v = src[i];
          if (!(std::fabs(v) < 56.0)) throw SmcError(SMC_ERR_ARG, "nucleon position beyond +-56 fm");
          raw[((size_t)e * 2 + nu) * kAMax * 2 + i] = (int)std::llround(v / kQ);
        }
      }
      bv[e] = b ? b[e] : 0.0;
    }
    HIPCHK(hipMemcpyAsync(c->buf.raw, raw.data(), raw.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->buf.bval, bv.data(), bv.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    run_event(c, nev, fluct_seed, 0, 1);
    return SMC_OK;
  } catch (const SmcError& e) { return fail(c, e); }
}

smc_status smc_collide(smc_ctx* c, int* npart_proj, int* npart_targ, int* ncoll) {
  if (!c) return SMC_ERR_ARG;
  try {
    if (!c->have_events) throw SmcError(SMC_ERR_STATE, "no events");
    std::vector<int> cnt((size_t)c->nev * 8);
    HIPCHK(hipMemcpyAsync(cnt.data(), c->buf.counts, cnt.size() * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    sync(c);
    int st = SMC_OK;
    for (int e = 0; e < c->nev; ++e) {
      if (npart_proj) npart_proj[e] = cnt[(size_t)e * 8 + 0];
      if (npart_targ) npart_targ[e] = cnt[(size_t)e * 8 + 1];
      if (ncoll) ncoll[e] = cnt[(size_t)e * 8 + 2];
      const int s = cnt[(size_t)e * 8 + 3];
      if (s & kStNoEvent) st = SMC_ERR_NO_EVENT;
      else if ((s & kStPairOverflow) && st == SMC_OK) st = SMC_ERR_CAPACITY;
    }
    if (st == SMC_ERR_NO_EVENT) c->err = "no accepted event within max_attempts";
    if (st == SMC_ERR_CAPACITY) c->err = "collision pairs exceed max_pairs";
    return st;
  } catch (const SmcError& e) { return fail(c, e); }
}

smc_status smc_get_nucleons(smc_ctx* c, int ev, double* xy_proj, int* wound_proj, float* fluct_proj,
                            double* xy_targ, int* wound_targ, float* fluct_targ) {
  if (!c) return SMC_ERR_ARG;
  try {
    if (!c->have_events || ev < 0 || ev >= c->nev) throw SmcError(SMC_ERR_ARG, "event index out of range");
    std::vector<int> pos(2 * kAMax * 2), wd(2 * kAMax);
    std::vector<float> fl(2 * kAMax);
    const size_t base = (size_t)ev * 2 * kAMax;
    HIPCHK(hipMemcpyAsync(pos.data(), c->buf.pos + base * 2, pos.size() * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(wd.data(), c->buf.wound + base, wd.size() * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(fl.data(), c->buf.fluct + base, fl.size() * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    sync(c);
    // a mixed batch's event has its system's nucleon counts (at most the
    // context's); the entries past them read NaN / -2 / 0
    const int sys = event_system(c, ev);
    for (int nu = 0; nu < 2; ++nu) {
      const int cap = nu ? c->p.Atarg : c->p.Aproj;
      const int n = sys >= 0 ? c->mix_sys[2 * sys + nu].A : cap;
      double* xy = nu ? xy_targ : xy_proj;
      int* w = nu ? wound_targ : wound_proj;
      float* f = nu ? fluct_targ : fluct_proj;
      for (int i = 0; i < cap; ++i) {
        const size_t k = (size_t)nu * kAMax + i;
        const bool in = i < n;
        if (xy) {
          xy[2 * i] = in ? pos[2 * k] * kQ : std::nan("");
          xy[2 * i + 1] = in ? pos[2 * k + 1] * kQ : std::nan("");
        }
        if (w) w[i] = in ? wd[k] : -2;
        if (f) f[i] = in ? fl[k] : 0.f;
      }
    }
    return SMC_OK;
  } catch (const SmcError& e) { return fail(c, e); }
}

smc_status smc_get_pairs(smc_ctx* c, int ev, int cap, int* pair_proj, int* pair_targ, float* fluct,
                         int* npairs) {
  if (!c) return SMC_ERR_ARG;
  try {
    if (!c->have_events || ev < 0 || ev >= c->nev) throw SmcError(SMC_ERR_ARG, "event index out of range");
    int cnt[8];
    HIPCHK(hipMemcpyAsync(cnt, c->buf.counts + (size_t)ev * 8, sizeof(cnt), hipMemcpyDeviceToHost, c->stream));
    sync(c);
    const int n = std::min(cnt[2], c->cfg.max_pairs);
    if (npairs) *npairs = cnt[2];
    const int m = std::min(n, cap);
    if (m > 0) {
      std::vector<unsigned> pr(m);
      std::vector<float> pf(m);
      const size_t base = (size_t)ev * c->cfg.max_pairs;
      HIPCHK(hipMemcpyAsync(pr.data(), c->buf.pairs + base, m * sizeof(unsigned), hipMemcpyDeviceToHost, c->stream));
      HIPCHK(hipMemcpyAsync(pf.data(), c->buf.pfluct + base, m * sizeof(float), hipMemcpyDeviceToHost, c->stream));
      sync(c);
      for (int k = 0; k < m; ++k) {
        if (pair_proj) pair_proj[k] = (int)(pr[k] & 0xFFFFu);
        if (pair_targ) pair_targ[k] = (int)(pr[k] >> 16);
        if (fluct) fluct[k] = pf[k];
      }
    }
    if (cnt[2] > cap) { c->err = "pair buffer too small"; return SMC_ERR_CAPACITY; }
    return SMC_OK;
  } catch (const SmcError& e) { return fail(c, e); }
}

// The event's source list: participants of the projectile and of the target
// in wound order, then the collision pairs in reference order (the order of
// getHotSpots, MCnucl.cpp:1303-1323) at their current positions.
smc_status smc_get_sources(smc_ctx* c, int ev, int cap, double* xy, int* kind, float* weight, int* nsrc) {
  if (!c) return SMC_ERR_ARG;
  try {
    if (!c->have_events || ev < 0 || ev >= c->nev) throw SmcError(SMC_ERR_ARG, "event index out of range");
    int cnt[8];
    HIPCHK(hipMemcpyAsync(cnt, c->buf.counts + (size_t)ev * 8, sizeof(cnt), hipMemcpyDeviceToHost, c->stream));
    sync(c);
    const int K = cnt[5];
    if (nsrc) *nsrc = K;
    const int m = std::min(K, std::max(cap, 0));
    if (m > 0) {
      std::vector<int4> sv(m);
      HIPCHK(hipMemcpyAsync(sv.data(), c->buf.src + (size_t)ev * c->buf.src_cap, m * sizeof(int4),
                            hipMemcpyDeviceToHost, c->stream));
      sync(c);
      for (int k = 0; k < m; ++k) {
        if (xy) {
          xy[2 * k] = sv[k].x * kH;
          xy[2 * k + 1] = sv[k].y * kH;
        }
        if (kind) kind[k] = sv[k].w;
        if (weight) {
          float f;
          std::memcpy(&f, &sv[k].z, sizeof f);
          weight[k] = f;
        }
      }
    }
    if (K > cap) {
      c->err = "source buffer too small";
      return SMC_ERR_CAPACITY;
    }
    return SMC_OK;
  } catch (const SmcError& e) { return fail(c, e); }
}

// MCnucl::recenterGrid / rotateGrid (MCnucl.cpp:1119-1174) for every event
// of the batch: the centre of mass of the current density and, with turn,
// the order-n angle -atan2(-Im, -Re) / n of sum r^n e^{in theta} rho
// (GlueDensity::calcCMAngle, GlueDensity.cpp:87-144); the sources move by
// -centre and turn by the angle.  The next density call smears the moved
// sources.  rot: host [nev][4] {xc, yc, angle, sum rho} (NULL allowed).
smc_status smc_recenter(smc_ctx* c, int n, int turn, double* rot) {
  if (!c) return SMC_ERR_ARG;
  try {
    if (!c->have_events) throw SmcError(SMC_ERR_STATE, "no events: call smc_sample or smc_run first");
    if (n < 1 || n > 9) throw SmcError(SMC_ERR_ARG, "order must satisfy 1 <= n <= 9");
    if (!c->avg_rot) {
      c->avg_rot = dalloc<double>((size_t)c->cap * 4);
      c->avg_acc = dalloc<unsigned char>((size_t)c->cap);
      c->avg_part = dalloc<double>((size_t)kAccumSegments * c->d.nx * c->d.ny);
    }
    if (!c->have_density) run_density(c);
    ensure_cm(c);
    check_launch(launch_cm_angle(c->cfg, c->buf, c->grid, c->cm, c->cm_nu, n, turn ? 1 : 0, c->nev, c->avg_rot,
                                 c->stream),
                 "k_cm_angle");
    check_launch(launch_rotate_src(c->cfg, c->buf, c->avg_rot, c->nev, c->stream), "k_rotate_src");
    c->have_density = c->have_thickness = c->have_moments = 0;
    c->cm_nu = 0;
    if (rot)
      HIPCHK(hipMemcpyAsync(rot, c->avg_rot, (size_t)c->nev * 4 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    sync(c);
    return SMC_OK;
  } catch (const SmcError& e) { return fail(c, e); }
}

smc_status smc_thickness(smc_ctx* c, float* TA, float* TB) {
  if (!c) return SMC_ERR_ARG;
  try {
    if (!c->have_events) throw SmcError(SMC_ERR_STATE, "no events");
    float* a = grid_buffer(c, &c->ta);
    float* b = grid_buffer(c, &c->tb);
    run_smear(c, thickness_spec(c), a, b);
    c->have_thickness = 1;
    const size_t n = (size_t)c->nev * c->d.nx * c->d.ny;
    if (TA) HIPCHK(hipMemcpyAsync(TA, a, n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    if (TB) HIPCHK(hipMemcpyAsync(TB, b, n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    sync(c);
    return SMC_OK;
  } catch (const SmcError& e) { return fail(c, e); }
}

smc_status smc_density(smc_ctx* c, float* s, double* dndy) {
  if (!c) return SMC_ERR_ARG;
  try {
    run_density(c);
    const size_t n = (size_t)c->nev * c->d.nx * c->d.ny;
    if (s) {
      densify_grid(c);
      HIPCHK(hipMemcpyAsync(s, c->grid, n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    }
    int st = SMC_OK;
    if (dndy) {
      run_moments(c, 1);
      std::vector<double> ev((size_t)c->nev * SMC_EV_STRIDE);
      HIPCHK(hipMemcpyAsync(ev.data(), c->evout, ev.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
      sync(c);
      for (int e = 0; e < c->nev; ++e) {
        dndy[e] = ev[(size_t)e * SMC_EV_STRIDE + SMC_EV_TOTAL];
        if (dndy[e] < 1e-15) st = SMC_ERR_DNDY_ZERO;  // MCnucl.cpp:812-816
      }
      if (st) c->err = "dndy = 0 for an event";
    }
    sync(c);
    return st;
  } catch (const SmcError& e) { return fail(c, e); }
}

smc_status smc_density_f64(smc_ctx* c, int ev0, int n, double* s) {
  if (!c || !s || ev0 < 0 || n < 0) return SMC_ERR_ARG;
  try {
    if (!c->have_events) throw SmcError(SMC_ERR_STATE, "no events: call smc_sample or smc_set_positions first");
    if (n > c->nev - ev0) throw SmcError(SMC_ERR_ARG, "event range beyond the current batch");
    if (c->p.which_mc_model != 5 && c->p.which_mc_model != 7)
      throw SmcError(SMC_ERR_UNSUPPORTED, "FP64 grids are built for which_mc_model 5 and 7");
    if (n == 0) return SMC_OK;
    // the reference's constants, computed as MCnucl.cpp / Particle.cpp write them
    const double w = c->d.width, wsq = w * w;
    Grid64Const gc{};
    gc.norm = 1 / (2 * M_PI * w * w);
    gc.two_w_w = 2 * w * w;
    gc.two_wsq = 2 * wsq;
    gc.r_max = 5 * w;
    gc.dc_max = 25. * wsq;
    gc.alpha = c->p.alpha;
    gc.prefactor = (1.0 - c->p.alpha) / 2.;
    gc.model = c->p.which_mc_model;
    gc.use_binary = c->p.which_mc_model == 5 && c->p.alpha > 1e-8;
    gc.uli = c->cfg.uli;
    gc.fluct = c->cfg.fluct;
    const size_t N = (size_t)c->d.nx * c->d.ny;
    // events per launch: at most kGrid64Chunk and about 1 GiB of device grids
    const int fit = (int)std::max<size_t>(1, ((size_t)1 << 30) / (N * sizeof(double)));
    const int chunk = std::min(n, std::min(kGrid64Chunk, fit));
    if (c->grid64_cap < chunk) {
      if (c->grid64) HIPCHK(hipFree(c->grid64));
      c->grid64 = nullptr;
      c->grid64_cap = 0;
      c->grid64 = dalloc<double>((size_t)chunk * N);
      c->grid64_cap = chunk;
    }
    for (int e = 0; e < n; e += chunk) {
      const int m = std::min(chunk, n - e);
      TimerScope ts(c, SMC_K_GRID64);
      check_launch(launch_grid64(c->cfg, c->buf, gc, ev0 + e, m, c->grid64, c->stream), "k_grid64");
      ts.done();
      HIPCHK(hipMemcpyAsync(s + (size_t)e * N, c->grid64, (size_t)m * N * sizeof(double), hipMemcpyDeviceToHost,
                            c->stream));
    }
    sync(c);
    return SMC_OK;
  } catch (const SmcError& e) { return fail(c, e); }
}

smc_status smc_moments(smc_ctx* c, double* ev_out) {
  if (!c) return SMC_ERR_ARG;
  try {
    if (!c->have_events) throw SmcError(SMC_ERR_STATE, "no events");
    run_moments(c, 0);
    c->have_moments = 1;
    if (ev_out)
      HIPCHK(hipMemcpyAsync(ev_out, c->evout, (size_t)c->nev * SMC_EV_STRIDE * sizeof(double),
                            hipMemcpyDeviceToHost, c->stream));
    sync(c);
    tag_systems(c, ev_out, c->nev);
    return SMC_OK;
  } catch (const SmcError& e) { return fail(c, e); }
}

smc_status smc_rho_binary(smc_ctx* c, float* rho) {
  if (!c) return SMC_ERR_ARG;
  try {
    if (!c->have_events) throw SmcError(SMC_ERR_STATE, "no events");
    float* g = grid_buffer(c, &c->aux0);
    run_rho_binary(c, g);
    if (rho)
      HIPCHK(hipMemcpyAsync(rho, g, (size_t)c->nev * c->d.nx * c->d.ny * sizeof(float), hipMemcpyDeviceToHost,
                            c->stream));
    sync(c);
    return SMC_OK;
  } catch (const SmcError& e) { return fail(c, e); }
}

smc_status smc_spectator_density(smc_ctx* c, float* spec_proj, float* spec_targ) {
  if (!c) return SMC_ERR_ARG;
  try {
    if (!c->have_events) throw SmcError(SMC_ERR_STATE, "no events");
    float* a = grid_buffer(c, &c->aux0);
    float* b = grid_buffer(c, &c->aux1);
    run_spectator_density(c, a, b);
    const size_t n = (size_t)c->nev * c->d.nx * c->d.ny * sizeof(float);
    if (spec_proj) HIPCHK(hipMemcpyAsync(spec_proj, a, n, hipMemcpyDeviceToHost, c->stream));
    if (spec_targ) HIPCHK(hipMemcpyAsync(spec_targ, b, n, hipMemcpyDeviceToHost, c->stream));
    sync(c);
    return SMC_OK;
  } catch (const SmcError& e) { return fail(c, e); }
}

smc_status smc_average_profiles(smc_ctx* c, int from_order, int to_order, const unsigned char* accept, double* sums,
                                double* rp_sums) {
  if (!c) return SMC_ERR_ARG;
  try {
    average_profiles(c, from_order, to_order, accept, sums, rp_sums, nullptr);
    return SMC_OK;
  } catch (const SmcError& e) { return fail(c, e); }
}

smc_status smc_average_profiles_extra(smc_ctx* c, int from_order, int to_order, const unsigned char* accept,
                                      double* sums, double* rp_sums, const smc_avg_extra* extra) {
  if (!c) return SMC_ERR_ARG;
  try {
    average_profiles(c, from_order, to_order, accept, sums, rp_sums, extra);
    return SMC_OK;
  } catch (const SmcError& e) { return fail(c, e); }
}

smc_status smc_run(smc_ctx* c, uint64_t seed, uint64_t event0, int nev, double* ev_out) {
  if (!c) return SMC_ERR_ARG;
  try {
    if (nev <= 0 || nev > c->cap) throw SmcError(SMC_ERR_CAPACITY, "nev outside [1, max_events]");
    run_event(c, nev, seed, event0, 0);
    run_moments(c, 0);
    c->have_moments = 1;
    if (ev_out) {
      HIPCHK(hipMemcpyAsync(ev_out, c->evout, (size_t)nev * SMC_EV_STRIDE * sizeof(double),
                            hipMemcpyDeviceToHost, c->stream));
      sync(c);
      tag_systems(c, ev_out, nev);
    }
    return SMC_OK;
  } catch (const SmcError& e) { return fail(c, e); }
}

namespace {

void graph_release(smc_ctx* c) {
  if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  c->graph_exec = nullptr;
  c->graph = nullptr;
  c->graph_event_node = nullptr;
  c->graph_args.clear();
}

// Captures smc_run's launches for nev events on the ctx's stream.  Every
// buffer they touch exists already (the eager first call made them), and
// the sequence has no memset node.
void graph_capture(smc_ctx* c, int nev, uint64_t seed, uint64_t event0) {
  const int timing = c->timing;
  c->timing = 0;
  HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  try {
    c->have_density = 0;
    run_event(c, nev, seed, event0, 0);
    run_moments(c, 0);
  } catch (...) {
    hipGraph_t g = nullptr;
    (void)hipStreamEndCapture(c->stream, &g);
    if (g) (void)hipGraphDestroy(g);
    c->timing = timing;
    throw;
  }
  c->timing = timing;
  HIPCHK(hipStreamEndCapture(c->stream, &c->graph));
  c->graph_cm_nu = c->cm_nu;
  size_t n = 0;
  HIPCHK(hipGraphGetNodes(c->graph, nullptr, &n));
  std::vector<hipGraphNode_t> nodes(n);
  HIPCHK(hipGraphGetNodes(c->graph, nodes.data(), &n));
  for (hipGraphNode_t node : nodes) {
    hipGraphNodeType t;
    HIPCHK(hipGraphNodeGetType(node, &t));
    if (t != hipGraphNodeTypeKernel) continue;
    hipKernelNodeParams kp{};
    HIPCHK(hipGraphKernelNodeGetParams(node, &kp));
    if (kp.func == event_kernel_symbol(c->cfg) && kp.kernelParams) {
      c->graph_event_node = node;
      c->graph_event_params = kp;
      c->graph_args.assign(kp.kernelParams, kp.kernelParams + kEventKernelArgs);
      break;
    }
  }
  if (!c->graph_event_node) throw SmcError(SMC_ERR_HIP, "captured batch has no event kernel node");
  c->graph_args[kEventArgKey] = &c->graph_key;
  c->graph_args[kEventArgEvent0] = &c->graph_event0;
  c->graph_event_params.kernelParams = c->graph_args.data();
  HIPCHK(hipGraphInstantiate(&c->graph_exec, c->graph, nullptr, nullptr, 0));
}

}  // namespace

// smc_run for callers that cannot allocate: the same launches, no timing
// events, the moment kernels' scratch in the static device array of
// smc_noalloc.hip, the records written to the caller's device buffer.  Every
// buffer the batch touches must exist already (an earlier smc_run on the
// context allocated them).  Asynchronous on the context's stream.
smc_status smc_run_noalloc(smc_ctx* c, uint64_t seed, uint64_t event0, int nev, double* ev_out_dev) {
  if (!c) return SMC_ERR_ARG;
  try {
    if (!ev_out_dev) throw SmcError(SMC_ERR_ARG, "ev_out_dev is NULL");
    if (nev <= 0 || nev > c->cap) throw SmcError(SMC_ERR_CAPACITY, "nev outside [1, max_events]");
    if (nev > kNoAllocMaxEvents)
      throw SmcError(SMC_ERR_CAPACITY, "nev above smc_run_noalloc's " + std::to_string(kNoAllocMaxEvents) +
                                           " events per call (the size of its device scratch)");
    if ((size_t)moments_band_doubles() * sizeof(double) * kNoAllocMaxEvents != kNoAllocBandBytes ||
        (size_t)moments_cm_units_min() * 4 * sizeof(double) * kNoAllocMaxEvents != kNoAllocCmBytes)
      throw SmcError(SMC_ERR_UNSUPPORTED, "moment scratch layout does not match the static scratch");
    const bool kln = c->p.which_mc_model == 1;
    if (kln && (c->p.kln_direct || c->p.smear_path == 2))
      throw SmcError(SMC_ERR_UNSUPPORTED, "smc_run_noalloc takes MC-KLN through the dN/dy table in one launch only");
    const SmearSpec spec = density_spec(c);
    const bool need_tmp = !kln && spec.combine && !density_cm_units(c);
    if (!c->grid || (kln && !c->kln_table) || (need_tmp && !c->tmp))
      throw SmcError(SMC_ERR_STATE, "call smc_run once on this context first: it creates the buffers of the batch");
    int rc = 0;
    unsigned char* base = static_cast<unsigned char*>(noalloc_scratch(&rc));
    if (!base) throw SmcError(SMC_ERR_HIP, std::string("scratch lookup: ") + hipGetErrorString((hipError_t)rc));
    double* band = reinterpret_cast<double*>(base);
    void* rects = base + kNoAllocBandBytes;
    double* cms = reinterpret_cast<double*>(base + kNoAllocBandBytes + kNoAllocRectBytes);
    int* nrect = reinterpret_cast<int*>(base + kNoAllocBandBytes + kNoAllocRectBytes + kNoAllocCmBytes);
    check_launch(launch_event(c->cfg, c->buf, nev, seed, event0, 0, c->stream), "k_event");
    c->batch_event0 = event0;
    c->batch_seed = seed;
    c->batch_mode = 0;
    c->nev = nev;
    c->have_events = 1;
    c->have_thickness = 0;
    density_smear_launch(c);
    c->cm_nu = density_cm_units(c);
    c->grid_sparse = c->cm_nu ? 1 : 0;
    c->have_density = 1;
    check_launch(launch_moments(c->cfg, c->buf, c->grid, nev, ev_out_dev, 0, c->cm_nu ? c->cm : nullptr, c->cm_nu, cms,
                                band, rects, nrect, c->stream),
                 "k_moments");
    c->have_moments = 0;  // the context's own record buffer holds an earlier batch
    return SMC_OK;
  } catch (const SmcError& e) { return fail(c, e); }
}

// The whole hot path for nev_total events, in batches of max_events through
// the ctx's device buffers (one grid batch, one source batch: memory is
// bounded whatever nev_total is).  Batch i's records go to pinned staging
// on a second stream while batch i + 1 runs (two record buffers alternate),
// then into ev_out on the host.
smc_status smc_run_stream(smc_ctx* c, uint64_t seed, uint64_t event0, int64_t nev_total, double* ev_out) {
  if (!c) return SMC_ERR_ARG;
  try {
    if (nev_total <= 0) throw SmcError(SMC_ERR_ARG, "nev_total must be positive");
    if (!ev_out) throw SmcError(SMC_ERR_ARG, "ev_out is NULL");
    const int cap = c->cap;
    const size_t rec = (size_t)SMC_EV_STRIDE * sizeof(double);
    if (!c->copy_stream) {
      HIPCHK(hipStreamCreateWithFlags(&c->copy_stream, hipStreamNonBlocking));
      for (int i = 0; i < 2; ++i) {
        HIPCHK(hipHostMalloc((void**)&c->stage[i], (size_t)cap * rec, hipHostMallocDefault));
        HIPCHK(hipEventCreateWithFlags(&c->ev_done[i], hipEventDisableTiming));
        HIPCHK(hipEventCreateWithFlags(&c->ev_copied[i], hipEventDisableTiming));
      }
      c->evout_alt = dalloc<double>((size_t)cap * SMC_EV_STRIDE);
    }
    double* const bufs[2] = {c->evout, c->evout_alt};
    const int64_t nb = (nev_total + cap - 1) / cap;
    int nbatch[2] = {0, 0};
    uint64_t e0batch[2] = {0, 0};
    auto drain = [&](int64_t i) {  // batch i's records from staging to ev_out
      const int k = (int)(i & 1);
      HIPCHK(hipEventSynchronize(c->ev_copied[k]));
      double* dst = ev_out + (size_t)(i * cap) * SMC_EV_STRIDE;
      std::memcpy(dst, c->stage[k], (size_t)nbatch[k] * rec);
      if (!c->mix_of.empty())
        for (int e = 0; e < nbatch[k]; ++e)
          dst[(size_t)e * SMC_EV_STRIDE + SMC_EV_SYSTEM] = c->mix_of[(e0batch[k] + (uint64_t)e) % c->mix_of.size()];
    };
    try {
      for (int64_t i = 0; i < nb; ++i) {
        const int k = (int)(i & 1);
        const int n = (int)std::min<int64_t>(cap, nev_total - i * cap);
        if (i >= 2) drain(i - 2);  // staging k and record buffer k are free again
        c->evout = bufs[k];
        run_event(c, n, seed, event0 + (uint64_t)(i * cap), 0);
        run_moments(c, 0);
        HIPCHK(hipEventRecord(c->ev_done[k], c->stream));
        HIPCHK(hipStreamWaitEvent(c->copy_stream, c->ev_done[k], 0));
        HIPCHK(hipMemcpyAsync(c->stage[k], bufs[k], (size_t)n * rec, hipMemcpyDeviceToHost, c->copy_stream));
        HIPCHK(hipEventRecord(c->ev_copied[k], c->copy_stream));
        nbatch[k] = n;
        e0batch[k] = event0 + (uint64_t)(i * cap);
      }
      for (int64_t i = std::max<int64_t>(0, nb - 2); i < nb; ++i) drain(i);
    } catch (...) {
      c->evout = bufs[0];
      throw;
    }
    // the ctx keeps the last batch; its records stay readable by
    // smc_fetch_events from the context's own record buffer
    const int kl = (int)((nb - 1) & 1);
    if (kl == 1)
      HIPCHK(hipMemcpyAsync(bufs[0], bufs[1], (size_t)nbatch[1] * rec, hipMemcpyDeviceToDevice, c->stream));
    c->evout = bufs[0];
    sync(c);
    c->have_moments = 1;
    graph_release(c);
    c->graph_warm = 0;
    return SMC_OK;
  } catch (const SmcError& e) { return fail(c, e); }
}

// smc_run as a replayed hipGraph.  The first call for a batch size runs
// eagerly (it allocates every buffer the batch uses); the second captures
// the launch sequence; later calls only set the event kernel's seed and
// event0 and launch the graph.  Results equal smc_run's.
smc_status smc_run_graph(smc_ctx* c, uint64_t seed, uint64_t event0, int nev, double* ev_out) {
  if (!c) return SMC_ERR_ARG;
  try {
    if (nev <= 0 || nev > c->cap) throw SmcError(SMC_ERR_CAPACITY, "nev outside [1, max_events]");
    if (c->graph_nev != nev) {
      graph_release(c);
      c->graph_nev = nev;
      c->graph_warm = 0;
    }
    if (!c->graph_warm) {
      run_event(c, nev, seed, event0, 0);
      run_moments(c, 0);
      c->graph_warm = 1;
    } else {
      if (!c->graph_exec) graph_capture(c, nev, seed, event0);
      c->graph_key = make_uint2((unsigned)seed, (unsigned)(seed >> 32));
      c->graph_event0 = event0;
      HIPCHK(hipGraphExecKernelNodeSetParams(c->graph_exec, c->graph_event_node, &c->graph_event_params));
      HIPCHK(hipGraphLaunch(c->graph_exec, c->stream));
      c->batch_event0 = event0;
      c->batch_seed = seed;
      c->batch_mode = 0;
      c->nev = nev;
      c->have_events = c->have_density = 1;
      c->have_thickness = 0;
      c->cm_nu = c->graph_cm_nu;
    }
    c->have_moments = 1;
    if (ev_out) {
      HIPCHK(hipMemcpyAsync(ev_out, c->evout, (size_t)nev * SMC_EV_STRIDE * sizeof(double),
                            hipMemcpyDeviceToHost, c->stream));
      sync(c);
      tag_systems(c, ev_out, nev);
    }
    return SMC_OK;
  } catch (const SmcError& e) { return fail(c, e); }
}

smc_status smc_set_tile_variant(smc_ctx* c, int variant) {
  if (!c) return SMC_ERR_ARG;
  if (variant < 0 || variant >= kTileVariants) {
    c->err = "tile variant out of range";
    return SMC_ERR_ARG;
  }
  c->cfg.tile_variant = variant;
  c->have_density = c->have_moments = 0;  // the centre-of-mass partials follow the tile shape
  graph_release(c);
  c->graph_warm = 0;
  return SMC_OK;
}

smc_status smc_calibrate_fma(double* tflops, double* ms) {
  if (!tflops) return SMC_ERR_ARG;
  const int blocks = 4096, iters = 2000;  // 16 workgroups per CU, about 7 ms at the FMA rate
  float* out = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t a = nullptr, b = nullptr;
  int rc = SMC_OK;
  if (hipMalloc(&out, (size_t)blocks * 256 * sizeof(float)) != hipSuccess || hipStreamCreate(&st) != hipSuccess ||
      hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) {
    rc = SMC_ERR_HIP;
  } else {
    float best = 0.f;
    for (int r = 0; r < 3 && rc == SMC_OK; ++r) {  // the fastest of three launches
      if (hipEventRecord(a, st) != hipSuccess || launch_fma_chain(out, blocks, iters, st) != 0 ||
          hipEventRecord(b, st) != hipSuccess || hipEventSynchronize(b) != hipSuccess) {
        rc = SMC_ERR_HIP;
        break;
      }
      float f = 0.f;
      if (hipEventElapsedTime(&f, a, b) != hipSuccess) rc = SMC_ERR_HIP;
      if (r == 0 || f < best) best = f;
    }
    if (rc == SMC_OK) {
      *tflops = 2.0 * blocks * 256.0 * 128.0 * iters / (best * 1e-3) / 1e12;
      if (ms) *ms = best;
    }
  }
  if (a) (void)hipEventDestroy(a);
  if (b) (void)hipEventDestroy(b);
  if (st) (void)hipStreamDestroy(st);
  if (out) (void)hipFree(out);
  return rc;
}

smc_status smc_set_attempts_per_pass(smc_ctx* c, int n) {
  if (!c) return SMC_ERR_ARG;
  if (n != 1 && n != 2) {
    c->err = "attempts per pass must be 1 or 2";
    return SMC_ERR_ARG;
  }
  c->cfg.spec_attempts = n == 2;
  update_spec(c);
  graph_release(c);  // a captured event node holds the old configuration
  c->graph_warm = 0;
  return SMC_OK;
}

// Times the event kernel on the ctx's current batch (sampled events, its
// seed and first event) with one and with two attempts per pass, reps
// launches each after an untimed one, and keeps the faster.  Both give the
// same events, so the relaunches leave the batch as it was.
smc_status smc_autotune_attempts(smc_ctx* c, int reps, int* best, double* ms) {
  if (!c) return SMC_ERR_ARG;
  try {
    if (!c->have_events || c->batch_mode != 0)
      throw SmcError(SMC_ERR_STATE, "no sampled batch: call smc_sample or smc_run first");
    if (reps < 1) throw SmcError(SMC_ERR_ARG, "reps must be positive");
    hipEvent_t a, b;
    HIPCHK(hipEventCreate(&a));
    HIPCHK(hipEventCreate(&b));
    double t[2];
    for (int v = 0; v < 2; ++v) {
      c->cfg.spec_attempts = v;
      update_spec(c);
      check_launch(launch_event(c->cfg, c->buf, c->nev, c->batch_seed, c->batch_event0, 0, c->stream), "k_event");
      HIPCHK(hipEventRecord(a, c->stream));
      for (int r = 0; r < reps; ++r)
        check_launch(launch_event(c->cfg, c->buf, c->nev, c->batch_seed, c->batch_event0, 0, c->stream), "k_event");
      HIPCHK(hipEventRecord(b, c->stream));
      HIPCHK(hipEventSynchronize(b));
      float f = 0.f;
      HIPCHK(hipEventElapsedTime(&f, a, b));
      t[v] = f / reps;
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    c->cfg.spec_attempts = t[1] < t[0];
    update_spec(c);
    graph_release(c);
    c->graph_warm = 0;
    if (best) *best = c->cfg.spec_attempts ? 2 : 1;
    if (ms) {
      ms[0] = t[0];
      ms[1] = t[1];
    }
    return SMC_OK;
  } catch (const SmcError& e) { return fail(c, e); }
}

smc_status smc_get_attempts_per_pass(const smc_ctx* c, int* n) {
  if (!c || !n) return SMC_ERR_ARG;
  *n = c->cfg.spec_attempts ? 2 : 1;
  return SMC_OK;
}

// Times the density step's smearing launch (density_smear_launch: exactly
// what run_density launches, centre-of-mass partials and the MC-KLN table
// step included) of the ctx's current events in every k_tile shape, reps
// launches each after one untimed launch, and keeps the fastest.  The grids
// do not depend on the shape.  Only the VALU tile path has shapes.
smc_status smc_autotune(smc_ctx* c, int reps, int* best, double* ms) {
  if (!c) return SMC_ERR_ARG;
  try {
    if (!c->have_events) throw SmcError(SMC_ERR_STATE, "no events: call smc_sample or smc_run first");
    if (reps < 1) throw SmcError(SMC_ERR_ARG, "reps must be positive");
    if (c->p.smear_path != 0) throw SmcError(SMC_ERR_UNSUPPORTED, "tile shapes exist on smear_path 0 only");
    if (c->p.which_mc_model == 1 && !c->p.kln_direct) ensure_kln_table(c);
    hipEvent_t a, b;
    HIPCHK(hipEventCreate(&a));
    HIPCHK(hipEventCreate(&b));
    double t[kTileVariants];
    int pick = 0;
    for (int v = 0; v < kTileVariants; ++v) {
      c->cfg.tile_variant = v;
      density_smear_launch(c);
      HIPCHK(hipEventRecord(a, c->stream));
      for (int r = 0; r < reps; ++r) density_smear_launch(c);
      HIPCHK(hipEventRecord(b, c->stream));
      HIPCHK(hipEventSynchronize(b));
      float f = 0.f;
      HIPCHK(hipEventElapsedTime(&f, a, b));
      t[v] = f / reps;
      if (t[v] < t[pick]) pick = v;
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    c->cfg.tile_variant = pick;
    c->have_density = c->have_thickness = c->have_moments = 0;
    graph_release(c);
    c->graph_warm = 0;
    if (best) *best = pick;
    if (ms)
      for (int v = 0; v < kTileVariants; ++v) ms[v] = t[v];
    return SMC_OK;
  } catch (const SmcError& e) { return fail(c, e); }
}

// Per-event collision systems in one batch (config "mixed systems"): each
// entry of systems is a full parameter set of which only the nuclei (Aproj,
// Atarg, WS and deformation keys) may differ from the context's; sampled
// event g uses systems[pattern[g % npattern]].  nsys = 0 restores the
// context's own system.
smc_status smc_set_systems(smc_ctx* c, int nsys, const smc_params* systems, int npattern, const int* pattern) {
  if (!c) return SMC_ERR_ARG;
  try {
    if (nsys < 0 || nsys > 64) throw SmcError(SMC_ERR_ARG, "nsys must be in [0, 64]");
    std::vector<NucSys> sys;
    std::vector<int> of;
    if (nsys > 0) {
      if (!systems || !pattern || npattern < 1) throw SmcError(SMC_ERR_ARG, "systems and a pattern are required");
      for (int i = 0; i < nsys; ++i) {
        smc_ctx t;  // host-side derivation only
        t.p = systems[i];
        t.cap = 1;
        try {
          validate_and_derive(&t);
        } catch (const SmcError& e) {
          throw SmcError(e.code, "system " + std::to_string(i) + ": " + e.what());
        }
        const smc_derived &a = t.d, &b = c->d;
        const smc_params &pa = t.p, &pb = c->p;
        const bool same = a.siginNN == b.siginNN && a.width == b.width && a.dsq == b.dsq && a.dmin == b.dmin &&
                          a.nx == b.nx && a.ny == b.ny && pa.dx == pb.dx && pa.dy == pb.dy && pa.maxx == pb.maxx &&
                          pa.maxy == pb.maxy && pa.which_mc_model == pb.which_mc_model && pa.alpha == pb.alpha &&
                          pa.cc_fluctuation_model == pb.cc_fluctuation_model &&
                          pa.cc_fluctuation_Gamma_theta == pb.cc_fluctuation_Gamma_theta && pa.bmin == pb.bmin &&
                          pa.bmax == pb.bmax && pa.Npmin == pb.Npmin && pa.Npmax == pb.Npmax;
        if (!same) throw SmcError(SMC_ERR_ARG, "system " + std::to_string(i) + " differs beyond its nuclei");
        if (pa.Aproj > pb.Aproj || pa.Atarg > pb.Atarg)
          throw SmcError(SMC_ERR_CAPACITY, "system " + std::to_string(i) + ": nuclei larger than the context's");
        sys.push_back(t.cfg.nuc[0]);
        sys.push_back(t.cfg.nuc[1]);
      }
      for (int k = 0; k < npattern; ++k) {
        if (pattern[k] < 0 || pattern[k] >= nsys) throw SmcError(SMC_ERR_ARG, "pattern entry outside [0, nsys)");
        of.push_back(pattern[k]);
      }
    }
    graph_release(c);
    c->graph_warm = 0;
    if (c->d_mix_sys) (void)hipFree(c->d_mix_sys);
    if (c->d_mix_of) (void)hipFree(c->d_mix_of);
    c->d_mix_sys = nullptr;
    c->d_mix_of = nullptr;
    c->cfg.sys = nullptr;
    c->cfg.sys_of = nullptr;
    c->cfg.sys_n = 0;
    if (nsys > 0) {
      c->d_mix_sys = dalloc<NucSys>(sys.size());
      c->d_mix_of = dalloc<int>(of.size());
      HIPCHK(hipMemcpy(c->d_mix_sys, sys.data(), sys.size() * sizeof(NucSys), hipMemcpyHostToDevice));
      HIPCHK(hipMemcpy(c->d_mix_of, of.data(), of.size() * sizeof(int), hipMemcpyHostToDevice));
      c->cfg.sys = c->d_mix_sys;
      c->cfg.sys_of = c->d_mix_of;
      c->cfg.sys_n = (int)of.size();
    }
    c->mix_sys = sys;
    c->mix_of = of;
    update_spec(c);
    c->have_events = c->have_density = c->have_thickness = c->have_moments = 0;
    return SMC_OK;
  } catch (const SmcError& e) { return fail(c, e); }
}

smc_status smc_get_event_system(smc_ctx* c, int ev, int* system) {
  if (!c || !system) return SMC_ERR_ARG;
  if (!c->have_events || ev < 0 || ev >= c->nev) return SMC_ERR_ARG;
  *system = event_system(c, ev);
  return SMC_OK;
}

smc_status smc_fetch_events(smc_ctx* c, int nev, double* ev_out) {
  if (!c || !ev_out) return SMC_ERR_ARG;
  try {
    if (!c->have_moments || nev > c->nev) throw SmcError(SMC_ERR_STATE, "no moments for these events");
    HIPCHK(hipMemcpyAsync(ev_out, c->evout, (size_t)nev * SMC_EV_STRIDE * sizeof(double), hipMemcpyDeviceToHost,
                          c->stream));
    sync(c);
    tag_systems(c, ev_out, nev);
    return SMC_OK;
  } catch (const SmcError& e) { return fail(c, e); }
}

smc_status smc_kln_dndy(smc_ctx* c, int n, const double* ta, const double* tb, double* out) {
  if (!c) return SMC_ERR_ARG;
  try {
    if (n <= 0 || !ta || !tb || !out) throw SmcError(SMC_ERR_ARG, "bad KLN point arrays");
    double* da = dalloc<double>((size_t)n);
    double* db = dalloc<double>((size_t)n);
    double* dout = dalloc<double>((size_t)n);
    HIPCHK(hipMemcpyAsync(da, ta, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(db, tb, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    kln_points(c, n, da, db, dout);
    HIPCHK(hipMemcpyAsync(out, dout, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    sync(c);
    (void)hipFree(da);
    (void)hipFree(db);
    (void)hipFree(dout);
    return SMC_OK;
  } catch (const SmcError& e) { return fail(c, e); }
}

smc_status smc_kln_make_table(smc_ctx* c, double* table) {
  if (!c) return SMC_ERR_ARG;
  try {
    ensure_kln_table(c);
    if (table) {
      const size_t n = (size_t)c->d.kln_tmax * c->d.kln_tmax;
      HIPCHK(hipMemcpyAsync(table, c->kln_table, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
      sync(c);
    }
    return SMC_OK;
  } catch (const SmcError& e) { return fail(c, e); }
}

smc_status smc_diff_setup(smc_ctx* c, uint64_t seed, uint64_t event0, int nev, double* theta0) {
  if (!c) return SMC_ERR_ARG;
  try {
    if (nev <= 0 || nev > c->cap) throw SmcError(SMC_ERR_CAPACITY, "nev outside [1, max_events]");
    diff_setup(c, seed, event0, nev);
    if (theta0) std::copy(c->dif_theta0, c->dif_theta0 + 5, theta0);
    return SMC_OK;
  } catch (const SmcError& e) { return fail(c, e); }
}

smc_status smc_diff_forward(smc_ctx* c, const double* theta, double* out) {
  if (!c || !theta) return SMC_ERR_ARG;
  try {
    diff_forward(c, theta);
    if (out)
      HIPCHK(hipMemcpyAsync(out, c->dif_out, (size_t)c->dif_nev * kDiffOut * sizeof(double), hipMemcpyDeviceToHost,
                            c->stream));
    sync(c);
    return SMC_OK;
  } catch (const SmcError& e) { return fail(c, e); }
}

smc_status smc_diff_backward(smc_ctx* c, const double* cot, double* grad) {
  if (!c || !cot || !grad) return SMC_ERR_ARG;
  try {
    diff_backward(c, cot, grad);
    return SMC_OK;
  } catch (const SmcError& e) { return fail(c, e); }
}

smc_status smc_diff_topology(smc_ctx* c, int ev, double* uvw_proj, double* uvw_targ, double* b, double* gamma_v,
                             int cap, int* nsrc) {
  if (!c) return SMC_ERR_ARG;
  try {
    if (!c->dif_nev) throw SmcError(SMC_ERR_STATE, "call smc_diff_setup first");
    if (ev < 0 || ev >= c->dif_nev) throw SmcError(SMC_ERR_ARG, "event index out of range");
    std::vector<double> du(2 * kAMax * 4);
    HIPCHK(hipMemcpyAsync(du.data(), c->dif.du + (size_t)ev * 2 * kAMax * 4, du.size() * sizeof(double),
                          hipMemcpyDeviceToHost, c->stream));
    int cnt[8];
    HIPCHK(hipMemcpyAsync(cnt, c->buf.counts + (size_t)ev * 8, sizeof(cnt), hipMemcpyDeviceToHost, c->stream));
    double bb = 0.0;
    HIPCHK(hipMemcpyAsync(&bb, c->dif.bval + ev, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    std::vector<double> gv((size_t)c->buf.src_cap);
    HIPCHK(hipMemcpyAsync(gv.data(), c->dif.gam_v + (size_t)ev * c->buf.src_cap, gv.size() * sizeof(double),
                          hipMemcpyDeviceToHost, c->stream));
    sync(c);
    for (int nu = 0; nu < 2; ++nu) {
      double* o = nu ? uvw_targ : uvw_proj;
      if (!o) continue;
      const int A = nu ? c->p.Atarg : c->p.Aproj;
      for (int i = 0; i < A; ++i)
        for (int q = 0; q < 3; ++q) o[3 * i + q] = du[((size_t)nu * kAMax + i) * 4 + q];
    }
    if (b) *b = bb;
    const int K = cnt[5];
    if (nsrc) *nsrc = K;
    if (gamma_v) {
      if (cap < K) throw SmcError(SMC_ERR_CAPACITY, "gamma_v capacity below the source count");
      for (int k = 0; k < K; ++k) gamma_v[k] = gv[k];
    }
    return SMC_OK;
  } catch (const SmcError& e) { return fail(c, e); }
}

smc_status smc_set_timing(smc_ctx* c, int on) {
  if (!c) return SMC_ERR_ARG;
  c->timing = on;
  return SMC_OK;
}

smc_status smc_get_timing(smc_ctx* c, int k, double* total_ms, int* launches) {
  if (!c || k < 0 || k >= SMC_K_COUNT) return SMC_ERR_ARG;
  try {
    sync(c);
    for (auto& pe : c->pending) {
      float ms = 0.f;
      HIPCHK(hipEventElapsedTime(&ms, pe.a, pe.b));
      c->tms[pe.k] += ms;
      c->tcount[pe.k] += 1;
      (void)hipEventDestroy(pe.a);
      (void)hipEventDestroy(pe.b);
    }
    c->pending.clear();
    if (total_ms) *total_ms = c->tms[k];
    if (launches) *launches = c->tcount[k];
    return SMC_OK;
  } catch (const SmcError& e) { return fail(c, e); }
}

smc_status smc_synchronize(smc_ctx* c) {
  if (!c) return SMC_ERR_ARG;
  try {
    sync(c);
    HIPCHK(hipDeviceSynchronize());
    return SMC_OK;
  } catch (const SmcError& e) { return fail(c, e); }
}

}  // extern "C"
