// smc_p3d_api.cpp -- host side of the 3D profiles (include/smc_p3d.h):
// the profile_3d constructor's constants and eta table
// (scripts/generate_3d_profiles/profile_3d.cpp:17-155) and the device calls.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/smc_p3d.h"
#include "smc_internal.h"

using namespace smc;

struct smc_p3d {
  smc_p3d_params p{};
  smc_p3d_derived d{};
  std::vector<double> eta, proj, targ;  // the eta table (host copy)
  double* tab = nullptr;                // device: eta, proj, targ
  hipStream_t stream = nullptr;
  std::string err;
};

namespace {

struct P3dError : std::runtime_error {
  int code;
  P3dError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define P3D_HIP(x)                                                                        \
  do {                                                                                    \
    hipError_t e_ = (x);                                                                  \
    if (e_ != hipSuccess) throw P3dError(SMC_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

smc_status p3d_fail(smc_p3d* h, const P3dError& e) {
  if (h) h->err = e.what();
  return e.code;
}

// device buffer freed at scope exit
template <typename T>
struct DevBuf {
  T* p = nullptr;
  explicit DevBuf(size_t n) {
    if (n == 0) n = 1;
    P3D_HIP(hipMalloc(&p, n * sizeof(T)));
  }
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
};

template <typename T>
void upload(T* dst, const T* src, size_t n, hipStream_t s) {
  if (n) P3D_HIP(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyHostToDevice, s));
}

P3dCfg cfg_of(const smc_p3d* h) {
  P3dCfg c{};
  c.nx = h->p.grid_nx;
  c.ny = h->p.grid_ny;
  c.neta = h->p.grid_neta;
  c.dx = h->p.grid_dx;
  c.dy = h->p.grid_dy;
  c.deta = h->p.grid_deta;
  c.random_flag = h->p.random_flag;
  c.y_beam = h->d.y_beam;
  c.eta0 = h->d.eta0;
  c.sigma_x0 = h->d.sigma_x0;
  c.sigma_eta0 = h->d.sigma_eta0;
  c.dsigma_x = 0.3;   // profile_3d.cpp:201-203
  c.dsigma_eta = 0.3;
  c.tab_eta = h->tab;
  c.tab_proj = h->tab + SMC_P3D_TABLE;
  c.tab_targ = h->tab + 2 * SMC_P3D_TABLE;
  c.tab_n = SMC_P3D_TABLE;
  c.proj_max = h->d.proj_max;
  c.targ_max = h->d.targ_max;
  return c;
}

// offsets: nev + 1 non-decreasing entries starting at 0
int check_offsets(const int* off, int nev, const char* what) {
  if (!off) throw P3dError(SMC_ERR_ARG, std::string(what) + ": offsets missing");
  if (off[0] != 0) throw P3dError(SMC_ERR_ARG, std::string(what) + ": offsets must start at 0");
  for (int e = 0; e < nev; ++e)
    if (off[e + 1] < off[e]) throw P3dError(SMC_ERR_ARG, std::string(what) + ": offsets decrease");
  return off[nev];
}

size_t grid_cells(const smc_p3d* h) {
  return (size_t)h->p.grid_neta * h->p.grid_nx * h->p.grid_ny;
}

void sample_dev(smc_p3d* h, int nev, const double* part_xy, const int* part_id, const int* part_off,
                const double* bin_xy, const int* bin_off, uint64_t seed, uint64_t event0, double* d_psrc,
                double* d_bsrc, int* d_poff, int* d_boff) {
  const int np = check_offsets(part_off, nev, "participants");
  const int nb = check_offsets(bin_off, nev, "binary collisions");
  if ((np && (!part_xy || !part_id)) || (nb && !bin_xy)) throw P3dError(SMC_ERR_ARG, "source arrays missing");
  DevBuf<double> pxy(2 * (size_t)np), bxy(2 * (size_t)nb);
  DevBuf<int> pid(np), st(1);
  upload(pxy.p, part_xy, 2 * (size_t)np, h->stream);
  upload(bxy.p, bin_xy, 2 * (size_t)nb, h->stream);
  upload(pid.p, part_id, (size_t)np, h->stream);
  upload(d_poff, part_off, (size_t)nev + 1, h->stream);
  upload(d_boff, bin_off, (size_t)nev + 1, h->stream);
  P3D_HIP(hipMemsetAsync(st.p, 0, sizeof(int), h->stream));
  P3D_HIP((hipError_t)launch_p3d_sample(cfg_of(h), seed, event0, nev, pxy.p, pid.p, d_poff, bxy.p, d_boff, d_psrc,
                                        d_bsrc, st.p, h->stream));
  int status = 0;
  P3D_HIP(hipMemcpyAsync(&status, st.p, sizeof(int), hipMemcpyDeviceToHost, h->stream));
  P3D_HIP(hipStreamSynchronize(h->stream));
  if (status) throw P3dError(SMC_ERR_ARG, "eta sampling did not accept within 65536 draws");
}

}  // namespace

extern "C" {

smc_status smc_p3d_create(smc_p3d** out, const smc_p3d_params* p) {
  if (!out || !p) return SMC_ERR_ARG;
  *out = nullptr;
  smc_p3d* h = new smc_p3d();
  try {
    h->p = *p;
    if (p->grid_nx < 1 || p->grid_ny < 1 || p->grid_neta < 1 || p->grid_nx > 0xffff || p->grid_ny > 0xffff ||
        p->grid_neta > 0xffff)
      throw P3dError(SMC_ERR_ARG, "grid sizes must be in [1, 65535]");
    if (!(p->grid_dx > 0 && p->grid_dy > 0 && p->grid_deta > 0)) throw P3dError(SMC_ERR_ARG, "bad grid spacing");
    if (p->random_flag < 0 || p->random_flag > 3) throw P3dError(SMC_ERR_ARG, "random_flag must be 0, 1, 2 or 3");
    if (!(p->ecm > 2.0)) throw P3dError(SMC_ERR_ARG, "ecm must exceed 2 GeV (y_beam)");
    // profile_3d.cpp:24-27
    smc_p3d_derived& d = h->d;
    d.y_beam = std::atanh(std::sqrt(1. - 1. / std::pow(p->ecm / 2., 2)));
    d.sigma_inelastic = regge_sigma_inel(p->ecm) * 0.1;
    // profile_3d.cpp:196-199
    d.eta0 = 2.0;
    d.sigma_x0 = std::sqrt(d.sigma_inelastic / (8 * M_PI));
    d.sigma_eta0 = 0.5;
    // profile_3d.cpp:133-155: 1000 points on [-y_beam, y_beam], linear in eta
    // (1 -/+ eta/y_beam), Gaussian fall-off of width 0.5 beyond |eta| = 2.5
    const double eta_peak = 2.5, sigma_eta_out = 0.5;
    const int n = SMC_P3D_TABLE;
    const double de = 2 * d.y_beam / (n - 1);
    h->eta.resize(n);
    h->proj.resize(n);
    h->targ.resize(n);
    d.proj_max = d.targ_max = 0.0;
    for (int i = 0; i < n; ++i) {
      const double eta = -d.y_beam + i * de;
      double f = 1.0;
      if (std::fabs(eta) > eta_peak)
        f = std::exp(-std::pow((std::fabs(eta) - eta_peak), 2) / (2. * sigma_eta_out * sigma_eta_out));
      const double left = (1. - eta / d.y_beam) * f, right = (1. + eta / d.y_beam) * f;
      d.proj_max = std::max(d.proj_max, left);
      d.targ_max = std::max(d.targ_max, right);
      h->eta[i] = eta;
      h->proj[i] = left;
      h->targ[i] = right;
    }
    P3D_HIP(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    P3D_HIP(hipMalloc(&h->tab, 3 * (size_t)n * sizeof(double)));
    upload(h->tab, h->eta.data(), n, h->stream);
    upload(h->tab + n, h->proj.data(), n, h->stream);
    upload(h->tab + 2 * n, h->targ.data(), n, h->stream);
    P3D_HIP(hipStreamSynchronize(h->stream));
    *out = h;
    return SMC_OK;
  } catch (const P3dError& e) {
    const int code = e.code;
    smc_p3d_destroy(h);
    return code;
  }
}

void smc_p3d_destroy(smc_p3d* h) {
  if (!h) return;
  if (h->tab) (void)hipFree(h->tab);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
}

const char* smc_p3d_last_error(const smc_p3d* h) { return h ? h->err.c_str() : "null handle"; }

smc_status smc_p3d_get_derived(const smc_p3d* h, smc_p3d_derived* d, double* eta, double* proj, double* targ) {
  if (!h) return SMC_ERR_ARG;
  if (d) *d = h->d;
  if (eta) std::copy(h->eta.begin(), h->eta.end(), eta);
  if (proj) std::copy(h->proj.begin(), h->proj.end(), proj);
  if (targ) std::copy(h->targ.begin(), h->targ.end(), targ);
  return SMC_OK;
}

smc_status smc_p3d_sample(smc_p3d* h, int nev, const double* part_xy, const int* part_id, const int* part_off,
                          const double* bin_xy, const int* bin_off, uint64_t seed, uint64_t event0,
                          double* part_src, double* bin_src) {
  if (!h || nev < 0) return SMC_ERR_ARG;
  if (nev == 0) return SMC_OK;
  try {
    const int np = check_offsets(part_off, nev, "participants"), nb = check_offsets(bin_off, nev, "binary collisions");
    DevBuf<double> ps(6 * (size_t)np), bs(6 * (size_t)nb);
    DevBuf<int> po((size_t)nev + 1), bo((size_t)nev + 1);
    sample_dev(h, nev, part_xy, part_id, part_off, bin_xy, bin_off, seed, event0, ps.p, bs.p, po.p, bo.p);
    if (part_src && np)
      P3D_HIP(hipMemcpyAsync(part_src, ps.p, 6 * (size_t)np * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    if (bin_src && nb)
      P3D_HIP(hipMemcpyAsync(bin_src, bs.p, 6 * (size_t)nb * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    P3D_HIP(hipStreamSynchronize(h->stream));
    return SMC_OK;
  } catch (const P3dError& e) { return p3d_fail(h, e); }
}

smc_status smc_p3d_smear(smc_p3d* h, int nev, const double* src, const int* src_off, float* rho) {
  if (!h || nev < 0 || !rho) return SMC_ERR_ARG;
  if (nev == 0) return SMC_OK;
  try {
    const int n = check_offsets(src_off, nev, "sources");
    if (n && !src) throw P3dError(SMC_ERR_ARG, "source records missing");
    for (int i = 0; i < n; ++i)
      if (!(src[6 * (size_t)i + 3] > 0 && src[6 * (size_t)i + 4] > 0 && src[6 * (size_t)i + 5] > 0))
        throw P3dError(SMC_ERR_ARG, "source widths must be positive");
    DevBuf<double> s(6 * (size_t)n);
    DevBuf<int> off((size_t)nev + 1);
    const size_t cells = grid_cells(h) * nev;
    DevBuf<float> g(cells);
    upload(s.p, src, 6 * (size_t)n, h->stream);
    upload(off.p, src_off, (size_t)nev + 1, h->stream);
    P3D_HIP((hipError_t)launch_p3d_smear(cfg_of(h), nev, s.p, off.p, g.p, h->stream));
    P3D_HIP(hipMemcpyAsync(rho, g.p, cells * sizeof(float), hipMemcpyDeviceToHost, h->stream));
    P3D_HIP(hipStreamSynchronize(h->stream));
    return SMC_OK;
  } catch (const P3dError& e) { return p3d_fail(h, e); }
}

smc_status smc_p3d_generate(smc_p3d* h, int nev, const double* part_xy, const int* part_id, const int* part_off,
                            const double* bin_xy, const int* bin_off, uint64_t seed, uint64_t event0,
                            float* rho_part, float* rho_binary, double* part_src, double* bin_src) {
  if (!h || nev < 0) return SMC_ERR_ARG;
  if (nev == 0) return SMC_OK;
  try {
    const int np = check_offsets(part_off, nev, "participants"), nb = check_offsets(bin_off, nev, "binary collisions");
    DevBuf<double> ps(6 * (size_t)np), bs(6 * (size_t)nb);
    DevBuf<int> po((size_t)nev + 1), bo((size_t)nev + 1);
    sample_dev(h, nev, part_xy, part_id, part_off, bin_xy, bin_off, seed, event0, ps.p, bs.p, po.p, bo.p);
    const size_t cells = grid_cells(h) * nev;
    const P3dCfg c = cfg_of(h);
    if (rho_part) {
      DevBuf<float> g(cells);
      P3D_HIP((hipError_t)launch_p3d_smear(c, nev, ps.p, po.p, g.p, h->stream));
      P3D_HIP(hipMemcpyAsync(rho_part, g.p, cells * sizeof(float), hipMemcpyDeviceToHost, h->stream));
      P3D_HIP(hipStreamSynchronize(h->stream));
    }
    if (rho_binary) {
      DevBuf<float> g(cells);
      P3D_HIP((hipError_t)launch_p3d_smear(c, nev, bs.p, bo.p, g.p, h->stream));
      P3D_HIP(hipMemcpyAsync(rho_binary, g.p, cells * sizeof(float), hipMemcpyDeviceToHost, h->stream));
      P3D_HIP(hipStreamSynchronize(h->stream));
    }
    if (part_src && np)
      P3D_HIP(hipMemcpyAsync(part_src, ps.p, 6 * (size_t)np * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    if (bin_src && nb)
      P3D_HIP(hipMemcpyAsync(bin_src, bs.p, 6 * (size_t)nb * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    P3D_HIP(hipStreamSynchronize(h->stream));
    return SMC_OK;
  } catch (const P3dError& e) { return p3d_fail(h, e); }
}

}  // extern "C"
