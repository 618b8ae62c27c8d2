// smc_p3d.hip -- the 3D (x, y, eta) extension of operation 2:
// scripts/generate_3d_profiles/profile_3d.cpp.
//
//  k_p3d_sample  one lane per source: the source's space-time rapidity from
//                the tabulated eta distribution by rejection
//                (profile_3d.cpp:157-191) and its widths per random_flag
//                (:194-300), Philox draws keyed by (event, list, index).
//  k_p3d_smear   one wave per (event, eta slice, 32 x 64 tile): every source
//                whose eta window holds the slice and whose x/y windows meet
//                the tile adds norm * exp(-dis_eta) exp(-dis_x) exp(-dis_y)
//                (:302-400) in list order -- one FMA per cell, the row
//                factors exchanged through LDS, the weight w exp(-dis_eta)
//                folded into each lane's column factor.
#include "smc_device.h"

namespace smc {

// counter streams of this path (beside smc_device.h's)
constexpr unsigned kStrP3dEta = 16, kStrP3dWidth = 17;

// profile_3d.cpp:402-434: index of the largest entry below v, -1 out of range
__device__ inline int p3d_search(const double* A, int n, double v) {
  int i = 0, f = n - 1;
  if (v > A[f] || v < A[i]) return -1;
  int m = (f + i) / 2;
  while (f - i > 1) {
    if (A[m] < v) i = m;
    else f = m;
    m = (f + i) / 2;
  }
  return i;
}

// profile_3d.cpp:157-191 for one source; draws (event, list, idx, iteration)
__device__ inline double p3d_eta(const P3dCfg& c, uint2 key, unsigned long long ev, unsigned list, unsigned idx,
                                 int id, int* status) {
  const double* tab = id == 1 ? c.tab_proj : c.tab_targ;
  const double pmax = id == 1 ? c.proj_max : c.targ_max;
  for (unsigned it = 0; it < 65536u; ++it) {
    const uint4 r = draw(key, ev, kStrP3dEta, list, idx, 0, it);
    const double eta = c.y_beam * (1. - 2. * u53(r.x, r.y));
    const int k = p3d_search(c.tab_eta, c.tab_n, eta);
    if (k < 0 || k + 1 >= c.tab_n) continue;  // outside the table: draw again
    const double frac = (eta - c.tab_eta[k]) / (c.tab_eta[k + 1] - c.tab_eta[k]);
    const double prob = tab[k] * (1. - frac) + tab[k + 1] * frac;
    if (!(pmax * u53(r.z, r.w) > prob)) return eta;
  }
  *status |= 1;
  return 0.0;
}

// src: 6 doubles per source {x, y, eta, sigma_x, sigma_y, sigma_eta}
__global__ __launch_bounds__(256) void k_p3d_sample(P3dCfg c, uint2 key, unsigned long long event0, int nev,
                                                    const double* part_xy, const int* part_id, const int* part_off,
                                                    const double* bin_xy, const int* bin_off, double* part_src,
                                                    double* bin_src, int* status) {
  const int e = blockIdx.y;
  if (e >= nev) return;
  const int np = part_off[e + 1] - part_off[e], nb = bin_off[e + 1] - bin_off[e];
  const unsigned long long ev = event0 + (unsigned long long)e;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < np + nb; i += gridDim.x * blockDim.x) {
    const bool bin = i >= np;
    const int k = bin ? i - np : i;  // index within its list
    const int g = bin ? bin_off[e] + k : part_off[e] + k;
    // binary collisions take the id of the participant with their index
    // (profile_3d.cpp:263-299, participant_list[i].id); past the end of the
    // participant list (undefined there) the index wraps around it
    const int id = np > 0 ? part_id[part_off[e] + (bin ? k % np : k)] : 2;
    const double* xy = bin ? bin_xy : part_xy;
    double* o = (bin ? bin_src : part_src) + (size_t)g * 6;
    o[0] = xy[2 * (size_t)g];
    o[1] = xy[2 * (size_t)g + 1];
    double eta, sx = c.sigma_x0, sy = c.sigma_x0, se = c.sigma_eta0;
    const unsigned list = bin ? 1u : 0u;
    if (c.random_flag == 0) {
      eta = bin ? c.eta0 : (id == 1 ? c.eta0 : -c.eta0);
    } else {
      int st = 0;
      eta = p3d_eta(c, key, ev, list, (unsigned)k, id, &st);
      if (st) atomicOr(status, 1);
      const uint4 r = draw(key, ev, kStrP3dWidth, list, (unsigned)k, 0, 0);
      if (c.random_flag == 2) {
        se = c.sigma_eta0 + c.dsigma_eta * (1. - 2. * u53(r.x, r.y));
        sx = sy = c.sigma_x0 + c.dsigma_x * (1. - 2. * u53(r.z, r.w));
      } else if (c.random_flag == 3) {
        const uint4 r2 = draw(key, ev, kStrP3dWidth, list, (unsigned)k, 0, 1);
        sx = c.sigma_x0 + c.dsigma_x * (1. - 2. * u53(r.x, r.y));
        sy = c.sigma_x0 + c.dsigma_x * (1. - 2. * u53(r.z, r.w));
        se = c.sigma_eta0 + c.dsigma_eta * (1. - 2. * u53(r2.x, r2.y));
      }
    }
    o[2] = eta;
    o[3] = sx;
    o[4] = sy;
    o[5] = se;
  }
}

struct P3dEntry {
  float x, y;           // source position [fm]
  float ax, ay;         // -log2(e) / (2 sigma^2) per axis
  float w;              // norm_x norm_y norm_eta exp(-dis_eta) at this slice
  int xw, yw;           // [lo, hi) cell windows, 16 bits each
  int pad;
};

constexpr int kP3dRows = 32;

// profile_3d.cpp:310-323: idx0 = (int)(p/d + (n - 1)/2) with integer
// (n - 1)/2, range = (int)(6 sigma/d), cells [idx0 - range, idx0 + range)
// clamped to [0, n)
__device__ __forceinline__ int p3d_window(double p, double s, double d, int n) {
  const int i0 = (int)(p / d + (n - 1) / 2);
  const int rg = (int)(6. * s / d);
  const int lo = max(i0 - rg, 0), hi = min(i0 + rg, n);
  return (hi > lo) ? (lo | (hi << 16)) : 0;
}

__global__ __launch_bounds__(64) void k_p3d_smear(P3dCfg c, int nev, const double* src, const int* src_off,
                                                  float* rho) {
  __shared__ P3dEntry wl[64];
  __shared__ float4 rd[kP3dRows / 4];
  const int lane = threadIdx.x;
  const int nx = c.nx, ny = c.ny, ne = c.neta;
  const int ntx = (nx + kP3dRows - 1) / kP3dRows, nty = (ny + 63) / 64;
  const int tile = blockIdx.x % (ntx * nty), j = blockIdx.x / (ntx * nty);  // j: eta slice
  const int e = blockIdx.y;
  if (j >= ne || e >= nev) return;
  const int r0 = kP3dRows * (tile % ntx), c0 = 64 * (tile / ntx);
  const int col = c0 + lane;
  // cell centres (profile_3d.cpp:80-91): (i - (n - 1)/2.) d
  const double xc = (r0 + (lane & 31) - (nx - 1) / 2.) * c.dx;
  const double yc = (col - (ny - 1) / 2.) * c.dy;
  const double etac = (j - (ne - 1) / 2.) * c.deta;
  const int s0 = src_off[e], n = src_off[e + 1] - s0;
  float* rdf = reinterpret_cast<float*>(&rd[0]);
  constexpr float kLog2e = 1.4426950408889634f;
  constexpr double kTwoPi = 6.283185307179586;

  float acc[kP3dRows];
#pragma unroll
  for (int r = 0; r < kP3dRows; ++r) acc[r] = 0.f;

  for (int sub = 0; sub < n; sub += 64) {
    const int k = sub + lane;
    bool sel = false;
    P3dEntry en{0.f, 0.f, 0.f, 0.f, 0.f, 0, 0, 0};
    if (k < n) {
      const double* s = src + (size_t)(s0 + k) * 6;
      const double px = s[0], py = s[1], pe = s[2], sx = s[3], sy = s[4], se = s[5];
      const int ew = p3d_window(pe, se, c.deta, ne);
      if ((ew & 0xffff) <= j && (ew >> 16) > j) {
        en.xw = p3d_window(px, sx, c.dx, nx);
        en.yw = p3d_window(py, sy, c.dy, ny);
        const int xlo = en.xw & 0xffff, xhi = en.xw >> 16, ylo = en.yw & 0xffff, yhi = en.yw >> 16;
        sel = xlo < r0 + kP3dRows && xhi > r0 && ylo < c0 + 64 && yhi > c0;
        if (sel) {
          const double de = etac - pe;
          const double norm = 1. / sqrt(kTwoPi * se * se) / sqrt(kTwoPi * sx * sx) / sqrt(kTwoPi * sy * sy);
          en.w = (float)(norm * exp(-de * de / (2. * se * se)));
          en.x = (float)px;
          en.y = (float)py;
          en.ax = (float)(-1. / (2. * sx * sx)) * kLog2e;
          en.ay = (float)(-1. / (2. * sy * sy)) * kLog2e;
        }
      }
    }
    const unsigned long long m = __ballot(sel);
    if (sel) wl[__popcll(m & ((1ull << lane) - 1ull))] = en;
    const int cnt = __popcll(m);
    __builtin_amdgcn_wave_barrier();
    for (int g = 0; g < cnt; ++g) {
      const P3dEntry q = wl[g];
      const int xlo = q.xw & 0xffff, xhi = q.xw >> 16, ylo = q.yw & 0xffff, yhi = q.yw >> 16;
      if (lane < kP3dRows) {
        const int row = r0 + lane;
        const float dx = (float)(xc - (double)q.x);
        rdf[lane] = (row >= xlo && row < xhi) ? __builtin_amdgcn_exp2f(dx * dx * q.ax) : 0.f;
      }
      const float dy = (float)(yc - (double)q.y);
      const float gy = (col >= ylo && col < yhi) ? q.w * __builtin_amdgcn_exp2f(dy * dy * q.ay) : 0.f;
      __builtin_amdgcn_wave_barrier();
#pragma unroll
      for (int cc = 0; cc < kP3dRows / 8; ++cc) {
        if (r0 + 8 * cc < xhi && r0 + 8 * cc + 8 > xlo) {  // 8-row chunks the window misses are skipped
#pragma unroll
          for (int i = 2 * cc; i < 2 * cc + 2; ++i) {
            const float4 v = rd[i];
            acc[4 * i] = fmaf(v.x, gy, acc[4 * i]);
            acc[4 * i + 1] = fmaf(v.y, gy, acc[4 * i + 1]);
            acc[4 * i + 2] = fmaf(v.z, gy, acc[4 * i + 2]);
            acc[4 * i + 3] = fmaf(v.w, gy, acc[4 * i + 3]);
          }
        }
      }
      __builtin_amdgcn_wave_barrier();
    }
    __builtin_amdgcn_wave_barrier();
  }
  if (col >= ny) return;
  float* o = rho + ((size_t)e * ne + j) * (size_t)nx * ny;
#pragma unroll
  for (int r = 0; r < kP3dRows; ++r) {
    const int ix = r0 + r;
    if (ix < nx) o[(size_t)ix * ny + col] = acc[r];
  }
}

int launch_p3d_sample(const P3dCfg& c, unsigned long long seed, unsigned long long event0, int nev,
                      const double* part_xy, const int* part_id, const int* part_off, const double* bin_xy,
                      const int* bin_off, double* part_src, double* bin_src, int* status, void* stream) {
  if (nev <= 0) return 0;
  const uint2 key = make_uint2((unsigned)seed, (unsigned)(seed >> 32));
  hipLaunchKernelGGL(k_p3d_sample, dim3(4, nev), dim3(256), 0, (hipStream_t)stream, c, key, event0, nev, part_xy,
                     part_id, part_off, bin_xy, bin_off, part_src, bin_src, status);
  return (int)hipGetLastError();
}

int launch_p3d_smear(const P3dCfg& c, int nev, const double* src, const int* src_off, float* rho, void* stream) {
  if (nev <= 0) return 0;
  if (c.nx < 1 || c.ny < 1 || c.neta < 1 || c.nx > 0xffff || c.ny > 0xffff || c.neta > 0xffff)
    return (int)hipErrorInvalidValue;
  const int tiles = ((c.nx + kP3dRows - 1) / kP3dRows) * ((c.ny + 63) / 64);
  hipLaunchKernelGGL(k_p3d_smear, dim3((unsigned)(tiles * c.neta), nev), dim3(64), 0, (hipStream_t)stream, c, nev,
                     src, src_off, rho);
  return (int)hipGetLastError();
}

}  // namespace smc
