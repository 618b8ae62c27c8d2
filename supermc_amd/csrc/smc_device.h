// smc_device.h -- device helpers shared by the gfx950 kernels: Philox
// counter-based RNG, exact footprint-table lookup, wave/block reductions.
#pragma once
#include <hip/hip_runtime.h>

#include "smc_internal.h"

namespace smc {

// --------------------------------------------------------------- Philox
// Philox4x32-10, Salmon, Moraes, Dror & Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC'11.  Counter-based: every draw is a pure function of
// (seed, event, stream, index, attempt/iteration).
__device__ __forceinline__ uint4 philox(uint4 c, uint2 k) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const unsigned lo0 = 0xD2511F53u * c.x, hi0 = __umulhi(0xD2511F53u, c.x);
    const unsigned lo1 = 0xCD9E8D57u * c.z, hi1 = __umulhi(0xCD9E8D57u, c.z);
    c = make_uint4(hi1 ^ c.y ^ k.x, lo1, hi0 ^ c.w ^ k.y, lo0);
    k.x += 0x9E3779B9u;
    k.y += 0xBB67AE85u;
  }
  return c;
}

enum Stream : unsigned {
  kStrB = 1,
  kStrOrient = 2,
  kStrNucleon = 3,
  kStrFluctA = 4,
  kStrFluctB = 5,
  kStrFluctPair = 6,
  kStrDeuteron = 7,
  kStrDiffB = 8,        // differentiable batches: impact parameter
  kStrDiffNucleon = 9,  //   nucleon candidates (radial quantile, direction)
  kStrDiffGamma = 10,   //   Gamma quantile uniforms
  kStrCollide = 11,     // collision criterion 2: one uniform per tested pair
};

__device__ __forceinline__ uint4 draw(uint2 key, unsigned long long ev, unsigned stream,
                                      unsigned nucleus, unsigned idx, unsigned attempt,
                                      unsigned iter) {
  uint4 c;
  c.x = (unsigned)ev;
  c.y = ((unsigned)(ev >> 32) & 0xFFFFu) | ((stream & 0xFFu) << 16) | ((nucleus & 0xFFu) << 24);
  c.z = idx;
  c.w = (attempt << 16) | (iter & 0xFFFFu);
  return philox(c, key);
}

// uniform in (0,1) from 24 bits
__device__ __forceinline__ float u01(unsigned r) { return ((float)(r >> 8) + 0.5f) * (1.0f / 16777216.0f); }
// uniform in (0,1) from 53 bits
__device__ __forceinline__ double u53(unsigned a, unsigned b) {
  unsigned long long m = ((unsigned long long)(a >> 5) << 26) | (b >> 6);
  return ((double)m + 0.5) * (1.0 / 9007199254740992.0);
}

constexpr float kPi = 3.14159265358979323846f;
constexpr float kQf = (float)kQ;
constexpr float kHf = (float)kH;

// ------------------------------------------------------------ helpers
__device__ __forceinline__ int step_eval(const StepTable& T, int p) {
  int k = (int)floorf(fmaf(T.a, (float)p, T.b));
  k = max(0, min(T.n, k));
  while (k < T.n && T.thr[k] <= p) ++k;
  while (k > 0 && T.thr[k - 1] > p) --k;
  return T.vlo + k;
}

__device__ __forceinline__ int quant(float v) { return (int)rintf(v * 16777216.0f); }

// Point3D::rotate (MathBasics.cpp:41-50)
__device__ __forceinline__ void rotate3(float cth, float phi, float& x, float& y, float& z) {
  const float x0 = x, y0 = y, z0 = z;
  float sphi, cphi;
  sincosf(phi, &sphi, &cphi);
  const float sth = sqrtf(fmaxf(0.f, 1.f - cth * cth));
  x = cth * cphi * x0 - sphi * y0 + sth * cphi * z0;
  y = cth * sphi * x0 + cphi * y0 + sth * sphi * z0;
  z = -sth * x0 + cth * z0;
}

// Nucleus::SphericalHarmonics (Nucleus.cpp:669-694)
__device__ __forceinline__ float y20(float c) { return (3.f * c * c - 1.f) * 0.31539156525252005f; }
__device__ __forceinline__ float y40(float c) {
  const float c2 = c * c;
  return (35.f * c2 * c2 - 30.f * c2 + 3.f) * 0.10578554691520431f;
}

// One Woods-Saxon candidate by rejection (Nucleus.cpp:578-621): the r
// distribution r^2/(1+exp((r-R)/a)) on [0, R+2.5), isotropic direction,
// deformed radius R(theta) with the rotation of the nucleus applied.
__device__ inline void ws_candidate(const NucSys& ns, uint2 key, unsigned long long ev, unsigned nuc,
                             unsigned cand, unsigned attempt, float ctr, float phir, float& x,
                             float& y, float& z) {
  for (unsigned it = 0; it < 0xFFFFu; ++it) {
    const uint4 r = draw(key, ev, kStrNucleon, nuc, cand, attempt, it);
    if (ns.deformed) {
      const float rr = ns.rmax * cbrtf(u01(r.x));
      const float cx = 1.f - 2.f * u01(r.y);
      const float rad1 = ns.R * (1.f + ns.beta2 * y20(cx) + ns.beta4 * y40(cx));
      const float rw1 = 1.f / (1.f + expf(-rad1 / ns.a));
      if (u01(r.z) * rw1 > 1.f / (1.f + expf((rr - rad1) / ns.a))) continue;
      const float sx = sqrtf(fmaxf(0.f, 1.f - cx * cx));
      float sp, cp;
      sincosf(2.f * kPi * u01(r.w), &sp, &cp);
      x = rr * sx * cp;
      y = rr * sx * sp;
      z = rr * cx;
      rotate3(ctr, phir, x, y, z);
      return;
    } else {
      const float rr = ns.rmax * cbrtf(u01(r.x));
      if (u01(r.y) * ns.rwmax > 1.f / (1.f + expf((rr - ns.R) / ns.a))) continue;
      const float cx = 1.f - 2.f * u01(r.z);
      const float sx = sqrtf(fmaxf(0.f, 1.f - cx * cx));
      float sp, cp;
      sincosf(2.f * kPi * u01(r.w), &sp, &cp);
      x = rr * sx * cp;
      y = rr * sx * sp;
      z = rr * cx;
      return;
    }
  }
  x = y = z = 0.f;
}

// One proposal of the rejection sampler above (Nucleus.cpp:578-621): one
// draw per proposal; returns whether it is accepted and its position.
__device__ inline bool ws_proposal(const NucSys& ns, uint2 key, unsigned long long ev, unsigned nuc, unsigned idx,
                                   unsigned attempt, unsigned step, float ctr, float phir, float& x, float& y,
                                   float& z) {
  const uint4 r = draw(key, ev, kStrNucleon, nuc, idx, attempt, step);
  const float rr = ns.rmax * cbrtf(u01(r.x));
  float cx, acc;
  if (ns.deformed) {
    cx = 1.f - 2.f * u01(r.y);
    const float rad1 = ns.R * (1.f + ns.beta2 * y20(cx) + ns.beta4 * y40(cx));
    const float rw1 = 1.f / (1.f + expf(-rad1 / ns.a));
    acc = (u01(r.z) * rw1 <= 1.f / (1.f + expf((rr - rad1) / ns.a))) ? 1.f : 0.f;
  } else {
    // u rho(0) <= 1/(1+e^((r-R)/a))  <=>  u rho(0) (1 + e^((r-R)/a)) <= 1
    acc = (u01(r.y) * ns.rwmax * (1.f + __expf((rr - ns.R) * ns.inv_a)) <= 1.f) ? 1.f : 0.f;
    cx = 1.f - 2.f * u01(r.z);
  }
  const float sx = sqrtf(fmaxf(0.f, 1.f - cx * cx));
  // sin and cos of 2 pi u: the hardware sine and cosine take revolutions
  const float rev = u01(r.w);
  const float sp = __builtin_amdgcn_sinf(rev), cp = __builtin_amdgcn_cosf(rev);
  x = rr * sx * cp;
  y = rr * sx * sp;
  z = rr * cx;
  if (ns.deformed) rotate3(ctr, phir, x, y, z);
  return acc != 0.f;
}

// Hulthen CDF (HulthenFunc.cpp:27-36) inverted by safeguarded Newton.
__device__ inline float hulthen_inv(float u) {
  const float al = 0.228f, be = 1.18f;
  const float c = (al * be * (al + be)) / ((al - be) * (al - be));
  float lo = 0.f, hi = 100.f, r = 2.f;
  for (int it = 0; it < 60; ++it) {
    const float e1 = expf(-r * (al + be)), e2 = expf(-2.f * al * r), e3 = expf(-2.f * be * r);
    const float F = 2.f * c * (2.f * e1 / (al + be) - 0.5f * e2 / al - 0.5f * e3 / be + 0.5f / al + 0.5f / be -
                               2.f / (al + be)) - u;
    if (F > 0.f) hi = r; else lo = r;
    const float d = expf(-al * r) - expf(-be * r);
    const float pdf = 2.f * c * d * d;
    float rn = (pdf > 0.f) ? r - F / pdf : 0.5f * (lo + hi);
    if (!(rn > lo && rn < hi)) rn = 0.5f * (lo + hi);
    if (fabsf(rn - r) < 1e-6f) { r = rn; break; }
    r = rn;
  }
  return r;
}

// Marsaglia & Tsang (2000) Gamma(k, scale); k < 1 via G(k+1) U^(1/k).
__device__ inline float gamma_draw(float k, float scale, uint2 key, unsigned long long ev, unsigned stream,
                            unsigned idx) {
  unsigned it = 0;
  uint4 r = draw(key, ev, stream, 0, idx, 0, it++);
  float boost = 1.f;
  if (k < 1.f) {
    boost = powf(u01(r.w), 1.f / k);
    k += 1.f;
  }
  const float d = k - 1.f / 3.f, c = 1.f / sqrtf(9.f * d);
  for (int guard = 0; guard < 4096; ++guard) {
    const float u1 = u01(r.x), u2 = u01(r.y), u3 = u01(r.z);
    const float xn = sqrtf(-2.f * logf(u1)) * cosf(2.f * kPi * u2);
    float v = 1.f + c * xn;
    if (v > 0.f) {
      v = v * v * v;
      const float x2 = xn * xn;
      if (u3 < 1.f - 0.0331f * x2 * x2) return scale * d * v * boost;
      if (logf(u3) < 0.5f * x2 + d * (1.f - v + logf(v))) return scale * d * v * boost;
    }
    r = draw(key, ev, stream, 0, idx, 0, it++);
  }
  return scale * d * boost;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// 256-thread exclusive scan through LDS (tmp has 256 ints).
__device__ inline int block_excl_scan(int v, int* tmp) {
  const int t = threadIdx.x;
  tmp[t] = v;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    const int add = (t >= off) ? tmp[t - off] : 0;
    __syncthreads();
    tmp[t] += add;
    __syncthreads();
  }
  const int incl = tmp[t];
  __syncthreads();
  return incl - v;
}

// Ascending bitonic sort of key[0][0..255] and key[1][0..255] by a
// 256-thread workgroup (threads 0-127 sort key[0], 128-255 key[1]); ends
// with a barrier.
__device__ inline void bitonic_sort_2x256(unsigned long long (*key)[256]) {
  const int t = threadIdx.x, a = t >> 7, p = t & 127;
  __syncthreads();
  for (int k = 2; k <= 256; k <<= 1)
    for (int j = k >> 1; j > 0; j >>= 1) {
      const int i = ((p & ~(j - 1)) << 1) | (p & (j - 1));
      const int l = i + j;
      const bool up = (i & k) == 0;
      const unsigned long long x = key[a][i], y = key[a][l];
      if ((x > y) == up) {
        key[a][i] = y;
        key[a][l] = x;
      }
      __syncthreads();
    }
}

// Exact footprints of a source at (ph, qh) half quanta: the participant box
// (MCnucl.cpp:831-844) and the +-5w window (MCnucl.cpp:450-457, 729-736),
// each packed as lo | hi << 16 per axis.
__device__ __forceinline__ int4 src_windows(const DevCfg& cfg, int ph, int qh, bool box) {
  int4 w;
  if (box) {
    w.x = step_eval(cfg.tab[kBoxLoX], ph) | (step_eval(cfg.tab[kBoxHiX], ph) << 16);
    w.y = step_eval(cfg.tab[kBoxLoY], qh) | (step_eval(cfg.tab[kBoxHiY], qh) << 16);
  } else {
    w.x = 0;
    w.y = 0;
  }
  w.z = step_eval(cfg.tab[kWinLoX], ph) | (step_eval(cfg.tab[kWinHiX], ph) << 16);
  w.w = step_eval(cfg.tab[kWinLoY], qh) | (step_eval(cfg.tab[kWinHiY], qh) << 16);
  return w;
}

// Sum of a double over a 256-thread block (red: 4 doubles of LDS).
__device__ __forceinline__ double block_sum_d(double v, double* red) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wv] = v;
  __syncthreads();
  return red[0] + red[1] + red[2] + red[3];
}


}  // namespace smc
