// smc_kernels.hip -- gfx950 kernels for superMC's per-event initial-condition
// hot path.
//
//  k_event    one workgroup per event: impact parameter, Philox sampling of
//             both nuclei (Woods-Saxon with hard core, deformed WS, Hulthen
//             deuteron, proton), CM recentering, x-sort, A x B disk test with
//             exact integer predicate, participant / pair compaction in the
//             reference's order, Gamma fluctuation factors, source list.
//             (MakeDensity.cpp:2143-2162, Nucleus.cpp:187-317,
//              MCnucl.cpp:217-352, MCnucl.cpp:1271-1301)
//  k_smear    2D Gaussian smearing of a source list onto the grid, one
//             workgroup per (event, 16x64 tile), exact reference footprints
//             via integer threshold tables.  (MCnucl.cpp:398-478, 626-866)
//  k_moments  FP64 centre of mass, <r^n> and masked eps_n moments, one
//             workgroup per event.  (MakeDensity.cpp:2244-2430,
//             MCnucl.cpp:1303-1323)
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
__device__ void ws_candidate(const NucSys& ns, uint2 key, unsigned long long ev, unsigned nuc,
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

// Hulthen CDF (HulthenFunc.cpp:27-36) inverted by safeguarded Newton.
__device__ float hulthen_inv(float u) {
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
__device__ float gamma_draw(float k, float scale, uint2 key, unsigned long long ev, unsigned stream,
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
__device__ int block_excl_scan(int v, int* tmp) {
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

// ------------------------------------------------------------- k_event
__global__ __launch_bounds__(256) void k_event(DevCfg cfg, EvBufs buf, uint2 key,
                                               unsigned long long event0, int fed) {
  const int e = blockIdx.x;
  const int t = threadIdx.x;
  const int lane = t & 63, wv = t >> 6;
  const unsigned long long gev = event0 + (unsigned long long)e;

  __shared__ int px[2][kAMax], py[2][kAMax];  // input/sample order (quanta)
  __shared__ float pz[2][kAMax];              // local z (hard core only)
  __shared__ int sx[2][kAMax], sy[2][kAMax];  // sorted
  __shared__ unsigned hits[kAMax][kAMax / 32];
  __shared__ int rowcnt[kAMax], fh[kAMax], scan_tmp[256];
  __shared__ int s_acc[2];
  __shared__ double s_b;
  __shared__ int s_xc;
  __shared__ double s_red[2][2];

  const int A = cfg.nuc[0].A, B = cfg.nuc[1].A;
  const int NN[2] = {A, B};
  int attempt = 0, status = kStOk;
  int npa = 0, npb = 0, ncoll = 0;

  for (;; ++attempt) {
    if (fed) {
      for (int i = t; i < A; i += 256) {
        px[0][i] = buf.raw[((size_t)e * 2 + 0) * kAMax * 2 + 2 * i];
        py[0][i] = buf.raw[((size_t)e * 2 + 0) * kAMax * 2 + 2 * i + 1];
      }
      for (int i = t; i < B; i += 256) {
        px[1][i] = buf.raw[((size_t)e * 2 + 1) * kAMax * 2 + 2 * i];
        py[1][i] = buf.raw[((size_t)e * 2 + 1) * kAMax * 2 + 2 * i + 1];
      }
      __syncthreads();
    } else {
      // b = sqrt((bmax^2-bmin^2) u + bmin^2)  (MakeDensity.cpp:2149); the
      // nuclei sit at +-b/2 rounded to the position quantum.
      if (t == 0) {
        const uint4 r = draw(key, gev, kStrB, 0, 0, attempt, 0);
        const double b = sqrt((cfg.bmax2 - cfg.bmin2) * u53(r.x, r.y) + cfg.bmin2);
        const int xc = (int)llrint(b * 0.5 / kQ);
        s_xc = xc;
        s_b = 2.0 * (double)xc * kQ;
        s_acc[0] = s_acc[1] = 0;
      }
      __syncthreads();
      // orientation of each nucleus (Nucleus.cpp:193-197)
      float ctr[2], phr[2];
      for (int nu = 0; nu < 2; ++nu) {
        const uint4 r = draw(key, gev, kStrOrient, nu, 0, attempt, 0);
        ctr[nu] = 1.f - 2.f * u01(r.x);
        phr[nu] = 2.f * kPi * u01(r.y);
      }
      // proton / deuteron fill directly (Nucleus.cpp:199-210, 342-359)
      if (wv < 2) {
        const NucSys& ns = cfg.nuc[wv];
        const int xcen = (wv == 0) ? s_xc : -s_xc;
        if (ns.kind == kProton) {
          if (lane == 0) { px[wv][0] = xcen; py[wv][0] = 0; pz[wv][0] = 0.f; s_acc[wv] = 1; }
        } else if (ns.kind == kDeuteron) {
          if (lane == 0) {
            const uint4 r = draw(key, gev, kStrDeuteron, wv, 0, attempt, 0);
            const float d = hulthen_inv(u01(r.x));
            float x = 0.5f * d, y = 0.f, z = 0.f;
            rotate3(ctr[wv], phr[wv], x, y, z);
            const int xq = quant(x), yq = quant(y);
            px[wv][0] = xq + xcen; py[wv][0] = yq; pz[wv][0] = z;
            px[wv][1] = -xq + xcen; py[wv][1] = -yq; pz[wv][1] = -z;
            s_acc[wv] = 2;
          }
        }
      }
      __syncthreads();
      // Woods-Saxon nuclei: sequential acceptance against all earlier
      // nucleons (hard core, Nucleus.cpp:276-294), 64 candidates per round
      // taken in stream order -- the same law as the reference's one-at-a-
      // time resampling.
      for (unsigned round = 0;; ++round) {
        const bool need0 = cfg.nuc[0].kind == kWoodsSaxon && s_acc[0] < A;
        const bool need1 = cfg.nuc[1].kind == kWoodsSaxon && s_acc[1] < B;
        if (!need0 && !need1) break;
        if (round > 4096) { status |= kStNoEvent; break; }
        if (wv < 2 && ((wv == 0) ? need0 : need1)) {
          const NucSys& ns = cfg.nuc[wv];
          const int An = NN[wv];
          const int acc0 = s_acc[wv];
          float x, y, z;
          ws_candidate(ns, key, gev, wv, round * 64u + lane, attempt, ctr[wv], phr[wv], x, y, z);
          const int xq = quant(x), yq = quant(y);
          bool ok = true;
          for (int j = 0; j < acc0 && ok; ++j) {
            const float dx = (float)(xq - px[wv][j]) * kQf, dy = (float)(yq - py[wv][j]) * kQf;
            const float dz = z - pz[wv][j];
            if (dx * dx + dy * dy + dz * dz < cfg.dmin2) ok = false;
          }
          bool mine = false;
          int slot = 0, nacc = acc0;
          for (int l = 0; l < 64 && nacc < An; ++l) {
            const int okl = __shfl((int)ok, l, 64);
            if (!okl) continue;
            const int xl = __shfl(xq, l, 64), yl = __shfl(yq, l, 64);
            const float zl = __shfl(z, l, 64);
            bool conflict = false;
            if (mine) {
              const float dx = (float)(xq - xl) * kQf, dy = (float)(yq - yl) * kQf, dz = z - zl;
              conflict = dx * dx + dy * dy + dz * dz < cfg.dmin2;
            }
            if (__any(conflict)) continue;
            if (lane == l) { mine = true; slot = nacc; }
            ++nacc;
          }
          if (mine) { px[wv][slot] = xq; py[wv][slot] = yq; pz[wv][slot] = z; }
          if (lane == 0) s_acc[wv] = nacc;
        }
        __syncthreads();
      }
      // CM recentering (Nucleus.cpp:296-309) in FP64, shift rounded to the
      // quantum, then the nucleus centre +-b/2.
      if (wv < 2 && cfg.nuc[wv].kind == kWoodsSaxon) {
        const int An = NN[wv];
        double sxm = 0.0, sym = 0.0;
        for (int i = lane; i < An; i += 64) { sxm += (double)px[wv][i]; sym += (double)py[wv][i]; }
        sxm = wave_sum(sxm);
        sym = wave_sum(sym);
        const int cxq = (int)llrint(sxm / An), cyq = (int)llrint(sym / An);
        const int xcen = (wv == 0) ? s_xc : -s_xc;
        for (int i = lane; i < An; i += 64) { px[wv][i] = px[wv][i] - cxq + xcen; py[wv][i] -= cyq; }
      }
      __syncthreads();
    }

    // ---- sort each nucleus by x, ties by input index (Nucleus.cpp:314)
    for (int nu = 0; nu < 2; ++nu) {
      const int An = NN[nu];
      for (int i = t; i < An; i += 256) {
        const int xi = px[nu][i];
        int rank = 0;
        for (int j = 0; j < An; ++j) {
          const int xj = px[nu][j];
          rank += (xj < xi) || (xj == xi && j < i);
        }
        sx[nu][rank] = xi;
        sy[nu][rank] = py[nu][i];
      }
    }
    __syncthreads();

    // ---- A x B disk test (MCnucl.cpp:359-366), exact in integers:
    // fl(fl(sqrt(d2))^2) <= dsq  <=>  dxq^2 + dyq^2 <= hit_n.
    {
      int cnt = 0;
      if (t < A) {
        unsigned bits[kAMax / 32];
#pragma unroll
        for (int w = 0; w < kAMax / 32; ++w) bits[w] = 0u;
        const int xa = sx[0][t], ya = sy[0][t];
        const int lim = cfg.hit_lim;
        int lo = 0, hi = B;
        while (lo < hi) {
          const int mid = (lo + hi) >> 1;
          if (sx[1][mid] < xa - lim) lo = mid + 1; else hi = mid;
        }
        for (int j = lo; j < B; ++j) {
          const int xb = sx[1][j];
          if (xb > xa + lim) break;
          const int dx = xa - xb, dy = ya - sy[1][j];
          if (dy > lim || dy < -lim) continue;
          const long long n = (long long)dx * dx + (long long)dy * dy;
          if (n <= cfg.hit_n) {
#pragma unroll
            for (int w = 0; w < kAMax / 32; ++w)
              if (w == (j >> 5)) bits[w] |= 1u << (j & 31);
            ++cnt;
          }
        }
#pragma unroll
        for (int w = 0; w < kAMax / 32; ++w) hits[t][w] = bits[w];
        rowcnt[t] = cnt;
      } else {
        rowcnt[t] = 0;
      }
      __syncthreads();
      // first hitting projectile row of each target nucleon: its wound order
      int first = -1;
      if (t < B) {
        const unsigned m = 1u << (t & 31);
        const int wd = t >> 5;
        for (int i = 0; i < A; ++i)
          if (hits[i][wd] & m) { first = i; break; }
      }
      fh[t] = first;
      npa = __syncthreads_count(t < A && cnt > 0);
      npb = __syncthreads_count(first >= 0);
      // total Ncoll
      int tot = wave_sum(cnt);
      if (lane == 0) scan_tmp[wv] = tot;
      __syncthreads();
      ncoll = scan_tmp[0] + scan_tmp[1] + scan_tmp[2] + scan_tmp[3];
      __syncthreads();
    }
    if (fed) break;
    const int np = npa + npb;
    if (ncoll > 0 && np >= cfg.npmin && np <= cfg.npmax && !(status & kStNoEvent)) break;
    if (attempt + 1 >= cfg.max_attempts || (status & kStNoEvent)) { status |= kStNoEvent; break; }
    __syncthreads();
  }

  // ---- outputs: sorted positions, wound order, pairs, fluct, sources
  const size_t nbase = (size_t)e * 2 * kAMax;
  int woundA = -1, woundB = -1;
  {
    const int wA = (t < A && rowcnt[t] > 0) ? 1 : 0;
    const int rA = block_excl_scan(wA, scan_tmp);
    if (wA) woundA = rA;
    if (t < B && fh[t] >= 0) {
      int r = 0;
      const int f = fh[t];
      for (int j = 0; j < B; ++j) {
        const int g = fh[j];
        r += (g >= 0) && (g < f || (g == f && j < t));
      }
      woundB = r;
    }
  }
  const int off = block_excl_scan(t < A ? rowcnt[t] : 0, scan_tmp);
  float fA = 1.f, fB = 1.f;
  if (cfg.fluct) {
    if (woundA >= 0) fA = gamma_draw(cfg.gk_part, cfg.gs_part, key, gev, kStrFluctA, (unsigned)t);
    if (woundB >= 0) fB = gamma_draw(cfg.gk_part, cfg.gs_part, key, gev, kStrFluctB, (unsigned)t);
  }
  if (t < A) {
    buf.pos[(nbase + t) * 2] = sx[0][t];
    buf.pos[(nbase + t) * 2 + 1] = sy[0][t];
    buf.wound[nbase + t] = woundA;
    buf.fluct[nbase + t] = fA;
  }
  if (t < B) {
    buf.pos[(nbase + kAMax + t) * 2] = sx[1][t];
    buf.pos[(nbase + kAMax + t) * 2 + 1] = sy[1][t];
    buf.wound[nbase + kAMax + t] = woundB;
    buf.fluct[nbase + kAMax + t] = fB;
  }
  const int nsrc_part = npa + npb;
  int4* src = buf.src + (size_t)e * buf.src_cap;
  const bool src_ok = nsrc_part + ncoll <= buf.src_cap;
  if (src_ok) {
    if (woundA >= 0) src[woundA] = make_int4(2 * sx[0][t], 2 * sy[0][t], __float_as_int(fA), kSrcPartA);
    if (woundB >= 0) src[npa + woundB] = make_int4(2 * sx[1][t], 2 * sy[1][t], __float_as_int(fB), kSrcPartB);
  }
  if (ncoll > cfg.max_pairs || !src_ok) status |= kStPairOverflow;
  if (t < A && rowcnt[t] > 0) {
    const unsigned long long pbase = (unsigned long long)e * cfg.max_pairs;
    int k = off;
    const int xa = sx[0][t], ya = sy[0][t];
    for (int w = 0; w < kAMax / 32; ++w) {
      unsigned m = hits[t][w];
      while (m) {
        const int b = __ffs(m) - 1;
        m &= m - 1;
        const int j = w * 32 + b;
        if (k < cfg.max_pairs) {
          float fp = 1.f;
          if (cfg.fluct) fp = gamma_draw(cfg.gk_bin, cfg.gs_bin, key, gev, kStrFluctPair, (unsigned)k);
          buf.pairs[pbase + k] = (unsigned)t | ((unsigned)j << 16);
          buf.pfluct[pbase + k] = fp;
          if (src_ok)
            src[nsrc_part + k] = make_int4(xa + sx[1][j], ya + sy[1][j], __float_as_int(fp), kSrcPair);
        }
        ++k;
      }
    }
  }
  if (t == 0) {
    int* c = buf.counts + (size_t)e * 8;
    c[0] = npa;
    c[1] = npb;
    c[2] = ncoll;
    c[3] = status;
    c[4] = attempt + 1;
    c[5] = src_ok ? nsrc_part + ncoll : 0;
    if (!fed) buf.bval[e] = s_b;
  }
}

// ------------------------------------------------------------- k_smear
// Gather formulation: each workgroup owns a 16 (x) x 64 (y) tile, culls the
// event's sources against it (ordered, deterministic compaction) and each
// lane accumulates 4 cells of one column.  Cell values use the reference's
// footprints: window [lo,hi) from the threshold tables and the disk cut
// d^2 <= 25 w^2.
constexpr int kTX = 16, kTY = 64;

__global__ __launch_bounds__(256) void k_smear(DevCfg cfg, EvBufs buf, SmearSpec spec, float* out0,
                                               float* out1) {
  const int e = blockIdx.y;
  const int ntx = (cfg.nx + kTX - 1) / kTX;
  const int tx0 = (blockIdx.x % ntx) * kTX, ty0 = (blockIdx.x / ntx) * kTY;
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const int iy = ty0 + lane;
  const int rx0 = tx0 + wv * 4;

  __shared__ int4 lsrc[256];
  __shared__ int4 lwin[256];
  __shared__ float lw[256];
  __shared__ int wcount[4];

  int xgh[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) xgh[r] = (int)llrint((cfg.xmin + (rx0 + r) * cfg.dx) / kH);
  const int ygh = (int)llrint((cfg.ymin + iy * cfg.dy) / kH);

  float a0[4] = {0.f, 0.f, 0.f, 0.f}, a1[4] = {0.f, 0.f, 0.f, 0.f};
  const int K = buf.counts[(size_t)e * 8 + 5];
  const int4* src = buf.src + (size_t)e * buf.src_cap;

  for (int base = 0; base < K; base += 256) {
    const int k = base + t;
    bool take = false;
    int4 s = make_int4(0, 0, 0, 0), win = make_int4(0, 0, 0, 0);
    float wgt = 0.f;
    if (k < K) {
      s = src[k];
      const int kind = s.w;
      if (spec.enabled[kind]) {
        if (spec.win[kind] == kWinBox) {
          win.x = step_eval(cfg.tab[kBoxLoX], s.x);
          win.y = step_eval(cfg.tab[kBoxHiX], s.x);
          win.z = step_eval(cfg.tab[kBoxLoY], s.y);
          win.w = step_eval(cfg.tab[kBoxHiY], s.y);
        } else {
          win.x = step_eval(cfg.tab[kWinLoX], s.x);
          win.y = step_eval(cfg.tab[kWinHiX], s.x);
          win.z = step_eval(cfg.tab[kWinLoY], s.y);
          win.w = step_eval(cfg.tab[kWinHiY], s.y);
        }
        take = win.x < win.y && win.z < win.w && win.x < tx0 + kTX && win.y > tx0 && win.z < ty0 + kTY &&
               win.w > ty0;
        wgt = spec.scale[kind] * (spec.use_fluct[kind] ? __int_as_float(s.z) : 1.f);
        s.z = spec.acc[kind];
      }
    }
    // ordered compaction
    const unsigned long long m = __ballot(take);
    const int before = __popcll(m & ((1ull << lane) - 1ull));
    if (lane == 0) wcount[wv] = __popcll(m);
    __syncthreads();
    int woff = 0;
    for (int q = 0; q < wv; ++q) woff += wcount[q];
    const int n = wcount[0] + wcount[1] + wcount[2] + wcount[3];
    if (take) {
      lsrc[woff + before] = s;
      lwin[woff + before] = win;
      lw[woff + before] = wgt;
    }
    __syncthreads();
    for (int q = 0; q < n; ++q) {
      const int4 ss = lsrc[q];
      const int4 ww = lwin[q];
      if (rx0 >= ww.y || rx0 + 4 <= ww.x) continue;  // wave-uniform
      const float dy = (float)(ygh - ss.y) * kHf;
      const float dy2 = dy * dy;
      const float gy = (iy >= ww.z && iy < ww.w) ? lw[q] * __expf(-dy2 * cfg.inv2w2) : 0.f;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int ix = rx0 + r;
        if (ix >= ww.x && ix < ww.y) {
          const float dx = (float)(xgh[r] - ss.x) * kHf;
          const float dx2 = dx * dx;
          const float g = (dx2 + dy2 <= cfg.cut2) ? __expf(-dx2 * cfg.inv2w2) * gy : 0.f;
          if (ss.z == 0) a0[r] += g; else a1[r] += g;
        }
      }
    }
    __syncthreads();
  }
  if (iy >= cfg.ny) return;
  const size_t N = (size_t)cfg.nx * cfg.ny;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int ix = rx0 + r;
    if (ix >= cfg.nx) continue;
    const size_t c = (size_t)e * N + (size_t)ix * cfg.ny + iy;
    if (spec.combine == 0) out0[c] = a0[r];
    else if (spec.combine == 1) out0[c] = sqrtf(a0[r] * a1[r]);
    else { out0[c] = a0[r]; out1[c] = a1[r]; }
  }
}

// ----------------------------------------------------------- k_moments
// One workgroup per event.  mode 0: full eps_n record; mode 1: only the
// grid sum (dndy).
__device__ __forceinline__ double block_sum_d(double v, double* red) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wv] = v;
  __syncthreads();
  return red[0] + red[1] + red[2] + red[3];
}

constexpr int kMaxOrd = 9;

__global__ __launch_bounds__(256) void k_moments(DevCfg cfg, EvBufs buf, const float* grid, double* evout,
                                                 int mode) {
  extern __shared__ unsigned mask[];  // nx * nwy words
  __shared__ int rect[2 * kAMax][4];
  __shared__ double red[4];
  __shared__ int ired[4][4];

  const int e = blockIdx.x, t = threadIdx.x;
  const int nx = cfg.nx, ny = cfg.ny, nwy = (ny + 31) >> 5;
  const size_t N = (size_t)nx * ny;
  const float* g = grid + (size_t)e * N;
  double* out = evout + (size_t)e * SMC_EV_STRIDE;
  const int* cnt = buf.counts + (size_t)e * 8;
  const int npa = cnt[0], npb = cnt[1], ncoll = cnt[2];
  const double ff = cfg.final_factor;

  // pass 1: centre of mass (MakeDensity.cpp:2273-2282)
  double sx = 0.0, sy = 0.0, st = 0.0;
  for (size_t c = t; c < N; c += 256) {
    const double d = (double)g[c] * ff;
    if (d == 0.0) continue;
    const int ix = (int)(c / ny), iy = (int)(c % ny);
    sx += (cfg.xmin + ix * cfg.dx) * d;
    sy += (cfg.ymin + iy * cfg.dy) * d;
    st += d;
  }
  const double total = block_sum_d(st, red);
  const double xsum = block_sum_d(sx, red);
  const double ysum = block_sum_d(sy, red);
  if (mode == 1) {
    if (t == 0) out[SMC_EV_TOTAL] = total;
    return;
  }
  const double xc = xsum / total, yc = ysum / total;

  // hot-spot mask (MCnucl.cpp:1303-1323, MakeDensity.cpp:2341-2359)
  const int np = npa + npb;
  const int hbq = cfg.hb_h >> 1;  // box half-width in quanta
  // the CollisionPair boxes are zero boxes at the origin (CollisionPair.h:12,
  // Box2D.cpp:14-22), so with Ncoll > 0 the region always contains (0,0)
  int rxl = (ncoll > 0) ? 0 : 0x7fffffff, rxr = (ncoll > 0) ? 0 : (int)0x80000000;
  int ryl = rxl, ryr = rxr;
  for (int i = t; i < 2 * kAMax; i += 256) {
    const int nu = i / kAMax, j = i % kAMax;
    if (j >= cfg.nuc[nu].A) continue;
    const int wr = buf.wound[(size_t)e * 2 * kAMax + i];
    if (wr < 0) continue;
    const int x = buf.pos[((size_t)e * 2 * kAMax + i) * 2], y = buf.pos[((size_t)e * 2 * kAMax + i) * 2 + 1];
    rxl = min(rxl, x - hbq);
    rxr = max(rxr, x + hbq);
    ryl = min(ryl, y - hbq);
    ryr = max(ryr, y + hbq);
  }
  {
    int v0 = rxl, v1 = ryl, v2 = -rxr, v3 = -ryr;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      v0 = min(v0, __shfl_xor(v0, off, 64));
      v1 = min(v1, __shfl_xor(v1, off, 64));
      v2 = min(v2, __shfl_xor(v2, off, 64));
      v3 = min(v3, __shfl_xor(v3, off, 64));
    }
    if ((t & 63) == 0) { ired[t >> 6][0] = v0; ired[t >> 6][1] = v1; ired[t >> 6][2] = v2; ired[t >> 6][3] = v3; }
    __syncthreads();
    rxl = min(min(ired[0][0], ired[1][0]), min(ired[2][0], ired[3][0]));
    ryl = min(min(ired[0][1], ired[1][1]), min(ired[2][1], ired[3][1]));
    rxr = -min(min(ired[0][2], ired[1][2]), min(ired[2][2], ired[3][2]));
    ryr = -min(min(ired[0][3], ired[1][3]), min(ired[2][3], ired[3][3]));
  }
  const int offx = step_eval(cfg.tab[kOffX], rxl), offy = step_eval(cfg.tab[kOffY], ryl);
  // participant rectangles in grid coordinates
  for (int i = t; i < 2 * kAMax; i += 256) {
    const int nu = i / kAMax, j = i % kAMax;
    int r0 = 0, r1 = 0, r2 = 0, r3 = 0, valid = 0;
    if (j < cfg.nuc[nu].A && buf.wound[(size_t)e * 2 * kAMax + i] >= 0) {
      const int x = buf.pos[((size_t)e * 2 * kAMax + i) * 2], y = buf.pos[((size_t)e * 2 * kAMax + i) * 2 + 1];
      r0 = step_eval(cfg.tab[kMaskX], x - hbq - rxl) + offx;
      r1 = step_eval(cfg.tab[kMaskX], x + hbq - rxl) + offx;
      r2 = step_eval(cfg.tab[kMaskY], y - hbq - ryl) + offy;
      r3 = step_eval(cfg.tab[kMaskX], y + hbq - ryl) + offy;  // divides by dx (MakeDensity.cpp:2357)
      valid = 1;
    }
    rect[i][0] = valid ? max(r0, 0) : 0;
    rect[i][1] = valid ? min(r1, nx) : 0;
    rect[i][2] = valid ? max(r2, 0) : 0;
    rect[i][3] = valid ? min(r3, ny) : 0;
  }
  __syncthreads();
  // rows owned by threads, no atomics
  for (int ix = t; ix < nx; ix += 256) {
    unsigned* row = mask + (size_t)ix * nwy;
    for (int w = 0; w < nwy; ++w) row[w] = 0u;
    for (int i = 0; i < 2 * kAMax; ++i) {
      const int x0 = rect[i][0], x1 = rect[i][1];
      if (ix < x0 || ix >= x1) continue;
      const int y0 = rect[i][2], y1 = rect[i][3];
      for (int y = y0; y < y1;) {
        const int w = y >> 5, b0 = y & 31;
        const int b1 = min(32, b0 + (y1 - y));
        const unsigned bits = (b1 - b0 == 32) ? 0xffffffffu : (((1u << (b1 - b0)) - 1u) << b0);
        row[w] |= bits;
        y += b1 - b0;
      }
    }
  }
  __syncthreads();

  // pass 2: <r^n> over the full grid and masked moments
  const int from = cfg.from_order, to = cfg.to_order;
  double rn[10];
  double mre[kMaxOrd], mim[kMaxOrd], mno[kMaxOrd], pre[kMaxOrd], pim[kMaxOrd], pno[kMaxOrd];
#pragma unroll
  for (int n = 0; n < 10; ++n) rn[n] = 0.0;
#pragma unroll
  for (int o = 0; o < kMaxOrd; ++o) mre[o] = mim[o] = mno[o] = pre[o] = pim[o] = pno[o] = 0.0;
  for (size_t c = t; c < N; c += 256) {
    const double d = (double)g[c] * ff;
    if (d == 0.0) continue;
    const int ix = (int)(c / ny), iy = (int)(c % ny);
    const double x = cfg.xmin + ix * cfg.dx - xc;
    const double y = cfg.ymin + iy * cfg.dy - yc;
    const double r2 = x * x + y * y;
    const double r = sqrt(r2);
    double pw[10];
    pw[0] = 1.0;
#pragma unroll
    for (int n = 1; n < 10; ++n) pw[n] = pw[n - 1] * r;
#pragma unroll
    for (int n = 0; n < 10; ++n) rn[n] += pw[n] * d;
    if (!((mask[(size_t)ix * nwy + (iy >> 5)] >> (iy & 31)) & 1u)) continue;
    if (r == 0.0) continue;  // every r^2 / r^n weighted term vanishes at r = 0
    double zr = x, zi = y;  // z^n = r^n e^{i n theta}
#pragma unroll
    for (int n = 1; n <= kMaxOrd; ++n) {
      if (n >= from && n <= to) {
        const int o = n - 1;
        const double w2 = r2 / pw[n];  // r^2 / r^n
        mre[o] += zr * w2 * d;
        mim[o] += zi * w2 * d;
        mno[o] += r2 * d;
        if (n == 1) {
          pre[o] += r2 * x * d;
          pim[o] += r2 * y * d;
          pno[o] += r2 * r * d;
        } else {
          pre[o] += zr * d;
          pim[o] += zi * d;
          pno[o] += pw[n] * d;
        }
      }
      const double nr = zr * x - zi * y, ni = zr * y + zi * x;
      zr = nr;
      zi = ni;
    }
  }
  double rnt[10];
#pragma unroll
  for (int n = 0; n < 10; ++n) rnt[n] = block_sum_d(rn[n], red);
  if (t == 0) {
    for (int n = 0; n < 10; ++n) out[SMC_EV_RN + n] = rnt[n] / total;
    out[SMC_EV_XC] = xc;
    out[SMC_EV_YC] = yc;
    out[SMC_EV_TOTAL] = total;
    out[SMC_EV_DSDY] = total * cfg.dx * cfg.dy;
    out[SMC_EV_NPART] = np;
    out[SMC_EV_NCOLL] = ncoll;
    out[SMC_EV_B] = buf.bval[e];
    out[SMC_EV_STATUS] = cnt[3];
  }
  const double eps = 1e-15;
#pragma unroll
  for (int o = 0; o < kMaxOrd; ++o) {
    const int n = o + 1;
    const double a = block_sum_d(mre[o], red), b = block_sum_d(mim[o], red), c = block_sum_d(mno[o], red);
    const double p = block_sum_d(pre[o], red), q = block_sum_d(pim[o], red), s = block_sum_d(pno[o], red);
    if (t == 0) {
      const bool on = n >= from && n <= to;
      out[SMC_EV_EPS_RE + n] = on ? -a / (c + eps) : 0.0;
      out[SMC_EV_EPS_IM + n] = on ? -b / (c + eps) : 0.0;
      out[SMC_EV_EPSP_RE + n] = on ? -p / (s + eps) : 0.0;
      out[SMC_EV_EPSP_IM + n] = on ? -q / (s + eps) : 0.0;
    }
  }
  if (t == 0) {
    out[SMC_EV_EPS_RE] = out[SMC_EV_EPS_IM] = out[SMC_EV_EPSP_RE] = out[SMC_EV_EPSP_IM] = 0.0;
  }
}

// ------------------------------------------------------------ launchers
static hipStream_t S(void* s) { return (hipStream_t)s; }

int launch_event(const DevCfg& cfg, const EvBufs& b, int nev, unsigned long long seed,
                 unsigned long long event0, int fed, void* stream) {
  if (nev <= 0) return 0;
  const uint2 key = make_uint2((unsigned)seed, (unsigned)(seed >> 32));
  hipLaunchKernelGGL(k_event, dim3(nev), dim3(256), 0, S(stream), cfg, b, key, event0, fed);
  return (int)hipGetLastError();
}

int launch_smear(const DevCfg& cfg, const EvBufs& b, const SmearSpec& s, int nev, float* out0,
                 float* out1, void* stream) {
  if (nev <= 0) return 0;
  const int ntx = (cfg.nx + kTX - 1) / kTX, nty = (cfg.ny + kTY - 1) / kTY;
  hipLaunchKernelGGL(k_smear, dim3(ntx * nty, nev), dim3(256), 0, S(stream), cfg, b, s, out0, out1);
  return (int)hipGetLastError();
}

int launch_moments(const DevCfg& cfg, const EvBufs& b, const float* grid, int nev, double* ev_out,
                   int mode, void* stream) {
  if (nev <= 0) return 0;
  const size_t shm = (size_t)cfg.nx * ((cfg.ny + 31) / 32) * sizeof(unsigned);
  hipLaunchKernelGGL(k_moments, dim3(nev), dim3(256), shm, S(stream), cfg, b, grid, ev_out, mode);
  return (int)hipGetLastError();
}

}  // namespace smc
