// smc_moments.hip -- eps_n moments of per-event grids (gfx950).
//
// One 512-thread workgroup per event (MakeDensity.cpp:2244-2430,
// MCnucl.cpp:1303-1323): FP64 centre of mass over the full grid, <r^n>
// (n = 0..9) over the full grid, and the r^2- and r^n-weighted eps_n sums
// over the hot-spot mask built from the participant boxes.  Grid rows are
// prefetched one row ahead; the mask is an LDS bitmap filled with one wave
// per participant box.
#include "smc_device.h"

namespace smc {

constexpr int kGW = 8;            // waves per workgroup
constexpr int kGT = kGW * 64;     // threads per workgroup

// LDS of the moment phase
struct MomLds {
  int rect[2 * kAMax][4];  // participant mask rectangles (grid cells)
  double red[kGW][64];     // per-wave partial sums
  double tot[64];          // block sums
  int ired[kGW][4];
};

// Sum NV doubles per thread over a workgroup of NW waves into L.tot[0..NV),
// in a fixed order (deterministic).
template <int NW, int NV>
__device__ __forceinline__ void block_sums(double (&v)[NV], MomLds& L) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  __syncthreads();
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    const double r = wave_sum(v[i]);
    if (lane == 0) L.red[wv][i] = r;
    __builtin_amdgcn_sched_barrier(0);  // one value in flight at a time
  }
  __syncthreads();
  if (threadIdx.x < NV) {
    double s = 0.0;
    for (int q = 0; q < NW; ++q) s += L.red[q][threadIdx.x];
    L.tot[threadIdx.x] = s;
  }
  __syncthreads();
}

template <int NORD>
struct MomAcc {
  double rn[10];
  double mre[NORD], mim[NORD], mno[NORD], pre[NORD], pim[NORD], pno[NORD];
};

// One cell of the moment sums (MakeDensity.cpp:2285-2423).  cos/sin(n theta)
// come from z^n = (x + i y)^n = r^n e^{i n theta}.
template <int NORD>
__device__ __forceinline__ void moment_cell(MomAcc<NORD>& M, int from, int to, double d, double x, double y,
                                            bool masked) {
  const double r2 = x * x + y * y;
  const double r = sqrt(r2);
  double pw = d;
#pragma unroll
  for (int n = 0; n < 10; ++n) {
    M.rn[n] += pw;
    pw *= r;
  }
  if (!masked || r == 0.0) return;  // every r^2 / r^n weighted term vanishes at r = 0
  const double inv_r = 1.0 / r;
  double zr = x, zi = y;  // z^n
  double irp = 1.0;       // r^(2-n), n >= 3
  double rp = r;          // r^n
#pragma unroll
  for (int n = 1; n <= 9; ++n) {
#pragma unroll
    for (int o = 0; o < NORD; ++o) {
      if (n != from + o || n > to) continue;
      const double wr = (n == 1) ? r : ((n == 2) ? 1.0 : irp);  // r^2 / r^n
      M.mre[o] += zr * wr * d;
      M.mim[o] += zi * wr * d;
      M.mno[o] += r2 * d;
      if (n == 1) {  // r^3 weight for n = 1 (MakeDensity.cpp:2408-2413)
        M.pre[o] += r2 * x * d;
        M.pim[o] += r2 * y * d;
        M.pno[o] += r2 * r * d;
      } else {
        M.pre[o] += zr * d;
        M.pim[o] += zi * d;
        M.pno[o] += rp * d;
      }
    }
    if (n >= 2) irp *= inv_r;
    rp *= r;
    const double nr = zr * x - zi * y, ni = zr * y + zi * x;
    zr = nr;
    zi = ni;
  }
}

// Rectangle [x0, x1) x [y0, y1) of grid cells that holds every non-zero
// cell of event e (the union of its source windows, from k_event).
struct Rect {
  int x0, x1, y0, y1;
};
__device__ __forceinline__ Rect event_rect(const EvBufs& buf, int e) {
  const int a = buf.counts[(size_t)e * 8 + 6], b = buf.counts[(size_t)e * 8 + 7];
  return Rect{a & 0xffff, a >> 16, b & 0xffff, b >> 16};
}

// Loads row ix of the grid inside the rectangle, MAXC values per lane
// (columns y0 + lane + 64 j); zero outside it.
template <int MAXC>
__device__ __forceinline__ void load_row(const float* g, int ix, const Rect& R, int ny, float (&v)[MAXC]) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int j = 0; j < MAXC; ++j) {
    const int iy = R.y0 + lane + 64 * j;
    v[j] = (ix < R.x1 && iy < R.y1) ? g[(size_t)ix * ny + iy] : 0.f;
  }
}

// Hot-spot mask of event e as an LDS bitmap mask[nx][(ny+31)/32]
// (MCnucl.cpp:1303-1323, MakeDensity.cpp:2341-2359).  The caller orders the
// atomicOr writes before any read with a barrier.
template <int NW>
__device__ void build_mask(const DevCfg& cfg, const EvBufs& buf, int e, unsigned* mask, MomLds& L) {
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const int nx = cfg.nx, ny = cfg.ny;
  const int ncoll = buf.counts[(size_t)e * 8 + 2];
  const int nwy = (ny + 31) >> 5;
  // union
  // of participant boxes in region coordinates.  CollisionPair boxes are
  // zero boxes at the origin (CollisionPair.h:12, Box2D.cpp:14-22).
  for (int i = t; i < nx * nwy; i += NW * 64) mask[i] = 0u;
  const int hbq = cfg.hb_h >> 1;
  int rxl = (ncoll > 0) ? 0 : 0x7fffffff, rxr = (ncoll > 0) ? 0 : (int)0x80000000;
  int ryl = rxl, ryr = rxr;
  const int* pos = buf.pos + (size_t)e * 2 * kAMax * 2;
  const int* wound = buf.wound + (size_t)e * 2 * kAMax;
  const int i0 = t;  // one nucleon slot per thread (2 * kAMax <= NW * 64)
  const bool part = i0 < 2 * kAMax && (i0 % kAMax) < cfg.nuc[i0 / kAMax].A && wound[i0] >= 0;
  int px = 0, py = 0;
  if (part) {
    px = pos[2 * i0];
    py = pos[2 * i0 + 1];
    rxl = px - hbq;
    rxr = px + hbq;
    ryl = py - hbq;
    ryr = py + hbq;
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
    if (lane == 0) { L.ired[wv][0] = v0; L.ired[wv][1] = v1; L.ired[wv][2] = v2; L.ired[wv][3] = v3; }
    __syncthreads();
    rxl = ryl = 0x7fffffff;
    int nrxr = 0x7fffffff, nryr = 0x7fffffff;
    for (int q = 0; q < NW; ++q) {
      rxl = min(rxl, L.ired[q][0]);
      ryl = min(ryl, L.ired[q][1]);
      nrxr = min(nrxr, L.ired[q][2]);
      nryr = min(nryr, L.ired[q][3]);
    }
    rxr = -nrxr;
    ryr = -nryr;
  }
  // participant rectangles in grid cells, compacted (OR is order-free)
  {
    const int offx = step_eval(cfg.tab[kOffX], rxl), offy = step_eval(cfg.tab[kOffY], ryl);
    const unsigned long long m = __ballot(part);
    if (lane == 0) L.ired[wv][0] = __popcll(m);
    __syncthreads();
    int base = 0, np = 0;
    for (int q = 0; q < NW; ++q) {
      if (q < wv) base += L.ired[q][0];
      np += L.ired[q][0];
    }
    if (part) {
      const int k = base + __popcll(m & ((1ull << lane) - 1ull));
      L.rect[k][0] = max(step_eval(cfg.tab[kMaskX], px - hbq - rxl) + offx, 0);
      L.rect[k][1] = min(step_eval(cfg.tab[kMaskX], px + hbq - rxl) + offx, nx);
      L.rect[k][2] = max(step_eval(cfg.tab[kMaskY], py - hbq - ryl) + offy, 0);
      L.rect[k][3] = min(step_eval(cfg.tab[kMaskX], py + hbq - ryl) + offy, ny);  // /dx (MakeDensity.cpp:2357)
    }
    __syncthreads();
    // one wave per rectangle, one lane per row
    for (int k = wv; k < np; k += NW) {
      const int x0 = L.rect[k][0], x1 = L.rect[k][1], y0 = L.rect[k][2], y1 = L.rect[k][3];
      for (int ix = x0 + lane; ix < x1; ix += 64) {
        for (int y = y0; y < y1;) {
          const int w = y >> 5, b0 = y & 31;
          const int b1 = min(32, b0 + (y1 - y));
          const unsigned bits = (b1 - b0 == 32) ? 0xffffffffu : (((1u << (b1 - b0)) - 1u) << b0);
          atomicOr(&mask[(size_t)ix * nwy + w], bits);
          y += b1 - b0;
        }
      }
    }
  }
}

// ------------------------------------------------------------------ moments
// Three steps, all deterministic (fixed summation orders):
//   centre of mass: per-unit partials {sum v, sum x v, sum y v} of the grid,
//     units = the smearing kernel's tiles (written as it stores the grid) or
//     kNB row bands of k_mom_cm;
//   k_mom_band: one wavefront per (event, row band b: rows R.x0 + b + kNB i)
//     builds the hot-spot mask of its rows in LDS and accumulates <r^n> and
//     the masked eps_n sums (MakeDensity.cpp:2285-2423) about that centre;
//   k_mom_final: one wavefront per event adds the kNB band partials and
//     writes the event record.
constexpr int kNB = 8;       // row bands per event
constexpr int kNVMax = 64;   // partial sums per band (10 + 6 * 9)
// rows of a band loaded ahead of the one being summed (Pb+Pb 261^2, 10^5
// events: 1 row 7.41 ms, 2 rows 7.16 ms, 3 rows 8.06 ms, same records)
constexpr int kMomPF = 2;

// Centre of mass of event e from its nu unit partials.
__device__ __forceinline__ void event_cm(const DevCfg& cfg, const double* cmp, int nu, int e, double& total,
                                         double& xc, double& yc) {
  const int lane = threadIdx.x & 63;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  for (int u = lane; u < nu; u += 64) {
    const double* p = cmp + ((size_t)e * nu + u) * 4;
    s0 += p[0];
    s1 += p[1];
    s2 += p[2];
  }
  s0 = wave_sum(s0);
  s1 = wave_sum(s1);
  s2 = wave_sum(s2);
  total = s0 * cfg.final_factor;  // MakeDensity.cpp:2273-2282 with setSd's finalFactor
  xc = s1 / s0;
  yc = s2 / s0;
}

// Grid-sum units of k_mom_cm: one wavefront per (row band, event).  SEG:
// the rectangle is taken in column segments of 64 MAXC (grids wider than
// one segment); otherwise a single pass over its first 64 MAXC columns.
template <int MAXC, int SEG>
__global__ __launch_bounds__(64) void k_mom_cm(DevCfg cfg, EvBufs buf, const float* grid, double* cmp) {
  const int b = blockIdx.x, e = blockIdx.y, lane = threadIdx.x;
  const float* g = grid + (size_t)e * cfg.nx * cfg.ny;
  const Rect R = event_rect(buf, e);
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  for (int cs = R.y0;; cs += 64 * MAXC) {
    Rect Rs = R;
    if (SEG) {
      Rs.y0 = cs;
      Rs.y1 = min(R.y1, cs + 64 * MAXC);
    }
    float cur[MAXC], nxt[MAXC];
    load_row<MAXC>(g, Rs.x0 + b, Rs, cfg.ny, cur);
    for (int ix = Rs.x0 + b; ix < Rs.x1; ix += kNB) {
      load_row<MAXC>(g, ix + kNB, Rs, cfg.ny, nxt);
      double rs = 0.0, ys = 0.0;
#pragma unroll
      for (int j = 0; j < MAXC; ++j) {
        const double d = (double)cur[j];
        rs += d;
        ys += (cfg.ymin + (Rs.y0 + lane + 64 * j) * cfg.dy) * d;
      }
      s0 += rs;
      s1 += (cfg.xmin + ix * cfg.dx) * rs;
      s2 += ys;
#pragma unroll
      for (int j = 0; j < MAXC; ++j) cur[j] = nxt[j];
    }
    if (!SEG || Rs.y1 >= R.y1) break;
  }
  s0 = wave_sum(s0);
  s1 = wave_sum(s1);
  s2 = wave_sum(s2);
  if (lane == 0) {
    double* p = cmp + ((size_t)e * kNB + b) * 4;
    p[0] = s0;
    p[1] = s1;
    p[2] = s2;
    p[3] = 0.0;
  }
}

// Hot-spot mask rectangles (MCnucl.cpp:1303-1323, MakeDensity.cpp:2341-2359)
// of every event, one 256-thread workgroup per event: the union box of the
// participant boxes (with the zero boxes of the collision pairs at the
// origin, CollisionPair.h:12, Box2D.cpp:14-22) fixes the offsets, then each
// participant's box becomes grid cells [x0, x1) x [y0, y1), packed as
// {x0 | x1 << 16, y0 | y1 << 16}; empty boxes are dropped.
__global__ __launch_bounds__(256) void k_mask_rects(DevCfg cfg, EvBufs buf, uint2* rects, int* nrect) {
  __shared__ int red[4][2];
  __shared__ int wcnt[4];
  const int e = blockIdx.x, t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const int nx = cfg.nx, ny = cfg.ny;
  const int ncoll = buf.counts[(size_t)e * 8 + 2];
  const int hbq = cfg.hb_h >> 1;
  const int* pos = buf.pos + (size_t)e * 2 * kAMax * 2;
  const int* wound = buf.wound + (size_t)e * 2 * kAMax;
  int rxl = (ncoll > 0) ? 0 : 0x7fffffff, ryl = rxl;
  for (int i = t; i < 2 * kAMax; i += 256) {
    if ((i % kAMax) < cfg.nuc[i / kAMax].A && wound[i] >= 0) {
      rxl = min(rxl, pos[2 * i] - hbq);
      ryl = min(ryl, pos[2 * i + 1] - hbq);
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    rxl = min(rxl, __shfl_xor(rxl, off, 64));
    ryl = min(ryl, __shfl_xor(ryl, off, 64));
  }
  if (lane == 0) {
    red[wv][0] = rxl;
    red[wv][1] = ryl;
  }
  __syncthreads();
  rxl = min(min(red[0][0], red[1][0]), min(red[2][0], red[3][0]));
  ryl = min(min(red[0][1], red[1][1]), min(red[2][1], red[3][1]));
  const int offx = step_eval(cfg.tab[kOffX], rxl), offy = step_eval(cfg.tab[kOffY], ryl);
  uint2* out = rects + (size_t)e * 2 * kAMax;
  int base = 0;
  for (int i0 = 0; i0 < 2 * kAMax; i0 += 256) {
    const int i = i0 + t;
    int x0 = 0, x1 = 0, y0 = 0, y1 = 0;
    if (i < 2 * kAMax && (i % kAMax) < cfg.nuc[i / kAMax].A && wound[i] >= 0) {
      const int px = pos[2 * i], py = pos[2 * i + 1];
      x0 = max(step_eval(cfg.tab[kMaskX], px - hbq - rxl) + offx, 0);
      x1 = min(step_eval(cfg.tab[kMaskX], px + hbq - rxl) + offx, nx);
      y0 = max(step_eval(cfg.tab[kMaskY], py - hbq - ryl) + offy, 0);
      y1 = min(step_eval(cfg.tab[kMaskX], py + hbq - ryl) + offy, ny);  // /dx (MakeDensity.cpp:2357)
    }
    const bool keep = x0 < x1 && y0 < y1;
    const unsigned long long m = __ballot(keep);
    __syncthreads();  // wcnt free
    if (lane == 0) wcnt[wv] = __popcll(m);
    __syncthreads();
    int k = base;
    for (int q = 0; q < wv; ++q) k += wcnt[q];
    if (keep) out[k + __popcll(m & ((1ull << lane) - 1ull))] = make_uint2(x0 | (x1 << 16), y0 | (y1 << 16));
    base += wcnt[0] + wcnt[1] + wcnt[2] + wcnt[3];
  }
  if (t == 0) nrect[e] = base;
}

// Hot-spot mask bits of the rows R.x0 + b + kNB lr, lr < nrow, as
// bm[lr][nwy], from the event's rectangles (OR is order-free).
__device__ void band_mask(const DevCfg& cfg, const uint2* rects, int np, const Rect& R, int b, int nrow,
                          unsigned* bm) {
  const int lane = threadIdx.x & 63;
  const int nwy = (cfg.ny + 31) >> 5;
  for (int i = lane; i < nrow * nwy; i += 64) bm[i] = 0u;
  __syncthreads();  // mask cleared
  for (int k = lane; k < np; k += 64) {
    const uint2 rc = rects[k];
    const int x0 = max((int)(rc.x & 0xffff), R.x0), x1 = min((int)(rc.x >> 16), R.x1);
    const int y0 = rc.y & 0xffff, y1 = rc.y >> 16;
    if (x0 >= x1) continue;
    for (int lr = (x0 - R.x0 - b + kNB - 1) / kNB; R.x0 + b + kNB * lr < x1; ++lr) {
      for (int y = y0; y < y1;) {
        const int w = y >> 5, b0 = y & 31;
        const int b1 = min(32, b0 + (y1 - y));
        const unsigned bits = (b1 - b0 == 32) ? 0xffffffffu : (((1u << (b1 - b0)) - 1u) << b0);
        atomicOr(&bm[lr * nwy + w], bits);
        y += b1 - b0;
      }
    }
  }
  __syncthreads();  // mask complete
}

// One cell of the fixed-order (FROM..TO, FROM >= 2) sums: r and 1/r from
// one reciprocal square root, d r^n shared between <r^n> and the
// r^n-weighted sums, one r^2 normalisation for all orders.
//   a[0..9] <r^n> numerators, a[10] sum_mask r^2 d, per order n:
//   a[11 + 5 (n - FROM) + {0,1}] = sum_mask z^n r^(2-n) d, {2,3} = z^n d,
//   {4} = r^n d (z = x + i y about the centre); for n = 1 (FROM = 1) the
//   reference weights z r d, z r^2 d and r^3 d instead.
template <int FROM, int TO>
__device__ __forceinline__ void cell_fixed(double* a, double x, double y, double d, bool masked) {
  const double r2 = fma(x, x, y * y);
  // 1/r: the hardware estimate and one Newton step (relative error ~1e-16)
  const double y0 = __builtin_amdgcn_rsq(r2);
  const double e0 = fma(-(r2 * y0), y0, 1.0);
  const double rs = (r2 > 0.0) ? fma(0.5 * y0, e0, y0) : 0.0;
  const double r = r2 * rs;
  double pw[TO + 1];
  pw[0] = d;
#pragma unroll
  for (int n = 1; n <= TO; ++n) pw[n] = pw[n - 1] * r;
  double q = pw[TO];
#pragma unroll
  for (int n = 0; n <= TO; ++n) a[n] += pw[n];
#pragma unroll
  for (int n = TO + 1; n < 10; ++n) {
    q *= r;
    a[n] += q;
  }
  if (masked && r2 > 0.0) {
    a[10] = fma(r2, d, a[10]);
    if constexpr (FROM == 1) {  // n = 1: z r and z r^2, r^3 (MakeDensity.cpp:2408-2413)
      double* o = a + 11;
      const double rd = r * d, r2d = r2 * d;
      o[0] = fma(x, rd, o[0]);
      o[1] = fma(y, rd, o[1]);
      o[2] = fma(x, r2d, o[2]);
      o[3] = fma(y, r2d, o[3]);
      o[4] += pw[3];
    }
    double zr = x, zi = y, dw = d;  // z^n and d r^(2-n)
#pragma unroll
    for (int n = 2; n <= TO; ++n) {
      const double nr = fma(zr, x, -zi * y), ni = fma(zr, y, zi * x);
      zr = nr;
      zi = ni;
      if (n > 2) dw *= rs;
      if (n >= FROM) {
        double* o = a + 11 + 5 * (n - FROM);
        o[0] = fma(zr, dw, o[0]);
        o[1] = fma(zi, dw, o[1]);
        o[2] = fma(zr, d, o[2]);
        o[3] = fma(zi, d, o[3]);
        o[4] += pw[n];
      }
    }
  }
}

// FIX = 0: any order range (moment_cell); FIX = f > 0: orders f .. f + NORD - 1
// fixed at compile time (cell_fixed), with one shared r^2 denominator.
template <int NORD, int FIX>
struct BandLayout {
  static constexpr int NV = FIX ? 11 + 5 * NORD : 10 + 6 * NORD;
};

// Pass over one row band of one event; SEG as in k_mom_cm.
template <int NORD, int MAXC, int FIX, int SEG>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(FIX ? (NORD > 4 ? 2 : (MAXC <= 5 ? 4 : 3)) : 1, 8))) void k_mom_band(
    DevCfg cfg, EvBufs buf, const float* grid, const double* cmp, int nu, const uint2* rects, const int* nrect,
    double* band) {
  extern __shared__ unsigned bm[];
  constexpr int NV = BandLayout<NORD, FIX>::NV;
  static_assert(NV <= kNVMax, "band partials");
  const int b = blockIdx.x, e = blockIdx.y, lane = threadIdx.x;
  const int nx = cfg.nx, ny = cfg.ny, nwy = (ny + 31) >> 5;
  const float* g = grid + (size_t)e * nx * ny;
  const Rect R = event_rect(buf, e);
  double total, xc, yc;
  event_cm(cfg, cmp, nu, e, total, xc, yc);
  const int nrow = max(0, (R.x1 - R.x0 - b + kNB - 1) / kNB);
  band_mask(cfg, rects + (size_t)e * 2 * kAMax, nrect[e], R, b, nrow, bm);
  const double ff = cfg.final_factor;
  const int from = cfg.from_order, to = cfg.to_order;
  double a[NV];
#pragma unroll
  for (int i = 0; i < NV; ++i) a[i] = 0.0;
  MomAcc<NORD> M;
  if constexpr (!FIX) {
#pragma unroll
    for (int n = 0; n < 10; ++n) M.rn[n] = 0.0;
#pragma unroll
    for (int o = 0; o < NORD; ++o) M.mre[o] = M.mim[o] = M.mno[o] = M.pre[o] = M.pim[o] = M.pno[o] = 0.0;
  }
  // cell values enter without finalFactor; the sums are scaled once below
  for (int cs = R.y0;; cs += 64 * MAXC) {
    Rect Rs = R;
    if (SEG) {
      Rs.y0 = cs;
      Rs.y1 = min(R.y1, cs + 64 * MAXC);
    }
    double yv[MAXC];
#pragma unroll
    for (int j = 0; j < MAXC; ++j) yv[j] = cfg.ymin + (Rs.y0 + lane + 64 * j) * cfg.dy - yc;
    // the band's rows in flight: q[0] is summed while q[1..kMomPF] load
    float q[kMomPF + 1][MAXC];
#pragma unroll
    for (int k = 0; k < kMomPF; ++k) load_row<MAXC>(g, Rs.x0 + b + kNB * k, Rs, ny, q[k]);
    for (int lr = 0; lr < nrow; ++lr) {
      const int ix = Rs.x0 + b + kNB * lr;
      load_row<MAXC>(g, ix + kNB * kMomPF, Rs, ny, q[kMomPF]);
      const double x = cfg.xmin + ix * cfg.dx - xc;
      const unsigned* row = bm + lr * nwy;
#pragma unroll
      for (int j = 0; j < MAXC; ++j) {
        if (q[0][j] != 0.f) {
          const int iy = Rs.y0 + lane + 64 * j;
          const bool masked = (row[iy >> 5] >> (iy & 31)) & 1u;
          if constexpr (FIX != 0)
            cell_fixed<FIX, FIX + NORD - 1>(a, x, yv[j], (double)q[0][j], masked);
          else
            moment_cell<NORD>(M, from, to, (double)q[0][j], x, yv[j], masked);
        }
      }
#pragma unroll
      for (int k = 0; k < kMomPF; ++k)
#pragma unroll
        for (int j = 0; j < MAXC; ++j) q[k][j] = q[k + 1][j];
    }
    if (!SEG || Rs.y1 >= R.y1) break;
  }
  if constexpr (!FIX) {
#pragma unroll
    for (int n = 0; n < 10; ++n) a[n] = M.rn[n];
#pragma unroll
    for (int o = 0; o < NORD; ++o) {
      a[10 + 6 * o + 0] = M.mre[o];
      a[10 + 6 * o + 1] = M.mim[o];
      a[10 + 6 * o + 2] = M.mno[o];
      a[10 + 6 * o + 3] = M.pre[o];
      a[10 + 6 * o + 4] = M.pim[o];
      a[10 + 6 * o + 5] = M.pno[o];
    }
  }
  double* out = band + ((size_t)e * kNB + b) * kNVMax;
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    const double r = wave_sum(a[i] * ff);
    if (lane == 0) out[i] = r;
    __builtin_amdgcn_sched_barrier(0);  // one value in flight at a time
  }
}

// Event record from the band sums T (layout of k_mom_band).
template <int NORD, int FIX>
__device__ void write_record(const DevCfg& cfg, const EvBufs& buf, int e, const double* T, double total, double xc,
                             double yc, double* evout) {
  const int* cnt = buf.counts + (size_t)e * 8;
  double* out = evout + (size_t)e * SMC_EV_STRIDE;
  for (int i = 0; i < 40; ++i) out[i] = 0.0;
  for (int n = 0; n < 10; ++n) out[SMC_EV_RN + n] = T[n] / total;
  const double eps = 1e-15;
  if constexpr (FIX != 0) {
    for (int n = FIX; n < FIX + NORD; ++n) {
      if (n < cfg.from_order || n > cfg.to_order) continue;  // sums of a wider fixed range
      const double* o = T + 11 + 5 * (n - FIX);
      out[SMC_EV_EPS_RE + n] = -o[0] / (T[10] + eps);
      out[SMC_EV_EPS_IM + n] = -o[1] / (T[10] + eps);
      out[SMC_EV_EPSP_RE + n] = -o[2] / (o[4] + eps);
      out[SMC_EV_EPSP_IM + n] = -o[3] / (o[4] + eps);
    }
  } else {
    for (int o = 0; o < NORD; ++o) {
      const int n = cfg.from_order + o;
      if (n > cfg.to_order) break;
      out[SMC_EV_EPS_RE + n] = -T[10 + 6 * o + 0] / (T[10 + 6 * o + 2] + eps);
      out[SMC_EV_EPS_IM + n] = -T[10 + 6 * o + 1] / (T[10 + 6 * o + 2] + eps);
      out[SMC_EV_EPSP_RE + n] = -T[10 + 6 * o + 3] / (T[10 + 6 * o + 5] + eps);
      out[SMC_EV_EPSP_IM + n] = -T[10 + 6 * o + 4] / (T[10 + 6 * o + 5] + eps);
    }
  }
  out[SMC_EV_XC] = xc;
  out[SMC_EV_YC] = yc;
  out[SMC_EV_TOTAL] = total;
  out[SMC_EV_DSDY] = total * cfg.dx * cfg.dy;
  out[SMC_EV_NPART] = cnt[0] + cnt[1];
  out[SMC_EV_NCOLL] = cnt[2];
  out[SMC_EV_B] = buf.bval[e];
  out[SMC_EV_STATUS] = cnt[3];
  for (int i = SMC_EV_STATUS + 1; i < SMC_EV_STRIDE; ++i) out[i] = 0.0;  // reserved slots
}

template <int NORD, int FIX>
__global__ __launch_bounds__(64) void k_mom_final(DevCfg cfg, EvBufs buf, const double* cmp, int nu,
                                                  const double* band, double* evout) {
  __shared__ double T[kNVMax];
  constexpr int NV = BandLayout<NORD, FIX>::NV;
  const int e = blockIdx.x, lane = threadIdx.x;
  double total, xc, yc;
  event_cm(cfg, cmp, nu, e, total, xc, yc);
  if (lane < NV) {
    double s = 0.0;
    for (int b = 0; b < kNB; ++b) s += band[((size_t)e * kNB + b) * kNVMax + lane];
    T[lane] = s;
  }
  __syncthreads();
  if (lane == 0) write_record<NORD, FIX>(cfg, buf, e, T, total, xc, yc, evout);
}

// Sum of the grid only (dndy, MCnucl.cpp:817-820), no finalFactor.
__global__ __launch_bounds__(kGT) void k_grid_sum(DevCfg cfg, EvBufs buf, const float* grid, double* evout) {
  __shared__ MomLds L;
  const int e = blockIdx.x;
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const float* g = grid + (size_t)e * cfg.nx * cfg.ny;
  double s[1] = {0.0};
  const Rect R = event_rect(buf, e);
  for (int ix = R.x0 + wv; ix < R.x1; ix += kGW)
    for (int iy = R.y0 + lane; iy < R.y1; iy += 64) s[0] += (double)g[(size_t)ix * cfg.ny + iy];
  block_sums<kGW, 1>(s, L);
  if (t == 0) evout[(size_t)e * SMC_EV_STRIDE + SMC_EV_TOTAL] = L.tot[0];
}

// Differentiable batches: per-event FP64 scalars for |eps_2|, |eps_3| and
// their adjoint (layout kDs* in smc_internal.h), the forward outputs, and
// the hot-spot mask (frozen at theta0: built from the stored positions).
//   A_n = sum_mask rho z^n r^(2-n), N = sum_mask rho r^2, z = (x-xc) + i(y-yc)
//   Mx, My = sum_mask rho x, rho y;  B3x, B3y = sum_mask rho d(z^3/r)/dx, /dy
__global__ __launch_bounds__(kGT) void k_diff_moments(DevCfg cfg, EvBufs buf, DiffBufs d, const float* grid,
                                                      double ffdxdy, double* out) {
  __shared__ MomLds L;
  extern __shared__ unsigned dyn_mask[];
  const int e = blockIdx.x;
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const int nx = cfg.nx, ny = cfg.ny, nwy = (ny + 31) >> 5;
  const float* g = grid + (size_t)e * nx * ny;
  // the density is zero outside the event's source rectangle (and not
  // stored there): each thread keeps its cells of the whole grid's stride
  // pattern inside the rectangle only, so every sum adds the same non-zero
  // terms in the same order
  const int rx = buf.counts[(size_t)e * 8 + 6], ry = buf.counts[(size_t)e * 8 + 7];
  const int x0 = rx & 0xffff, x1 = min(rx >> 16, nx), y0 = ry & 0xffff, y1 = min(ry >> 16, ny);
  const int ix0 = wv + (x0 > wv ? (x0 - wv + kGW - 1) / kGW * kGW : 0);
  const int iy0 = lane + (y0 > lane ? (y0 - lane + 63) / 64 * 64 : 0);
  double s1[3] = {0.0, 0.0, 0.0};
  for (int ix = ix0; ix < x1; ix += kGW) {
    const double xg = cfg.xmin + ix * cfg.dx;
    for (int iy = iy0; iy < y1; iy += 64) {
      const double v = g[(size_t)ix * ny + iy];
      s1[0] += v;
      s1[1] += xg * v;
      s1[2] += (cfg.ymin + iy * cfg.dy) * v;
    }
  }
  build_mask<kGW>(cfg, buf, e, dyn_mask, L);
  block_sums<kGW, 3>(s1, L);
  const double S = L.tot[0], xc = L.tot[1] / S, yc = L.tot[2] / S;
  double a[11];
#pragma unroll
  for (int i = 0; i < 11; ++i) a[i] = 0.0;
  for (int ix = ix0; ix < x1; ix += kGW) {
    const double x = cfg.xmin + ix * cfg.dx - xc;
    const unsigned* row = dyn_mask + (size_t)ix * nwy;
    for (int iy = iy0; iy < y1; iy += 64) {
      const double v = g[(size_t)ix * ny + iy];
      if (v == 0.0 || !((row[iy >> 5] >> (iy & 31)) & 1u)) continue;
      const double y = cfg.ymin + iy * cfg.dy - yc;
      const double r2 = x * x + y * y;
      if (r2 == 0.0) continue;
      const double ir = 1.0 / sqrt(r2), ir3 = ir * ir * ir;
      const double z2r = x * x - y * y, z2i = 2.0 * x * y;
      const double z3r = z2r * x - z2i * y, z3i = z2r * y + z2i * x;
      a[0] += v * z2r;
      a[1] += v * z2i;
      a[2] += v * z3r * ir;
      a[3] += v * z3i * ir;
      a[4] += v * r2;
      a[5] += v * x;
      a[6] += v * y;
      a[7] += v * (3.0 * z2r * ir - z3r * x * ir3);
      a[8] += v * (3.0 * z2i * ir - z3i * x * ir3);
      a[9] += v * (-3.0 * z2i * ir - z3r * y * ir3);
      a[10] += v * (3.0 * z2r * ir - z3i * y * ir3);
    }
  }
  block_sums<kGW, 11>(a, L);
  if (t == 0) {
    const double* T = L.tot;
    double* sc = d.dsc + (size_t)e * kDsStride;
    sc[kDsS] = S;
    sc[kDsXc] = xc;
    sc[kDsYc] = yc;
    sc[kDsA2r] = T[0];
    sc[kDsA2i] = T[1];
    sc[kDsA3r] = T[2];
    sc[kDsA3i] = T[3];
    sc[kDsN] = T[4];
    sc[kDsMx] = T[5];
    sc[kDsMy] = T[6];
    sc[kDsB3xr] = T[7];
    sc[kDsB3xi] = T[8];
    sc[kDsB3yr] = T[9];
    sc[kDsB3yi] = T[10];
    const double N = T[4] + 1e-15;
    double* o = out + (size_t)e * kDiffOut;
    o[0] = sqrt(T[0] * T[0] + T[1] * T[1]) / N;
    o[1] = sqrt(T[2] * T[2] + T[3] * T[3]) / N;
    o[2] = S * ffdxdy;
    o[3] = -T[0] / N;
    o[4] = -T[1] / N;
    o[5] = -T[2] / N;
    o[6] = -T[3] / N;
    o[7] = buf.counts[(size_t)e * 8 + 3];
  }
  for (int i = t; i < nx * nwy; i += kGT) d.mask[(size_t)e * nx * nwy + i] = dyn_mask[i];
}

int launch_diff_moments(const DevCfg& cfg, const EvBufs& b, const DiffBufs& d, const float* grid, int nev,
                        double ffdxdy, double* out, void* stream) {
  if (nev <= 0) return 0;
  const size_t shm = (size_t)cfg.nx * ((cfg.ny + 31) / 32) * sizeof(unsigned);
  hipLaunchKernelGGL(k_diff_moments, dim3(nev), dim3(kGT), shm, (hipStream_t)stream, cfg, b, d, grid, ffdxdy, out);
  return (int)hipGetLastError();
}

int launch_moments(const DevCfg& cfg, const EvBufs& b, const float* grid, int nev, double* ev_out, int mode,
                   const double* cm_units, int nu, double* cm_scratch, double* band, void* rects, int* nrect,
                   void* stream) {
  if (nev <= 0) return 0;
  hipStream_t st = (hipStream_t)stream;
  if (mode != 0) {
    hipLaunchKernelGGL(k_grid_sum, dim3(nev), dim3(kGT), 0, st, cfg, b, grid, ev_out);
    return (int)hipGetLastError();
  }
  // rectangles up to 320 columns in one pass of 5 columns per lane, up to
  // 576 in one pass of 9, wider ones in segments of 576
  const int width = cfg.ny <= 320 ? 0 : (cfg.ny <= 64 * 9 ? 1 : 2);
  if (!cm_units) {  // centre-of-mass partials from the grid itself
    const int err = launch_cm_partials(cfg, b, grid, nev, cm_scratch, stream);
    if (err) return err;
    cm_units = cm_scratch;
    nu = kNB;
  }
  uint2* rc = static_cast<uint2*>(rects);
  hipLaunchKernelGGL(k_mask_rects, dim3(nev), dim3(256), 0, st, cfg, b, rc, nrect);
  const int nord = cfg.to_order - cfg.from_order + 1;
  const size_t shm = (size_t)((cfg.nx + kNB - 1) / kNB) * ((cfg.ny + 31) / 32) * sizeof(unsigned);
  const dim3 gb(kNB, nev);
#define SMC_BAND(NORD_, FIX_)                                                                                       \
  do {                                                                                                              \
    if (width == 0)                                                                                                 \
      hipLaunchKernelGGL((k_mom_band<NORD_, 5, FIX_, 0>), gb, dim3(64), shm, st, cfg, b, grid, cm_units, nu, rc,      \
                         nrect, band);                                                                              \
    else if (width == 1)                                                                                            \
      hipLaunchKernelGGL((k_mom_band<NORD_, 9, FIX_, 0>), gb, dim3(64), shm, st, cfg, b, grid, cm_units, nu, rc,      \
                         nrect, band);                                                                              \
    else                                                                                                            \
      hipLaunchKernelGGL((k_mom_band<NORD_, 9, FIX_, 1>), gb, dim3(64), shm, st, cfg, b, grid, cm_units, nu, rc,      \
                         nrect, band);                                                                              \
  } while (0)
  // fixed-order forms for orders 2..5 (the eccentricity table) and 1..9
  // (the parameter file's default, and every range of more than four
  // orders, whose records keep only the requested ones); ranges of up to
  // four other orders take the general form
  if (cfg.from_order == 2 && cfg.to_order == 5) {
    SMC_BAND(4, 2);
    hipLaunchKernelGGL((k_mom_final<4, 2>), dim3(nev), dim3(64), 0, st, cfg, b, cm_units, nu, band, ev_out);
  } else if (nord > 4) {
    SMC_BAND(9, 1);
    hipLaunchKernelGGL((k_mom_final<9, 1>), dim3(nev), dim3(64), 0, st, cfg, b, cm_units, nu, band, ev_out);
  } else {
    SMC_BAND(4, 0);
    hipLaunchKernelGGL((k_mom_final<4, 0>), dim3(nev), dim3(64), 0, st, cfg, b, cm_units, nu, band, ev_out);
  }
#undef SMC_BAND
  return (int)hipGetLastError();
}

int launch_cm_partials(const DevCfg& cfg, const EvBufs& b, const float* grid, int nev, double* cm, void* stream) {
  if (nev <= 0) return 0;
  hipStream_t st = (hipStream_t)stream;
  if (cfg.ny <= 320) hipLaunchKernelGGL((k_mom_cm<5, 0>), dim3(kNB, nev), dim3(64), 0, st, cfg, b, grid, cm);
  else if (cfg.ny <= 64 * 9) hipLaunchKernelGGL((k_mom_cm<9, 0>), dim3(kNB, nev), dim3(64), 0, st, cfg, b, grid, cm);
  else hipLaunchKernelGGL((k_mom_cm<9, 1>), dim3(kNB, nev), dim3(64), 0, st, cfg, b, grid, cm);
  return (int)hipGetLastError();
}

int moments_band_doubles() { return kNB * kNVMax; }
int moments_cm_units_min() { return kNB; }

}  // namespace smc
