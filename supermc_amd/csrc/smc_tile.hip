// smc_tile.hip -- exact-footprint Gaussian smearing on VALU tiles.
//
// The smearing of MCnucl.cpp:432-478 (T_A / T_B), :689-760 (MC-Glauber
// entropy) and :762-811 (sqrt(T_A T_B)) adds, for every source k inside its
// integer footprint window and inside the disk dx^2 + dy^2 <= 25 w^2
// (Particle.cpp:122-132),
//     T[x][y] += w_k gx_k(x) gy_k(y),   g(d) = exp(-d^2 / (2 w^2)).
//
// The disk cut without a compare.  Inside the disk gx gy >= theta =
// exp(-cut2 / (2 w^2)), outside it gx gy < theta.  The row factor is stored
// as K gx with K = 2^-126 / theta, so the product p = (K gx) gy is a normal
// float exactly when the cell is inside the disk and a denormal otherwise.
// This file is compiled with f32 denormals flushed (the kernel's MODE
// register flushes the outputs of v_mul_f32), so cells outside the disk
// get p = 0 and every cell costs one multiply and one FMA:
//     acc += p * (w_k 2^64),    T = acc * 2^-64 / K
// (2^64 keeps w_k p far above the denormal range for any w_k > 2^-64).
// The boundary is that of the product in float, within a few ulp of the
// exact disk.  Without a cut (cut2 = 1e30) K = 2^64, nothing flushes and
// the footprint is the integer window alone.
//
// One wavefront owns a tile of TR rows x 64 columns in TR accumulators per
// lane (lane = column).  It selects the sources whose window meets the tile
// in source order (ballot compaction, so the summation order is fixed).
// Row factors K gx(row) go through LDS (broadcast float4 reads); with PAIR
// the two half-waves compute them for two sources at once.  Every lane
// computes gy(column) itself, and chunks of CH rows that a source's window
// misses are skipped by a uniform branch.  The launch uses TR = 32, CH = 8,
// PAIR (measured against TR 16/64, CH 4/16/32, unpaired, and a 32-column
// layout with interleaved half-wave rows).
#include "smc_device.h"

namespace smc {

struct TileEntry {
  int xh, yh;
  float w;
  int xw, yw;
  int pad;
};

// The MC-KLN density of MCnucl::setDensity (which_mc_model 1,
// MCnucl.cpp:654-687) fused into the smearing: T_A and T_B accumulate in two
// register sets of one wave, and each cell's dN/dy is the sixPoint2dInterp
// (arsenal.cpp:33-54) of the table at (T_A/dT, T_B/dT) in FP64, from the same
// float T values the separate k_tile + k_kln_table pair reads back.
struct KlnTab {
  const double* table;  // tmax x tmax, row = T_A index
  int tmax;
  double dT;
};

__device__ __forceinline__ float kln_cell(float A, float B, const KlnTab& kt, bool& out_of_table) {
  const double di = (double)A / kt.dT, dj = (double)B / kt.dT;
  if (!(di >= 0 && di < kt.tmax - 2) || !(dj >= 0 && dj < kt.tmax - 2)) {  // also catches NaN
    out_of_table = true;
    return 0.f;
  }
  // row and column 0 of the table are 0, so T_A = 0 or T_B = 0 gives 0
  if (A == 0.f || B == 0.f) return 0.f;
  const int a = (int)floor(di), b = (int)floor(dj);
  const double x = di - a, y = dj - b;
  const double* t = kt.table;
  const int m = kt.tmax;
  const double v00 = t[a * m + b], v01 = t[a * m + b + 1], v02 = t[a * m + b + 2];
  const double v10 = t[(a + 1) * m + b], v11 = t[(a + 1) * m + b + 1];
  const double v20 = t[(a + 2) * m + b];
  const double axx = 1.0 / 2.0 * (v00 - 2 * v10 + v20);
  const double axy = v00 - v01 - v10 + v11;
  const double ayy = 1.0 / 2.0 * (v00 - 2 * v01 + v02);
  const double bx = 1.0 / 2.0 * (-3.0 * v00 + 4 * v10 - v20);
  const double by = 1.0 / 2.0 * (-3.0 * v00 + 4 * v01 - v02);
  return (float)(axx * x * x + axy * x * y + ayy * y * y + bx * x + by * y + v00);
}

// EPI 1 (MC-KLN) and 2 (sqrt(T_A T_B), MCnucl.cpp:762-811): the sources of
// accumulator 0 (T_A), then those of accumulator 1 (T_B), each set in source
// order, and the output is the cell's dN/dy from the table, or
// sqrtf(T_A T_B) of the stored floats.  EPI 0: the sources of accumulator
// acc_sel only.
template <int TR, int CH, bool PAIR, int EPI>
__global__ __launch_bounds__(64) void k_tile(DevCfg cfg, EvBufs buf, SmearSpec spec, int acc_sel, int nev, float* out,
                                             KlnTab kt) {
  static_assert(TR % CH == 0 && CH % 4 == 0 && TR <= 64 && (!PAIR || TR <= 32), "tile rows");
  __shared__ TileEntry wl[64];
  __shared__ float4 rd[2][TR / 4];  // K gx(row) of the current source(s), 0 outside the window
  const int lane = threadIdx.x;
  const int nx = cfg.nx, ny = cfg.ny;
  const int ntx = (nx + TR - 1) / TR;
  // Workgroups go to the 8 XCDs round-robin by linear id; this mapping
  // gives all tiles of an event one XCD (one L2 holds its source list, and
  // the partial cache lines at tile edges meet in that L2).
  const int T = ntx * ((ny + 63) >> 6);
  const unsigned L = blockIdx.x, rg = L % (8u * T);
  const int e = (int)(8u * (L / (8u * T)) + rg % 8u), tile = (int)(rg / 8u);
  if (e >= nev) return;
  const int r0 = TR * (tile % ntx), c0 = 64 * (tile / ntx);
  // exp2 argument -(d_h)^2 kH^2 inv2w2 log2 e with d_h in half-quanta (kH^2 is a power of two)
  const float cq = cfg.inv2w2 * 1.4426950408889634f * (kHf * kHf);
  const float kk = cfg.tile_k;
  const int col = c0 + lane;
  const int ygh = (int)llrint((cfg.ymin + col * cfg.dy) / kH);
  const int rl = PAIR ? (lane & 31) : lane;     // row this lane computes a factor for
  const int rs = PAIR ? (lane >> 5) : 0;        // which of the pair
  const int myrow = r0 + (rl < TR ? rl : 0);
  const int rowh = (int)llrint((cfg.xmin + myrow * cfg.dx) / kH);
  const int K = buf.counts[(size_t)e * 8 + 5];
  const int4* src = buf.src + (size_t)e * buf.src_cap;
  const int4* swin = buf.swin + (size_t)e * buf.src_cap;
  float* rdf = reinterpret_cast<float*>(&rd[0][0]);

  constexpr bool TWO = EPI != 0;
  float acc[TR], accA[TWO ? TR : 1];  // accA: T_A while the second pass fills acc with T_B
#pragma unroll
  for (int r = 0; r < TR; ++r) acc[r] = 0.f;
  // tiles outside the event's source rectangle stay zero
  const int rx = buf.counts[(size_t)e * 8 + 6], ry = buf.counts[(size_t)e * 8 + 7];
  const bool live = (rx & 0xffff) < r0 + TR && (rx >> 16) > r0 && (ry & 0xffff) < c0 + 64 && (ry >> 16) > c0;
  if (!live && spec.sparse) {  // nothing to store: zero centre-of-mass partials only
    if (spec.cm_out && lane < 4) spec.cm_out[((size_t)e * T + tile) * 4 + lane] = 0.0;
    return;
  }

  auto fma_phase = [&](const float4* rr, float gy, float wq, int xw) {
    const int xlo = xw & 0xffff, xhi = xw >> 16;
#pragma unroll
    for (int c = 0; c < TR / CH; ++c) {
      // chunks of CH rows that the source's window misses are skipped
      if (CH == TR || (r0 + CH * c < xhi && r0 + CH * c + CH > xlo)) {
#pragma unroll
        for (int i = c * CH / 4; i < (c + 1) * CH / 4; ++i) {
          const float4 v = rr[i];
          acc[4 * i] = fmaf(v.x * gy, wq, acc[4 * i]);
          acc[4 * i + 1] = fmaf(v.y * gy, wq, acc[4 * i + 1]);
          acc[4 * i + 2] = fmaf(v.z * gy, wq, acc[4 * i + 2]);
          acc[4 * i + 3] = fmaf(v.w * gy, wq, acc[4 * i + 3]);
        }
      }
    }
  };
  auto col_factor = [&](const TileEntry& q) {
    const int ylo = q.yw & 0xffff, yhi = q.yw >> 16;
    const float dy = (float)(ygh - q.yh);
    return (col >= ylo && col < yhi) ? __builtin_amdgcn_exp2f(-(dy * dy) * cq) : 0.f;
  };

  for (int pass = 0; pass < (TWO ? 2 : 1); ++pass) {
   if (TWO && pass == 1) {
#pragma unroll
    for (int r = 0; r < TR; ++r) {
      accA[TWO ? r : 0] = acc[r];
      acc[r] = 0.f;
    }
   }
   const int want = TWO ? pass : acc_sel;
   for (int sub = 0; live && sub < K; sub += 64) {
    const int k = sub + lane;
    bool sel = false;
    TileEntry en{0, 0, 0.f, 0, 0, 0};
    if (k < K) {
      const int4 s = src[k];
      const int kind = s.w;
      if (spec.enabled[kind] && spec.acc[kind] == want) {
        const int4 wv = swin[k];
        en.xh = s.x;
        en.yh = s.y;
        en.xw = (spec.win[kind] == kWinBox) ? wv.x : wv.z;
        en.yw = (spec.win[kind] == kWinBox) ? wv.y : wv.w;
        en.w = spec.scale[kind] * (spec.use_fluct[kind] ? __int_as_float(s.z) : 1.f) * cfg.tile_w;
        const int xlo = en.xw & 0xffff, xhi = en.xw >> 16, ylo = en.yw & 0xffff, yhi = en.yw >> 16;
        sel = xlo < r0 + TR && xhi > r0 && ylo < c0 + 64 && yhi > c0;
      }
    }
    const unsigned long long m = __ballot(sel);
    if (sel) wl[__popcll(m & ((1ull << lane) - 1ull))] = en;
    const int n = __popcll(m);
    __builtin_amdgcn_wave_barrier();
    for (int g = 0; g < n; g += (PAIR ? 2 : 1)) {
      // row factors: lanes rl < TR of half rs for source g + rs
      {
        const int gi = (PAIR && g + rs < n) ? g + rs : g;
        const TileEntry q = wl[gi];
        if (rl < TR) {
          const float dx = (float)(rowh - q.xh);
          const int xlo = q.xw & 0xffff, xhi = q.xw >> 16;
          const float gx = kk * __builtin_amdgcn_exp2f(-(dx * dx) * cq);
          rdf[rs * TR + rl] = (myrow >= xlo && myrow < xhi) ? gx : 0.f;
        }
      }
      const TileEntry q0 = wl[g];
      const float gy0 = col_factor(q0);
      const int xw0 = __builtin_amdgcn_readfirstlane(q0.xw);
      __builtin_amdgcn_wave_barrier();
      fma_phase(rd[0], gy0, q0.w, xw0);
      if constexpr (PAIR) {
        if (g + 1 < n) {
          const TileEntry q1 = wl[g + 1];
          const float gy1 = col_factor(q1);
          const int xw1 = __builtin_amdgcn_readfirstlane(q1.xw);
          fma_phase(rd[1], gy1, q1.w, xw1);
        }
      }
      __builtin_amdgcn_wave_barrier();
    }
    __builtin_amdgcn_wave_barrier();
   }
  }
  const float so = cfg.tile_out;
  if constexpr (EPI == 2) {
    // sqrt(T_A T_B) of the stored floats; the product is formed in double,
    // so the result does not depend on the f32 denormal mode of this file
#pragma unroll
    for (int r = 0; r < TR; ++r) acc[r] = (float)sqrt((double)(accA[r] * so) * (double)(acc[r] * so));
  }
  if constexpr (EPI == 1) {  // dN/dy from (T_A, T_B) = (accA, acc) as stored floats
    bool off = false;
#pragma unroll
    for (int r = 0; r < TR; ++r)
      acc[r] = (live && col < ny && r0 + r < nx) ? kln_cell(accA[r] * so, acc[r] * so, kt, off) : 0.f;
    if (off) atomicOr(&buf.counts[(size_t)e * 8 + 3], kStKlnTable);
  }
  const float os = TWO ? 1.f : so;  // stored value = acc * os
  if (spec.cm_out) {  // centre-of-mass partials of the stored grid values
    double s0 = 0.0, s1 = 0.0;
#pragma unroll
    for (int r = 0; r < TR; ++r) {
      if (col < ny && r0 + r < nx) {
        const double v = (double)(acc[r] * os);
        s0 += v;
        s1 = fma(cfg.xmin + (r0 + r) * cfg.dx, v, s1);
      }
    }
    double s2 = (col < ny) ? (cfg.ymin + col * cfg.dy) * s0 : 0.0;
    s0 = wave_sum(s0);
    s1 = wave_sum(s1);
    s2 = wave_sum(s2);
    if (lane == 0) {
      double* p = spec.cm_out + ((size_t)e * T + tile) * 4;
      p[0] = s0;
      p[1] = s1;
      p[2] = s2;
      p[3] = 0.0;
    }
  }
  if (col >= ny) return;
  const size_t N = (size_t)nx * ny;
#pragma unroll
  for (int r = 0; r < TR; ++r) {
    const int ix = r0 + r;
    if (ix < nx) out[(size_t)e * N + (size_t)ix * ny + col] = acc[r] * os;
  }
}

constexpr int kTileRows = 32;

// Tile shapes the autotuner chooses from (rows x 64 columns, row chunk,
// paired row factors).  Every shape adds the same products in the same
// source order per cell, so the grids do not depend on the choice.
//   0: 32 rows, chunks of 8, paired   (the default)
//   1: 16 rows, chunks of 8, paired
//   2: 32 rows, chunks of 4, paired
//   3: 64 rows, chunks of 8, unpaired
constexpr int kVariantRows[kTileVariants] = {32, 16, 32, 64};

static int units_for_rows(const DevCfg& cfg, int rows) { return ((cfg.nx + rows - 1) / rows) * ((cfg.ny + 63) / 64); }
static int variant_of(const DevCfg& cfg) {
  return (cfg.tile_variant >= 0 && cfg.tile_variant < kTileVariants) ? cfg.tile_variant : 0;
}
int tile_units(const DevCfg& cfg) { return units_for_rows(cfg, kVariantRows[variant_of(cfg)]); }
int tile_units_max(const DevCfg& cfg) {
  int m = 0;
  for (int v = 0; v < kTileVariants; ++v) m = max(m, units_for_rows(cfg, kVariantRows[v]));
  return m;
}

template <int TR, int CH, bool PAIR>
static void launch_tr(const DevCfg& cfg, const EvBufs& b, const SmearSpec& s, int nev, float* out0, float* out1,
                      hipStream_t st) {
  // one dimension: tiles x events padded to a multiple of the 8 XCDs
  const dim3 grid((unsigned)(((cfg.nx + TR - 1) / TR) * ((cfg.ny + 63) / 64)) * (unsigned)((nev + 7) / 8 * 8));
  const KlnTab none{nullptr, 0, 1.0};
  if (s.combine == 1) {  // sqrt(T_A T_B) in one launch, with the partials of the combined grid
    hipLaunchKernelGGL((k_tile<TR, CH, PAIR, 2>), grid, dim3(64), 0, st, cfg, b, s, 0, nev, out0, none);
    return;
  }
  SmearSpec s1 = s;  // partials only for a single accumulator
  if (s.combine != 0) s1.cm_out = nullptr;
  hipLaunchKernelGGL((k_tile<TR, CH, PAIR, 0>), grid, dim3(64), 0, st, cfg, b, s1, 0, nev, out0, none);
  if (s.combine != 0)
    hipLaunchKernelGGL((k_tile<TR, CH, PAIR, 0>), grid, dim3(64), 0, st, cfg, b, s1, 1, nev, out1, none);
}

template <int TR, int CH, bool PAIR>
static void launch_tr_kln(const DevCfg& cfg, const EvBufs& b, const SmearSpec& s, int nev, float* out,
                          const KlnTab& kt, hipStream_t st) {
  const dim3 grid((unsigned)(((cfg.nx + TR - 1) / TR) * ((cfg.ny + 63) / 64)) * (unsigned)((nev + 7) / 8 * 8));
  hipLaunchKernelGGL((k_tile<TR, CH, PAIR, 1>), grid, dim3(64), 0, st, cfg, b, s, 0, nev, out, kt);
}

// The fused MC-KLN launch in the ctx's tile shape (the same shapes as
// launch_tile; its centre-of-mass partials have tile_units(cfg) units).
int launch_tile_kln(const DevCfg& cfg, const EvBufs& b, const SmearSpec& s, int nev, float* out,
                    const double* table, int tmax, double dT, void* stream) {
  if (nev <= 0) return 0;
  if (!(cfg.tile_k > 0.f && cfg.tile_w > 0.f && cfg.tile_out > 0.f) || !table || tmax < 3 || !(dT > 0))
    return (int)hipErrorInvalidValue;
  const KlnTab kt{table, tmax, dT};
  hipStream_t st = (hipStream_t)stream;
  switch (variant_of(cfg)) {
    case 1: launch_tr_kln<16, 8, true>(cfg, b, s, nev, out, kt, st); break;
    case 2: launch_tr_kln<32, 4, true>(cfg, b, s, nev, out, kt, st); break;
    case 3: launch_tr_kln<64, 8, false>(cfg, b, s, nev, out, kt, st); break;
    default: launch_tr_kln<kTileRows, 8, true>(cfg, b, s, nev, out, kt, st); break;
  }
  return (int)hipGetLastError();
}

// Cells outside each event's source rectangle set to zero (the grid of a
// sparse launch, completed before whole grids are read).
__global__ __launch_bounds__(256) void k_zero_outside(float* grid, const int* counts, int nx, int ny) {
  const int e = blockIdx.y;
  const int rx = counts[(size_t)e * 8 + 6], ry = counts[(size_t)e * 8 + 7];
  const int x0 = rx & 0xffff, x1 = rx >> 16, y0 = ry & 0xffff, y1 = ry >> 16;
  float* g = grid + (size_t)e * nx * ny;
  const int n = nx * ny;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    const int ix = i / ny, iy = i - ix * ny;
    if (ix < x0 || ix >= x1 || iy < y0 || iy >= y1) g[i] = 0.f;
  }
}

int launch_zero_outside(float* g, const int* counts, int nx, int ny, int nev, void* stream) {
  if (nev <= 0) return 0;
  const int per = (nx * ny + 255) / 256;
  const dim3 grid((unsigned)(per < 64 ? per : 64), (unsigned)nev);
  hipLaunchKernelGGL(k_zero_outside, grid, dim3(256), 0, (hipStream_t)stream, g, counts, nx, ny);
  return (int)hipGetLastError();
}

int launch_tile(const DevCfg& cfg, const EvBufs& b, const SmearSpec& s, int nev, float* out0, float* out1,
                void* stream) {
  if (nev <= 0) return 0;
  if (!(cfg.tile_k > 0.f && cfg.tile_w > 0.f && cfg.tile_out > 0.f)) return (int)hipErrorInvalidValue;
  hipStream_t st = (hipStream_t)stream;
  switch (variant_of(cfg)) {
    case 1: launch_tr<16, 8, true>(cfg, b, s, nev, out0, out1, st); break;
    case 2: launch_tr<32, 4, true>(cfg, b, s, nev, out0, out1, st); break;
    case 3: launch_tr<64, 8, false>(cfg, b, s, nev, out0, out1, st); break;
    default: launch_tr<kTileRows, 8, true>(cfg, b, s, nev, out0, out1, st); break;
  }
  return (int)hipGetLastError();
}

}  // namespace smc
