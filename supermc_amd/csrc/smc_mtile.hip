// smc_mtile.hip -- Gaussian smearing as the separable contraction
// T = Gx . diag(w) . Gy^T on MFMA, with the reference's exact footprint.
//
// The smearing of MCnucl.cpp:432-478 (T_A / T_B), :689-760 (MC-Glauber
// entropy) and :762-811 (sqrt(T_A T_B)) adds, for every source k inside its
// integer footprint window and inside the disk dx^2 + dy^2 <= 25 w^2
// (Particle.cpp:122-132),
//     T[x][y] += w_k gx_k(x) gy_k(y),   g(d) = exp(-d^2 / (2 w^2)).
// Over the square window the sum is a matrix product: rows of Gx are the
// sources' row factors, columns of Gy^T their weighted column factors.  The
// disk is not a product, so the kernel does two things per source:
//   1. MFMA over the square window (all sources of a group at once);
//   2. a VALU pass that subtracts, cell by cell, the products the MFMA added
//      outside the disk.  It visits only the 4 x 16 patches that hold such
//      cells (flagged per source by a distance test, (source, patch) pairs
//      spread over the lanes), and decides each cell by
//      gx gy < exp(-cut2 / (2 w^2)) in float, within a few ulp of the FP64
//      disk -- the cells there carry e^-12.5 of a source's peak.
// So the grids are those of the exact footprint, to the parity bars, on any
// grid width.
//
// Layout.  One wavefront owns a 32 x 64 tile (the tile grid and the
// event-to-XCD mapping of k_tile, so the centre-of-mass partials have the
// same units) as 2 x 4 MFMA blocks of 16 x 16.  The MFMA row index i of a
// block is mapped to the physical row 4 (i % 4) + i / 4, which makes each
// accumulator register v of lane l the cell (4 v + l / 16, l % 16): a
// contiguous 4 x 16 patch, the unit of the exact-cut pass.
//
// Precision (PREC):
//   0  FP32: v_mfma_f32_16x16x4_f32, 4 sources per instruction;
//   1  BF16x3: factors split f = hi + lo in bf16, three products
//      Ah Bl + Al Bh + Ah Bh per v_mfma_f32_16x16x32_bf16 (32 sources);
//      each term carries up to 3 * 2^-16 relative error;
//   2  BF16x6: f = hi + mid + lo, six products (all with at most one
//      non-leading piece, plus mid mid); each term within about 2^-22.
// All accumulate in FP32.  Operands carry powers of two (SA, SB) that keep
// every piece a normal number.
#include "smc_device.h"

namespace smc {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

struct MtEntry {
  int xh, yh;
  float w;  // weight (scale x fluct)
  int xw, yw;
  int pad;
};

constexpr float kMtSA = 0x1p32f;  // row factor scale
constexpr float kMtSB = 0x1p32f;  // weighted column factor scale

// The same table interpolation as k_tile's MC-KLN epilogue (arsenal.cpp:33-54)
__device__ __forceinline__ float mt_kln_cell(float A, float B, const double* t, int m, double dT, bool& off) {
  const double di = (double)A / dT, dj = (double)B / dT;
  if (!(di >= 0 && di < m - 2) || !(dj >= 0 && dj < m - 2)) {
    off = true;
    return 0.f;
  }
  if (A == 0.f || B == 0.f) return 0.f;
  const int a = (int)floor(di), b = (int)floor(dj);
  const double x = di - a, y = dj - b;
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

// bf16 pieces of v: hi = bf16(v), then the pieces of the remainders
template <int NP>
__device__ __forceinline__ void mt_split(float v, __bf16 (&p)[NP]) {
#pragma unroll
  for (int q = 0; q < NP; ++q) {
    p[q] = (__bf16)v;
    v -= (float)p[q];
  }
}

template <int PREC, int EPI>
__global__ __launch_bounds__(64) void k_mtile(DevCfg cfg, EvBufs buf, SmearSpec spec, int acc_sel, int nev,
                                              float* out, const double* ktab, int ktmax, double kdT) {
  constexpr int G = PREC == 0 ? 4 : 32;     // sources per MFMA group
  constexpr int NP = PREC == 2 ? 3 : 2;     // bf16 pieces per factor
  constexpr bool TWO = EPI != 0;
  __shared__ MtEntry wl[64];
  // factors for the exact-cut pass, one row per physical row / column, a
  // padded stride so that the 16 columns of a read fall in distinct banks
  constexpr int GS = G + 1;
  __shared__ float la[2][16][GS];  // SA gx(row): [row block][physical row][source]
  __shared__ float lg[4][16][GS];  // gy(col):    [column block][column][source]
  __shared__ unsigned lmask[G];    // patches (ri * 16 + v * 4 + ci) of each source that need the cut
  const int lane = threadIdx.x, q = lane >> 4, j16 = lane & 15;
  const int nx = cfg.nx, ny = cfg.ny;
  const int ntx = (nx + 31) / 32;
  const int T = ntx * ((ny + 63) >> 6);
  const unsigned L = blockIdx.x, rg = L % (8u * T);
  const int e = (int)(8u * (L / (8u * T)) + rg % 8u), tile = (int)(rg / 8u);
  if (e >= nev) return;
  const int r0 = 32 * (tile % ntx), c0 = 64 * (tile / ntx);
  const int rx = buf.counts[(size_t)e * 8 + 6], ry = buf.counts[(size_t)e * 8 + 7];
  const bool live = (rx & 0xffff) < r0 + 32 && (rx >> 16) > r0 && (ry & 0xffff) < c0 + 64 && (ry >> 16) > c0;
  if (!live && spec.sparse) {
    if (spec.cm_out && lane < 4) spec.cm_out[((size_t)e * T + tile) * 4 + lane] = 0.0;
    return;
  }
  // exp2 argument -(d_h)^2 kH^2 inv2w2 log2 e with d_h in half-quanta
  const float cq = cfg.inv2w2 * 1.4426950408889634f * (kHf * kHf);
  // disk threshold: SA gx gy < thr  <=>  outside (cut2 in fm^2, huge without a cut)
  const float thr = kMtSA * __builtin_amdgcn_exp2f(-cfg.cut2 * cfg.inv2w2 * 1.4426950408889634f);
  const float cut2 = cfg.cut2;
  const int K = buf.counts[(size_t)e * 8 + 5];
  const int4* src = buf.src + (size_t)e * buf.src_cap;
  const int4* swin = buf.swin + (size_t)e * buf.src_cap;
  // this lane's MFMA A row (physical row of block row index i = lane % 16)
  const int prow = 4 * (j16 & 3) + (j16 >> 2);
  int rowh[2], colh[4];
#pragma unroll
  for (int ri = 0; ri < 2; ++ri) rowh[ri] = (int)llrint((cfg.xmin + (r0 + 16 * ri + prow) * cfg.dx) / kH);
#pragma unroll
  for (int ci = 0; ci < 4; ++ci) colh[ci] = (int)llrint((cfg.ymin + (c0 + 16 * ci + j16) * cfg.dy) / kH);

  v4f acc[2][4], accA[TWO ? 2 : 1][TWO ? 4 : 1];
#pragma unroll
  for (int ri = 0; ri < 2; ++ri)
#pragma unroll
    for (int ci = 0; ci < 4; ++ci) acc[ri][ci] = v4f{0.f, 0.f, 0.f, 0.f};

  for (int pass = 0; pass < (TWO ? 2 : 1); ++pass) {
    if (TWO && pass == 1) {
#pragma unroll
      for (int ri = 0; ri < 2; ++ri)
#pragma unroll
        for (int ci = 0; ci < 4; ++ci) {
          accA[TWO ? ri : 0][TWO ? ci : 0] = acc[ri][ci];
          acc[ri][ci] = v4f{0.f, 0.f, 0.f, 0.f};
        }
    }
    const int want = TWO ? pass : acc_sel;
    for (int sub = 0; live && sub < K; sub += 64) {
      // sources of this accumulator whose window meets the tile, in source order
      const int k = sub + lane;
      bool sel = false;
      MtEntry en{0, 0, 0.f, 0, 0, 0};
      if (k < K) {
        const int4 s = src[k];
        const int kind = s.w;
        if (spec.enabled[kind] && spec.acc[kind] == want) {
          const int4 wv = swin[k];
          en.xh = s.x;
          en.yh = s.y;
          en.xw = (spec.win[kind] == kWinBox) ? wv.x : wv.z;
          en.yw = (spec.win[kind] == kWinBox) ? wv.y : wv.w;
          en.w = spec.scale[kind] * (spec.use_fluct[kind] ? __int_as_float(s.z) : 1.f);
          const int xlo = en.xw & 0xffff, xhi = en.xw >> 16, ylo = en.yw & 0xffff, yhi = en.yw >> 16;
          sel = xlo < r0 + 32 && xhi > r0 && ylo < c0 + 64 && yhi > c0;
        }
      }
      const unsigned long long m = __ballot(sel);
      if (sel) wl[__popcll(m & ((1ull << lane) - 1ull))] = en;
      const int n = __popcll(m);
      __builtin_amdgcn_wave_barrier();
      for (int g = 0; g < n; g += G) {
        // ---- operands: row factors (A) and weighted column factors (B); the
        // unweighted factors also go to LDS for the exact-cut pass
        if constexpr (PREC == 0) {
          const int s = g + q;
          const bool ok = s < n;
          const MtEntry en2 = wl[ok ? s : 0];
          const int xlo = en2.xw & 0xffff, xhi = en2.xw >> 16, ylo = en2.yw & 0xffff, yhi = en2.yw >> 16;
          float a[2], b[4];
          unsigned rm = 0u, cmk = 0u;
#pragma unroll
          for (int ri = 0; ri < 2; ++ri) {
            const int row = r0 + 16 * ri + prow;
            const float dx = (float)(rowh[ri] - en2.xh);
            a[ri] = (ok && row >= xlo && row < xhi) ? kMtSA * __builtin_amdgcn_exp2f(-(dx * dx) * cq) : 0.f;
            la[ri][prow][q] = a[ri];
            if (__ballot(a[ri] != 0.f)) rm |= 1u << ri;
          }
#pragma unroll
          for (int ci = 0; ci < 4; ++ci) {
            const int col = c0 + 16 * ci + j16;
            const float dy = (float)(colh[ci] - en2.yh);
            const float gy = (ok && col >= ylo && col < yhi) ? __builtin_amdgcn_exp2f(-(dy * dy) * cq) : 0.f;
            lg[ci][j16][q] = gy;
            b[ci] = gy * (en2.w * kMtSB);
            if (__ballot(b[ci] != 0.f)) cmk |= 1u << ci;
          }
#pragma unroll
          for (int ri = 0; ri < 2; ++ri)
#pragma unroll
            for (int ci = 0; ci < 4; ++ci)
              if (((rm >> ri) & 1u) && ((cmk >> ci) & 1u))
                acc[ri][ci] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[ri], b[ci], acc[ri][ci], 0, 0, 0);
        } else {
          // lane (i = lane % 16, q = lane / 16) holds sources g + 8 q + t, t < 8
          float va[2][8], vb[4][8], wt[8];
#pragma unroll
          for (int t = 0; t < 8; ++t) {
            const int s = g + 8 * q + t;
            const bool ok = s < n;
            const MtEntry en2 = wl[ok ? s : 0];
            const int xlo = en2.xw & 0xffff, xhi = en2.xw >> 16, ylo = en2.yw & 0xffff, yhi = en2.yw >> 16;
            wt[t] = en2.w * kMtSB;
#pragma unroll
            for (int ri = 0; ri < 2; ++ri) {
              const int row = r0 + 16 * ri + prow;
              const float dx = (float)(rowh[ri] - en2.xh);
              va[ri][t] = (ok && row >= xlo && row < xhi) ? kMtSA * __builtin_amdgcn_exp2f(-(dx * dx) * cq) : 0.f;
            }
#pragma unroll
            for (int ci = 0; ci < 4; ++ci) {
              const int col = c0 + 16 * ci + j16;
              const float dy = (float)(colh[ci] - en2.yh);
              vb[ci][t] = (ok && col >= ylo && col < yhi) ? __builtin_amdgcn_exp2f(-(dy * dy) * cq) : 0.f;
            }
          }
          v8bf ap[2][NP];
          unsigned rm = 0u;
#pragma unroll
          for (int ri = 0; ri < 2; ++ri) {
#pragma unroll
            for (int t = 0; t < 8; ++t) la[ri][prow][8 * q + t] = va[ri][t];
            bool any = false;
#pragma unroll
            for (int t = 0; t < 8; ++t) {
              any |= va[ri][t] != 0.f;
              __bf16 pc[NP];
              mt_split<NP>(va[ri][t], pc);
#pragma unroll
              for (int p = 0; p < NP; ++p) ap[ri][p][t] = pc[p];
            }
            if (__ballot(any)) rm |= 1u << ri;
          }
#pragma unroll
          for (int ci = 0; ci < 4; ++ci) {
#pragma unroll
            for (int t = 0; t < 8; ++t) lg[ci][j16][8 * q + t] = vb[ci][t];
            v8bf bp[NP];
            bool any = false;
#pragma unroll
            for (int t = 0; t < 8; ++t) {
              const float v = vb[ci][t] * wt[t];
              any |= v != 0.f;
              __bf16 pc[NP];
              mt_split<NP>(v, pc);
#pragma unroll
              for (int p = 0; p < NP; ++p) bp[p][t] = pc[p];
            }
            if (!__ballot(any) || !rm) continue;
#pragma unroll
            for (int ri = 0; ri < 2; ++ri) {
              if (!((rm >> ri) & 1u)) continue;
              v4f d = acc[ri][ci];
              if constexpr (NP == 3) {  // smallest terms first
                d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ap[ri][2], bp[0], d, 0, 0, 0);
                d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ap[ri][0], bp[2], d, 0, 0, 0);
                d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ap[ri][1], bp[1], d, 0, 0, 0);
              }
              d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ap[ri][1], bp[0], d, 0, 0, 0);
              d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ap[ri][0], bp[1], d, 0, 0, 0);
              d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ap[ri][0], bp[0], d, 0, 0, 0);
              acc[ri][ci] = d;
            }
          }
        }
        // ---- which 4 x 16 patches of which source hold cells of its square
        // window outside its disk: lane-parallel, (source, patch) pairs spread
        // over the lanes; a patch whose farthest cell (with a margin) lies
        // inside the disk needs nothing
        const int ng = min(G, n - g);
        {
          constexpr int PER = 32 * G / 64;  // pairs per lane
          // G = 32: lane l takes source l / 2, patches 16 (l % 2) + k, k < 16;
          // G = 4: lane l takes source l / 16, patches l % 16 and l % 16 + 16
          const int s = PER == 2 ? (lane >> 4) : (lane >> 1);
          const int p0 = PER == 2 ? (lane & 15) : 16 * (lane & 1), pstep = PER == 2 ? 16 : 1;
          unsigned bits = 0u;
          if (s < ng) {
            const MtEntry es = wl[g + s];
            const int xlo = es.xw & 0xffff, xhi = es.xw >> 16, ylo = es.yw & 0xffff, yhi = es.yw >> 16;
            const float sx = (float)((double)es.xh * kH), sy = (float)((double)es.yh * kH);
#pragma unroll
            for (int k = 0; k < PER; ++k) {
              const int pp = p0 + k * pstep, ri = pp >> 4, v = (pp >> 2) & 3, ci = pp & 3;
              const int pr0 = max(r0 + 16 * ri + 4 * v, xlo), pr1 = min(r0 + 16 * ri + 4 * v + 4, xhi);
              const int pc0 = max(c0 + 16 * ci, ylo), pc1 = min(c0 + 16 * ci + 16, yhi);
              if (pr0 < pr1 && pc0 < pc1) {
                const float fx0 = (float)(cfg.xmin + pr0 * cfg.dx) - sx, fx1 = (float)(cfg.xmin + (pr1 - 1) * cfg.dx) - sx;
                const float fy0 = (float)(cfg.ymin + pc0 * cfg.dy) - sy, fy1 = (float)(cfg.ymin + (pc1 - 1) * cfg.dy) - sy;
                if (fmaxf(fx0 * fx0, fx1 * fx1) + fmaxf(fy0 * fy0, fy1 * fy1) > cut2 * 0.9999f) bits |= 1u << k;
              }
            }
          }
          if constexpr (PER == 2) {  // bit i of source s's mask sits in lane 16 s + i of b0 (i < 16) or b1
            const unsigned long long b0 = __ballot(bits & 1u), b1 = __ballot(bits & 2u);
            if (lane < G) lmask[lane] = (unsigned)((b0 >> (16 * lane)) & 0xffffu) | ((unsigned)((b1 >> (16 * lane)) & 0xffffu) << 16);
          } else {  // G = 32: lanes 2 s and 2 s + 1 hold source s's low and high 16 patches
            const unsigned hi = __shfl_xor(bits, 1, 64);
            if (!(lane & 1)) lmask[lane >> 1] = bits | (hi << 16);
          }
        }
        __builtin_amdgcn_wave_barrier();
        // ---- exact cut: subtract, in the flagged patches, the products the
        // MFMA added outside the disk
        for (int s = 0; s < ng; ++s) {
          const unsigned mk = __builtin_amdgcn_readfirstlane(lmask[s]);
          if (!mk) continue;
          const float wsb = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(wl[g + s].w))) * kMtSB;
#pragma unroll
          for (int pp = 0; pp < 32; ++pp) {
            if (!((mk >> pp) & 1u)) continue;
            const int ri = pp >> 4, v = (pp >> 2) & 3, ci = pp & 3;
            const float t = la[ri][4 * v + q][s] * lg[ci][j16][s];
            acc[ri][ci][v] -= (t < thr) ? t * wsb : 0.f;
          }
        }
        __builtin_amdgcn_wave_barrier();
      }
    }
  }

  // ---- epilogue: stored value = acc * so per cell.  Where no disk reaches
  // the true value is 0 and the square-window sum minus its subtraction
  // leaves rounding noise of either sign (a few ulp of products below
  // e^-12.5 of a peak): sums are clamped at 0 from below, so that sqrt and
  // the table see non-negative thicknesses.
  const float so = 1.f / (kMtSA * kMtSB);
  bool off = false;
#pragma unroll
  for (int ri = 0; ri < 2; ++ri)
#pragma unroll
    for (int ci = 0; ci < 4; ++ci)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        float x = fmaxf(acc[ri][ci][v] * so, 0.f);
        if constexpr (EPI == 2) {
          x = (float)sqrt((double)fmaxf(accA[TWO ? ri : 0][TWO ? ci : 0][v] * so, 0.f) * (double)x);
        } else if constexpr (EPI == 1) {
          const int row = r0 + 16 * ri + 4 * v + q, col = c0 + 16 * ci + j16;
          x = (live && row < nx && col < ny)
                  ? mt_kln_cell(fmaxf(accA[TWO ? ri : 0][TWO ? ci : 0][v] * so, 0.f), x, ktab, ktmax, kdT, off)
                  : 0.f;
        }
        acc[ri][ci][v] = x;
      }
  if constexpr (EPI == 1) {
    if (off) atomicOr(&buf.counts[(size_t)e * 8 + 3], kStKlnTable);
  }
  if (spec.cm_out) {  // centre-of-mass partials of the stored values
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
#pragma unroll
    for (int ci = 0; ci < 4; ++ci) {
      const int col = c0 + 16 * ci + j16;
      if (col >= ny) continue;
      double cs = 0.0;
#pragma unroll
      for (int ri = 0; ri < 2; ++ri)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int row = r0 + 16 * ri + 4 * v + q;
          if (row < nx) {
            const double d = (double)acc[ri][ci][v];
            cs += d;
            s1 = fma(cfg.xmin + row * cfg.dx, d, s1);
          }
        }
      s0 += cs;
      s2 = fma(cfg.ymin + col * cfg.dy, cs, s2);
    }
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
  const size_t N = (size_t)nx * ny;
#pragma unroll
  for (int ri = 0; ri < 2; ++ri)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int row = r0 + 16 * ri + 4 * v + q;
      if (row >= nx) continue;
#pragma unroll
      for (int ci = 0; ci < 4; ++ci) {
        const int col = c0 + 16 * ci + j16;
        if (col < ny) out[(size_t)e * N + (size_t)row * ny + col] = acc[ri][ci][v];
      }
    }
}

int mtile_units(const DevCfg& cfg) { return ((cfg.nx + 31) / 32) * ((cfg.ny + 63) / 64); }

template <int PREC>
static void launch_mt(const DevCfg& cfg, const EvBufs& b, const SmearSpec& s, int nev, float* out0, float* out1,
                      const double* ktab, int ktmax, double kdT, hipStream_t st) {
  const dim3 grid((unsigned)mtile_units(cfg) * (unsigned)((nev + 7) / 8 * 8));
  if (ktab) {
    hipLaunchKernelGGL((k_mtile<PREC, 1>), grid, dim3(64), 0, st, cfg, b, s, 0, nev, out0, ktab, ktmax, kdT);
    return;
  }
  if (s.combine == 1) {
    hipLaunchKernelGGL((k_mtile<PREC, 2>), grid, dim3(64), 0, st, cfg, b, s, 0, nev, out0, nullptr, 0, 1.0);
    return;
  }
  SmearSpec s1 = s;
  if (s.combine != 0) s1.cm_out = nullptr;
  hipLaunchKernelGGL((k_mtile<PREC, 0>), grid, dim3(64), 0, st, cfg, b, s1, 0, nev, out0, nullptr, 0, 1.0);
  if (s.combine != 0)
    hipLaunchKernelGGL((k_mtile<PREC, 0>), grid, dim3(64), 0, st, cfg, b, s1, 1, nev, out1, nullptr, 0, 1.0);
}

// prec: 0 FP32, 1 BF16x3, 2 BF16x6.  ktab (tmax x tmax dN/dy table) selects
// the fused MC-KLN epilogue (T_A, T_B of s, then the table per cell).
int launch_mtile(const DevCfg& cfg, const EvBufs& b, const SmearSpec& s, int nev, float* out0, float* out1, int prec,
                 const double* ktab, int ktmax, double kdT, void* stream) {
  if (nev <= 0) return 0;
  if (prec < 0 || prec > 2 || (ktab && (ktmax < 3 || !(kdT > 0)))) return (int)hipErrorInvalidValue;
  hipStream_t st = (hipStream_t)stream;
  if (prec == 0) launch_mt<0>(cfg, b, s, nev, out0, out1, ktab, ktmax, kdT, st);
  else if (prec == 1) launch_mt<1>(cfg, b, s, nev, out0, out1, ktab, ktmax, kdT, st);
  else launch_mt<2>(cfg, b, s, nev, out0, out1, ktab, ktmax, kdT, st);
  return (int)hipGetLastError();
}

}  // namespace smc
