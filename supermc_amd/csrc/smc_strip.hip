// smc_strip.hip -- separable Gaussian smearing on MFMA strips, FP32 and
// BF16x3 split precision (square window; the exact footprint is
// smc_tile.hip's).
//
// The separable form of the Gaussian smearing (MCnucl.cpp:432-478 T_A/T_B,
// :689-760 MC-Glauber, :762-811 sqrt(T_A T_B)):
//     T[x][y] = sum_k w_k gx_k(x) gy_k(y) = (Gx . diag(w) . Gy^T)[x][y]
// with gx_k / gy_k zero outside source k's exact integer window (computed
// once per source by k_event).  Each wavefront owns the 16 x (16*17) cells
// of one strip in 17 MFMA accumulators, selects the sources whose x-window
// meets the strip, orders them by column tile (ballot compaction, so the
// order is deterministic), and issues v_mfma_f32_16x16x4_f32 over four
// sources at a time:
//     A (16 x 4)  lane = row + 16 * source : gx_source(row)
//     B (4 x 16)  lane = col + 16 * source : w_source * gy_source(col)
//     D (16 x 16) lane l, element v        : cell (4*(l/16) + v, l%16)
// The r <= 5w disk cut inside the window is not applied on this path: it
// changes a cell by at most e^-12.5 of a source's peak.
#include "smc_device.h"

namespace smc {

constexpr int kSNT = 17;              // column tiles per strip (grids up to 272 wide)
constexpr int kSMaxN = kSNT * 16;

typedef float v4f __attribute__((ext_vector_type(4)));

struct StripEntry {
  int xh, yh;
  float w;
  int xw, yw;
  int acc;
};

template <int NACC>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(NACC == 1 ? 4 : 2, 8))) void k_strip(DevCfg cfg, EvBufs buf, SmearSpec spec, float* out0, float* out1) {
  __shared__ int ygh[kSMaxN];
  __shared__ StripEntry wl[64];
  const int lane = threadIdx.x;
  const int R = blockIdx.x, e = blockIdx.y;
  const int r0 = 16 * R;
  const int nx = cfg.nx, ny = cfg.ny;
  for (int i = lane; i < kSMaxN; i += 64) ygh[i] = (int)llrint((cfg.ymin + i * cfg.dy) / kH);
  const int arow = r0 + (lane & 15);
  const int arowh = (int)llrint((cfg.xmin + arow * cfg.dx) / kH);
  const float c2 = cfg.inv2w2 * 1.4426950408889634f;
  const int K = buf.counts[(size_t)e * 8 + 5];
  const int4* src = buf.src + (size_t)e * buf.src_cap;
  const int4* swin = buf.swin + (size_t)e * buf.src_cap;

  v4f acc0[kSNT], acc1[NACC == 2 ? kSNT : 1];
#pragma unroll
  for (int c = 0; c < kSNT; ++c) acc0[c] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int c = 0; c < (NACC == 2 ? kSNT : 1); ++c) acc1[c] = v4f{0.f, 0.f, 0.f, 0.f};
  __builtin_amdgcn_wave_barrier();

  for (int sub = 0; sub < K; sub += 64) {
    const int k = sub + lane;
    bool sel = false;
    int bucket = kSNT;
    StripEntry en{0, 0, 0.f, 0, 0, 0};
    if (k < K) {
      const int4 s = src[k];
      const int kind = s.w;
      if (spec.enabled[kind]) {
        const int4 wv = swin[k];
        en.xh = s.x;
        en.yh = s.y;
        en.xw = (spec.win[kind] == kWinBox) ? wv.x : wv.z;
        en.yw = (spec.win[kind] == kWinBox) ? wv.y : wv.w;
        en.w = spec.scale[kind] * (spec.use_fluct[kind] ? __int_as_float(s.z) : 1.f);
        en.acc = spec.acc[kind];
        const int xlo = en.xw & 0xffff, xhi = en.xw >> 16, ylo = en.yw & 0xffff, yhi = en.yw >> 16;
        sel = xlo < r0 + 16 && xhi > r0 && ylo < yhi;
        bucket = ylo >> 4;
      }
    }
    // ordered compaction, grouped by the column tile of the window start
    int n = 0;
    for (int b = 0; b < kSNT; ++b) {
      const bool mine = sel && bucket == b;
      const unsigned long long m = __ballot(mine);
      if (m == 0ull) continue;
      if (mine) wl[n + __popcll(m & ((1ull << lane) - 1ull))] = en;
      n += __popcll(m);
    }
    __builtin_amdgcn_wave_barrier();
    for (int g = 0; g < n; g += 4) {
      const int sidx = g + (lane >> 4);
      const bool valid = sidx < n;
      const StripEntry q = wl[valid ? sidx : 0];
      const int xlo = q.xw & 0xffff, xhi = q.xw >> 16, ylo = q.yw & 0xffff, yhi = q.yw >> 16;
      const int clo = valid ? (ylo >> 4) : kSNT, chi = valid ? ((yhi - 1) >> 4) : -1;
      const int c0 = min(min(__builtin_amdgcn_readlane(clo, 0), __builtin_amdgcn_readlane(clo, 16)),
                         min(__builtin_amdgcn_readlane(clo, 32), __builtin_amdgcn_readlane(clo, 48)));
      const int c1 = max(max(__builtin_amdgcn_readlane(chi, 0), __builtin_amdgcn_readlane(chi, 16)),
                         max(__builtin_amdgcn_readlane(chi, 32), __builtin_amdgcn_readlane(chi, 48)));
      float a = 0.f;
      if (valid && arow >= xlo && arow < xhi) {
        const float dx = (float)(arowh - q.xh) * kHf;
        a = __builtin_amdgcn_exp2f(-(dx * dx) * c2);
      }
      const float w0 = (valid && q.acc == 0) ? q.w : 0.f;
      const float w1 = (valid && q.acc == 1) ? q.w : 0.f;
#pragma unroll
      for (int c = 0; c < kSNT; ++c) {
        if (c >= c0 && c <= c1) {
          const int col = c * 16 + (lane & 15);
          float gy = 0.f;
          if (valid && col >= ylo && col < yhi) {
            const float dy = (float)(ygh[col] - q.yh) * kHf;
            gy = __builtin_amdgcn_exp2f(-(dy * dy) * c2);
          }
          acc0[c] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, w0 * gy, acc0[c], 0, 0, 0);
          if constexpr (NACC == 2) acc1[c] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, w1 * gy, acc1[c], 0, 0, 0);
        }
      }
    }
    __builtin_amdgcn_wave_barrier();
  }

  const size_t N = (size_t)nx * ny;
  const int rsub = 4 * (lane >> 4), csub = lane & 15;
#pragma unroll
  for (int c = 0; c < kSNT; ++c)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int ix = r0 + rsub + v, iy = c * 16 + csub;
      if (ix >= nx || iy >= ny) continue;
      const size_t o = (size_t)e * N + (size_t)ix * ny + iy;
      if constexpr (NACC == 1) {
        out0[o] = acc0[c][v];
      } else {
        if (spec.combine == 1) out0[o] = sqrtf(acc0[c][v] * acc1[c][v]);
        else { out0[o] = acc0[c][v]; out1[o] = acc1[c][v]; }
      }
    }
}

// BF16x3 split precision (smear_path 3).  The same strips, sources and
// square windows as k_strip, contracted 16 sources at a time with
// v_mfma_f32_16x16x16_bf16.  Each factor f (gx, and w gy) is split as
// f = hi + lo, hi = bf16(f), lo = bf16(f - hi), and a tile takes three
// products with FP32 accumulation:
//     T += Ah Bh + Ah Bl + Al Bh     (the dropped Al Bl is below 2^-16 of AB)
// so each term carries about 2^-16 relative error instead of bf16's 2^-8.
//     A (16 x 16)  lane = row + 16 * q, slot j : gx of source 4q + j
//     B (16 x 16)  lane = col + 16 * q, slot j : w gy of source 4q + j
//     D (16 x 16)  lane l, element v           : cell (4*(l/16) + v, l%16)
typedef short v4s __attribute__((ext_vector_type(4)));

__device__ __forceinline__ void split_bf16(const float (&v)[4], v4s& hi, v4s& lo) {
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const __bf16 h = (__bf16)v[j];
    const __bf16 l = (__bf16)(v[j] - (float)h);
    hi[j] = __builtin_bit_cast(short, h);
    lo[j] = __builtin_bit_cast(short, l);
  }
}

template <int NACC>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(NACC == 1 ? 4 : 2, 8))) void k_strip3(
    DevCfg cfg, EvBufs buf, SmearSpec spec, float* out0, float* out1) {
  __shared__ int ygh[kSMaxN];
  __shared__ StripEntry wl[64];
  const int lane = threadIdx.x, q4 = lane >> 4;
  const int R = blockIdx.x, e = blockIdx.y;
  const int r0 = 16 * R;
  const int nx = cfg.nx, ny = cfg.ny;
  for (int i = lane; i < kSMaxN; i += 64) ygh[i] = (int)llrint((cfg.ymin + i * cfg.dy) / kH);
  const int arow = r0 + (lane & 15);
  const int arowh = (int)llrint((cfg.xmin + arow * cfg.dx) / kH);
  const float c2 = cfg.inv2w2 * 1.4426950408889634f;
  const int K = buf.counts[(size_t)e * 8 + 5];
  const int4* src = buf.src + (size_t)e * buf.src_cap;
  const int4* swin = buf.swin + (size_t)e * buf.src_cap;

  v4f acc0[kSNT], acc1[NACC == 2 ? kSNT : 1];
#pragma unroll
  for (int c = 0; c < kSNT; ++c) acc0[c] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int c = 0; c < (NACC == 2 ? kSNT : 1); ++c) acc1[c] = v4f{0.f, 0.f, 0.f, 0.f};
  __builtin_amdgcn_wave_barrier();

  for (int sub = 0; sub < K; sub += 64) {
    const int k = sub + lane;
    bool sel = false;
    int bucket = kSNT;
    StripEntry en{0, 0, 0.f, 0, 0, 0};
    if (k < K) {
      const int4 s = src[k];
      const int kind = s.w;
      if (spec.enabled[kind]) {
        const int4 wv = swin[k];
        en.xh = s.x;
        en.yh = s.y;
        en.xw = (spec.win[kind] == kWinBox) ? wv.x : wv.z;
        en.yw = (spec.win[kind] == kWinBox) ? wv.y : wv.w;
        en.w = spec.scale[kind] * (spec.use_fluct[kind] ? __int_as_float(s.z) : 1.f);
        en.acc = spec.acc[kind];
        const int xlo = en.xw & 0xffff, xhi = en.xw >> 16, ylo = en.yw & 0xffff, yhi = en.yw >> 16;
        sel = xlo < r0 + 16 && xhi > r0 && ylo < yhi;
        bucket = ylo >> 4;
      }
    }
    // ordered compaction, grouped by the column tile of the window start
    int n = 0;
    for (int b = 0; b < kSNT; ++b) {
      const bool mine = sel && bucket == b;
      const unsigned long long m = __ballot(mine);
      if (m == 0ull) continue;
      if (mine) wl[n + __popcll(m & ((1ull << lane) - 1ull))] = en;
      n += __popcll(m);
    }
    __builtin_amdgcn_wave_barrier();
    for (int g = 0; g < n; g += 16) {
      // this lane's four sources of the group, and their row factors
      StripEntry qs[4];
      float a[4];
      int clo = kSNT, chi = -1;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int sidx = g + 4 * q4 + j;
        const bool valid = sidx < n;
        qs[j] = wl[valid ? sidx : 0];
        if (!valid) qs[j].w = 0.f;
        const int xlo = qs[j].xw & 0xffff, xhi = qs[j].xw >> 16, ylo = qs[j].yw & 0xffff, yhi = qs[j].yw >> 16;
        if (valid) {
          clo = min(clo, ylo >> 4);
          chi = max(chi, (yhi - 1) >> 4);
        }
        a[j] = 0.f;
        if (valid && arow >= xlo && arow < xhi) {
          const float dx = (float)(arowh - qs[j].xh) * kHf;
          a[j] = __builtin_amdgcn_exp2f(-(dx * dx) * c2);
        }
      }
      const int c0 = min(min(__builtin_amdgcn_readlane(clo, 0), __builtin_amdgcn_readlane(clo, 16)),
                         min(__builtin_amdgcn_readlane(clo, 32), __builtin_amdgcn_readlane(clo, 48)));
      const int c1 = max(max(__builtin_amdgcn_readlane(chi, 0), __builtin_amdgcn_readlane(chi, 16)),
                         max(__builtin_amdgcn_readlane(chi, 32), __builtin_amdgcn_readlane(chi, 48)));
      v4s ah, al;
      split_bf16(a, ah, al);
#pragma unroll
      for (int c = 0; c < kSNT; ++c) {
        if (c >= c0 && c <= c1) {
          const int col = c * 16 + (lane & 15);
          float b0[4], b1[4];
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            const int ylo = qs[j].yw & 0xffff, yhi = qs[j].yw >> 16;
            float gy = 0.f;
            if (col >= ylo && col < yhi) {
              const float dy = (float)(ygh[col] - qs[j].yh) * kHf;
              gy = __builtin_amdgcn_exp2f(-(dy * dy) * c2);
            }
            b0[j] = (qs[j].acc == 0) ? qs[j].w * gy : 0.f;
            b1[j] = (NACC == 2 && qs[j].acc == 1) ? qs[j].w * gy : 0.f;
          }
          v4s bh, bl;
          split_bf16(b0, bh, bl);
          acc0[c] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(al, bh, acc0[c], 0, 0, 0);
          acc0[c] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, bl, acc0[c], 0, 0, 0);
          acc0[c] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, bh, acc0[c], 0, 0, 0);
          if constexpr (NACC == 2) {
            split_bf16(b1, bh, bl);
            acc1[c] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(al, bh, acc1[c], 0, 0, 0);
            acc1[c] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, bl, acc1[c], 0, 0, 0);
            acc1[c] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, bh, acc1[c], 0, 0, 0);
          }
        }
      }
    }
    __builtin_amdgcn_wave_barrier();
  }

  const size_t N = (size_t)nx * ny;
  const int rsub = 4 * q4, csub = lane & 15;
#pragma unroll
  for (int c = 0; c < kSNT; ++c)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int ix = r0 + rsub + v, iy = c * 16 + csub;
      if (ix >= nx || iy >= ny) continue;
      const size_t o = (size_t)e * N + (size_t)ix * ny + iy;
      if constexpr (NACC == 1) {
        out0[o] = acc0[c][v];
      } else {
        out0[o] = acc0[c][v];
        out1[o] = acc1[c][v];
      }
    }
}

int launch_strip(const DevCfg& cfg, const EvBufs& b, const SmearSpec& s, int nev, float* out0, float* out1,
                 bool bf16x3, void* stream) {
  if (nev <= 0) return 0;
  if (cfg.ny > kSMaxN) return (int)hipErrorInvalidValue;
  if (bf16x3 && s.combine == 1) return (int)hipErrorInvalidValue;  // sqrt(T_A T_B) is not on the split path
  const dim3 grid((cfg.nx + 15) / 16, nev);
  hipStream_t st = (hipStream_t)stream;
  if (bf16x3) {
    if (s.combine == 0) hipLaunchKernelGGL((k_strip3<1>), grid, dim3(64), 0, st, cfg, b, s, out0, out1);
    else hipLaunchKernelGGL((k_strip3<2>), grid, dim3(64), 0, st, cfg, b, s, out0, out1);
  } else {
    if (s.combine == 0) hipLaunchKernelGGL((k_strip<1>), grid, dim3(64), 0, st, cfg, b, s, out0, out1);
    else hipLaunchKernelGGL((k_strip<2>), grid, dim3(64), 0, st, cfg, b, s, out0, out1);
  }
  return (int)hipGetLastError();
}

}  // namespace smc
