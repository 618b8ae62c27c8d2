// smc_strip.hip -- separable Gaussian smearing on MFMA strips (square
// window; the exact footprint is smc_tile.hip's).
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

__global__ void k_combine_sqrt(float* a, const float* b, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = sqrtf(a[i] * b[i]);
}

int launch_combine_sqrt(float* a, const float* b, size_t n, void* stream) {
  hipLaunchKernelGGL(k_combine_sqrt, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, a, b, n);
  return (int)hipGetLastError();
}

int launch_strip(const DevCfg& cfg, const EvBufs& b, const SmearSpec& s, int nev, float* out0, float* out1,
                 void* stream) {
  if (nev <= 0) return 0;
  if (cfg.ny > kSMaxN) return (int)hipErrorInvalidValue;
  const dim3 grid((cfg.nx + 15) / 16, nev);
  hipStream_t st = (hipStream_t)stream;
  if (s.combine == 0) hipLaunchKernelGGL((k_strip<1>), grid, dim3(64), 0, st, cfg, b, s, out0, out1);
  else hipLaunchKernelGGL((k_strip<2>), grid, dim3(64), 0, st, cfg, b, s, out0, out1);
  return (int)hipGetLastError();
}

}  // namespace smc
