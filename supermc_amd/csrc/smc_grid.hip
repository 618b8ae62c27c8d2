// smc_grid.hip -- fused per-event grid kernel for gfx950.
//
// One 512-thread workgroup per event.  The event's 2D grid (up to 272 x 272
// cells = 17 x 17 tiles of 16 x 16) lives in the accumulator registers of its
// eight waves for the whole kernel:
//
//   1. smearing (MCnucl.cpp:626-866 / 398-478): the source list written by
//      k_event is staged in LDS with its exact integer footprints; every
//      wave owns two full tile rows plus two or three tiles of the last tile
//      row.  Per tile row it selects the sources whose x-window meets the
//      row, sorts them by column tile, and applies them in groups of four:
//        MFMA path  T_tile += Gx(16 x 4) . diag(w) Gy^T(4 x 16)
//                   (v_mfma_f32_16x16x4_f32; Gx/Gy carry the integer
//                   windows, the r <= 5w corner cut is not applied)
//        VALU path  the same cells with the reference's disk cut, exact
//   2. optional write-out of the grid (smc_density / smc_thickness),
//   3. eps_n moments (MakeDensity.cpp:2244-2430) straight from the
//      registers: FP64 centre of mass, <r^n> over the full grid and the
//      hot-spot-masked eps_n sums.
//
// No HBM traffic besides the source list and the per-event record.
#include "smc_device.h"

namespace smc {

constexpr int kGW = 8;            // waves per workgroup
constexpr int kGT = kGW * 64;     // threads per workgroup
constexpr int kNT = 17;           // tiles per grid dimension
constexpr int kGMaxN = kNT * 16;  // cells per dimension
constexpr int kChunk = 1024;      // sources staged per pass

typedef float v4f __attribute__((ext_vector_type(4)));

// LDS of the moment phase
struct MomLds {
  int rect[2 * kAMax][4];  // participant mask rectangles (grid cells)
  double red[kGW][64];     // per-wave partial sums
  double tot[64];          // block sums
  int ired[kGW][4];
};

// LDS of the fused kernel
struct GridLds {
  int xgh[kGMaxN], ygh[kGMaxN];  // cell centres in half quanta
  int sxh[kChunk], syh[kChunk];  // source positions (half quanta)
  float sw[kChunk];              // source weights
  int sxw[kChunk], syw[kChunk];  // windows, lo | hi << 16
  unsigned short wl[kGW][64];    // per-wave selection list
  MomLds m;
};

__device__ __forceinline__ int rl(int v, int l) { return __builtin_amdgcn_readlane(v, l); }

// Applies the staged sources to one tile row.  cols[t] is the column tile
// held in acc[t]; ncol of them are live.
template <int NC, bool MFMA>
__device__ __forceinline__ void smear_row(const GridLds& L, unsigned short* wl, int R, v4f (&acc)[NC],
                                          const int (&cols)[NC], int nchunk, float c2, float cut2) {
  const int lane = threadIdx.x & 63;
  const int r0 = R * 16;
  const int arow = min(r0 + (lane & 15), kGMaxN - 1);
  const int arowh = L.xgh[arow];
  for (int sub = 0; sub < nchunk; sub += 64) {
    const int k = sub + lane;
    bool sel = false;
    int bucket = kNT;
    if (k < nchunk) {
      const int xw = L.sxw[k], yw = L.syw[k];
      const int xlo = xw & 0xffff, xhi = xw >> 16, ylo = yw & 0xffff, yhi = yw >> 16;
      sel = xlo < r0 + 16 && xhi > r0 && ylo < yhi;
      bucket = ylo >> 4;
    }
    // ordered compaction into the wave list, grouped by column tile
    int n = 0;
    for (int b = 0; b < kNT; ++b) {
      const bool mine = sel && bucket == b;
      const unsigned long long m = __ballot(mine);
      if (m == 0ull) continue;
      if (mine) wl[n + __popcll(m & ((1ull << lane) - 1ull))] = (unsigned short)k;
      n += __popcll(m);
    }
    __builtin_amdgcn_wave_barrier();
    for (int g = 0; g < n; g += 4) {
      const int s = g + (lane >> 4);
      const bool valid = s < n;
      const int kk = valid ? wl[s] : 0;
      const int xh = L.sxh[kk], yh = L.syh[kk];
      const float w = valid ? L.sw[kk] : 0.f;
      const int xw = L.sxw[kk], yw = L.syw[kk];
      const int xlo = xw & 0xffff, xhi = xw >> 16, ylo = yw & 0xffff, yhi = yw >> 16;
      const int clo = valid ? (ylo >> 4) : kNT, chi = valid ? ((yhi - 1) >> 4) : -1;
      const int c0 = min(min(rl(clo, 0), rl(clo, 16)), min(rl(clo, 32), rl(clo, 48)));
      const int c1 = max(max(rl(chi, 0), rl(chi, 16)), max(rl(chi, 32), rl(chi, 48)));
      if constexpr (MFMA) {
        // A: Gx[row][source], lane = row + 16 * source
        float a = 0.f;
        if (valid && arow >= xlo && arow < xhi) {
          const float dx = (float)(arowh - xh) * kHf;
          a = __builtin_amdgcn_exp2f(-(dx * dx) * c2);
        }
#pragma unroll
        for (int t = 0; t < NC; ++t) {
          const int c = cols[t];
          if (c >= c0 && c <= c1) {
            // B: w Gy[source][col], lane = col + 16 * source
            const int col = c * 16 + (lane & 15);
            float b = 0.f;
            if (valid && col >= ylo && col < yhi) {
              const float dy = (float)(L.ygh[col] - yh) * kHf;
              b = w * __builtin_amdgcn_exp2f(-(dy * dy) * c2);
            }
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc[t], 0, 0, 0);
          }
        }
      } else {
        // exact reference footprint: each lane owns 4 cells of the tile
        // (rows 4*(lane/16)+v, column lane%16), disk cut d^2 <= 25 w^2
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int sl = 16 * q;
          if (rl((int)valid, sl) == 0) continue;
          const int qxh = rl(xh, sl), qyh = rl(yh, sl);
          const float qw = __int_as_float(rl(__float_as_int(w), sl));
          const int qxw = rl(xw, sl), qyw = rl(yw, sl);
          const int qxlo = qxw & 0xffff, qxhi = qxw >> 16, qylo = qyw & 0xffff, qyhi = qyw >> 16;
          const int qc0 = qylo >> 4, qc1 = (qyhi - 1) >> 4;
#pragma unroll
          for (int t = 0; t < NC; ++t) {
            const int c = cols[t];
            if (c >= qc0 && c <= qc1) {
              const int col = c * 16 + (lane & 15);
              const float dy = (float)(L.ygh[min(col, kGMaxN - 1)] - qyh) * kHf;
              const float dy2 = dy * dy;
              const float gy = (col >= qylo && col < qyhi) ? qw * __builtin_amdgcn_exp2f(-dy2 * c2) : 0.f;
#pragma unroll
              for (int v = 0; v < 4; ++v) {
                const int row = r0 + 4 * (lane >> 4) + v;
                const float dx = (float)(L.xgh[min(row, kGMaxN - 1)] - qxh) * kHf;
                const float dx2 = dx * dx;
                const bool in = row >= qxlo && row < qxhi && dx2 + dy2 <= cut2;
                acc[t][v] += in ? __builtin_amdgcn_exp2f(-dx2 * c2) * gy : 0.f;
              }
            }
          }
        }
      }
    }
    __builtin_amdgcn_wave_barrier();
  }
}

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

// Loads row ix of the grid, MAXC values per lane (columns lane + 64 j).
template <int MAXC>
__device__ __forceinline__ void load_row(const float* g, int ix, int nx, int ny, float (&v)[MAXC]) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int j = 0; j < MAXC; ++j) {
    const int iy = lane + 64 * j;
    v[j] = (ix < nx && iy < ny) ? g[(size_t)ix * ny + iy] : 0.f;
  }
}

// Moments of one event's grid g[nx][ny] (global memory) for a workgroup of
// NW waves.  s1 holds this thread's partial (sum rho, sum x rho, sum y rho)
// when the caller computed them (have_s1), otherwise pass 1 reads the grid.
// ff = finalFactor (setSd, MakeDensity.cpp:2514-2522).
template <int NW, int NORD, int MAXC>
__device__ void grid_moments(const DevCfg& cfg, const EvBufs& buf, int e, const float* g, double* evout,
                             MomLds& L, unsigned* mask, bool have_s1, double (&s1)[3]) {
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const int nx = cfg.nx, ny = cfg.ny;
  const int* cnt = buf.counts + (size_t)e * 8;
  const int npa = cnt[0], npb = cnt[1], ncoll = cnt[2];
  const double ff = cfg.final_factor;
  const int nwy = (ny + 31) >> 5;

  // pass 1: centre of mass (MakeDensity.cpp:2273-2282), rows prefetched
  if (!have_s1) {
    float cur[MAXC], nxt[MAXC];
    load_row<MAXC>(g, wv, nx, ny, cur);
    for (int ix = wv; ix < nx; ix += NW) {
      load_row<MAXC>(g, ix + NW, nx, ny, nxt);
      const double xg = cfg.xmin + ix * cfg.dx;
      double rs = 0.0, ys = 0.0;
#pragma unroll
      for (int j = 0; j < MAXC; ++j) {
        const double d = (double)cur[j] * ff;
        rs += d;
        ys += (cfg.ymin + (lane + 64 * j) * cfg.dy) * d;
      }
      s1[0] += rs;
      s1[1] += xg * rs;
      s1[2] += ys;
#pragma unroll
      for (int j = 0; j < MAXC; ++j) cur[j] = nxt[j];
    }
  }

  // hot-spot mask (MCnucl.cpp:1303-1323, MakeDensity.cpp:2341-2359): union
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
  block_sums<NW, 3>(s1, L);  // also orders the mask writes before the reads below
  const double total = L.tot[0];
  const double xc = L.tot[1] / total, yc = L.tot[2] / total;

  // pass 2: <r^n> over the full grid and the masked eps_n sums, rows prefetched
  MomAcc<NORD> M;
#pragma unroll
  for (int n = 0; n < 10; ++n) M.rn[n] = 0.0;
#pragma unroll
  for (int o = 0; o < NORD; ++o) M.mre[o] = M.mim[o] = M.mno[o] = M.pre[o] = M.pim[o] = M.pno[o] = 0.0;
  const int from = cfg.from_order, to = cfg.to_order;
  {
    float cur[MAXC], nxt[MAXC];
    load_row<MAXC>(g, wv, nx, ny, cur);
    for (int ix = wv; ix < nx; ix += NW) {
      load_row<MAXC>(g, ix + NW, nx, ny, nxt);
      const double x = cfg.xmin + ix * cfg.dx - xc;
      const unsigned* row = mask + (size_t)ix * nwy;
#pragma unroll
      for (int j = 0; j < MAXC; ++j) {
        if (cur[j] != 0.f) {
          const int iy = lane + 64 * j;
          const double y = cfg.ymin + iy * cfg.dy - yc;
          const bool masked = (row[iy >> 5] >> (iy & 31)) & 1u;
          moment_cell<NORD>(M, from, to, (double)cur[j] * ff, x, y, masked);
        }
      }
#pragma unroll
      for (int j = 0; j < MAXC; ++j) cur[j] = nxt[j];
    }
  }
  constexpr int NV = 10 + 6 * NORD;
  double all[NV];
#pragma unroll
  for (int n = 0; n < 10; ++n) all[n] = M.rn[n];
#pragma unroll
  for (int o = 0; o < NORD; ++o) {
    all[10 + 6 * o + 0] = M.mre[o];
    all[10 + 6 * o + 1] = M.mim[o];
    all[10 + 6 * o + 2] = M.mno[o];
    all[10 + 6 * o + 3] = M.pre[o];
    all[10 + 6 * o + 4] = M.pim[o];
    all[10 + 6 * o + 5] = M.pno[o];
  }
  block_sums<NW, NV>(all, L);
  if (t == 0) {
    const double* T = L.tot;
    double* out = evout + (size_t)e * SMC_EV_STRIDE;
    for (int i = 0; i < 40; ++i) out[i] = 0.0;
    for (int n = 0; n < 10; ++n) out[SMC_EV_RN + n] = T[n] / total;
    const double eps = 1e-15;
    for (int o = 0; o < NORD; ++o) {
      const int n = from + o;
      if (n > to) break;
      out[SMC_EV_EPS_RE + n] = -T[10 + 6 * o + 0] / (T[10 + 6 * o + 2] + eps);
      out[SMC_EV_EPS_IM + n] = -T[10 + 6 * o + 1] / (T[10 + 6 * o + 2] + eps);
      out[SMC_EV_EPSP_RE + n] = -T[10 + 6 * o + 3] / (T[10 + 6 * o + 5] + eps);
      out[SMC_EV_EPSP_IM + n] = -T[10 + 6 * o + 4] / (T[10 + 6 * o + 5] + eps);
    }
    out[SMC_EV_XC] = xc;
    out[SMC_EV_YC] = yc;
    out[SMC_EV_TOTAL] = total;
    out[SMC_EV_DSDY] = total * cfg.dx * cfg.dy;
    out[SMC_EV_NPART] = npa + npb;
    out[SMC_EV_NCOLL] = ncoll;
    out[SMC_EV_B] = buf.bval[e];
    out[SMC_EV_STATUS] = cnt[3];
  }
}

// Sum of the grid only (dndy, MCnucl.cpp:817-820), no finalFactor.
template <int NW>
__device__ void grid_sum(const DevCfg& cfg, int e, const float* g, double* evout, MomLds& L) {
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  double s[1] = {0.0};
  for (int ix = wv; ix < cfg.nx; ix += NW)
    for (int iy = lane; iy < cfg.ny; iy += 64) s[0] += (double)g[(size_t)ix * cfg.ny + iy];
  block_sums<NW, 1>(s, L);
  if (t == 0) evout[(size_t)e * SMC_EV_STRIDE + SMC_EV_TOTAL] = L.tot[0];
}

// flags: bit 0 write the grid (always the case: it is the moments' input),
// bit 1 eps_n record, bit 2 dndy only.
template <bool MFMA, int NORD>
__global__ __launch_bounds__(kGT) void k_grid(DevCfg cfg, EvBufs buf, SmearSpec spec, float* grid, double* evout,
                                              int flags) {
  __shared__ GridLds L;
  extern __shared__ unsigned dyn_mask[];
  const int e = blockIdx.x, t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const int nx = cfg.nx, ny = cfg.ny;
  const int* cnt = buf.counts + (size_t)e * 8;
  const int K = cnt[5];
  const int4* src = buf.src + (size_t)e * buf.src_cap;
  const float c2 = cfg.inv2w2 * 1.4426950408889634f;

  for (int i = t; i < kGMaxN; i += kGT) {
    L.xgh[i] = (int)llrint((cfg.xmin + i * cfg.dx) / kH);
    L.ygh[i] = (int)llrint((cfg.ymin + i * cfg.dy) / kH);
  }
  // tiles owned by this wave: rows wv and wv+8 (all 17 columns) and two or
  // three columns of the last tile row
  int colsF[kNT];
#pragma unroll
  for (int c = 0; c < kNT; ++c) colsF[c] = c;
  const int cols16[3] = {wv, wv + 8, (wv == 0) ? 16 : -100};
  v4f accA[kNT], accB[kNT], acc16[3];
#pragma unroll
  for (int c = 0; c < kNT; ++c) {
    accA[c] = v4f{0.f, 0.f, 0.f, 0.f};
    accB[c] = v4f{0.f, 0.f, 0.f, 0.f};
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) acc16[c] = v4f{0.f, 0.f, 0.f, 0.f};

  for (int base = 0; base < K; base += kChunk) {
    const int nchunk = min(kChunk, K - base);
    __syncthreads();
    for (int k = t; k < nchunk; k += kGT) {
      const int4 s = src[base + k];
      const int kind = s.w;
      int xlo = 0, xhi = 0, ylo = 0, yhi = 0;
      float wgt = 0.f;
      if (spec.enabled[kind]) {
        if (spec.win[kind] == kWinBox) {
          xlo = step_eval(cfg.tab[kBoxLoX], s.x);
          xhi = step_eval(cfg.tab[kBoxHiX], s.x);
          ylo = step_eval(cfg.tab[kBoxLoY], s.y);
          yhi = step_eval(cfg.tab[kBoxHiY], s.y);
        } else {
          xlo = step_eval(cfg.tab[kWinLoX], s.x);
          xhi = step_eval(cfg.tab[kWinHiX], s.x);
          ylo = step_eval(cfg.tab[kWinLoY], s.y);
          yhi = step_eval(cfg.tab[kWinHiY], s.y);
        }
        wgt = spec.scale[kind] * (spec.use_fluct[kind] ? __int_as_float(s.z) : 1.f);
        if (!(xlo < xhi && ylo < yhi)) { xlo = xhi = ylo = yhi = 0; }
      }
      L.sxh[k] = s.x;
      L.syh[k] = s.y;
      L.sw[k] = wgt;
      L.sxw[k] = xlo | (xhi << 16);
      L.syw[k] = ylo | (yhi << 16);
    }
    __syncthreads();
    smear_row<kNT, MFMA>(L, L.wl[wv], wv, accA, colsF, nchunk, c2, cfg.cut2);
    smear_row<kNT, MFMA>(L, L.wl[wv], wv + 8, accB, colsF, nchunk, c2, cfg.cut2);
    smear_row<3, MFMA>(L, L.wl[wv], 16, acc16, cols16, nchunk, c2, cfg.cut2);
  }

  // accumulator element v of tile (R, c) is cell (16R + 4(lane/16) + v, 16c + lane%16)
  const int rsub = 4 * (lane >> 4), csub = lane & 15;
  const size_t N = (size_t)nx * ny;
  float* g = grid + (size_t)e * N;
#pragma unroll
  for (int c = 0; c < kNT; ++c)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int iy = c * 16 + csub;
      const int ixa = wv * 16 + rsub + v, ixb = (wv + 8) * 16 + rsub + v;
      if (iy < ny && ixa < nx) g[(size_t)ixa * ny + iy] = accA[c][v];
      if (iy < ny && ixb < nx) g[(size_t)ixb * ny + iy] = accB[c][v];
    }
#pragma unroll
  for (int q = 0; q < 3; ++q)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int iy = cols16[q] * 16 + csub, ix = 256 + rsub + v;
      if (cols16[q] >= 0 && iy < ny && ix < nx) g[(size_t)ix * ny + iy] = acc16[q][v];
    }
  if (!(flags & 6)) return;
  if (flags & 4) {
    __syncthreads();  // the workgroup's grid writes are visible to all its threads
    grid_sum<kGW>(cfg, e, g, evout, L.m);
    return;
  }
  // pass-1 partial sums straight from the registers
  double s1[3] = {0.0, 0.0, 0.0};
  const double ff = cfg.final_factor;
  auto acc1 = [&](float val, int ix, int iy) {
    if (ix < nx && iy < ny) {
      const double d = (double)val * ff;
      s1[0] += d;
      s1[1] += (cfg.xmin + ix * cfg.dx) * d;
      s1[2] += (cfg.ymin + iy * cfg.dy) * d;
    }
  };
#pragma unroll
  for (int c = 0; c < kNT; ++c)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      acc1(accA[c][v], wv * 16 + rsub + v, c * 16 + csub);
      acc1(accB[c][v], (wv + 8) * 16 + rsub + v, c * 16 + csub);
    }
#pragma unroll
  for (int q = 0; q < 3; ++q)
#pragma unroll
    for (int v = 0; v < 4; ++v)
      if (cols16[q] >= 0) acc1(acc16[q][v], 256 + rsub + v, cols16[q] * 16 + csub);
  __syncthreads();  // the workgroup's grid writes are visible to all its threads
  grid_moments<kGW, NORD, 5>(cfg, buf, e, g, evout, L.m, dyn_mask, true, s1);
}

// Moments of grids already in memory (non-fused densities), same code path.
template <int NORD>
__global__ __launch_bounds__(kGT) void k_moments2(DevCfg cfg, EvBufs buf, const float* grid, double* evout,
                                                  int flags) {
  __shared__ MomLds L;
  extern __shared__ unsigned dyn_mask[];
  const int e = blockIdx.x;
  const float* g = grid + (size_t)e * cfg.nx * cfg.ny;
  double s1[3] = {0.0, 0.0, 0.0};
  if (flags & 2) grid_moments<kGW, NORD, 9>(cfg, buf, e, g, evout, L, dyn_mask, false, s1);
  else grid_sum<kGW>(cfg, e, g, evout, L);
}

int launch_grid(const DevCfg& cfg, const EvBufs& b, const SmearSpec& s, int nev, float* grid, double* ev_out,
                int flags, int mfma, void* stream) {
  if (nev <= 0) return 0;
  if (cfg.nx > kGMaxN || cfg.ny > kGMaxN || !grid) return (int)hipErrorInvalidValue;
  const int nord = cfg.to_order - cfg.from_order + 1;
  hipStream_t st = (hipStream_t)stream;
  const size_t shm = (size_t)cfg.nx * ((cfg.ny + 31) / 32) * sizeof(unsigned);
  if (nord <= 4) {
    if (mfma) hipLaunchKernelGGL((k_grid<true, 4>), dim3(nev), dim3(kGT), shm, st, cfg, b, s, grid, ev_out, flags);
    else hipLaunchKernelGGL((k_grid<false, 4>), dim3(nev), dim3(kGT), shm, st, cfg, b, s, grid, ev_out, flags);
  } else {
    if (mfma) hipLaunchKernelGGL((k_grid<true, 9>), dim3(nev), dim3(kGT), shm, st, cfg, b, s, grid, ev_out, flags);
    else hipLaunchKernelGGL((k_grid<false, 9>), dim3(nev), dim3(kGT), shm, st, cfg, b, s, grid, ev_out, flags);
  }
  return (int)hipGetLastError();
}

int launch_moments(const DevCfg& cfg, const EvBufs& b, const float* grid, int nev, double* ev_out, int mode,
                   void* stream) {
  if (nev <= 0) return 0;
  if (cfg.nx > kGMaxN * 2 || cfg.ny > kGMaxN * 2) return (int)hipErrorInvalidValue;
  const int flags = (mode == 0) ? 2 : 4;
  const int nord = cfg.to_order - cfg.from_order + 1;
  hipStream_t st = (hipStream_t)stream;
  const size_t shm = (size_t)cfg.nx * ((cfg.ny + 31) / 32) * sizeof(unsigned);
  if (nord <= 4) hipLaunchKernelGGL((k_moments2<4>), dim3(nev), dim3(kGT), shm, st, cfg, b, grid, ev_out, flags);
  else hipLaunchKernelGGL((k_moments2<9>), dim3(nev), dim3(kGT), shm, st, cfg, b, grid, ev_out, flags);
  return (int)hipGetLastError();
}

}  // namespace smc
