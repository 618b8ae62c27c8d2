// smc_grid64.hip -- the entropy-density grid in FP64 with the reference's
// double arithmetic: MCnucl::setDensity for which_mc_model 5
// (MCnucl.cpp:689-760: (1-alpha)/2 sum_part f T_n + alpha sum_coll f G) and
// 7 (MCnucl.cpp:762-811: sqrt(rho_A rho_B)), participants through
// Particle::getSmoothDensity (Particle.cpp:122-132, 168-174) over their
// boxes, collisions over the +-5w windows with the dc <= 25 w^2 cut.
//
// Per (source, cell) in FP64, as the reference writes it:
//   participant  r = sqrt((xg-x)^2 + (yg-y)^2); r > 5w -> 0, else
//                (1/(2 pi w^2)) exp(-r r / (2 w w)), times f
//   collision    dc = (x-xg)^2 + (y-yg)^2 <= 25 w^2:
//                f (1/(2 pi w^2)) exp(-dc / (2 w^2)) (alpha + (1 - alpha) aw)
//                with aw = 0, or for sub_model 2 (MCnucl.cpp:713-717: no
//                participant term) the pair's weight of MCnucl.cpp:345-348
// summed per cell in source order (participants of A and B in wound order,
// then the pairs in pair order), the participant sum scaled by (1-alpha)/2
// at the end.  The file is built with -ffp-contract=off so that no product
// and sum fuse: every operation rounds as the reference's x86-64 doubles do,
// and only exp may differ from the host's by an ulp.
//
// One workgroup of 4 waves per (16 x 64 tile, event); lane = column, each
// wave 4 rows.  Tiles outside the event's source rectangle are written as
// zeros without a source pass.
#include "smc_device.h"

namespace smc {

namespace {
constexpr int kGX = 16, kGY = 64;
}

__global__ __launch_bounds__(256) void k_grid64(DevCfg cfg, EvBufs buf, Grid64Const gc, int ev0, double* out) {
  const int el = blockIdx.y;
  const int e = ev0 + el;
  const int ntx = (cfg.nx + kGX - 1) / kGX;
  const int tx0 = (blockIdx.x % ntx) * kGX, ty0 = (blockIdx.x / ntx) * kGY;
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const int iy = ty0 + lane;
  const int rx0 = tx0 + wv * 4;

  __shared__ double2 lpos[256];
  __shared__ int4 lwin[256];  // x lo, x hi, y lo, y hi
  __shared__ double lf[256];
  __shared__ double lw[256];  // pair weight: alpha, or alpha + (1 - alpha) aw for sub_model 2
  __shared__ int lacc[256];   // 0: participant sum, 1: collision sum / nucleus B
  __shared__ int lpair[256];  // 1: collision term
  __shared__ int wcount[4];

  const int rxp = buf.counts[(size_t)e * 8 + 6], ryp = buf.counts[(size_t)e * 8 + 7];
  const bool live = tx0 < (rxp >> 16) && tx0 + kGX > (rxp & 0xffff) && ty0 < (ryp >> 16) &&
                    ty0 + kGY > (ryp & 0xffff);

  double xg[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) xg[r] = cfg.xmin + (double)(rx0 + r) * cfg.dx;
  const double yg = cfg.ymin + (double)iy * cfg.dy;

  double a0[4] = {0.0, 0.0, 0.0, 0.0}, a1[4] = {0.0, 0.0, 0.0, 0.0};
  const int K = live ? buf.counts[(size_t)e * 8 + 5] : 0;  // uniform over the workgroup
  const int nsrc_part = buf.counts[(size_t)e * 8] + buf.counts[(size_t)e * 8 + 1];
  const size_t pbase = (size_t)e * cfg.max_pairs;
  const int4* src = buf.src + (size_t)e * buf.src_cap;
  const int4* swin = buf.swin + (size_t)e * buf.src_cap;

  for (int base = 0; base < K; base += 256) {
    const int k = base + t;
    bool take = false;
    double2 pos = make_double2(0.0, 0.0);
    int4 win = make_int4(0, 0, 0, 0);
    double f = 1.0, wgt = gc.alpha;
    int acc = 0, pair = 0;
    if (k < K) {
      const int4 s = src[k];
      const int4 w4 = swin[k];
      pair = s.w == kSrcPair;
      const int wx = pair ? w4.z : w4.x, wy = pair ? w4.w : w4.y;  // 5w windows for collisions, boxes otherwise
      win = make_int4(wx & 0xffff, wx >> 16, wy & 0xffff, wy >> 16);
      take = (pair ? gc.use_binary : !gc.uli) && win.x < win.y && win.z < win.w && win.x < tx0 + kGX &&
             win.y > tx0 && win.z < ty0 + kGY && win.w > ty0;
      pos = make_double2((double)s.x * kH, (double)s.y * kH);
      f = gc.fluct ? (double)__int_as_float(s.z) : 1.0;
      if (pair && gc.uli) {  // the pair's Gamma factor and (Alpha + (1 - Alpha) aw) of MCnucl.cpp:755
        const size_t pk = pbase + (size_t)(k - nsrc_part);
        f = gc.fluct ? (double)buf.pfluct[pk] : 1.0;
        wgt = gc.alpha + (1. - gc.alpha) * (double)buf.paw[pk];
      }
      acc = gc.model == 7 ? (s.w == kSrcPartB) : pair;
    }
    // ordered compaction: the tile's sources in list order
    const unsigned long long m = __ballot(take);
    const int before = __popcll(m & ((1ull << lane) - 1ull));
    if (lane == 0) wcount[wv] = __popcll(m);
    __syncthreads();
    int woff = 0;
    for (int q = 0; q < wv; ++q) woff += wcount[q];
    const int n = wcount[0] + wcount[1] + wcount[2] + wcount[3];
    if (take) {
      lpos[woff + before] = pos;
      lwin[woff + before] = win;
      lf[woff + before] = f;
      lw[woff + before] = wgt;
      lacc[woff + before] = acc;
      lpair[woff + before] = pair;
    }
    __syncthreads();
    for (int q = 0; q < n; ++q) {
      const int4 ww = lwin[q];
      if (rx0 >= ww.y || rx0 + 4 <= ww.x) continue;  // wave-uniform
      if (iy < ww.z || iy >= ww.w) continue;
      const double2 p = lpos[q];
      const double fq = lf[q];
      const bool to1 = lacc[q] != 0;
      if (lpair[q]) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int ix = rx0 + r;
          if (ix < ww.x || ix >= ww.y) continue;
          const double ddx = p.x - xg[r], ddy = p.y - yg;
          const double dc = ddx * ddx + ddy * ddy;
          if (dc <= gc.dc_max) {
            const double v = fq * gc.norm * exp(-dc / gc.two_wsq) * lw[q];
            if (to1) a1[r] += v; else a0[r] += v;
          }
        }
      } else {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int ix = rx0 + r;
          if (ix < ww.x || ix >= ww.y) continue;
          const double ddx = xg[r] - p.x, ddy = yg - p.y;
          const double rr = sqrt(ddx * ddx + ddy * ddy);
          const double tn = (rr > gc.r_max) ? 0.0 : gc.norm * exp(-rr * rr / gc.two_w_w);
          const double v = tn * fq;
          if (to1) a1[r] += v; else a0[r] += v;
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
    const double v = gc.model == 7 ? sqrt(a0[r] * a1[r]) : a0[r] * gc.prefactor + a1[r];
    out[(size_t)el * N + (size_t)ix * cfg.ny + iy] = v;
  }
}

int launch_grid64(const DevCfg& cfg, const EvBufs& b, const Grid64Const& gc, int ev0, int nev, double* out,
                  void* stream) {
  if (nev <= 0) return 0;
  const int ntx = (cfg.nx + kGX - 1) / kGX, nty = (cfg.ny + kGY - 1) / kGY;
  hipLaunchKernelGGL(k_grid64, dim3(ntx * nty, nev), dim3(256), 0, (hipStream_t)stream, cfg, b, gc, ev0, out);
  return (int)hipGetLastError();
}

}  // namespace smc
