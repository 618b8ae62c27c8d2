// smc_avg.hip -- averaged, rotated profiles (operation 3 of the reference,
// MakeDensity.cpp:736-1650 pT-integrated branch), and the spectator source
// lists of operation 2 (MakeDensity.cpp:148-500).
//
// Per accepted event and order n (in the reference's order, rotations
// accumulating from one order to the next):
//   k_cm_angle    GlueDensity::calcCMAngle (GlueDensity.cpp:87-144): centre
//                 of mass of the current density and the order-n angle
//                 -atan2(-Im, -Re)/n of sum r^n e^{i n theta} rho about it
//                 (full grid, no mask), one 256-thread workgroup per event;
//   k_rotate_src  MCnucl::recenterGrid / rotateGrid (MCnucl.cpp:1119-1174):
//                 every source (participants and collision pairs) moves by
//                 -CM and turns by the angle (Point3D::rotate with theta = 0,
//                 MathBasics.cpp:41-50), back to half-quanta; footprints and
//                 the source rectangle are recomputed from the new
//                 positions (Particle::calculateBounds);
//   k_accum_part, k_accum_fin
//                 sum over the batch's accepted events of finalFactor * rho
//                 into an FP64 grid: per (cell, segment of events) partials
//                 in event order, then the segments in order.
#include "smc_device.h"

namespace smc {

namespace {
struct ARect {
  int x0, x1, y0, y1;
};
__device__ __forceinline__ ARect arect(const EvBufs& buf, int e) {
  const int a = buf.counts[(size_t)e * 8 + 6], b = buf.counts[(size_t)e * 8 + 7];
  return ARect{a & 0xffff, a >> 16, b & 0xffff, b >> 16};
}
}  // namespace

// rot[e] = {xc, yc, angle, total}.  cm: nu centre-of-mass unit partials per
// event {sum v, sum x v, sum y v, 0}.  angle_on = 0 gives angle 0 (recentre
// only, the reaction-plane profile).
__global__ __launch_bounds__(256) void k_cm_angle(DevCfg cfg, EvBufs buf, const float* grid, const double* cm, int nu,
                                                  int n, int angle_on, double* rot) {
  __shared__ double red[4];
  const int e = blockIdx.x, t = threadIdx.x, lane = t & 63, wv = t >> 6;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  for (int u = t; u < nu; u += 256) {
    const double* p = cm + ((size_t)e * nu + u) * 4;
    s0 += p[0];
    s1 += p[1];
    s2 += p[2];
  }
  s0 = block_sum_d(s0, red);
  s1 = block_sum_d(s1, red);
  s2 = block_sum_d(s2, red);
  const double xc = s1 / s0, yc = s2 / s0;
  double re = 0.0, im = 0.0;
  if (angle_on) {
    const ARect R = arect(buf, e);
    const float* g = grid + (size_t)e * cfg.nx * cfg.ny;
    for (int ix = R.x0 + wv; ix < R.x1; ix += 4) {
      const double x = cfg.xmin + ix * cfg.dx - xc;
      for (int iy = R.y0 + lane; iy < R.y1; iy += 64) {
        const double d = (double)g[(size_t)ix * cfg.ny + iy];
        if (d == 0.0) continue;
        const double y = cfg.ymin + iy * cfg.dy - yc;
        // r^n cos(n theta), r^n sin(n theta) = Re, Im (x + i y)^n
        double zr = x, zi = y;
        for (int k = 1; k < n; ++k) {
          const double nr = zr * x - zi * y, ni = zr * y + zi * x;
          zr = nr;
          zi = ni;
        }
        re += zr * d;
        im += zi * d;
      }
    }
  }
  re = block_sum_d(re, red);
  im = block_sum_d(im, red);
  if (t == 0) {
    double* o = rot + (size_t)e * 4;
    o[0] = xc;
    o[1] = yc;
    o[2] = angle_on ? -atan2(-im, -re) / n : 0.0;
    o[3] = s0;
  }
}

__global__ __launch_bounds__(256) void k_rotate_src(DevCfg cfg, EvBufs buf, const double* rot) {
  __shared__ int s_bb[4];
  const int e = blockIdx.x, t = threadIdx.x, lane = t & 63;
  const int K = buf.counts[(size_t)e * 8 + 5];
  const double* r = rot + (size_t)e * 4;
  const double xc = r[0], yc = r[1], phi = r[2];
  const double cp = cos(phi), sp = sin(phi);
  int4* src = buf.src + (size_t)e * buf.src_cap;
  int4* swin = buf.swin + (size_t)e * buf.src_cap;
  if (t < 4) s_bb[t] = (t & 1) ? -1 : 0x7FFFFFFF;
  __syncthreads();
  int bb[4] = {0x7FFFFFFF, -1, 0x7FFFFFFF, -1};
  for (int k = t; k < K; k += 256) {
    int4 s = src[k];
    const double x0 = s.x * kH - xc, y0 = s.y * kH - yc;
    // Point3D::rotate(cos 0, phi): x = cphi x0 - sphi y0, y = sphi x0 + cphi y0
    const double x = cp * x0 - sp * y0, y = sp * x0 + cp * y0;
    s.x = (int)llrint(x / kH);
    s.y = (int)llrint(y / kH);
    src[k] = s;
    const bool box = s.w != kSrcPair;
    const int4 w = src_windows(cfg, s.x, s.y, box);
    swin[k] = w;
    const int v[4] = {w.z, w.w, box ? w.x : w.z, box ? w.y : w.w};
    for (int q = 0; q < 4; q += 2) {
      const int x0i = v[q] & 0xffff, x1i = v[q] >> 16, y0i = v[q + 1] & 0xffff, y1i = v[q + 1] >> 16;
      if (x0i < x1i && y0i < y1i) {
        bb[0] = min(bb[0], x0i);
        bb[1] = max(bb[1], x1i);
        bb[2] = min(bb[2], y0i);
        bb[3] = max(bb[3], y1i);
      }
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    bb[0] = min(bb[0], __shfl_xor(bb[0], off, 64));
    bb[1] = max(bb[1], __shfl_xor(bb[1], off, 64));
    bb[2] = min(bb[2], __shfl_xor(bb[2], off, 64));
    bb[3] = max(bb[3], __shfl_xor(bb[3], off, 64));
  }
  if (lane == 0) {
    atomicMin(&s_bb[0], bb[0]);
    atomicMax(&s_bb[1], bb[1]);
    atomicMin(&s_bb[2], bb[2]);
    atomicMax(&s_bb[3], bb[3]);
  }
  __syncthreads();
  if (t == 0) {
    int* c = buf.counts + (size_t)e * 8;
    const bool none = s_bb[0] >= s_bb[1];
    c[6] = none ? 0 : (s_bb[0] | (s_bb[1] << 16));
    c[7] = none ? 0 : (s_bb[2] | (s_bb[3] << 16));
  }
}

// Pass 1 of the sum: thread (cell, segment) adds the cell over the events
// of its segment in order, skipping excluded events and events whose source
// rectangle misses the cell (the smearing writes exact zeros there).  The
// rectangles of 256 events at a time are staged in LDS (an excluded event
// gets an empty one), and eight loads are issued before their adds, so each
// wave keeps eight in flight.
// With grid2 the added value is the product of the two grids' cells, in
// double (T_A T_B of the optional averages).
template <bool PROD>
__global__ __launch_bounds__(256) void k_accum_part(const float* grid, const float* grid2, const int* counts, int nx,
                                                    int ny, int nev, int per_seg, const unsigned char* accept,
                                                    double* part) {
  __shared__ int2 s_rect[256];
  const int ncell = nx * ny, t = threadIdx.x;
  const int c = blockIdx.x * 256 + t, seg = blockIdx.y;
  const bool on = c < ncell;
  const int ix = on ? c / ny : 0, iy = on ? c - ix * ny : 0;
  const int e0 = seg * per_seg, e1 = min(nev, e0 + per_seg);
  const float* g = grid + (on ? c : 0);
  const float* g2 = PROD ? grid2 + (on ? c : 0) : nullptr;
  auto cell = [&](int e, int k) -> double {
    const int2 r = s_rect[k];
    const bool in = ix >= (r.x & 0xffff) && ix < (r.x >> 16) && iy >= (r.y & 0xffff) && iy < (r.y >> 16);
    if (!in) return 0.0;
    if constexpr (PROD) return (double)g[(size_t)e * ncell] * (double)g2[(size_t)e * ncell];
    return (double)g[(size_t)e * ncell];
  };
  double s = 0.0;
  for (int base = e0; base < e1; base += 256) {
    const int n = min(256, e1 - base);
    __syncthreads();  // s_rect free
    if (t < n) {
      const int e = base + t;
      s_rect[t] = accept[e] ? make_int2(counts[(size_t)e * 8 + 6], counts[(size_t)e * 8 + 7]) : make_int2(0, 0);
    }
    __syncthreads();
    if (!on) continue;
    int k = 0;
    for (; k + 8 <= n; k += 8) {
      double v[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) v[j] = cell(base + k + j, k + j);
#pragma unroll
      for (int j = 0; j < 8; ++j) s += v[j];
    }
    for (; k < n; ++k) s += cell(base + k, k);
  }
  if (on) part[(size_t)seg * ncell + c] = s;
}

// Pass 2: sum[c] += ff * (segment partials added in segment order).
__global__ __launch_bounds__(256) void k_accum_fin(const double* part, int nseg, int ncell, double ff, double* sum) {
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= ncell) return;
  double s = 0.0;
  for (int g = 0; g < nseg; ++g) s += part[(size_t)g * ncell + c];
  sum[c] += ff * s;
}

// Spectators (MCnucl::getSpectators, MCnucl.cpp:1222-1245): the nucleons
// without a collision, projectile ones then target ones in collision order,
// as sources of kind A / B with the +-5w window of MCnucl.cpp:546-560, into
// their own source list (sb: the counts / src / swin of that list).
__global__ __launch_bounds__(256) void k_spectator_src(DevCfg cfg, EvBufs buf, EvBufs sb) {
  __shared__ int s_bb[4], s_cnt[4];
  const int e = blockIdx.x, t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const int* pos = buf.pos + (size_t)e * 2 * kAMax * 2;
  const int* wound = buf.wound + (size_t)e * 2 * kAMax;
  int4* src = sb.src + (size_t)e * sb.src_cap;
  int4* swin = sb.swin + (size_t)e * sb.src_cap;
  if (t < 4) s_bb[t] = (t & 1) ? -1 : 0x7FFFFFFF;
  int bb[4] = {0x7FFFFFFF, -1, 0x7FFFFFFF, -1};
  int base = 0;
  for (int nu = 0; nu < 2; ++nu) {
    const int A = cfg.nuc[nu].A;
    for (int i0 = 0; i0 < A; i0 += 256) {
      const int i = i0 + t;
      const bool spec = i < A && wound[nu * kAMax + i] == -1;  // -2: no nucleon in this slot
      const unsigned long long m = __ballot(spec);
      __syncthreads();
      if (lane == 0) s_cnt[wv] = __popcll(m);
      __syncthreads();
      int k = base;
      for (int q = 0; q < wv; ++q) k += s_cnt[q];
      k += __popcll(m & ((1ull << lane) - 1ull));
      if (spec) {
        const int xh = 2 * pos[(nu * kAMax + i) * 2], yh = 2 * pos[(nu * kAMax + i) * 2 + 1];
        src[k] = make_int4(xh, yh, __float_as_int(1.f), nu == 0 ? kSrcPartA : kSrcPartB);
        const int4 w = src_windows(cfg, xh, yh, false);
        swin[k] = w;
        const int x0 = w.z & 0xffff, x1 = w.z >> 16, y0 = w.w & 0xffff, y1 = w.w >> 16;
        if (x0 < x1 && y0 < y1) {
          bb[0] = min(bb[0], x0);
          bb[1] = max(bb[1], x1);
          bb[2] = min(bb[2], y0);
          bb[3] = max(bb[3], y1);
        }
      }
      base += s_cnt[0] + s_cnt[1] + s_cnt[2] + s_cnt[3];
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    bb[0] = min(bb[0], __shfl_xor(bb[0], off, 64));
    bb[1] = max(bb[1], __shfl_xor(bb[1], off, 64));
    bb[2] = min(bb[2], __shfl_xor(bb[2], off, 64));
    bb[3] = max(bb[3], __shfl_xor(bb[3], off, 64));
  }
  if (lane == 0) {
    atomicMin(&s_bb[0], bb[0]);
    atomicMax(&s_bb[1], bb[1]);
    atomicMin(&s_bb[2], bb[2]);
    atomicMax(&s_bb[3], bb[3]);
  }
  __syncthreads();
  if (t == 0) {
    int* c = sb.counts + (size_t)e * 8;
    for (int q = 0; q < 5; ++q) c[q] = buf.counts[(size_t)e * 8 + q];
    c[5] = base;
    const bool none = s_bb[0] >= s_bb[1];
    c[6] = none ? 0 : (s_bb[0] | (s_bb[1] << 16));
    c[7] = none ? 0 : (s_bb[2] | (s_bb[3] << 16));
  }
}

int launch_spectator_src(const DevCfg& cfg, const EvBufs& b, const EvBufs& sb, int nev, void* stream) {
  if (nev <= 0) return 0;
  hipLaunchKernelGGL(k_spectator_src, dim3(nev), dim3(256), 0, (hipStream_t)stream, cfg, b, sb);
  return (int)hipGetLastError();
}

int launch_cm_angle(const DevCfg& cfg, const EvBufs& b, const float* grid, const double* cm, int nu, int n,
                    int angle_on, int nev, double* rot, void* stream) {
  if (nev <= 0) return 0;
  hipLaunchKernelGGL(k_cm_angle, dim3(nev), dim3(256), 0, (hipStream_t)stream, cfg, b, grid, cm, nu, n, angle_on,
                     rot);
  return (int)hipGetLastError();
}

int launch_rotate_src(const DevCfg& cfg, const EvBufs& b, const double* rot, int nev, void* stream) {
  if (nev <= 0) return 0;
  hipLaunchKernelGGL(k_rotate_src, dim3(nev), dim3(256), 0, (hipStream_t)stream, cfg, b, rot);
  return (int)hipGetLastError();
}

int launch_accum_avg(const float* grid, const int* counts, int nx, int ny, int nev, const unsigned char* accept,
                     double ff, double* part, double* sum, void* stream, const float* grid2) {
  if (nev <= 0) return 0;
  const int ncell = nx * ny;
  const int per_seg = (nev + kAccumSegments - 1) / kAccumSegments;
  const int nseg = (nev + per_seg - 1) / per_seg;  // <= kAccumSegments: part holds kAccumSegments x ncell
  const unsigned bx = (unsigned)((ncell + 255) / 256);
  if (grid2)
    hipLaunchKernelGGL(k_accum_part<true>, dim3(bx, (unsigned)nseg), dim3(256), 0, (hipStream_t)stream, grid, grid2,
                       counts, nx, ny, nev, per_seg, accept, part);
  else
    hipLaunchKernelGGL(k_accum_part<false>, dim3(bx, (unsigned)nseg), dim3(256), 0, (hipStream_t)stream, grid,
                       nullptr, counts, nx, ny, nev, per_seg, accept, part);
  hipLaunchKernelGGL(k_accum_fin, dim3(bx), dim3(256), 0, (hipStream_t)stream, part, nseg, ncell, ff, sum);
  return (int)hipGetLastError();
}

}  // namespace smc
