// smc_kln.hip -- MC-KLN per-cell dN/dy on gfx950 (KLNModel.cpp:97-125,
// :219-277, :360-399; KLNfunc.h:10-17; KLNModel.h:91-96).
//
// At y = 0 the k_T-factorised integrand factorises as
//   func = P(pT) * alpha_s(max(k1^2, k2^2, pT^2)) * min(1, Qa^2/k1^2)
//          * min(1, Qb^2/k2^2) / (alpha_s(Qa^2) alpha_s(Qb^2))
// with Qa^2 = T_A g(pT), Qb^2 = T_B g(pT), g(pT) = (2/1.53)(0.01/x)^lambda
// and x = pT/sqrt(s).  Everything but the two min() factors and the two
// alpha_s(Q^2) factors is independent of the cell, so the host precomputes
// per node n of the (pT, kT, phi) product rule
//   c_n = w_n alpha_s(max(k1^2,k2^2,pT^2)),  u_n = 1/k1^2,  v_n = 1/k2^2
// (phi and 2 pi - phi merged) and per pT node P and g.  A cell then costs
//   sum_pT P / (alpha(a) alpha(b)) sum_n c_n min(1, a u_n) min(1, b v_n).
//
//  k_kln_direct  FP32 per cell (smc_kln_direct.hip).
//  k_kln_points  FP64 for arbitrary (T_A, T_B): KLNModel::getdNdy and the
//                dN/dy table of MCnucl::makeTable (MCnucl.cpp:911-960).
//  k_kln_table   MCnucl::setDensity for which_mc_model 1 (MCnucl.cpp:654-687):
//                sixPoint2dInterp of the table at (T_A/dT, T_B/dT), FP64.
#include "smc_device.h"

namespace smc {

__device__ __forceinline__ double alpha_s_d(double q2, double lq2, double amax, double inv_beta0) {
  if (q2 <= lq2) return amax;
  return fmin(amax, inv_beta0 / log(q2 / lq2));
}

// node layout: ptd[i] = {P, g} (double), nodes[i * nn + n] = {c, u, v} (double)
__global__ __launch_bounds__(256) void k_kln_points(const double* ta, const double* tb, double* out, int n,
                                                    const double2* ptd, const double4* nodes, int npt, int nn,
                                                    KlnConst kc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double A = ta[i], B = tb[i];
  if (!(A > 0.0 && B > 0.0)) {
    out[i] = 0.0;
    return;
  }
  double sum = 0.0;
  for (int p = 0; p < npt; ++p) {
    const double2 pg = ptd[p];
    const double a = A * pg.y, b = B * pg.y;
    double s = 0.0;
    for (int m = 0; m < nn; ++m) {
      const double4 q = nodes[(size_t)p * nn + m];
      s += q.x * fmin(1.0, a * q.y) * fmin(1.0, b * q.z);
    }
    sum += pg.x * s / (alpha_s_d(a, kc.lq2, kc.amax, kc.inv_beta0) * alpha_s_d(b, kc.lq2, kc.amax, kc.inv_beta0));
  }
  out[i] = sum;
}

// MCnucl.cpp:654-687: status bit set when a cell leaves the table (exit(1)
// in the reference).
__global__ __launch_bounds__(256) void k_kln_table(const float* ta, const float* tb, float* out, int ncell_ev,
                                                   int nev, int ny, const double* table, int tmax, double dT,
                                                   int* counts) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)ncell_ev * nev) return;
  // cells outside the event's source rectangle have T_A = T_B = 0, and the
  // table's row and column 0 are 0
  const size_t e = i / ncell_ev;
  const int c = (int)(i - e * ncell_ev), ix = c / ny, iy = c - ix * ny;
  const int rx = counts[e * 8 + 6], ry = counts[e * 8 + 7];
  if (ix < (rx & 0xffff) || ix >= (rx >> 16) || iy < (ry & 0xffff) || iy >= (ry >> 16)) {
    out[i] = 0.f;
    return;
  }
  const double di = (double)ta[i] / dT, dj = (double)tb[i] / dT;
  if (!(di >= 0 && di < tmax - 2) || !(dj >= 0 && dj < tmax - 2)) {  // also catches NaN
    out[i] = 0.f;
    atomicOr(&counts[(i / ncell_ev) * 8 + 3], kStKlnTable);
    return;
  }
  const int a = (int)floor(di), b = (int)floor(dj);
  const double x = di - a, y = dj - b;
  const double v00 = table[a * tmax + b], v01 = table[a * tmax + b + 1], v02 = table[a * tmax + b + 2];
  const double v10 = table[(a + 1) * tmax + b], v11 = table[(a + 1) * tmax + b + 1];
  const double v20 = table[(a + 2) * tmax + b];
  // arsenal.cpp:33-54
  const double axx = 1.0 / 2.0 * (v00 - 2 * v10 + v20);
  const double axy = v00 - v01 - v10 + v11;
  const double ayy = 1.0 / 2.0 * (v00 - 2 * v01 + v02);
  const double bx = 1.0 / 2.0 * (-3.0 * v00 + 4 * v10 - v20);
  const double by = 1.0 / 2.0 * (-3.0 * v00 + 4 * v01 - v02);
  out[i] = (float)(axx * x * x + axy * x * y + ayy * y * y + bx * x + by * y + v00);
}

int launch_kln_points(const double* ta, const double* tb, double* out, int n, const void* ptd, const void* nodes,
                      int npt, int nn, const KlnConst& kc, void* stream) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_kln_points, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, ta, tb, out, n,
                     (const double2*)ptd, (const double4*)nodes, npt, nn, kc);
  return (int)hipGetLastError();
}

int launch_kln_table(const float* ta, const float* tb, float* out, int ncell_ev, int nev, int ny, const double* table,
                     int tmax, double dT, int* counts, void* stream) {
  const size_t n = (size_t)ncell_ev * nev;
  if (n == 0) return 0;
  hipLaunchKernelGGL(k_kln_table, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, ta, tb, out,
                     ncell_ev, nev, ny, table, tmax, dT, counts);
  return (int)hipGetLastError();
}

}  // namespace smc
