// smc_diff.hip -- differentiable MC-Glauber batches at frozen collision
// topology (no counterpart in the reference, which has no derivative code).
//
// Parameters theta = (w, R, a, alpha, k): nucleon width, Woods-Saxon radius
// and diffuseness (both nuclei), two-component mixing, Gamma k = 1/theta.
// Frozen at theta0: the uniforms of every accepted nucleon, the impact
// parameter, participants and collision pairs, every source's integer
// footprint window with its r <= 5 w0 disk cut, and the hot-spot mask.
//
//  positions  r = F^{-1}(u; R, a), F the Woods-Saxon radial CDF truncated at
//             R + 2.5 (the law of Nucleus.cpp:578-621), times a fixed unit
//             direction; CM subtracted (Nucleus.cpp:296-309); dr/dR, dr/da
//             by implicit differentiation of F(r) = u.
//  weights    participant (1-alpha)/2 f, pair alpha f, f = G^{-1}(v; s)/s
//             with s the Gamma shapes of MCnucl.cpp:1271-1301; df/ds from
//             dG^{-1}/ds = -(dP/ds)/pdf.
//  density    the exact-footprint tile kernel with w in the Gaussian and the
//             frozen windows.
//  adjoint    cell cotangents G = dL/drho inside each event's source
//             rectangle (k_diff_cot_coef, k_diff_cot), one wavefront per
//             source gathering G against the recomputed Gaussians
//             (k_diff_gather), per-event chain rule to theta (k_diff_chain).
#include "smc_device.h"

namespace smc {

// ---------------------------------------------------------------- math
// N(r) = int_0^r t^2 rho(t) dt and its R and a derivatives, GL on [0, r].
struct WsInt {
  double n, nR, na;
};

__device__ inline WsInt ws_integrals(double r, double R, double a, const double2* gl, int ngl) {
  WsInt s{0.0, 0.0, 0.0};
  for (int j = 0; j < ngl; ++j) {
    const double t = r * gl[j].x, wt = r * gl[j].y;
    const double rho = 1.0 / (1.0 + exp((t - R) / a));
    const double dR = rho * (1.0 - rho) / a;
    const double t2w = t * t * wt;
    s.n += t2w * rho;
    s.nR += t2w * dR;
    s.na += t2w * dR * (t - R) / a;
  }
  return s;
}

// Normalisation Z = N(R + 2.5) and its R and a derivatives (the upper limit
// moves with R).
struct WsNorm {
  double n, R, a;
};
__device__ inline WsNorm ws_norm(double R, double a, const double2* gl, int ngl) {
  const double rmax = R + 2.5;
  const WsInt Z = ws_integrals(rmax, R, a, gl, ngl);
  const double rhomax = 1.0 / (1.0 + exp(2.5 / a));
  return WsNorm{Z.n, Z.nR + rmax * rmax * rhomax, Z.na};
}

// r = F^{-1}(u; R, a) on [0, R + 2.5] and dr/dR, dr/da.  tab (optional):
// kWsTab quantiles r at v = u^(1/3) = i/(kWsTab-1) for this (R, a), then
// {Z, Z_R, Z_a} (k_ws_table).  With it, Newton starts from the cubic
// interpolation of the table (error ~1e-12 fm) and usually stops after one
// step.  Newton stops once a step is below 1e-9 (R + 2.5): in the quadratic
// regime the root is then within ~1e-16 fm of the new point.  The R and a
// derivatives come from the last evaluation, within that step of r.
__device__ inline void ws_finv(double u, double R, double a, const double2* gl, int ngl, double& r, double& rR,
                               double& ra, const double* tab = nullptr) {
  const double rmax = R + 2.5;
  const WsNorm Z = tab ? WsNorm{tab[kWsTab], tab[kWsTab + 1], tab[kWsTab + 2]} : ws_norm(R, a, gl, ngl);
  const double target = u * Z.n;
  double lo = 0.0, hi = rmax, x = rmax * cbrt(u);
  if (tab) {
    const double v = cbrt(u) * (kWsTab - 1);
    const int i = min((int)v, kWsTab - 2);
    const double f = v - i;
    if (i >= 1 && i + 2 < kWsTab) {  // Lagrange cubic through nodes i-1 .. i+2
      const double wm = -f * (f - 1.0) * (f - 2.0) / 6.0, w0 = (f + 1.0) * (f - 1.0) * (f - 2.0) / 2.0;
      const double w1 = -(f + 1.0) * f * (f - 2.0) / 2.0, w2 = (f + 1.0) * f * (f - 1.0) / 6.0;
      x = wm * tab[i - 1] + w0 * tab[i] + w1 * tab[i + 1] + w2 * tab[i + 2];
    } else {
      x = tab[i] + f * (tab[i + 1] - tab[i]);
    }
  }
  WsInt N{0.0, 0.0, 0.0};
  double xe = x, rho = 0.0;  // last evaluation point and rho there
  for (int it = 0; it < 100; ++it) {
    N = ws_integrals(x, R, a, gl, ngl);
    const double fv = N.n - target;
    if (fv < 0.0) lo = x;
    else hi = x;
    rho = 1.0 / (1.0 + exp((x - R) / a));
    double xn = x - fv / (x * x * rho);
    // a step below the resolution of x (xn == x) is a converged root
    const bool newton = (xn > lo && xn < hi) || xn == x;
    if (!newton) xn = 0.5 * (lo + hi);
    const bool done = fabs(xn - x) <= (newton ? 1e-9 : 1e-14) * rmax;
    xe = x;
    x = xn;
    if (done) break;
  }
  r = x;
  const double Fr = xe * xe * rho / Z.n;
  const double FR = N.nR / Z.n - N.n * Z.R / (Z.n * Z.n);
  const double Fa = N.na / Z.n - N.n * Z.a / (Z.n * Z.n);
  rR = -FR / Fr;
  ra = -Fa / Fr;
}

// digamma by upward recurrence to z >= 12 and the asymptotic series
__device__ inline double digamma_d(double z) {
  double acc = 0.0;
  while (z < 12.0) {
    acc -= 1.0 / z;
    z += 1.0;
  }
  const double iz = 1.0 / z, iz2 = iz * iz;
  return acc + log(z) - 0.5 * iz -
         iz2 * (1.0 / 12 - iz2 * (1.0 / 120 - iz2 * (1.0 / 252 - iz2 * (1.0 / 240 - iz2 * (1.0 / 132)))));
}

// Regularised lower incomplete gamma P(s, x) by its power series
// P = e^{-x} x^s / Gamma(s+1) sum_n x^n / ((s+1)...(s+n)) and, when dPds is
// requested, dP/ds = P (ln x - psi(s+1)) - prefactor sum_n t_n H_n with
// H_n = sum_{m<=n} 1/(s+m).
__device__ inline double gamma_p(double s, double x, double lg1, double psi1, double* dPds) {
  if (x <= 0.0) {
    if (dPds) *dPds = 0.0;
    return 0.0;
  }
  const double lx = log(x);
  const double pref = exp(-x + s * lx - lg1);
  double t = 1.0, sum = 1.0, H = 0.0, dsum = 0.0;
  for (int n = 1; n < 2000; ++n) {
    const double sn = s + n;
    t *= x / sn;
    H += 1.0 / sn;
    sum += t;
    dsum += t * H;
    if (t < 1e-17 * sum && n > x) break;
  }
  const double P = pref * sum;
  if (dPds) *dPds = P * (lx - psi1) - pref * dsum;
  return P;
}

// Unit-scale Gamma quantile x(v; s) (P(s, x) = v) and dx/ds.  Halley's
// iteration (P' = pdf, P'' = pdf ((s-1)/x - 1)) from the small-x form
// x0 = (v Gamma(s+1))^(1/s), which lies left of the root for every s > 0
// (e^-x sum_n x^n / ((s+1)...(s+n)) < 1).  The bracket [lo, hi] is built
// from the evaluations; a step that leaves it bisects (or doubles while hi
// is open).  A Halley step below 1e-5 x leaves an error ~1e-15 x, so the
// iteration stops there: about 3-4 evaluations of P per quantile.
__device__ inline void gamma_qf(double v, double s, double& x, double& dxds) {
  if (!(v > 0.0)) {
    x = 0.0;
    dxds = 0.0;
    return;
  }
  const double lg1 = lgamma(s + 1.0), psi1 = digamma_d(s + 1.0), lgs = lg1 - log(s);
  double lo = 0.0, hi = HUGE_VAL;
  double xx = exp((log(v) + lg1) / s);
  for (int it = 0; it < 200; ++it) {
    const double f = gamma_p(s, xx, lg1, psi1, nullptr) - v;
    if (f < 0.0) lo = xx;
    else hi = xx;
    const double pdf = exp(-xx + (s - 1.0) * log(xx) - lgs);
    const double d2 = pdf * ((s - 1.0) / xx - 1.0);
    const double den = 2.0 * pdf * pdf - f * d2;
    double xn = (den > 0.0) ? xx - 2.0 * f * pdf / den : xx - f / pdf;
    const bool halley = (xn > lo && xn < hi) || xn == xx;
    if (!halley) xn = (hi < HUGE_VAL) ? 0.5 * (lo + hi) : 2.0 * xx;
    const bool done = fabs(xn - xx) <= (halley ? 1e-5 : 1e-15) * xx;
    xx = xn;
    if (done) break;
  }
  x = xx;
  double dP = 0.0;
  gamma_p(s, x, lg1, psi1, &dP);
  const double pdf = exp(-x + (s - 1.0) * log(x) - lgs);
  dxds = -dP / pdf;
}

// Gamma shape of a source kind (MCnucl.cpp:1271-1301; k = 1/theta)
__device__ __forceinline__ double diff_shape(int kind, double alpha, double kg) {
  return (kind == kSrcPair) ? (alpha + 1e-8) * kg : (1.0 - alpha + 1e-8) * 0.5 * kg;
}

// The radial quantile at v = u^(1/3) = i/(kWsTab-1), i < kWsTab; thread 0
// also stores {Z, Z_R, Z_a} after the table.
__global__ __launch_bounds__(256) void k_ws_table(double R, double a, const double2* gl, int ngl, double* tab) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= kWsTab) return;
  if (i == 0) {
    const WsNorm z = ws_norm(R, a, gl, ngl);
    tab[kWsTab] = z.n;
    tab[kWsTab + 1] = z.R;
    tab[kWsTab + 2] = z.a;
  }
  const double v = (double)i / (kWsTab - 1);
  double r, rR, ra;
  ws_finv(v * v * v, R, a, gl, ngl, r, rR, ra);
  tab[i] = r;
}

// ------------------------------------------------------------- topology
// One workgroup per pending event: b, then each nucleus (one lane each)
// from candidates (u, cos theta, phi) in stream order with the hard core,
// r from the radial quantile at theta0.  Writes the quantised positions
// sorted by (x, index) to raw (fed-position input of k_event) and the
// candidates' uniforms in the same order.
__global__ __launch_bounds__(64) void k_diff_topo(DevCfg cfg, EvBufs buf, DiffBufs d, uint2 key,
                                                  unsigned long long event0, double R0, double a0) {
  const int e = blockIdx.x;
  const int lane = threadIdx.x;
  if (!d.pending[e]) return;
  const unsigned attempt = (unsigned)d.attempt[e];
  const unsigned long long gev = event0 + (unsigned long long)e;
  __shared__ double px[2][kAMax], py[2][kAMax], pz[2][kAMax];
  __shared__ double uu[2][kAMax][3];
  __shared__ int ok[2];
  __shared__ int s_xc;
  __shared__ int ord[2][kAMax], xq[2][kAMax];
  if (lane == 0) {
    const uint4 r = draw(key, gev, kStrDiffB, 0, 0, attempt, 0);
    const double b = sqrt((cfg.bmax2 - cfg.bmin2) * u53(r.x, r.y) + cfg.bmin2);
    s_xc = (int)llrint(b * 0.5 / kQ);
    d.bval[e] = 2.0 * (double)s_xc * kQ;
    buf.bval[e] = 2.0 * (double)s_xc * kQ;
  }
  __syncthreads();
  if (lane < 2) {
    const int nu = lane;
    const int A = cfg.nuc[nu].A;
    int acc = 0;
    const int max_cand = 64 * A + 64;
    for (int j = 0; j < max_cand && acc < A; ++j) {
      const uint4 r1 = draw(key, gev, kStrDiffNucleon, nu, j, attempt, 0);
      const uint4 r2 = draw(key, gev, kStrDiffNucleon, nu, j, attempt, 1);
      const double u = u53(r1.x, r1.y);
      const double c = 2.0 * u53(r2.x, r2.y) - 1.0;
      const double ph = 2.0 * M_PI * u53(r2.z, r2.w);
      double r, rR, ra;
      ws_finv(u, R0, a0, d.gl, d.ngl, r, rR, ra, d.wstab);
      const double st = sqrt(fmax(0.0, 1.0 - c * c));
      const double x = r * st * cos(ph), y = r * st * sin(ph), z = r * c;
      bool good = true;
      for (int m = 0; m < acc && good; ++m) {
        const double dx = x - px[nu][m], dy = y - py[nu][m], dz = z - pz[nu][m];
        if (dx * dx + dy * dy + dz * dz < (double)cfg.dmin2) good = false;
      }
      if (!good) continue;
      px[nu][acc] = x;
      py[nu][acc] = y;
      pz[nu][acc] = z;
      uu[nu][acc][0] = u;
      uu[nu][acc][1] = c;
      uu[nu][acc][2] = ph;
      ++acc;
    }
    ok[nu] = acc == A;
    if (acc < A) {
      // no topology this attempt: nuclei far apart (no collision), so the
      // collision step rejects the event and it is drawn again
      int* raw = buf.raw + ((size_t)e * 2 + nu) * kAMax * 2;
      const double x0 = (nu == 0) ? -55.0 : 5.0;
      for (int q = 0; q < A; ++q) {
        raw[2 * q] = (int)llrint((x0 + 0.19 * q) / kQ);
        raw[2 * q + 1] = 0;
      }
    } else {
      double sx = 0.0, sy = 0.0;
      for (int i = 0; i < A; ++i) {
        sx += px[nu][i];
        sy += py[nu][i];
      }
      const double xcen = (nu == 0 ? 1.0 : -1.0) * (double)s_xc * kQ;
      int* raw = buf.raw + ((size_t)e * 2 + nu) * kAMax * 2;
      double* du = d.du + ((size_t)e * 2 + nu) * kAMax * 4;
      // quantised positions, then a stable insertion sort by x
      for (int i = 0; i < A; ++i) {
        xq[nu][i] = (int)llrint((px[nu][i] - sx / A + xcen) / kQ);
        int p = i;
        while (p > 0 && xq[nu][ord[nu][p - 1]] > xq[nu][i]) {
          ord[nu][p] = ord[nu][p - 1];
          --p;
        }
        ord[nu][p] = i;
      }
      for (int q = 0; q < A; ++q) {
        const int i = ord[nu][q];
        raw[2 * q] = xq[nu][i];
        raw[2 * q + 1] = (int)llrint((py[nu][i] - sy / A) / kQ);
        du[4 * q] = uu[nu][i][0];
        du[4 * q + 1] = uu[nu][i][1];
        du[4 * q + 2] = uu[nu][i][2];
        du[4 * q + 3] = 0.0;
      }
    }
  }
  __syncthreads();
  if (lane == 0) d.topo_ok[e] = ok[0] && ok[1];
}

// ------------------------------------------------------------- forward
// Positions at theta and their R, a derivatives (CM included).
// dpos[((e*2+nu)*kAMax + i)*6 + {x, y, x_R, y_R, x_a, y_a}]
__global__ __launch_bounds__(256) void k_diff_pos(DevCfg cfg, DiffBufs d, double R, double a) {
  const int e = blockIdx.x, t = threadIdx.x;
  __shared__ double red[4];
  for (int nu = 0; nu < 2; ++nu) {
    const int A = cfg.nuc[nu].A;
    double v[6] = {0, 0, 0, 0, 0, 0};
    if (t < A) {
      const double* du = d.du + (((size_t)e * 2 + nu) * kAMax + t) * 4;
      double r, rR, ra;
      ws_finv(du[0], R, a, d.gl, d.ngl, r, rR, ra, d.wstab);
      const double st = sqrt(fmax(0.0, 1.0 - du[1] * du[1]));
      const double nx = st * cos(du[2]), ny = st * sin(du[2]);
      v[0] = r * nx;
      v[1] = r * ny;
      v[2] = rR * nx;
      v[3] = rR * ny;
      v[4] = ra * nx;
      v[5] = ra * ny;
    }
    double m[6];
    for (int q = 0; q < 6; ++q) m[q] = block_sum_d(v[q], red) / A;
    if (t < A) {
      const double xcen = (nu == 0 ? 0.5 : -0.5) * d.bval[e];
      double* o = d.dpos + (((size_t)e * 2 + nu) * kAMax + t) * 6;
      o[0] = v[0] - m[0] + xcen;
      o[1] = v[1] - m[1];
      for (int q = 2; q < 6; ++q) o[q] = v[q] - m[q];
    }
  }
}

// Source positions (half quanta) and Gamma weights at theta into src; f
// and df/ds per source into dsrc.
__global__ __launch_bounds__(256) void k_diff_src(DevCfg cfg, EvBufs buf, DiffBufs d, uint2 key,
                                                  unsigned long long event0, double alpha, double kg, int fluct) {
  const int e = blockIdx.x, t = threadIdx.x;
  const int* cnt = buf.counts + (size_t)e * 8;
  const int npa = cnt[0], npb = cnt[1], ncoll = cnt[2], K = cnt[5];
  __shared__ int inv[2][kAMax];
  for (int nu = 0; nu < 2; ++nu) {
    if (t < cfg.nuc[nu].A) {
      const int w = buf.wound[(size_t)e * 2 * kAMax + nu * kAMax + t];
      if (w >= 0) inv[nu][w] = t;
    }
  }
  __syncthreads();
  const unsigned long long gev = event0 + (unsigned long long)e;
  int4* src = buf.src + (size_t)e * buf.src_cap;
  double2* ds = d.dsrc + (size_t)e * buf.src_cap;
  const double* pos = d.dpos + (size_t)e * 2 * kAMax * 6;
  for (int k = t; k < K && k < npa + npb + ncoll; k += 256) {
    int4 s = src[k];
    double x, y;
    if (k < npa + npb) {
      const int nu = (k < npa) ? 0 : 1;
      const int i = inv[nu][nu ? k - npa : k];
      x = pos[((size_t)nu * kAMax + i) * 6];
      y = pos[((size_t)nu * kAMax + i) * 6 + 1];
    } else {
      const unsigned pr = buf.pairs[(size_t)e * cfg.max_pairs + (k - npa - npb)];
      const int i = (int)(pr & 0xFFFFu), j = (int)(pr >> 16);
      x = 0.5 * (pos[(size_t)i * 6] + pos[((size_t)kAMax + j) * 6]);
      y = 0.5 * (pos[(size_t)i * 6 + 1] + pos[((size_t)kAMax + j) * 6 + 1]);
    }
    s.x = (int)llrint(x / kH);
    s.y = (int)llrint(y / kH);
    double f = 1.0, dfds = 0.0;
    if (fluct) {
      const uint4 r = draw(key, gev, kStrDiffGamma, (unsigned)s.w, (unsigned)k, 0, 0);
      const double v = u53(r.x, r.y);
      const double sh = diff_shape(s.w, alpha, kg);
      double q, dq;
      gamma_qf(v, sh, q, dq);
      f = q / sh;
      dfds = dq / sh - q / (sh * sh);
      d.gam_v[(size_t)e * buf.src_cap + k] = v;
    }
    s.z = __float_as_int((float)f);
    src[k] = s;
    ds[k] = make_double2(f, dfds);
  }
}

// ------------------------------------------------------------- backward
// Cell cotangents G = dL/drho for L = sum_e lam2 |eps2| + lam3 |eps3| +
// mu dS/dy, from the per-event scalars of k_diff_moments (layout kDs*).
// Per event: the coefficients of the cell cotangents -- lam cx / S,
// lam cy / S (the centre's dependence), lam Ar / (Am N), lam Ai / (Am N),
// lam Am / N^2 (the masked g_n terms) for n = 2, 3, and c0 (dS/dy).
__global__ __launch_bounds__(64) void k_diff_cot_coef(DiffBufs d, const double* cot, double ffdxdy, int nev) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nev) return;
  const double* sc = d.dsc + (size_t)e * kDsStride;
  double* out = d.cotc + (size_t)e * kCotStride;
  const double S = sc[kDsS], N = sc[kDsN] + 1e-15;
  const double dNx = -2.0 * sc[kDsMx], dNy = -2.0 * sc[kDsMy];
  out[0] = cot[(size_t)e * 3 + 2] * ffdxdy;
  for (int q = 0; q < 2; ++q) {
    const double lam = cot[(size_t)e * 3 + q];
    const double Ar = sc[kDsA2r + 2 * q], Ai = sc[kDsA2i + 2 * q];
    const double Am = sqrt(Ar * Ar + Ai * Ai);
    double* f = out + 1 + 5 * q;
    f[0] = f[1] = f[2] = f[3] = f[4] = 0.0;
    if (lam == 0.0 || !(Am > 0.0)) continue;
    double dAxr, dAxi, dAyr, dAyi;
    if (q == 0) {  // d(z^2)/dx = 2z, d(z^2)/dy = 2iz
      dAxr = -2.0 * sc[kDsMx];
      dAxi = -2.0 * sc[kDsMy];
      dAyr = 2.0 * sc[kDsMy];
      dAyi = -2.0 * sc[kDsMx];
    } else {
      dAxr = -sc[kDsB3xr];
      dAxi = -sc[kDsB3xi];
      dAyr = -sc[kDsB3yr];
      dAyi = -sc[kDsB3yi];
    }
    const double cx = (Ar * dAxr + Ai * dAxi) / (Am * N) - Am * dNx / (N * N);
    const double cy = (Ar * dAyr + Ai * dAyi) / (Am * N) - Am * dNy / (N * N);
    f[0] = lam * cx / S;
    f[1] = lam * cy / S;
    f[2] = lam * Ar / (Am * N);
    f[3] = lam * Ai / (Am * N);
    f[4] = lam * Am / (N * N);
  }
}

// Cell cotangents G = dL/drho inside each event's source rectangle -- the
// cells the gather reads (every source window lies in it); kCotBlocks
// workgroups per event stride over the rectangle's cells.
constexpr int kCotBlocks = 64;
__global__ __launch_bounds__(256) void k_diff_cot(DevCfg cfg, EvBufs buf, DiffBufs d, float* G) {
  __shared__ double cf[kCotStride];
  const int e = blockIdx.y;
  const int nx = cfg.nx, ny = cfg.ny, nwy = (ny + 31) >> 5;
  const double* sc = d.dsc + (size_t)e * kDsStride;
  if (threadIdx.x < kCotStride) cf[threadIdx.x] = d.cotc[(size_t)e * kCotStride + threadIdx.x];
  __syncthreads();
  const int rx = buf.counts[(size_t)e * 8 + 6], ry = buf.counts[(size_t)e * 8 + 7];
  const int x0 = rx & 0xffff, x1 = rx >> 16, y0 = ry & 0xffff, y1 = ry >> 16;
  const int wy = y1 - y0;
  const size_t ncr = (x1 > x0 && wy > 0) ? (size_t)(x1 - x0) * wy : 0;
  const double c0 = cf[0], ax = cf[1] + cf[6], ay = cf[2] + cf[7];
  const double xc = sc[kDsXc], yc = sc[kDsYc];
  const bool third = cf[8] != 0.0 || cf[9] != 0.0 || cf[10] != 0.0;
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < ncr; k += (size_t)kCotBlocks * blockDim.x) {
    const int ix = x0 + (int)(k / wy), iy = y0 + (int)(k % wy);
    const double x = cfg.xmin + ix * cfg.dx - xc, y = cfg.ymin + iy * cfg.dy - yc;
    const double r2 = x * x + y * y;
    const bool m = (d.mask[(size_t)e * nx * nwy + (size_t)ix * nwy + (iy >> 5)] >> (iy & 31)) & 1u;
    double g = c0;
    g += ax * x + ay * y;
    if (m && r2 > 0.0) {
      // g_n(z) = z^n r^(2-n)
      const double gr = x * x - y * y, gi = 2.0 * x * y;
      g += cf[3] * gr + cf[4] * gi - cf[5] * r2;
      if (third) {
        const double ir = 1.0 / sqrt(r2);
        const double tr = (gr * x - gi * y) * ir, ti = (gr * y + gi * x) * ir;
        g += cf[8] * tr + cf[9] * ti - cf[10] * r2;
      }
    }
    G[(size_t)e * nx * ny + (size_t)ix * ny + iy] = (float)g;
  }
}

// One wavefront per source: sums of G E, G E dx, G E dy, G E d^2 over the
// frozen footprint (E = exp(-d^2 / 2w^2)), then dL/dW, dL/dx, dL/dy, dL/dw.
// kGatherBlocks workgroups per event loop over its sources; workgroup ids
// are mapped so that all of an event's workgroups run on one XCD (one L2
// holds its cotangent grid).  Per chunk of 64 window rows, lane j computes
// row j's offset and Gaussian once; the row loop reads them with readlane.
constexpr int kGatherBlocks = 32;
__global__ __launch_bounds__(256) void k_diff_gather(DevCfg cfg, EvBufs buf, SmearSpec spec, DiffBufs d,
                                                     const float* G, double w, double alpha, int nev) {
  const unsigned L = blockIdx.x, rg = L % (8u * kGatherBlocks);
  const int e = (int)(8u * (L / (8u * kGatherBlocks)) + rg % 8u), kb = (int)(rg / 8u);
  if (e >= nev) return;
  const int lane = threadIdx.x & 63;
  const int K = buf.counts[(size_t)e * 8 + 5];
  const int nx = cfg.nx, ny = cfg.ny;
  const float c2 = (float)(1.0 / (2.0 * w * w) * 1.4426950408889634);
  const float cut2 = cfg.cut2;
  const float* g = G + (size_t)e * nx * ny;
  for (int k = kb * 4 + (threadIdx.x >> 6); k < K; k += 4 * kGatherBlocks) {
    const int4 s = buf.src[(size_t)e * buf.src_cap + k];
    const int4 wv = buf.swin[(size_t)e * buf.src_cap + k];
    const int kind = s.w;
    const int xw = (spec.win[kind] == kWinBox) ? wv.x : wv.z;
    const int yw = (spec.win[kind] == kWinBox) ? wv.y : wv.w;
    const int xlo = xw & 0xffff, xhi = xw >> 16, ylo = yw & 0xffff, yhi = yw >> 16;
    // The Gaussian factorises, G(dx) G(dy), and dx depends on the row only,
    // dy on the column only: each lane (a column) sums over the rows
    //   P = sum_r Ex_r g_rc,  Q = sum_r Ex_r dx_r g_rc,  R = sum_r Ex_r dx_r^2 g_rc
    // and folds them with its own Ey and dy once per column block:
    //   e0 = Ey P,  e1 = Ey Q,  e2 = Ey dy P,  e3 = Ey (R + dy^2 P).
    float e0 = 0.f, e1 = 0.f, e2 = 0.f, e3 = 0.f;
    for (int r0 = xlo; r0 < xhi; r0 += 64) {
      const int nr = min(64, xhi - r0);
      float dxl, exl, liml;  // row r0 + lane: offset from the source, its Gaussian, cut2 - dx^2
      {
        const int rowh = (int)llrint((cfg.xmin + (r0 + lane) * cfg.dx) / kH);
        dxl = (float)(rowh - s.x) * kHf;
        exl = __builtin_amdgcn_exp2f(-(dxl * dxl) * c2);
        liml = cut2 - dxl * dxl;
      }
      for (int c0 = ylo; c0 < yhi; c0 += 64) {
        const int col = c0 + lane;
        const bool on = col < yhi;
        const int ygh = (int)llrint((cfg.ymin + col * cfg.dy) / kH);
        const float dy = (float)(ygh - s.y) * kHf;
        const float dy2 = dy * dy;
        const float* gr = g + (size_t)r0 * ny + (on ? col : ylo);
        float P = 0.f, Q = 0.f, R = 0.f;
        auto row = [&](int j, float gv) {
          const float dx = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(dxl), j));
          const float Ex = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(exl), j));
          const float lim = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(liml), j));
          const float t = (on && dy2 <= lim) ? gv * Ex : 0.f;
          P += t;
          const float tx = t * dx;
          Q += tx;
          R = fmaf(tx, dx, R);
        };
        int j = 0;
        for (; j + 4 <= nr; j += 4) {  // four rows' loads in flight, added in row order
          float gv[4];
#pragma unroll
          for (int q = 0; q < 4; ++q) gv[q] = gr[(size_t)(j + q) * ny];
#pragma unroll
          for (int q = 0; q < 4; ++q) row(j + q, gv[q]);
        }
        for (; j < nr; ++j) row(j, gr[(size_t)j * ny]);
        const float Ey = __builtin_amdgcn_exp2f(-dy2 * c2);
        e0 = fmaf(Ey, P, e0);
        e1 = fmaf(Ey, Q, e1);
        e2 = fmaf(Ey * dy, P, e2);
        e3 = fmaf(Ey, fmaf(dy2, P, R), e3);
      }
    }
    const double s0 = wave_sum((double)e0), s1 = wave_sum((double)e1), s2 = wave_sum((double)e2),
                 s3 = wave_sum((double)e3);
    if (lane == 0) {
      const double norm = 1.0 / (2.0 * M_PI * w * w);
      const double f = d.dsrc[(size_t)e * buf.src_cap + k].x;
      const double W = ((kind == kSrcPair) ? alpha : 0.5 * (1.0 - alpha)) * f;
      double4 o;
      o.x = norm * s0;                                     // dL/dW
      o.y = W * norm * s1 / (w * w);                       // dL/dx_source
      o.z = W * norm * s2 / (w * w);                       // dL/dy_source
      o.w = W * norm * (s3 / (w * w * w) - 2.0 * s0 / w);  // dL/dw through this source
      d.gsrc[(size_t)e * buf.src_cap + k] = o;
    }
  }
}

// Per-event chain rule: gev[e][5] = dL/d(w, R, a, alpha, k).
__global__ __launch_bounds__(256) void k_diff_chain(DevCfg cfg, EvBufs buf, DiffBufs d, double alpha, double kg) {
  const int e = blockIdx.x, t = threadIdx.x;
  __shared__ double red[4];
  const int* cnt = buf.counts + (size_t)e * 8;
  const int npa = cnt[0], npb = cnt[1], ncoll = cnt[2], K = cnt[5];
  const double4* gs = d.gsrc + (size_t)e * buf.src_cap;
  const double2* ds = d.dsrc + (size_t)e * buf.src_cap;
  double gw = 0.0, ga = 0.0, gk = 0.0;
  for (int k = t; k < K; k += 256) {
    const double4 o = gs[k];
    gw += o.w;
    const double f = ds[k].x, fs = ds[k].y;
    if (k >= npa + npb) {  // W = alpha f(s), s = (alpha + eps) k
      ga += o.x * (f + alpha * fs * kg);
      gk += o.x * alpha * fs * (alpha + 1e-8);
    } else {  // W = (1 - alpha)/2 f(s), s = (1 - alpha + eps) k / 2
      ga += o.x * (-0.5 * f - 0.5 * (1.0 - alpha) * fs * 0.5 * kg);
      gk += o.x * 0.5 * (1.0 - alpha) * fs * 0.5 * (1.0 - alpha + 1e-8);
    }
  }
  // positions: participants carry their source, pairs half to each nucleon;
  // the pair list goes through LDS 256 entries at a time and every thread
  // (nucleon t of each nucleus) adds its pairs in pair order
  __shared__ unsigned lp[256];
  const unsigned* pairs = buf.pairs + (size_t)e * cfg.max_pairs;
  double gx[2] = {0.0, 0.0}, gy[2] = {0.0, 0.0};
  int wnd[2];
  for (int nu = 0; nu < 2; ++nu) {
    wnd[nu] = (t < cfg.nuc[nu].A) ? buf.wound[(size_t)e * 2 * kAMax + nu * kAMax + t] : -1;
    if (wnd[nu] >= 0) {
      const double4 o = gs[nu ? npa + wnd[nu] : wnd[nu]];
      gx[nu] += o.y;
      gy[nu] += o.z;
    }
  }
  const int np = min(ncoll, cfg.max_pairs);
  for (int q0 = 0; q0 < np; q0 += 256) {
    __syncthreads();
    if (q0 + t < np) lp[t] = pairs[q0 + t];
    __syncthreads();
    const int m = min(256, np - q0);
    for (int q = 0; q < m; ++q) {
      const unsigned pr = lp[q];
      const bool a = wnd[0] >= 0 && (int)(pr & 0xFFFFu) == t, b = wnd[1] >= 0 && (int)(pr >> 16) == t;
      if (a || b) {
        const double4 o = gs[npa + npb + q0 + q];
        if (a) {
          gx[0] += 0.5 * o.y;
          gy[0] += 0.5 * o.z;
        }
        if (b) {
          gx[1] += 0.5 * o.y;
          gy[1] += 0.5 * o.z;
        }
      }
    }
  }
  double gR = 0.0, gA = 0.0;
  for (int nu = 0; nu < 2; ++nu) {
    if (t < cfg.nuc[nu].A) {
      const double* p = d.dpos + (((size_t)e * 2 + nu) * kAMax + t) * 6;
      gR += gx[nu] * p[2] + gy[nu] * p[3];
      gA += gx[nu] * p[4] + gy[nu] * p[5];
    }
  }
  const double v[5] = {gw, gR, gA, ga, gk};
  for (int q = 0; q < 5; ++q) {
    const double sum = block_sum_d(v[q], red);
    if (t == 0) d.gev[(size_t)e * 5 + q] = sum;
  }
}

// ------------------------------------------------------------ launchers
int launch_diff_topo(const DevCfg& cfg, const EvBufs& b, const DiffBufs& d, int nev, unsigned long long seed,
                     unsigned long long event0, double R0, double a0, void* stream) {
  const uint2 key = make_uint2((unsigned)seed, (unsigned)(seed >> 32));
  hipLaunchKernelGGL(k_ws_table, dim3((kWsTab + 255) / 256), dim3(256), 0, (hipStream_t)stream, R0, a0, d.gl, d.ngl,
                     d.wstab);
  hipLaunchKernelGGL(k_diff_topo, dim3(nev), dim3(64), 0, (hipStream_t)stream, cfg, b, d, key, event0, R0, a0);
  return (int)hipGetLastError();
}

int launch_diff_forward_pre(const DevCfg& cfg, const EvBufs& b, const DiffBufs& d, int nev, unsigned long long seed,
                            unsigned long long event0, const double* theta, int fluct, void* stream) {
  const uint2 key = make_uint2((unsigned)seed, (unsigned)(seed >> 32));
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_ws_table, dim3((kWsTab + 255) / 256), dim3(256), 0, st, theta[1], theta[2], d.gl, d.ngl,
                     d.wstab);
  hipLaunchKernelGGL(k_diff_pos, dim3(nev), dim3(256), 0, st, cfg, d, theta[1], theta[2]);
  hipLaunchKernelGGL(k_diff_src, dim3(nev), dim3(256), 0, st, cfg, b, d, key, event0, theta[3], theta[4], fluct);
  return (int)hipGetLastError();
}

int launch_diff_backward(const DevCfg& cfg, const EvBufs& b, const SmearSpec& spec, const DiffBufs& d, int nev,
                         const double* cot_dev, double ffdxdy, float* G, const double* theta, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_diff_cot_coef, dim3((unsigned)((nev + 63) / 64)), dim3(64), 0, st, d, cot_dev, ffdxdy, nev);
  hipLaunchKernelGGL(k_diff_cot, dim3(kCotBlocks, nev), dim3(256), 0, st, cfg, b, d, G);
  const unsigned nev8 = (unsigned)((nev + 7) / 8 * 8);  // k_diff_gather: whole groups of 8 events (one per XCD)
  hipLaunchKernelGGL(k_diff_gather, dim3(nev8 * kGatherBlocks), dim3(256), 0, st, cfg, b, spec, d, (const float*)G,
                     theta[0], theta[3], nev);
  hipLaunchKernelGGL(k_diff_chain, dim3(nev), dim3(256), 0, st, cfg, b, d, theta[3], theta[4]);
  return (int)hipGetLastError();
}

}  // namespace smc
