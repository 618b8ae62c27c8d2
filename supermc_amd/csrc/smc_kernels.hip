// smc_kernels.hip -- gfx950 kernels for superMC's per-event initial-condition
// hot path.
//
//  k_event    one workgroup per event: impact parameter, Philox sampling of
//             both nuclei (Woods-Saxon with hard core, deformed WS, Hulthen
//             deuteron, proton), CM recentering, x-sort, A x B disk test with
//             exact integer predicate, participant / pair compaction in the
//             reference's order, Gamma fluctuation factors, source list.
//             (MakeDensity.cpp:2143-2162, Nucleus.cpp:187-317,
//              MCnucl.cpp:217-352, MCnucl.cpp:1271-1301)
//  k_smear    2D Gaussian smearing of a source list onto the grid, one
//             workgroup per (event, 16x64 tile), exact reference footprints
//             via integer threshold tables.  (MCnucl.cpp:398-478, 626-866)
//  (moments: smc_moments.hip; the default smearing: smc_tile.hip)
#include <hip/hip_runtime.h>

#include "smc_device.h"

namespace smc {

// ------------------------------------------------------------- k_event
// GAUSS: collision criterion 2 (the disk criterion's instance carries none
// of its code).  SPEC: the instance with the second sampling slot for
// proton / deuteron projectiles (launched only when a system of the batch
// has one; the other instance keeps the registers that slot costs).
template <bool GAUSS, bool SPEC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5, 8))) void k_event(DevCfg cfg, EvBufs buf, uint2 key,
                                               unsigned long long event0, int mode) {
  // mode 0: impact parameter sampled, events accepted as MakeDensity.cpp:2147-2162
  // mode 1: fed positions; mode 2: sampled nuclei at the given b, one attempt
  const int fed = mode == 1;
  const int e = blockIdx.x;
  const int t = threadIdx.x;
  const int lane = t & 63, wv = t >> 6;
  const unsigned long long gev = event0 + (unsigned long long)e;

  __shared__ __align__(16) int px[2][kAMax], py[2][kAMax];  // input/sample order (quanta)
  __shared__ __align__(16) float pz[2][kAMax];              // local z * 2^24 (hard core only)
  __shared__ __align__(16) int cqx[2][64], cqy[2][64];       // candidates of a round
  __shared__ __align__(16) float cqz[2][64];
  __shared__ int s_nc[2], s_wc[2][4];
  __shared__ unsigned long long skey[2][kAMax];  // sort keys
  __shared__ int wrank[kAMax];
  __shared__ int s_bb[4];
  __shared__ unsigned long long s_okm[4];
  __shared__ unsigned long long s_cm[4][64];  // conflict masks of a round, per wave
  __shared__ __align__(16) int sx[2][kAMax], sy[2][kAMax];  // sorted
  __shared__ unsigned hits[kAMax][kAMax / 32];
  __shared__ int rowcnt[kAMax], fh[kAMax], scan_tmp[256];
  __shared__ int s_acc[2];
  __shared__ double s_b;
  __shared__ int s_xc;
  __shared__ int s_xc2;    // b / 2 of the next attempt (two attempts per pass)
  __shared__ double s_b2;
  __shared__ int lpx[2], lpy[2];  // the light projectile when slot 0 holds a target
  __shared__ double s_red[2][2];
  __shared__ NucSys s_nuc[2];  // this event's nuclei

  // sampled events of a mixed batch take their system's nuclei; fed
  // positions keep the context's
  if (t < 2) {
    if (cfg.sys_n > 0 && !fed) s_nuc[t] = cfg.sys[2 * cfg.sys_of[gev % (unsigned long long)cfg.sys_n] + t];
    else s_nuc[t] = t ? cfg.nuc[1] : cfg.nuc[0];
  }
  __syncthreads();
  const int A = s_nuc[0].A, B = s_nuc[1].A;
  const int NN[2] = {A, B};
  int attempt = 0, status = kStOk;
  int npa = 0, npb = 0, ncoll = 0;

  // Two attempts per sampling pass for a proton or deuteron on a
  // Woods-Saxon target (sampled b, the SPEC instance).  The two waves that
  // serve nucleus 0 would sit idle while the target is drawn; instead they
  // draw, into slot 0, the target of the next attempt from that attempt's
  // own counters, and the projectile lives in lpx / lpy.  A rejected
  // attempt hands slot 0 to the next one, which then needs no sampling
  // pass.  Every event is the one the attempts taken one by one give.
  const bool spec = SPEC && mode == 0 && s_nuc[0].kind != kWoodsSaxon && s_nuc[1].kind == kWoodsSaxon;
  const int NS[2] = {spec ? B : A, B};  // nucleons drawn into each slot
  const int sys0 = spec ? 1 : 0;        // nucleus whose counters slot 0 draws
  bool pending = false;                 // slot 0 holds this attempt's target

  // proton / deuteron nucleus nu_ of attempt att centred at xcen
  // (Nucleus.cpp:199-210, 342-359): one lane writes X, Y (and Z) and
  // returns the nucleon count
  const auto fill_light = [&](int nu_, unsigned att, int xcen, int* X, int* Y, float* Z) -> int {
    if (s_nuc[nu_].kind == kProton) {
      X[0] = xcen; Y[0] = 0; if (Z) Z[0] = 0.f;
      return 1;
    }
    // deuteron: Hulthen separation along the nucleus' orientation
    // (Nucleus.cpp:193-197)
    const uint4 ro = draw(key, gev, kStrOrient, nu_, 0, att, 0);
    const float c = 1.f - 2.f * u01(ro.x), ph = 2.f * kPi * u01(ro.y);
    const uint4 r = draw(key, gev, kStrDeuteron, nu_, 0, att, 0);
    const float d = hulthen_inv(u01(r.x));
    float x = 0.5f * d, y = 0.f, z = 0.f;
    rotate3(c, ph, x, y, z);
    const int xq = quant(x), yq = quant(y);
    X[0] = xq + xcen; Y[0] = yq;
    X[1] = -xq + xcen; Y[1] = -yq;
    if (Z) { Z[0] = z * 0x1p24f; Z[1] = -z * 0x1p24f; }
    return 2;
  };

  for (;; ++attempt) {
    bool pair_drawn = false;  // this pass drew the next attempt's target too
    if (fed) {
      for (int i = t; i < A; i += 256) {
        px[0][i] = buf.raw[((size_t)e * 2 + 0) * kAMax * 2 + 2 * i];
        py[0][i] = buf.raw[((size_t)e * 2 + 0) * kAMax * 2 + 2 * i + 1];
      }
      for (int i = t; i < B; i += 256) {
        px[1][i] = buf.raw[((size_t)e * 2 + 1) * kAMax * 2 + 2 * i];
        py[1][i] = buf.raw[((size_t)e * 2 + 1) * kAMax * 2 + 2 * i + 1];
      }
      __syncthreads();
    } else if (pending) {
      // this attempt's target was drawn into slot 0 by the previous pass
      for (int i = t; i < B; i += 256) {
        px[1][i] = px[0][i];
        py[1][i] = py[0][i];
      }
      if (t == 0) {
        s_xc = s_xc2;
        s_b = s_b2;
        fill_light(0, attempt, s_xc2, lpx, lpy, nullptr);
      }
      __syncthreads();
      pending = false;
    } else {
      // b = sqrt((bmax^2-bmin^2) u + bmin^2)  (MakeDensity.cpp:2149); the
      // nuclei sit at +-b/2 rounded to the position quantum.
      if (t == 0) {
        const uint4 r = draw(key, gev, kStrB, 0, 0, attempt, 0);
        const double b = (mode == 2) ? buf.bval[e] : sqrt((cfg.bmax2 - cfg.bmin2) * u53(r.x, r.y) + cfg.bmin2);
        const int xc = (int)llrint(b * 0.5 / kQ);
        s_xc = xc;
        s_b = 2.0 * (double)xc * kQ;
        s_acc[0] = s_acc[1] = 0;
        if (spec) {
          const uint4 r2 = draw(key, gev, kStrB, 0, 0, attempt + 1, 0);
          const int xc2 = (int)llrint(sqrt((cfg.bmax2 - cfg.bmin2) * u53(r2.x, r2.y) + cfg.bmin2) * 0.5 / kQ);
          s_xc2 = xc2;
          s_b2 = 2.0 * (double)xc2 * kQ;
          fill_light(0, attempt, xc, lpx, lpy, nullptr);
        }
      }
      __syncthreads();
      // orientation of the nucleus drawn into each slot (Nucleus.cpp:193-197)
      float ctr[2], phr[2];
      for (int nu = 0; nu < 2; ++nu) {
        const uint4 r = draw(key, gev, kStrOrient, nu ? 1 : sys0, 0, (spec && nu == 0) ? attempt + 1 : attempt, 0);
        ctr[nu] = 1.f - 2.f * u01(r.x);
        phr[nu] = 2.f * kPi * u01(r.y);
      }
      // proton / deuteron fill directly
      if (wv < 2 && lane == 0 && s_nuc[wv ? 1 : sys0].kind != kWoodsSaxon)
        s_acc[wv] = fill_light(wv, attempt, (wv == 0) ? s_xc : -s_xc, px[wv], py[wv], pz[wv]);
      __syncthreads();
      // Woods-Saxon nuclei: sequential acceptance against all earlier
      // nucleons (hard core, Nucleus.cpp:276-294), 64 candidates per round
      // taken in stream order -- the same law as the reference's one-at-a-
      // time resampling.
      // Rounds of 64 candidates per nucleus.  Proposals come from all four
      // waves (waves nu and nu + 2 serve nucleus nu), one draw per lane per
      // step, accepted ones compacted in stream order (step, wave, lane)
      // into the round's candidate list; then the hard core against the
      // nucleons accepted so far (split over the two waves) and the
      // sequential acceptance among the round's candidates in list order.
      const int nu = wv & 1, sub = wv >> 1;
      const unsigned att_nu = (spec && nu == 0) ? attempt + 1 : attempt;
      bool spec_bad = false;
      for (unsigned round = 0;; ++round) {
        const bool need0 = s_nuc[sys0].kind == kWoodsSaxon && s_acc[0] < NS[0];
        const bool need1 = s_nuc[1].kind == kWoodsSaxon && s_acc[1] < B;
        if (!need0 && !need1) break;
        if (round > 4096) {
          // this attempt's nuclei unfinished: no event; only the next
          // attempt's target unfinished: that attempt draws it again
          if (!spec || need1) status |= kStNoEvent;
          else spec_bad = true;
          break;
        }
        const bool need = nu ? need1 : need0;
        const NucSys ns = s_nuc[nu ? 1 : sys0];
        int nc[2] = {0, 0};  // candidates so far (uniform over the workgroup)
        // candidates wanted this round: the nucleons still missing plus a
        // margin for the hard core, at most 64
        int want[2];
        for (int q = 0; q < 2; ++q) {
          const int miss = NS[q] - s_acc[q];
          want[q] = min(64, miss + (miss >> 3) + 4);
        }
        for (unsigned it = 0;; ++it) {
          bool take = false;
          int xq = 0, yq = 0;
          float z = 0.f;
          if (need && nc[nu] < want[nu]) {
            float x, y;
            take = ws_proposal(ns, key, gev, nu ? 1 : sys0, round * 128u + sub * 64u + lane, att_nu, it, ctr[nu],
                               phr[nu], x, y, z);
            xq = quant(x);
            yq = quant(y);
          }
          const unsigned long long m = __ballot(take);
          int* wc = s_wc[it & 1];  // double-buffered: one barrier per step
          if (lane == 0) wc[wv] = __popcll(m);
          __syncthreads();
          if (take) {
            const int pos = nc[nu] + (sub ? wc[nu] : 0) + __popcll(m & ((1ull << lane) - 1ull));
            if (pos < 64) {
              cqx[nu][pos] = xq;
              cqy[nu][pos] = yq;
              cqz[nu][pos] = z * 0x1p24f;
            }
          }
          nc[0] = min(64, nc[0] + wc[0] + wc[2]);
          nc[1] = min(64, nc[1] + wc[1] + wc[3]);
          const bool more0 = need0 && nc[0] < want[0], more1 = need1 && nc[1] < want[1];
          if (!(more0 || more1) || it >= 0xFFFEu) break;
        }
        if (t < 2) s_nc[t] = nc[t];
        __syncthreads();
        // Distances in position quanta: x, y differences exact in int32, z
        // held as z * 2^24, so d^2 < dmin^2 reads d2q < dmin^2 * 2^48.
        const float dmin2q = cfg.dmin2 * 0x1p48f;
        // hard core against the nucleons accepted in earlier rounds: wave
        // nu + 2*sub checks every candidate against half of them
        const int acc0 = s_acc[nu];
        const int ncand = s_nc[nu];  // a full list unless the step cap was hit
        bool ok = lane < ncand;
        const int xq = cqx[nu][lane], yq = cqy[nu][lane];
        const float zq = cqz[nu][lane];
        if (need && ok) {
          const int half = ((acc0 + 7) >> 3) << 2;  // multiple of 4
          const int j0 = sub ? half : 0, j1 = sub ? acc0 : min(half, acc0);
          const int4* px4 = reinterpret_cast<const int4*>(px[nu]);
          const int4* py4 = reinterpret_cast<const int4*>(py[nu]);
          const float4* pz4 = reinterpret_cast<const float4*>(pz[nu]);
          int j = j0;
          for (; j + 8 <= j1 && ok; j += 8) {
            const int4 ax0 = px4[j >> 2], ay0 = py4[j >> 2], ax1 = px4[(j >> 2) + 1], ay1 = py4[(j >> 2) + 1];
            const float4 az0 = pz4[j >> 2], az1 = pz4[(j >> 2) + 1];
            const int axs[8] = {ax0.x, ax0.y, ax0.z, ax0.w, ax1.x, ax1.y, ax1.z, ax1.w};
            const int ays[8] = {ay0.x, ay0.y, ay0.z, ay0.w, ay1.x, ay1.y, ay1.z, ay1.w};
            const float azs[8] = {az0.x, az0.y, az0.z, az0.w, az1.x, az1.y, az1.z, az1.w};
#pragma unroll
            for (int u = 0; u < 8; ++u) {
              const float fx = (float)(xq - axs[u]), fy = (float)(yq - ays[u]), fz = zq - azs[u];
              if (fmaf(fx, fx, fmaf(fy, fy, fz * fz)) < dmin2q) ok = false;
            }
          }
          for (; j < j1 && ok; ++j) {
            const float fx = (float)(xq - px[nu][j]), fy = (float)(yq - py[nu][j]), fz = zq - pz[nu][j];
            if (fmaf(fx, fx, fmaf(fy, fy, fz * fz)) < dmin2q) ok = false;
          }
        }
        // conflicts with the earlier candidates of this round: bit l of cm
        // for l < lane within the hard core; wave sub takes l in [32 sub, 32 sub + 32)
        unsigned long long cm = 0ull;
        if (need && lane < ncand) {
          const int4* cx4 = reinterpret_cast<const int4*>(cqx[nu]);
          const int4* cy4 = reinterpret_cast<const int4*>(cqy[nu]);
          const float4* cz4 = reinterpret_cast<const float4*>(cqz[nu]);
#pragma unroll 4
          for (int l4 = 8 * sub; l4 < 8 * sub + 8; ++l4) {
            if (4 * l4 >= lane) break;
            const int4 bx = cx4[l4], by = cy4[l4];
            const float4 bz = cz4[l4];
            const int bxs[4] = {bx.x, bx.y, bx.z, bx.w}, bys[4] = {by.x, by.y, by.z, by.w};
            const float bzs[4] = {bz.x, bz.y, bz.z, bz.w};
#pragma unroll
            for (int u = 0; u < 4; ++u) {
              const int l = 4 * l4 + u;
              const float fx = (float)(xq - bxs[u]), fy = (float)(yq - bys[u]), fz = zq - bzs[u];
              if (l < lane && fmaf(fx, fx, fmaf(fy, fy, fz * fz)) < dmin2q) cm |= 1ull << l;
            }
          }
        }
        const unsigned long long okw = __ballot(ok);
        if (lane == 0) s_okm[wv] = okw;
        s_cm[wv][lane] = cm;
        __syncthreads();
        if (need && sub == 0) {
          cm |= s_cm[wv + 2][lane];
          // sequential acceptance in list order: a candidate without conflicts
          // is accepted iff it passed the hard core; the conflicted ones are
          // decided in order against the accepted set
          const unsigned long long okm = s_okm[nu] & s_okm[nu + 2];
          const unsigned long long conf = __ballot(cm != 0ull);
          unsigned long long accm = okm & ~conf;
          unsigned long long rest = okm & conf;
          while (rest) {
            const int l = __ffsll((long long)rest) - 1;
            rest &= rest - 1;
            const unsigned lo32 = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)cm, l);
            const unsigned hi32 = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(cm >> 32), l);
            if (!((((unsigned long long)hi32 << 32) | lo32) & accm)) accm |= 1ull << l;
          }
          // keep the first An - acc0 (later candidates never affect earlier ones)
          const int An = NS[nu];
          int nok = __popcll(accm);
          while (nok > An - acc0) {
            accm &= ~(1ull << (63 - __clzll(accm)));
            --nok;
          }
          if ((accm >> lane) & 1ull) {
            const int slot = acc0 + __popcll(accm & ((1ull << lane) - 1ull));
            px[nu][slot] = xq;
            py[nu][slot] = yq;
            pz[nu][slot] = zq;
          }
          if (lane == 0) s_acc[nu] = acc0 + nok;
        }
        __syncthreads();
      }
      // CM recentering (Nucleus.cpp:296-309) in FP64, shift rounded to the
      // quantum, then the nucleus centre +-b/2.
      if (wv < 2 && s_nuc[wv ? 1 : sys0].kind == kWoodsSaxon) {
        const int An = NS[wv];
        double sxm = 0.0, sym = 0.0;
        for (int i = lane; i < An; i += 64) { sxm += (double)px[wv][i]; sym += (double)py[wv][i]; }
        sxm = wave_sum(sxm);
        sym = wave_sum(sym);
        const int cxq = (int)llrint(sxm / An), cyq = (int)llrint(sym / An);
        const int xcen = (wv == 0) ? (spec ? -s_xc2 : s_xc) : -s_xc;
        for (int i = lane; i < An; i += 64) { px[wv][i] = px[wv][i] - cxq + xcen; py[wv][i] -= cyq; }
      }
      __syncthreads();
      pair_drawn = spec && !spec_bad;
    }

    // ---- sort each nucleus by x, ties by input index (Nucleus.cpp:314):
    // bitonic sort of the keys (x, index)
    // (with two attempts per pass the projectile is in lpx / lpy)
    const int* X0 = spec ? lpx : px[0];
    const int* Y0 = spec ? lpy : py[0];
    for (int nu = 0; nu < 2; ++nu)
      skey[nu][t] = (t < NN[nu]) ? (((unsigned long long)((unsigned)(nu ? px[1] : X0)[t] ^ 0x80000000u) << 32) |
                                    (unsigned)t)
                                 : ~0ull;
    bitonic_sort_2x256(skey);
    for (int nu = 0; nu < 2; ++nu)
      if (t < NN[nu]) {
        const int i = (int)(skey[nu][t] & 0xFFFFFFFFull);
        sx[nu][t] = (nu ? px[1] : X0)[i];
        sy[nu][t] = (nu ? py[1] : Y0)[i];
      }
    if (t < B) fh[t] = 0x7FFFFFFF;
    __syncthreads();

    // ---- A x B disk test (MCnucl.cpp:359-366), exact in integers:
    // fl(fl(sqrt(d2))^2) <= dsq  <=>  dxq^2 + dyq^2 <= hit_n.
    {
      int cnt = 0;
      if (t < A) {
        unsigned bits[kAMax / 32];
#pragma unroll
        for (int w = 0; w < kAMax / 32; ++w) bits[w] = 0u;
        const int xa = sx[0][t], ya = sy[0][t];
        // criterion 2 tests every target whose 8w box overlaps the
        // projectile's (MCnucl.cpp:244-289), with one uniform each
        constexpr bool g2 = GAUSS;
        const int lim = g2 ? cfg.box_lim : cfg.hit_lim;
        int lo = 0, hi = B;
        while (lo < hi) {
          const int mid = (lo + hi) >> 1;
          if (sx[1][mid] < xa - lim) lo = mid + 1; else hi = mid;
        }
        // targets in x order, four per LDS load, until x exceeds xa + lim
        const int4* sx4 = reinterpret_cast<const int4*>(sx[1]);
        const int4* sy4 = reinterpret_cast<const int4*>(sy[1]);
        for (int j4 = lo >> 2; 4 * j4 < B; ++j4) {
          const int4 xb4 = sx4[j4], yb4 = sy4[j4];
          const int xbs[4] = {xb4.x, xb4.y, xb4.z, xb4.w}, ybs[4] = {yb4.x, yb4.y, yb4.z, yb4.w};
          uint4 rr = make_uint4(0u, 0u, 0u, 0u);
          if constexpr (g2) rr = draw(key, gev, kStrCollide, (unsigned)t, (unsigned)j4, (unsigned)attempt, 0u);
          const unsigned rs[4] = {rr.x, rr.y, rr.z, rr.w};
          bool past = false;
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const int j = 4 * j4 + u;
            if (j < lo || j >= B || past) continue;
            if (xbs[u] > xa + lim) {
              past = true;
              // the reference's loop tests one more target, the first past
              // the box, when it tested the one before (its condition reads
              // the previous target's box); only a draw can make it collide
              if (!g2 || j == lo) continue;
            }
            const int dx = xa - xbs[u], dy = ya - ybs[u];
            if (dy > lim || dy < -lim) continue;
            const long long n = (long long)dx * dx + (long long)dy * dy;
            bool h;
            if constexpr (g2) {  // u < 1 - exp(-sigma_gg T_pp(b)) (GaussianNucleonsCal.cpp:59-67)
              const float q = cfg.gauss_A * __builtin_amdgcn_exp2f(-(float)n * cfg.gauss_k);
              const float P = (q < 1e-3f) ? q * (1.f - 0.5f * q * (1.f - q * (1.f / 3.f))) : 1.f - __expf(-q);
              h = ((float)rs[u] + 0.5f) * 0x1p-32f < P;
            } else {
              h = n <= cfg.hit_n;
            }
            if (h) {
#pragma unroll
              for (int w = 0; w < kAMax / 32; ++w)
                if (w == (j >> 5)) bits[w] |= 1u << (j & 31);
              ++cnt;
              atomicMin(&fh[j], t);  // first hitting projectile row
            }
          }
          if (past) break;
        }
#pragma unroll
        for (int w = 0; w < kAMax / 32; ++w) hits[t][w] = bits[w];
        rowcnt[t] = cnt;
      } else {
        rowcnt[t] = 0;
      }
      __syncthreads();
      // first hitting projectile row of each target nucleon: its wound order
      const int first = (t < B && fh[t] != 0x7FFFFFFF) ? fh[t] : -1;
      fh[t] = first;
      npa = __syncthreads_count(t < A && cnt > 0);
      npb = __syncthreads_count(first >= 0);
      // total Ncoll
      int tot = wave_sum(cnt);
      if (lane == 0) scan_tmp[wv] = tot;
      __syncthreads();
      ncoll = scan_tmp[0] + scan_tmp[1] + scan_tmp[2] + scan_tmp[3];
      __syncthreads();
    }
    if (mode != 0) break;
    const int np = npa + npb;
    if (ncoll > 0 && np >= cfg.npmin && np <= cfg.npmax && !(status & kStNoEvent)) break;
    if (attempt + 1 >= cfg.max_attempts || (status & kStNoEvent)) { status |= kStNoEvent; break; }
    pending = pair_drawn;
    __syncthreads();
  }

  // ---- outputs: sorted positions, wound order, pairs, fluct, sources
  const size_t nbase = (size_t)e * 2 * kAMax;
  int woundA = -1, woundB = -1;
  {
    const int wA = (t < A && rowcnt[t] > 0) ? 1 : 0;
    const int rA = block_excl_scan(wA, scan_tmp);
    if (wA) woundA = rA;
    // target wound order: by first hitting row, then index (bitonic sort)
    skey[0][t] = ~0ull;
    skey[1][t] = (t < B && fh[t] >= 0) ? (((unsigned long long)fh[t] << 32) | (unsigned)t) : ~0ull;
    bitonic_sort_2x256(skey);
    if (skey[1][t] != ~0ull) wrank[skey[1][t] & 0xFFFFu] = t;
    __syncthreads();
    if (t < B && fh[t] >= 0) woundB = wrank[t];
  }
  const int off = block_excl_scan(t < A ? rowcnt[t] : 0, scan_tmp);
  if (cfg.uli) {
    // sub_model 2: the collisions of each target nucleon (its column of the
    // hit bitmap), into wrank, which the wound order no longer needs
    __syncthreads();
    int n = 0;
    if (t < B)
      for (int i = 0; i < A; ++i) n += (hits[i][t >> 5] >> (t & 31)) & 1u;
    wrank[t] = n;
    __syncthreads();
  }
  float fA = 1.f, fB = 1.f;
  if (cfg.fluct) {
    if (woundA >= 0) fA = gamma_draw(cfg.gk_part, cfg.gs_part, key, gev, kStrFluctA, (unsigned)t);
    if (woundB >= 0) fB = gamma_draw(cfg.gk_part, cfg.gs_part, key, gev, kStrFluctB, (unsigned)t);
  }
  // slots past this event's nucleus sizes read -2 (no nucleon; -1 is a
  // nucleon that was not hit): readers bound their scans by the context's
  // sizes, which a mixed batch's event may not fill
  static_assert(kAMax <= 256, "one slot per thread");
  if (t < A) {
    buf.pos[(nbase + t) * 2] = sx[0][t];
    buf.pos[(nbase + t) * 2 + 1] = sy[0][t];
    buf.wound[nbase + t] = woundA;
    buf.fluct[nbase + t] = fA;
  } else if (t < kAMax) {
    buf.wound[nbase + t] = -2;
  }
  if (t < B) {
    buf.pos[(nbase + kAMax + t) * 2] = sx[1][t];
    buf.pos[(nbase + kAMax + t) * 2 + 1] = sy[1][t];
    buf.wound[nbase + kAMax + t] = woundB;
    buf.fluct[nbase + kAMax + t] = fB;
  } else if (t < kAMax) {
    buf.wound[nbase + kAMax + t] = -2;
  }
  const int nsrc_part = npa + npb;
  int4* src = buf.src + (size_t)e * buf.src_cap;
  // pairs past max_pairs are not stored (status bit 2): the source list
  // holds the participants and the stored pairs only
  const int npairs_st = min(ncoll, cfg.max_pairs);
  const bool src_ok = nsrc_part + npairs_st <= buf.src_cap;
  int4* swin = buf.swin + (size_t)e * buf.src_cap;
  // bounding rectangle of every source window: all grid cells outside it
  // are zero for every density of this event
  int bb[4] = {0x7FFFFFFF, -1, 0x7FFFFFFF, -1};  // x0, x1, y0, y1
  auto grow = [&bb](int4 w, bool box) {
    const int v[4] = {w.z, w.w, box ? w.x : w.z, box ? w.y : w.w};
    for (int q = 0; q < 4; q += 2) {
      const int x0 = v[q] & 0xffff, x1 = v[q] >> 16, y0 = v[q + 1] & 0xffff, y1 = v[q + 1] >> 16;
      if (x0 < x1 && y0 < y1) {
        bb[0] = min(bb[0], x0);
        bb[1] = max(bb[1], x1);
        bb[2] = min(bb[2], y0);
        bb[3] = max(bb[3], y1);
      }
    }
  };
  if (t < 4) s_bb[t] = (t & 1) ? -1 : 0x7FFFFFFF;
  if (src_ok) {
    if (woundA >= 0) {
      src[woundA] = make_int4(2 * sx[0][t], 2 * sy[0][t], __float_as_int(fA), kSrcPartA);
      const int4 w = src_windows(cfg, 2 * sx[0][t], 2 * sy[0][t], true);
      swin[woundA] = w;
      grow(w, true);
    }
    if (woundB >= 0) {
      src[npa + woundB] = make_int4(2 * sx[1][t], 2 * sy[1][t], __float_as_int(fB), kSrcPartB);
      const int4 w = src_windows(cfg, 2 * sx[1][t], 2 * sy[1][t], true);
      swin[npa + woundB] = w;
      grow(w, true);
    }
  }
  if (ncoll > cfg.max_pairs || !src_ok) status |= kStPairOverflow;
  // pair list in reference order (projectile rows, then target index)
  const unsigned long long pbase = (unsigned long long)e * cfg.max_pairs;
  if (t < A && rowcnt[t] > 0) {
    int k = off;
    for (int w = 0; w < kAMax / 32; ++w) {
      unsigned m = hits[t][w];
      while (m) {
        const int j = w * 32 + __ffs(m) - 1;
        m &= m - 1;
        if (k < cfg.max_pairs) buf.pairs[pbase + k] = (unsigned)t | ((unsigned)j << 16);
        ++k;
      }
    }
  }
  __syncthreads();  // the pair list (global) is read back by every thread below
  // Gamma factors, sources and windows of the pairs, spread over the workgroup
  const int npairs = min(ncoll, cfg.max_pairs);
  for (int k = t; k < npairs; k += 256) {
    const unsigned pr = buf.pairs[pbase + k];
    const int i = (int)(pr & 0xFFFFu), j = (int)(pr >> 16);
    float fp = 1.f;
    if (cfg.fluct) fp = gamma_draw(cfg.gk_bin, cfg.gs_bin, key, gev, kStrFluctPair, (unsigned)k);
    buf.pfluct[pbase + k] = fp;
    float wsrc = fp;  // the pair source's weight
    if (cfg.uli) {
      // MCnucl.cpp:345-348: additional_weight += 1/n for both nucleons, in
      // int arithmetic (n >= 1 collisions), so aw = [n_a == 1] + [n_b == 1];
      // the density weighs the pair by Alpha + (1 - Alpha) aw (:750-755)
      const int aw = (rowcnt[i] == 1) + (wrank[j] == 1);
      buf.paw[pbase + k] = (unsigned char)aw;
      wsrc = fp * (cfg.alpha + (1.f - cfg.alpha) * (float)aw);
    }
    if (src_ok) {
      const int xh = sx[0][i] + sx[1][j], yh = sy[0][i] + sy[1][j];
      src[nsrc_part + k] = make_int4(xh, yh, __float_as_int(wsrc), kSrcPair);
      const int4 w = src_windows(cfg, xh, yh, false);
      swin[nsrc_part + k] = w;
      grow(w, false);
    }
  }
  if (!src_ok) {  // windows not stored: the whole grid
    bb[0] = bb[2] = 0;
    bb[1] = cfg.nx;
    bb[3] = cfg.ny;
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
    c[0] = npa;
    c[1] = npb;
    c[2] = ncoll;
    c[3] = status;
    c[4] = attempt + 1;
    c[5] = src_ok ? nsrc_part + npairs_st : 0;
    const bool none = s_bb[0] >= s_bb[1];
    c[6] = none ? 0 : (s_bb[0] | (s_bb[1] << 16));
    c[7] = none ? 0 : (s_bb[2] | (s_bb[3] << 16));
    if (!fed) buf.bval[e] = s_b;
  }
}

// ------------------------------------------------------------- k_smear
// Gather formulation: each workgroup owns a 16 (x) x 64 (y) tile, culls the
// event's sources against it (ordered, deterministic compaction) and each
// lane accumulates 4 cells of one column.  Cell values use the reference's
// footprints: window [lo,hi) from the threshold tables and the disk cut
// d^2 <= 25 w^2.
constexpr int kTX = 16, kTY = 64;

__global__ __launch_bounds__(256) void k_smear(DevCfg cfg, EvBufs buf, SmearSpec spec, float* out0,
                                               float* out1) {
  const int e = blockIdx.y;
  const int ntx = (cfg.nx + kTX - 1) / kTX;
  const int tx0 = (blockIdx.x % ntx) * kTX, ty0 = (blockIdx.x / ntx) * kTY;
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const int iy = ty0 + lane;
  const int rx0 = tx0 + wv * 4;

  __shared__ int4 lsrc[256];
  __shared__ int4 lwin[256];
  __shared__ float lw[256];
  __shared__ int wcount[4];

  int xgh[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) xgh[r] = (int)llrint((cfg.xmin + (rx0 + r) * cfg.dx) / kH);
  const int ygh = (int)llrint((cfg.ymin + iy * cfg.dy) / kH);

  float a0[4] = {0.f, 0.f, 0.f, 0.f}, a1[4] = {0.f, 0.f, 0.f, 0.f};
  const int K = buf.counts[(size_t)e * 8 + 5];
  const int4* src = buf.src + (size_t)e * buf.src_cap;

  for (int base = 0; base < K; base += 256) {
    const int k = base + t;
    bool take = false;
    int4 s = make_int4(0, 0, 0, 0), win = make_int4(0, 0, 0, 0);
    float wgt = 0.f;
    if (k < K) {
      s = src[k];
      const int kind = s.w;
      if (spec.enabled[kind]) {
        if (spec.win[kind] == kWinBox) {
          win.x = step_eval(cfg.tab[kBoxLoX], s.x);
          win.y = step_eval(cfg.tab[kBoxHiX], s.x);
          win.z = step_eval(cfg.tab[kBoxLoY], s.y);
          win.w = step_eval(cfg.tab[kBoxHiY], s.y);
        } else {
          win.x = step_eval(cfg.tab[kWinLoX], s.x);
          win.y = step_eval(cfg.tab[kWinHiX], s.x);
          win.z = step_eval(cfg.tab[kWinLoY], s.y);
          win.w = step_eval(cfg.tab[kWinHiY], s.y);
        }
        take = win.x < win.y && win.z < win.w && win.x < tx0 + kTX && win.y > tx0 && win.z < ty0 + kTY &&
               win.w > ty0;
        wgt = spec.scale[kind] * (spec.use_fluct[kind] ? __int_as_float(s.z) : 1.f);
        s.z = spec.acc[kind];
      }
    }
    // ordered compaction
    const unsigned long long m = __ballot(take);
    const int before = __popcll(m & ((1ull << lane) - 1ull));
    if (lane == 0) wcount[wv] = __popcll(m);
    __syncthreads();
    int woff = 0;
    for (int q = 0; q < wv; ++q) woff += wcount[q];
    const int n = wcount[0] + wcount[1] + wcount[2] + wcount[3];
    if (take) {
      lsrc[woff + before] = s;
      lwin[woff + before] = win;
      lw[woff + before] = wgt;
    }
    __syncthreads();
    for (int q = 0; q < n; ++q) {
      const int4 ss = lsrc[q];
      const int4 ww = lwin[q];
      if (rx0 >= ww.y || rx0 + 4 <= ww.x) continue;  // wave-uniform
      const float dy = (float)(ygh - ss.y) * kHf;
      const float dy2 = dy * dy;
      const float gy = (iy >= ww.z && iy < ww.w) ? lw[q] * __expf(-dy2 * cfg.inv2w2) : 0.f;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int ix = rx0 + r;
        if (ix >= ww.x && ix < ww.y) {
          const float dx = (float)(xgh[r] - ss.x) * kHf;
          const float dx2 = dx * dx;
          const float g = (dx2 + dy2 <= cfg.cut2) ? __expf(-dx2 * cfg.inv2w2) * gy : 0.f;
          if (ss.z == 0) a0[r] += g; else a1[r] += g;
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
    const size_t c = (size_t)e * N + (size_t)ix * cfg.ny + iy;
    if (spec.combine == 0) out0[c] = a0[r];
    else if (spec.combine == 1) out0[c] = sqrtf(a0[r] * a1[r]);
    else { out0[c] = a0[r]; out1[c] = a1[r]; }
  }
}

// ------------------------------------------------------------ launchers
static hipStream_t S(void* s) { return (hipStream_t)s; }

const void* event_kernel_symbol(const DevCfg& cfg) {
  if (cfg.spec_live)
    return cfg.crit_gauss ? reinterpret_cast<const void*>(&k_event<true, true>)
                          : reinterpret_cast<const void*>(&k_event<false, true>);
  return cfg.crit_gauss ? reinterpret_cast<const void*>(&k_event<true, false>)
                        : reinterpret_cast<const void*>(&k_event<false, false>);
}

int launch_event(const DevCfg& cfg, const EvBufs& b, int nev, unsigned long long seed,
                 unsigned long long event0, int mode, void* stream) {
  if (nev <= 0) return 0;
  const uint2 key = make_uint2((unsigned)seed, (unsigned)(seed >> 32));
  const void* fn = event_kernel_symbol(cfg);
  if (fn == reinterpret_cast<const void*>(&k_event<true, true>))
    hipLaunchKernelGGL((k_event<true, true>), dim3(nev), dim3(256), 0, S(stream), cfg, b, key, event0, mode);
  else if (fn == reinterpret_cast<const void*>(&k_event<false, true>))
    hipLaunchKernelGGL((k_event<false, true>), dim3(nev), dim3(256), 0, S(stream), cfg, b, key, event0, mode);
  else if (fn == reinterpret_cast<const void*>(&k_event<true, false>))
    hipLaunchKernelGGL((k_event<true, false>), dim3(nev), dim3(256), 0, S(stream), cfg, b, key, event0, mode);
  else
    hipLaunchKernelGGL((k_event<false, false>), dim3(nev), dim3(256), 0, S(stream), cfg, b, key, event0, mode);
  return (int)hipGetLastError();
}

int launch_smear(const DevCfg& cfg, const EvBufs& b, const SmearSpec& s, int nev, float* out0,
                 float* out1, void* stream) {
  if (nev <= 0) return 0;
  const int ntx = (cfg.nx + kTX - 1) / kTX, nty = (cfg.ny + kTY - 1) / kTY;
  hipLaunchKernelGGL(k_smear, dim3(ntx * nty, nev), dim3(256), 0, S(stream), cfg, b, s, out0, out1);
  return (int)hipGetLastError();
}

}  // namespace smc
