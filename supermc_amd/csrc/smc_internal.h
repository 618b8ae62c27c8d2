// smc_internal.h -- structures shared by the host side (smc_api.cpp) and the
// gfx950 kernels (smc_kernels.hip).  Not part of the C-ABI.
#pragma once
#include <stdint.h>

#include "../../include/smc.h"

namespace smc {

constexpr int kAMax = SMC_MAX_A;       // nucleons per nucleus held in LDS
constexpr double kQ = SMC_POS_QUANTUM_FM;  // position quantum (2^-24 fm)
constexpr double kH = kQ * 0.5;            // half quantum: pair midpoints
constexpr double kPosLimitFm = 60.0;       // |x|,|y| bound for int32 math

// Nucleus kinds (Nucleus.cpp:199-316)
enum NucKind : int { kWoodsSaxon = 0, kProton = 1, kDeuteron = 2 };

struct NucSys {
  int A;
  int kind;
  int deformed;
  float R, a, beta2, beta4;
  float rmax;   // R + 2.5 (Nucleus.cpp:118)
  float rwmax;  // 1/(1+exp(-R/a)) (Nucleus.cpp:119)
  float inv_a;  // 1/a
};

// Monotone step function f: int32 -> [vlo, vlo+n] given by thresholds:
// f(p) = vlo + #{k : thr[k] <= p}.  Guess: vlo-relative index ~ a*p + b.
struct StepTable {
  const int* thr;  // device pointer, n entries, non-decreasing
  int n;
  int vlo;
  float a, b;
};

// Footprint tables (exact reproduction of the reference's (int) casts).
enum TableId : int {
  kWinLoX = 0,  // (int)((x - 5w - Xmin)/dx)   clamped [0,nx]   p in half-quanta
  kWinHiX,      // (int)((x + 5w - Xmin)/dx)
  kWinLoY,
  kWinHiY,
  kBoxLoX,      // (int)((fl(x - hb) - Xmin)/dx)  hb = 4*max(w,qw)
  kBoxHiX,
  kBoxLoY,
  kBoxHiY,
  kMaskX,       // (int)(d/dx)   d = box edge - region edge >= 0, quanta
  kMaskY,       // (int)(d/dy)
  kOffX,        // (int)((rXL - Xmin)/dx)   rXL in quanta (may be < Xmin)
  kOffY,
  kNumTables
};

struct DevCfg {
  NucSys nuc[2];
  float dmin2;           // hard core squared [fm^2]
  double bmin2, bmax2;   // impact parameter range squared
  int npmin, npmax;
  int max_attempts;
  int max_pairs;
  long long hit_n;       // collide iff dxq^2+dyq^2 <= hit_n (quanta^2)
  int hit_lim;           // |dxq| > hit_lim cannot collide
  int crit_gauss;        // collision criterion 2: a draw per pair of overlapping 8w boxes
  int spec_attempts;     // 1: a proton / deuteron projectile's workgroup draws two attempts' targets per pass
  int spec_live;         // spec_attempts and some system of the context has such a projectile: the
                         // event kernel instance that carries the second slot
  int box_lim;           // 8w in quanta: boxes overlap iff |dxq|, |dyq| <= box_lim
  float gauss_k;         // log2(e) q^2 / (4 w^2): e^{-b^2/4w^2} = 2^(-n gauss_k), n in quanta^2
  float gauss_A;         // sigma_gg / (4 pi w^2)
  int fluct;             // Gamma factors on
  float gk_part, gs_part, gk_bin, gs_bin;  // MCnucl.cpp:1271-1301
  // grid
  int nx, ny;
  double xmin, ymin, dx, dy;
  float w, inv2w2, norm, cut2;  // width, 1/(2w^2), 1/(2 pi w^2), 25 w^2
  // flush-to-zero disk cut of the tile smearing (smc_tile.hip): row factor
  // scale K = 2^-126 / exp2(-cut2 * inv2w2 * log2 e), weight scale, and the
  // output scale 1 / (K tile_w)
  float tile_k, tile_w, tile_out;
  int tile_variant;      // k_tile shape, 0 .. kTileVariants - 1 (smc_tile.hip)
  // collision systems mixed in one batch (smc_set_systems): sampled event g
  // uses nuclei sys[2 s], sys[2 s + 1] with s = sys_of[g % sys_n]; sys_n = 0:
  // every event uses nuc[]
  const NucSys* sys;
  const int* sys_of;
  int sys_n;
  int hb_h;              // box half-width 4*max(w,qw) in half-quanta (rounded)
  float alpha, pref_part;
  int use_binary;        // alpha > 1e-8 (MCnucl.cpp:724)
  int uli;               // MC-Glauber sub_model 2 (MCnucl.cpp:713-717): pairs only, pair k weighted
                         // alpha + (1 - alpha) aw_k (MCnucl.cpp:345-348, 750-755)
  double final_factor;
  int from_order, to_order;
  StepTable tab[kNumTables];
};

// Per-event device buffers (structure of arrays).
struct EvBufs {
  int* pos;        // [ev][2][kAMax][2] sorted positions (quanta)
  float* fluct;    // [ev][2][kAMax]
  int* wound;      // [ev][2][kAMax] participant rank (wound order) or -1
  unsigned* pairs; // [ev][max_pairs] proj index | targ index << 16
  float* pfluct;   // [ev][max_pairs]
  unsigned char* paw;  // [ev][max_pairs] sub_model 2 only: aw = [n_a == 1] + [n_b == 1], n = a nucleon's
                       // collisions (the reference's integer 1/n, MCnucl.cpp:346-347); null otherwise
  int* counts;     // [ev][8]: npartA, npartB, ncoll, status, attempts, nsrc,
                   //   source rectangle x0 | x1 << 16, y0 | y1 << 16 (cells)
  double* bval;    // [ev]
  int* raw;        // [ev][2][kAMax][2] fed positions (quanta), input order
  int4* src;       // [ev][src_cap] {x_h, y_h, fluct bits, kind}
  int4* swin;      // [ev][src_cap] footprints {box x, box y, 5w x, 5w y}, lo | hi << 16
  int src_cap;
};

// Event status bits
constexpr int kStOk = 0;
constexpr int kStNoEvent = 1;
constexpr int kStPairOverflow = 2;
constexpr int kStKlnTable = 8;

// Differentiable batches (frozen topology): device buffers of a context
struct DiffBufs {
  int* pending;       // [ev] topology still to be drawn
  int* attempt;       // [ev] attempt counter of the topology draw
  int* topo_ok;       // [ev] both nuclei filled
  double* bval;       // [ev] impact parameter used
  double* du;         // [ev][2][kAMax][4] (u, cos theta, phi, 0), collision order
  double* dpos;       // [ev][2][kAMax][6] x, y, x_R, y_R, x_a, y_a at theta
  double2* dsrc;      // [ev][src_cap] Gamma factor f and df/ds
  double* gam_v;      // [ev][src_cap] Gamma quantile uniforms
  double* dsc;        // [ev][kDsStride] per-event moment scalars
  unsigned* mask;     // [ev][nx][(ny+31)/32] hot-spot mask at theta0
  double4* gsrc;      // [ev][src_cap] dL/dW, dL/dx, dL/dy, dL/dw per source
  double* gev;        // [ev][5] dL/dtheta per event
  const double2* gl;  // Gauss-Legendre nodes/weights on [0, 1]
  int ngl;
  double* wstab;      // [kWsTabLen] radial quantiles of the current (R, a) (Newton starts), then Z, Z_R, Z_a
  double* cotc;       // [ev][kCotStride] coefficients of the cell cotangents (k_diff_cot_coef)
};
constexpr int kCotStride = 12;  // c0, then (lam cx/S, lam cy/S, lam Ar/(Am N), lam Ai/(Am N), lam Am/N^2) for n = 2, 3
constexpr int kWsTab = 2049;              // radial quantile nodes, uniform in u^(1/3)
constexpr int kWsTabLen = kWsTab + 3;     // + the normalisation and its R, a derivatives
// per-event scalars of the differentiable moments
enum : int {
  kDsS = 0, kDsXc, kDsYc, kDsA2r, kDsA2i, kDsA3r, kDsA3i, kDsN, kDsMx, kDsMy,
  kDsB3xr, kDsB3xi, kDsB3yr, kDsB3yi, kDsStride = 16
};
constexpr int kDiffOut = 8;  // per-event forward outputs: |e2|, |e3|, dSdy, e2 re/im, e3 re/im, status

// running-coupling constants of KLNModel::getAlphaStrong (KLNModel.h:91-96)
struct KlnConst {
  double lq2, amax, inv_beta0;
};

// Smear kernel source kinds and accumulation modes
enum SrcKind : int { kSrcPartA = 0, kSrcPartB = 1, kSrcPair = 2 };
enum WinType : int { kWinBox = 0, kWin5 = 1 };

struct SmearSpec {
  int enabled[3];   // per kind
  int win[3];       // WinType per kind
  float scale[3];   // weight multiplier per kind
  int use_fluct[3];
  int acc[3];       // accumulator (0/1) per kind
  int combine;      // 0: out0 = acc0; 1: out0 = sqrt(acc0*acc1); 2: out0=acc0,out1=acc1
  double* cm_out;   // combine 0, tile path: per-tile {sum v, sum x v, sum y v, 0} of out0, or null
  int sparse;       // 1: tiles outside the event's source rectangle are not stored (every
                    // cell there is zero; readers that need whole grids call
                    // launch_zero_outside first); 0: whole grids are written
};

// FP64 density grids (smc_grid64.hip): the reference's double expressions,
// evaluated per (source, cell); constants computed on the host as the
// reference writes them
struct Grid64Const {
  double norm;       // 1 / (2 pi w w)
  double two_w_w;    // 2 w w   (Particle.cpp:122-132)
  double two_wsq;    // 2 (w w) (MCnucl.cpp:732-756)
  double r_max;      // 5 w
  double dc_max;     // 25 (w w)
  double alpha, prefactor;  // alpha, (1 - alpha) / 2
  int model;         // which_mc_model: 5 or 7
  int use_binary;    // alpha > 1e-8
  int fluct;         // Gamma factors on
  int uli;           // sub_model 2: no participant term, pair weights alpha + (1 - alpha) aw
};
constexpr int kGrid64Chunk = 2048;  // most events per k_grid64 launch (device buffer 8 B x cells each)
int launch_grid64(const DevCfg& cfg, const EvBufs& b, const Grid64Const& gc, int ev0, int nev, double* out,
                  void* stream);

// plain FP32 FMA rate (smc_probe.hip): blocks x 256 lanes x 128 x iters FMAs
int launch_fma_chain(float* out, int blocks, int iters, void* stream);

// Launch wrappers (smc_kernels.hip)
// k_event's host handle (matches the func of its captured graph node); its
// kernel parameters are (cfg, buf, key, event0, mode)
const void* event_kernel_symbol(const DevCfg& cfg);  // the k_event instance cfg launches
constexpr int kEventKernelArgs = 5, kEventArgKey = 2, kEventArgEvent0 = 3;
int launch_event(const DevCfg& cfg, const EvBufs& b, int nev, unsigned long long seed,
                 unsigned long long event0, int mode, void* stream);
int launch_smear(const DevCfg& cfg, const EvBufs& b, const SmearSpec& s, int nev,
                 float* out0, float* out1, void* stream);
// mode 0: eps_n record; centre-of-mass partials from cm_units (nu per
// event) or, if null, computed into cm_scratch (moments_cm_units_min() per
// event); band holds moments_band_doubles() per event, rects 2 kAMax
// 8-byte mask rectangles and nrect one int per event.  mode 1: grid sums.
int launch_moments(const DevCfg& cfg, const EvBufs& b, const float* grid, int nev, double* ev_out, int mode,
                   const double* cm_units, int nu, double* cm_scratch, double* band, void* rects, int* nrect,
                   void* stream);
int moments_band_doubles();
int moments_cm_units_min();
// centre-of-mass partials of grids from another producer: moments_cm_units_min() units per event
int launch_cm_partials(const DevCfg& cfg, const EvBufs& b, const float* grid, int nev, double* cm, void* stream);
// averaged rotated profiles (smc_avg.hip)
int launch_cm_angle(const DevCfg& cfg, const EvBufs& b, const float* grid, const double* cm, int nu, int n,
                    int angle_on, int nev, double* rot, void* stream);
int launch_rotate_src(const DevCfg& cfg, const EvBufs& b, const double* rot, int nev, void* stream);
// spectators of every event as a source list of their own (sb.counts / src / swin)
int launch_spectator_src(const DevCfg& cfg, const EvBufs& b, const EvBufs& sb, int nev, void* stream);
// part: kAccumSegments x nx x ny doubles of scratch (segment partials)
constexpr int kAccumSegments = 32;
// with grid2, the sum of the cells' products grid * grid2
int launch_accum_avg(const float* grid, const int* counts, int nx, int ny, int nev, const unsigned char* accept,
                     double ff, double* part, double* sum, void* stream, const float* grid2 = nullptr);
constexpr int kTileVariants = 4;
int tile_units(const DevCfg& cfg);  // smearing tiles per event of cfg's variant (units of SmearSpec::cm_out)
int tile_units_max(const DevCfg& cfg);  // the most over all variants (buffer sizing)
// exact-footprint smearing as an MFMA contraction (smc_mtile.hip): prec 0
// FP32, 1 BF16x3, 2 BF16x6; ktab (tmax x tmax) selects the fused MC-KLN
// epilogue.  Its centre-of-mass partials have mtile_units(cfg) units (the
// 32 x 64 tiles of k_tile's default shape).
int mtile_units(const DevCfg& cfg);
int launch_mtile(const DevCfg& cfg, const EvBufs& b, const SmearSpec& s, int nev, float* out0, float* out1, int prec,
                 const double* ktab, int ktmax, double kdT, void* stream);
int launch_tile(const DevCfg& cfg, const EvBufs& b, const SmearSpec& s, int nev, float* out0, float* out1,
                void* stream);
// smc_run_noalloc's device scratch (smc_noalloc.hip): per event the moment
// band sums (moments_band_doubles() doubles), the hot-spot rectangles
// (2 kAMax 8-byte entries), centre-of-mass units (moments_cm_units_min() x 4
// doubles) and the rectangle count, for at most kNoAllocMaxEvents events
constexpr int kNoAllocMaxEvents = 2048;
constexpr size_t kNoAllocBandBytes = (size_t)kNoAllocMaxEvents * 8 * 64 * sizeof(double);
constexpr size_t kNoAllocRectBytes = (size_t)kNoAllocMaxEvents * 2 * kAMax * 8;
constexpr size_t kNoAllocCmBytes = (size_t)kNoAllocMaxEvents * 8 * 4 * sizeof(double);
constexpr size_t kNoAllocNrectBytes = (size_t)kNoAllocMaxEvents * sizeof(int);
constexpr size_t kNoAllocScratchBytes = kNoAllocBandBytes + kNoAllocRectBytes + kNoAllocCmBytes + kNoAllocNrectBytes;
// device address of the current device's scratch (looked up once per
// device); nullptr with the HIP error in *rc on failure
void* noalloc_scratch(int* rc);
// zeroes every cell outside each event's source rectangle (counts[6..7]):
// completes grids written with SmearSpec::sparse
int launch_zero_outside(float* grid, const int* counts, int nx, int ny, int nev, void* stream);
// 3D profiles (smc_p3d.hip), scripts/generate_3d_profiles/profile_3d.cpp
struct P3dCfg {
  int nx, ny, neta;
  double dx, dy, deta;
  int random_flag;
  double y_beam, eta0, sigma_x0, sigma_eta0, dsigma_x, dsigma_eta;
  const double* tab_eta;  // tab_n entries each (device)
  const double* tab_proj;
  const double* tab_targ;
  int tab_n;
  double proj_max, targ_max;
};
int launch_p3d_sample(const P3dCfg& c, unsigned long long seed, unsigned long long event0, int nev,
                      const double* part_xy, const int* part_id, const int* part_off, const double* bin_xy,
                      const int* bin_off, double* part_src, double* bin_src, int* status, void* stream);
int launch_p3d_smear(const P3dCfg& c, int nev, const double* src, const int* src_off, float* rho, void* stream);
// Regge96.cpp:27-50 sigma_tot - sigma_el [mb] (smc_api.cpp)
double regge_sigma_inel(double ecm);

// MC-KLN density in one pass: T_A (acc 0) and T_B (acc 1) of s, then dN/dy
// from the tmax x tmax table at (T_A/dT, T_B/dT), in the ctx's tile shape;
// s.cm_out gets the output's centre-of-mass partials (tile_units per event)
// when set
int launch_tile_kln(const DevCfg& cfg, const EvBufs& b, const SmearSpec& s, int nev, float* out,
                    const double* table, int tmax, double dT, void* stream);
void set_tile_scales(DevCfg& g);
int launch_diff_topo(const DevCfg& cfg, const EvBufs& b, const DiffBufs& d, int nev, unsigned long long seed,
                     unsigned long long event0, double R0, double a0, void* stream);
int launch_diff_forward_pre(const DevCfg& cfg, const EvBufs& b, const DiffBufs& d, int nev, unsigned long long seed,
                            unsigned long long event0, const double* theta, int fluct, void* stream);
int launch_diff_moments(const DevCfg& cfg, const EvBufs& b, const DiffBufs& d, const float* grid, int nev,
                        double ffdxdy, double* out, void* stream);
int launch_diff_backward(const DevCfg& cfg, const EvBufs& b, const SmearSpec& spec, const DiffBufs& d, int nev,
                         const double* cot_dev, double ffdxdy, float* G, const double* theta, void* stream);
int launch_kln_direct(const float* ta, const float* tb, float* out, size_t ncell, const void* ptf,
                      const void* nodes, int npt, int nn, const KlnConst& kc, unsigned* list, unsigned* count,
                      void* stream);  // list: ncell entries, count: one (device scratch of the caller)
int launch_kln_points(const double* ta, const double* tb, double* out, int n, const void* ptd,
                      const void* nodes, int npt, int nn, const KlnConst& kc, void* stream);
int launch_kln_table(const float* ta, const float* tb, float* out, int ncell_ev, int nev, int ny,
                     const double* table, int tmax, double dT, int* counts, void* stream);

}  // namespace smc
