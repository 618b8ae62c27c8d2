/*
 * smc.h -- C-ABI of libsmc, the MI355X (gfx950) implementation of
 * superMC's per-event initial-condition hot path.
 *
 * The reference exposes this path only as C++ classes, called from one
 * thread inside superMC.e; there is no FFI.  Each entry point below replaces
 * a reference member function (cited beside it); the C++ facade in
 * supermc_amd/host/ re-exposes MCnucl / MakeDensity / KLNModel /
 * ParameterReader with the reference's signatures on top of these calls.
 *
 * Conventions
 *  - Plain pointers and sizes only.  Host arrays are caller-owned and only
 *    borrowed for the duration of a call.  Device buffers belong to the ctx.
 *  - Grids are [event][ix][iy] with iy contiguous (the reference's
 *    double[Maxx][Maxy], MCnucl.cpp:125-136); cell centre x = -maxx + ix*dx.
 *  - Nucleon positions are quantised to multiples of SMC_POS_QUANTUM_FM
 *    (2^-24 fm) on the way in; all index arithmetic on them is exact.
 *  - Every call returns an smc_status; the reference's fatal exits map to
 *    the codes below (it printed and called exit()).
 *  - A ctx owns one HIP stream and its device buffers; it is not
 *    thread-safe.  Use one ctx per host thread.
 */
#ifndef SMC_H
#define SMC_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define SMC_VERSION 1
#define SMC_POS_QUANTUM_FM (1.0 / 16777216.0) /* 2^-24 fm */
#define SMC_MAX_A 256                         /* nucleons per nucleus */

typedef int smc_status;
#define SMC_OK 0
#define SMC_ERR_ARG 1          /* bad argument / parameter value */
#define SMC_ERR_HIP 2          /* HIP runtime failure */
#define SMC_ERR_CAPACITY 3     /* batch or pair capacity exceeded */
#define SMC_ERR_NO_EVENT 4     /* no accepted event within max_attempts (MakeDensity.cpp:2154-2161) */
#define SMC_ERR_KLN_TABLE 5    /* thickness outside the dN/dy table (MCnucl.cpp:424-427, :665-669) */
#define SMC_ERR_DNDY_ZERO 6    /* dndy < 1e-15 (MCnucl.cpp:812-816) */
#define SMC_ERR_UNSUPPORTED 7  /* option outside the implemented path */
#define SMC_ERR_STATE 8        /* call made before the state it needs exists */
#define SMC_ERR_MISSING_KEY 9  /* parameter missing (ParameterReader.cpp:149-156 exit(-1)) */

/* Parameters.  Names and meanings are those of the reference's
 * parameters.dat keys (ParameterReader semantics); the "extension" block
 * adds keys the reference hard-codes, whose defaults reproduce it. */
typedef struct smc_params {
  /* model (parameters.dat:6-21) */
  int which_mc_model;          /* 1 MC-KLN, 5 MC-Glauber, 7 sqrt(TA TB) */
  int sub_model;               /* Glauber: 1 (classical), 2 ("Ulrich": pairs only,
                                  MCnucl.cpp:713-717); 7 (KLN) */
  double lambda;               /* KLN saturation exponent */
  int tmax, tmax_subdivision;  /* KLN table size (MCnucl.cpp:915-919) */
  double alpha;                /* two-component mixing */
  /* nuclei (parameters.dat:29-35) */
  int Aproj, Atarg, proj_deformed, targ_deformed;
  /* nucleons (parameters.dat:38-60) */
  /* collision_criterion (MCnucl.cpp:357-385): 1 disk b^2 <= dsq; 2 Gaussian
   * probability 1 - exp(-sigma_gg e^{-b^2/4w^2} / (4 pi w^2)) per pair whose
   * 8w boxes overlap (GaussianNucleonsCal.cpp:59-67); other values follow
   * shape_of_entropy (2 -> Gaussian, 4..9 -> disk).  3 and 4 (quark /
   * density collisions) are not on this path.  shape_of_entropy: 2 or 4..9
   * (Gaussian entropy, MCnucl.cpp:752-756); 1 (disk) and 3 (quarks) are not
   * on this path. */
  int shape_of_nucleons, collision_criterion, shape_of_entropy;  /* shape 2, 3 or 4 (Gaussian nucleons) */
  double quark_width, gauss_nucl_width;
  /* collision (parameters.dat:64-68) */
  double ecm, bmin, bmax;
  int Npmin, Npmax;
  /* fluctuations (parameters.dat:136-139) */
  int cc_fluctuation_model;    /* 0 none, 6 Gamma per source */
  double cc_fluctuation_Gamma_theta;
  /* grid (parameters.dat:128-133) */
  double maxx, maxy, dx, dy;
  /* output (parameters.dat:97-110) */
  double finalFactor;
  int ecc_from_order, ecc_to_order;
  /* extension keys: <= 0 (or NaN for betas) selects the reference value */
  double sigma_nn_mb;          /* else Regge96 at ecm (MCnucl.cpp:58-64) */
  double proj_ws_R, proj_ws_a, targ_ws_R, targ_ws_a;   /* Nucleus.cpp:65-111 */
  double proj_beta2, proj_beta4, targ_beta2, targ_beta4; /* Nucleus.cpp:122-148 */
  double nucleon_min_dist;     /* hard core, reference 0.9 fm (Nucleus.cpp:190); <0 = reference */
  /* KLN quadrature extension (replaces the BASES table integral) */
  int kln_n_pt, kln_n_kt, kln_n_phi;
  double kln_pt_min, kln_pt_max;
  int kln_direct;              /* 1: per-cell quadrature; 0: table + sixPoint2dInterp */
  /* capacities */
  int max_events;              /* events per batch held on the device */
  int max_pairs;               /* per-event collision-pair capacity (<=0: Aproj*Atarg) */
  int max_attempts;            /* per-event acceptance attempts */
  int smear_path;              /* every path applies the reference footprint (integer window
                                  and r <= 5w, Particle.cpp:122-132) and takes every model:
                                  0 separable VALU tiles [default]
                                  1 FP32 MFMA: Gx.diag(w).Gy^T over the square window, then
                                    the products outside the disk subtracted per cell
                                  2 per-cell exponentials
                                  3 as 1 with BF16x3 split-precision MFMA (bf16 hi + lo
                                    factors, three products, FP32 accumulation; each term
                                    within 3 * 2^-16 relative)
                                  4 as 1 with BF16x6 (hi + mid + lo, six products; each
                                    term within about 2^-22 relative) */
  double gaussian_lambda;      /* shape_of_nucleons 3: width and sigma_gg from lambda
                                  (GaussianNucleonsCal.cpp:41-46); parameter file 4.14 */
} smc_params;

/* Derived quantities (for the facade and tests) */
typedef struct smc_derived {
  double siginNN;      /* mb */
  double width;        /* nucleon Gaussian width [fm] (GaussianNucleonsCal.cpp:34-54) */
  double dsq;          /* 0.1 sigma / pi [fm^2] (MCnucl.cpp:120) */
  int nx, ny;          /* Maxx, Maxy */
  double xmin, ymin;
  int64_t hit_n_max;   /* largest dx^2+dy^2 (in quanta^2) that collides */
  double kln_dT;       /* table step (MCnucl.cpp:915-919) */
  int kln_tmax;
  double ws_R[2], ws_a[2], beta2[2], beta4[2], dmin;
  int criterion;       /* collision criterion in effect (1 disk, 2 Gaussian) */
  double sigma_gg;     /* effective cross section [fm^2] of the Gaussian test
                          (GaussianNucleonsCal.cpp:130-163) */
  int64_t box_lim;     /* half the tested box, 8w, in quanta (Particle.cpp:20) */
} smc_derived;

/* Per-event result record written by smc_moments / smc_run.
 * Index layout (doubles). */
#define SMC_EV_STRIDE 64
#define SMC_EV_EPS_RE 0    /* [10] -Re sum r^2 e^{in theta} rho / (sum r^2 rho + 1e-15) */
#define SMC_EV_EPS_IM 10   /* [10] */
#define SMC_EV_EPSP_RE 20  /* [10] r^n weight (r^3 for n=1) */
#define SMC_EV_EPSP_IM 30  /* [10] */
#define SMC_EV_RN 40       /* [10] <r^n>, n = 0..9 */
#define SMC_EV_XC 50
#define SMC_EV_YC 51
#define SMC_EV_TOTAL 52    /* sum of finalFactor * s over the grid */
#define SMC_EV_DSDY 53     /* total * dx * dy (MakeDensity.cpp:2442) */
#define SMC_EV_NPART 54
#define SMC_EV_NCOLL 55
#define SMC_EV_B 56
#define SMC_EV_STATUS 57
#define SMC_EV_SYSTEM 58   /* index into smc_set_systems' list (0 when the batch is unmixed) */

/* Fills p with the values of the reference's parameters.dat. */
smc_status smc_params_default(smc_params* p);
/* sizeof(smc_params) and sizeof(smc_derived), for bindings to check layout. */
size_t smc_params_sizeof(void);
size_t smc_derived_sizeof(void);

typedef struct smc_ctx smc_ctx;

/* MCnucl::MCnucl (MCnucl.cpp:22-155) + MakeDensity::MakeDensity set-up
 * (MakeDensity.cpp:21-84): validates parameters, derives sigma/width/dsq,
 * builds the exact footprint tables, allocates the batch on the device.
 * Grids of up to 2048 x 2048 cells (SMC_ERR_ARG above). */
smc_status smc_create(smc_ctx** ctx, const smc_params* p);
void smc_destroy(smc_ctx* ctx);
const char* smc_last_error(const smc_ctx* ctx);
smc_status smc_get_derived(const smc_ctx* ctx, smc_derived* out);

/* Event generation with acceptance, MakeDensity.cpp:2143-2162 +
 * MCnucl::generateNuclei (MCnucl.cpp:208-214) + Nucleus::populate
 * (Nucleus.cpp:187-317) + MCnucl::getBinaryCollision (MCnucl.cpp:217-308).
 * Counter-based: event i of the batch is event (event0 + i) of stream
 * `seed`, independent of batch boundaries. */
smc_status smc_sample(smc_ctx* ctx, uint64_t seed, uint64_t event0, int nev);

/* MCnucl::generateNuclei(b) + getBinaryCollision() (MCnucl.cpp:208-308)
 * for nev events at the given impact parameters [fm], one attempt each (the
 * caller applies the acceptance of MakeDensity.cpp:2147-2162).  b is rounded
 * to twice the position quantum; smc_moments reports the value used. */
smc_status smc_sample_b(smc_ctx* ctx, uint64_t seed, uint64_t event0, int nev, const double* b);

/* Fed positions (oracle-parity path): nev events of Aproj / Atarg nucleons,
 * xy interleaved in fm, in any order; b per event is recorded only.
 * Runs the collision step on them.  fluct_seed drives the Gamma factors. */
smc_status smc_set_positions(smc_ctx* ctx, int nev, const double* xy_proj,
                             const double* xy_targ, const double* b, uint64_t fluct_seed);

/* Npart per nucleus and Ncoll per event (MCnucl::getNpart1/2, getNcoll). */
smc_status smc_collide(smc_ctx* ctx, int* npart_proj, int* npart_targ, int* ncoll);

/* Nucleons of event ev in collision order (sorted by x, Nucleus.cpp:314):
 * xy [A][2] fm, wound order (participant rank, -1 if spectator) and fluct
 * factor.  Either nucleus pointer may be NULL. */
smc_status smc_get_nucleons(smc_ctx* ctx, int ev, double* xy_proj, int* wound_proj,
                            float* fluct_proj, double* xy_targ, int* wound_targ,
                            float* fluct_targ);

/* Collision pairs of event ev in reference order (MCnucl.cpp:326-352):
 * indices into the sorted nucleon lists and the pair fluct factor. */
smc_status smc_get_pairs(smc_ctx* ctx, int ev, int cap, int* pair_proj, int* pair_targ,
                         float* fluct, int* npairs);

/* The sources of event ev as MCnucl holds them (Nucleus::getParticipants
 * and binaryCollision, read by MCnucl::getHotSpots, MCnucl.cpp:1303-1323):
 * projectile participants in wound order, then target participants in
 * wound order, then the collision pairs in reference order, at their
 * current positions (after smc_recenter: moved and turned).  xy [cap][2] fm,
 * kind 0 projectile / 1 target participant / 2 pair, weight = fluct factor
 * (any pointer may be NULL); nsrc = the event's source count.  Returns
 * SMC_ERR_CAPACITY (after filling cap entries) when cap < nsrc. */
smc_status smc_get_sources(smc_ctx* ctx, int ev, int cap, double* xy, int* kind, float* weight, int* nsrc);

/* MCnucl::recenterGrid (turn = 0) and MCnucl::rotateGrid (turn = 1)
 * (MCnucl.cpp:1119-1174) for every event of the batch: the sources move by
 * minus the centre of mass of the current density and, with turn, rotate by
 * the order-n angle -atan2(-Im, -Re)/n of sum r^n e^{i n theta} rho
 * (GlueDensity::calcCMAngle, GlueDensity.cpp:87-144).  The next density
 * call smears the moved sources.  rot: host [nev][4] = {xc, yc, angle,
 * sum rho} (NULL allowed).  1 <= n <= 9. */
smc_status smc_recenter(smc_ctx* ctx, int n, int turn, double* rot);

/* MCnucl::calculateThickness (MCnucl.cpp:398-478).  TA/TB host [nev][nx][ny]
 * (NULL: keep on device only). */
smc_status smc_thickness(smc_ctx* ctx, float* TA, float* TB);

/* MCnucl::setDensity(iy=0, ipt=-1) (MCnucl.cpp:626-820) for the configured
 * model; s host [nev][nx][ny] (NULL: device only); dndy per event
 * (sum of s, no finalFactor; NULL allowed). */
smc_status smc_density(smc_ctx* ctx, float* s, double* dndy);

/* MCnucl::setDensity(iy=0, ipt=-1) in the reference's own precision: the
 * double grids of MCnucl.cpp:689-760 (which_mc_model 5) and :762-811 (7),
 * each (source, cell) term evaluated in FP64 with the reference's
 * expressions (Particle.cpp:122-132, 168-174) and summed in its order.
 * Events [ev0, ev0 + n) of the current batch; s host [n][nx][ny].
 * SMC_ERR_UNSUPPORTED for other models, SMC_ERR_ARG for a range outside
 * the batch. */
smc_status smc_density_f64(smc_ctx* ctx, int ev0, int n, double* s);

/* MakeDensity::setSd + dumpEccentricities (MakeDensity.cpp:2244-2522):
 * per-event records, SMC_EV_STRIDE doubles each (NULL: device only). */
smc_status smc_moments(smc_ctx* ctx, double* ev_out);

/* The whole hot path for nev events: sample -> collide -> density ->
 * moments, inputs and outputs resident on the device.  ev_out as above
 * (NULL: keep on device, fetch with smc_fetch_events). */
smc_status smc_run(smc_ctx* ctx, uint64_t seed, uint64_t event0, int nev, double* ev_out);
/* smc_run for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops).  No counterpart in the reference.
 * The same events and records as smc_run, written to ev_out_dev, a device
 * buffer of [nev][SMC_EV_STRIDE] doubles the caller owns (slot
 * SMC_EV_SYSTEM is 0).  The call allocates nothing: the moment kernels'
 * per-event scratch lives in a static device array of the library (about
 * 17 MB per GPU, held until the process ends), so nev is at most 2048
 * (SMC_ERR_CAPACITY above), and the context must already hold the buffers
 * of a batch (run smc_run on it once first; SMC_ERR_STATE otherwise).
 * MC-KLN only through the dN/dy table (smear_path 0, 1, 3 or 4).  It returns
 * without waiting: call smc_synchronize before reading ev_out_dev.  Every
 * call on one GPU shares that GPU's scratch: calls that overlap in time on
 * one GPU (contexts on different threads or streams) race. */
smc_status smc_run_noalloc(smc_ctx* ctx, uint64_t seed, uint64_t event0, int nev, double* ev_out_dev);

/* The generateEccTable event loop (MakeDensity.cpp:2143-2224) for nev_total
 * events of stream seed from event0, any number of them, in batches of
 * max_events through the ctx's own device buffers: device memory stays that
 * of one batch.  ev_out: host [nev_total][SMC_EV_STRIDE] (required); batch
 * i's records are copied out on a second stream while batch i + 1 runs.
 * Events equal smc_run's for the same indices.  The ctx afterwards holds
 * the last batch. */
smc_status smc_run_stream(smc_ctx* ctx, uint64_t seed, uint64_t event0, int64_t nev_total, double* ev_out);
/* smc_run replayed as a hipGraph (no counterpart in the reference: for
 * callers that launch many same-sized batches).  The first call for a batch
 * size runs eagerly, the second captures the launch sequence, later calls
 * set the event kernel's seed / event0 and launch the graph.  Results are
 * smc_run's; per-kernel timing covers the eager calls only. */
smc_status smc_run_graph(smc_ctx* ctx, uint64_t seed, uint64_t event0, int nev, double* ev_out);
smc_status smc_fetch_events(smc_ctx* ctx, int nev, double* ev_out);

/* Operation 2 grids (MakeDensity::generate_profile_ebe_Jet).
 * smc_rho_binary: MCnucl::calculate_rho_binary (MCnucl.cpp:481-527), the sum
 * over collision pairs of the normalised nucleon Gaussian (+-5w window,
 * d^2 <= 25 w^2).  smc_spectator_density: MCnucl::calculate_spectator_density
 * (MCnucl.cpp:530-614), the same Gaussians on the nucleons without a
 * collision, projectile and target apart.  Host grids [nev][nx][ny] (NULL:
 * device only). */
smc_status smc_rho_binary(smc_ctx* ctx, float* rho);
smc_status smc_spectator_density(smc_ctx* ctx, float* spec_proj, float* spec_targ);

/* Averaged, rotated profiles: the event loop body of
 * MakeDensity::generate_profile_average (operation 3, MakeDensity.cpp:736-
 * 1650, pT-integrated) for the ctx's events.  For each order n = from..to in
 * turn (the sources stay moved from one order to the next, as in the
 * reference): if rp_sums is non-NULL the sources are recentred on the
 * density's centre of mass (MCnucl::recenterGrid, MCnucl.cpp:1119-1145) and
 * finalFactor * rho is added to rp_sums[n-from]; then the sources are
 * recentred and turned by the order-n angle -atan2(-Im, -Re)/n of
 * sum r^n e^{i n theta} rho (GlueDensity::calcCMAngle, GlueDensity.cpp:
 * 87-144; MCnucl::rotateGrid, MCnucl.cpp:1147-1174) and finalFactor * rho is
 * added to sums[n-from].  sums, rp_sums: host [to-from+1][nx][ny], added
 * to.  accept: host [nev], nonzero = event included (NULL: status 0).  The
 * ctx's sources and density are left turned. */
smc_status smc_average_profiles(smc_ctx* ctx, int from_order, int to_order, const unsigned char* accept,
                                double* sums, double* rp_sums);

/* The optional averages of operation 3 (parameters.dat output_TATB,
 * output_rho_binary, output_TA, output_spectator_density;
 * MakeDensity.cpp:887-1200, 1290-1385), each host [to-from+1][nx][ny],
 * added to; NULL skips it.  After each order's turn (and, for the *_rp
 * ones, after the recentring of the reaction-plane pass, which needs
 * rp_sums): T_A T_B of the moved participants, the binary-collision density
 * of the moved pairs (MCnucl::calculate_rho_binary), T_A and T_B; after the
 * turn only, the spectator densities of the projectile and the target
 * (MCnucl::calculate_spectator_density), the spectators moved and turned
 * with the sources (MCnucl.cpp:1139-1143, 1168-1173).  No finalFactor. */
typedef struct smc_avg_extra {
  double* tatb;
  double* rho_binary;
  double* ta;
  double* tb;
  double* spec_a;
  double* spec_b;
  double* tatb_rp;
  double* rho_binary_rp;
  double* ta_rp;
  double* tb_rp;
} smc_avg_extra;
/* smc_average_profiles with the optional averages of extra (NULL: none). */
smc_status smc_average_profiles_extra(smc_ctx* ctx, int from_order, int to_order, const unsigned char* accept,
                                      double* sums, double* rp_sums, const smc_avg_extra* extra);

/* MC-KLN: KLNModel::getdNdy (KLNModel.cpp:97-125) at y=0 for n (TA,TB)
 * points, by the Gauss-Legendre product rule on the device. */
smc_status smc_kln_dndy(smc_ctx* ctx, int n, const double* ta, const double* tb, double* out);
/* MCnucl::makeTable (MCnucl.cpp:911-960) into the ctx (and to host table
 * [tmax][tmax] if non-NULL). */
smc_status smc_kln_make_table(smc_ctx* ctx, double* table);

/* Differentiable batches at frozen collision topology.  No counterpart in
 * the reference (it has no derivative code); for calibration callers.
 * theta = {w, R, a, alpha, k}: Gaussian nucleon width [fm], Woods-Saxon
 * radius and diffuseness of both nuclei [fm], two-component alpha, Gamma
 * k = 1/cc_fluctuation_Gamma_theta.  Requires which_mc_model 5, spherical
 * Woods-Saxon nuclei with equal R and a, cc_fluctuation_model 0 or 6.
 *
 * smc_diff_setup draws nev events at theta0 (the ctx's parameters, returned
 * in theta0[5] if non-NULL) with the acceptance of MakeDensity.cpp:2147-2162
 * and freezes their uniforms, impact parameters, participants, pairs,
 * footprint windows and hot-spot masks.  Positions are r = F^-1(u; R, a) of
 * the Woods-Saxon radial CDF, Gamma factors the quantiles of fixed uniforms.
 * smc_diff_forward evaluates the batch at theta; out [nev][8] (optional):
 * |eps_2|, |eps_3| (r^2 weight), dS/dy, Re eps_2, Im eps_2, Re eps_3,
 * Im eps_3, status.  smc_diff_backward takes cot [nev][3] = dL/d(|eps_2|,
 * |eps_3|, dS/dy) per event and returns grad[5] = dL/dtheta at the last
 * forward's theta.  smc_diff_topology exports event ev's frozen uniforms:
 * uvw [A][3] = (u, cos theta, phi) per nucleon in collision order, b, and
 * the Gamma uniform of each of its nsrc sources (participants in wound
 * order, then pairs). */
smc_status smc_diff_setup(smc_ctx* ctx, uint64_t seed, uint64_t event0, int nev, double* theta0);
smc_status smc_diff_forward(smc_ctx* ctx, const double* theta, double* out);
smc_status smc_diff_backward(smc_ctx* ctx, const double* cot, double* grad);
smc_status smc_diff_topology(smc_ctx* ctx, int ev, double* uvw_proj, double* uvw_targ, double* b,
                             double* gamma_v, int cap, int* nsrc);

/* Collision systems mixed in one batch (no counterpart in the reference,
 * which runs one system per job).  systems[i] is a full parameter set of
 * which only the nuclei (Aproj, Atarg and the WS / deformation keys) may
 * differ from the context's, with Aproj / Atarg at most the context's.
 * Sampled event g (smc_sample / smc_run) uses systems[pattern[g %
 * npattern]]; fed positions keep the context's system.  An event of system
 * s is the event a context made for s alone gives for the same seed and
 * index.  nsys = 0 restores the context's own system.  Records get the
 * system in slot SMC_EV_SYSTEM; smc_get_nucleons returns the event's
 * nucleons (entries past its counts: NaN, wound -2). */
smc_status smc_set_systems(smc_ctx* ctx, int nsys, const smc_params* systems, int npattern, const int* pattern);
smc_status smc_get_event_system(smc_ctx* ctx, int ev, int* system);  /* -1: unmixed batch */

/* Tile shapes of the exact-footprint smearing kernel (no counterpart in
 * the reference): 0 = 32 x 64 tiles, 8-row chunks (default); 1 = 16 x 64;
 * 2 = 32 x 64, 4-row chunks; 3 = 64 x 64.  Every shape gives the same grids.
 * smc_autotune times each on the ctx's current events (reps launches after
 * a warm-up), keeps the fastest, and returns it with the per-shape times in
 * ms[SMC_TILE_VARIANTS] (either pointer may be NULL). */
#define SMC_TILE_VARIANTS 4
smc_status smc_set_tile_variant(smc_ctx* ctx, int variant);
smc_status smc_autotune(smc_ctx* ctx, int reps, int* best, double* ms);

/* Sampling passes of the event kernel (no counterpart in the reference).
 * An event whose projectile is a proton or deuteron and whose target is a
 * Woods-Saxon nucleus redraws b until a collision happens
 * (MakeDensity.cpp:2143-2162), about six times for p+Pb at b < 20 fm.
 * With n = 2 (the default) its workgroup draws the targets of two
 * consecutive attempts per pass, each from that attempt's own counters, so
 * the events are bit-for-bit those of n = 1; n = 1 draws one attempt per
 * pass.  Other systems are not affected. */
smc_status smc_set_attempts_per_pass(smc_ctx* ctx, int n);
smc_status smc_get_attempts_per_pass(const smc_ctx* ctx, int* n);
/* Times the event kernel on the ctx's current sampled batch with n = 1 and
 * n = 2 (reps launches each after a warm-up), keeps the faster and returns
 * it, with the per-n times in ms[2] (either pointer may be NULL).  The
 * batch's events are unchanged. */
smc_status smc_autotune_attempts(smc_ctx* ctx, int reps, int* best, double* ms);

/* Plain FP32 FMA rate of the current device for this build (no counterpart
 * in the reference): independent v_fma_f32 chains over 4096 workgroups,
 * the fastest of three launches; *tflops = 2 x FMAs / time, *ms its time
 * (may be NULL).  The VALU kernels' rooflines are read against it. */
smc_status smc_calibrate_fma(double* tflops, double* ms);

/* Timing of the ctx's kernels with HIP events on its stream. */
#define SMC_K_EVENT 0
#define SMC_K_SMEAR 1
#define SMC_K_MOMENTS 2
#define SMC_K_KLN 3
#define SMC_K_ADJOINT 4
#define SMC_K_GRID64 5
#define SMC_K_COUNT 8
smc_status smc_set_timing(smc_ctx* ctx, int on);
smc_status smc_get_timing(smc_ctx* ctx, int kernel, double* total_ms, int* launches);
smc_status smc_synchronize(smc_ctx* ctx);

#ifdef __cplusplus
}
#endif
#endif
