/*
 * smc_oracle.h -- FP64 CPU restatement of superMC's per-event
 * initial-condition hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product path (supermc_amd/,
 * include/) may link, load or call this code; only tests/, the smoke() entry
 * and the cpu_baseline leg of bench.py use it, and only as the checker.
 *
 * Every function cites the reference file:line it restates.  Grids are
 * double[nx][ny], x-major with y contiguous (MCnucl.cpp:125-136).
 */
#ifndef SMC_ORACLE_H
#define SMC_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  double xmin, ymin, dx, dy;
  int nx, ny;
} orc_grid;

/* MCnucl.cpp:33-40: Xmin=-maxx, Maxx=(int)((Xmax-Xmin)/dx+0.1)+1 */
void orc_grid_init(orc_grid* g, double maxx, double maxy, double dx, double dy);

/* Regge96.cpp:27-50 and MCnucl.cpp:58-64: sigma_inel = sigma_tot - sigma_el [mb] */
double orc_sigma_inel(double ecm);
/* GaussianNucleonsCal.cpp:34-54 (shapes 1, 2, 4) */
double orc_nucleon_width(int shape_of_nucleons, double siginNN, double gauss_nucl_width);
/* MCnucl.cpp:120 */
double orc_dsq(double siginNN);

/* Particle.cpp:16-24 + Box2D.cpp:24-46: bounding box of a nucleon whose
 * quarks sit at its centre (box = +-4w square).  box = {xL, xR, yL, yR}. */
void orc_particle_box(double x, double y, double w, double box[4]);

/* Nucleus.h:104-105 sort key with the deterministic tie-break of this
 * build: (xLeft, input index).  Writes the permutation to perm. */
void orc_sort_nucleus(int n, const double* x, const double* y, double w, int* perm);

/* MCnucl.cpp:217-308 (sweep with box pruning, disk criterion 1),
 * MCnucl.cpp:326-352 (pair order), Nucleus.cpp:325-331 (wound order).
 * Inputs are nucleons in sorted order.  Outputs:
 *   wound_a/b[i]  = 1 if nucleon i is a participant
 *   order_a/b     = participant indices in wound order (length npart)
 *   pair_a/b[k]   = indices of the k-th collision pair (reference order)
 * Returns Ncoll, or -1 if more than max_pairs pairs. */
int orc_collide(int na, const double* xa, const double* ya,
                int nb, const double* xb, const double* yb,
                double w, double dsq,
                int* wound_a, int* wound_b, int* order_a, int* order_b,
                int* npart_a, int* npart_b,
                int* pair_a, int* pair_b, int max_pairs);

/* Brute-force A x B disk test (MCnucl.cpp:357-366 predicate only). */
int orc_collide_bruteforce(int na, const double* xa, const double* ya,
                           int nb, const double* xb, const double* yb,
                           double dsq, int* wound_a, int* wound_b);

/* MCnucl.cpp:432-478: T += sum over participants of Particle::getSmoothTn
 * (Particle.cpp:122-132) over the +-5w window. */
void orc_thickness(const orc_grid* g, double w, int n, const double* x,
                   const double* y, double* T);

/* MCnucl.cpp:822-866 addDensity (shape_of_entropy 2): participant Gaussians
 * times fluct factor over the bounding-box footprint, added into dens. */
void orc_add_density(const orc_grid* g, double w, int n, const double* x,
                     const double* y, const double* fluct, double* dens);

/* MCnucl.cpp:626-820, which_mc_model 5 / sub_model 1, shape_of_entropy 2.
 * Participants of A and B (any order) with fluct factors, pair midpoints
 * with fluct factors.  s is written (not accumulated).  Returns dndy = sum s. */
double orc_density_glauber(const orc_grid* g, double w, double alpha, int sub_model,
                           int npa, const double* xa, const double* ya, const double* fa,
                           int npb, const double* xb, const double* yb, const double* fb,
                           int ncoll, const double* xc, const double* yc, const double* fc,
                           const double* aw, double* s);

/* MCnucl.cpp:762-811, which_mc_model 7: s = sqrt(rho_A rho_B). */
double orc_density_trento(const orc_grid* g, double w,
                          int npa, const double* xa, const double* ya, const double* fa,
                          int npb, const double* xb, const double* yb, const double* fb,
                          double* s);

/* MakeDensity.cpp:2244-2511 + MCnucl.cpp:1303-1323.  density is the grid
 * after setSd (MakeDensity.cpp:2514-2522).  boxes: participant boxes
 * (4 doubles each) in hot-spot order; ncoll > 0 adds the degenerate
 * CollisionPair boxes at the origin (CollisionPair.h:12 + Box2D.cpp:14-22).
 * Outputs (arrays of 10, index = order): mom_re/mom_im (r^2 weight),
 * momp_re/momp_im (r^n weight, r^3 for n=1), rn (<r^n>, n=0..9).
 * cm[0..2] = xc, yc, total.  Returns 0, or -1 when the mask is empty. */
int orc_moments(const orc_grid* g, const double* density,
                int nbox, const double* boxes, int ncoll,
                int from_order, int to_order,
                double* mom_re, double* mom_im, double* momp_re,
                double* momp_im, double* rn, double* cm);

/* cells of the eps_n hot-spot mask as a 0/1 [nx][ny] grid */
int orc_hotspot_mask(const orc_grid* g, int nbox, const double* boxes, int ncoll, unsigned char* mask);

/* ---- MC-KLN (which_mc_model 1, sub_model 7) ---- */
typedef struct {
  double ecm;      /* sqrt(s) [GeV] */
  double lambda;   /* saturation exponent (setLambdaSaturation) */
  double ptmin, ptmax;
  double g2hfac;   /* KLNModel.cpp:65 x setNormalization (MakeDensity.cpp:105) */
  double lambda_qcd;
  double alpha_max;
} orc_kln;

void orc_kln_defaults(orc_kln* k, double ecm, double lambda);
/* KLNModel.cpp:219-277 integrand on the unit cube (BASES variables). */
double orc_kln_func(const orc_kln* k, double ta, double tb, const double* u);
/* getdNdy (KLNModel.cpp:97-110) with the BASES integral replaced by a
 * Gauss-Legendre product rule of n_pt x n_kt x n_phi nodes on the unit cube. */
double orc_kln_dndy_gl(const orc_kln* k, double ta, double tb,
                       int n_pt, int n_kt, int n_phi);
/* Gauss-Legendre nodes/weights on [0,1] (Golub-Welsch free Newton form). */
void orc_gauss_legendre01(int n, double* x, double* w);
/* arsenal.cpp:33-54 */
double orc_six_point(double x, double y, double v00, double v01, double v02,
                     double v10, double v11, double v20);
/* MCnucl.cpp:654-687: dndy(cell) from the table[tmax][tmax]. Returns dndy
 * sum, or a negative value when a cell leaves the table (reference exit(1)). */
double orc_density_kln_table(const orc_grid* g, const double* ta, const double* tb,
                             const double* table, int tmax, double dT, double* s);

/* ---- samplers (distribution-level restatements, own RNG) ---- */
typedef struct { uint64_t s; } orc_rng;
void orc_rng_seed(orc_rng* r, uint64_t seed);
double orc_rng_uniform(orc_rng* r);       /* (0,1) */
/* Nucleus.cpp:578-621 (spherical + deformed) incl. Point3D::rotate */
void orc_sample_ws(orc_rng* r, double R, double a, double beta2, double beta4,
                   double ctr, double phir, int deformed, double* x, double* y, double* z);
/* Nucleus.cpp:187-317: A nucleons with hard core dmin, CM subtracted, shifted */
void orc_populate(orc_rng* r, int A, double R, double a, double beta2, double beta4,
                  int deformed, double dmin, double xcenter, double* x, double* y);
/* HulthenFunc.cpp:27-36 */
double orc_hulthen_cdf(double r);
/* Marsaglia-Tsang (2000) gamma(shape, scale) */
double orc_gamma(orc_rng* r, double shape, double scale);
/* Point3D::rotate (MathBasics.cpp:41-50) */
void orc_rotate(double cth, double phi, double* x, double* y, double* z);
/* Nucleus.cpp:669-694 */
double orc_ylm(int l, double ct);

/* Whole CPU event loop (MakeDensity.cpp:2143-2224, Glauber model 5, ecc
 * only, no file output) for the CPU baseline.  Returns events generated. */
typedef struct {
  int A, B;
  double R_a, a_a, R_b, a_b, dmin;
  double siginNN, w, alpha, gamma_theta;
  int fluct;               /* cc_fluctuation_model == 6 */
  double bmin, bmax;
  int npmin, npmax;
  double maxx, dx;
  int from_order, to_order;
} orc_event_cfg;
/* eps2_out (may be NULL): nev x 5 per accepted event: Re eps_2, Im eps_2
 * (r^2 weight), b, Npart, Ncoll */
int orc_run_events(const orc_event_cfg* c, uint64_t seed, int nev, double* eps2_out);

/* ---- 3D profiles, scripts/generate_3d_profiles/profile_3d.cpp ---- */
/* profile_3d.cpp:24-27 and 133-155: y_beam, sigma_inelastic [fm^2] and the
 * n-point eta table (eta, projectile and target weights); maxes = {proj, targ}. */
void orc_p3d_setup(double ecm, int n, double* y_beam, double* sigma_inel_fm2, double* eta, double* proj,
                   double* targ, double* maxes);
/* profile_3d.cpp:402-434: index of the largest entry below v (-1 out of range
 * when skip_out_of_range, which is the only mode the sampler uses). */
long orc_p3d_search(const double* A, int n, double v);
/* profile_3d.cpp:302-344 (participants) = :346-400 (binary collisions):
 * rho[neta][nx][ny] += norm exp(-dis_eta - dis_x - dis_y) for n sources of 6
 * doubles {x, y, eta, sigma_x, sigma_y, sigma_eta}; rho is not cleared. */
void orc_p3d_smear(int nx, int ny, int neta, double dx, double dy, double deta, int n, const double* src,
                   double* rho);

#ifdef __cplusplus
}
#endif
#endif
