/*
 * smc_oracle.c -- FP64 CPU restatement of superMC's per-event
 * initial-condition hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may use it.
 *
 * Each function restates the reference lines cited beside it, keeping the
 * reference's floating-point expression order where parity depends on it
 * (footprint indices, box arithmetic, the collision predicate).
 */
#include "smc_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

static int imax_(int a, int b) { return a > b ? a : b; }
static int imin_(int a, int b) { return a < b ? a : b; }
static double dmin_(double a, double b) { return (b < a) ? b : a; }  /* std::min */
static double dmax_(double a, double b) { return (a < b) ? b : a; }  /* std::max */

void orc_grid_init(orc_grid* g, double maxx, double maxy, double dx, double dy)
{
  /* MCnucl.cpp:33-40 */
  double Xmax = maxx, Ymax = maxy;
  double Xmin = -Xmax, Ymin = -Ymax;
  g->xmin = Xmin;
  g->ymin = Ymin;
  g->dx = dx;
  g->dy = dy;
  g->nx = (int)((Xmax - Xmin) / dx + 0.1) + 1;
  g->ny = (int)((Ymax - Ymin) / dy + 0.1) + 1;
}

double orc_sigma_inel(double ecm)
{
  /* Regge96.cpp:27-50 with i=0 (pp) and i1=i2=0 (baryons); MCnucl.cpp:58-64 */
  double s = ecm * ecm;
  double sig = 22.0 * pow(s, 0.079) + 56.1 * pow(s, -0.46);
  double bel = 2.0 * 2.3 + 2.0 * 2.3 + 4.0 * pow(s, 0.0808) - 4.2;
  double sigel = 0.0511 * sig * sig / bel;
  return sig - sigel;
}

double orc_nucleon_width(int shape, double siginNN, double gauss_nucl_width)
{
  /* GaussianNucleonsCal.cpp:34-54 */
  if (shape == 1) return sqrt(0.1 * siginNN / (M_PI)) / 2.0;
  if (shape == 2) return sqrt(0.1 * siginNN / M_PI) / sqrt(8);
  if (shape == 4) return gauss_nucl_width;
  return -1.0;
}

double orc_dsq(double siginNN)
{
  /* MCnucl.cpp:120 with overSample = 1 */
  return 0.1 * siginNN / M_PI / 1;
}

void orc_particle_box(double x, double y, double w, double box[4])
{
  /* Particle.cpp:16-21: baseBox.setCenter(x,y) on the zero box
   * (Box2D.cpp:24-33), then setSquareDimensions(8w) (Box2D.cpp:35-46). */
  double size = 8 * w;
  double xL = x - size / 2, xR = x + size / 2;
  double yL = y - size / 2, yR = y + size / 2;
  /* calculateBounds (Particle.cpp:94-99) with every quark at the nucleon
   * centre and quark_width <= w leaves the base box edges unchanged. */
  box[0] = xL;
  box[1] = xR;
  box[2] = yL;
  box[3] = yR;
}

typedef struct { double key; int idx; } sort_item;

static int cmp_item(const void* a, const void* b)
{
  const sort_item* p = (const sort_item*)a;
  const sort_item* q = (const sort_item*)b;
  if (p->key < q->key) return -1;
  if (p->key > q->key) return 1;
  return (p->idx < q->idx) ? -1 : (p->idx > q->idx);
}

void orc_sort_nucleus(int n, const double* x, const double* y, double w, int* perm)
{
  /* Nucleus.cpp:314 + Nucleus.h:104-105: sort by bounding-box xLeft.  The
   * reference uses std::sort (unspecified order for equal keys); ties are
   * broken here by input index. */
  sort_item* it = (sort_item*)malloc(sizeof(sort_item) * (n > 0 ? n : 1));
  for (int i = 0; i < n; i++) {
    double box[4];
    orc_particle_box(x[i], y[i], w, box);
    it[i].key = box[0];
    it[i].idx = i;
  }
  qsort(it, n, sizeof(sort_item), cmp_item);
  for (int i = 0; i < n; i++) perm[i] = it[i].idx;
  free(it);
  (void)y;
}

static int hit_disk(double x1, double y1, double x2, double y2, double dsq)
{
  /* MCnucl.cpp:359-366 (criterion 1) */
  double b = sqrt((x2 - x1) * (x2 - x1) + (y2 - y1) * (y2 - y1));
  return (b * b <= dsq) ? 1 : 0;
}

int orc_collide(int na, const double* xa, const double* ya,
                int nb, const double* xb, const double* yb,
                double w, double dsq,
                int* wound_a, int* wound_b, int* order_a, int* order_b,
                int* npart_a, int* npart_b,
                int* pair_a, int* pair_b, int max_pairs)
{
  /* who_hit_me lists (Particle.h:24), stored densely */
  int* cnt_a = (int*)calloc(na > 0 ? na : 1, sizeof(int));
  int* cnt_b = (int*)calloc(nb > 0 ? nb : 1, sizeof(int));
  int* hits_a = (int*)malloc(sizeof(int) * (size_t)(na > 0 ? na : 1) * (nb > 0 ? nb : 1));
  int na_part = 0, nb_part = 0;
  double (*bxa)[4] = malloc(sizeof(double[4]) * (na > 0 ? na : 1));
  double (*bxb)[4] = malloc(sizeof(double[4]) * (nb > 0 ? nb : 1));
  for (int i = 0; i < na; i++) orc_particle_box(xa[i], ya[i], w, bxa[i]);
  for (int j = 0; j < nb; j++) orc_particle_box(xb[j], yb[j], w, bxb[j]);

  /* MCnucl.cpp:247-292: x-sorted sweep with box pruning */
  int startingIndex = 0;
  for (int iproj = 0; iproj < na; iproj++) {
    const double* projBox = bxa[iproj];
    double targBox[4] = {0, 0, 0, 0};  /* Box2D() is all zeros */
    while (startingIndex < nb) {
      memcpy(targBox, bxb[startingIndex], sizeof(targBox));
      if (targBox[1] >= projBox[0]) break;
      startingIndex++;
    }
    int i = startingIndex;
    while (i < nb && projBox[1] >= targBox[0]) {
      memcpy(targBox, bxb[i], sizeof(targBox));
      if (projBox[2] <= targBox[3]) {
        if (projBox[3] >= targBox[2]) {
          if (hit_disk(xa[iproj], ya[iproj], xb[i], yb[i], dsq)) {
            hits_a[(size_t)iproj * nb + cnt_a[iproj]] = i;
            cnt_a[iproj]++;
            cnt_b[i]++;
            /* Nucleus.cpp:325-331: wounded on the first hit */
            if (cnt_a[iproj] == 1) order_a[na_part++] = iproj;
            if (cnt_b[i] == 1) order_b[nb_part++] = i;
          }
        }
      }
      i++;
    }
  }

  /* MCnucl.cpp:326-352: proj participants in wound order x who_hit_me */
  int ncoll = 0;
  for (int k = 0; k < na_part; k++) {
    int ip = order_a[k];
    for (int m = 0; m < cnt_a[ip]; m++) {
      if (ncoll < max_pairs) {
        pair_a[ncoll] = ip;
        pair_b[ncoll] = hits_a[(size_t)ip * nb + m];
      }
      ncoll++;
    }
  }
  for (int i = 0; i < na; i++) wound_a[i] = cnt_a[i] > 0;
  for (int j = 0; j < nb; j++) wound_b[j] = cnt_b[j] > 0;
  *npart_a = na_part;
  *npart_b = nb_part;
  free(cnt_a); free(cnt_b); free(hits_a); free(bxa); free(bxb);
  return ncoll > max_pairs ? -1 : ncoll;
}

int orc_collide_bruteforce(int na, const double* xa, const double* ya,
                           int nb, const double* xb, const double* yb,
                           double dsq, int* wound_a, int* wound_b)
{
  int ncoll = 0;
  for (int i = 0; i < na; i++) wound_a[i] = 0;
  for (int j = 0; j < nb; j++) wound_b[j] = 0;
  for (int i = 0; i < na; i++)
    for (int j = 0; j < nb; j++)
      if (hit_disk(xa[i], ya[i], xb[j], yb[j], dsq)) {
        ncoll++;
        wound_a[i] = 1;
        wound_b[j] = 1;
      }
  return ncoll;
}

static double smooth_tn(double x, double y, double xg, double yg, double width)
{
  /* Particle.cpp:122-132 (no gluon field) */
  double r = sqrt((xg - x) * (xg - x) + (yg - y) * (yg - y));
  if (r > 5 * width) return 0;
  double Tn = (1 / (2 * M_PI * width * width)) * exp(-r * r / (2 * width * width));
  return Tn;
}

void orc_thickness(const orc_grid* g, double w, int n, const double* x,
                   const double* y, double* T)
{
  /* MCnucl.cpp:432-478, shape_of_nucleons 2..9 */
  double d_max = 5. * w;
  for (int ip = 0; ip < n; ip++) {
    int x_idx_left = (int)((x[ip] - d_max - g->xmin) / g->dx);
    int x_idx_right = (int)((x[ip] + d_max - g->xmin) / g->dx);
    int y_idx_left = (int)((y[ip] - d_max - g->ymin) / g->dy);
    int y_idx_right = (int)((y[ip] + d_max - g->ymin) / g->dy);
    x_idx_left = imax_(0, x_idx_left);
    x_idx_right = imin_(g->nx, x_idx_right);
    y_idx_left = imax_(0, y_idx_left);
    y_idx_right = imin_(g->ny, y_idx_right);
    for (int ix = x_idx_left; ix < x_idx_right; ix++) {
      double xg = g->xmin + ix * g->dx;
      for (int iy = y_idx_left; iy < y_idx_right; iy++) {
        double yg = g->ymin + iy * g->dy;
        T[(size_t)ix * g->ny + iy] += smooth_tn(x[ip], y[ip], xg, yg, w);
      }
    }
  }
}

void orc_add_density(const orc_grid* g, double w, int n, const double* x,
                     const double* y, const double* fluct, double* dens)
{
  /* MCnucl.cpp:822-866, shape_of_entropy 2 */
  for (int ip = 0; ip < n; ip++) {
    double box[4];
    orc_particle_box(x[ip], y[ip], w, box);
    int x_idx_left = (int)((box[0] - g->xmin) / g->dx);
    int x_idx_right = (int)((box[1] - g->xmin) / g->dx);
    int y_idx_left = (int)((box[2] - g->ymin) / g->dy);
    int y_idx_right = (int)((box[3] - g->ymin) / g->dy);
    x_idx_left = imax_(0, x_idx_left);
    x_idx_right = imin_(g->nx, x_idx_right);
    y_idx_left = imax_(0, y_idx_left);
    y_idx_right = imin_(g->ny, y_idx_right);
    double f = fluct ? fluct[ip] : 1.0;
    for (int ir = x_idx_left; ir < x_idx_right; ir++) {
      double xg = g->xmin + ir * g->dx;
      for (int jr = y_idx_left; jr < y_idx_right; jr++) {
        double yg = g->ymin + jr * g->dy;
        /* Particle::getSmoothDensity (Particle.cpp:168-174) */
        dens[(size_t)ir * g->ny + jr] += smooth_tn(x[ip], y[ip], xg, yg, w) * f;
      }
    }
  }
}

double orc_density_glauber(const orc_grid* g, double w, double alpha, int sub_model,
                           int npa, const double* xa, const double* ya, const double* fa,
                           int npb, const double* xb, const double* yb, const double* fb,
                           int ncoll, const double* xc, const double* yc, const double* fc,
                           const double* aw, double* s)
{
  /* MCnucl.cpp:689-760, which_mc_model 5, shape_of_entropy 2; sub_model 1
   * ("classical": wounded term times (1-alpha)/2, :704-712) or 2 ("Ulrich":
   * no wounded term, :713-717); aw: each pair's additional_weight
   * (MCnucl.cpp:345-348), NULL for 0 */
  size_t N = (size_t)g->nx * g->ny;
  double* rhop = (double*)calloc(N, sizeof(double));
  double* tab = (double*)calloc(N, sizeof(double));
  if (sub_model == 1) {
    orc_add_density(g, w, npa, xa, ya, fa, rhop);
    orc_add_density(g, w, npb, xb, yb, fb, rhop);
    double prefactor = (1.0 - alpha) / 2.;
    for (size_t c = 0; c < N; c++) rhop[c] = rhop[c] * prefactor;
  }
  double wsq = w * w;
  double dc_sq_max_gaussian = 25. * wsq;
  double d_max = 5. * w;
  if (alpha > 1e-8) {
    for (int icoll = 0; icoll < ncoll; icoll++) {
      double x = xc[icoll], y = yc[icoll];
      double fluctfactor = fc ? fc[icoll] : 1.0;
      double additional_weight = aw ? aw[icoll] : 0.;
      int x_idx_left = (int)((x - d_max - g->xmin) / g->dx);
      int x_idx_right = (int)((x + d_max - g->xmin) / g->dx);
      int y_idx_left = (int)((y - d_max - g->ymin) / g->dy);
      int y_idx_right = (int)((y + d_max - g->ymin) / g->dy);
      x_idx_left = imax_(0, x_idx_left);
      x_idx_right = imin_(g->nx, x_idx_right);
      y_idx_left = imax_(0, y_idx_left);
      y_idx_right = imin_(g->ny, y_idx_right);
      for (int ir = x_idx_left; ir < x_idx_right; ir++) {
        double xg = g->xmin + ir * g->dx;
        for (int jr = y_idx_left; jr < y_idx_right; jr++) {
          double yg = g->ymin + jr * g->dy;
          double dc = (x - xg) * (x - xg) + (y - yg) * (y - yg);
          if (dc <= dc_sq_max_gaussian)
            tab[(size_t)ir * g->ny + jr] += fluctfactor * (1 / (2 * M_PI * w * w)) *
                                            exp(-dc / (2 * wsq)) * (alpha + (1. - alpha) * additional_weight);
        }
      }
    }
  }
  double dndy = 0.0;
  for (size_t c = 0; c < N; c++) {
    double density = rhop[c] + tab[c];
    s[c] = density;
    dndy += density;
  }
  free(rhop);
  free(tab);
  return dndy;
}

double orc_density_trento(const orc_grid* g, double w,
                          int npa, const double* xa, const double* ya, const double* fa,
                          int npb, const double* xb, const double* yb, const double* fb,
                          double* s)
{
  /* MCnucl.cpp:762-811 */
  size_t N = (size_t)g->nx * g->ny;
  double* ra = (double*)calloc(N, sizeof(double));
  double* rb = (double*)calloc(N, sizeof(double));
  orc_add_density(g, w, npa, xa, ya, fa, ra);
  orc_add_density(g, w, npb, xb, yb, fb, rb);
  double dndy = 0;
  for (size_t c = 0; c < N; c++) {
    s[c] = sqrt(ra[c] * rb[c]);
    dndy += s[c];
  }
  free(ra);
  free(rb);
  return dndy;
}

int orc_hotspot_mask(const orc_grid* g, int nbox, const double* boxes, int ncoll, unsigned char* mask)
{
  /* The cells orc_moments weights for eps_n (MCnucl.cpp:1303-1323,
   * MakeDensity.cpp:2341-2423), as a [nx][ny] 0/1 grid. */
  int nhot = nbox + ncoll;
  memset(mask, 0, (size_t)g->nx * g->ny);
  if (nhot <= 0) return -1;
  double region[4];
  if (nbox > 0) memcpy(region, boxes, sizeof(region));
  else region[0] = region[1] = region[2] = region[3] = 0.0;
  for (int h = 1; h < nhot; h++) {
    double zero[4] = {0, 0, 0, 0};
    const double* o = (h < nbox) ? boxes + 4 * h : zero;
    region[0] = dmin_(o[0], region[0]);
    region[1] = dmax_(o[1], region[1]);
    region[2] = dmin_(o[2], region[2]);
    region[3] = dmax_(o[3], region[3]);
  }
  double dx = g->dx, dy = g->dy;
  int nYpoints = (int)((region[3] - region[2]) / dy);
  for (int h = 0; h < nhot; h++) {
    double zero[4] = {0, 0, 0, 0};
    const double* hs = (h < nbox) ? boxes + 4 * h : zero;
    for (int x_i = (int)((hs[0] - region[0]) / dx); x_i < (int)((hs[1] - region[0]) / dx); x_i++)
      for (int y_i = (int)((hs[2] - region[2]) / dy); y_i < (int)((hs[3] - region[2]) / dx); y_i++) {
        if (y_i >= nYpoints) continue;
        int i = x_i + (int)((region[0] - g->xmin) / dx);
        int j = y_i + (int)((region[2] - g->ymin) / dy);
        if (i < g->nx && i >= 0 && j < g->ny && j >= 0) mask[(size_t)i * g->ny + j] = 1;
      }
  }
  return 0;
}

int orc_moments(const orc_grid* g, const double* density,
                int nbox, const double* boxes, int ncoll,
                int from_order, int to_order,
                double* mom_real, double* mom_imag, double* momp_real,
                double* momp_imag, double* rn, double* cm)
{
  /* MakeDensity.cpp:2244-2430 */
  const int max_order = 10;
  const double eps = 1e-15;
  double norm[10], normp[10];
  for (int i = 0; i < max_order; i++) {
    mom_real[i] = mom_imag[i] = norm[i] = 0.0;
    momp_real[i] = momp_imag[i] = normp[i] = 0.0;
    rn[i] = 0.0;
  }
  int Maxx = g->nx, Maxy = g->ny;
  double Xmin = g->xmin, Ymin = g->ymin, dx = g->dx, dy = g->dy;

  double xc = 0.0, yc = 0.0, total = 0.0;
  for (int i = 0; i < Maxx; i++)
    for (int j = 0; j < Maxy; j++) {
      double x = Xmin + i * dx, y = Ymin + j * dy;
      double d = density[(size_t)i * Maxy + j];
      xc += x * d;
      yc += y * d;
      total += d;
    }
  xc /= total;
  yc /= total;
  cm[0] = xc;
  cm[1] = yc;
  cm[2] = total;

  double density_norm = 0.;
  for (int ix = 0; ix < Maxx; ix++) {
    double x_o = Xmin + ix * dx - xc;
    for (int j = 0; j < Maxy; j++) {
      double y_o = Ymin + j * dy - yc;
      double r_o = sqrt(x_o * x_o + y_o * y_o);
      double d = density[(size_t)ix * Maxy + j];
      for (int iorder = 0; iorder < max_order; iorder++) rn[iorder] += pow(r_o, iorder) * d;
      density_norm += d;
    }
  }
  for (int iorder = 0; iorder < max_order; iorder++) rn[iorder] /= density_norm;

  /* MCnucl.cpp:1303-1323 getHotSpots: participant boxes, then one
   * degenerate zero box per collision pair. */
  int nhot = nbox + ncoll;
  if (nhot <= 0) return -1;
  double region[4];
  if (nbox > 0) memcpy(region, boxes, sizeof(region));
  else region[0] = region[1] = region[2] = region[3] = 0.0;
  for (int h = 1; h < nhot; h++) {
    double zero[4] = {0, 0, 0, 0};
    const double* o = (h < nbox) ? boxes + 4 * h : zero;
    /* Box2D::overUnion (Box2D.h:49-55) */
    region[0] = dmin_(o[0], region[0]);
    region[1] = dmax_(o[1], region[1]);
    region[2] = dmin_(o[2], region[2]);
    region[3] = dmax_(o[3], region[3]);
  }
  /* MakeDensity.cpp:2341-2359 */
  int nYpoints = (int)((region[3] - region[2]) / dy);
  int nXpoints = (int)((region[1] - region[0]) / dx);
  unsigned char* important = (unsigned char*)calloc((size_t)(nXpoints > 0 ? nXpoints : 1) *
                                                        (nYpoints > 0 ? nYpoints : 1), 1);
  for (int h = 0; h < nhot; h++) {
    double zero[4] = {0, 0, 0, 0};
    const double* hs = (h < nbox) ? boxes + 4 * h : zero;
    for (int x_i = (int)((hs[0] - region[0]) / dx); x_i < (int)((hs[1] - region[0]) / dx); x_i++)
      for (int y_i = (int)((hs[2] - region[2]) / dy); y_i < (int)((hs[3] - region[2]) / dx); y_i++)
        important[(size_t)x_i * nYpoints + y_i] = 1;
  }

  for (int order = from_order; order <= to_order; order++) {
    for (int iRegion = 0; iRegion < nXpoints; iRegion++) {
      for (int jRegion = 0; jRegion < nYpoints; jRegion++) {
        if (!important[(size_t)iRegion * nYpoints + jRegion]) continue;
        int i = iRegion + (int)((region[0] - Xmin) / dx);
        int j = jRegion + (int)((region[2] - Ymin) / dy);
        double x = Xmin + i * dx - xc;
        double y = Ymin + j * dy - yc;
        double r = sqrt(x * x + y * y);
        double theta = atan2(y, x);
        if (i < Maxx && i >= 0 && j < Maxy && j >= 0) {
          double d = density[(size_t)i * Maxy + j];
          mom_real[order] += r * r * cos(order * theta) * d;
          mom_imag[order] += r * r * sin(order * theta) * d;
          norm[order] += r * r * d;
          if (order == 1) {
            momp_real[order] += pow(r, 3) * cos(order * theta) * d;
            momp_imag[order] += pow(r, 3) * sin(order * theta) * d;
            normp[order] += pow(r, 3) * d;
          } else {
            momp_real[order] += pow(r, order) * cos(order * theta) * d;
            momp_imag[order] += pow(r, order) * sin(order * theta) * d;
            normp[order] += pow(r, order) * d;
          }
        }
      }
    }
    mom_real[order] = -mom_real[order] / (norm[order] + eps);
    mom_imag[order] = -mom_imag[order] / (norm[order] + eps);
    momp_real[order] = -momp_real[order] / (normp[order] + eps);
    momp_imag[order] = -momp_imag[order] / (normp[order] + eps);
  }
  free(important);
  return 0;
}

/* ---------------------------------------------------------------- KLN */

static const double kHbarC = 0.197327053;            /* KLNModel.cpp:15 */

void orc_kln_defaults(orc_kln* k, double ecm, double lambda)
{
  /* KLNModel.cpp:36-72 and MakeDensity.cpp:97-105 */
  k->ecm = ecm;
  k->lambda = lambda;
  k->ptmin = 0.1;
  k->ptmax = 12.0;
  k->g2hfac = 2.0 * 1.0;
  k->lambda_qcd = 0.2;
  k->alpha_max = 0.5;
}

static double kln_alpha(const orc_kln* k, double q2)
{
  /* KLNModel.h:91-96, Beta0 = (33-2Nf)/(12 pi) (KLNModel.cpp:21) */
  double lq2 = k->lambda_qcd * k->lambda_qcd;
  double beta0 = (33.0 - 2.0 * 3.0) / (12 * M_PI);
  if (q2 <= lq2) return k->alpha_max;
  return dmin_(k->alpha_max, 1.0 / (beta0 * log(q2 / lq2)));
}

static double kln_wave(const orc_kln* k, double qssq, double x, double kt2)
{
  /* KLNModel.cpp:381-399 (model KLN) with KLNfunc::getFunc (KLNfunc.h:10-17) */
  double CF = (3.0 * 3.0 - 1.0) / (2 * 3.0);
  double fac = CF * 2. / (3. * M_PI * M_PI);
  double rapdep = pow(1.0 - x, 4.);
  double alp = kln_alpha(k, qssq);
  double f = (kt2 <= qssq) ? fac / alp : fac * qssq / kt2 / alp;
  return f * rapdep;
}

double orc_kln_func(const orc_kln* k, double ta, double tb, const double* u)
{
  /* KLNModel.cpp:219-277 at y = 0, dNdeta = 0, dEtdy = 0, PT_order 1 */
  double pt = k->ptmin + u[0] * (k->ptmax - k->ptmin);
  double ktmax = pt;
  double kt = ktmax * u[1];
  double phi = 2 * M_PI * u[2];
  double ktsq1 = 0.25 * (pt * pt + kt * kt + 2 * kt * pt * cos(phi));
  double ktsq2 = 0.25 * (pt * pt + kt * kt - 2 * pt * kt * cos(phi));
  double mt = pt;
  double x1 = mt / k->ecm * exp(0.0);
  double x2 = mt / k->ecm / exp(0.0);
  if (x1 > 1.0 || x2 > 1.0) return 0.0;
  /* SaturationScale (KLNModel.cpp:360-376, case KLN) */
  double qs2a = ta * 2. / 1.53 * pow(0.01 / x1, k->lambda);
  double qs2b = tb * 2. / 1.53 * pow(0.01 / x2, k->lambda);
  double fnc1 = kln_wave(k, qs2a, x1, ktsq1);
  double fnc2 = kln_wave(k, qs2b, x2, ktsq2);
  double fnc = fnc1 * fnc2;
  double scale = dmax_(ktsq1, ktsq2);
  double alp = kln_alpha(k, dmax_(scale, mt * mt));
  double result = alp * fnc;
  result *= 2.0 * M_PI * (k->ptmax - k->ptmin) * pt;
  result /= mt * mt;
  result *= 2.0 * M_PI * kt * ktmax;
  result /= 4.0;
  return result;
}

void orc_gauss_legendre01(int n, double* x, double* w)
{
  /* Legendre roots by Newton iteration on P_n, mapped to [0,1]. */
  for (int i = 0; i < (n + 1) / 2; i++) {
    double z = cos(M_PI * (i + 0.75) / (n + 0.5));
    double pp = 0.0;
    for (int it = 0; it < 100; it++) {
      double p1 = 1.0, p2 = 0.0;
      for (int j = 1; j <= n; j++) {
        double p3 = p2;
        p2 = p1;
        p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
      }
      pp = n * (z * p1 - p2) / (z * z - 1.0);
      double z1 = z;
      z = z1 - p1 / pp;
      if (fabs(z - z1) < 1e-16) break;
    }
    double wt = 2.0 / ((1.0 - z * z) * pp * pp);
    x[i] = 0.5 * (1.0 - z);
    x[n - 1 - i] = 0.5 * (1.0 + z);
    w[i] = 0.5 * wt;
    w[n - 1 - i] = 0.5 * wt;
  }
}

double orc_kln_dndy_gl(const orc_kln* k, double ta, double tb,
                       int n_pt, int n_kt, int n_phi)
{
  /* KLNModel.cpp:97-110: trEtaY * g2hfac * Norm * (integral) * 9/32,
   * Norm = 2/CF/hbarc^2 (KLNModel.cpp:20). */
  double* xp = malloc(sizeof(double) * n_pt);
  double* wp = malloc(sizeof(double) * n_pt);
  double* xk = malloc(sizeof(double) * n_kt);
  double* wk = malloc(sizeof(double) * n_kt);
  double* xf = malloc(sizeof(double) * n_phi);
  double* wf = malloc(sizeof(double) * n_phi);
  orc_gauss_legendre01(n_pt, xp, wp);
  orc_gauss_legendre01(n_kt, xk, wk);
  orc_gauss_legendre01(n_phi, xf, wf);
  double sum = 0.0;
  for (int a = 0; a < n_pt; a++)
    for (int b = 0; b < n_kt; b++)
      for (int c = 0; c < n_phi; c++) {
        double u[3] = {xp[a], xk[b], xf[c]};
        sum += wp[a] * wk[b] * wf[c] * orc_kln_func(k, ta, tb, u);
      }
  free(xp); free(wp); free(xk); free(wk); free(xf); free(wf);
  double CF = (3.0 * 3.0 - 1.0) / (2 * 3.0);
  double Norm = 2. / CF / (kHbarC * kHbarC);
  return 1.0 * k->g2hfac * Norm * sum * 9. / 32.;
}

double orc_six_point(double x, double y, double v00, double v01, double v02,
                     double v10, double v11, double v20)
{
  /* arsenal.cpp:33-54 */
  double axx = 1.0 / 2.0 * (v00 - 2 * v10 + v20);
  double axy = v00 - v01 - v10 + v11;
  double ayy = 1.0 / 2.0 * (v00 - 2 * v01 + v02);
  double bx = 1.0 / 2.0 * (-3.0 * v00 + 4 * v10 - v20);
  double by = 1.0 / 2.0 * (-3.0 * v00 + 4 * v01 - v02);
  double c = v00;
  return axx * x * x + axy * x * y + ayy * y * y + bx * x + by * y + c;
}

double orc_density_kln_table(const orc_grid* g, const double* ta, const double* tb,
                             const double* table, int tmax, double dT, double* s)
{
  /* MCnucl.cpp:654-687 */
  double dndy = 0.0;
  for (int ir = 0; ir < g->nx; ir++)
    for (int jr = 0; jr < g->ny; jr++) {
      size_t c = (size_t)ir * g->ny + jr;
      double di = ta[c] / dT;
      double dj = tb[c] / dT;
      if ((di < 0 || di >= tmax - 2) || (dj < 0 || dj >= tmax - 2)) return -1.0;
      int i = (int)floor(di), j = (int)floor(dj);
      const double* t = table;
      double v = orc_six_point(di - i, dj - j, t[i * tmax + j], t[i * tmax + j + 1],
                               t[i * tmax + j + 2], t[(i + 1) * tmax + j],
                               t[(i + 1) * tmax + j + 1], t[(i + 2) * tmax + j]);
      s[c] = v;
      dndy += v;
    }
  return dndy;
}

/* ------------------------------------------------------------ samplers */

void orc_rng_seed(orc_rng* r, uint64_t seed) { r->s = seed * 0x9E3779B97F4A7C15ull + 1; }

static uint64_t splitmix(orc_rng* r)
{
  /* splitmix64 (Steele, Lea, Flood 2014) */
  uint64_t z = (r->s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

double orc_rng_uniform(orc_rng* r) { return ((splitmix(r) >> 11) + 0.5) * (1.0 / 9007199254740992.0); }

double orc_ylm(int l, double ct)
{
  /* Nucleus.cpp:669-694 */
  if (l == 2) return (3.0 * ct * ct - 1.0) * 0.31539156525252005;
  if (l == 4) {
    double y = 35.0 * ct * ct * ct * ct;
    y -= 30.0 * ct * ct;
    y += 3.0;
    return y * 0.10578554691520431;
  }
  return 0.0;
}

void orc_rotate(double cth, double phi, double* x, double* y, double* z)
{
  /* MathBasics.cpp:41-50 */
  double x0 = *x, y0 = *y, z0 = *z;
  double cphi = cos(phi);
  double sth = sqrt(1. - cth * cth), sphi = sin(phi);
  *x = cth * cphi * x0 - sphi * y0 + sth * cphi * z0;
  *y = cth * sphi * x0 + cphi * y0 + sth * sphi * z0;
  *z = -sth * x0 + 0. * y0 + cth * z0;
}

void orc_sample_ws(orc_rng* r, double R, double a, double beta2, double beta4,
                   double ctr, double phir, int deformed, double* x, double* y, double* z)
{
  /* Nucleus.cpp:578-621, rmaxCut = R + 2.5 and rwMax (Nucleus.cpp:118-119) */
  double rmaxCut = R + 2.5;
  double rr, cx;
  if (deformed) {
    double rad1, rwMax1;
    do {
      rr = rmaxCut * pow(orc_rng_uniform(r), 1.0 / 3.0);
      cx = 1.0 - 2.0 * orc_rng_uniform(r);
      rad1 = R * (1.0 + beta2 * orc_ylm(2, cx) + beta4 * orc_ylm(4, cx));
      rwMax1 = 1.0 / (1.0 + exp(-rad1 / a));
    } while (orc_rng_uniform(r) * rwMax1 > 1.0 / (1.0 + exp((rr - rad1) / a)));
    double sx = sqrt(1.0 - cx * cx);
    double phi = 2 * M_PI * orc_rng_uniform(r);
    *x = rr * sx * cos(phi);
    *y = rr * sx * sin(phi);
    *z = rr * cx;
    orc_rotate(ctr, phir, x, y, z);
  } else {
    double rwMax = 1.0 / (1.0 + exp(-R / a));
    do {
      rr = rmaxCut * pow(orc_rng_uniform(r), 1.0 / 3.0);
    } while (orc_rng_uniform(r) * rwMax > 1.0 / (1.0 + exp((rr - R) / a)));
    cx = 1.0 - 2.0 * orc_rng_uniform(r);
    double sx = sqrt(1.0 - cx * cx);
    double phi = 2 * M_PI * orc_rng_uniform(r);
    *x = rr * sx * cos(phi);
    *y = rr * sx * sin(phi);
    *z = rr * cx;
  }
}

void orc_populate(orc_rng* r, int A, double R, double a, double beta2, double beta4,
                  int deformed, double dmin, double xcenter, double* x, double* y)
{
  /* Nucleus.cpp:187-309 (A >= 3, no NN correlation tables) */
  double rmin = dmin * dmin;
  double cx = 1.0 - 2.0 * orc_rng_uniform(r);
  double phi = 2 * M_PI * orc_rng_uniform(r);
  double* z = (double*)malloc(sizeof(double) * (A > 0 ? A : 1));
  double xcm = 0, ycm = 0, zcm = 0;
  for (int ia = 0; ia < A; ia++) {
    double px, py, pz;
    int icon;
    do {
      orc_sample_ws(r, R, a, beta2, beta4, cx, phi, deformed, &px, &py, &pz);
      icon = 0;
      for (int i = 0; i < ia; i++) {
        double r2 = (px - x[i]) * (px - x[i]) + (py - y[i]) * (py - y[i]) + (pz - z[i]) * (pz - z[i]);
        if (r2 < rmin) { icon = 1; break; }
      }
    } while (icon == 1);
    xcm += px; ycm += py; zcm += pz;
    x[ia] = px; y[ia] = py; z[ia] = pz;
  }
  for (int ia = 0; ia < A; ia++) {
    x[ia] = x[ia] - xcm / A + xcenter;
    y[ia] = y[ia] - ycm / A + 0.0;
  }
  free(z);
}

double orc_hulthen_cdf(double r)
{
  /* HulthenFunc.cpp:27-36 */
  double alpha = .228;
  double beta = 1.18;
  if (r <= 0) return 0;
  const double c = (alpha * beta * (alpha + beta)) / ((alpha - beta) * (alpha - beta));
  return 2 * c * (2 * (exp(-r * (alpha + beta)) / (alpha + beta)) - .5 * exp(-2 * alpha * r) / alpha -
                  .5 * exp(-2 * beta * r) / beta + .5 / alpha + .5 / beta - 2 / (alpha + beta));
}

static double std_normal(orc_rng* r)
{
  /* Marsaglia polar method */
  double u, v, s;
  do {
    u = 2.0 * orc_rng_uniform(r) - 1.0;
    v = 2.0 * orc_rng_uniform(r) - 1.0;
    s = u * u + v * v;
  } while (s >= 1.0 || s == 0.0);
  return u * sqrt(-2.0 * log(s) / s);
}

double orc_gamma(orc_rng* r, double shape, double scale)
{
  /* Marsaglia & Tsang, ACM TOMS 26 (2000) 363; shape < 1 by the
   * G(k) = G(k+1) U^(1/k) boost. */
  if (shape < 1.0) {
    double u = orc_rng_uniform(r);
    return orc_gamma(r, 1.0 + shape, scale) * pow(u, 1.0 / shape);
  }
  double d = shape - 1.0 / 3.0, c = 1.0 / sqrt(9.0 * d);
  for (;;) {
    double xx, v;
    do {
      xx = std_normal(r);
      v = 1.0 + c * xx;
    } while (v <= 0.0);
    v = v * v * v;
    double u = orc_rng_uniform(r);
    if (u < 1.0 - 0.0331 * xx * xx * xx * xx) return scale * d * v;
    if (log(u) < 0.5 * xx * xx + d * (1.0 - v + log(v))) return scale * d * v;
  }
}

int orc_run_events(const orc_event_cfg* c, uint64_t seed, int nev, double* eps2_out)
{
  /* MakeDensity.cpp:2143-2224 for which_mc_model 5, operation 9, use_sd 1,
   * use_ed 0; results are kept in memory instead of appended to files. */
  orc_rng rng;
  orc_rng_seed(&rng, seed);
  orc_grid g;
  orc_grid_init(&g, c->maxx, c->maxx, c->dx, c->dx);
  size_t N = (size_t)g.nx * g.ny;
  double* s = (double*)malloc(N * sizeof(double));
  double* TA = (double*)malloc(N * sizeof(double));
  double* TB = (double*)malloc(N * sizeof(double));
  int A = c->A, B = c->B;
  double *xa = malloc(sizeof(double) * A), *ya = malloc(sizeof(double) * A);
  double *xb = malloc(sizeof(double) * B), *yb = malloc(sizeof(double) * B);
  double *sxa = malloc(sizeof(double) * A), *sya = malloc(sizeof(double) * A);
  double *sxb = malloc(sizeof(double) * B), *syb = malloc(sizeof(double) * B);
  int *perm = malloc(sizeof(int) * (A > B ? A : B));
  int *wa = malloc(sizeof(int) * A), *wb = malloc(sizeof(int) * B);
  int *oa = malloc(sizeof(int) * A), *ob = malloc(sizeof(int) * B);
  int maxp = A * B;
  int *pa = malloc(sizeof(int) * maxp), *pb = malloc(sizeof(int) * maxp);
  double *pxa = malloc(sizeof(double) * A), *pya = malloc(sizeof(double) * A), *pfa = malloc(sizeof(double) * A);
  double *pxb = malloc(sizeof(double) * B), *pyb = malloc(sizeof(double) * B), *pfb = malloc(sizeof(double) * B);
  double *cx = malloc(sizeof(double) * maxp), *cy = malloc(sizeof(double) * maxp), *cf = malloc(sizeof(double) * maxp);
  double *boxes = malloc(sizeof(double) * 4 * (A + B));
  double dsq = orc_dsq(c->siginNN);
  double theta = c->gamma_theta, eps = 1e-8;
  int done = 0;
  while (done < nev) {
    int npa = 0, npb = 0, ncoll = 0;
    double b = 0;
    for (;;) {
      b = sqrt((c->bmax * c->bmax - c->bmin * c->bmin) * orc_rng_uniform(&rng) + c->bmin * c->bmin);
      orc_populate(&rng, A, c->R_a, c->a_a, 0, 0, 0, c->dmin, b / 2.0, xa, ya);
      orc_populate(&rng, B, c->R_b, c->a_b, 0, 0, 0, c->dmin, -b / 2.0, xb, yb);
      orc_sort_nucleus(A, xa, ya, c->w, perm);
      for (int i = 0; i < A; i++) { sxa[i] = xa[perm[i]]; sya[i] = ya[perm[i]]; }
      orc_sort_nucleus(B, xb, yb, c->w, perm);
      for (int i = 0; i < B; i++) { sxb[i] = xb[perm[i]]; syb[i] = yb[perm[i]]; }
      ncoll = orc_collide(A, sxa, sya, B, sxb, syb, c->w, dsq, wa, wb, oa, ob, &npa, &npb, pa, pb, maxp);
      int np = npa + npb;
      if (ncoll > 0 && np <= c->npmax && np >= c->npmin) break;
    }
    for (int k = 0; k < npa; k++) {
      pxa[k] = sxa[oa[k]]; pya[k] = sya[oa[k]];
      pfa[k] = c->fluct ? orc_gamma(&rng, (1 - c->alpha + eps) / 2. / theta, 2. / (1 - c->alpha + eps) * theta) : 1.0;
      orc_particle_box(pxa[k], pya[k], c->w, boxes + 4 * k);
    }
    for (int k = 0; k < npb; k++) {
      pxb[k] = sxb[ob[k]]; pyb[k] = syb[ob[k]];
      pfb[k] = c->fluct ? orc_gamma(&rng, (1 - c->alpha + eps) / 2. / theta, 2. / (1 - c->alpha + eps) * theta) : 1.0;
      orc_particle_box(pxb[k], pyb[k], c->w, boxes + 4 * (npa + k));
    }
    for (int k = 0; k < ncoll; k++) {
      cx[k] = (sxa[pa[k]] + sxb[pb[k]]) / 2.0;
      cy[k] = (sya[pa[k]] + syb[pb[k]]) / 2.0;
      cf[k] = c->fluct ? orc_gamma(&rng, (c->alpha + eps) / theta, 1. / (c->alpha + eps) * theta) : 1.0;
    }
    /* MakeDensity.cpp:2167: calculateThickness runs every event */
    memset(TA, 0, N * sizeof(double));
    memset(TB, 0, N * sizeof(double));
    orc_thickness(&g, c->w, npa, pxa, pya, TA);
    orc_thickness(&g, c->w, npb, pxb, pyb, TB);
    orc_density_glauber(&g, c->w, c->alpha, 1, npa, pxa, pya, pfa, npb, pxb, pyb, pfb,
                        ncoll, cx, cy, cf, NULL, s);
    double mr[10], mi[10], mpr[10], mpi[10], rn[10], cm[3];
    orc_moments(&g, s, npa + npb, boxes, ncoll, c->from_order, c->to_order, mr, mi, mpr, mpi, rn, cm);
    if (eps2_out) {
      double* o = eps2_out + 5 * (size_t)done;
      o[0] = mr[2];
      o[1] = mi[2];
      o[2] = b;
      o[3] = npa + npb;
      o[4] = ncoll;
    }
    done++;
  }
  free(s); free(TA); free(TB); free(xa); free(ya); free(xb); free(yb);
  free(sxa); free(sya); free(sxb); free(syb); free(perm); free(wa); free(wb); free(oa); free(ob);
  free(pa); free(pb); free(pxa); free(pya); free(pfa); free(pxb); free(pyb); free(pfb);
  free(cx); free(cy); free(cf); free(boxes);
  return done;
}

/* ------------------------------------------------------------------ 3D profiles */
void orc_p3d_setup(double ecm, int n, double* y_beam, double* sigma_inel_fm2, double* eta, double* proj,
                   double* targ, double* maxes)
{
  /* profile_3d.cpp:24-27 */
  const double yb = atanh(sqrt(1. - 1. / pow(ecm / 2., 2)));
  *y_beam = yb;
  *sigma_inel_fm2 = orc_sigma_inel(ecm) * 0.1;
  /* profile_3d.cpp:133-155 */
  const double eta_peak = 2.5, sigma_eta_out = 0.5;
  const double deta = 2 * yb / (n - 1);
  maxes[0] = maxes[1] = 0.0;
  for (int i = 0; i < n; i++) {
    const double e = -yb + i * deta;
    double f = 1.0;
    if (fabs(e) > eta_peak) f = exp(-pow((fabs(e) - eta_peak), 2) / (2. * sigma_eta_out * sigma_eta_out));
    const double l = (1. - e / yb) * f, r = (1. + e / yb) * f;
    if (l > maxes[0]) maxes[0] = l;
    if (r > maxes[1]) maxes[1] = r;
    eta[i] = e;
    proj[i] = l;
    targ[i] = r;
  }
}

long orc_p3d_search(const double* A, int n, double v)
{
  /* profile_3d.cpp:402-434 with skip_out_of_range */
  int i = 0, f = n - 1, m;
  if (v > A[f]) return -1;
  if (v < A[i]) return -1;
  m = (int)floor((f + i) / 2.);
  while ((f - i) > 1) {
    if (A[m] < v) i = m;
    else f = m;
    m = (int)floor((f + i) / 2.);
  }
  return i;
}

void orc_p3d_smear(int nx, int ny, int neta, double dx, double dy, double deta, int n, const double* src,
                   double* rho)
{
  /* profile_3d.cpp:75-91: cell centres */
  for (int i = 0; i < n; i++) {
    const double* s = src + 6 * (size_t)i;
    const double px = s[0], py = s[1], pe = s[2], sx = s[3], sy = s[4], se = s[5];
    /* profile_3d.cpp:310-323: integer (n - 1)/2 in the index, clamped windows */
    const int ix0 = (int)(px / dx + (nx - 1) / 2);
    const int iy0 = (int)(py / dy + (ny - 1) / 2);
    const int ie0 = (int)(pe / deta + (neta - 1) / 2);
    const int rx = (int)(6 * sx / dx), ry = (int)(6 * sy / dy), re = (int)(6 * se / deta);
    const int xl = ix0 - rx > 0 ? ix0 - rx : 0, xr = ix0 + rx < nx ? ix0 + rx : nx;
    const int yl = iy0 - ry > 0 ? iy0 - ry : 0, yr = iy0 + ry < ny ? iy0 + ry : ny;
    const int el = ie0 - re > 0 ? ie0 - re : 0, er = ie0 + re < neta ? ie0 + re : neta;
    /* profile_3d.cpp:325-343 */
    for (int j = el; j < er; j++) {
      const double ej = (j - (neta - 1) / 2.) * deta;
      const double dis_eta = (ej - pe) * (ej - pe) / (2. * se * se);
      const double norm_eta = 1. / sqrt(2. * M_PI * se * se);
      for (int k = xl; k < xr; k++) {
        const double xk = (k - (nx - 1) / 2.) * dx;
        const double dis_x = (xk - px) * (xk - px) / (2. * sx * sx);
        const double norm_x = 1. / sqrt(2. * M_PI * sx * sx);
        for (int l = yl; l < yr; l++) {
          const double yl_ = (l - (ny - 1) / 2.) * dy;
          const double dis_y = (yl_ - py) * (yl_ - py) / (2. * sy * sy);
          const double norm_y = 1. / sqrt(2. * M_PI * sy * sy);
          rho[((size_t)j * nx + k) * ny + l] += exp(-dis_eta - dis_x - dis_y) * norm_eta * norm_x * norm_y;
        }
      }
    }
  }
}
