// ref_harness.cpp -- extern "C" probes into the parts of the reference that
// build from their own sources (no GSL, no missing blobs): Regge96, KLNModel
// + Bases (the BASES/VEGAS integrator), arsenal, Box2D, MathBasics,
// HulthenFunc, ParameterReader.  Built by oracle/Makefile into oracle/_ref/
// from the reference sources where they lie; used only by tests to pin the
// oracle restatement (oracle/smc_oracle.c).  TEST INFRASTRUCTURE ONLY.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "Box2D.h"
#include "HulthenFunc.h"
#include "KLNModel.h"
#include "KLNfunc.h"
#include "MathBasics.h"
#include "ParameterReader.h"
#include "Regge96.h"
#include "arsenal.h"

namespace {

// Captures KLNModel::func values at caller-chosen points.  func depends on
// the private thickness state that getdNdy sets, so the probe runs one
// getdNdy and evaluates the base-class integrand on its first callback.
class ProbeKLN : public KLNModel {
 public:
  ProbeKLN(double srt, UnintegPartonDist* f) : KLNModel(srt, 7, f), n_(0), u_(0), out_(0), done_(false) {}
  void arm(int n, const double* u, double* out) { n_ = n; u_ = u; out_ = out; done_ = false; }
  double func(double* x) {
    if (!done_) {
      done_ = true;
      for (int i = 0; i < n_; i++) {
        double p[3] = {u_[3 * i], u_[3 * i + 1], u_[3 * i + 2]};
        out_[i] = KLNModel::func(p);
      }
    }
    return KLNModel::func(x);
  }
 private:
  int n_;
  const double* u_;
  double* out_;
  bool done_;
};

}  // namespace

extern "C" {

double ref_sigma_inel(double ecm) {
  // MCnucl.cpp:58-64
  double sig = hadronxsec::totalXsection(ecm, 0);
  double sigel = hadronxsec::elasticXsection(sig, ecm, 0, 0);
  return sig - sigel;
}

double ref_kln_dndy(double ecm, double lambda, double ta, double tb) {
  // MakeDensity.cpp:88-105 set-up, MCnucl.cpp:939 call
  KLNfunc f;
  KLNModel k(ecm, 7, &f);
  k.setLambdaSaturation(lambda);
  k.setNormalization(1.0);
  return k.getdNdy(0.0, ta, tb, -1, 1);
}

void ref_kln_func(double ecm, double lambda, double ta, double tb, int n,
                  const double* u, double* out) {
  KLNfunc f;
  ProbeKLN k(ecm, &f);
  k.setLambdaSaturation(lambda);
  k.setNormalization(1.0);
  k.arm(n, u, out);
  k.getdNdy(0.0, ta, tb, -1, 1);
}

double ref_six_point(double x, double y, double v00, double v01, double v02,
                     double v10, double v11, double v20) {
  return sixPoint2dInterp(x, y, v00, v01, v02, v10, v11, v20);
}

// Box of a nucleon built at (x0,y0) and moved to (x1,y1), following the call
// sequence of Particle.cpp:16-21 and :94-99 (quarks at the centre with width
// qw, boxes from Quark.h:28-35 and Quark.cpp:5-10) and Particle::setX/setY
// (Particle.cpp:176-185) as used by Nucleus::populate (Nucleus.cpp:296-309).
void ref_particle_box(double x0, double y0, int moved, double x1, double y1,
                      double w, double qw, double* out) {
  Box2D base;
  base.setCenter(x0, y0);
  base.setSquareDimensions(8 * w);
  Box2D bounding = base;
  for (int q = 0; q < 3; q++) {
    Box2D qbox;
    qbox.setCenter(0.0, 0.0);
    qbox.setSquareDimensions(8 * qw);
    Box2D placed = qbox;
    placed.setCenter(x0 + 0.0, y0 + 0.0);
    bounding.overUnion(placed);
  }
  if (moved) {
    bounding.setX(x1);
    bounding.setY(y1);
  }
  out[0] = bounding.getXL();
  out[1] = bounding.getXR();
  out[2] = bounding.getYL();
  out[3] = bounding.getYR();
}

void ref_overunion(const double* a, const double* b, double* out) {
  Box2D A(a[0], a[1], a[2], a[3]);
  Box2D B(b[0], b[1], b[2], b[3]);
  A.overUnion(B);
  out[0] = A.getXL();
  out[1] = A.getXR();
  out[2] = A.getYL();
  out[3] = A.getYR();
}

void ref_rotate(double cth, double phi, double* xyz) {
  Point3D p(xyz[0], xyz[1], xyz[2]);
  p.rotate(cth, phi);
  xyz[0] = p.x;
  xyz[1] = p.y;
  xyz[2] = p.z;
}

double ref_hulthen_cdf(double r) { return HulthenFunc::CDF(r); }

double ref_hulthen_invcdf(double u) {
  HulthenFunc h;
  return h.invCDF(u);
}

// Reads parameters the way main.cpp:22-24 does; returns the value of `name`
// or writes 0 to *found when absent.
double ref_param_value(const char* path, int argc, const char** argv,
                       const char* name, int* found) {
  ParameterReader p;
  p.readFromFile(path);
  std::vector<char*> args;
  args.push_back(const_cast<char*>("superMC.e"));
  for (int i = 0; i < argc; i++) args.push_back(const_cast<char*>(argv[i]));
  p.readFromArguments((long)args.size(), &args[0]);
  if (!p.exist(name)) {
    *found = 0;
    return 0.0;
  }
  *found = 1;
  return p.getVal(name);
}

}  // extern "C"
