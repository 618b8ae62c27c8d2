#include "KLNModel.h"

#include <cstdlib>
#include <iostream>

#include "smc_host.h"

KLNModel::KLNModel(double srt, int mode, void*) : ecm_(srt), mode_(mode) {
  if (mode != 7) {
    std::cerr << "KLNModel: only the KLN unintegrated gluon distribution (model 7) is on this path" << std::endl;
    std::exit(1);
  }
}

KLNModel::~KLNModel() { smc_destroy(ctx_); }

void KLNModel::setLambdaSaturation(double l) {
  lambda_ = l;
  smc_destroy(ctx_);
  ctx_ = nullptr;
}

void KLNModel::setPtRange(double a, double b) {
  ptmin_ = a;
  ptmax_ = b;
  smc_destroy(ctx_);
  ctx_ = nullptr;
}

void KLNModel::setQuadrature(int n_pt, int n_kt, int n_phi) {
  npt_ = n_pt;
  nkt_ = n_kt;
  nphi_ = n_phi;
  smc_destroy(ctx_);
  ctx_ = nullptr;
}

void KLNModel::ensure() {
  if (ctx_) return;
  smc_params p;
  smc_params_default(&p);
  p.which_mc_model = 1;
  p.sub_model = 7;
  p.ecm = ecm_;
  p.lambda = lambda_;
  p.kln_pt_min = ptmin_;
  p.kln_pt_max = ptmax_;
  p.kln_n_pt = npt_;
  p.kln_n_kt = nkt_;
  p.kln_n_phi = nphi_;
  p.shape_of_nucleons = 4;
  p.collision_criterion = 1;
  p.max_events = 1;
  smc_host::check(nullptr, smc_create(&ctx_, &p), "KLNModel");
}

void KLNModel::getdNdyBatch(int n, const double* ta, const double* tb, double* out) {
  ensure();
  smc_host::check(ctx_, smc_kln_dndy(ctx_, n, ta, tb, out), "KLNModel::getdNdy");
  // setNormalization scales g2hfac (KLNModel.h:55), a linear factor of dN/dy
  for (int i = 0; i < n; ++i) out[i] *= g2hfac_;
}

double KLNModel::getdNdy(double y, double npart1, double npart2, double pt, int pt_order) {
  if (y != 0.0 || pt >= 0.0 || pt_order != 1) {
    std::cerr << "KLNModel::getdNdy: only y = 0, pT-integrated, pt_order 1 is on this path" << std::endl;
    std::exit(1);
  }
  double out = 0.0;
  getdNdyBatch(1, &npart1, &npart2, &out);
  return out;
}
