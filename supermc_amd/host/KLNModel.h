// KLNModel -- the reference's KLNModel interface (KLNModel.h:46-72) for
// pT-integrated dN/dy at y = 0 (KLNModel::getdNdy, KLNModel.cpp:97-125),
// evaluated on the device by a Gauss-Legendre product rule over (pT, kT,
// phi) instead of the reference's BASES Monte-Carlo integral.
//
// Supported: y = 0, pt < 0 (pT-integrated), pt_order = 1, the KLN uGD
// (mode 7).  Other arguments print a message and exit, like the reference's
// fatal paths.  getdNdyBatch evaluates many (T_A, T_B) points in one call.
#ifndef SMC_HOST_KLNMODEL_H
#define SMC_HOST_KLNMODEL_H

#include <vector>

#include "smc.h"

class KLNModel {
 public:
  // f (the reference's UnintegPartonDist*) is unused: the KLN uGD is built in
  KLNModel(double srt, int mode, void* f = nullptr);
  virtual ~KLNModel();
  KLNModel(const KLNModel&) = delete;
  KLNModel& operator=(const KLNModel&) = delete;

  double getdNdy(double y, double npart1, double npart2, double pt = -1.0, int pt_order = 1);
  void getdNdyBatch(int n, const double* ta, const double* tb, double* out);

  void setNormalization(double c) { g2hfac_ *= c; }
  void setg2hfac(double c) { g2hfac_ *= c; }
  void setLambdaSaturation(double l);
  void setPtRange(double a, double b);
  void getPtRange(double* a, double* b) { *a = ptmin_; *b = ptmax_; }
  void setQuadrature(int n_pt, int n_kt, int n_phi);  // extension: node counts
  double getEcm() { return ecm_; }
  int getModel() { return mode_; }

 private:
  void ensure();
  smc_ctx* ctx_ = nullptr;
  double ecm_, lambda_ = 0.138, ptmin_ = 0.1, ptmax_ = 12.0, g2hfac_ = 1.0;
  int mode_, npt_ = 32, nkt_ = 32, nphi_ = 48;
};

#endif
