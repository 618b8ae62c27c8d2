#include "smc_host.h"

#include <sys/time.h>

#include <cstdlib>
#include <iostream>

namespace smc_host {

smc_params params_from_reader(ParameterReader& r, int max_events) {
  smc_params p;
  smc_params_default(&p);
  // MCnucl.cpp:22-155, MakeDensity.cpp:21-84
  p.which_mc_model = (int)r.getVal("which_mc_model");
  p.sub_model = (int)r.getVal("sub_model");
  p.alpha = r.getVal("alpha");
  p.Aproj = (int)r.getVal("Aproj");
  p.Atarg = (int)r.getVal("Atarg");
  p.proj_deformed = (int)r.getVal("proj_deformed");
  p.targ_deformed = (int)r.getVal("targ_deformed");
  p.shape_of_nucleons = (int)r.getVal("shape_of_nucleons");
  p.collision_criterion = (int)r.getVal("collision_criterion");
  p.shape_of_entropy = (int)r.getVal("shape_of_entropy");
  p.quark_width = r.getVal("quark_width");
  p.gauss_nucl_width = r.getVal("gauss_nucl_width");
  p.ecm = r.getVal("ecm");
  p.bmin = r.getVal("bmin");
  p.bmax = r.getVal("bmax");
  p.Npmin = (int)r.getVal("Npmin");
  p.Npmax = (int)r.getVal("Npmax");
  p.cc_fluctuation_model = (int)r.getVal("cc_fluctuation_model");
  p.maxx = r.getVal("maxx");
  p.maxy = r.getVal("maxy");
  p.dx = r.getVal("dx");
  p.dy = r.getVal("dy");
  p.finalFactor = r.getVal("finalFactor");
  if (p.which_mc_model == 1) {
    p.lambda = r.getVal("lambda");
    p.tmax = (int)r.getVal("tmax");
    p.tmax_subdivision = (int)r.getVal("tmax_subdivision");
  }
  if (p.cc_fluctuation_model == 6) p.cc_fluctuation_Gamma_theta = r.getVal("cc_fluctuation_Gamma_theta");
  if (r.exist("ecc_from_order")) p.ecc_from_order = (int)r.getVal("ecc_from_order");
  if (r.exist("ecc_to_order")) p.ecc_to_order = (int)r.getVal("ecc_to_order");
  // extension keys (absent: the reference's hard-coded values)
  struct D { const char* k; double* v; } dk[] = {
      {"sigma_nn_mb", &p.sigma_nn_mb},   {"proj_ws_R", &p.proj_ws_R},   {"proj_ws_a", &p.proj_ws_a},
      {"targ_ws_R", &p.targ_ws_R},       {"targ_ws_a", &p.targ_ws_a},   {"proj_beta2", &p.proj_beta2},
      {"proj_beta4", &p.proj_beta4},     {"targ_beta2", &p.targ_beta2}, {"targ_beta4", &p.targ_beta4},
      {"nucleon_min_dist", &p.nucleon_min_dist}, {"kln_pt_min", &p.kln_pt_min}, {"kln_pt_max", &p.kln_pt_max}};
  for (auto& e : dk)
    if (r.exist(e.k)) *e.v = r.getVal(e.k);
  struct I { const char* k; int* v; } ik[] = {
      {"kln_n_pt", &p.kln_n_pt}, {"kln_n_kt", &p.kln_n_kt}, {"kln_n_phi", &p.kln_n_phi},
      {"kln_direct", &p.kln_direct}, {"max_pairs", &p.max_pairs}, {"max_attempts", &p.max_attempts},
      {"smear_path", &p.smear_path}};
  for (auto& e : ik)
    if (r.exist(e.k)) *e.v = (int)r.getVal(e.k);
  p.max_events = max_events;
  return p;
}

uint64_t seed_from_reader(ParameterReader& r) {
  long s = r.exist("randomSeed") ? (long)r.getVal("randomSeed") : -1;
  if (s < 0) {
    timeval t;
    gettimeofday(&t, nullptr);
    s = t.tv_usec;
  }
  return (uint64_t)s;
}

void check(smc_ctx* ctx, smc_status st, const char* where) {
  if (st == SMC_OK) return;
  std::cerr << where << ": " << (ctx ? smc_last_error(ctx) : "smc_create failed") << " (status " << st << ")"
            << std::endl;
  std::exit(st == SMC_ERR_MISSING_KEY ? -1 : 1);
}

}  // namespace smc_host
