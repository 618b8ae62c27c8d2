#include "smc_host.h"

#include <sys/time.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <vector>

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
  p.gaussian_lambda = r.getVal("gaussian_lambda");  // read unconditionally, as GaussianNucleonsCal.cpp:27
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

namespace {

struct EventNucleons {
  std::vector<double> xy[2];
  std::vector<int> wound[2];
  std::vector<int> order[2];  // participant indices in wound order
  std::vector<int> pa, pb;    // pair indices
};

EventNucleons fetch_event(smc_ctx* ctx, const smc_params& p, int ev, bool pairs) {
  EventNucleons n;
  for (int nu = 0; nu < 2; ++nu) {
    n.xy[nu].assign(2 * SMC_MAX_A, 0.0);
    n.wound[nu].assign(SMC_MAX_A, -1);
  }
  check(ctx, smc_get_nucleons(ctx, ev, n.xy[0].data(), n.wound[0].data(), nullptr, n.xy[1].data(),
                              n.wound[1].data(), nullptr),
        "dumpBinaryTable");
  for (int nu = 0; nu < 2; ++nu) {
    int np = 0;
    for (int i = 0; i < SMC_MAX_A; ++i) np = std::max(np, n.wound[nu][i] + 1);
    n.order[nu].assign(np, -1);
    for (int i = 0; i < SMC_MAX_A; ++i)
      if (n.wound[nu][i] >= 0) n.order[nu][n.wound[nu][i]] = i;
  }
  if (pairs) {
    int np = 0;
    const int cap = p.Aproj * p.Atarg;  // every possible pair
    n.pa.resize(cap);
    n.pb.resize(cap);
    check(ctx, smc_get_pairs(ctx, ev, cap, n.pa.data(), n.pb.data(), nullptr, &np), "dumpBinaryTable");
    n.pa.resize(np);
    n.pb.resize(np);
  }
  return n;
}

// Nucleus::dumpParticipants (Nucleus.cpp:753-764)
void write_participants(std::ofstream& of, const EventNucleons& n) {
  for (int nu = 0; nu < 2; ++nu)
    for (int i : n.order[nu])
      of << std::setprecision(3) << std::setw(10) << n.xy[nu][2 * i] << "   " << std::setprecision(3)
         << std::setw(10) << n.xy[nu][2 * i + 1] << "   " << nu + 1 << std::endl;
}

}  // namespace

void dump_binary_table(smc_ctx* ctx, const smc_params& p, int ev, const char* binary_file) {
  const EventNucleons n = fetch_event(ctx, p, ev, true);
  {
    // CollisionPair position: the midpoint of its two nucleons
    std::ofstream of(binary_file, std::ios_base::app);
    for (size_t k = 0; k < n.pa.size(); ++k) {
      const double x = (n.xy[0][2 * n.pa[k]] + n.xy[1][2 * n.pb[k]]) / 2;
      const double y = (n.xy[0][2 * n.pa[k] + 1] + n.xy[1][2 * n.pb[k] + 1]) / 2;
      of << std::setprecision(3) << std::setw(10) << x << std::setprecision(3) << std::setw(10) << y << std::endl;
    }
  }
  {
    std::ofstream of("data/wounded.data");
    write_participants(of, n);
  }
  {
    // Nucleus::dumpQuarks (Nucleus.cpp:780-797): three quarks per
    // participant; box = Box2D of side 8 quark_width centred on the quark
    std::ofstream of("data/quarks.data", std::ios_base::app);
    const double h = 8 * p.quark_width / 2;
    for (int nu = 0; nu < 2; ++nu)
      for (int i : n.order[nu]) {
        const double x = n.xy[nu][2 * i], y = n.xy[nu][2 * i + 1];
        std::ostringstream box;
        box << (-h) + x << " " << h + x << " " << (-h) + y << " " << h + y;
        for (int q = 0; q < 3; ++q)
          of << std::setprecision(3) << std::setw(10) << x << std::setprecision(3) << std::setw(10) << y << " "
             << box.str() << std::endl;
      }
  }
  // Nucleus::dumpNucleons (Nucleus.cpp:766-778): every nucleon, collision order
  const int A[2] = {p.Aproj, p.Atarg};
  for (int nu = 0; nu < 2; ++nu) {
    std::ofstream of(nu == 0 ? "data/nucl1.data" : "data/nucl2.data");
    for (int i = 0; i < A[nu]; ++i)
      of << std::setprecision(3) << std::setw(10) << n.xy[nu][2 * i] << std::setprecision(3) << std::setw(10)
         << n.xy[nu][2 * i + 1] << std::endl;
  }
}

double spectator_rapidity(double ecm) {
  const double v_z = std::sqrt(1. - 1. / ((ecm / 2.) * (ecm / 2.)));
  return 0.5 * std::log((1. + v_z) / (1. - v_z + 1e-100));
}

int dump_spectators_table(smc_ctx* ctx, const smc_params& p, int ev, int event) {
  const EventNucleons n = fetch_event(ctx, p, ev, false);
  const double Y = spectator_rapidity(p.ecm);
  std::ostringstream name;
  name << "data/Spectators_event_" << event << ".dat";
  std::ofstream of(name.str());
  const int A[2] = {p.Aproj, p.Atarg};
  int count = 0;
  for (int nu = 0; nu < 2; ++nu)
    for (int i = 0; i < A[nu]; ++i) {
      if (n.wound[nu][i] >= 0) continue;
      of << std::scientific << std::setprecision(4) << std::setw(10) << n.xy[nu][2 * i] << "  " << n.xy[nu][2 * i + 1]
         << "  " << (nu == 0 ? Y : -Y) << std::endl;
      ++count;
    }
  return count;
}

void dump_participant_table(smc_ctx* ctx, const smc_params& p, int ev, const char* filename) {
  const EventNucleons n = fetch_event(ctx, p, ev, false);
  std::ofstream of(filename, std::ios_base::app);
  write_participants(of, n);
}

}  // namespace smc_host
