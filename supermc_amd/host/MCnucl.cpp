#include "MCnucl.h"

#include <algorithm>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <iostream>

#include "smc_host.h"

using smc_host::check;

MCnucl::MCnucl(ParameterReader* paraRdr) {
  p_ = smc_host::params_from_reader(*paraRdr, 1);
  seed_ = smc_host::seed_from_reader(*paraRdr);
  check(nullptr, smc_create(&ctx_, &p_), "MCnucl::MCnucl");
  smc_derived d;
  check(ctx_, smc_get_derived(ctx_, &d), "MCnucl::MCnucl");
  nx_ = d.nx;
  ny_ = d.ny;
  box_half_ = 8 * std::max(d.width, p_.quark_width) / 2;
  dx_ = p_.dx;
  dy_ = p_.dy;
  npmin_ = p_.Npmin;
  npmax_ = p_.Npmax;
  const size_t n = (size_t)nx_ * ny_;
  ta_.assign(n, 0.0);
  tb_.assign(n, 0.0);
  rho_.assign(n, 0.0);
}

MCnucl::~MCnucl() { smc_destroy(ctx_); }

// MCnucl.cpp:208-214 with Nucleus::populate: both nuclei at +-b/2, sampled
// on the device (event counter_ of stream seed_), then the collision step.
void MCnucl::generateNuclei(double b) {
  check(ctx_, smc_sample_b(ctx_, seed_, counter_++, 1, &b), "MCnucl::generateNuclei");
  have_event_ = true;
  ncoll_ = npart1_ = npart2_ = 0;
}

void MCnucl::deleteNucleus() { have_event_ = false; }

// MCnucl.cpp:217-308: Npart1/2 and the pair list; returns Ncoll.
int MCnucl::getBinaryCollision() {
  if (!have_event_) {
    std::cerr << "MCnucl::getBinaryCollision: call generateNuclei first" << std::endl;
    std::exit(1);
  }
  check(ctx_, smc_collide(ctx_, &npart1_, &npart2_, &ncoll_), "MCnucl::getBinaryCollision");
  return ncoll_;
}

// MCnucl.cpp:388-393
int MCnucl::CentralityCut() {
  const int n = npart1_ + npart2_;
  return (n <= npmax_ && n >= npmin_) ? 1 : 0;
}

// MCnucl.cpp:398-478
void MCnucl::calculateThickness() {
  std::vector<float> a(ta_.size()), b(tb_.size());
  check(ctx_, smc_thickness(ctx_, a.data(), b.data()), "MCnucl::calculateThickness");
  for (size_t i = 0; i < a.size(); ++i) {
    ta_[i] = a[i];
    tb_[i] = b[i];
  }
}

// MCnucl.cpp:626-820 for ipt < 0
void MCnucl::setDensity(int iy, int ipt) {
  if (iy != 0 || ipt >= 0) {
    std::cerr << "MCnucl::setDensity: only iy = 0, ipt < 0 (pT-integrated, y = 0) is on this path" << std::endl;
    std::exit(1);
  }
  std::vector<float> s(rho_.size());
  check(ctx_, smc_density(ctx_, s.data(), &dndy_), "MCnucl::setDensity");
  for (size_t i = 0; i < s.size(); ++i) rho_[i] = s[i];
}

// MCnucl.cpp:911-960 (built on the device, kept in the context)
void MCnucl::makeTable() { check(ctx_, smc_kln_make_table(ctx_, nullptr), "MCnucl::makeTable"); }

namespace {
[[noreturn]] void off_path(const char* what) {
  std::cout << what << ": pT-resolved densities and tables are not built in this library." << std::endl;
  std::exit(1);
}
}  // namespace

double MCnucl::getRho(int i, int x, int y, int pt) {
  if (pt >= 0) off_path("MCnucl::getRho(i, x, y, pt)");
  return getRho(i, x, y);
}

void MCnucl::makeTable(double, double, int) { off_path("MCnucl::makeTable(double, double, int)"); }

void MCnucl::dumpdNdydptTable5Col(std::string, double****, const int) { off_path("MCnucl::dumpdNdydptTable5Col"); }

void MCnucl::dumpdNdyTable4Col(std::string filename, double*** dNdyTable, const int iy) {
  smc_derived d{};
  check(ctx_, smc_get_derived(ctx_, &d), "MCnucl::dumpdNdyTable4Col");
  std::ofstream of(filename.c_str(), std::ios_base::app);
  const double y = 0.0;  // the one rapidity bin at y = 0 (ny = 1, ymax = 0)
  for (int i = 1; i < d.kln_tmax; ++i)
    for (int j = 1; j < d.kln_tmax; ++j)
      of << std::fixed << std::setprecision(3) << std::setw(10) << y << std::fixed << std::setprecision(3)
         << std::setw(10) << d.kln_dT * i << std::fixed << std::setprecision(3) << std::setw(10) << d.kln_dT * j
         << std::setprecision(12) << std::setw(22) << dNdyTable[iy][i][j] << std::endl;
}

// MCnucl.cpp:1177-1211 / 1213-1220
void MCnucl::dumpBinaryTable(std::string filename) { smc_host::dump_binary_table(ctx_, p_, 0, filename.c_str()); }
void MCnucl::dumpparticipantTable(std::string filename) {
  smc_host::dump_participant_table(ctx_, p_, 0, filename.c_str());
}

int MCnucl::getSpectators() {
  std::vector<double> xa, xb;
  std::vector<int> wa, wb;
  getNucleons(xa, wa, xb, wb);
  int n = 0;
  for (int w : wa) n += w < 0;
  for (int w : wb) n += w < 0;
  return n;
}

void MCnucl::dumpSpectatorsTable(int event) { smc_host::dump_spectators_table(ctx_, p_, 0, event); }

void MCnucl::calculate_rho_binary() {
  std::vector<float> r(rho_.size());
  check(ctx_, smc_rho_binary(ctx_, r.data()), "MCnucl::calculate_rho_binary");
  rho_binary_.assign(r.begin(), r.end());
}

void MCnucl::calculate_spectator_density() {
  std::vector<float> a(rho_.size()), b(rho_.size());
  check(ctx_, smc_spectator_density(ctx_, a.data(), b.data()), "MCnucl::calculate_spectator_density");
  spec1_.assign(a.begin(), a.end());
  spec2_.assign(b.begin(), b.end());
}

void MCnucl::getNucleons(std::vector<double>& xy_proj, std::vector<int>& wound_proj, std::vector<double>& xy_targ,
                         std::vector<int>& wound_targ) {
  xy_proj.assign(2 * (size_t)p_.Aproj, 0.0);
  xy_targ.assign(2 * (size_t)p_.Atarg, 0.0);
  wound_proj.assign(p_.Aproj, -1);
  wound_targ.assign(p_.Atarg, -1);
  std::vector<float> fp(p_.Aproj), ft(p_.Atarg);
  check(ctx_,
        smc_get_nucleons(ctx_, 0, xy_proj.data(), wound_proj.data(), fp.data(), xy_targ.data(), wound_targ.data(),
                         ft.data()),
        "MCnucl::getNucleons");
}

// MCnucl.cpp:1303-1323.  A participant's box is its 8 w square (with the
// quarks at the nucleon centre, Particle.cpp:16-21, 94-99) about its
// current position; a collision pair contributes the all-zero box.
Box2D MCnucl::getHotSpots(std::vector<Box2D>& hotSpots) {
  hotSpots.clear();
  int n = 0;
  const smc_status st = smc_get_sources(ctx_, 0, 0, nullptr, nullptr, nullptr, &n);  // the count only
  if (st != SMC_ERR_CAPACITY) check(ctx_, st, "MCnucl::getHotSpots");
  std::vector<double> xy(2 * (size_t)std::max(n, 1));
  std::vector<int> kind(std::max(n, 1));
  check(ctx_, smc_get_sources(ctx_, 0, n, xy.data(), kind.data(), nullptr, &n), "MCnucl::getHotSpots");
  for (int k = 0; k < n; ++k) {
    Box2D b;
    if (kind[k] != 2) {  // baseBox.setCenter(x, y) on the zero box, then setSquareDimensions(8 w)
      b.setCenter(xy[2 * k], xy[2 * k + 1]);
      b.setSquareDimensions(2 * box_half_);
    }
    hotSpots.push_back(b);
  }
  if (hotSpots.empty()) {
    std::cerr << "MCnucl::getHotSpots: the event has no sources" << std::endl;
    std::exit(1);
  }
  Box2D u = hotSpots[0];
  for (size_t i = 1; i < hotSpots.size(); ++i) u.overUnion(hotSpots[i]);
  return u;
}

void MCnucl::recenterGrid(int iy, int n) {
  (void)iy;
  check(ctx_, smc_recenter(ctx_, n, 0, nullptr), "MCnucl::recenterGrid");
}

void MCnucl::rotateGrid(int iy, int n) {
  (void)iy;
  check(ctx_, smc_recenter(ctx_, n, 1, nullptr), "MCnucl::rotateGrid");
}
