// MCnucl -- one-event-at-a-time interface of the reference's MCnucl
// (MCnucl.h:87-146) on top of libsmc.  Each call drives the gfx950 kernels
// for a batch of one event; batched event loops live in MakeDensity.
//
// Differences from the reference class (the rest keeps its signatures and
// meaning):
//  - generateNuclei(b) samples both nuclei and runs the collision step on
//    the device; getBinaryCollision() returns that event's Ncoll.
//  - setDensity(iy, ipt) supports ipt < 0 (pT-integrated) and iy = 0.
//  - KLN (which_mc_model 1) uses the library's dN/dy table, so setKLN /
//    setLgX are not needed.
//  - recenterGrid / rotateGrid move the participants and collision pairs
//    (the density's sources); spectators keep their positions.
//  - lastCx1 .. lastPh2 are 0 (the reference never assigns them).
#ifndef SMC_HOST_MCNUCL_H
#define SMC_HOST_MCNUCL_H

#include <cstdint>
#include <string>
#include <vector>

#include "Box2D.h"
#include "ParameterReader.h"
#include "smc.h"

class MCnucl {
 public:
  explicit MCnucl(ParameterReader* paraRdr);
  ~MCnucl();
  MCnucl(const MCnucl&) = delete;
  MCnucl& operator=(const MCnucl&) = delete;

  double lastCx1 = 0, lastPh1 = 0, lastCx2 = 0, lastPh2 = 0;

  double getTA1(int x, int y) { return ta_[(size_t)x * ny_ + y]; }
  double getTA2(int x, int y) { return tb_[(size_t)x * ny_ + y]; }
  double getRho(int i, int x, int y) { (void)i; return rho_[(size_t)x * ny_ + y]; }
  // pt < 0: the pT-integrated density; pT-resolved densities are not built
  // (the program stops with a message, as for the other options off this path)
  double getRho(int i, int x, int y, int pt);
  void setRho(int i, int x, int y, double val) { (void)i; rho_[(size_t)x * ny_ + y] = val; }
  int getNcoll() { return ncoll_; }
  int getNpart1() { return npart1_; }
  int getNpart2() { return npart2_; }
  double getdNdy() { return dndy_ * dx_ * dy_; }

  void generateNuclei(double b);
  void deleteNucleus();
  void setDensity(int iy, int ipt);
  void calculateThickness();
  int getBinaryCollision();
  int CentralityCut();
  void setCentralityCut(int Nmin, int Nmax) { npmax_ = Nmax; npmin_ = Nmin; }
  void makeTable();
  // MCnucl.cpp:967-1025, the dN/dy d2pT table: not built (stops with a message)
  void makeTable(double ptmin, double dpt, int iPtmax);
  // MCnucl.cpp:1027-1048: rows (y, T_A, T_B, dNdyTable[iy][i][j]) for
  // i, j >= 1 on this context's table grid (tmax points of dT), appended
  // to filename; dNdyTable is the caller's [iy][tmax][tmax] array
  void dumpdNdyTable4Col(std::string filename, double*** dNdyTable, const int iy);
  // MCnucl.cpp:1050-1076, the pT-resolved table: not built (stops with a message)
  void dumpdNdydptTable5Col(std::string filename, double**** dNdydptTable, const int iy);
  // MCnucl.cpp:1303-1323: the participant bounding boxes (projectile, then
  // target, in wound order) and the collision pairs' boxes (a CollisionPair
  // never sets its box: all edges 0), and their union
  Box2D getHotSpots(std::vector<Box2D>& hotSpots);
  // MCnucl.cpp:1119-1174: move the sources by minus the density's centre of
  // mass; rotateGrid also turns them by the order-n angle.  Call setDensity
  // again for the moved profile.
  void recenterGrid(int iy, int n = 2);
  void rotateGrid(int iy, int n = 2);

  // MCnucl.cpp:1177-1220: collision tables of the current event (formats of
  // the reference; quarks at their nucleon's centre)
  void dumpBinaryTable(std::string filename);
  void dumpparticipantTable(std::string filename);
  // MCnucl.cpp:1222-1260: spectators of the current event (count, table)
  int getSpectators();
  void dumpSpectatorsTable(int event);
  // MCnucl.cpp:481-614: binary-collision and spectator densities
  void calculate_rho_binary();
  double get_rho_binary(int x, int y) { return rho_binary_[(size_t)x * ny_ + y]; }
  void calculate_spectator_density();
  double get_spectator_density(int nucleus_id, int x, int y) {
    return (nucleus_id == 1 ? spec1_ : spec2_)[(size_t)x * ny_ + y];
  }

  // the event's nucleons in collision order (x, y in fm; wound rank or -1)
  void getNucleons(std::vector<double>& xy_proj, std::vector<int>& wound_proj, std::vector<double>& xy_targ,
                   std::vector<int>& wound_targ);
  smc_ctx* context() { return ctx_; }

 private:
  smc_ctx* ctx_ = nullptr;
  smc_params p_;
  uint64_t seed_ = 0, counter_ = 0;
  int nx_ = 0, ny_ = 0;
  double dx_ = 0, dy_ = 0;
  double box_half_ = 0;  // half the participant box: 4 max(width, quark_width) (Particle.cpp:16-21, 94-99)
  int npart1_ = 0, npart2_ = 0, ncoll_ = 0, npmin_ = 0, npmax_ = 0;
  double dndy_ = 0;
  bool have_event_ = false;
  std::vector<double> ta_, tb_, rho_, rho_binary_, spec1_, spec2_;
};

#endif
