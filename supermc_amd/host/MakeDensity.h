// MakeDensity -- the reference's event-loop driver (MakeDensity.h:43-62)
// over libsmc.  Events run in device batches of `gpu_batch_events`
// (extension key, default 16384) with everything resident on the GPU; the
// per-event lines and files are written in the reference's formats and
// order.
//
//  generateEccTable(n)      operation 9 (MakeDensity.cpp:2108-2240):
//                           data/sn_ecc_eccp_<order>.dat (and en_, use_ed)
//  generate_profile_ebe(n)  operation 1 (MakeDensity.cpp:502-730), pT-
//                           integrated: data/sd_event_<i>_block.dat / _4col,
//                           and per event the collision tables of
//                           MCnucl::dumpBinaryTable (data/binary.dat,
//                           wounded.data, quarks.data, nucl1/2.data)
//  generate_profile_ebe_Jet(n)  operation 2 (MakeDensity.cpp:148-500),
//                           pT-integrated: per event the participant,
//                           spectator and binary-collision tables, the
//                           eccentricity lines (sn_ecc_eccp_<n>_event_<i>),
//                           the profiles, and with output_rho_binary /
//                           output_TA / output_rhob / output_spectator_density
//                           the binary density, T_A and T_B, T_A + T_B and the
//                           spectator densities.  Events failing the dS/dy
//                           cut write nothing (the reference writes their
//                           tables before the cut).
//  generate_profile_average(n)  operation 3 (MakeDensity.cpp:736-1650),
//                           pT-integrated: averages over n accepted events
//                           of the profiles recentred and turned to each
//                           order's angle, data/sdAvg_order_<n>_block.dat /
//                           _4col (edAvg_ with use_ed; sdAvg_RP_ with
//                           generate_reaction_plane_avg_profile), and with
//                           output_TATB / output_rho_binary / output_TA /
//                           output_spectator_density the averages of T_A T_B,
//                           the binary density, T_A and T_B (also
//                           _RP) and the spectator densities of the moved
//                           sources (TATB_fromSd_order_<n>, ...).  The dS/dy
//                           cut is applied once per event, on the unturned
//                           profile.
//
// Random numbers: Philox stream randomSeed; event i of a run is counter i,
// so a run is reproducible and independent of gpu_batch_events.
#ifndef SMC_HOST_MAKEDENSITY_H
#define SMC_HOST_MAKEDENSITY_H

#include <cstdint>
#include <string>
#include <vector>

#include "ParameterReader.h"
#include "smc.h"

class MakeDensity {
 public:
  explicit MakeDensity(ParameterReader* paraRdr);
  ~MakeDensity();
  MakeDensity(const MakeDensity&) = delete;
  MakeDensity& operator=(const MakeDensity&) = delete;

  void generateEccTable(int nevent);
  void generate_profile_ebe(int nevent);
  void generate_profile_average(int nevent);
  void generate_profile_ebe_Jet(int nevent);

  // Writes the eccentricity lines of one event record (the SMC_EV_* layout
  // of smc.h) in the format of MakeDensity.cpp:2432-2500.
  void dumpEccentricities(const char* base_filename, const double* record, int from_order, int to_order,
                          double Npart_current, double Nbin_current, double b);
  void dumpDensityBlock(const char* filename, const float* s, double factor);
  void dumpDensity4Col(const char* filename, const float* s, double factor, double npart);
  void dumpDensityBlock(const char* filename, const double* s);
  void dumpDensity4Col(const char* filename, const double* s, double npart);

  // events accepted by the last generate call, and their records
  const std::vector<double>& records() const { return records_; }

 private:
  // one accepted event passes (status 0, and the dS/dy cut if enabled)
  bool accept(const double* rec) const;
  int next_batch(int want);  // runs one device batch, returns its size

  ParameterReader* rdr_;
  smc_ctx* ctx_ = nullptr;
  smc_params p_;
  uint64_t seed_ = 0, counter_ = 0;
  int batch_ = 16384;
  int nx_ = 0, ny_ = 0;
  double xmin_ = 0, ymin_ = 0;
  int cut_dsdy_ = 0;
  double cut_lo_ = 0, cut_hi_ = 0;
  bool deformed_ = false;
  std::vector<double> batch_rec_, records_;
  int64_t dropped_ = 0;
};

#endif
