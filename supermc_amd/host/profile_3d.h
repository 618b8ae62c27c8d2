// profile_3d.h -- the 3D (x, y, eta) profile class of
// scripts/generate_3d_profiles (profile_3d.h) on the device path of
// include/smc_p3d.h.  Its public constructor, methods and participant_info
// record are the reference's interface; the state behind them is this
// facade's own: a handle on the library, the eta tables it returns, and the
// two grids.  GSL's time-seeded stream becomes Philox keyed by the same
// time-derived seed (printed as "random seed: ...").
#ifndef profile_3d_H
#define profile_3d_H

#include <cstdint>
#include <string>
#include <vector>

#include "smc_p3d.h"

using namespace std;

struct participant_info {
  int id;
  double x, y, eta;
  double sigma_x, sigma_y, sigma_eta;
};

class profile_3d {
 public:
  profile_3d(vector<participant_info> part_list_in, vector<participant_info> binary_list_in, int grid_nx_in,
             int grid_ny_in, int grid_neta_in, double grid_dx_in, double grid_dy_in, double grid_deta_in,
             double ecm_in, int random_flag_in);
  ~profile_3d();
  profile_3d(const profile_3d&) = delete;
  profile_3d& operator=(const profile_3d&) = delete;

  void set_variables();
  void set_eta_distribution();
  void generate_3d_profile();
  void generate_3d_profile_binary();
  void output_3d_entropy_profile(string filename, double alpha);
  void output_3d_rhob_profile(string filename);
  double sample_eta_distribution_from_array(int participant_id);

  long binarySearch(vector<double>* A, double value, bool skip_out_of_range);

  // the sources after set_variables (eta and widths filled in)
  const vector<participant_info>& participants() const { return sources_[0]; }
  const vector<participant_info>& binary_collisions() const { return sources_[1]; }
  uint64_t random_seed() const { return seed_; }

 private:
  // library handles: the profile's own, and one with random_flag 1 for the
  // single table draws of sample_eta_distribution_from_array
  smc_p3d* dev_ = nullptr;
  smc_p3d* dev_draw_ = nullptr;
  smc_p3d_params grid_{};
  smc_p3d_derived derived_{};  // y_beam, sigma_inel, table maxima
  // eta, projectile and target columns of the 1000-point table
  vector<double> table_[3];
  // sources: [0] participants, [1] binary collisions
  vector<participant_info> sources_[2];
  // densities of both source lists, [neta][nx][ny]
  vector<float> density_[2];
  uint64_t seed_ = 0;
  uint64_t draws_ = 0;  // single draws made so far

  void require(int rc, smc_p3d* h, const char* where) const;
  void add_smeared(int which);
  void write_grid(const string& filename, double w_part, double w_bin) const;
};

#endif
