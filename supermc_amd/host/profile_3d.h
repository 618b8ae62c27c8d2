// profile_3d.h -- the reference's profile_3d class
// (scripts/generate_3d_profiles/profile_3d.h) over the device path of
// include/smc_p3d.h: same constructor, methods and output files.  The grids
// are computed on the GPU; GSL's time-seeded stream is replaced by Philox
// keyed by the same time-derived seed (printed as the reference prints it).
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
 private:
  vector<participant_info> participant_list;
  vector<participant_info> binary_collision_list;

  int grid_nx, grid_ny, grid_neta;
  double grid_dx, grid_dy, grid_deta;
  vector<float> rho_part, rho_binary;  // [neta][nx][ny]

  double eta_distribution_proj_max, eta_distribution_targ_max;
  vector<double> eta_distribution_array_eta;
  vector<double> eta_distribution_array_projectile;
  vector<double> eta_distribution_array_target;

  double ecm, y_beam, sigma_inelastic;

  int random_flag;
  uint64_t seed;
  uint64_t single_draws;  // draws made by sample_eta_distribution_from_array
  smc_p3d* h;
  smc_p3d* h_draw;  // random_flag 1 handle for single table draws

  void check(int rc, const char* where);
  void smear(const vector<participant_info>& list, vector<float>& rho);

 public:
  profile_3d(vector<participant_info> part_list_in, vector<participant_info> binary_list_in, int grid_nx_in,
             int grid_ny_in, int grid_neta_in, double grid_dx_in, double grid_dy_in, double grid_deta_in,
             double ecm_in, int random_flag_in);
  ~profile_3d();

  void set_variables();
  void set_eta_distribution();
  void generate_3d_profile();
  void generate_3d_profile_binary();
  void output_3d_entropy_profile(string filename, double alpha);
  void output_3d_rhob_profile(string filename);
  double sample_eta_distribution_from_array(int participant_id);

  long binarySearch(vector<double>* A, double value, bool skip_out_of_range);

  // the sources after set_variables (eta and widths filled in)
  const vector<participant_info>& participants() const { return participant_list; }
  const vector<participant_info>& binary_collisions() const { return binary_collision_list; }
  uint64_t random_seed() const { return seed; }
};

#endif
