// profile_3d.cpp -- facade of scripts/generate_3d_profiles/profile_3d.cpp:
// set-up and file output on the host, sampling and smearing through the
// C-ABI (include/smc_p3d.h) on the GPU.
#include "profile_3d.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <ctime>
#include <fstream>
#include <iomanip>
#include <iostream>

namespace {

// events of this path: set_variables is event 0 of the seed's stream, single
// draws of sample_eta_distribution_from_array take events 2^40 + n
constexpr uint64_t kSingleDrawEvent0 = 1ull << 40;

}  // namespace

void profile_3d::check(int rc, const char* where) {
  if (rc == SMC_OK) return;
  cout << "profile_3d::" << where << ": " << smc_p3d_last_error(h) << " (status " << rc << ")" << endl;
  exit(1);
}

// profile_3d.cpp:17-110
profile_3d::profile_3d(vector<participant_info> part_list_in, vector<participant_info> binary_list_in,
                       int grid_nx_in, int grid_ny_in, int grid_neta_in, double grid_dx_in, double grid_dy_in,
                       double grid_deta_in, double ecm_in, int random_flag_in)
    : h(nullptr), h_draw(nullptr) {
  ecm = ecm_in;
  grid_nx = grid_nx_in;
  grid_ny = grid_ny_in;
  grid_neta = grid_neta_in;
  grid_dx = grid_dx_in;
  grid_dy = grid_dy_in;
  grid_deta = grid_deta_in;
  random_flag = random_flag_in;
  smc_p3d_params p{grid_nx, grid_ny, grid_neta, grid_dx, grid_dy, grid_deta, ecm, random_flag};
  const int rc = smc_p3d_create(&h, &p);
  if (rc != SMC_OK) {
    cout << "profile_3d: invalid set-up (status " << rc << "): grid sizes, spacings, ecm > 2 GeV and "
         << "random_flag in 0..3 are required" << endl;
    exit(1);
  }
  smc_p3d_derived d{};
  check(smc_p3d_get_derived(h, &d, nullptr, nullptr, nullptr), "profile_3d");
  y_beam = d.y_beam;
  sigma_inelastic = d.sigma_inelastic;

  cout << "ecm = " << ecm << " A GeV, "
       << "y_beam = " << y_beam << ", "
       << "sigmaNN = " << sigma_inelastic * 10. << " mb. " << endl;

  for (size_t i = 0; i < part_list_in.size(); i++) {
    participant_info t{};
    t.x = part_list_in[i].x;
    t.y = part_list_in[i].y;
    t.id = part_list_in[i].id;
    participant_list.push_back(t);
  }
  for (size_t i = 0; i < binary_list_in.size(); i++) {
    participant_info t{};
    t.x = binary_list_in[i].x;
    t.y = binary_list_in[i].y;
    t.id = -1;
    binary_collision_list.push_back(t);
  }
  rho_part.assign((size_t)grid_neta * grid_nx * grid_ny, 0.f);
  rho_binary.assign((size_t)grid_neta * grid_nx * grid_ny, 0.f);

  seed = (uint64_t)time(NULL);
  single_draws = 0;
  cout << "random seed: " << seed << endl;

  set_eta_distribution();
  set_variables();
}

profile_3d::~profile_3d() {
  if (h) smc_p3d_destroy(h);
  if (h_draw) smc_p3d_destroy(h_draw);
}

// profile_3d.cpp:133-155 (tabulated by the library at construction)
void profile_3d::set_eta_distribution() {
  eta_distribution_array_eta.assign(SMC_P3D_TABLE, 0.0);
  eta_distribution_array_projectile.assign(SMC_P3D_TABLE, 0.0);
  eta_distribution_array_target.assign(SMC_P3D_TABLE, 0.0);
  smc_p3d_derived d{};
  check(smc_p3d_get_derived(h, &d, eta_distribution_array_eta.data(), eta_distribution_array_projectile.data(),
                            eta_distribution_array_target.data()),
        "set_eta_distribution");
  eta_distribution_proj_max = d.proj_max;
  eta_distribution_targ_max = d.targ_max;
}

// profile_3d.cpp:157-191: one draw on the device per call, from the table
// whatever random_flag is (a handle with random_flag 1, made on first use)
double profile_3d::sample_eta_distribution_from_array(int participant_id) {
  if (!h_draw) {
    smc_p3d_params p{grid_nx, grid_ny, grid_neta, grid_dx, grid_dy, grid_deta, ecm, 1};
    check(smc_p3d_create(&h_draw, &p), "sample_eta_distribution_from_array");
  }
  const double xy[2] = {0.0, 0.0};
  const int id[1] = {participant_id};
  const int off1[2] = {0, 1}, off0[2] = {0, 0};
  double src[6];
  const int rc = smc_p3d_sample(h_draw, 1, xy, id, off1, nullptr, off0, seed, kSingleDrawEvent0 + single_draws++,
                                src, nullptr);
  if (rc != SMC_OK) {
    cout << "profile_3d::sample_eta_distribution_from_array: " << smc_p3d_last_error(h_draw) << endl;
    exit(1);
  }
  return src[2];
}

// profile_3d.cpp:194-300: eta and widths of every source, on the device
void profile_3d::set_variables() {
  const size_t np = participant_list.size(), nb = binary_collision_list.size();
  vector<double> pxy(2 * np), bxy(2 * nb), ps(6 * np), bs(6 * nb);
  vector<int> pid(np);
  for (size_t i = 0; i < np; i++) {
    pxy[2 * i] = participant_list[i].x;
    pxy[2 * i + 1] = participant_list[i].y;
    pid[i] = participant_list[i].id;
  }
  for (size_t i = 0; i < nb; i++) {
    bxy[2 * i] = binary_collision_list[i].x;
    bxy[2 * i + 1] = binary_collision_list[i].y;
  }
  const int poff[2] = {0, (int)np}, boff[2] = {0, (int)nb};
  check(smc_p3d_sample(h, 1, pxy.data(), pid.data(), poff, bxy.data(), boff, seed, 0, ps.data(), bs.data()),
        "set_variables");
  auto fill = [](vector<participant_info>& l, const vector<double>& s) {
    for (size_t i = 0; i < l.size(); i++) {
      l[i].eta = s[6 * i + 2];
      l[i].sigma_x = s[6 * i + 3];
      l[i].sigma_y = s[6 * i + 4];
      l[i].sigma_eta = s[6 * i + 5];
    }
  };
  fill(participant_list, ps);
  fill(binary_collision_list, bs);
}

void profile_3d::smear(const vector<participant_info>& list, vector<float>& rho) {
  vector<double> src(6 * list.size());
  for (size_t i = 0; i < list.size(); i++) {
    const participant_info& p = list[i];
    const double r[6] = {p.x, p.y, p.eta, p.sigma_x, p.sigma_y, p.sigma_eta};
    for (int k = 0; k < 6; k++) src[6 * i + k] = r[k];
  }
  const int off[2] = {0, (int)list.size()};
  vector<float> add(rho.size());
  check(smc_p3d_smear(h, 1, src.data(), off, add.data()), "generate_3d_profile");
  for (size_t i = 0; i < rho.size(); i++) rho[i] += add[i];  // the reference accumulates (+=)
}

// profile_3d.cpp:302-344
void profile_3d::generate_3d_profile() { smear(participant_list, rho_part); }

// profile_3d.cpp:346-400
void profile_3d::generate_3d_profile_binary() { smear(binary_collision_list, rho_binary); }

// profile_3d.cpp:402-444: one line per (eta, x), ny values each
void profile_3d::output_3d_entropy_profile(string filename, double alpha) {
  ofstream of(filename.c_str());
  for (int i = 0; i < grid_neta; i++) {
    for (int j = 0; j < grid_nx; j++) {
      for (int k = 0; k < grid_ny; k++) {
        const size_t c = ((size_t)i * grid_nx + j) * grid_ny + k;
        double sd_local = ((double)rho_part[c] * (1. - alpha) / 2. + (double)rho_binary[c] * alpha);
        of << scientific << setw(18) << setprecision(8) << sd_local << "   ";
      }
      of << endl;
    }
  }
  of.close();
}

void profile_3d::output_3d_rhob_profile(string filename) {
  ofstream of(filename.c_str());
  for (int i = 0; i < grid_neta; i++) {
    for (int j = 0; j < grid_nx; j++) {
      for (int k = 0; k < grid_ny; k++) {
        double rhob_local = rho_part[((size_t)i * grid_nx + j) * grid_ny + k];
        of << scientific << setw(18) << setprecision(8) << rhob_local << "   ";
      }
      of << endl;
    }
  }
  of.close();
}

// profile_3d.cpp:402-434 (the reference's helper; the device sampler runs
// the same search)
long profile_3d::binarySearch(vector<double>* A, double value, bool skip_out_of_range) {
  int length = A->size();
  int idx_i = 0, idx_f = length - 1, idx;
  if (value > (*A)[idx_f]) {
    if (skip_out_of_range) return -1;
    cout << "binarySearch: desired value is too large, exceeding the end of the table." << endl;
    exit(-1);
  }
  if (value < (*A)[idx_i]) {
    if (skip_out_of_range) return -1;
    cout << "binarySearch: desired value is too small, exceeding the beginning of table." << endl;
    exit(-1);
  }
  idx = (int)floor((idx_f + idx_i) / 2.);
  while ((idx_f - idx_i) > 1) {
    if ((*A)[idx] < value) idx_i = idx;
    else idx_f = idx;
    idx = (int)floor((idx_f + idx_i) / 2.);
  }
  return idx_i;
}
