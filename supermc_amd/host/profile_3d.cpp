// profile_3d.cpp -- the 3D profile facade: set-up and file output on the
// host, eta / width sampling and the separable 3D smearing through the
// C-ABI (include/smc_p3d.h) on the GPU.  The methods keep the meaning of
// scripts/generate_3d_profiles/profile_3d.cpp (cited per method); the code
// is organised around a library handle and per-list arrays.
#include "profile_3d.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <ctime>
#include <iostream>

namespace {

// counter ranges of the seed's stream: the sources of set_variables are
// event 0, single draws of sample_eta_distribution_from_array start here
constexpr uint64_t kFirstSingleDraw = 1ull << 40;

// the six per-source numbers the library reads and writes
void pack(const participant_info& p, double* r) {
  r[0] = p.x;
  r[1] = p.y;
  r[2] = p.eta;
  r[3] = p.sigma_x;
  r[4] = p.sigma_y;
  r[5] = p.sigma_eta;
}

}  // namespace

void profile_3d::require(int rc, smc_p3d* h, const char* where) const {
  if (rc == SMC_OK) return;
  std::printf("profile_3d::%s failed: %s (status %d)\n", where, h ? smc_p3d_last_error(h) : "no handle", rc);
  std::exit(1);
}

// profile_3d.cpp:17-110: grid and beam set-up, the sources' transverse
// positions (participants keep their nucleus id, collisions get -1), the
// eta tables, then eta and widths of every source.
profile_3d::profile_3d(vector<participant_info> part_list_in, vector<participant_info> binary_list_in,
                       int grid_nx_in, int grid_ny_in, int grid_neta_in, double grid_dx_in, double grid_dy_in,
                       double grid_deta_in, double ecm_in, int random_flag_in) {
  grid_ = smc_p3d_params{grid_nx_in, grid_ny_in, grid_neta_in, grid_dx_in, grid_dy_in, grid_deta_in, ecm_in,
                         random_flag_in};
  const int rc = smc_p3d_create(&dev_, &grid_);
  if (rc != SMC_OK) {
    std::printf("profile_3d: set-up refused (status %d); it needs positive grid sizes and spacings, ecm above "
                "2 GeV and random_flag 0..3\n", rc);
    std::exit(1);
  }
  require(smc_p3d_get_derived(dev_, &derived_, nullptr, nullptr, nullptr), dev_, "profile_3d");
  // the reference's first console line (tools parse it)
  std::cout << "ecm = " << ecm_in << " A GeV, y_beam = " << derived_.y_beam
            << ", sigmaNN = " << derived_.sigma_inelastic * 10. << " mb. " << std::endl;

  for (const participant_info& p : part_list_in) sources_[0].push_back(participant_info{p.id, p.x, p.y, 0, 0, 0, 0});
  for (const participant_info& p : binary_list_in) sources_[1].push_back(participant_info{-1, p.x, p.y, 0, 0, 0, 0});
  const size_t cells = (size_t)grid_.grid_neta * grid_.grid_nx * grid_.grid_ny;
  density_[0].assign(cells, 0.f);
  density_[1].assign(cells, 0.f);

  seed_ = (uint64_t)std::time(nullptr);
  std::cout << "random seed: " << seed_ << std::endl;

  set_eta_distribution();
  set_variables();
}

profile_3d::~profile_3d() {
  if (dev_) smc_p3d_destroy(dev_);
  if (dev_draw_) smc_p3d_destroy(dev_draw_);
}

// profile_3d.cpp:133-155: the 1000-point eta table and its two columns
// (tabulated by the library when the handle is made)
void profile_3d::set_eta_distribution() {
  for (vector<double>& col : table_) col.assign(SMC_P3D_TABLE, 0.0);
  require(smc_p3d_get_derived(dev_, &derived_, table_[0].data(), table_[1].data(), table_[2].data()), dev_,
          "set_eta_distribution");
}

// profile_3d.cpp:157-191: one draw from the table for a source of nucleus
// participant_id, on the device, whatever the profile's random_flag
double profile_3d::sample_eta_distribution_from_array(int participant_id) {
  if (!dev_draw_) {
    smc_p3d_params one = grid_;
    one.random_flag = 1;
    require(smc_p3d_create(&dev_draw_, &one), dev_draw_, "sample_eta_distribution_from_array");
  }
  const double xy[2] = {0.0, 0.0};
  const int id = participant_id;
  const int one_source[2] = {0, 1}, no_source[2] = {0, 0};
  double out[6];
  require(smc_p3d_sample(dev_draw_, 1, xy, &id, one_source, nullptr, no_source, seed_, kFirstSingleDraw + draws_++,
                         out, nullptr),
          dev_draw_, "sample_eta_distribution_from_array");
  return out[2];
}

// profile_3d.cpp:194-300: eta and the three widths of every source, drawn
// on the device from the seed's stream
void profile_3d::set_variables() {
  vector<double> xy[2], out[2];
  vector<int> ids(sources_[0].size());
  int off[2][2];
  for (int w = 0; w < 2; ++w) {
    const vector<participant_info>& l = sources_[w];
    xy[w].resize(2 * l.size());
    out[w].resize(6 * l.size());
    for (size_t i = 0; i < l.size(); ++i) {
      xy[w][2 * i] = l[i].x;
      xy[w][2 * i + 1] = l[i].y;
    }
    off[w][0] = 0;
    off[w][1] = (int)l.size();
  }
  for (size_t i = 0; i < ids.size(); ++i) ids[i] = sources_[0][i].id;
  require(smc_p3d_sample(dev_, 1, xy[0].data(), ids.data(), off[0], xy[1].data(), off[1], seed_, 0, out[0].data(),
                         out[1].data()),
          dev_, "set_variables");
  for (int w = 0; w < 2; ++w)
    for (size_t i = 0; i < sources_[w].size(); ++i) {
      const double* r = &out[w][6 * i];
      sources_[w][i].eta = r[2];
      sources_[w][i].sigma_x = r[3];
      sources_[w][i].sigma_y = r[4];
      sources_[w][i].sigma_eta = r[5];
    }
}

// smears source list `which` and adds it to that list's density (the
// reference's generate_* add into their arrays)
void profile_3d::add_smeared(int which) {
  const vector<participant_info>& l = sources_[which];
  vector<double> src(6 * l.size());
  for (size_t i = 0; i < l.size(); ++i) pack(l[i], &src[6 * i]);
  const int off[2] = {0, (int)l.size()};
  vector<float> grid(density_[which].size());
  require(smc_p3d_smear(dev_, 1, src.data(), off, grid.data()), dev_, "generate_3d_profile");
  std::transform(grid.begin(), grid.end(), density_[which].begin(), density_[which].begin(),
                 [](float a, float b) { return a + b; });
}

// profile_3d.cpp:302-344
void profile_3d::generate_3d_profile() { add_smeared(0); }

// profile_3d.cpp:346-400
void profile_3d::generate_3d_profile_binary() { add_smeared(1); }

// The files of profile_3d.cpp:402-434: w_part * participant density +
// w_bin * collision density per cell, one line per (eta slice, x row) with
// the row's ny values, each as %18.8e followed by three spaces (the
// reference's iostream scientific, width 18, precision 8).
void profile_3d::write_grid(const string& filename, double w_part, double w_bin) const {
  std::FILE* f = std::fopen(filename.c_str(), "w");
  if (!f) {
    std::printf("profile_3d: cannot write %s\n", filename.c_str());
    std::exit(1);
  }
  const size_t ny = (size_t)grid_.grid_ny, rows = (size_t)grid_.grid_neta * grid_.grid_nx;
  for (size_t r = 0; r < rows; ++r) {
    for (size_t k = 0; k < ny; ++k) {
      const size_t c = r * ny + k;
      const double v = w_part * (double)density_[0][c] + (w_bin != 0.0 ? w_bin * (double)density_[1][c] : 0.0);
      std::fprintf(f, "%18.8e   ", v);
    }
    std::fputc('\n', f);
  }
  std::fclose(f);
}

// entropy density (1 - alpha)/2 participants + alpha collisions
void profile_3d::output_3d_entropy_profile(string filename, double alpha) {
  write_grid(filename, (1. - alpha) / 2., alpha);
}

// net-baryon density: the participants' density
void profile_3d::output_3d_rhob_profile(string filename) { write_grid(filename, 1.0, 0.0); }

// The table search of profile_3d.cpp:436-470, on a table sorted upwards:
// the largest index i with A[i] < value (0 when value equals A[0]); a value
// outside [A[0], A[end]] gives -1 with skip_out_of_range, else the program
// stops (exit -1).
long profile_3d::binarySearch(vector<double>* A, double value, bool skip_out_of_range) {
  const vector<double>& t = *A;
  if (t.empty() || value < t.front() || value > t.back()) {
    if (skip_out_of_range) return -1;
    std::printf("profile_3d::binarySearch: %g lies outside the table [%g, %g]\n", value,
                t.empty() ? 0.0 : t.front(), t.empty() ? 0.0 : t.back());
    std::exit(-1);
  }
  const long first_not_below = std::lower_bound(t.begin(), t.end(), value) - t.begin();
  return std::max(0L, first_not_below - 1);
}
