#include "MakeDensity.h"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <iostream>

#include "smc_host.h"

using smc_host::check;

MakeDensity::MakeDensity(ParameterReader* paraRdr) : rdr_(paraRdr) {
  if (rdr_->exist("gpu_batch_events")) batch_ = std::max(1, (int)rdr_->getVal("gpu_batch_events"));
  p_ = smc_host::params_from_reader(*rdr_, batch_);
  seed_ = smc_host::seed_from_reader(*rdr_);
  check(nullptr, smc_create(&ctx_, &p_), "MakeDensity::MakeDensity");
  smc_derived d;
  check(ctx_, smc_get_derived(ctx_, &d), "MakeDensity::MakeDensity");
  nx_ = d.nx;
  ny_ = d.ny;
  xmin_ = d.xmin;
  ymin_ = d.ymin;
  // MakeDensity.cpp:2173-2181 (optional dS/dy cut)
  cut_dsdy_ = rdr_->exist("cutdSdy") ? (int)rdr_->getVal("cutdSdy") : 0;
  if (cut_dsdy_ == 1) {
    cut_lo_ = rdr_->getVal("cutdSdy_lowerBound");
    cut_hi_ = rdr_->getVal("cutdSdy_upperBound");
  }
  deformed_ = p_.proj_deformed == 1 || p_.targ_deformed == 1;
  // MC-KLN: the dN/dy table once, as MakeDensity's constructor does
  if (p_.which_mc_model == 1 && !p_.kln_direct) check(ctx_, smc_kln_make_table(ctx_, nullptr), "makeTable");
}

MakeDensity::~MakeDensity() { smc_destroy(ctx_); }

int MakeDensity::next_batch(int want) {
  const int n = std::min(want, batch_);
  batch_rec_.resize((size_t)n * SMC_EV_STRIDE);
  check(ctx_, smc_run(ctx_, seed_, counter_, n, batch_rec_.data()), "MakeDensity: event batch");
  counter_ += (uint64_t)n;
  return n;
}

bool MakeDensity::accept(const double* rec) const {
  if (rec[SMC_EV_STATUS] != 0.0) return false;  // no accepted configuration / capacity
  if (cut_dsdy_ == 1) {
    // gettotaldSdy: sum rho dx dy without finalFactor (MakeDensity.cpp:2557-2567)
    const double ff = p_.finalFactor;
    const double s = (ff != 0.0) ? rec[SMC_EV_TOTAL] / ff * p_.dx * p_.dy : 0.0;
    if (s < cut_lo_ || s > cut_hi_) return false;
  }
  return true;
}

void MakeDensity::dumpEccentricities(const char* base_filename, const double* rec, int from_order, int to_order,
                                     double Npart_current, double Nbin_current, double b) {
  const int max_order = 10;
  char buffer[512];
  const double dsdy = rec[SMC_EV_DSDY];
  auto tail = [&](std::ofstream& of) {
    of << std::setprecision(5) << std::setw(10) << Npart_current << std::setprecision(5) << std::setw(10)
       << Nbin_current << std::setprecision(8) << std::setw(16) << dsdy << std::setprecision(8) << std::setw(16)
       << b;
    if (deformed_)
      for (int k = 0; k < 4; ++k) of << std::setprecision(8) << std::setw(16) << 0.0;
    of << std::endl;
  };
  for (int order = from_order; order <= to_order; ++order) {
    std::snprintf(buffer, sizeof(buffer), base_filename, order);
    std::ofstream of(buffer, std::ios_base::app);
    of << std::setprecision(8) << std::setw(16) << rec[SMC_EV_EPS_RE + order] << std::setprecision(8)
       << std::setw(16) << rec[SMC_EV_EPS_IM + order] << std::setprecision(8) << std::setw(16)
       << rec[SMC_EV_EPSP_RE + order] << std::setprecision(8) << std::setw(16) << rec[SMC_EV_EPSP_IM + order]
       << std::setprecision(8) << std::setw(16) << rec[SMC_EV_RN + order];
    tail(of);
  }
  std::snprintf(buffer, sizeof(buffer), base_filename, max_order);
  std::ofstream of(buffer, std::ios_base::app);
  for (int order = 1; order < max_order; ++order)
    of << std::setprecision(8) << std::setw(16) << rec[SMC_EV_EPS_RE + order] << std::setprecision(8)
       << std::setw(16) << rec[SMC_EV_EPS_IM + order] << std::setprecision(8) << std::setw(16)
       << rec[SMC_EV_EPSP_RE + order] << std::setprecision(8) << std::setw(16) << rec[SMC_EV_EPSP_IM + order]
       << std::setprecision(8) << std::setw(16) << rec[SMC_EV_RN + order];
  tail(of);
}

void MakeDensity::generateEccTable(int nevent) {
  const int from = p_.ecc_from_order, to = p_.ecc_to_order;
  const bool use_sd = rdr_->getVal("use_sd") != 0, use_ed = rdr_->getVal("use_ed") != 0;
  records_.clear();
  int written = 0;
  while (written < nevent) {
    const int n = next_batch(nevent - written);
    for (int e = 0; e < n && written < nevent; ++e) {
      const double* rec = &batch_rec_[(size_t)e * SMC_EV_STRIDE];
      if (!accept(rec)) {
        ++dropped_;
        continue;
      }
      // setSd / setEd give the same profile here (MakeDensity.cpp:2526-2535)
      if (use_sd) dumpEccentricities("data/sn_ecc_eccp_%d.dat", rec, from, to, rec[SMC_EV_NPART],
                                     rec[SMC_EV_NCOLL], rec[SMC_EV_B]);
      if (use_ed) dumpEccentricities("data/en_ecc_eccp_%d.dat", rec, from, to, rec[SMC_EV_NPART],
                                     rec[SMC_EV_NCOLL], rec[SMC_EV_B]);
      records_.insert(records_.end(), rec, rec + SMC_EV_STRIDE);
      ++written;
    }
  }
}

namespace {
// MakeDensity.cpp:2595-2608
template <class T>
void write_block(const char* filename, const T* s, double factor, int nx, int ny) {
  std::ofstream of(filename, std::ios_base::out);
  for (int i = 0; i < nx; ++i) {
    for (int j = 0; j < ny; ++j) of << std::setprecision(12) << std::setw(22) << (double)s[(size_t)i * ny + j] * factor;
    of << std::endl;
  }
}

// MakeDensity.cpp:2573-2593 (one rapidity bin at y = 0)
template <class T>
void write_4col(const char* filename, const T* s, double factor, double npart, int nx, int ny, double xmin,
                double ymin, double dx, double dy) {
  std::ofstream of(filename, std::ios_base::out);
  of << "# <npart>= " << npart << " xmax= " << nx << " ymax= " << ny << std::endl;
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j) {
      const double x = xmin + i * dx, y = ymin + j * dy, rap = 0.0;
      of << std::setprecision(3) << std::setw(10) << rap << std::setprecision(3) << std::setw(10) << x
         << std::setprecision(3) << std::setw(10) << y << std::setprecision(12) << std::setw(22)
         << (double)s[(size_t)i * ny + j] * factor << std::endl;
    }
}
}  // namespace

void MakeDensity::dumpDensityBlock(const char* filename, const float* s, double factor) {
  write_block(filename, s, factor, nx_, ny_);
}

void MakeDensity::dumpDensity4Col(const char* filename, const float* s, double factor, double npart) {
  write_4col(filename, s, factor, npart, nx_, ny_, xmin_, ymin_, p_.dx, p_.dy);
}

void MakeDensity::dumpDensityBlock(const char* filename, const double* s) { write_block(filename, s, 1.0, nx_, ny_); }

void MakeDensity::dumpDensity4Col(const char* filename, const double* s, double npart) {
  write_4col(filename, s, 1.0, npart, nx_, ny_, xmin_, ymin_, p_.dx, p_.dy);
}

void MakeDensity::generate_profile_ebe(int nevent) {
  const bool use_sd = rdr_->getVal("use_sd") != 0, use_ed = rdr_->getVal("use_ed") != 0;
  const bool use_block = rdr_->getVal("use_block") != 0, use_4col = rdr_->getVal("use_4col") != 0;
  const size_t ncell = (size_t)nx_ * ny_;
  // the reference's grids are doubles: for the two-component and
  // sqrt(T_A T_B) models the files are written from the library's FP64
  // grids (its double expressions), MC-KLN from the FP32 grid
  const bool fp64 = p_.which_mc_model == 5 || p_.which_mc_model == 7;
  std::vector<float> grid;
  std::vector<double> grid64(fp64 ? ncell : 0);
  std::vector<double> dndy;
  records_.clear();
  int event = 1;
  char buffer[512];
  while (event <= nevent) {
    const int n = next_batch(nevent - event + 1);
    grid.resize(fp64 ? 0 : (size_t)n * ncell);
    dndy.resize(n);
    const smc_status st = smc_density(ctx_, fp64 ? nullptr : grid.data(), dndy.data());
    if (st != SMC_ERR_DNDY_ZERO) check(ctx_, st, "MakeDensity::generate_profile_ebe");
    for (int e = 0; e < n && event <= nevent; ++e) {
      const double* rec = &batch_rec_[(size_t)e * SMC_EV_STRIDE];
      if (!accept(rec)) {
        ++dropped_;
        continue;
      }
      // MakeDensity.cpp:597: the collision tables of every event
      smc_host::dump_binary_table(ctx_, p_, e, "data/binary.dat");
      if (fp64) check(ctx_, smc_density_f64(ctx_, e, 1, grid64.data()), "MakeDensity::generate_profile_ebe");
      const float* s = fp64 ? nullptr : &grid[(size_t)e * ncell];
      const char* kinds[2] = {"sd", "ed"};
      const bool on[2] = {use_sd, use_ed};
      for (int k = 0; k < 2; ++k) {
        if (!on[k]) continue;
        if (use_4col) {
          std::snprintf(buffer, sizeof(buffer), "data/%s_event_%d_4col.dat", kinds[k], event);
          if (fp64) write_4col(buffer, grid64.data(), p_.finalFactor, rec[SMC_EV_NPART], nx_, ny_, xmin_, ymin_,
                               p_.dx, p_.dy);
          else dumpDensity4Col(buffer, s, p_.finalFactor, rec[SMC_EV_NPART]);
        }
        if (use_block) {
          std::snprintf(buffer, sizeof(buffer), "data/%s_event_%d_block.dat", kinds[k], event);
          if (fp64) write_block(buffer, grid64.data(), p_.finalFactor, nx_, ny_);
          else dumpDensityBlock(buffer, s, p_.finalFactor);
        }
      }
      records_.insert(records_.end(), rec, rec + SMC_EV_STRIDE);
      ++event;
    }
  }
}

void MakeDensity::generate_profile_average(int nevent) {
  const int from = (int)rdr_->getVal("average_from_order"), to = (int)rdr_->getVal("average_to_order");
  const bool use_sd = rdr_->getVal("use_sd") != 0, use_ed = rdr_->getVal("use_ed") != 0;
  const bool use_block = rdr_->getVal("use_block") != 0, use_4col = rdr_->getVal("use_4col") != 0;
  const bool rp = rdr_->getVal("generate_reaction_plane_avg_profile") != 0;
  // the optional averages (MakeDensity.cpp:749-752)
  const bool out_tatb = rdr_->getVal("output_TATB") != 0, out_rb = rdr_->getVal("output_rho_binary") != 0;
  const bool out_ta = rdr_->getVal("output_TA") != 0, out_spec = rdr_->getVal("output_spectator_density") != 0;
  if (from < 1 || to < from || to > 9) {
    std::cerr << "generate_profile_average: average orders must satisfy 1 <= from <= to <= 9" << std::endl;
    std::exit(1);
  }
  const int nord = to - from + 1;
  const size_t ncell = (size_t)nx_ * ny_, n_all = nord * ncell;
  std::vector<double> sums(n_all, 0.0), rp_sums(rp ? n_all : 0, 0.0);
  // one array per optional average, in the order of smc_avg_extra
  enum { kTatb, kRb, kTa, kTb, kSa, kSb, kTatbRp, kRbRp, kTaRp, kTbRp, kN };
  const bool on_ex[kN] = {out_tatb, out_rb, out_ta, out_ta, out_spec, out_spec,
                          out_tatb && rp, out_rb && rp, out_ta && rp, out_ta && rp};
  std::vector<double> ex[kN];
  for (int k = 0; k < kN; ++k) ex[k].assign(on_ex[k] ? n_all : 0, 0.0);
  auto ptr = [&](int k) { return on_ex[k] ? ex[k].data() : nullptr; };
  const smc_avg_extra extra{ptr(kTatb), ptr(kRb), ptr(kTa),    ptr(kTb),  ptr(kSa),
                            ptr(kSb),   ptr(kTatbRp), ptr(kRbRp), ptr(kTaRp), ptr(kTbRp)};
  std::vector<unsigned char> acc;
  records_.clear();
  int done = 0;
  double npart_last = 0.0;
  while (done < nevent) {
    const int n = next_batch(nevent - done);
    acc.assign(n, 0);
    for (int e = 0; e < n && done < nevent; ++e) {
      const double* rec = &batch_rec_[(size_t)e * SMC_EV_STRIDE];
      if (!accept(rec)) {
        ++dropped_;
        continue;
      }
      acc[e] = 1;
      npart_last = rec[SMC_EV_NPART];
      records_.insert(records_.end(), rec, rec + SMC_EV_STRIDE);
      ++done;
    }
    check(ctx_,
          smc_average_profiles_extra(ctx_, from, to, acc.data(), sums.data(), rp ? rp_sums.data() : nullptr, &extra),
          "MakeDensity::generate_profile_average");
  }
  // the reference's running average over accepted events
  for (double& v : sums) v /= nevent;
  for (double& v : rp_sums) v /= nevent;
  for (std::vector<double>& e : ex)
    for (double& v : e) v /= nevent;
  char buffer[512];
  auto dump = [&](const char* stem, const double* g, int order) {
    if (use_4col) {
      std::snprintf(buffer, sizeof(buffer), "data/%s_order_%d_4col.dat", stem, order);
      dumpDensity4Col(buffer, g, npart_last);
    }
    if (use_block) {
      std::snprintf(buffer, sizeof(buffer), "data/%s_order_%d_block.dat", stem, order);
      dumpDensityBlock(buffer, g);
    }
  };
  // entropy and energy profiles (setEd equals setSd, MakeDensity.cpp:2524-2533),
  // then TA*TB, rho_binary, TA / TB (also after the recentring) and the
  // spectator densities (MakeDensity.cpp:1650-1900 file names)
  const char* from_tag[2] = {"Sd", "Ed"};
  const char* avg_tag[2] = {"sd", "ed"};
  const bool on[2] = {use_sd, use_ed};
  char stem[256];
  for (int o = 0; o < nord; ++o) {
    const int order = from + o;
    for (int k = 0; k < 2; ++k) {
      if (!on[k]) continue;
      for (int r = 0; r < (rp ? 2 : 1); ++r) {
        std::snprintf(stem, sizeof(stem), "%sAvg%s", avg_tag[k], r ? "_RP" : "");
        dump(stem, (r ? rp_sums.data() : sums.data()) + o * ncell, order);
      }
      struct { int slot, rp_slot; const char* fmt; } items[] = {
          {kTatb, kTatbRp, "TATB_from%s%s"},
          {kRb, kRbRp, "rho_binary_from%s%s"},
          {kTa, kTaRp, "nuclear_thickness_TA_from%s%s"},
          {kTb, kTbRp, "nuclear_thickness_TB_from%s%s"},
          {kSa, -1, "spectator_density_A_from%s%s"},
          {kSb, -1, "spectator_density_B_from%s%s"}};
      for (const auto& it : items) {
        if (!on_ex[it.slot]) continue;
        std::snprintf(stem, sizeof(stem), it.fmt, from_tag[k], "");
        dump(stem, ex[it.slot].data() + o * ncell, order);
        if (it.rp_slot >= 0 && on_ex[it.rp_slot]) {
          std::snprintf(stem, sizeof(stem), it.fmt, from_tag[k], "_RP");
          dump(stem, ex[it.rp_slot].data() + o * ncell, order);
        }
      }
    }
  }
}

void MakeDensity::generate_profile_ebe_Jet(int nevent) {
  // MakeDensity.cpp:148-500 (pT-integrated, one rapidity bin)
  const int from = p_.ecc_from_order, to = p_.ecc_to_order;
  const bool use_sd = rdr_->getVal("use_sd") != 0, use_ed = rdr_->getVal("use_ed") != 0;
  const bool use_block = rdr_->getVal("use_block") != 0, use_4col = rdr_->getVal("use_4col") != 0;
  const bool out_rho_binary = rdr_->getVal("output_rho_binary") != 0, out_ta = rdr_->getVal("output_TA") != 0;
  const bool out_rhob = rdr_->getVal("output_rhob") != 0;
  const bool out_spec = rdr_->getVal("output_spectator_density") != 0;
  const size_t ncell = (size_t)nx_ * ny_;
  const int chunk = std::min(batch_, 512);  // per-event grids come back to the host
  std::vector<float> grid, ta, tb, rho, spa, spb;
  std::vector<double> dndy;
  records_.clear();
  int event = 1;
  char buffer[512];
  auto grids = [&](const char* stem, const float* g, double factor, double npart) {
    if (use_4col) {
      std::snprintf(buffer, sizeof(buffer), "data/%s_event_%d_4col.dat", stem, event);
      dumpDensity4Col(buffer, g, factor, npart);
    }
    if (use_block) {
      std::snprintf(buffer, sizeof(buffer), "data/%s_event_%d_block.dat", stem, event);
      dumpDensityBlock(buffer, g, factor);
    }
  };
  while (event <= nevent) {
    const int n = next_batch(std::min(chunk, nevent - event + 1));
    grid.resize((size_t)n * ncell);
    dndy.resize(n);
    const smc_status st = smc_density(ctx_, grid.data(), dndy.data());
    if (st != SMC_ERR_DNDY_ZERO) check(ctx_, st, "MakeDensity::generate_profile_ebe_Jet");
    if (out_ta || out_rhob) {
      ta.resize((size_t)n * ncell);
      tb.resize((size_t)n * ncell);
      check(ctx_, smc_thickness(ctx_, ta.data(), tb.data()), "MakeDensity::generate_profile_ebe_Jet");
    }
    if (out_rho_binary) {
      rho.resize((size_t)n * ncell);
      check(ctx_, smc_rho_binary(ctx_, rho.data()), "MakeDensity::generate_profile_ebe_Jet");
    }
    if (out_spec) {
      spa.resize((size_t)n * ncell);
      spb.resize((size_t)n * ncell);
      check(ctx_, smc_spectator_density(ctx_, spa.data(), spb.data()), "MakeDensity::generate_profile_ebe_Jet");
    }
    for (int e = 0; e < n && event <= nevent; ++e) {
      const double* rec = &batch_rec_[(size_t)e * SMC_EV_STRIDE];
      if (!accept(rec)) {
        ++dropped_;
        continue;
      }
      const double npart = rec[SMC_EV_NPART];
      std::snprintf(buffer, sizeof(buffer), "data/ParticipantTable_event_%d.dat", event);
      smc_host::dump_participant_table(ctx_, p_, e, buffer);
      smc_host::dump_spectators_table(ctx_, p_, e, event);
      std::snprintf(buffer, sizeof(buffer), "data/BinaryCollisionTable_event_%d.dat", event);
      smc_host::dump_binary_table(ctx_, p_, e, buffer);
      const size_t off = (size_t)e * ncell;
      if (use_sd) {
        std::snprintf(buffer, sizeof(buffer), "data/sn_ecc_eccp_%%d_event_%d.dat", event);
        const std::string base = buffer;
        dumpEccentricities(base.c_str(), rec, from, to, npart, rec[SMC_EV_NCOLL], rec[SMC_EV_B]);
        grids("sd", &grid[off], p_.finalFactor, npart);
      }
      if (use_ed) {
        std::snprintf(buffer, sizeof(buffer), "data/en_ecc_eccp_%%d_event_%d.dat", event);
        const std::string base = buffer;
        dumpEccentricities(base.c_str(), rec, from, to, npart, rec[SMC_EV_NCOLL], rec[SMC_EV_B]);
        grids("ed", &grid[off], p_.finalFactor, npart);
      }
      if (out_rho_binary) grids("rho_binary", &rho[off], 1.0, npart);
      if (out_ta) {
        grids("nuclear_thickness_TA", &ta[off], 1.0, npart);
        grids("nuclear_thickness_TB", &tb[off], 1.0, npart);
      }
      if (out_rhob) {
        std::vector<float> sum(ncell);
        for (size_t i = 0; i < ncell; ++i) sum[i] = ta[off + i] + tb[off + i];
        grids("rhob", sum.data(), 1.0, npart);
      }
      if (out_spec) {
        grids("spectator_density_A", &spa[off], 1.0, npart);
        grids("spectator_density_B", &spb[off], 1.0, npart);
      }
      records_.insert(records_.end(), rec, rec + SMC_EV_STRIDE);
      ++event;
    }
  }
}
