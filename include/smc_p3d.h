/*
 * smc_p3d.h -- C-ABI of the 3D (x, y, eta) profiles that extend operation 2
 * (reference: scripts/generate_3d_profiles/profile_3d.cpp, class profile_3d).
 *
 * Plain pointers and sizes; status codes and SMC_OK as in smc.h.  A handle
 * holds what the profile_3d constructor sets up (profile_3d.cpp:17-110):
 * the grid, y_beam, sigma_inelastic, the tabulated eta distribution.
 *
 * Sources are the operation-2 participant list (x, y, id: 1 projectile,
 * otherwise target; MCnucl::dumpparticipantTable) and binary-collision list
 * (x, y; MCnucl::dumpBinaryTable) of each event, events concatenated with
 * offsets: part_off[nev + 1], bin_off[nev + 1].  A source record is 6
 * doubles {x, y, eta, sigma_x, sigma_y, sigma_eta} (participant_info).
 * Grids are float [nev][grid_neta][grid_nx][grid_ny] (rho[j][k][l] of the
 * reference, y contiguous).
 */
#ifndef SMC_P3D_H
#define SMC_P3D_H
#include <stdint.h>

#include "smc.h"

#ifdef __cplusplus
extern "C" {
#endif

#define SMC_P3D_TABLE 1000 /* entries of the eta distribution (profile_3d.cpp:136) */

typedef struct {
  int grid_nx, grid_ny, grid_neta;    /* profile_3d(...) arguments (profile_3d.cpp:17-22) */
  double grid_dx, grid_dy, grid_deta;
  double ecm;                         /* sqrt(s_NN) [GeV] */
  int random_flag;                    /* 0..3 (profile_3d.cpp:194-300) */
} smc_p3d_params;

typedef struct {
  double y_beam;           /* atanh(sqrt(1 - 1/(ecm/2)^2)) (profile_3d.cpp:24) */
  double sigma_inelastic;  /* Regge96 sigma_tot - sigma_el, fm^2 (profile_3d.cpp:25-27) */
  double eta0;             /* 2.0 (profile_3d.cpp:196) */
  double sigma_x0;         /* sqrt(sigma_inelastic / (8 pi)) (profile_3d.cpp:197-198) */
  double sigma_eta0;       /* 0.5 (profile_3d.cpp:199) */
  double proj_max, targ_max; /* maxima of the two eta tables (profile_3d.cpp:135-155) */
} smc_p3d_derived;

typedef struct smc_p3d smc_p3d;

/* profile_3d::profile_3d + set_eta_distribution (profile_3d.cpp:17-110,
 * 133-155): validates the grid and random_flag, derives y_beam and
 * sigma_inelastic, tabulates the eta distribution on the device. */
smc_status smc_p3d_create(smc_p3d** h, const smc_p3d_params* p);
void smc_p3d_destroy(smc_p3d* h);
const char* smc_p3d_last_error(const smc_p3d* h);
/* The derived constants and (each optional, SMC_P3D_TABLE entries) the eta
 * table: eta, projectile and target weights. */
smc_status smc_p3d_get_derived(const smc_p3d* h, smc_p3d_derived* d, double* eta, double* proj, double* targ);

/* profile_3d::set_variables (profile_3d.cpp:194-300) with
 * sample_eta_distribution_from_array (:157-191) on the device.  Each
 * source's draws are Philox keyed by (seed, event0 + event, list, index):
 * the same law as the reference's one GSL stream, independent of batch
 * boundaries.  Binary collision i takes the id of participant i (the
 * reference's participant_list[i].id), wrapping past the list's end.
 * part_src / bin_src receive the source records (either may be null). */
smc_status smc_p3d_sample(smc_p3d* h, int nev, const double* part_xy, const int* part_id, const int* part_off,
                          const double* bin_xy, const int* bin_off, uint64_t seed, uint64_t event0,
                          double* part_src, double* bin_src);

/* profile_3d::generate_3d_profile / generate_3d_profile_binary
 * (profile_3d.cpp:302-400) for given source records: rho [nev][neta][nx][ny]
 * += norm exp(-dis_eta - dis_x - dis_y) over each source's +-6 sigma cell
 * windows, in list order. */
smc_status smc_p3d_smear(smc_p3d* h, int nev, const double* src, const int* src_off, float* rho);

/* Both steps without a host round trip: sampling, then rho_part from the
 * participants and rho_binary from the binary collisions (each grid
 * optional, and the source records if part_src / bin_src are given). */
smc_status smc_p3d_generate(smc_p3d* h, int nev, const double* part_xy, const int* part_id, const int* part_off,
                            const double* bin_xy, const int* bin_off, uint64_t seed, uint64_t event0,
                            float* rho_part, float* rho_binary, double* part_src, double* bin_src);

#ifdef __cplusplus
}
#endif
#endif
