// generate_profile_3d.e -- drop-in for the driver of
// scripts/generate_3d_profiles (main.cpp:9-68).  Arguments: an operation-2
// participant table (x y nucleus-id per line) and binary-collision table
// (x y per line).  It builds the participant density on the reference's
// fixed set-up -- 261 x 261 cells of 0.1 fm, 101 eta slices of 0.1,
// sqrt(s) = 19.6 GeV, random_flag 1 -- on the GPU and writes it to test.dat
// in the working directory.
#include <cstdio>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "profile_3d.h"

namespace {

// the reference driver's fixed profile set-up (main.cpp:46-57)
struct Setup {
  int nx = 261, ny = 261, neta = 101;
  double dx = 0.1, dy = 0.1, deta = 0.1;
  double ecm = 19.6;
  int random_flag = 1;
};

// whole records of `fields` numbers per line; x, y and (with three fields)
// the nucleus id
std::vector<participant_info> read_table(const char* path, int fields) {
  std::vector<participant_info> out;
  std::ifstream in(path);
  participant_info r{};
  while (in >> r.x >> r.y && (fields == 2 || in >> r.id)) {
    out.push_back(r);
    r = participant_info{};
  }
  return out;
}

}  // namespace

int main(int argc, char* argv[]) {
  if (argc != 3) {
    std::cout << "Usage: generate_profile_3d.e participant_list binary_collision_list" << std::endl;
    return 0;
  }
  std::cout << "read in files: " << argv[1] << " and " << argv[2] << std::endl;
  const std::vector<participant_info> parts = read_table(argv[1], 3);
  const std::vector<participant_info> collisions = read_table(argv[2], 2);
  std::cout << "Npart = " << parts.size() << std::endl;
  std::cout << "Ncoll = " << collisions.size() << std::endl;

  const Setup s;
  profile_3d profile(parts, collisions, s.nx, s.ny, s.neta, s.dx, s.dy, s.deta, s.ecm, s.random_flag);
  profile.generate_3d_profile();
  profile.output_3d_rhob_profile("test.dat");
  return 0;
}
