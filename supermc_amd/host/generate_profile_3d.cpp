// generate_profile_3d.e -- the driver of scripts/generate_3d_profiles
// (main.cpp): reads an operation-2 participant list (x y id per line) and
// binary-collision list (x y per line), builds the 261 x 261 x 101 profile
// at ecm = 19.6 GeV with random_flag 1 on the GPU and writes the
// participant density to test.dat.
#include <fstream>
#include <iostream>
#include <vector>

#include "profile_3d.h"

using namespace std;

int main(int argc, char* argv[]) {
  if (argc != 3) {
    cout << "Usage: generate_profile_3d.e participant_list binary_collision_list" << endl;
    exit(0);
  }
  string participant_filename = argv[1];
  string binary_collision_filename = argv[2];
  cout << "read in files: " << participant_filename << " and " << binary_collision_filename << endl;

  // one record per line; a trailing partial record is dropped (main.cpp:24-43)
  vector<participant_info> part_list;
  ifstream part_if(participant_filename.c_str());
  participant_info t{};
  while (part_if >> t.x >> t.y >> t.id) part_list.push_back(t);
  vector<participant_info> binary_list;
  ifstream binary_if(binary_collision_filename.c_str());
  participant_info b{};
  while (binary_if >> b.x >> b.y) binary_list.push_back(b);

  cout << "Npart = " << part_list.size() << endl;
  cout << "Ncoll = " << binary_list.size() << endl;

  // main.cpp:46-57
  int grid_nx = 261;
  int grid_ny = 261;
  int grid_neta = 101;
  double grid_dx = 0.1;
  double grid_dy = 0.1;
  double grid_deta = 0.1;
  double ecm = 19.6;
  int random_flag = 1;

  profile_3d test(part_list, binary_list, grid_nx, grid_ny, grid_neta, grid_dx, grid_dy, grid_deta, ecm,
                  random_flag);
  test.generate_3d_profile();
  test.output_3d_rhob_profile("test.dat");
  return 0;
}
