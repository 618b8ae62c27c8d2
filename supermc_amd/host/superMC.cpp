// superMC.e -- the reference's driver (main.cpp) on the MI355X path:
// parameters.dat, then "key=value" arguments, then operation 9
// (eccentricity table), 1 (event-by-event profiles), 2 (event-by-event
// profiles with the jet-calculation tables and grids) or 3 (averaged rotated
// profiles).
#include <cstdlib>
#include <ctime>
#include <iostream>

#include "MakeDensity.h"
#include "ParameterReader.h"

int main(int argc, char* argv[]) {
  ParameterReader paraRdr;
  paraRdr.readFromFile("parameters.dat");
  paraRdr.readFromArguments(argc, argv);
  paraRdr.echo();

  const int nevent = (int)paraRdr.getVal("nev");
  const int operation = (int)paraRdr.getVal("operation");
  const clock_t start = clock();
  {
    MakeDensity dens(&paraRdr);
    switch (operation) {
      case 1:
        dens.generate_profile_ebe(nevent);
        break;
      case 2:
        dens.generate_profile_ebe_Jet(nevent);
        break;
      case 3:
        dens.generate_profile_average(nevent);
        break;
      case 9:
        dens.generateEccTable(nevent);
        break;
      default:
        std::cout << "Error: operation choice " << operation << " not recognized (this build runs 1, 2, 3 and 9)."
                  << std::endl;
    }
  }
  std::cout << "Time elapsed (in seconds): " << (double)(clock() - start) / CLOCKS_PER_SEC << std::endl;
  return 0;
}
