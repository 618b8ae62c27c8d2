// facade_check -- drives the reference-interface facade for the tests.
//   facade_check reader <params-file> [key=value ...]   parse + print names/values (no GPU)
//   facade_check mcnucl <params-file> b0 b1 ...         one event per b through MCnucl
//   facade_check hotspots <params-file> b0 b1 ...       MCnucl::getHotSpots, then after rotateGrid
//   facade_check kln <ecm> <lambda> ta0 tb0 ta1 tb1 ... KLNModel::getdNdy per point
//   facade_check dndytable <params-file> <out>          MCnucl::makeTable, dumpdNdyTable4Col
//   facade_check p3d <parts> <bins> nx ny neta dx deta ecm random_flag alpha <out>
//                                                       profile_3d: both grids, entropy file
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include <fstream>

#include "KLNModel.h"
#include "MCnucl.h"
#include "ParameterReader.h"
#include "profile_3d.h"

int main(int argc, char* argv[]) {
  if (argc < 3) return 2;
  const char* mode = argv[1];
  if (!std::strcmp(mode, "reader")) {
    ParameterReader r;
    r.readFromFile(argv[2]);
    r.readFromArguments(argc, argv, "#", 3);
    const char* keys[] = {"which_mc_model", "ecm", "maxx", "gauss_nucl_width", "aproj", "finalfactor"};
    for (const char* k : keys) std::printf("%s %d %.17g\n", k, r.exist(k) ? 1 : 0, r.exist(k) ? r.getVal(k) : 0.0);
    std::printf("missing %d\n", r.exist("no_such_key") ? 1 : 0);
    r.getVal("no_such_key");  // exits with -1 and the reference's message
    return 0;
  }
  if (!std::strcmp(mode, "mcnucl")) {
    ParameterReader r;
    r.readFromFile(argv[2]);
    MCnucl mc(&r);
    for (int i = 3; i < argc; ++i) {
      const double b = std::atof(argv[i]);
      mc.generateNuclei(b);
      const int nc = mc.getBinaryCollision();
      const int cut = mc.CentralityCut();
      double ta = 0, tb = 0, rho = 0;
      if (nc > 0) {
        mc.calculateThickness();
        mc.setDensity(0, -1);
        for (int x = 0; x < 261; ++x)
          for (int y = 0; y < 261; ++y) {
            ta += mc.getTA1(x, y);
            tb += mc.getTA2(x, y);
            rho += mc.getRho(0, x, y);
          }
      }
      std::printf("%d %d %d %d %d %.17g %.17g %.17g %.17g\n", i - 3, mc.getNpart1(), mc.getNpart2(), nc, cut,
                  nc > 0 ? mc.getdNdy() : 0.0, ta, tb, rho);
      mc.deleteNucleus();
    }
    return 0;
  }
  if (!std::strcmp(mode, "p3d") && argc == 13) {
    std::vector<participant_info> parts, bins;
    std::ifstream pf(argv[2]), bf(argv[3]);
    participant_info t{};
    while (pf >> t.x >> t.y >> t.id) parts.push_back(t);
    while (bf >> t.x >> t.y) bins.push_back(t);
    const int nx = std::atoi(argv[4]), ny = std::atoi(argv[5]), neta = std::atoi(argv[6]);
    const double dx = std::atof(argv[7]), deta = std::atof(argv[8]), ecm = std::atof(argv[9]);
    profile_3d p(parts, bins, nx, ny, neta, dx, dx, deta, ecm, std::atoi(argv[10]));
    p.generate_3d_profile();
    p.generate_3d_profile_binary();
    p.output_3d_entropy_profile(argv[12], std::atof(argv[11]));
    std::printf("seed %llu\n", (unsigned long long)p.random_seed());
    for (const participant_info& q : p.participants())
      std::printf("part %.17g %.17g %.17g %.17g\n", q.eta, q.sigma_x, q.sigma_y, q.sigma_eta);
    std::printf("draw %.17g\n", p.sample_eta_distribution_from_array(1));
    std::vector<double> tab{-1.0, 0.0, 0.5, 2.0};
    std::printf("search %ld %ld %ld\n", p.binarySearch(&tab, 0.25, true), p.binarySearch(&tab, 3.0, true),
                p.binarySearch(&tab, -1.0, true));
    return 0;
  }
  if (!std::strcmp(mode, "hotspots")) {  // MCnucl::getHotSpots before and after rotateGrid
    ParameterReader r;
    r.readFromFile(argv[2]);
    MCnucl mc(&r);
    for (int i = 3; i < argc; ++i) {
      mc.generateNuclei(std::atof(argv[i]));
      const int nc = mc.getBinaryCollision();
      std::vector<double> xa, xb;
      std::vector<int> wa, wb;
      mc.getNucleons(xa, wa, xb, wb);
      std::printf("E %d %d\n", i - 3, nc);
      for (size_t k = 0; k < wa.size(); ++k) std::printf("N 0 %d %.17g %.17g\n", wa[k], xa[2 * k], xa[2 * k + 1]);
      for (size_t k = 0; k < wb.size(); ++k) std::printf("N 1 %d %.17g %.17g\n", wb[k], xb[2 * k], xb[2 * k + 1]);
      if (nc == 0) continue;
      for (const char* tag : {"B", "R"}) {
        if (tag[0] == 'R') {
          mc.setDensity(0, -1);
          mc.rotateGrid(0, 2);
        }
        std::vector<Box2D> hs;
        const Box2D u = mc.getHotSpots(hs);
        for (const Box2D& b : hs) std::printf("%s %.17g %.17g %.17g %.17g\n", tag, b.getXL(), b.getXR(), b.getYL(), b.getYR());
        std::printf("U%s %.17g %.17g %.17g %.17g\n", tag, u.getXL(), u.getXR(), u.getYL(), u.getYR());
      }
    }
    return 0;
  }
  if (!std::strcmp(mode, "dndytable") && argc == 4) {  // MCnucl::makeTable + dumpdNdyTable4Col
    ParameterReader r;
    r.readFromFile(argv[2]);
    MCnucl mc(&r);
    mc.makeTable();
    smc_derived d{};
    if (smc_get_derived(mc.context(), &d) != SMC_OK) return 3;
    const int n = d.kln_tmax;
    std::vector<double> flat((size_t)n * n);
    if (smc_kln_make_table(mc.context(), flat.data()) != SMC_OK) return 3;
    std::vector<double*> rows(n);
    for (int i = 0; i < n; ++i) rows[i] = &flat[(size_t)i * n];
    double** one = rows.data();
    mc.dumpdNdyTable4Col(argv[3], &one, 0);
    return 0;
  }
  if (!std::strcmp(mode, "kln")) {
    KLNModel k(std::atof(argv[2]), 7);
    k.setLambdaSaturation(std::atof(argv[3]));
    for (int i = 4; i + 1 < argc; i += 2)
      std::printf("%.17g\n", k.getdNdy(0.0, std::atof(argv[i]), std::atof(argv[i + 1])));
    return 0;
  }
  return 2;
}
