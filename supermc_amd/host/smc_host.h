// smc_host.h -- glue between the reference-style classes of this facade and
// the libsmc C-ABI: parameters.dat keys -> smc_params, and the fatal-error
// convention of the reference (message, then exit).
#ifndef SMC_HOST_H
#define SMC_HOST_H

#include <cstdint>

#include "ParameterReader.h"
#include "smc.h"

namespace smc_host {

// smc_params from the reader.  Keys the reference reads unconditionally are
// read with getVal (exit(-1) when missing, ParameterReader.cpp:149-156);
// model-specific keys only for that model; extension keys only if present.
smc_params params_from_reader(ParameterReader& r, int max_events);

// Philox stream key: randomSeed (< 0: microseconds of the clock, main.cpp).
uint64_t seed_from_reader(ParameterReader& r);

// Prints the library's message and exits (the reference's fatal paths print
// and call exit).  Returns normally when st == SMC_OK.
void check(smc_ctx* ctx, smc_status st, const char* where);

// Collision tables of event ev of the context, as MCnucl::dumpBinaryTable
// (MCnucl.cpp:1177-1211) writes them: binary_file (appended) gets the pair
// midpoints; data/wounded.data the participants of both nuclei in wound
// order with the nucleus id (1, 2); data/quarks.data (appended) the three
// valence quarks of each participant with their boxes; data/nucl1.data and
// data/nucl2.data every nucleon.  Quarks sit at their nucleon's centre (the
// quark table of the reference is not part of its checkout), so each quark
// box is the nucleon position +- 4 quark_width.  p: the context's parameters.
void dump_binary_table(smc_ctx* ctx, const smc_params& p, int ev, const char* binary_file);
// MCnucl::dumpparticipantTable (MCnucl.cpp:1213-1220): participants of both
// nuclei, appended to filename.
void dump_participant_table(smc_ctx* ctx, const smc_params& p, int ev, const char* filename);
// MCnucl::getSpectators + dumpSpectatorsTable (MCnucl.cpp:1222-1260): the
// nucleons without a collision (projectile then target, collision order)
// with rapidity +-Y of the beams, to data/Spectators_event_<event>.dat.
// Returns the number of spectators.
int dump_spectators_table(smc_ctx* ctx, const smc_params& p, int ev, int event);
// beam rapidity of the spectators (MCnucl.cpp:1225-1228)
double spectator_rapidity(double ecm);

}  // namespace smc_host

#endif
