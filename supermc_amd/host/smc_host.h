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

}  // namespace smc_host

#endif
