// srs_internal.h -- declarations shared by the host C++ files and the HIP
// translation unit of libsrs_amd.  Not part of the ABI.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "srs_abi.h"

namespace srs {

// Records the message for srs_last_error() and returns code.
int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
void clear_error();

// Species indices, cpu/parameters.h:136-141.
enum { CA = 0, ATP = 1, CMDN = 2, TRPN = 3, CSQN = 4, DYE = 5, NS = 6 };

// Reverse Cuthill-McKee order of the element graph given by the packed
// operator (nij, self entries ignored).  order[i] = old index of new row i.
void rcm_order(int32_t n, const int32_t* nij, std::vector<int32_t>& order);

// Host replica of the reference loop bookkeeping: true when a states output
// is due at time t, cpu/RunSimulation.cpp:552 (floor crossing of accumulated t).
bool output_due(double t, double dt_last, double interval);

}  // namespace srs
