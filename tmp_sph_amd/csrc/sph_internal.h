// Internal declarations shared by the host and HIP translation units of the
// MI355X delta-SPH library.  Nothing here is part of the C ABI.
#pragma once

#include <cstddef>
#include <cstdint>

#include "../../include/sph_capi.h"

namespace sph {

// Records `msg` as the calling thread's last error and returns `code`.
int set_error(int code, const char* msg);
// Maps a HIP status to SPH_OK or a positive code and records its message.
int hip_status(int hip_err, const char* where);

float cutoff_rr(float h);

// Number of key bits needed for cell hashes in [0, numCells).
inline int key_bits(int numCells) {
  int b = 1;
  while (b < 31 && (1u << b) < (unsigned)numCells) ++b;
  return b;
}

}  // namespace sph
