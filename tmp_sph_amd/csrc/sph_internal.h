// Internal declarations shared by the host and HIP translation units of the
// MI355X delta-SPH library.  Nothing here is part of the C ABI.
#pragma once

#include <cstddef>
#include <cstdint>

#include "../../include/sph_capi.h"

#ifdef __HIPCC__
#define SPH_HD __host__ __device__
#else
#define SPH_HD
#endif

namespace sph {

// Stencil reach of a grid: the number of cells K on each side of a target's
// cell that can hold a neighbour, the smallest K >= 1 with K * gridSize >= 2h
// (the reference's grid, gridSize = 2h, has K = 1; a grid subdivided by d has
// K = d).
SPH_HD inline int stencil_reach(float kh, float gridSize) {
  int k = 1;
  while (k < 8 && (float)k * gridSize < kh * 0.999999f) ++k;
  return k;
}

// Records `msg` as the calling thread's last error and returns `code`.
int set_error(int code, const char* msg);
// Maps a HIP status to SPH_OK or a positive code and records its message.
int hip_status(int hip_err, const char* where);

float cutoff_rr(float h);

// Validates a parameter block (SPH_OK or SPH_E_INVAL with the reason recorded)
// without touching the process-wide block of the reference call surface.
int check_params(const sph_params* p);

// Number of key bits needed for cell hashes in [0, numCells).
inline int key_bits(int numCells) {
  int b = 1;
  while (b < 31 && (1u << b) < (unsigned)numCells) ++b;
  return b;
}

}  // namespace sph
