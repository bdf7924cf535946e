// Device helpers shared by the fast-path kernels (instance selection,
// tile mapping, stencil row ranges).
#pragma once

#include "sph_device.h"
#include "sph_kernels.h"

namespace sph {
namespace fast {

__device__ __forceinline__ float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float fast_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

__device__ __forceinline__ int inst_of(int ty) { return ty >> 8; }
__device__ __forceinline__ int type_of(int ty) { return ty & 0xff; }

// Calls body(P, cell_offset) once for this lane with its instance's
// parameters.  Each pass serves the lanes that share the first active lane's
// instance, so P is wave-uniform inside body.
template <class F>
__device__ __forceinline__ void with_instance(const InstTable& T, int inst, F&& body) {
  for (;;) {
    const int u = __builtin_amdgcn_readfirstlane(inst);
    if (inst == u) {
      body(T.P[u], T.cell_off[u]);
      break;
    }
  }
}

// Workgroups are dispatched round-robin over the 8 XCDs (block b runs on XCD
// b % 8), each with its own L2.  Neighbouring target tiles share most of
// their candidates, so the block id is remapped to give every XCD one
// contiguous run of tiles instead of every eighth tile.
__device__ __forceinline__ int tile_id(bool remap) {
  const int b = blockIdx.x;
  if (!remap) return b;
  const int nb = gridDim.x;
  const int q = nb >> 3, r = nb & 7;
  const int x = b & 7, k = b >> 3;
  return x * q + min(x, r) + k;
}

struct RowRange {
  int s, e;
};

// Sorted index ranges of the stencil rows of the cell with global key `key`,
// in the reference order (z outer, y inner; the x-adjacent cells of a row are
// contiguous in sorted order, so each row is one range).
template <int DIM>
__device__ __forceinline__ void row_ranges(const sph_params& P, int coff, int key,
                                           const int* __restrict__ cellStart,
                                           const int* __restrict__ cellEnd,
                                           RowRange (&rows)[DIM == 3 ? 9 : 3]) {
  const int X = P.gridxdim, Y = P.gridydim, Z = P.gridzdim;
  const int XY = X * Y;
  const int hash = key - coff;
  const int cz = hash / XY;
  const int rem = hash - cz * XY;
  const int cy = rem / X;
  const int cx = rem - cy * X;
  const int x0 = cx > 0 ? cx - 1 : 0;
  const int x1 = cx < X - 1 ? cx + 1 : X - 1;
  int nr = 0;
#pragma unroll
  for (int dz = -1; dz <= 1; ++dz) {
#pragma unroll
    for (int dy = (DIM == 3 ? -1 : 0); dy <= (DIM == 3 ? 1 : 0); ++dy) {
      const int z = cz + dz, y = cy + dy;
      int s = 0, e = 0;
      if (z >= 0 && z < Z && y >= 0 && y < Y) {
        const int rowBase = coff + z * XY + y * X;
        s = -1;
        for (int c = rowBase + x0; c <= rowBase + x1; ++c) {
          const int st = cellStart[c];
          if (st != -1) {
            if (s < 0) s = st;
            e = cellEnd[c];
          }
        }
        if (s < 0) s = e = 0;
      }
      rows[nr].s = s;
      rows[nr].e = e;
      ++nr;
    }
  }
}

// rows[r] for a run-time r without indexing the array dynamically (which
// would move it to scratch or LDS): a chain of selects on registers.
template <int N>
__device__ __forceinline__ RowRange pick_row(const RowRange (&rows)[N], int r) {
  int s = 0, e = 0;
#pragma unroll
  for (int k = 0; k < N; ++k) {
    const int m = -(int)(r == k);  // all ones for the wanted row
    s |= rows[k].s & m;
    e |= rows[k].e & m;
  }
  return RowRange{s, e};
}

__device__ __forceinline__ bool outside64(const sph_params& P, double x, double y, double z) {
  return x > P.domain_max[0] || x < P.domain_min[0] || y > P.domain_max[1] ||
         y < P.domain_min[1] || z > P.domain_max[2] || z < P.domain_min[2];
}

}  // namespace fast
}  // namespace sph
