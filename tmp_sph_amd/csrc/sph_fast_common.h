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
// their candidates, so the block id may be remapped:
//   remap 0   tile = block (every XCD sees every eighth tile);
//   remap 1   every XCD gets one contiguous run of tiles (a eighth of the
//             domain each: best reuse, but the fluid is not spread evenly);
//   remap G   (a power of two >= 2) runs of G consecutive tiles are dealt to
//             the XCDs in turn, so each XCD's L2 serves whole runs while the
//             eight XCDs still work on neighbouring parts of the domain.  The
//             last partial round of 8 G tiles keeps the identity map.
__device__ __forceinline__ int tile_id(int remap) {
  const int b = blockIdx.x;
  if (remap == 0) return b;
  const int nb = gridDim.x;
  const int x = b & 7, k = b >> 3;
  if (remap == 1) {
    const int q = nb >> 3, r = nb & 7;
    return x * q + min(x, r) + k;
  }
  const int sh = __builtin_ctz((unsigned)remap);  // G = 1 << sh
  const int full = nb >> (sh + 3) << (sh + 3);     // tiles in whole rounds of 8 G
  if (b >= full) return b;
  return ((((k >> sh) << 3) + x) << sh) + (k & (remap - 1));
}

struct RowRange {
  int s, e;
};

// Sorted index ranges of the stencil rows of the cell with global key `key`,
// in the reference order (z outer, y inner; the x-adjacent cells of a row are
// contiguous in sorted order, so each row is one range).
//
// Cells whose box lies provably beyond the cut-off from the target p are left
// out: a whole row when its y/z face distance exceeds it, or the end cell of
// a row (the row stays one contiguous range).  Such cells hold no pair the
// loops would accept, so the visited pairs and their order are unchanged; in
// 3-D about a quarter of the candidates go.  Box distances are shrunk by
// gridSize/1024, far above the rounding of the cell bounds, and nothing is
// culled for a target outside the grid (its cell was clamped).
//
// With fluidcells (not null), each row is also trimmed to its first..last
// cell that holds a fluid particle: kernel A only pairs with fluid
// neighbours, so the cells cut away hold no pair it accepts, and a wall target
// with no fluid in its stencil skips its scan altogether.
template <int DIM>
__device__ __forceinline__ void row_ranges(const sph_params& P, int coff, int key, const float4& p,
                                           const int* __restrict__ cellStart,
                                           const int* __restrict__ cellEnd,
                                           const int* __restrict__ fluidcells,
                                           RowRange (&rows)[DIM == 3 ? 9 : 3]) {
  const int X = P.gridxdim, Y = P.gridydim, Z = P.gridzdim;
  const int XY = X * Y;
  const int hash = key - coff;
  const int cz = hash / XY;
  const int rem = hash - cz * XY;
  const int cy = rem / X;
  const int cx = rem - cy * X;
  const float gs = P.gridSize;
  const float m = gs * (1.0f / 1024.0f);
  const float cut = P.rr_cut;
  // distance from p to the lower / upper face of its own cell along each axis
  const float lx = P.gridxmin + (float)cx * gs, ly = P.gridymin + (float)cy * gs, lz = P.gridzmin + (float)cz * gs;
  const bool inside = p.x >= P.gridxmin && p.x <= P.gridxmin + (float)X * gs && p.z >= P.gridzmin &&
                      p.z <= P.gridzmin + (float)Z * gs &&
                      (DIM == 2 || (p.y >= P.gridymin && p.y <= P.gridymin + (float)Y * gs));
  auto face = [&](float d) { return inside ? fmaxf(d - m, 0.0f) : 0.0f; };
  const float xl = face(p.x - lx), xr = face(lx + gs - p.x);
  const float yl = face(p.y - ly), yr = face(ly + gs - p.y);
  const float zl = face(p.z - lz), zr = face(lz + gs - p.z);
  int nr = 0;
#pragma unroll
  for (int dz = -1; dz <= 1; ++dz) {
#pragma unroll
    for (int dy = (DIM == 3 ? -1 : 0); dy <= (DIM == 3 ? 1 : 0); ++dy) {
      const int z = cz + dz, y = cy + dy;
      const float fz = dz < 0 ? zl : (dz > 0 ? zr : 0.0f);
      const float fy = dy < 0 ? yl : (dy > 0 ? yr : 0.0f);
      const float dyz = fz * fz + fy * fy;
      int s = 0, e = 0;
      if (z >= 0 && z < Z && y >= 0 && y < Y && dyz <= cut) {
        const int x0 = (cx > 0 && xl * xl + dyz <= cut) ? cx - 1 : cx;
        const int x1 = (cx < X - 1 && xr * xr + dyz <= cut) ? cx + 1 : cx;
        const int rowBase = coff + z * XY + y * X;
        s = -1;
        for (int c = rowBase + x0; c <= rowBase + x1; ++c) {
          const int st = cellStart[c];
          if (st != -1 && (!fluidcells || fluidcells[c])) {
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

// 16-byte and 4-byte loads at a 32-bit element offset from a wave-uniform
// base: the address is base (SGPRs) + one 32-bit VGPR offset, and the unrolled
// candidate loads share that offset through the instruction's immediate.
__device__ __forceinline__ float4 ld16(const float4* base, unsigned byte_off) {
  return *reinterpret_cast<const float4*>(reinterpret_cast<const char*>(base) + byte_off);
}
__device__ __forceinline__ float ld4(const float* base, unsigned byte_off) {
  return *reinterpret_cast<const float*>(reinterpret_cast<const char*>(base) + byte_off);
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
