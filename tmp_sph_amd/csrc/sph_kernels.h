// Host-side launchers of the gfx950 kernels (internal C++ interface).
#pragma once

#include <hip/hip_runtime.h>

#include "sph_internal.h"

namespace sph {

// ----------------------------------------------------------------- radix sort
// Stable LSD radix sort of (key, value) int pairs on keys in [0, 2^bits).
// Scratch: keys_alt/vals_alt of n ints, hist of radix_hist_words(n) words,
// totals of 256 words.  The sorted pairs end in (keys, vals) or in
// (keys_alt, vals_alt); *result_in_alt tells which.
int radix_hist_words(int n);
hipError_t radix_sort_pairs(int* keys, int* vals, int* keys_alt, int* vals_alt, unsigned* hist,
                            unsigned* totals, int n, int bits, hipStream_t s, bool* result_in_alt);

hipError_t fill_int(int* p, int value, long long n, hipStream_t s);

// --------------------------------------------------- reference-surface kernels
hipError_t launch_hash_aos(const sph_params& P, const float* pos, int* hash, int* idx, int n,
                           hipStream_t s);
hipError_t launch_reorder_aos(const sph_params& P, const float* oPos, const float* oVel,
                              const float* oRho, const float* oP, const int* oType,
                              const int* oZone, float* sPos, float* sVel, float* sRho, float* sP,
                              int* sType, int* sZone, const int* hash, const int* idx,
                              int* cellStart, int* cellEnd, int n, int* table_l,
                              const int* table_r, hipStream_t s);
hipError_t launch_recover_aos(float* pos, float* vel, float* rho, float* p, int* type, int* zone, const int* idx,
                              const float* sPos, const float* sVel, const float* sRho, const float* sP,
                              const int* sType, const int* sZone, int n, hipStream_t s);
hipError_t launch_ns_aos(const sph_params& P, const float* sPos, const float* sVel, float* sRho,
                         float* sP, const int* sType, float* drho, float* acc,
                         const int* cellStart, const int* cellEnd, const int* hash, float* dofv,
                         int n, hipStream_t s);
hipError_t launch_pc_aos(const sph_params& P, bool predictor, float* sPos, float* sVel,
                         float* sRho, float* sP, const float* drho, const float* acc,
                         const int* idx, float* tPos, float* tVel, float* tRho, int n,
                         int* sType, hipStream_t s);

// ------------------------------------------------------------- fast path
// One sorted particle set.  PR = (x, y, z, rho signed: + fluid, - otherwise),
// VP = (vx, vy, vz, p); X = FP64 positions (SoA); ID = original id; TY = type.
// Spare entries after PR of a fast set: the neighbour loops issue their
// unrolled candidate loads without clamping to the row end.
constexpr int CAND_PAD = 32;

struct FastSet {
  float4* PR;
  float4* VP;
  double* X[3];
  int* ID;
  int* TY;
};

// Per-half-step work arrays shared by both particle sets.
struct FastWork {
  int* keys;
  int* idx;
  int* keys_alt;
  int* idx_alt;
  unsigned* hist;
  unsigned* totals;
  int* cellStart;
  int* cellEnd;
  float* DV;       // velocity divergence (acoustic damper)
  float4* RATE;    // (drho/dt, ax, ay, az)
  double* TX[3];   // positions at the start of the step (predictor slot order)
  float4* TV;      // (vx, vy, vz, rho) at the start of the step
  int* cellFluid;  // 1 for cells holding a fluid particle (rebuilt with the cell ranges)
  int* NL;         // neighbour lists, slot-major: NL[k * n + i] (k < nl_cap); null until list mode is used
  int* NC;         // neighbour count of each target (> nl_cap: list overflowed)
  int nl_cap;
  // dense cell table of the fine-row loops (nbr_lds 4): cellFirst[c] = first
  // sorted index with key >= c, for every cell and a sentinel; fluidC/nearC:
  // coarse-cell fluid flags that let kernel A skip walls away from the water;
  // runs / runCount: queue of the longest runs of cells (runs_cap entries of
  // 3 ints)
  int* cellFirst;
  unsigned char* fluidC;  // per cell of the instances' own (2h) grid: holds fluid
  unsigned char* nearC;   // the same: fluid within one cell
  int* runs;
  int* runCount;
  int runs_cap;
};

// Neighbour-list capacity per target: 3-D supports hold ~74 neighbours on a
// lattice, 2-D ~21; targets with more use the scan path.
constexpr int NL_CAP_3D = 128;
constexpr int NL_CAP_2D = 48;

// Device table of the instances of one context (K >= 1).  Instance k owns the
// global cells [cell_off[k], cell_off[k] + P[k].numCells); all share `dim`.
struct InstTable {
  const sph_params* P;
  const int* cell_off;
  int K;
  int dim;
  int total_cells;
};

// The reference surface's SPH_NS_simpleversion on the fast loops: the
// caller's sorted AoS arrays packed into st (PR/VP/TY, n + CAND_PAD entries)
// with the per-cell fluid flags (`cells` entries, cleared here), and the
// rates, dofv and the walls' pressure and density unpacked back.
hipError_t launch_ref_pack(const float* sPos, const float* sVel, const float* sRho, const float* sP,
                           const int* sType, const int* hash, const FastSet& st, int* cellFluid, int cells, int n,
                           hipStream_t s);
hipError_t launch_ref_unpack(const FastSet& st, const FastWork& w, const int* sType, float* drho, float* acc,
                             float* dofv, float* sRho, float* sP, int n, hipStream_t s);

// zero (may be null): one word set to 0 by the launch (the dense table's run counter).
// zero: an int set to 0 on the way; cellStart/cellFluid: cell tables cleared
// on the way (for launch_cells_gather_fast with cleared = true).
hipError_t launch_hash_fast(const InstTable& T, const FastSet& in, int* keys, int* idx, int n, hipStream_t s,
                            int* zero = nullptr, int* cellStart = nullptr, int* cellFluid = nullptr);
// Dense cell table (cellFirst) of the sorted keys, the gather into sorted
// order and the coarse fluid flags (Tc: the instances' own grids): the cell
// build of the fine-row loops.
hipError_t launch_cells_dense(const InstTable& T, const InstTable& Tc, const FastSet& in, const FastSet& out,
                              const int* keys, const int* idx, const FastWork& w, int n, hipStream_t s);
// Capacity of the run queue for a table of `cells` cells.
int dense_runs_cap(long long cells);
// cleared: the hash already cleared cellStart and cellFluid.
hipError_t launch_cells_gather_fast(const InstTable& T, const FastSet& in, const FastSet& out, const int* keys,
                                    const int* idx, int* cellStart, int* cellEnd, int* cellFluid, int n,
                                    hipStream_t s, bool cleared = false);
hipError_t launch_nsA_fast(const InstTable& T, const FastSet& st, const FastWork& w, int n, const sph_tuning& tu,
                           hipStream_t s);
// Neighbour-list mode: one scan records each target's in-support neighbours
// (cut-off and j != i, fluid and wall) in visit order; kernels A and B then
// iterate the lists.  Same pairs in the same order as the scan kernels.
hipError_t launch_nbrlist_fast(const InstTable& T, const FastSet& st, const FastWork& w, int n, const sph_tuning& tu,
                               hipStream_t s);
// walls: also the rates of non-fluid targets (the reference computes them;
// the integrator ignores them); otherwise their RATE is zero.
hipError_t launch_nsB_fast(const InstTable& T, const FastSet& st, const FastWork& w, int n, const sph_tuning& tu,
                           hipStream_t s, bool walls = false);
// Fine-row loops (tuning nbr_lds 4): any stencil reach, per-lane trimmed
// candidate ranges, one cell-table read per wave and stencil row.
hipError_t launch_nsA_fine(const InstTable& T, const InstTable& Tc, const FastSet& st, const FastWork& w, int n,
                           const sph_tuning& tu, hipStream_t s);
hipError_t launch_nsB_fine(const InstTable& T, const FastSet& st, const FastWork& w, int n, const sph_tuning& tu,
                           hipStream_t s);
hipError_t launch_pair_stats_fine(const InstTable& T, const InstTable& Tc, const FastSet& st, const FastWork& w,
                                  unsigned long long* out4, int n, hipStream_t s);
hipError_t launch_pc_fast(const InstTable& T, bool predictor, const FastSet& st, const FastWork& w,
                          const int* perm, int n, hipStream_t s);
hipError_t launch_upload_fast(int dim, const FastSet& st, const float* pos, const float* vel, const float* rho,
                              const float* prs, const int* type, const int* inst, int n, hipStream_t s);
hipError_t launch_download_fast(const FastSet& st, float* pos, float* vel, float* rho, float* prs, int* type,
                                double* pos64, int n, hipStream_t s);
hipError_t launch_download_rates(const FastSet& st, const FastWork& w, float* drho, float* acc, float* dofv, int n,
                                 hipStream_t s);
hipError_t launch_pair_stats(const InstTable& T, const FastSet& st, const FastWork& w, unsigned long long* out4,
                             int n, hipStream_t s);
hipError_t launch_gauges_fast(const InstTable& T, const FastSet& st, const float* gx, const float* gy,
                              const float* zcap, int ng, float hw, unsigned* out, int n, hipStream_t s);
hipError_t launch_pressure_at_fast(const FastSet& st, const int* ids, int nid, float* out, int n, hipStream_t s);
hipError_t launch_probe_fast(const InstTable& T, const FastSet& st, double band, unsigned long long* out, int n,
                             hipStream_t s);

}  // namespace sph
