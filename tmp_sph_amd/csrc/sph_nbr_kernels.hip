// Neighbour-loop kernels of the fast path: kernel A (wall pressure
// extrapolation, velocity divergence; opt_src_v1/equation/ns.cu:30-134) and
// kernel B (continuity with delta-SPH diffusion, momentum;
// opt_src_v1/equation/ns.cu:137-230).
//
// One thread per target in sorted order; a target visits its stencil rows
// (z outer, y inner; the x-adjacent cells of a row are one contiguous sorted
// range) in the reference's order and its candidates in ascending index.
// Launch variants, chosen by the autotuner (every variant visits the same
// pairs in the same order, so all give bit-identical results):
//   staged   one wave per block: when the wave's 64 targets lie in one cell
//            row of one instance, the union of their stencil rows is copied
//            once into LDS and candidates are read from there; other waves
//            read HBM/L2 ("global").
//   planes   the same with one z-plane of three stencil rows in LDS at a
//            time (a quarter of the LDS, so more waves per CU); the wave
//            restages at the first row of each plane.
//   compact  only ~15% of the candidates of a 3-D cell stencil are within the
//            cutoff, and a wave runs the pair body whenever any lane accepts,
//            so the plain loop pays the full pair cost on almost every
//            candidate.  The compact loop scans candidates cheaply, appends
//            each lane's accepted ones to its own LDS list and, whenever a
//            list is nearly full (and at the end), runs the pair body densely
//            over all lists.
//   rows     one stencil row at a time in LDS with the neighbour data the
//            pair body reads (velocity/pressure, dofv), so the loop makes no
//            global loads: the wave waits on memory once per row.
//   masks    tests a group of 4 x UNROLL candidates into a per-lane bit mask,
//            then runs the pair body over each lane's set bits in ascending
//            order: a wave executes as many bodies per group as its busiest
//            lane accepted, not one per candidate.
// A context starts with rows, unroll 4 and prefetch (neighbour data read with
// every candidate), the fastest shape measured for the 2-D and 3-D dam breaks
// from 0.5 M particles up.
#include "sph_pair.h"

namespace sph {

using namespace fast;

namespace {

inline unsigned blocks_for(long long n, int b) { return (unsigned)((n + b - 1) / b); }

constexpr int STAGE_CAP = 1024;   // float4 entries per staged wave (16 KB of LDS)
constexpr int STAGE_PAD = 8;      // spare tile entries read by unrolled loads past a row's end
constexpr int PLANE_CAP = 448;    // float4 entries of one staged plane (7 KB of LDS)
constexpr int PLANE_PAD = 32;     // spare plane entries read by a mask group past a row's end
constexpr int ROW_CAP = 192;      // entries of one staged row (with its neighbour data: 7 KB of LDS)
constexpr int COMPACT_CAP = 16;   // accepted candidates per lane per dense phase

template <int DIM>
constexpr int nrows() { return DIM == 3 ? 9 : 3; }

// ------------------------------------------------------------ row providers
// rows(r) returns the sorted range of stencil row r.
template <int NR>
struct RegRows {  // in registers, selected without dynamic indexing
  RowRange v[NR];
  __device__ __forceinline__ RowRange operator()(int r) const { return pick_row(v, r); }
};
struct LdsRows {  // in the thread's own LDS slots
  const RowRange* slot;
  int stride;
  __device__ __forceinline__ RowRange operator()(int r) const { return slot[r * stride]; }
};

// ------------------------------------------------------- candidate sources
// src.row(r) is called at the start of row r; src.at(j0, u) returns
// PR[j0 + u] for u < UNROLL (4 x UNROLL for the mask loop), reading up to that
// many entries minus one past a row's end (PR is allocated with CAND_PAD spare
// entries; results there are discarded).
// src.ext(extra, j) returns candidate j's neighbour data (extra(j) reads it
// from global memory).
struct GlobalExtras {
  template <class E>
  __device__ __forceinline__ auto ext(const E& extra, int j) const { return extra(j); }
};

struct GlobalSrc {
  const float4* PR;
  template <class E>
  __device__ __forceinline__ auto ext(const E& extra, int j) const { return extra(j); }
  __device__ __forceinline__ void row(int) {}
  __device__ __forceinline__ float4 at(int j0, int u) const { return ld16(PR, (unsigned)j0 * 16u + 16u * u); }
};

// Candidates of row r live at tile[j + shift[r]]; the shifts are wave-uniform.
template <int NR>
struct StagedSrc : GlobalExtras {
  const float4* tile;
  int shift[NR];
  int cur;
  __device__ __forceinline__ void row(int r) {
    int v = 0;
#pragma unroll
    for (int k = 0; k < NR; ++k) v |= shift[k] & -(int)(r == k);
    cur = v;
  }
  __device__ __forceinline__ float4 at(int j0, int u) const { return tile[j0 + cur + u]; }
};

// Candidates of the current plane (rows 3p..3p+2) at tile[j + sh[r - 3p]]:
// uni[r] is the wave's union of row r (wave-uniform).  row() is called by
// all 64 lanes of the (single-wave) block at every row; at the first row of
// a plane it copies the plane's three unions into the tile.
template <int NR>
struct PlaneSrc : GlobalExtras {
  const float4* PR;
  float4* tile;
  RowRange uni[NR];
  int sh[3];
  int cur;
  __device__ __forceinline__ void row(int r) {
    const int k = r % 3;
    if (k == 0) {
      __syncthreads();  // every lane is done with the previous plane
      const int lane = threadIdx.x;
      int base = 0;
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        const RowRange u = pick_row(uni, r + q);
        const int len = u.e - u.s;
        for (int t = lane; t < len; t += 128) {
          const float4 a = PR[u.s + t];
          const bool two = t + 64 < len;
          const float4 b = PR[u.s + (two ? t + 64 : t)];
          tile[base + t] = a;
          if (two) tile[base + t + 64] = b;
        }
        sh[q] = base - u.s;
        base += len;
      }
      __syncthreads();
    }
    cur = k == 0 ? sh[0] : (k == 1 ? sh[1] : sh[2]);
  }
  __device__ __forceinline__ float4 at(int j0, int u) const { return tile[j0 + cur + u]; }
};

// The current row's candidates at tp[j + cur] and their neighbour data at
// tv/td[j + cur] (td only when DV is given): uni[r] is the wave's union of row
// r.  row() is called by all 64 lanes of the (single-wave) block at every row.
template <int NR>
struct RowSrc {
  const float4* PR;
  const float4* VP;
  const float* DV;
  float4* tp;
  float4* tv;
  float* td;
  RowRange uni[NR];
  int cur;
  __device__ __forceinline__ void row(int r) {
    __syncthreads();  // every lane is done with the previous row
    const RowRange u = pick_row(uni, r);
    const int len = u.e - u.s;
    const int lane = threadIdx.x;
    for (int t = lane; t < len; t += 128) {  // two entries per lane per pass: loads first
      const bool two = t + 64 < len;
      const int t2 = two ? t + 64 : t;
      const float4 a0 = PR[u.s + t], a1 = PR[u.s + t2];
      const float4 b0 = VP[u.s + t], b1 = VP[u.s + t2];
      float c0 = 0.0f, c1 = 0.0f;
      if (DV) {
        c0 = DV[u.s + t];
        c1 = DV[u.s + t2];
      }
      tp[t] = a0;
      tv[t] = b0;
      if (DV) td[t] = c0;
      if (two) {
        tp[t2] = a1;
        tv[t2] = b1;
        if (DV) td[t2] = c1;
      }
    }
    cur = -u.s;
    __syncthreads();
  }
  __device__ __forceinline__ float4 at(int j0, int u) const { return tp[j0 + cur + u]; }
  template <class E>
  __device__ __forceinline__ auto ext(const E& extra, int j) const { return extra.from(tv, td, j + cur); }
};

// ------------------------------------------------------------ visit loops
// visit(accept, extra, pair): accept(j, pj) is the candidate test, extra(j)
// loads what the pair body needs of candidate j beyond PR (velocity etc.),
// pair(j, pj, e) is the body for an accepted candidate; candidates are seen
// in ascending order.  The loads for the accepted candidates of one unrolled
// group are issued together before the first pair body, so the group waits
// one memory latency, not one per accepted candidate.  With PF (prefetch)
// they are issued for every candidate together with its position (more
// loads, no wait after the cut-off test).
template <int NR, int U, bool PF, class Rows, class Src>
struct PlainVisit {
  Rows rows;
  Src src;
  template <class Accept, class Extra, class Pair>
  __device__ __forceinline__ void operator()(Accept&& accept, Extra&& extra, Pair&& pair) {
#pragma unroll 1
    for (int r = 0; r < NR; ++r) {
      const RowRange g = rows(r);
      src.row(r);
      for (int j0 = g.s; j0 < g.e; j0 += U) {
        float4 pjs[U];
        decltype(extra(0)) es[U];
#pragma unroll
        for (int u = 0; u < U; ++u) pjs[u] = src.at(j0, u);  // issued together
        if constexpr (PF) {
#pragma unroll
          for (int u = 0; u < U; ++u) es[u] = src.ext(extra, j0 + u);
        }
        bool acc[U];
#pragma unroll
        for (int u = 0; u < U; ++u) acc[u] = (j0 + u < g.e) & accept(j0 + u, pjs[u]);
        if constexpr (!PF) {
          // the accepted candidates' loads go out together: one wait per group
#pragma unroll
          for (int u = 0; u < U; ++u)
            if (acc[u]) es[u] = src.ext(extra, j0 + u);
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
          if (acc[u]) pair(j0 + u, pjs[u], es[u]);
      }
    }
  }
};

template <int NR, int U, bool PF, class Rows, class Src>
struct CompactVisit {
  Rows rows;
  Src src;
  int* list;  // this wave's COMPACT_CAP x 64 slots, [slot][lane]
  const float4* PR;
  template <class Accept, class Extra, class Pair>
  __device__ __forceinline__ void operator()(Accept&& accept, Extra&& extra, Pair&& pair) {
    const int lane = threadIdx.x & 63;
    int cnt = 0;
    // Dense phase over every lane's list, in order.  Called by the whole
    // wave (the conditions are wave-uniform); lanes keep their own order, so
    // flushing early changes nothing in the sums.  With PF the loads of four
    // list entries are issued before the first of them is used.
    auto flush = [&]() {
      if constexpr (PF) {
#pragma unroll 1
        for (int k = 0; k < cnt; k += 4) {
          int jj[4];
          float4 pj[4];
          decltype(extra(0)) e[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) jj[u] = list[min(k + u, cnt - 1) * 64 + lane];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            pj[u] = ld16(PR, (unsigned)jj[u] * 16u);
            e[u] = src.ext(extra, jj[u]);
          }
#pragma unroll
          for (int u = 0; u < 4; ++u)
            if (k + u < cnt) pair(jj[u], pj[u], e[u]);
        }
      } else {
#pragma unroll 1
        for (int k = 0; k < cnt; ++k) {
          const int jj = list[k * 64 + lane];  // [slot][lane]: conflict-free for any per-lane count
          pair(jj, ld16(PR, (unsigned)jj * 16u), src.ext(extra, jj));
        }
      }
      cnt = 0;
    };
#pragma unroll 1
    for (int r = 0; r < NR; ++r) {  // same row for all lanes
      const RowRange g = rows(r);
      src.row(r);
      int j = g.s;
#pragma unroll 1
      while (__ballot(j < g.e) != 0ull) {
        if (__ballot(cnt > COMPACT_CAP - U) != 0ull) flush();
        if (j < g.e) {
          float4 pjs[U];
#pragma unroll
          for (int u = 0; u < U; ++u) pjs[u] = src.at(j, u);
#pragma unroll
          for (int u = 0; u < U; ++u)
            if (j + u < g.e && accept(j + u, pjs[u])) {
              list[cnt * 64 + lane] = j + u;
              ++cnt;
            }
        }
        j += U;
      }
    }
    flush();
  }
};

template <int NR, int U, class Rows, class Src>
struct MaskVisit {
  Rows rows;
  Src src;
  template <class Accept, class Extra, class Pair>
  __device__ __forceinline__ void operator()(Accept&& accept, Extra&& extra, Pair&& pair) {
    constexpr int G = 4 * U;  // candidates per group (<= 32 mask bits)
#pragma unroll 1
    for (int r = 0; r < NR; ++r) {
      const RowRange g = rows(r);
      src.row(r);
      for (int j0 = g.s; j0 < g.e; j0 += G) {
        unsigned m = 0u;
#pragma unroll
        for (int b = 0; b < G; b += 4) {
          float4 pjs[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) pjs[u] = src.at(j0, b + u);  // issued together
#pragma unroll
          for (int u = 0; u < 4; ++u)
            if (j0 + b + u < g.e && accept(j0 + b + u, pjs[u])) m |= 1u << (b + u);
        }
        if (m != 0u) {
          // accepted candidates in ascending order; the loads of the next one
          // go out before the body of the current one
          int k = __builtin_ctz(m);
          auto e = src.ext(extra, j0 + k);
          for (;;) {
            m &= m - 1u;
            auto en = e;
            int kn = k;
            if (m != 0u) {
              kn = __builtin_ctz(m);
              en = src.ext(extra, j0 + kn);
            }
            pair(j0 + k, src.at(j0, k), e);
            if (m == 0u) break;
            k = kn;
            e = en;
          }
        }
      }
    }
  }
};

// The visit loop of a variant: COMPACT 1 = LDS lists, 2 = mask groups, else plain.
template <int NR, int U, int COMPACT, bool PF, class Rows, class Src>
__device__ __forceinline__ auto make_visit(Rows rows, Src src, int* list, const float4* PR) {
  if constexpr (COMPACT == 1)
    return CompactVisit<NR, U, PF, Rows, Src>{rows, src, list, PR};
  else if constexpr (COMPACT == 2)
    return MaskVisit<NR, U, Rows, Src>{rows, src};
  else
    return PlainVisit<NR, U, PF, Rows, Src>{rows, src};
}

// Iterates target i's recorded neighbour list (slot-major, NL[k * stride + i])
// in recorded order; the loads of four entries go out before the first is
// used.
struct ListVisit {
  const int* NL;
  int stride;
  int i;
  int cnt;
  const float4* PR;
  template <class Accept, class Extra, class Pair>
  __device__ __forceinline__ void operator()(Accept&& accept, Extra&& extra, Pair&& pair) {
#pragma unroll 1
    for (int k = 0; k < cnt; k += 4) {
      int jj[4];
      float4 pj[4];
      decltype(extra(0)) e[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) jj[u] = NL[(size_t)min(k + u, cnt - 1) * stride + i];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        pj[u] = ld16(PR, (unsigned)jj[u] * 16u);
        e[u] = extra(jj[u]);
      }
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (k + u < cnt && accept(jj[u], pj[u])) pair(jj[u], pj[u], e[u]);
    }
  }
};

// Decides whether the wave can stage its candidates (all 64 targets in one
// cell row of one instance, union of the row ranges within STAGE_CAP) and, if
// so, copies the union into `tile` and fills src.  Every lane of the
// (single-wave) block must call it.
template <int DIM>
__device__ __forceinline__ bool stage_rows(const sph_params& P, int coff, int key, bool same_inst,
                                           const RowRange (&rows)[DIM == 3 ? 9 : 3], const float4* PR,
                                           float4* tile, StagedSrc<nrows<DIM>()>& src) {
  const int row_id = (key - coff) / P.gridxdim;  // z * Y + y of the target's cell
  const int u_row = __builtin_amdgcn_readfirstlane(row_id);
  if (!same_inst || __ballot(row_id != u_row) != 0ull) return false;
  int lo[nrows<DIM>()], len[nrows<DIM>()];
  int total = 0;
#pragma unroll
  for (int r = 0; r < nrows<DIM>(); ++r) {
    const bool ne = rows[r].e > rows[r].s;
    const int l = wave_min(ne ? rows[r].s : 0x7fffffff);
    const int h = wave_max(ne ? rows[r].e : 0);
    lo[r] = l;
    len[r] = l == 0x7fffffff ? 0 : h - l;
    src.shift[r] = total - l;
    total += len[r];
  }
  if (total > STAGE_CAP) return false;
  const int lane = threadIdx.x;
  int base = 0;
#pragma unroll
  for (int r = 0; r < nrows<DIM>(); ++r) {
    for (int t = lane; t < len[r]; t += 64) tile[base + t] = PR[lo[r] + t];
    base += len[r];
  }
  src.tile = tile;
  src.cur = 0;
  __syncthreads();
  return true;
}

// Plane staging: the wave stages when its active targets lie in one cell row
// of one instance and every plane's three row unions fit PLANE_CAP.  Inactive
// lanes come with empty rows and only help with the copies.
template <int DIM>
__device__ __forceinline__ bool plan_planes(const sph_params& P, int coff, int key, bool active, bool same_inst,
                                            const RowRange (&rows)[DIM == 3 ? 9 : 3], const float4* PR,
                                            float4* tile, PlaneSrc<nrows<DIM>()>& src) {
  constexpr int NR = nrows<DIM>();
  const unsigned long long act = __ballot(active);
  const int row_id = (key - coff) / P.gridxdim;
  const int u_row = __builtin_amdgcn_readlane(row_id, __builtin_ctzll(act));
  if (!same_inst || __ballot(active && row_id != u_row) != 0ull) return false;
  bool fits = true;
#pragma unroll
  for (int p0 = 0; p0 < NR; p0 += 3) {
    int total = 0;
#pragma unroll
    for (int r = p0; r < p0 + 3; ++r) {
      const bool ne = rows[r].e > rows[r].s;
      const int l = wave_min(ne ? rows[r].s : 0x7fffffff);
      const int h = wave_max(ne ? rows[r].e : 0);
      src.uni[r] = l == 0x7fffffff ? RowRange{0, 0} : RowRange{l, h};
      total += src.uni[r].e - src.uni[r].s;
    }
    fits = fits && total <= PLANE_CAP;
  }
  src.PR = PR;
  src.tile = tile;
  src.cur = 0;
  return fits;
}

// Row staging: as plan_planes, with every row's union within ROW_CAP.
template <int DIM>
__device__ __forceinline__ bool plan_rows(const sph_params& P, int coff, int key, bool active, bool same_inst,
                                          const RowRange (&rows)[DIM == 3 ? 9 : 3], RowSrc<nrows<DIM>()>& src) {
  constexpr int NR = nrows<DIM>();
  const unsigned long long act = __ballot(active);
  const int row_id = (key - coff) / P.gridxdim;
  const int u_row = __builtin_amdgcn_readlane(row_id, __builtin_ctzll(act));
  if (!same_inst || __ballot(active && row_id != u_row) != 0ull) return false;
  bool fits = true;
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    const bool ne = rows[r].e > rows[r].s;
    const int l = wave_min(ne ? rows[r].s : 0x7fffffff);
    const int h = wave_max(ne ? rows[r].e : 0);
    src.uni[r] = l == 0x7fffffff ? RowRange{0, 0} : RowRange{l, h};
    fits = fits && src.uni[r].e - src.uni[r].s <= ROW_CAP;
  }
  src.cur = 0;
  return fits;
}

// ------------------------------------------------------------ kernels
// STAGED: 64-thread blocks with LDS staging (1: whole stencil, 2: plane by
// plane, 3: row by row with the neighbour data); else up to 256 threads.
// COMPACT: 1 scan/dense loop, 2 mask groups; else the plain loop.
template <int DIM, int U, int STAGED, int COMPACT>
struct Cfg {
  static constexpr int NR = nrows<DIM>();
  static constexpr int LIST = COMPACT == 1 ? COMPACT_CAP * (STAGED ? 64 : 256) : 1;
  static constexpr int ROW_PAD = COMPACT == 2 ? 4 * U : U;  // entries read past a row's end
  static constexpr int TILE = STAGED == 3   ? ROW_CAP + ROW_PAD
                             : STAGED == 2 ? PLANE_CAP + PLANE_PAD
                             : STAGED      ? STAGE_CAP + STAGE_PAD
                                           : 1;
  static constexpr int TILEX = STAGED == 3 ? ROW_CAP + ROW_PAD : 1;  // staged neighbour data
  static constexpr int SLOTS = (!STAGED && COMPACT != 1) ? NR * 256 : 1;
};

// Runs body(P, visit) for this lane with a visit loop over its stencil.  The
// global path serves lanes instance by instance (wave-uniform parameters).
template <int DIM, int U, int STAGED, int COMPACT, bool PF, class Body>
__device__ __forceinline__ void with_visit(const InstTable& T, int inst, int key, const float4& p, bool active,
                                           const float4* PR, const float4* VPx, const float* DVx,
                                           const int* __restrict__ cs, const int* __restrict__ ce,
                                           const int* __restrict__ fc, float4* tile, float4* tilev, float* tiled,
                                           int* lists, RowRange* slots, Body&& body) {
  constexpr int NR = nrows<DIM>();
  int* list = lists + (threadIdx.x >> 6) * COMPACT_CAP * 64;
  auto global_path = [&](const sph_params& P, int coff) {
    if constexpr (!STAGED && COMPACT != 1) {
      // row ranges parked in the thread's own LDS slots (no barrier needed):
      // keeps 18 VGPRs free in 3-D, which buys occupancy for the plain loop
      RowRange rows[NR];
      row_ranges<DIM>(P, coff, key, p, cs, ce, fc, rows);
      const int t = threadIdx.x;
#pragma unroll
      for (int r = 0; r < NR; ++r) slots[r * 256 + t] = rows[r];
      body(P, make_visit<NR, U, COMPACT, PF>(LdsRows{slots + t, 256}, GlobalSrc{PR}, list, PR), true);
    } else {
      RegRows<NR> rows;
      row_ranges<DIM>(P, coff, key, p, cs, ce, fc, rows.v);
      body(P, make_visit<NR, U, COMPACT, PF>(rows, GlobalSrc{PR}, list, PR), true);
    }
  };
  if constexpr (STAGED == 3) {
    // as the plane path below, restaging at every row
    if (__ballot(active) == 0ull) return;
    const unsigned long long act = __ballot(active);
    const int u = __builtin_amdgcn_readlane(inst, __builtin_ctzll(act));
    const bool same_inst = __ballot(active && inst != u) == 0ull;
    const sph_params& P = T.P[u];
    const int coff = T.cell_off[u];
    RegRows<NR> rows;
    row_ranges<DIM>(P, coff, key, p, cs, ce, fc, rows.v);
    if (!active) {
#pragma unroll
      for (int r = 0; r < NR; ++r) rows.v[r] = RowRange{0, 0};
    }
    RowSrc<NR> src;
    src.PR = PR;
    src.VP = VPx;
    src.DV = DVx;
    src.tp = tile;
    src.tv = tilev;
    src.td = tiled;
    if (plan_rows<DIM>(P, coff, key, active, same_inst, rows.v, src)) {
      body(P, make_visit<NR, U, COMPACT, PF>(rows, src, list, PR), active);
    } else {
      if (!active) return;
      with_instance(T, inst, global_path);
    }
  } else if constexpr (STAGED == 2) {
    // the whole wave stays in the loop (inactive lanes with empty rows) for
    // the per-plane copies
    if (__ballot(active) == 0ull) return;
    const unsigned long long act = __ballot(active);
    const int u = __builtin_amdgcn_readlane(inst, __builtin_ctzll(act));
    const bool same_inst = __ballot(active && inst != u) == 0ull;
    const sph_params& P = T.P[u];
    const int coff = T.cell_off[u];
    RegRows<NR> rows;
    row_ranges<DIM>(P, coff, key, p, cs, ce, fc, rows.v);
    if (!active) {
#pragma unroll
      for (int r = 0; r < NR; ++r) rows.v[r] = RowRange{0, 0};
    }
    PlaneSrc<NR> src;
    if (plan_planes<DIM>(P, coff, key, active, same_inst, rows.v, PR, tile, src)) {
      body(P, make_visit<NR, U, COMPACT, PF>(rows, src, list, PR), active);
    } else {
      if (!active) return;
      with_instance(T, inst, global_path);
    }
  } else if constexpr (STAGED) {
    // the whole wave takes part in the staging decision and copy
    const int u = __builtin_amdgcn_readfirstlane(inst);
    const bool same_inst = __ballot(inst != u) == 0ull;
    const sph_params& P = T.P[u];
    const int coff = T.cell_off[u];
    RegRows<NR> rows;
    row_ranges<DIM>(P, coff, key, p, cs, ce, fc, rows.v);
    StagedSrc<NR> src;
    const bool staged = stage_rows<DIM>(P, coff, key, same_inst, rows.v, PR, tile, src);
    if (!active) return;
    if (staged) {
      body(P, make_visit<NR, U, COMPACT, PF>(rows, src, list, PR), true);
    } else {
      with_instance(T, inst, global_path);
    }
  } else {
    if (!active) return;
    with_instance(T, inst, global_path);
  }
}

template <int DIM, int U, int STAGED, int COMPACT, bool PF>
__global__ __launch_bounds__(STAGED ? 64 : 256, STAGED >= 2 ? 6 : 1) void nsA_kernel(InstTable T, float4* PR, float4* VP,
                                                                const int* __restrict__ TY, float* __restrict__ DV,
                                                                const int* __restrict__ keys,
                                                                const int* __restrict__ cs,
                                                                const int* __restrict__ ce,
                                                                const int* __restrict__ fc, int n, int remap) {
  using C = Cfg<DIM, U, STAGED, COMPACT>;
  __shared__ float4 tile[C::TILE];
  __shared__ float4 tilev[C::TILEX];
  __shared__ float tiled[1];  // kernel A stages no dofv
  __shared__ int lists[C::LIST];
  __shared__ RowRange slots[C::SLOTS];
  const int i = tile_id(remap) * blockDim.x + threadIdx.x;
  const bool valid = i < n;
  if (!STAGED && !valid) return;
  const int ic = valid ? i : n - 1;  // staged: lane 0 is always valid (the grid covers n)
  const float4 pi = PR[ic];
  const float4 vi = VP[ic];
  const int key = keys[ic];
  // kernel A reads no neighbour's dofv (it writes its own)
  with_visit<DIM, U, STAGED, COMPACT, PF>(T, inst_of(TY[ic]), key, pi, valid, PR, VP, nullptr, cs, ce, fc, tile,
                                          tilev, tiled, lists, slots,
                                      [&](const sph_params& P, auto visit, bool store) {
                                        nsA_target<DIM>(P, i, pi, vi, visit, PR, VP, DV, store);
                                      });
}

// 5 waves per SIMD (up to 96 VGPRs): faster for kernel B than 6 (80 VGPRs);
// kernel A, with the smaller body, runs best at 6
template <int DIM, int U, int STAGED, int COMPACT, bool PF>
__global__ __launch_bounds__(STAGED ? 64 : 256, STAGED >= 2 ? 5 : 1) void nsB_kernel(InstTable T, const float4* __restrict__ PR,
                                                                const float4* __restrict__ VP,
                                                                const int* __restrict__ TY,
                                                                const float* __restrict__ DV,
                                                                const int* __restrict__ keys,
                                                                const int* __restrict__ cs,
                                                                const int* __restrict__ ce,
                                                                float4* __restrict__ RATE, int n, int remap,
                                                                int walls) {
  using C = Cfg<DIM, U, STAGED, COMPACT>;
  __shared__ float4 tile[C::TILE];
  __shared__ float4 tilev[C::TILEX];
  __shared__ float tiled[C::TILEX];
  __shared__ int lists[C::LIST];
  __shared__ RowRange slots[C::SLOTS];
  const int i = tile_id(remap) * blockDim.x + threadIdx.x;
  const bool valid = i < n;
  if (!STAGED && !valid) return;
  const int ic = valid ? i : n - 1;
  const float4 pi = PR[ic];
  // walls != 0: non-fluid targets get their rates too (the reference surface)
  const bool target = pi.w > 0.0f || walls != 0;
  if (valid && !target) RATE[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if (!STAGED && !target) return;
  const float4 vi = VP[ic];
  const float dvi = DV[ic];
  const int key = keys[ic];
  with_visit<DIM, U, STAGED, COMPACT, PF>(T, inst_of(TY[ic]), key, pi, valid && target, PR, VP, DV, cs, ce,
                                          nullptr, tile, tilev, tiled, lists, slots,
                                      [&](const sph_params& P, auto visit, bool store) {
                                        nsB_target<DIM>(P, i, pi, vi, dvi, visit, VP, DV, RATE, store);
                                      });
}

// ------------------------------------------------------------ neighbour lists
// Records every target's in-support neighbours (rr <= cut, j != i; fluid and
// wall) in the order the scan kernels visit them.
template <int DIM, int U>
__global__ __launch_bounds__(256) void nbrlist_kernel(InstTable T, const float4* __restrict__ PR,
                                                      const int* __restrict__ TY, const int* __restrict__ keys,
                                                      const int* __restrict__ cs, const int* __restrict__ ce,
                                                      int* __restrict__ NL, int* __restrict__ NC, int cap, int n,
                                                      int remap) {
  constexpr int NR = nrows<DIM>();
  __shared__ RowRange slots[NR * 256];
  const int i = tile_id(remap) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 pi = PR[i];
  const int key = keys[i];
  int cnt = 0;
  with_instance(T, inst_of(TY[i]), [&](const sph_params& P, int coff) {
    RowRange rows[NR];
    row_ranges<DIM>(P, coff, key, pi, cs, ce, nullptr, rows);
    const int t = threadIdx.x;
#pragma unroll
    for (int r = 0; r < NR; ++r) slots[r * 256 + t] = rows[r];
    const float cut = P.rr_cut;
    PlainVisit<NR, U, false, LdsRows, GlobalSrc>{LdsRows{slots + t, 256}, GlobalSrc{PR}}(
        [&](int j, const float4& pj) {
          float dx, dy, dz;
          return geom<DIM>(pi, pj, dx, dy, dz) <= cut && j != i;
        },
        ExtraNone{},
        [&](int j, const float4&, const NoExtra&) {
          if (cnt < cap) NL[(size_t)cnt * n + i] = j;
          ++cnt;
        });
  });
  NC[i] = cnt;
}

// Kernels A and B over the recorded lists; a target whose list overflowed
// scans its stencil instead.
template <int DIM>
__global__ __launch_bounds__(256) void nsA_list_kernel(InstTable T, float4* PR, float4* VP, const int* __restrict__ TY,
                                                       float* __restrict__ DV, const int* __restrict__ keys,
                                                       const int* __restrict__ cs, const int* __restrict__ ce,
                                                       const int* __restrict__ fc, const int* __restrict__ NL,
                                                       const int* __restrict__ NC, int cap, int n, int remap) {
  constexpr int NR = nrows<DIM>();
  const int i = tile_id(remap) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 pi = PR[i];
  const float4 vi = VP[i];
  const int key = keys[i];
  const int cnt = NC[i];
  with_instance(T, inst_of(TY[i]), [&](const sph_params& P, int coff) {
    if (cnt <= cap) {
      nsA_target<DIM>(P, i, pi, vi, ListVisit{NL, n, i, cnt, PR}, PR, VP, DV);
    } else {
      RegRows<NR> rows;
      row_ranges<DIM>(P, coff, key, pi, cs, ce, fc, rows.v);
      nsA_target<DIM>(P, i, pi, vi, PlainVisit<NR, 2, false, RegRows<NR>, GlobalSrc>{rows, GlobalSrc{PR}}, PR, VP,
                      DV);
    }
  });
}

template <int DIM>
__global__ __launch_bounds__(256) void nsB_list_kernel(InstTable T, const float4* __restrict__ PR,
                                                       const float4* __restrict__ VP, const int* __restrict__ TY,
                                                       const float* __restrict__ DV, const int* __restrict__ keys,
                                                       const int* __restrict__ cs, const int* __restrict__ ce,
                                                       const int* __restrict__ NL, const int* __restrict__ NC,
                                                       float4* __restrict__ RATE, int cap, int n, int remap) {
  constexpr int NR = nrows<DIM>();
  const int i = tile_id(remap) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 pi = PR[i];
  if (!(pi.w > 0.0f)) {
    RATE[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    return;
  }
  const float4 vi = VP[i];
  const float dvi = DV[i];
  const int key = keys[i];
  const int cnt = NC[i];
  with_instance(T, inst_of(TY[i]), [&](const sph_params& P, int coff) {
    if (cnt <= cap) {
      nsB_target<DIM>(P, i, pi, vi, dvi, ListVisit{NL, n, i, cnt, PR}, VP, DV, RATE);
    } else {
      RegRows<NR> rows;
      row_ranges<DIM>(P, coff, key, pi, cs, ce, nullptr, rows.v);
      nsB_target<DIM>(P, i, pi, vi, dvi, PlainVisit<NR, 2, false, RegRows<NR>, GlobalSrc>{rows, GlobalSrc{PR}}, VP,
                      DV, RATE);
    }
  });
}

// ------------------------------------------------------------ launchers
template <int U, int STAGED, int COMPACT, bool PF>
hipError_t nsA_v(const InstTable& T, const FastSet& st, const FastWork& w, int n, int block, int remap,
                 hipStream_t s) {
  const int b = STAGED ? 64 : block;
  const dim3 g(blocks_for(n, b)), bl(b);
  if (T.dim == 3)
    hipLaunchKernelGGL((nsA_kernel<3, U, STAGED, COMPACT, PF>), g, bl, 0, s, T, st.PR, st.VP, (const int*)st.TY, w.DV,
                       (const int*)w.keys, (const int*)w.cellStart, (const int*)w.cellEnd, (const int*)w.cellFluid,
                       n, remap);
  else
    hipLaunchKernelGGL((nsA_kernel<2, U, STAGED, COMPACT, PF>), g, bl, 0, s, T, st.PR, st.VP, (const int*)st.TY, w.DV,
                       (const int*)w.keys, (const int*)w.cellStart, (const int*)w.cellEnd, (const int*)w.cellFluid,
                       n, remap);
  return hipGetLastError();
}

template <int U, int STAGED, int COMPACT, bool PF>
hipError_t nsB_v(const InstTable& T, const FastSet& st, const FastWork& w, int n, int block, int remap,
                 int walls, hipStream_t s) {
  const int b = STAGED ? 64 : block;
  const dim3 g(blocks_for(n, b)), bl(b);
  if (T.dim == 3)
    hipLaunchKernelGGL((nsB_kernel<3, U, STAGED, COMPACT, PF>), g, bl, 0, s, T, (const float4*)st.PR,
                       (const float4*)st.VP, (const int*)st.TY, (const float*)w.DV, (const int*)w.keys,
                       (const int*)w.cellStart, (const int*)w.cellEnd, w.RATE, n, remap, walls);
  else
    hipLaunchKernelGGL((nsB_kernel<2, U, STAGED, COMPACT, PF>), g, bl, 0, s, T, (const float4*)st.PR,
                       (const float4*)st.VP, (const int*)st.TY, (const float*)w.DV, (const int*)w.keys,
                       (const int*)w.cellStart, (const int*)w.cellEnd, w.RATE, n, remap, walls);
  return hipGetLastError();
}

template <int U>
hipError_t nsA_u(const InstTable& T, const FastSet& st, const FastWork& w, int n, int block, int remap, int mode,
                 hipStream_t s) {
  switch (mode) {
    case 1: return nsA_v<U, true, false, false>(T, st, w, n, block, remap, s);
    case 2: return nsA_v<U, false, true, false>(T, st, w, n, block, remap, s);
    case 3: return nsA_v<U, true, true, false>(T, st, w, n, block, remap, s);
    case 4: return nsA_v<U, false, false, true>(T, st, w, n, block, remap, s);
    case 5: return nsA_v<U, true, false, true>(T, st, w, n, block, remap, s);
    case 6: return nsA_v<U, false, true, true>(T, st, w, n, block, remap, s);
    case 7: return nsA_v<U, true, true, true>(T, st, w, n, block, remap, s);
    case 8: case 10: return nsA_v<U, 2, false, false>(T, st, w, n, block, remap, s);
    case 12: case 14: return nsA_v<U, 2, false, true>(T, st, w, n, block, remap, s);
    case 16: return nsA_v<U, 0, 2, false>(T, st, w, n, block, remap, s);
    case 24: return nsA_v<U, 2, 2, false>(T, st, w, n, block, remap, s);
    case 32: return nsA_v<U, 3, 0, false>(T, st, w, n, block, remap, s);
    case 36: return nsA_v<U, 3, 0, true>(T, st, w, n, block, remap, s);
    case 48: return nsA_v<U, 3, 2, false>(T, st, w, n, block, remap, s);
    default: return nsA_v<U, false, false, false>(T, st, w, n, block, remap, s);
  }
}

template <int U>
hipError_t nsB_u(const InstTable& T, const FastSet& st, const FastWork& w, int n, int block, int remap, int mode,
                 int walls, hipStream_t s) {
  switch (mode) {
    case 1: return nsB_v<U, true, false, false>(T, st, w, n, block, remap, walls, s);
    case 2: return nsB_v<U, false, true, false>(T, st, w, n, block, remap, walls, s);
    case 3: return nsB_v<U, true, true, false>(T, st, w, n, block, remap, walls, s);
    case 4: return nsB_v<U, false, false, true>(T, st, w, n, block, remap, walls, s);
    case 5: return nsB_v<U, true, false, true>(T, st, w, n, block, remap, walls, s);
    case 6: return nsB_v<U, false, true, true>(T, st, w, n, block, remap, walls, s);
    case 7: return nsB_v<U, true, true, true>(T, st, w, n, block, remap, walls, s);
    case 8: case 10: return nsB_v<U, 2, false, false>(T, st, w, n, block, remap, walls, s);
    case 12: case 14: return nsB_v<U, 2, false, true>(T, st, w, n, block, remap, walls, s);
    case 16: return nsB_v<U, 0, 2, false>(T, st, w, n, block, remap, walls, s);
    case 24: return nsB_v<U, 2, 2, false>(T, st, w, n, block, remap, walls, s);
    case 32: return nsB_v<U, 3, 0, false>(T, st, w, n, block, remap, walls, s);
    case 36: return nsB_v<U, 3, 0, true>(T, st, w, n, block, remap, walls, s);
    case 48: return nsB_v<U, 3, 2, false>(T, st, w, n, block, remap, walls, s);
    default: return nsB_v<U, false, false, false>(T, st, w, n, block, remap, walls, s);
  }
}

int norm_block(int b) { return (b >= 64 && b <= 256 && b % 64 == 0) ? b : 256; }
int mode_of(const sph_tuning& tu) {
  // row staging runs the plain loop (with or without prefetch) or mask groups;
  // mask groups run unstaged, plane- or row-staged, with their own
  // next-neighbour loads
  if (tu.nbr_lds == 3) return 32 | (tu.nbr_compact == 2 ? 16 : (tu.nbr_prefetch ? 4 : 0));
  if (tu.nbr_compact == 2) return 16 | (tu.nbr_lds == 2 ? 8 : 0);
  return (tu.nbr_lds == 2 ? 8 : tu.nbr_lds ? 1 : 0) | (tu.nbr_compact ? 2 : 0) | (tu.nbr_prefetch ? 4 : 0);
}

template <int U>
hipError_t nbrlist_u(const InstTable& T, const FastSet& st, const FastWork& w, int n, int block, int remap,
                     hipStream_t s) {
  const dim3 g(blocks_for(n, block)), bl(block);
  if (T.dim == 3)
    hipLaunchKernelGGL((nbrlist_kernel<3, U>), g, bl, 0, s, T, (const float4*)st.PR, (const int*)st.TY,
                       (const int*)w.keys, (const int*)w.cellStart, (const int*)w.cellEnd, w.NL, w.NC, w.nl_cap, n,
                       remap);
  else
    hipLaunchKernelGGL((nbrlist_kernel<2, U>), g, bl, 0, s, T, (const float4*)st.PR, (const int*)st.TY,
                       (const int*)w.keys, (const int*)w.cellStart, (const int*)w.cellEnd, w.NL, w.NC, w.nl_cap, n,
                       remap);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_nbrlist_fast(const InstTable& T, const FastSet& st, const FastWork& w, int n, const sph_tuning& tu,
                               hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const int b = norm_block(tu.nbr_block), rm = tu.xcd_remap;
  switch (tu.nbr_unroll) {
    case 1: return nbrlist_u<1>(T, st, w, n, b, rm, s);
    case 2: return nbrlist_u<2>(T, st, w, n, b, rm, s);
    case 8: return nbrlist_u<8>(T, st, w, n, b, rm, s);
    default: return nbrlist_u<4>(T, st, w, n, b, rm, s);
  }
}

namespace {

hipError_t nsA_list(const InstTable& T, const FastSet& st, const FastWork& w, int n, int block, int remap,
                    hipStream_t s) {
  const dim3 g(blocks_for(n, block)), bl(block);
  if (T.dim == 3)
    hipLaunchKernelGGL(nsA_list_kernel<3>, g, bl, 0, s, T, st.PR, st.VP, (const int*)st.TY, w.DV, (const int*)w.keys,
                       (const int*)w.cellStart, (const int*)w.cellEnd, (const int*)w.cellFluid, (const int*)w.NL,
                       (const int*)w.NC, w.nl_cap, n, remap);
  else
    hipLaunchKernelGGL(nsA_list_kernel<2>, g, bl, 0, s, T, st.PR, st.VP, (const int*)st.TY, w.DV, (const int*)w.keys,
                       (const int*)w.cellStart, (const int*)w.cellEnd, (const int*)w.cellFluid, (const int*)w.NL,
                       (const int*)w.NC, w.nl_cap, n, remap);
  return hipGetLastError();
}

hipError_t nsB_list(const InstTable& T, const FastSet& st, const FastWork& w, int n, int block, int remap,
                    hipStream_t s) {
  const dim3 g(blocks_for(n, block)), bl(block);
  if (T.dim == 3)
    hipLaunchKernelGGL(nsB_list_kernel<3>, g, bl, 0, s, T, (const float4*)st.PR, (const float4*)st.VP,
                       (const int*)st.TY, (const float*)w.DV, (const int*)w.keys, (const int*)w.cellStart,
                       (const int*)w.cellEnd, (const int*)w.NL, (const int*)w.NC, w.RATE, w.nl_cap, n, remap);
  else
    hipLaunchKernelGGL(nsB_list_kernel<2>, g, bl, 0, s, T, (const float4*)st.PR, (const float4*)st.VP,
                       (const int*)st.TY, (const float*)w.DV, (const int*)w.keys, (const int*)w.cellStart,
                       (const int*)w.cellEnd, (const int*)w.NL, (const int*)w.NC, w.RATE, w.nl_cap, n, remap);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_nsA_fast(const InstTable& T, const FastSet& st, const FastWork& w, int n, const sph_tuning& tu,
                           hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const int b = norm_block(tu.nbr_block), rm = tu.xcd_remap, md = mode_of(tu);
  if (tu.nbr_list) return nsA_list(T, st, w, n, b, rm, s);
  switch (tu.nbr_unroll) {
    case 1: return nsA_u<1>(T, st, w, n, b, rm, md, s);
    case 2: return nsA_u<2>(T, st, w, n, b, rm, md, s);
    case 8: return nsA_u<8>(T, st, w, n, b, rm, md, s);
    default: return nsA_u<4>(T, st, w, n, b, rm, md, s);
  }
}

hipError_t launch_nsB_fast(const InstTable& T, const FastSet& st, const FastWork& w, int n, const sph_tuning& tu,
                           hipStream_t s, bool walls) {
  if (n <= 0) return hipSuccess;
  const int b = norm_block(tu.nbr_block), rm = tu.xcd_remap, md = mode_of(tu), wl = walls ? 1 : 0;
  if (tu.nbr_list) {
    if (walls) return hipErrorInvalidValue;  // the list kernels serve fluid targets only
    return nsB_list(T, st, w, n, b, rm, s);
  }
  switch (tu.nbr_unroll) {
    case 1: return nsB_u<1>(T, st, w, n, b, rm, md, wl, s);
    case 2: return nsB_u<2>(T, st, w, n, b, rm, md, wl, s);
    case 8: return nsB_u<8>(T, st, w, n, b, rm, md, wl, s);
    default: return nsB_u<4>(T, st, w, n, b, rm, md, wl, s);
  }
}

}  // namespace sph
