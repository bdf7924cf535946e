// Fine-row neighbour loops (tuning nbr_lds 4): kernels A and B
// (opt_src_v1/equation/ns.cu:30-134, :137-230) on a cell grid of any stencil
// reach K (K cells on each side of the target's cell; K = 1 on the
// reference's 2h grid, K = 2 on a grid of cell edge h).
//
// On a fine grid most of a (2K+1)^3 cell block lies outside the support, so
// each target visits, per stencil row (dz, dy), only the cells of that row its
// support can reach: the row is culled when its y-z box is beyond the cut-off
// and otherwise trimmed in x to the cells within sqrt(rr_cut - d_yz^2).  The
// reference's last version searches a grid of one spacing with a fixed +-3
// row range (opt_src_v4/equation/ns.cu:62-93); here the range follows the
// target, which on a grid of edge h leaves about 185 candidates per target in
// 3-D instead of about 350 on the 2h grid.
//
// One wave (one-wave block) per 64 consecutive sorted targets.  A wave serves
// its targets group by group: the lanes of one instance, one cell row (z, y)
// and a window of at most 63 - 2K cells in x (waterfall; one group when the
// 64 targets share a row, the common case).  Per stencil row the group
//   1. reads the cell table of the window's cells, one lane per cell, and
//      turns it into dense offsets with a wave suffix minimum: off[t] is the
//      first sorted index at or after window cell t (empty cells take the
//      next non-empty cell's start, the tail the last cell's end);
//   2. gives each lane [off[x0], off[x1 + 1]) for its trimmed cells x0..x1
//      (cross-lane reads, no table loads per lane);
//   3. copies the union of the lanes' ranges, with the neighbour data the
//      pair bodies read, into LDS in chunks of FINE_CAP entries, and runs the
//      pair loop there.
// Rows are visited z outer, y inner, and a lane's candidates in ascending
// sorted index, which is the oracle's order on the same grid; culled cells
// hold no pair inside the support, so the pair set is the oracle's too.
#include "sph_pair.h"

namespace sph {

using namespace fast;

namespace {

inline unsigned blocks_for(long long n, int b) { return (unsigned)((n + b - 1) / b); }

constexpr int FINE_CAP = 128;  // staged candidates per chunk
constexpr int FINE_PAD = 8;    // entries read past a chunk's end by the unrolled candidate loads

// Candidate visit of one group of lanes (see above).  Called by all 64 lanes
// of the wave; lanes outside the group (on = false) get empty ranges and only
// help with the table reads and the copies.
template <int DIM, int U, bool PF, bool KA>
struct FineVisit {
  // wave-uniform
  int X, Y, Z, XY, coff, K, ucy, ucz, cA, cB;
  float gs, inv_gs, gxmin, cut, mx;
  // per lane
  bool on, inside;
  int cx;
  float px, yl, yr, zl, zr;  // target x; its distances to its cell's y/z faces, shrunk (0 outside the grid)
  const int* __restrict__ cs;
  const int* __restrict__ ce;
  const int* __restrict__ fc;  // kernel A: cells holding fluid
  const float4* PR;
  const float4* VP;
  const float* DV;
  float4* tp;
  float4* tv;
  float* td;

  template <class Accept, class Extra, class Pair>
  __device__ __forceinline__ void operator()(Accept&& accept, Extra&& extra, Pair&& pair) {
    const int lane = threadIdx.x & 63;
    const int L = cB - cA + 1;  // window cells (L + 1 <= 64)
#pragma unroll 1
    for (int dz = -K; dz <= K; ++dz) {
      const int z = ucz + dz;
      if (z < 0 || z >= Z) continue;
      const float fz = dz < 0 ? zl + (float)(-dz - 1) * gs : (dz > 0 ? zr + (float)(dz - 1) * gs : 0.0f);
#pragma unroll 1
      for (int dy = (DIM == 3 ? -K : 0); dy <= (DIM == 3 ? K : 0); ++dy) {
        const int y = ucy + dy;
        if (y < 0 || y >= Y) continue;
        const float fy = dy < 0 ? yl + (float)(-dy - 1) * gs : (dy > 0 ? yr + (float)(dy - 1) * gs : 0.0f);
        const float dyz = inside ? fz * fz + fy * fy : 0.0f;
        const bool lon = on && dyz <= cut;
        if (__ballot(lon) == 0ull) continue;
        // 1. cell table of the window, dense offsets
        const int rb = coff + z * XY + y * X + cA;
        int st = -1, en = -1;
        if (lane < L) {
          st = cs[rb + lane];
          if (KA && st != -1 && fc[rb + lane] == 0) st = -1;  // kernel A pairs with fluid only
          if (st != -1) en = ce[rb + lane];
        }
        const int E = wave_max(en);  // end of the window's last (fluid) cell
        if (E < 0) continue;
        int v = st != -1 ? st : 0x7fffffff;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
          const int w = __shfl_down(v, o);
          if (lane + o < 64) v = min(v, w);
        }
        const int off = min(v, E);
        // 2. this lane's cells and range
        int x0 = cx - K, x1 = cx + K;
        if (inside) {
          const float w = __builtin_sqrtf(fmaxf(cut - dyz, 0.0f));
          x0 = max(x0, (int)floorf((px - w - mx - gxmin) * inv_gs));
          x1 = min(x1, (int)floorf((px + w + mx - gxmin) * inv_gs));
        }
        x0 = max(x0, cA);
        x1 = min(x1, cB);
        const int a = __shfl(off, min(max(x0 - cA, 0), 63));
        const int b = __shfl(off, min(max(x1 + 1 - cA, 0), 63));
        int s = 0, e = 0;
        if (lon && x0 <= x1) {
          s = a;
          e = b;
        }
        // 3. stage the union chunk by chunk, run the pair loop from LDS
        const int us = wave_min(s < e ? s : 0x7fffffff);
        const int ue = wave_max(s < e ? e : -1);
#pragma unroll 1
        for (int c0 = us; c0 < ue; c0 += FINE_CAP) {
          const int len = min(FINE_CAP, ue - c0);
          __syncthreads();  // the previous chunk is consumed
          {
            // FINE_CAP = 2 x 64: every load is issued before the first store
            const int t0 = lane, t1 = lane + 64;
            const bool h0 = t0 < len, h1 = t1 < len;
            float4 a0, a1, b0, b1;
            float d0 = 0.0f, d1 = 0.0f;
            if (h0) {
              a0 = PR[c0 + t0];
              b0 = VP[c0 + t0];
              if (!KA) d0 = DV[c0 + t0];
            }
            if (h1) {
              a1 = PR[c0 + t1];
              b1 = VP[c0 + t1];
              if (!KA) d1 = DV[c0 + t1];
            }
            if (h0) {
              tp[t0] = a0;
              tv[t0] = b0;
              if (!KA) td[t0] = d0;
            }
            if (h1) {
              tp[t1] = a1;
              tv[t1] = b1;
              if (!KA) td[t1] = d1;
            }
          }
          __syncthreads();
          const int js = max(s, c0), je = min(e, c0 + len);
          for (int j0 = js; j0 < je; j0 += U) {
            const int t = j0 - c0;
            float4 pjs[U];
            decltype(extra(0)) es[U];
#pragma unroll
            for (int u = 0; u < U; ++u) pjs[u] = tp[t + u];
            if constexpr (PF) {
#pragma unroll
              for (int u = 0; u < U; ++u) es[u] = extra.from(tv, td, t + u);
            }
            bool acc[U];
#pragma unroll
            for (int u = 0; u < U; ++u) acc[u] = j0 + u < je && accept(j0 + u, pjs[u]);
            if constexpr (!PF) {
#pragma unroll
              for (int u = 0; u < U; ++u)
                if (acc[u]) es[u] = extra.from(tv, td, t + u);
            }
#pragma unroll
            for (int u = 0; u < U; ++u)
              if (acc[u]) pair(j0 + u, pjs[u], es[u]);
          }
        }
      }
    }
  }
};

// Splits the wave's targets into groups (one instance, one cell row, an x
// window of at most 63 - 2K cells) and calls body(P, visit, in_group) for
// each with every lane of the wave.  want: the lane has a target to serve.
template <int DIM, int U, bool PF, bool KA, class Body>
__device__ __forceinline__ void fine_groups(const InstTable& T, int inst, int key, const float4& p, bool want,
                                            const float4* PR, const float4* VP, const float* DV,
                                            const int* __restrict__ cs, const int* __restrict__ ce,
                                            const int* __restrict__ fc, float4* tp, float4* tv, float* td,
                                            Body&& body) {
  bool todo = want;
  for (;;) {
    const unsigned long long pend = __ballot(todo);
    if (pend == 0ull) break;
    const int first = __builtin_ctzll(pend);
    const int u = __builtin_amdgcn_readlane(inst, first);
    const sph_params& P = T.P[u];
    const int coff = T.cell_off[u];
    const int X = P.gridxdim, Y = P.gridydim;
    const int hh = key - coff;  // meaningful for the lanes of instance u
    const int row = hh / X;     // z * Y + y
    const int cx = hh - row * X;
    const int urow = __builtin_amdgcn_readlane(row, first);
    const int ucx = __builtin_amdgcn_readlane(cx, first);
    const int K = stencil_reach(P.kh, P.gridSize);
    // lanes of one row come in ascending cx, so ucx is the group's smallest
    const bool grp = todo && inst == u && row == urow && cx - ucx <= 62 - 2 * K;
    const int cxhi = wave_max(grp ? cx : -1);
    FineVisit<DIM, U, PF, KA> vis;
    vis.X = X;
    vis.Y = Y;
    vis.Z = P.gridzdim;
    vis.XY = X * Y;
    vis.coff = coff;
    vis.K = K;
    vis.ucz = urow / Y;
    vis.ucy = urow - vis.ucz * Y;
    vis.cA = max(ucx - K, 0);
    vis.cB = min(cxhi + K, X - 1);
    const float gs = P.gridSize;
    vis.gs = gs;
    vis.inv_gs = 1.0f / gs;
    vis.gxmin = P.gridxmin;
    vis.cut = P.rr_cut;
    // margin of the x trim: covers the rounding of the cell bounds and of the
    // multiply by 1/gs, so no cell holding a pair inside the support is cut
    vis.mx = gs * (1.0f / 256.0f);
    vis.on = grp;
    vis.cx = cx;
    vis.px = p.x;
    const float m = gs * (1.0f / 1024.0f);
    const int Z = P.gridzdim;
    const bool inside = p.x >= P.gridxmin && p.x <= P.gridxmin + (float)X * gs && p.z >= P.gridzmin &&
                        p.z <= P.gridzmin + (float)Z * gs &&
                        (DIM == 2 || (p.y >= P.gridymin && p.y <= P.gridymin + (float)Y * gs));
    vis.inside = inside;
    const float ly = P.gridymin + (float)vis.ucy * gs, lz = P.gridzmin + (float)vis.ucz * gs;
    vis.yl = inside && DIM == 3 ? fmaxf(p.y - ly - m, 0.0f) : 0.0f;
    vis.yr = inside && DIM == 3 ? fmaxf(ly + gs - p.y - m, 0.0f) : 0.0f;
    vis.zl = inside ? fmaxf(p.z - lz - m, 0.0f) : 0.0f;
    vis.zr = inside ? fmaxf(lz + gs - p.z - m, 0.0f) : 0.0f;
    vis.cs = cs;
    vis.ce = ce;
    vis.fc = fc;
    vis.PR = PR;
    vis.VP = VP;
    vis.DV = DV;
    vis.tp = tp;
    vis.tv = tv;
    vis.td = td;
    body(P, vis, grp);
    todo = todo && !grp;
  }
}

template <int DIM, int U, bool PF>
__global__ __launch_bounds__(64) void nsA_fine_kernel(InstTable T, float4* PR, float4* VP, const int* __restrict__ TY,
                                                      float* __restrict__ DV, const int* __restrict__ keys,
                                                      const int* __restrict__ cs, const int* __restrict__ ce,
                                                      const int* __restrict__ fc, int n, int remap) {
  __shared__ float4 tp[FINE_CAP + FINE_PAD];
  __shared__ float4 tv[FINE_CAP + FINE_PAD];
  __shared__ float td[1];  // kernel A reads no neighbour's dofv
  const int i = tile_id(remap) * 64 + threadIdx.x;
  const bool valid = i < n;
  const int ic = valid ? i : n - 1;
  const float4 pi = PR[ic];
  const float4 vi = VP[ic];
  const int key = keys[ic];
  fine_groups<DIM, U, PF, true>(T, inst_of(TY[ic]), key, pi, valid, PR, VP, nullptr, cs, ce, fc, tp, tv, td,
                                [&](const sph_params& P, auto visit, bool grp) {
                                  nsA_target<DIM>(P, i, pi, vi, visit, PR, VP, DV, grp);
                                });
}

template <int DIM, int U, bool PF>
__global__ __launch_bounds__(64) void nsB_fine_kernel(InstTable T, const float4* __restrict__ PR,
                                                      const float4* __restrict__ VP, const int* __restrict__ TY,
                                                      const float* __restrict__ DV, const int* __restrict__ keys,
                                                      const int* __restrict__ cs, const int* __restrict__ ce,
                                                      float4* __restrict__ RATE, int n, int remap) {
  __shared__ float4 tp[FINE_CAP + FINE_PAD];
  __shared__ float4 tv[FINE_CAP + FINE_PAD];
  __shared__ float td[FINE_CAP + FINE_PAD];
  const int i = tile_id(remap) * 64 + threadIdx.x;
  const bool valid = i < n;
  const int ic = valid ? i : n - 1;
  const float4 pi = PR[ic];
  const bool fluid = pi.w > 0.0f;
  if (valid && !fluid) RATE[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  const float4 vi = VP[ic];
  const float dvi = DV[ic];
  const int key = keys[ic];
  fine_groups<DIM, U, PF, false>(T, inst_of(TY[ic]), key, pi, valid && fluid, PR, VP, DV, cs, ce, nullptr, tp, tv,
                                 td, [&](const sph_params& P, auto visit, bool grp) {
                                   nsB_target<DIM>(P, i, pi, vi, dvi, visit, VP, DV, RATE, grp);
                                 });
}

// Candidate / in-support pair counts of the fine-row visits (statistics
// only): candA over every target's kernel-A rows, hitsA its in-support fluid
// neighbours, candB / hitsB over fluid targets' kernel-B rows.  Candidates
// exclude the target itself, as the 2h-grid statistics do.
template <int DIM>
__global__ __launch_bounds__(64) void pair_stats_fine_kernel(InstTable T, const float4* __restrict__ PR,
                                                             const float4* __restrict__ VP,
                                                             const float* __restrict__ DV,
                                                             const int* __restrict__ TY, const int* __restrict__ keys,
                                                             const int* __restrict__ cs, const int* __restrict__ ce,
                                                             const int* __restrict__ fc,
                                                             unsigned long long* __restrict__ out, int n) {
  __shared__ float4 tp[FINE_CAP + FINE_PAD];
  __shared__ float4 tv[FINE_CAP + FINE_PAD];
  __shared__ float td[FINE_CAP + FINE_PAD];
  const int i = blockIdx.x * 64 + threadIdx.x;
  const bool valid = i < n;
  const int ic = valid ? i : n - 1;
  const float4 pi = PR[ic];
  const bool fluid = pi.w > 0.0f;
  const int key = keys[ic];
  const int inst = inst_of(TY[ic]);
  unsigned long long cA = 0, hA = 0, c = 0, hB = 0;
  fine_groups<DIM, 1, false, true>(T, inst, key, pi, valid, PR, VP, nullptr, cs, ce, fc, tp, tv, td,
                                   [&](const sph_params&, auto visit, bool) {
                                     visit(
                                         [&](int j, const float4&) {
                                           cA += j != i;
                                           return false;
                                         },
                                         ExtraNone{}, [&](int, const float4&, const NoExtra&) {});
                                   });
  fine_groups<DIM, 1, false, false>(T, inst, key, pi, valid, PR, VP, DV, cs, ce, nullptr, tp, tv, td,
                                    [&](const sph_params& P, auto visit, bool) {
                                      const float cut = P.rr_cut;
                                      visit(
                                          [&](int j, const float4& pj) {
                                            if (j == i) return false;
                                            ++c;
                                            float dx, dy, dz;
                                            if (geom<DIM>(pi, pj, dx, dy, dz) <= cut) {
                                              ++hB;
                                              if (pj.w > 0.0f) ++hA;
                                            }
                                            return false;
                                          },
                                          ExtraNone{}, [&](int, const float4&, const NoExtra&) {});
                                    });
  unsigned long long v[4] = {cA, hA, fluid ? c : 0ull, fluid ? hB : 0ull};
  for (int k = 0; k < 4; ++k) {
    unsigned long long x = v[k];
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
    if (threadIdx.x == 0) atomicAdd(&out[k], x);
  }
}

template <int U, bool PF>
hipError_t nsA_fine_v(const InstTable& T, const FastSet& st, const FastWork& w, int n, int remap, hipStream_t s) {
  const dim3 g(blocks_for(n, 64)), b(64);
  if (T.dim == 3)
    hipLaunchKernelGGL((nsA_fine_kernel<3, U, PF>), g, b, 0, s, T, st.PR, st.VP, (const int*)st.TY, w.DV,
                       (const int*)w.keys, (const int*)w.cellStart, (const int*)w.cellEnd, (const int*)w.cellFluid,
                       n, remap);
  else
    hipLaunchKernelGGL((nsA_fine_kernel<2, U, PF>), g, b, 0, s, T, st.PR, st.VP, (const int*)st.TY, w.DV,
                       (const int*)w.keys, (const int*)w.cellStart, (const int*)w.cellEnd, (const int*)w.cellFluid,
                       n, remap);
  return hipGetLastError();
}

template <int U, bool PF>
hipError_t nsB_fine_v(const InstTable& T, const FastSet& st, const FastWork& w, int n, int remap, hipStream_t s) {
  const dim3 g(blocks_for(n, 64)), b(64);
  if (T.dim == 3)
    hipLaunchKernelGGL((nsB_fine_kernel<3, U, PF>), g, b, 0, s, T, (const float4*)st.PR, (const float4*)st.VP,
                       (const int*)st.TY, (const float*)w.DV, (const int*)w.keys, (const int*)w.cellStart,
                       (const int*)w.cellEnd, w.RATE, n, remap);
  else
    hipLaunchKernelGGL((nsB_fine_kernel<2, U, PF>), g, b, 0, s, T, (const float4*)st.PR, (const float4*)st.VP,
                       (const int*)st.TY, (const float*)w.DV, (const int*)w.keys, (const int*)w.cellStart,
                       (const int*)w.cellEnd, w.RATE, n, remap);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_nsA_fine(const InstTable& T, const FastSet& st, const FastWork& w, int n, const sph_tuning& tu,
                           hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const int rm = tu.xcd_remap ? 1 : 0;
  const bool pf = tu.nbr_prefetch != 0;
  switch (tu.nbr_unroll) {
    case 1: return pf ? nsA_fine_v<1, true>(T, st, w, n, rm, s) : nsA_fine_v<1, false>(T, st, w, n, rm, s);
    case 4: return pf ? nsA_fine_v<4, true>(T, st, w, n, rm, s) : nsA_fine_v<4, false>(T, st, w, n, rm, s);
    default: return pf ? nsA_fine_v<2, true>(T, st, w, n, rm, s) : nsA_fine_v<2, false>(T, st, w, n, rm, s);
  }
}

hipError_t launch_nsB_fine(const InstTable& T, const FastSet& st, const FastWork& w, int n, const sph_tuning& tu,
                           hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const int rm = tu.xcd_remap ? 1 : 0;
  const bool pf = tu.nbr_prefetch != 0;
  switch (tu.nbr_unroll) {
    case 1: return pf ? nsB_fine_v<1, true>(T, st, w, n, rm, s) : nsB_fine_v<1, false>(T, st, w, n, rm, s);
    case 4: return pf ? nsB_fine_v<4, true>(T, st, w, n, rm, s) : nsB_fine_v<4, false>(T, st, w, n, rm, s);
    default: return pf ? nsB_fine_v<2, true>(T, st, w, n, rm, s) : nsB_fine_v<2, false>(T, st, w, n, rm, s);
  }
}

hipError_t launch_pair_stats_fine(const InstTable& T, const FastSet& st, const FastWork& w, unsigned long long* out4,
                                  int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const dim3 g(blocks_for(n, 64)), b(64);
  if (T.dim == 3)
    hipLaunchKernelGGL(pair_stats_fine_kernel<3>, g, b, 0, s, T, (const float4*)st.PR, (const float4*)st.VP,
                       (const float*)w.DV, (const int*)st.TY, (const int*)w.keys, (const int*)w.cellStart,
                       (const int*)w.cellEnd, (const int*)w.cellFluid, out4, n);
  else
    hipLaunchKernelGGL(pair_stats_fine_kernel<2>, g, b, 0, s, T, (const float4*)st.PR, (const float4*)st.VP,
                       (const float*)w.DV, (const int*)st.TY, (const int*)w.keys, (const int*)w.cellStart,
                       (const int*)w.cellEnd, (const int*)w.cellFluid, out4, n);
  return hipGetLastError();
}

}  // namespace sph
