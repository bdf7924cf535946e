// Fine-row neighbour loops (tuning nbr_lds 4): kernels A and B
// (opt_src_v1/equation/ns.cu:30-134, :137-230) on a cell grid of any stencil
// reach K (K cells on each side of the target's cell; K = 1 on the
// reference's 2h grid, K = 2 on a grid of cell edge h), and the dense cell
// table they read.
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
// Cell table: first[c] = the first sorted index whose key is >= c, for every
// cell c of the context and one sentinel (= n).  The x-adjacent cells of a
// row are adjacent keys, so the candidates of cells x0..x1 of a row are the
// sorted range [first[row + x0], first[row + x1 + 1]) -- two loads per lane
// and row, whatever the cells hold (the reference's fill_cell builds a dense
// form of its table too, opt_src_v4/lib/mesh.cu:204-253).  Kernel A pairs with
// fluid only: a bit per cell marks cells that hold fluid, and a group of wall
// targets whose stencil window holds none skips its visit (most of the walls:
// those away from the water).
//
// One wave (one-wave block) per 64 consecutive sorted targets, served group
// by group: the lanes of one instance, one cell row (z, y) and at most
// GROUP_WIN cells of it (waterfall; one group when the 64 targets share a
// row, the common case).  Per stencil row
// the group copies the union of its lanes' ranges, with the neighbour data
// the pair bodies read, into LDS in chunks of FINE_CAP entries and runs the
// pair loop there; the next row's range loads are issued before the current
// row's copy, so a row costs one memory round trip.  Rows are visited z
// outer, y inner, and a lane's candidates in ascending sorted index, which is
// the oracle's order on the same grid; culled cells hold no pair inside the
// support, so the pair set is the oracle's too.
#include "sph_pair.h"

extern "C" __device__ int __ockl_wfred_min_i32(int);
extern "C" __device__ int __ockl_wfred_max_i32(int);

namespace sph {

using namespace fast;

namespace {

inline unsigned blocks_for(long long n, int b) { return (unsigned)((n + b - 1) / b); }

constexpr int FINE_CAP = 128;  // staged candidates per chunk (2 per lane)
constexpr int FINE_PAD = 8;    // entries read past a chunk's end by the unrolled candidate loads
constexpr int GAP_SMALL = 32;  // longer runs of cells are filled block-wide
constexpr int GROUP_WIN = 48;  // cells of its row one group of lanes may span

static_assert(FINE_CAP == 128, "the copy moves two entries per lane");

__device__ __forceinline__ int wmin(int v) { return __ockl_wfred_min_i32(v); }
__device__ __forceinline__ int wmax(int v) { return __ockl_wfred_max_i32(v); }

// The wave's LDS accesses are processed in order; this keeps the compiler
// from moving them across the copy / pair-loop boundary of a one-wave block.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ------------------------------------------------------------ dense table
// Sorted index i (0..n) owns the cells (keys[i-1], keys[i]] (keys[-1] = -1,
// keys[n] = cells): first[c] = i for them.  The runs partition 0..cells, so
// every entry is written once.  A thread writes its run itself when it is
// short; longer runs (the empty rest of a row, empty rows) are queued in LDS
// and written by the whole block; the few runs of GAP_GIANT cells or more
// (the padding planes above and below the particles) go to a global queue
// that fill_runs_kernel writes with the whole grid.  The global counter is
// zeroed by the hash kernel of the same half-step.
constexpr int GAP_GIANT = 1 << 16;

__global__ __launch_bounds__(256) void dense_first_kernel(const int* __restrict__ keys, int n, int cells,
                                                          int* __restrict__ first, int* __restrict__ q,
                                                          int* __restrict__ qn, int qcap) {
  __shared__ int lq[3 * 256];
  __shared__ int lqn;
  if (threadIdx.x == 0) lqn = 0;
  __syncthreads();
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i <= n) {
    const int lo = i == 0 ? 0 : keys[i - 1] + 1;
    const int hi = i == n ? cells : keys[i];
    const int len = hi - lo + 1;
    if (len > 0 && len <= GAP_SMALL) {
      for (int c = lo; c <= hi; ++c) first[c] = i;
    } else if (len > GAP_SMALL) {
      int slot = -1;
      if (len >= GAP_GIANT) {
        slot = atomicAdd(qn, 1);
        if (slot < qcap) {
          q[3 * slot] = lo;
          q[3 * slot + 1] = hi;
          q[3 * slot + 2] = i;
        } else {
          slot = -1;  // cannot happen (qcap bounds the giant runs); the block writes it instead
        }
      }
      if (slot < 0) {
        const int k = atomicAdd(&lqn, 1);
        lq[3 * k] = lo;
        lq[3 * k + 1] = hi;
        lq[3 * k + 2] = i;
      }
    }
  }
  __syncthreads();
  const int m = lqn;
  for (int k = 0; k < m; ++k) {
    const int lo = lq[3 * k], hi = lq[3 * k + 1], v = lq[3 * k + 2];
    for (int c = lo + (int)threadIdx.x; c <= hi; c += blockDim.x) first[c] = v;
  }
}

// The giant runs of the queue, with the whole grid; also clears the coarse
// fluid flags for the gather that follows.
__global__ void fill_runs_kernel(int* __restrict__ first, const int* __restrict__ q, const int* __restrict__ qn,
                                 int qcap, unsigned char* __restrict__ fluidC, int ccells) {
  const int cnt = min(*qn, qcap);
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  const int stride = gridDim.x * blockDim.x;
  for (int g = 0; g < cnt; ++g) {
    const int lo = q[3 * g], hi = q[3 * g + 1], v = q[3 * g + 2];
    for (int c = lo + tid; c <= hi; c += stride) first[c] = v;
  }
  for (int c = tid; c < ccells; c += stride) fluidC[c] = 0;
}

// Coarse cell (the instances' own 2h grid, table Tc) of a cell of the grid in
// use (table T, the same or a subdivision by 2): global indices both.
__device__ __forceinline__ int coarse_cell(const sph_params& P, int coff, const sph_params& Pc, int coffc, int key) {
  const int sh = P.gridxdim == Pc.gridxdim ? 0 : 1;
  const int X = P.gridxdim, XY = P.gridxdim * P.gridydim;
  const int h = key - coff;
  const int z = h / XY, r = h - z * XY;
  const int y = r / X, x = r - y * X;
  return coffc + (z >> sh) * Pc.gridxdim * Pc.gridydim + (y >> sh) * Pc.gridxdim + (x >> sh);
}

// Gather into sorted order (as cells_gather_kernel) and flag the coarse
// cells that hold fluid (every writer stores the same value).
template <int DIM>
__global__ void gather_fluidc_kernel(InstTable T, InstTable Tc, FastSet in, FastSet out, const int* __restrict__ keys,
                                     const int* __restrict__ idx, unsigned char* __restrict__ fluidC, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int src = idx[i];
  const float4 pr = in.PR[src];
  const int ty = in.TY[src];
  out.PR[i] = pr;
  out.VP[i] = in.VP[src];
  out.X[0][i] = in.X[0][src];
  if (DIM == 3) out.X[1][i] = in.X[1][src];
  out.X[2][i] = in.X[2][src];
  out.ID[i] = in.ID[src];
  out.TY[i] = ty;
  if (pr.w > 0.0f) {
    const int key = keys[i];
    with_instance(T, inst_of(ty), [&](const sph_params& P, int coff) {
      const int u = inst_of(ty);
      fluidC[coarse_cell(P, coff, Tc.P[u], Tc.cell_off[u], key)] = 1;
    });
  }
}

// nearC[c] = 1 when a coarse cell within one cell of c (same instance) holds
// fluid.  Every cell of a fine stencil of reach K <= 2 around a target lies
// within one coarse cell of the target's coarse cell, so a wall target whose
// coarse cell has nearC = 0 has no fluid neighbour.
template <int DIM>
__global__ void near_fluid_kernel(InstTable Tc, const unsigned char* __restrict__ fluidC,
                                  unsigned char* __restrict__ nearC) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= Tc.total_cells) return;
  int lo = 0, hi = Tc.K - 1;  // instance of cell c
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (Tc.cell_off[mid] <= c)
      lo = mid;
    else
      hi = mid - 1;
  }
  const sph_params& P = Tc.P[lo];
  const int X = P.gridxdim, Y = P.gridydim, Z = P.gridzdim;
  const int h = c - Tc.cell_off[lo];
  const int z = h / (X * Y), r = h - z * X * Y;
  const int y = r / X, x = r - y * X;
  unsigned char hit = 0;
  for (int dz = -1; dz <= 1; ++dz)
    for (int dy = (DIM == 3 ? -1 : 0); dy <= (DIM == 3 ? 1 : 0); ++dy)
      for (int dx = -1; dx <= 1; ++dx) {
        const int zz = z + dz, yy = y + dy, xx = x + dx;
        if (zz < 0 || zz >= Z || yy < 0 || yy >= Y || xx < 0 || xx >= X) continue;
        hit |= fluidC[c + dz * X * Y + dy * X + dx];
      }
  nearC[c] = hit;
}

// ------------------------------------------------------------ visit
// This lane's candidate range in one stencil row, loads issued and not yet
// waited for.
struct Pending {
  int s, e;
};

// Candidate visit of one group of lanes (see above).  Called by all 64 lanes
// of the wave; lanes outside the group (on = false) get empty ranges and only
// help with the copies.
template <int DIM, int U, bool PF, bool KA>
struct FineVisit {
  // wave-uniform
  int X, Y, Z, XY, coff, K, ucy, ucz, n;
  float gs, inv_gs, gxmin, cut, mx;
  // per lane
  bool on, inside;
  int cx;
  float px, yl, yr, zl, zr;  // target x; its distances to its cell's y/z faces, shrunk (0 outside the grid)
  const int* __restrict__ first;
  const float4* PR;
  const float4* VP;
  const float* DV;
  float4* tp;
  float4* tv;
  float* td;

  // Issues the loads of this lane's range in stencil row (dz, dy).
  __device__ __forceinline__ Pending issue(int dz, int dy) const {
    Pending p{0, 0};
    const int z = ucz + dz, y = ucy + dy;
    if (z < 0 || z >= Z || y < 0 || y >= Y) return p;  // wave-uniform
    const float fz = dz < 0 ? zl + (float)(-dz - 1) * gs : (dz > 0 ? zr + (float)(dz - 1) * gs : 0.0f);
    const float fy = dy < 0 ? yl + (float)(-dy - 1) * gs : (dy > 0 ? yr + (float)(dy - 1) * gs : 0.0f);
    const float dyz = inside ? fz * fz + fy * fy : 0.0f;
    int x0 = cx - K, x1 = cx + K;
    if (inside) {
      const float w = __builtin_sqrtf(fmaxf(cut - dyz, 0.0f));
      x0 = max(x0, (int)floorf((px - w - mx - gxmin) * inv_gs));
      x1 = min(x1, (int)floorf((px + w + mx - gxmin) * inv_gs));
    }
    x0 = max(x0, 0);
    x1 = min(x1, X - 1);
    if (!on || dyz > cut || x0 > x1) return p;
    const int rb = coff + z * XY + y * X;
    // table entries lie in 0..n; the clamp keeps every staged load inside the
    // particle arrays even for a table that was not built for this sort
    p.s = min(max(first[rb + x0], 0), n);
    p.e = min(max(first[rb + x1 + 1], 0), n);
    return p;
  }

  // Copies the union chunk [c0, c0 + len) (len <= FINE_CAP) into LDS, every
  // load issued before the first store (entries past the chunk re-read its
  // first one and are not stored).
  __device__ __forceinline__ void stage(int c0, int len) const {
    const int lane = threadIdx.x & 63;
    const int t0 = lane, t1 = lane + 64;
    const bool h0 = t0 < len, h1 = t1 < len;
    const unsigned o0 = (unsigned)(c0 + (h0 ? t0 : 0)), o1 = (unsigned)(c0 + (h1 ? t1 : 0));
    const float4 a0 = ld16(PR, o0 * 16u), a1 = ld16(PR, o1 * 16u);
    const float4 b0 = ld16(VP, o0 * 16u), b1 = ld16(VP, o1 * 16u);
    float d0 = 0.0f, d1 = 0.0f;
    if (!KA) {
      d0 = ld4(DV, o0 * 4u);
      d1 = ld4(DV, o1 * 4u);
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

  // Pair loop of this lane over its range [s, e) within the staged chunk
  // [c0, c0 + len).
  template <class Accept, class Extra, class Pair>
  __device__ __forceinline__ void run_chunk(int s, int e, int c0, int len, Accept& accept, Extra& extra,
                                            Pair& pair) const {
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
      for (int u = 0; u < U; ++u) acc[u] = (j0 + u < je) & accept(j0 + u, pjs[u]);
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

  // Rows in the oracle's order (z outer); the next row's range loads are in
  // flight while the current row is copied and visited.
  template <class Accept, class Extra, class Pair>
  __device__ __forceinline__ void operator()(Accept&& accept, Extra&& extra, Pair&& pair) {
    if (__ballot(on) == 0ull) return;  // e.g. wall targets with no fluid in reach (kernel A)
    int dz = -K, dy = DIM == 3 ? -K : 0;
    Pending nx = issue(dz, dy);
#pragma unroll 1
    for (;;) {
      const int s = nx.s, e = nx.e;
      int nz = dz, ny = dy;
      if (DIM == 3 && ny < K) {
        ++ny;
      } else {
        ++nz;
        ny = DIM == 3 ? -K : 0;
      }
      const bool more = nz <= K;
      if (more) nx = issue(nz, ny);
      if (__ballot(s < e) != 0ull) {  // a row culled for every lane costs no copy
        const int us = wmin(s < e ? s : 0x7fffffff);
        const int ue = wmax(s < e ? e : -1);
#pragma unroll 1
        for (int c0 = us; c0 < ue; c0 += FINE_CAP) {
          const int len = min(FINE_CAP, ue - c0);
          stage(c0, len);
          wave_lds_sync();
          run_chunk(s, e, c0, len, accept, extra, pair);
          wave_lds_sync();  // every lane is done with the chunk before the next copy
        }
      }
      if (!more) break;
      dz = nz;
      dy = ny;
    }
  }
};

// Splits the wave's targets into groups (one instance, one cell row, at most
// GROUP_WIN cells of it) and
// calls body(P, visit, in_group) for each with every lane of the wave.
// want: the lane has a target to serve.
template <int DIM, int U, bool PF, bool KA, class Body>
__device__ __forceinline__ void fine_groups(const InstTable& T, int n, int inst, int key, const float4& p, bool want,
                                            const float4* PR, const float4* VP, const float* DV,
                                            const int* __restrict__ first, float4* tp, float4* tv, float* td,
                                            Body&& body) {
  bool todo = want;
  for (;;) {
    const unsigned long long pend = __ballot(todo);
    if (pend == 0ull) break;
    const int lead = __builtin_ctzll(pend);
    const int u = __builtin_amdgcn_readlane(inst, lead);
    const sph_params& P = T.P[u];
    const int coff = T.cell_off[u];
    const int X = P.gridxdim, Y = P.gridydim, Z = P.gridzdim;
    const int hh = key - coff;  // meaningful for the lanes of instance u
    const int row = hh / X;     // z * Y + y
    const int cx = hh - row * X;
    const int urow = __builtin_amdgcn_readlane(row, lead);
    const int ucx = __builtin_amdgcn_readlane(cx, lead);  // lanes of a row come in ascending cx
    // a group spans at most GROUP_WIN cells of its row: a sparse row (the two
    // end walls of a tank, spray) splits into compact groups instead of one
    // whose copies would span everything in between
    bool grp = todo && inst == u && row == urow && cx - ucx < GROUP_WIN;
    const int K = stencil_reach(P.kh, P.gridSize);
    const int ucz = urow / Y, ucy = urow - ucz * Y;
    FineVisit<DIM, U, PF, KA> vis;
    vis.X = X;
    vis.Y = Y;
    vis.Z = Z;
    vis.XY = X * Y;
    vis.coff = coff;
    vis.n = n;
    vis.K = K;
    vis.ucz = ucz;
    vis.ucy = ucy;
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
    const bool inside = p.x >= P.gridxmin && p.x <= P.gridxmin + (float)X * gs && p.z >= P.gridzmin &&
                        p.z <= P.gridzmin + (float)Z * gs &&
                        (DIM == 2 || (p.y >= P.gridymin && p.y <= P.gridymin + (float)Y * gs));
    vis.inside = inside;
    const float ly = P.gridymin + (float)vis.ucy * gs, lz = P.gridzmin + (float)vis.ucz * gs;
    vis.yl = inside && DIM == 3 ? fmaxf(p.y - ly - m, 0.0f) : 0.0f;
    vis.yr = inside && DIM == 3 ? fmaxf(ly + gs - p.y - m, 0.0f) : 0.0f;
    vis.zl = inside ? fmaxf(p.z - lz - m, 0.0f) : 0.0f;
    vis.zr = inside ? fmaxf(lz + gs - p.z - m, 0.0f) : 0.0f;
    vis.first = first;
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

// ------------------------------------------------------------ kernels
// Whether a target of kernel A has anything to sum: fluid targets always; a
// wall target only when fluid lies within one coarse cell of its own.  A wall
// target without gets p = 0 and dofv = 0 here, the value of the empty sums.
__device__ __forceinline__ bool kernelA_wants(const InstTable& T, const InstTable& Tc, int inst, int key,
                                              const float4& pi, const unsigned char* __restrict__ nearC) {
  if (pi.w > 0.0f) return true;
  bool near = false;
  with_instance(T, inst, [&](const sph_params& P, int coff) {
    near = nearC[coarse_cell(P, coff, Tc.P[inst], Tc.cell_off[inst], key)] != 0;
  });
  return near;
}

template <int DIM, int U, bool PF>
__global__ __launch_bounds__(64) void nsA_fine_kernel(InstTable T, InstTable Tc, float4* PR, float4* VP,
                                                      const int* __restrict__ TY, float* __restrict__ DV,
                                                      const int* __restrict__ keys, const int* __restrict__ first,
                                                      const unsigned char* __restrict__ nearC, int n, int remap) {
  __shared__ float4 tp[FINE_CAP + FINE_PAD];
  __shared__ float4 tv[FINE_CAP + FINE_PAD];
  __shared__ float td[1];  // kernel A reads no neighbour's dofv
  const int i = tile_id(remap) * 64 + threadIdx.x;
  const bool valid = i < n;
  const int ic = valid ? i : n - 1;
  const float4 pi = PR[ic];
  const float4 vi = VP[ic];
  const int key = keys[ic];
  const int inst = inst_of(TY[ic]);
  const bool want = valid && kernelA_wants(T, Tc, inst, key, pi, nearC);
  if (valid && !want) {
    DV[i] = 0.0f;
    reinterpret_cast<float*>(VP)[4 * i + 3] = 0.0f;
  }
  fine_groups<DIM, U, PF, true>(T, n, inst, key, pi, want, PR, VP, nullptr, first, tp, tv, td,
                                [&](const sph_params& P, auto visit, bool grp) {
                                  nsA_target<DIM, false>(P, i, pi, vi, visit, PR, VP, DV, grp);
                                });
}

// Density of every non-fluid particle from the pressure kernel A gave it
// (the inverse EOS; opt_src_v1/equation/ns.cu:131 for the linear form).  Kept
// out of kernel A: the double-precision Tait inverse would set that kernel's
// register count and halve its occupancy.
__global__ void wall_density_kernel(InstTable T, float4* __restrict__ PR, const float4* __restrict__ VP,
                                    const int* __restrict__ TY, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (PR[i].w > 0.0f) return;  // fluid
  with_instance(T, inst_of(TY[i]), [&](const sph_params& P, int) {
    reinterpret_cast<float*>(PR)[4 * i + 3] = -eos_density(P, VP[i].w);
  });
}

template <int DIM, int U, bool PF>
__global__ __launch_bounds__(64) void nsB_fine_kernel(InstTable T, const float4* __restrict__ PR,
                                                      const float4* __restrict__ VP, const int* __restrict__ TY,
                                                      const float* __restrict__ DV, const int* __restrict__ keys,
                                                      const int* __restrict__ first, float4* __restrict__ RATE,
                                                      int n, int remap) {
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
  fine_groups<DIM, U, PF, false>(T, n, inst_of(TY[ic]), key, pi, valid && fluid, PR, VP, DV, first, tp, tv, td,
                                 [&](const sph_params& P, auto visit, bool grp) {
                                   nsB_target<DIM>(P, i, pi, vi, dvi, visit, VP, DV, RATE, grp);
                                 });
}

// Candidate / in-support pair counts of the fine-row visits (statistics
// only): candA over the kernel-A rows of the targets kernel A visits (fluid,
// and walls with fluid near), hitsA every target's in-support fluid
// neighbours, candB / hitsB over fluid targets' kernel-B rows.  Candidates
// exclude the target itself, as the 2h-grid statistics do.
template <int DIM>
__global__ __launch_bounds__(64) void pair_stats_fine_kernel(InstTable T, InstTable Tc, const float4* __restrict__ PR,
                                                             const float4* __restrict__ VP,
                                                             const float* __restrict__ DV,
                                                             const int* __restrict__ TY, const int* __restrict__ keys,
                                                             const int* __restrict__ first,
                                                             const unsigned char* __restrict__ nearC,
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
  const bool wantA = valid && kernelA_wants(T, Tc, inst, key, pi, nearC);
  fine_groups<DIM, 1, false, true>(T, n, inst, key, pi, wantA, PR, VP, nullptr, first, tp, tv, td,
                                   [&](const sph_params&, auto visit, bool) {
                                     visit(
                                         [&](int j, const float4&) {
                                           cA += j != i;
                                           return false;
                                         },
                                         ExtraNone{}, [&](int, const float4&, const NoExtra&) {});
                                   });
  fine_groups<DIM, 1, false, false>(T, n, inst, key, pi, valid, PR, VP, DV, first, tp, tv, td,
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
hipError_t nsA_fine_v(const InstTable& T, const InstTable& Tc, const FastSet& st, const FastWork& w, int n, int remap,
                      hipStream_t s) {
  const dim3 g(blocks_for(n, 64)), b(64);
  if (T.dim == 3)
    hipLaunchKernelGGL((nsA_fine_kernel<3, U, PF>), g, b, 0, s, T, Tc, st.PR, st.VP, (const int*)st.TY, w.DV,
                       (const int*)w.keys, (const int*)w.cellFirst, (const unsigned char*)w.nearC, n, remap);
  else
    hipLaunchKernelGGL((nsA_fine_kernel<2, U, PF>), g, b, 0, s, T, Tc, st.PR, st.VP, (const int*)st.TY, w.DV,
                       (const int*)w.keys, (const int*)w.cellFirst, (const unsigned char*)w.nearC, n, remap);
  return hipGetLastError();
}

template <int U, bool PF>
hipError_t nsB_fine_v(const InstTable& T, const FastSet& st, const FastWork& w, int n, int remap, hipStream_t s) {
  const dim3 g(blocks_for(n, 64)), b(64);
  if (T.dim == 3)
    hipLaunchKernelGGL((nsB_fine_kernel<3, U, PF>), g, b, 0, s, T, (const float4*)st.PR, (const float4*)st.VP,
                       (const int*)st.TY, (const float*)w.DV, (const int*)w.keys, (const int*)w.cellFirst, w.RATE,
                       n, remap);
  else
    hipLaunchKernelGGL((nsB_fine_kernel<2, U, PF>), g, b, 0, s, T, (const float4*)st.PR, (const float4*)st.VP,
                       (const int*)st.TY, (const float*)w.DV, (const int*)w.keys, (const int*)w.cellFirst, w.RATE,
                       n, remap);
  return hipGetLastError();
}

}  // namespace

int dense_runs_cap(long long cells) { return (int)((cells + 1) / GAP_GIANT + 2); }

hipError_t launch_cells_dense(const InstTable& T, const InstTable& Tc, const FastSet& in, const FastSet& out,
                              const int* keys, const int* idx, const FastWork& w, int n, hipStream_t s) {
  const int cells = T.total_cells;
  hipLaunchKernelGGL(dense_first_kernel, dim3(blocks_for((long long)n + 1, 256)), dim3(256), 0, s, keys, n, cells,
                     w.cellFirst, w.runs, w.runCount, w.runs_cap);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(fill_runs_kernel, dim3(1024), dim3(256), 0, s, w.cellFirst, (const int*)w.runs,
                     (const int*)w.runCount, w.runs_cap, w.fluidC, Tc.total_cells);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (n > 0) {
    if (T.dim == 3)
      hipLaunchKernelGGL(gather_fluidc_kernel<3>, dim3(blocks_for(n, 256)), dim3(256), 0, s, T, Tc, in, out, keys,
                         idx, w.fluidC, n);
    else
      hipLaunchKernelGGL(gather_fluidc_kernel<2>, dim3(blocks_for(n, 256)), dim3(256), 0, s, T, Tc, in, out, keys,
                         idx, w.fluidC, n);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  const dim3 gc(blocks_for(Tc.total_cells, 256));
  if (T.dim == 3)
    hipLaunchKernelGGL(near_fluid_kernel<3>, gc, dim3(256), 0, s, Tc, (const unsigned char*)w.fluidC, w.nearC);
  else
    hipLaunchKernelGGL(near_fluid_kernel<2>, gc, dim3(256), 0, s, Tc, (const unsigned char*)w.fluidC, w.nearC);
  return hipGetLastError();
}

hipError_t launch_nsA_fine(const InstTable& T, const InstTable& Tc, const FastSet& st, const FastWork& w, int n,
                           const sph_tuning& tu, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const int rm = tu.xcd_remap;
  const bool pf = tu.nbr_prefetch != 0;
  hipError_t e;
  switch (tu.nbr_unroll) {
    case 1: e = pf ? nsA_fine_v<1, true>(T, Tc, st, w, n, rm, s) : nsA_fine_v<1, false>(T, Tc, st, w, n, rm, s); break;
    case 4: e = pf ? nsA_fine_v<4, true>(T, Tc, st, w, n, rm, s) : nsA_fine_v<4, false>(T, Tc, st, w, n, rm, s); break;
    default: e = pf ? nsA_fine_v<2, true>(T, Tc, st, w, n, rm, s) : nsA_fine_v<2, false>(T, Tc, st, w, n, rm, s); break;
  }
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(wall_density_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, T, st.PR,
                     (const float4*)st.VP, (const int*)st.TY, n);
  return hipGetLastError();
}

hipError_t launch_nsB_fine(const InstTable& T, const FastSet& st, const FastWork& w, int n, const sph_tuning& tu,
                           hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const int rm = tu.xcd_remap;
  const bool pf = tu.nbr_prefetch != 0;
  switch (tu.nbr_unroll) {
    case 1: return pf ? nsB_fine_v<1, true>(T, st, w, n, rm, s) : nsB_fine_v<1, false>(T, st, w, n, rm, s);
    case 4: return pf ? nsB_fine_v<4, true>(T, st, w, n, rm, s) : nsB_fine_v<4, false>(T, st, w, n, rm, s);
    default: return pf ? nsB_fine_v<2, true>(T, st, w, n, rm, s) : nsB_fine_v<2, false>(T, st, w, n, rm, s);
  }
}

hipError_t launch_pair_stats_fine(const InstTable& T, const InstTable& Tc, const FastSet& st, const FastWork& w,
                                  unsigned long long* out4, int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const dim3 g(blocks_for(n, 64)), b(64);
  if (T.dim == 3)
    hipLaunchKernelGGL(pair_stats_fine_kernel<3>, g, b, 0, s, T, Tc, (const float4*)st.PR, (const float4*)st.VP,
                       (const float*)w.DV, (const int*)st.TY, (const int*)w.keys, (const int*)w.cellFirst,
                       (const unsigned char*)w.nearC, out4, n);
  else
    hipLaunchKernelGGL(pair_stats_fine_kernel<2>, g, b, 0, s, T, Tc, (const float4*)st.PR, (const float4*)st.VP,
                       (const float*)w.DV, (const int*)st.TY, (const int*)w.keys, (const int*)w.cellFirst,
                       (const unsigned char*)w.nearC, out4, n);
  return hipGetLastError();
}

}  // namespace sph
