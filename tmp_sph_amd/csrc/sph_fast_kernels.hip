// Fast-path kernels of the delta-SPH step for gfx950.
//
// One launch covers every dam-break instance of a context (K >= 1): the
// particles of all instances are sorted together by a global cell key
// (instance cell offset + local cell hash), so each instance keeps its own
// contiguous sorted range and its own neighbour order, and the per-particle
// work reads its instance's parameters.  A wave selects parameters with a
// waterfall loop over the distinct instances among its lanes, so they stay
// wave-uniform (SGPRs); waves inside one instance run the loop body once.
//
// Layout (one particle set, sorted by cell):
//   PR[i] = (x, y, z, +rho | -rho)  float4; the sign of w marks fluid (+)
//   VP[i] = (vx, vy, vz, p)         float4
//   X[k][i]                         FP64 positions, SoA; PR.xyz == (float)X
//   ID[i]                           original (concatenated) particle index
//   TY[i]                           instance << 8 | reference type
// A neighbour candidate costs one 16-byte load (PR); VP and dofv are loaded
// only for pairs inside the support.  Positions are FP64 only in the
// integrator; hashing and the neighbour loops use the FP32 copy.
//
// Physics per kernel (reference formulas, reorganised for fewer divides):
//   nsA  opt_src_v1/equation/ns.cu:30-134   wall p extrapolation, dofv
//   nsB  opt_src_v1/equation/ns.cu:137-230  drho/dt, dv/dt
//   pc   opt_src_v0/sph/sph_run.cu:388-484  predictor / corrector + EOS
#include "sph_fast_common.h"

namespace sph {

using namespace fast;

namespace {

inline unsigned blocks_for(long long n, int b) { return (unsigned)((n + b - 1) / b); }


// With cellStart (not null) it also clears the cell tables for the gather
// that follows the sort (cellStart = -1, cellFluid = 0 for all ncells cells),
// which saves that launch.
__global__ void hash_fast_kernel(InstTable T, const float4* __restrict__ PR, const int* __restrict__ TY,
                                 int* __restrict__ keys, int* __restrict__ idx, int n, int* __restrict__ zero,
                                 int* __restrict__ cellStart, int* __restrict__ cellFluid, long long ncells) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (zero && i == 0) *zero = 0;
  if (cellStart) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long c = i; c < ncells; c += stride) {
      cellStart[c] = -1;
      cellFluid[c] = 0;
    }
  }
  if (i >= n) return;
  const float4 p = PR[i];
  with_instance(T, inst_of(TY[i]), [&](const sph_params& P, int coff) {
    keys[i] = coff + cell_hash(P, p.x, p.y, p.z);
  });
  idx[i] = i;
}

// cellStart = -1 (empty) and cellFluid = 0 for every cell, grid-stride.
__global__ void clear_cells_kernel(int* __restrict__ cellStart, int* __restrict__ cellFluid, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (; i < n; i += stride) {
    cellStart[i] = -1;
    cellFluid[i] = 0;
  }
}

template <int DIM>
__global__ void cells_gather_kernel(FastSet in, FastSet out, const int* __restrict__ keys,
                                    const int* __restrict__ idx, int* __restrict__ cellStart,
                                    int* __restrict__ cellEnd, int* __restrict__ cellFluid, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int h = keys[i];
  if (i == 0) {
    cellStart[h] = 0;
  } else {
    const int hp = keys[i - 1];
    if (h != hp) {
      cellStart[h] = i;
      cellEnd[hp] = i;
    }
  }
  if (i == n - 1) cellEnd[h] = n;
  const int src = idx[i];
  const float4 pr = in.PR[src];
  out.PR[i] = pr;
  if (pr.w > 0.0f) cellFluid[h] = 1;  // fluid particle (every writer stores the same value)
  out.VP[i] = in.VP[src];
  out.X[0][i] = in.X[0][src];
  if (DIM == 3) out.X[1][i] = in.X[1][src];
  out.X[2][i] = in.X[2][src];
  out.ID[i] = in.ID[src];
  out.TY[i] = in.TY[src];
}



// Predictor (PRED) / corrector of the reference PC scheme with FP64
// positions.  The predictor stores the step-start state at its own sorted
// slot; the corrector finds it through perm (the half-step-2 sort order).
template <int DIM, bool PRED>
__global__ void pc_fast_kernel(InstTable T, FastSet st, FastWork w, const int* __restrict__ perm, int n) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 pr = st.PR[i];
  if (!(pr.w > 0.0f)) return;  // fluid <=> type 1 <=> positive signed density
  const int ty = st.TY[i];
  with_instance(T, inst_of(ty), [&](const sph_params& P, int) {
    const double x0 = st.X[0][i];
    const double y0 = DIM == 3 ? st.X[1][i] : 0.0;
    const double z0 = st.X[2][i];
    if (outside64(P, x0, y0, z0)) {
      st.X[0][i] = P.park_pos[0];
      if (DIM == 3) st.X[1][i] = P.park_pos[1];
      st.X[2][i] = P.park_pos[2];
      const float rs = P.park_type == 1 ? P.restDensity : -P.restDensity;
      st.PR[i] = make_float4(P.park_pos[0], DIM == 3 ? P.park_pos[1] : 0.0f, P.park_pos[2], rs);
      st.VP[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      st.TY[i] = (inst_of(ty) << 8) | (P.park_type & 0xff);
      return;
    }
    const float4 vp = st.VP[i];
    const float4 rate = w.RATE[i];
    const float dt = P.timeStep;
    const double hdt = 0.5 * (double)dt;
    double bx, by, bz;
    float4 bv;  // (vx, vy, vz, rho) at step start
    if (PRED) {
      bx = x0;
      by = y0;
      bz = z0;
      bv = make_float4(vp.x, vp.y, vp.z, pr.w);
      w.TX[0][i] = bx;
      if (DIM == 3) w.TX[1][i] = by;
      w.TX[2][i] = bz;
      w.TV[i] = bv;
    } else {
      const int r = perm[i];
      bx = w.TX[0][r];
      by = DIM == 3 ? w.TX[1][r] : 0.0;
      bz = w.TX[2][r];
      bv = w.TV[r];
    }
    float rho = (float)((double)bv.w + (double)(rate.x * dt) / 2.0);
    if (rho < P.restDensity) rho = P.restDensity;
    double nx = bx + (double)vp.x * hdt;
    double ny = DIM == 3 ? by + (double)vp.y * hdt : 0.0;
    double nz = bz + (double)vp.z * hdt;
    float vx = (float)((double)bv.x + (double)(rate.y * dt) / 2.0);
    float vy = DIM == 3 ? (float)((double)bv.y + (double)(rate.z * dt) / 2.0) : 0.0f;
    float vz = (float)((double)bv.z + (double)(rate.w * dt) / 2.0);
    if (!PRED) {
      rho = 2 * rho - bv.w;
      if (rho < P.restDensity) rho = P.restDensity;
      nx = 2.0 * nx - bx;
      ny = DIM == 3 ? 2.0 * ny - by : 0.0;
      nz = 2.0 * nz - bz;
      vx = 2 * vx - bv.x;
      vy = DIM == 3 ? 2 * vy - bv.y : 0.0f;
      vz = 2 * vz - bv.z;
    }
    st.X[0][i] = nx;
    if (DIM == 3) st.X[1][i] = ny;
    st.X[2][i] = nz;
    st.PR[i] = make_float4((float)nx, (float)ny, (float)nz, rho);
    st.VP[i] = make_float4(vx, vy, vz, eos_pressure(P, rho));
  });
}

__global__ void upload_kernel(int dim, FastSet st, const float* __restrict__ pos, const float* __restrict__ vel,
                              const float* __restrict__ rho, const float* __restrict__ prs,
                              const int* __restrict__ type, const int* __restrict__ inst, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float x = pos[3 * i], y = dim == 3 ? pos[3 * i + 1] : 0.0f, z = pos[3 * i + 2];
  const int t = type[i];
  st.X[0][i] = (double)x;
  if (st.X[1]) st.X[1][i] = (double)y;
  st.X[2][i] = (double)z;
  st.PR[i] = make_float4(x, y, z, t == 1 ? fabsf(rho[i]) : -fabsf(rho[i]));
  st.VP[i] = make_float4(vel[3 * i], dim == 3 ? vel[3 * i + 1] : 0.0f, vel[3 * i + 2], prs[i]);
  st.ID[i] = i;
  st.TY[i] = (inst[i] << 8) | (t & 0xff);
}

__global__ void download_kernel(FastSet st, float* __restrict__ pos, float* __restrict__ vel,
                                float* __restrict__ rho, float* __restrict__ prs, int* __restrict__ type,
                                double* __restrict__ pos64, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int o = st.ID[i];
  const float4 pr = st.PR[i];
  const float4 vp = st.VP[i];
  const double x = st.X[0][i], y = st.X[1] ? st.X[1][i] : 0.0, z = st.X[2][i];
  if (pos) {
    pos[3 * o] = (float)x;
    pos[3 * o + 1] = (float)y;
    pos[3 * o + 2] = (float)z;
  }
  if (pos64) {
    pos64[3 * o] = x;
    pos64[3 * o + 1] = y;
    pos64[3 * o + 2] = z;
  }
  if (vel) {
    vel[3 * o] = vp.x;
    vel[3 * o + 1] = vp.y;
    vel[3 * o + 2] = vp.z;
  }
  if (rho) rho[o] = fabsf(pr.w);
  if (prs) prs[o] = vp.w;
  if (type) type[o] = type_of(st.TY[i]);
}

// Candidate / in-support pair counts of kernels A and B (statistics only).
template <int DIM>
__global__ void pair_stats_kernel(InstTable T, const float4* __restrict__ PR, const int* __restrict__ TY,
                                  const int* __restrict__ keys, const int* __restrict__ cellStart,
                                  const int* __restrict__ cellEnd, const int* __restrict__ cellFluid,
                                  unsigned long long* __restrict__ out, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long c = 0, cA = 0, hA = 0, hB = 0;
  bool fi = false;
  if (i < n) {
    const float4 pi = PR[i];
    fi = pi.w > 0.0f;
    with_instance(T, inst_of(TY[i]), [&](const sph_params& P, int coff) {
      RowRange rows[DIM == 3 ? 9 : 3], rowsA[DIM == 3 ? 9 : 3];
      row_ranges<DIM>(P, coff, keys[i], pi, cellStart, cellEnd, nullptr, rows);
      row_ranges<DIM>(P, coff, keys[i], pi, cellStart, cellEnd, cellFluid, rowsA);
      for (int r = 0; r < (DIM == 3 ? 9 : 3); ++r)
        for (int j = rowsA[r].s; j < rowsA[r].e; ++j) cA += j != i;
      for (int r = 0; r < (DIM == 3 ? 9 : 3); ++r)
        for (int j = rows[r].s; j < rows[r].e; ++j) {
          if (j == i) continue;
          ++c;
          const float4 pj = PR[j];
          const float dx = pi.x - pj.x, dy = DIM == 3 ? pi.y - pj.y : 0.0f, dz = pi.z - pj.z;
          const float rr = dx * dx + dy * dy + dz * dz;
          if (rr <= P.rr_cut) {
            ++hB;
            if (pj.w > 0.0f) ++hA;
          }
        }
    });
  }
  unsigned long long v[4] = {cA, hA, fi ? c : 0ull, fi ? hB : 0ull};
  for (int k = 0; k < 4; ++k) {
    unsigned long long x = v[k];
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
    if ((threadIdx.x & 63) == 0) atomicAdd(&out[k], x);
  }
}

__global__ void download_rates_kernel(FastSet st, FastWork w, float* __restrict__ drho, float* __restrict__ acc,
                                      float* __restrict__ dofv, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int o = st.ID[i];
  const float4 r = w.RATE[i];
  if (drho) drho[o] = r.x;
  if (acc) {
    acc[3 * o] = r.y;
    acc[3 * o + 1] = r.z;
    acc[3 * o + 2] = r.w;
  }
  if (dofv) dofv[o] = w.DV[i];
}

// Surge front (max x of in-domain fluid) and column height (max z of fluid
// with x <= band) per instance.  Both are >= 0 inside the domain, so the IEEE
// bit patterns order like the values and an integer atomicMax reduces them.
// Order-preserving map of a double onto an unsigned 64-bit key (negative
// values: all bits flipped; others: sign bit set), so an unsigned atomicMax on
// keys is a max on the doubles for any sign.  Key 0 (the cleared word) lies
// below every finite value and means "no particle".
__device__ __forceinline__ unsigned long long ordered_key(double v) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

__global__ void probe_kernel(InstTable T, FastSet st, double band, unsigned long long* __restrict__ out, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const double none = -__builtin_huge_val();
  double front = none, height = none;
  int k = 0;
  if (i < n) {
    const int ty = st.TY[i];
    k = inst_of(ty);
    if (type_of(ty) == 1) {
      const double x = st.X[0][i], z = st.X[2][i];
      const double y = st.X[1] ? st.X[1][i] : 0.0;
      if (!outside64(T.P[k], x, y, z)) {
        front = x;
        if (x <= band) height = z;
      }
    }
  }
  // reduce over the lanes of each distinct instance in the wave (one pass
  // when the wave lies inside one instance), then one atomic per instance
  // every lane of the wave runs the loop (no early exit above)
  bool todo = i < n;
  for (unsigned long long pending = __ballot(todo); pending; pending = __ballot(todo)) {
    const int first = __ffsll((long long)pending) - 1;
    const int u = __shfl(k, first);
    const bool mine = todo && k == u;
    double f = mine ? front : none, h = mine ? height : none;
    for (int off = 32; off > 0; off >>= 1) {
      f = fmax(f, __shfl_xor(f, off));
      h = fmax(h, __shfl_xor(h, off));
    }
    if ((int)__lane_id() == first) {
      if (f != none) atomicMax(&out[2 * u], ordered_key(f));
      if (h != none) atomicMax(&out[2 * u + 1], ordered_key(h));
    }
    if (mine) todo = false;
  }
}

template <class K2, class K3, class... A>
hipError_t launch_dim(int dim, K2 k2, K3 k3, dim3 g, dim3 b, hipStream_t s, A... args) {
  if (dim == 3)
    hipLaunchKernelGGL(k3, g, b, 0, s, args...);
  else
    hipLaunchKernelGGL(k2, g, b, 0, s, args...);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_hash_fast(const InstTable& T, const FastSet& in, int* keys, int* idx, int n, hipStream_t s,
                            int* zero, int* cellStart, int* cellFluid) {
  if (n <= 0 && !zero && !cellStart) return hipSuccess;
  const long long nc = cellStart ? T.total_cells : 0;
  hipLaunchKernelGGL(hash_fast_kernel, dim3(blocks_for(n > 0 ? n : 1, 256)), dim3(256), 0, s, T, in.PR, in.TY, keys,
                     idx, n, zero, cellStart, cellFluid, nc);
  return hipGetLastError();
}

hipError_t launch_cells_gather_fast(const InstTable& T, const FastSet& in, const FastSet& out, const int* keys,
                                    const int* idx, int* cellStart, int* cellEnd, int* cellFluid, int n,
                                    hipStream_t s, bool cleared) {
  const long long nc = T.total_cells;
  if (nc > 0 && !cleared) {  // both cell tables cleared by one launch
    const long long blocks = std::min<long long>((nc + 255) / 256, 4096);
    hipLaunchKernelGGL(clear_cells_kernel, dim3((unsigned)blocks), dim3(256), 0, s, cellStart, cellFluid, nc);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  if (n <= 0) return hipSuccess;
  return launch_dim(T.dim, cells_gather_kernel<2>, cells_gather_kernel<3>, dim3(blocks_for(n, 256)), dim3(256), s,
                    in, out, keys, idx, cellStart, cellEnd, cellFluid, n);
}

hipError_t launch_pc_fast(const InstTable& T, bool predictor, const FastSet& st, const FastWork& w,
                          const int* perm, int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const dim3 g(blocks_for(n, 256)), b(256);
  if (predictor)
    return launch_dim(T.dim, pc_fast_kernel<2, true>, pc_fast_kernel<3, true>, g, b, s, T, st, w, perm, n);
  return launch_dim(T.dim, pc_fast_kernel<2, false>, pc_fast_kernel<3, false>, g, b, s, T, st, w, perm, n);
}

hipError_t launch_upload_fast(int dim, const FastSet& st, const float* pos, const float* vel, const float* rho,
                              const float* prs, const int* type, const int* inst, int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(upload_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, dim, st, pos, vel, rho, prs, type,
                     inst, n);
  return hipGetLastError();
}

hipError_t launch_download_fast(const FastSet& st, float* pos, float* vel, float* rho, float* prs, int* type,
                                double* pos64, int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(download_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, st, pos, vel, rho, prs, type,
                     pos64, n);
  return hipGetLastError();
}

hipError_t launch_download_rates(const FastSet& st, const FastWork& w, float* drho, float* acc, float* dofv, int n,
                                 hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(download_rates_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, st, w, drho, acc, dofv, n);
  return hipGetLastError();
}

hipError_t launch_pair_stats(const InstTable& T, const FastSet& st, const FastWork& w, unsigned long long* out4,
                             int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  return launch_dim(T.dim, pair_stats_kernel<2>, pair_stats_kernel<3>, dim3(blocks_for(n, 256)), dim3(256), s, T,
                    (const float4*)st.PR, (const int*)st.TY, (const int*)w.keys, (const int*)w.cellStart,
                    (const int*)w.cellEnd, (const int*)w.cellFluid, out4, n);
}

// Height gauges (the reference's measuring bands, src/io/io.cu:236-255):
// gauge g reports the largest z of an instance-0 fluid particle with
// |x - gx[g]| < hw (and |y - gy[g]| < hw in 3-D) and z < zcap[g].  Heights
// are stored as the bits of non-negative floats, whose integer order is the
// float order.
__global__ void gauge_kernel(InstTable T, FastSet st, const float* __restrict__ gx, const float* __restrict__ gy,
                             const float* __restrict__ zcap, int ng, float hw, unsigned* __restrict__ out, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int ty = st.TY[i];
  if (inst_of(ty) != 0 || type_of(ty) != 1) return;
  const float4 p = st.PR[i];
  for (int g = 0; g < ng; ++g) {
    const bool in_x = fabsf(p.x - gx[g]) < hw;
    const bool in_y = T.dim == 2 || fabsf(p.y - gy[g]) < hw;
    if (in_x && in_y && p.z < zcap[g]) atomicMax(&out[g], __float_as_uint(fmaxf(p.z, 0.0f)));
  }
}

// Pressure of the particles with the given original ids (the reference's
// measuring nodes, src/io/io.cu:268-276).
__global__ void pressure_at_kernel(FastSet st, const int* __restrict__ ids, int nid, float* __restrict__ out, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int id = st.ID[i];
  for (int k = 0; k < nid; ++k)
    if (ids[k] == id) out[k] = st.VP[i].w;
}

hipError_t launch_gauges_fast(const InstTable& T, const FastSet& st, const float* gx, const float* gy,
                              const float* zcap, int ng, float hw, unsigned* out, int n, hipStream_t s) {
  if (n <= 0 || ng <= 0) return hipSuccess;
  hipLaunchKernelGGL(gauge_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, T, st, gx, gy, zcap, ng, hw, out, n);
  return hipGetLastError();
}

hipError_t launch_pressure_at_fast(const FastSet& st, const int* ids, int nid, float* out, int n, hipStream_t s) {
  if (n <= 0 || nid <= 0) return hipSuccess;
  hipLaunchKernelGGL(pressure_at_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, st, ids, nid, out, n);
  return hipGetLastError();
}

hipError_t launch_probe_fast(const InstTable& T, const FastSet& st, double band, unsigned long long* out, int n,
                             hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(probe_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, T, st, band, out, n);
  return hipGetLastError();
}

}  // namespace sph
