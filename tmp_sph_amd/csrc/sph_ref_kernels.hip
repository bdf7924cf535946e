// Kernels behind the reference call surface: AoS float[3N] positions,
// velocities and accelerations, int types, one thread per particle.
// They implement the semantics of
//   calcHashD                        src/lib/mesh.cu:26-43
//   reorderD (+ permutation table)   opt_src_v0/lib/mesh.cu:106-159
//   computeBoundary_Delta_acoustic_D opt_src_v1/equation/ns.cu:30-134
//   computeGovering_equationD        opt_src_v1/equation/ns.cu:137-230
//   PC_predictionD / PC_correctionD  opt_src_v0/sph/sph_run.cu:388-484
// with the reference's FP32 evaluation order (SPH_NS_simpleversion in its
// reference-order mode), plus the pack / unpack kernels that let the
// default SPH_NS_simpleversion run the fast path's row-staged loops on the
// caller's arrays.  The fast path (sph_fast_kernels.hip) has its own layout
// and launch shapes.
#include "sph_device.h"
#include "sph_kernels.h"

namespace sph {

namespace {

__global__ void hash_aos_kernel(sph_params P, const float* __restrict__ pos, int* __restrict__ hash,
                                int* __restrict__ idx, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  hash[i] = cell_hash(P, pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
  idx[i] = i;
}

// Cell ranges at key changes plus the gather into sorted order
// (opt_src_v0/lib/mesh.cu:106-159).  cellStart was set to -1 beforehand.
__global__ void reorder_aos_kernel(const float* __restrict__ oPos, const float* __restrict__ oVel,
                                   const float* __restrict__ oRho, const float* __restrict__ oP,
                                   const int* __restrict__ oType, const int* __restrict__ oZone,
                                   float* __restrict__ sPos, float* __restrict__ sVel,
                                   float* __restrict__ sRho, float* __restrict__ sP,
                                   int* __restrict__ sType, int* __restrict__ sZone,
                                   const int* __restrict__ hash, const int* __restrict__ idx,
                                   int* __restrict__ cellStart, int* __restrict__ cellEnd, int n,
                                   int* __restrict__ table_l, const int* __restrict__ table_r) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int h = hash[i];
  if (i == 0 || h != hash[i - 1]) {
    cellStart[h] = i;
    if (i > 0) cellEnd[hash[i - 1]] = i;
  }
  if (i == n - 1) cellEnd[h] = n;
  const int src = idx[i];
  sPos[3 * i] = oPos[3 * src];
  sPos[3 * i + 1] = oPos[3 * src + 1];
  sPos[3 * i + 2] = oPos[3 * src + 2];
  sVel[3 * i] = oVel[3 * src];
  sVel[3 * i + 1] = oVel[3 * src + 1];
  sVel[3 * i + 2] = oVel[3 * src + 2];
  sRho[i] = oRho[src];
  sP[i] = oP[src];
  sType[i] = oType[src];
  if (table_l && table_r) table_l[i] = table_r[src] + (src - i);
  if (sZone && oZone) sZone[i] = oZone[src];
}

// Scatter from sorted back to original order (recoverD, src/lib/mesh.cu:162-179):
// row i of the sorted arrays goes to row idx[i].
__global__ void recover_aos_kernel(float* __restrict__ pos, float* __restrict__ vel, float* __restrict__ rho,
                                   float* __restrict__ p, int* __restrict__ type, int* __restrict__ zone,
                                   const int* __restrict__ idx, const float* __restrict__ sPos,
                                   const float* __restrict__ sVel, const float* __restrict__ sRho,
                                   const float* __restrict__ sP, const int* __restrict__ sType,
                                   const int* __restrict__ sZone, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int o = idx[i];
  pos[3 * o] = sPos[3 * i];
  pos[3 * o + 1] = sPos[3 * i + 1];
  pos[3 * o + 2] = sPos[3 * i + 2];
  vel[3 * o] = sVel[3 * i];
  vel[3 * o + 1] = sVel[3 * i + 1];
  vel[3 * o + 2] = sVel[3 * i + 2];
  rho[o] = sRho[i];
  p[o] = sP[i];
  type[o] = sType[i];
  if (zone && sZone) zone[o] = sZone[i];
}

// Visits the stencil rows of the cell `hash` in the reference order (z outer,
// y, then the x-adjacent cells, which are contiguous in sorted order).
// f(start, end) receives the sorted index range of one row.  Per-axis bounds
// replace the reference's single `0 <= hash <= hash_max` test.
template <int DIM, class F>
__device__ __forceinline__ void visit_rows(const sph_params& P, int hash,
                                           const int* __restrict__ cellStart,
                                           const int* __restrict__ cellEnd, F&& f) {
  const int X = P.gridxdim, Y = P.gridydim, Z = P.gridzdim;
  const int XY = X * Y;
  const int cz = hash / XY;
  const int rem = hash - cz * XY;
  const int cy = rem / X;
  const int cx = rem - cy * X;
  const int x0 = cx > 0 ? cx - 1 : 0;
  const int x1 = cx < X - 1 ? cx + 1 : X - 1;
  for (int dz = -1; dz <= 1; ++dz) {
    const int z = cz + dz;
    if (z < 0 || z >= Z) continue;
    for (int dy = (DIM == 3 ? -1 : 0); dy <= (DIM == 3 ? 1 : 0); ++dy) {
      const int y = cy + dy;
      if (y < 0 || y >= Y) continue;
      const int rowBase = z * XY + y * X;
      int s = -1, e = -1;
      for (int c = rowBase + x0; c <= rowBase + x1; ++c) {
        const int st = cellStart[c];
        if (st != -1) {
          if (s < 0) s = st;
          e = cellEnd[c];
        }
      }
      if (s >= 0) f(s, e);
    }
  }
}

template <int DIM>
__global__ void nsA_aos_kernel(sph_params P, const float* __restrict__ sPos,
                               const float* __restrict__ sVel, float* sRho, float* sP,
                               const int* __restrict__ sType, const int* __restrict__ cellStart,
                               const int* __restrict__ cellEnd, const int* __restrict__ hash,
                               float* __restrict__ dofv, int n) {
  // no FMA contraction: every operation rounds once, in the reference's order
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float px = sPos[3 * i], py = sPos[3 * i + 1], pz = sPos[3 * i + 2];
  const float vx = sVel[3 * i], vy = sVel[3 * i + 1], vz = sVel[3 * i + 2];
  const int ti = sType[i];
  float rhop_sum = 0.0f, w_sum = 0.0f, dofv_acc = 0.0f;
  visit_rows<DIM>(P, hash[i], cellStart, cellEnd, [&](int s, int e) {
    for (int j = s; j < e; ++j) {
      if (j == i) continue;
      const float dx = px - sPos[3 * j];
      const float dy = py - sPos[3 * j + 1];
      const float dz = pz - sPos[3 * j + 2];
      const float rr = dx * dx + dy * dy + dz * dz;
      if (rr <= P.rr_cut) {  // same pairs as sqrtf(rr)/h <= 2
        const float q = sqrtf(rr) / P.h;
        float fr = (1.0f - q / 2.0f) * (1.0f - q / 2.0f) * (1.0f - q / 2.0f);
        float w = fr * (1.0f - q / 2.0f);
        w *= (2.0f * q + 1) * P.adh;
        fr *= -5.0f * P.adh / (P.h * P.h);
        const int tj = sType[j];
        if (ti != 1 && tj == 1) {
          rhop_sum += (sP[j] - sRho[j] * ((0.0f - P.gravity) * dz)) * w;
          w_sum += w;
        } else if (ti == 1 && tj == 1) {
          const float dvdx = (vx - sVel[3 * j]) * dx + (vy - sVel[3 * j + 1]) * dy +
                             (vz - sVel[3 * j + 2]) * dz;
          dofv_acc -= fr * dvdx * P.particleMass / sRho[j];
        }
      }
    }
  });
  dofv[i] = dofv_acc;
  if (ti != 1) {
    float pw = fabsf(w_sum) > 1.0E-8f ? rhop_sum / w_sum : 0.0f;
    if (pw < 0.0f) pw = 0.0f;
    sP[i] = pw;
    sRho[i] = eos_density(P, pw);
  }
}

template <int DIM>
__global__ void nsB_aos_kernel(sph_params P, const float* __restrict__ sPos,
                               const float* __restrict__ sVel, const float* __restrict__ sRho,
                               const float* __restrict__ sP, const int* __restrict__ sType,
                               const float* __restrict__ dofv, const int* __restrict__ cellStart,
                               const int* __restrict__ cellEnd, const int* __restrict__ hash,
                               float* __restrict__ drho, float* __restrict__ acc, int n) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float px = sPos[3 * i], py = sPos[3 * i + 1], pz = sPos[3 * i + 2];
  const float vx = sVel[3 * i], vy = sVel[3 * i + 1], vz = sVel[3 * i + 2];
  const float pres0 = sP[i], dens0 = sRho[i], dofv0 = dofv[i];
  const int ti = sType[i];
  float drho_acc = 0.0f, ax = 0.0f, ay = 0.0f, az = 0.0f;
  visit_rows<DIM>(P, hash[i], cellStart, cellEnd, [&](int s, int e) {
    for (int j = s; j < e; ++j) {
      if (j == i) continue;
      const float dx = px - sPos[3 * j];
      const float dy = py - sPos[3 * j + 1];
      const float dz = pz - sPos[3 * j + 2];
      const float rr = dx * dx + dy * dy + dz * dz;
      if (rr <= P.rr_cut) {
        const float dvdx = (vx - sVel[3 * j]) * dx + (vy - sVel[3 * j + 1]) * dy +
                           (vz - sVel[3 * j + 2]) * dz;
        const float pres1 = sP[j], dens1 = sRho[j];
        const float q = sqrtf(rr) / P.h;
        const float fr = -5.0f * P.adh * (1.0f - q / 2.0f) * (1.0f - q / 2.0f) *
                         (1.0f - q / 2.0f) / (P.h * P.h);
        drho_acc += (dens0 * dvdx * fr + (dens0 - dens1) * rr * fr * P.delta * P.h * P.cs /
                                             (rr + P.eta * P.eta)) *
                    P.particleMass / dens1;
        float vtmp = -(pres0 + pres1) * P.particleMass * fr / (dens0 * dens1);
        if (ti == 1 && sType[j] == 1) {
          vtmp += P.h * P.cs * P.restDensity * P.particleMass * fr *
                  (P.damper * (dofv0 + dofv[j]) + P.afa * dvdx / (rr + P.eta * P.eta)) /
                  (dens0 * dens1);
        }
        ax += vtmp * dx;
        ay += vtmp * dy;
        az += vtmp * dz;
      }
    }
  });
  drho[i] = drho_acc;
  acc[3 * i] = ax;
  acc[3 * i + 1] = ay;
  acc[3 * i + 2] = az + P.gravity;
}

// Escape test with the float position widened to double, as the reference's
// comparisons against double literals do (src/sph/sph_run.cu:437).
__device__ __forceinline__ bool outside(const sph_params& P, float x, float y, float z) {
  return (double)x > P.domain_max[0] || (double)x < P.domain_min[0] ||
         (double)y > P.domain_max[1] || (double)y < P.domain_min[1] ||
         (double)z > P.domain_max[2] || (double)z < P.domain_min[2];
}

// Predictor / corrector of opt_src_v0/sph/sph_run.cu:388-484.  The increments
// `rate * timeStep / 2.0` are a float product halved in double, added to the
// float base in double and rounded back to float, as in the reference.
template <bool PRED>
__global__ void pc_aos_kernel(sph_params P, float* __restrict__ sPos, float* __restrict__ sVel,
                              float* __restrict__ sRho, float* __restrict__ sP,
                              const float* __restrict__ drho, const float* __restrict__ acc,
                              const int* __restrict__ idx, float* __restrict__ tPos,
                              float* __restrict__ tVel, float* __restrict__ tRho, int n,
                              int* __restrict__ sType) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (sType[i] != 1) return;
  const float dt = P.timeStep;
  if (outside(P, sPos[3 * i], sPos[3 * i + 1], sPos[3 * i + 2])) {
    sPos[3 * i] = P.park_pos[0];
    sPos[3 * i + 1] = P.park_pos[1];
    sPos[3 * i + 2] = P.park_pos[2];
    sVel[3 * i] = 0.0f;
    sVel[3 * i + 1] = 0.0f;
    sVel[3 * i + 2] = 0.0f;
    sP[i] = 0.0f;
    sRho[i] = P.restDensity;
    sType[i] = P.park_type;
    return;
  }
  const int r = PRED ? i : idx[i];
  if (PRED) {
    for (int k = 0; k < 3; ++k) {
      tPos[3 * r + k] = sPos[3 * i + k];
      tVel[3 * r + k] = sVel[3 * i + k];
    }
    tRho[r] = sRho[i];
  }
  float rho = (float)((double)tRho[r] + (double)(drho[i] * dt) / 2.0);
  if (rho < P.restDensity) rho = P.restDensity;
  float x[3], v[3];
  for (int k = 0; k < 3; ++k) x[k] = (float)((double)tPos[3 * r + k] + (double)(sVel[3 * i + k] * dt) / 2.0);
  for (int k = 0; k < 3; ++k) v[k] = (float)((double)tVel[3 * r + k] + (double)(acc[3 * i + k] * dt) / 2.0);
  if (!PRED) {
    rho = 2 * rho - tRho[r];
    if (rho < P.restDensity) rho = P.restDensity;
    for (int k = 0; k < 3; ++k) {
      x[k] = 2 * x[k] - tPos[3 * r + k];
      v[k] = 2 * v[k] - tVel[3 * r + k];
    }
  }
  sRho[i] = rho;
  for (int k = 0; k < 3; ++k) {
    sPos[3 * i + k] = x[k];
    sVel[3 * i + k] = v[k];
  }
  sP[i] = eos_pressure(P, rho);
}

inline unsigned blocks_for(int n, int b) { return (unsigned)((n + b - 1) / b); }

}  // namespace

hipError_t launch_hash_aos(const sph_params& P, const float* pos, int* hash, int* idx, int n,
                           hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(hash_aos_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, P, pos, hash, idx,
                     n);
  return hipGetLastError();
}

hipError_t launch_reorder_aos(const sph_params& P, const float* oPos, const float* oVel,
                              const float* oRho, const float* oP, const int* oType,
                              const int* oZone, float* sPos, float* sVel, float* sRho, float* sP,
                              int* sType, int* sZone, const int* hash, const int* idx,
                              int* cellStart, int* cellEnd, int n, int* table_l,
                              const int* table_r, hipStream_t s) {
  hipError_t e = fill_int(cellStart, -1, P.numCells, s);
  if (e != hipSuccess || n <= 0) return e;
  hipLaunchKernelGGL(reorder_aos_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, oPos, oVel,
                     oRho, oP, oType, oZone, sPos, sVel, sRho, sP, sType, sZone, hash, idx,
                     cellStart, cellEnd, n, table_l, table_r);
  return hipGetLastError();
}

hipError_t launch_recover_aos(float* pos, float* vel, float* rho, float* p, int* type, int* zone, const int* idx,
                              const float* sPos, const float* sVel, const float* sRho, const float* sP,
                              const int* sType, const int* sZone, int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(recover_aos_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, pos, vel, rho, p, type, zone,
                     idx, sPos, sVel, sRho, sP, sType, sZone, n);
  return hipGetLastError();
}

hipError_t launch_ns_aos(const sph_params& P, const float* sPos, const float* sVel, float* sRho,
                         float* sP, const int* sType, float* drho, float* acc,
                         const int* cellStart, const int* cellEnd, const int* hash, float* dofv,
                         int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const dim3 g(blocks_for(n, 256)), b(256);
  if (P.dim == 2) {
    hipLaunchKernelGGL(nsA_aos_kernel<2>, g, b, 0, s, P, sPos, sVel, sRho, sP, sType, cellStart,
                       cellEnd, hash, dofv, n);
    hipLaunchKernelGGL(nsB_aos_kernel<2>, g, b, 0, s, P, sPos, sVel, sRho, sP, sType, dofv,
                       cellStart, cellEnd, hash, drho, acc, n);
  } else {
    hipLaunchKernelGGL(nsA_aos_kernel<3>, g, b, 0, s, P, sPos, sVel, sRho, sP, sType, cellStart,
                       cellEnd, hash, dofv, n);
    hipLaunchKernelGGL(nsB_aos_kernel<3>, g, b, 0, s, P, sPos, sVel, sRho, sP, sType, dofv,
                       cellStart, cellEnd, hash, drho, acc, n);
  }
  return hipGetLastError();
}

// The caller's sorted AoS state in the fast layout (one instance): PR =
// (x, y, z, density signed + fluid / - otherwise), VP = (v, p), TY = type;
// the cells holding fluid flagged for kernel A's row trimming (cleared
// before).
__global__ void ref_pack_kernel(const float* __restrict__ sPos, const float* __restrict__ sVel,
                                const float* __restrict__ sRho, const float* __restrict__ sP,
                                const int* __restrict__ sType, const int* __restrict__ hash, float4* __restrict__ PR,
                                float4* __restrict__ VP, int* __restrict__ TY, int* __restrict__ cellFluid, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int t = sType[i];
  const float r = fabsf(sRho[i]);
  PR[i] = make_float4(sPos[3 * i], sPos[3 * i + 1], sPos[3 * i + 2], t == 1 ? r : -r);
  VP[i] = make_float4(sVel[3 * i], sVel[3 * i + 1], sVel[3 * i + 2], sP[i]);
  TY[i] = t & 0xff;
  if (t == 1) cellFluid[hash[i]] = 1;  // every writer stores the same value
}

// Rates and the walls' extrapolated pressure and density back into the
// caller's arrays (reference layouts).
__global__ void ref_unpack_kernel(const float4* __restrict__ PR, const float4* __restrict__ VP,
                                  const float* __restrict__ DV, const float4* __restrict__ RATE,
                                  const int* __restrict__ sType, float* __restrict__ drho, float* __restrict__ acc,
                                  float* __restrict__ dofv, float* __restrict__ sRho, float* __restrict__ sP, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 r = RATE[i];
  drho[i] = r.x;
  acc[3 * i] = r.y;
  acc[3 * i + 1] = r.z;
  acc[3 * i + 2] = r.w;
  dofv[i] = DV[i];
  if (sType[i] != 1) {
    sP[i] = VP[i].w;
    sRho[i] = -PR[i].w;
  }
}

hipError_t launch_ref_pack(const float* sPos, const float* sVel, const float* sRho, const float* sP,
                           const int* sType, const int* hash, const FastSet& st, int* cellFluid, int cells, int n,
                           hipStream_t s) {
  hipError_t e = fill_int(cellFluid, 0, cells, s);
  if (e != hipSuccess || n <= 0) return e;
  hipLaunchKernelGGL(ref_pack_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, sPos, sVel, sRho, sP, sType, hash,
                     st.PR, st.VP, st.TY, cellFluid, n);
  return hipGetLastError();
}

hipError_t launch_ref_unpack(const FastSet& st, const FastWork& w, const int* sType, float* drho, float* acc,
                             float* dofv, float* sRho, float* sP, int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(ref_unpack_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, (const float4*)st.PR,
                     (const float4*)st.VP, (const float*)w.DV, (const float4*)w.RATE, sType, drho, acc, dofv, sRho,
                     sP, n);
  return hipGetLastError();
}

hipError_t launch_pc_aos(const sph_params& P, bool predictor, float* sPos, float* sVel,
                         float* sRho, float* sP, const float* drho, const float* acc,
                         const int* idx, float* tPos, float* tVel, float* tRho, int n,
                         int* sType, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const dim3 g(blocks_for(n, 256)), b(256);
  if (predictor)
    hipLaunchKernelGGL(pc_aos_kernel<true>, g, b, 0, s, P, sPos, sVel, sRho, sP, drho, acc, idx,
                       tPos, tVel, tRho, n, sType);
  else
    hipLaunchKernelGGL(pc_aos_kernel<false>, g, b, 0, s, P, sPos, sVel, sRho, sP, drho, acc, idx,
                       tPos, tVel, tRho, n, sType);
  return hipGetLastError();
}

}  // namespace sph
