// Device-side helpers of the delta-SPH step for gfx950: cell coordinates and
// hashes, equations of state and the Wendland C2 kernel constants.
#pragma once

#include <hip/hip_runtime.h>

#include "sph_internal.h"

namespace sph {

// Cell coordinate along one axis: floorf((p - min) / size) as in calcGridPos
// (src/lib/mesh.cu:6-17).  FP32 subtraction and correctly rounded FP32
// division, so the value equals the reference's for every in-grid position.
// The result is clamped into [0, dim); the reference instead relies on 0.1 m
// of padding and reads out of bounds for particles that leave the grid.
__device__ __forceinline__ int cell_coord(float p, float mn, float size, int dim) {
  float f = floorf((p - mn) / size);
  f = fminf(fmaxf(f, 0.0f), (float)(dim - 1));
  return (int)f;
}

// calcGridHash (src/lib/mesh.cu:20-23): z*X*Y + y*X + x.
__device__ __forceinline__ int cell_hash(const sph_params& P, float x, float y, float z) {
  int cx = cell_coord(x, P.gridxmin, P.gridSize, P.gridxdim);
  int cy = cell_coord(y, P.gridymin, P.gridSize, P.gridydim);
  int cz = cell_coord(z, P.gridzmin, P.gridSize, P.gridzdim);
  return cz * P.gridxdim * P.gridydim + cy * P.gridxdim + cx;
}

// Equation of state.  Linear form: p = pow(cs, 2.0) * (rho - rho0) evaluated
// in double and stored as float (src/sph/sph_run.cu:467).  Tait form:
// p = B ((rho/rho0)^gamma - 1), evaluated in double because (rho/rho0)^7 - 1
// cancels heavily near rest density.
__device__ __forceinline__ double pow_gamma(double x, float gamma) {
  if (gamma == 7.0f) {
    double x2 = x * x;
    double x4 = x2 * x2;
    return x4 * x2 * x;
  }
  return pow(x, (double)gamma);
}

__device__ __forceinline__ float eos_pressure(const sph_params& P, float rho) {
  if (P.eos == SPH_EOS_TAIT) {
    double x = (double)rho / (double)P.restDensity;
    return (float)((double)P.tait_B * (pow_gamma(x, P.gamma) - 1.0));
  }
  double cs2 = (double)P.cs * (double)P.cs;
  return (float)(cs2 * (double)(rho - P.restDensity));
}

// Inverse EOS for wall particles.  Linear form as in kernel A
// (opt_src_v1/equation/ns.cu:131): p / (cs*cs) + rho0 in FP32.
__device__ __forceinline__ float eos_density(const sph_params& P, float p) {
  if (P.eos == SPH_EOS_TAIT) {
    double r = (double)P.restDensity * pow(1.0 + (double)p / (double)P.tait_B, 1.0 / (double)P.gamma);
    return (float)r;
  }
  return p / (P.cs * P.cs) + P.restDensity;
}

}  // namespace sph
