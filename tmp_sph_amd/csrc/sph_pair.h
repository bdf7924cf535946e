// Per-target pair physics of the neighbour-loop kernels, shared by every
// visit order: kernel A (wall pressure extrapolation, velocity divergence;
// opt_src_v1/equation/ns.cu:30-134) and kernel B (continuity with delta-SPH
// diffusion, momentum; opt_src_v1/equation/ns.cu:137-230).  A target's
// function calls visit(accept, extra, pair) once; the visit loop decides which
// candidates are seen and in which order.
#pragma once

#include "sph_fast_common.h"

namespace sph {
namespace fast {

struct NoExtra {};

// What the pair bodies read of a neighbour beyond PR: from global memory by
// index, or from a staged row (from(tv, td, t)).
struct ExtraNone {
  __device__ __forceinline__ NoExtra operator()(int) const { return NoExtra{}; }
  __device__ __forceinline__ NoExtra from(const float4*, const float*, int) const { return NoExtra{}; }
};
struct ExtraA {  // kernel A: velocity, pressure
  const float4* VP;
  __device__ __forceinline__ float4 operator()(int j) const { return ld16(VP, (unsigned)j * 16u); }
  __device__ __forceinline__ float4 from(const float4* tv, const float*, int t) const { return tv[t]; }
};

// ------------------------------------------------------------ per target
// The pair arithmetic below is written with explicit fmaf and compiled with
// contraction off, so that every loop variant (whose code the compiler
// schedules differently) rounds identically.  The squared distance is an fma
// chain (three operations instead of five per candidate); it differs from the
// reference's unfused sum by roundings of rr, which can move only a pair at
// the cut-off radius (q = 2 within a few ulps) in or out of the set, and
// there the kernel and its gradient vanish (t = 1 - q/2 = 0), so such a
// pair's terms are zero to within 1e-20 of the others.
template <int DIM>
__device__ __forceinline__ float geom(const float4& pi, const float4& pj, float& dx, float& dy, float& dz) {
#pragma clang fp contract(off)
  dx = pi.x - pj.x;
  dy = DIM == 3 ? pi.y - pj.y : 0.0f;
  dz = pi.z - pj.z;
  return DIM == 3 ? fmaf(dx, dx, fmaf(dy, dy, dz * dz)) : fmaf(dx, dx, dz * dz);
}

// Kernel A for target i (fluid neighbours only).  WALL_RHO: also store a
// wall target's density from its pressure (the inverse EOS, in double for the
// Tait form); otherwise only the pressure, and wall_density_kernel follows.
template <int DIM, bool WALL_RHO = true, class Visit>
__device__ __forceinline__ void nsA_target(const sph_params& P, int i, const float4& pi, const float4& vi,
                                           Visit visit, float4* PR, float4* VP, float* __restrict__ DV,
                                           bool store = true) {
#pragma clang fp contract(off)
  const bool fi = pi.w > 0.0f;
  const float inv_h = 1.0f / P.h;
  const float cfr = -5.0f * P.adh / (P.h * P.h);
  const float mg = 0.0f - P.gravity;
  const float m = P.particleMass;
  const float cut = P.rr_cut;
  const float tq = -0.5f * inv_h;
  const float uq = 2.0f * inv_h;
  const float ncm = -(cfr * m);  // applied once to the divergence sum
  // The wall pressure is the ratio of the two wall sums, so the kernel
  // constant adh cancels: neither sum carries it, and the emptiness threshold
  // of the weight sum is scaled instead.
  const float swmin = 1.0E-8f / P.adh;
  float sp = 0.0f, sw = 0.0f, dv = 0.0f;
  auto accept = [&](int, const float4& pj) {
    // masks, no short-circuit branches; a fluid target meets itself (rr = 0)
    // and that pair adds exactly zero (dvdx = 0)
    float dx, dy, dz;
    const float rr = geom<DIM>(pi, pj, dx, dy, dz);
    return (rr <= cut) & (pj.w > 0.0f);
  };
  auto wall_pair = [&](int, const float4& pj, const float4& vj) {  // pressure extrapolation
    float dx, dy, dz;
    const float rr = geom<DIM>(pi, pj, dx, dy, dz);
    const float s = fast_sqrt(rr);
    const float t = fmaf(s, tq, 1.0f);  // 1 - q/2
    const float t3 = t * t * t;
    const float w = (t3 * t) * fmaf(s, uq, 1.0f);  // (1 - q/2)^4 (1 + 2q)
    sp = fmaf(fmaf(-pj.w, mg * dz, vj.w), w, sp);
    sw += w;
  };
  auto fluid_pair = [&](int, const float4& pj, const float4& vj) {  // velocity divergence
    // t = 1 - q/2 straight from sqrt(rr), -cfr m applied to the sum
    // (reassociated against ns.cu:95-104: rounding moves in the last bits)
    float dx, dy, dz;
    const float rr = geom<DIM>(pi, pj, dx, dy, dz);
    const float t = fmaf(fast_sqrt(rr), tq, 1.0f);
    const float dvdx = DIM == 3 ? fmaf(vi.x - vj.x, dx, fmaf(vi.y - vj.y, dy, (vi.z - vj.z) * dz))
                                : fmaf(vi.x - vj.x, dx, (vi.z - vj.z) * dz);
    dv = fmaf(t * t * t, dvdx * fast_rcp(pj.w), dv);
  };
  // Waves of one target kind (most of them) run a body without the branch on
  // the kind; the choice is uniform over the lanes that run the visit.
  const bool walls_here = __ballot(store && !fi) != 0ull;
  const bool fluid_here = __ballot(store && fi) != 0ull;
  if (!walls_here) {
    visit(accept, ExtraA{VP}, fluid_pair);
  } else if (!fluid_here) {
    visit(accept, ExtraA{VP}, wall_pair);
  } else {
    visit(accept, ExtraA{VP}, [&](int j, const float4& pj, const float4& vj) {
      if (!fi)
        wall_pair(j, pj, vj);
      else
        fluid_pair(j, pj, vj);
    });
  }
  if (!store) return;
  DV[i] = dv * ncm;
  if (!fi) {
    float pw = fabsf(sw) > swmin ? sp / sw : 0.0f;
    if (pw < 0.0f) pw = 0.0f;
    reinterpret_cast<float*>(VP)[4 * i + 3] = pw;
    if (WALL_RHO) reinterpret_cast<float*>(PR)[4 * i + 3] = -eos_density(P, pw);
  }
}

// What kernel B reads of a neighbour beyond PR.
struct NbrB {
  float4 v;  // velocity, pressure
  float dv;  // dofv
};
struct ExtraB {
  const float4* VP;
  const float* DV;
  __device__ __forceinline__ NbrB operator()(int j) const {
    return NbrB{ld16(VP, (unsigned)j * 16u), ld4(DV, (unsigned)j * 4u)};
  }
  __device__ __forceinline__ NbrB from(const float4* tv, const float* td, int t) const { return NbrB{tv[t], td[t]}; }
};

// Kernel B for target i (all neighbours): fluid targets always; wall targets
// when the caller wants their rates too (the reference computes them; its
// integrator ignores them).  A wall target's density is |PR.w| and it has no
// artificial-viscosity or damper pairs (fluid-fluid only).
template <int DIM, class Visit>
__device__ __forceinline__ void nsB_target(const sph_params& P, int i, const float4& pi, const float4& vi, float dvi,
                                           Visit visit, const float4* __restrict__ VP,
                                           const float* __restrict__ DV, float4* __restrict__ RATE,
                                           bool store = true) {
#pragma clang fp contract(off)
  // Target constants folded so that a pair costs as few VALU operations as
  // possible: t = 1 - q/2 straight from sqrt(rr); the kernel factor cfr and
  // the mass m ride on the per-target factors of the momentum terms and
  // multiply the density sum once; the artificial-viscosity coefficient afa
  // rides on avc.  The products are
  // reassociated against the reference's order (ns.cu:205-215): rounding
  // moves in the last bits, within the parity contract.
  const bool fi = pi.w > 0.0f;
  const float rho_i = fabsf(pi.w);
  const float inv_h = 1.0f / P.h;
  const float tq = -0.5f * inv_h;
  const float cfr = -5.0f * P.adh / (P.h * P.h);
  const float m = P.particleMass;
  const float cm = cfr * m;
  const float dhc = P.delta * P.h * P.cs;
  const float eta2 = P.eta * P.eta;
  const float cut = P.rr_cut;
  const float damper = P.damper;
  const float dofv_i = damper * dvi;
  const float inv_rho_i = 1.0f / rho_i;
  const float pmc = m * inv_rho_i * cfr;                                  // pressure term factor
  const float pic = -(vi.w * pmc);                                         // its target part
  const float avcc = P.h * P.cs * P.restDensity * m * inv_rho_i * cfr;   // damper term factor
  const float avca = avcc * P.afa;                                         // viscosity term factor
  // the density rate as two sums, sum w dvdx and sum w (rho_i - rho_j) rr / (rr + eta^2),
  // combined once per target with rho_i, dhc and the constants
  float dr1 = 0.0f, dr2 = 0.0f, ax = 0.0f, ay = 0.0f, az = 0.0f;
  visit(
      [&](int, const float4& pj) {
        // the target meets itself (rr = 0): dx = dy = dz = 0, dvdx = 0 and
        // rho_i - rho_j = 0 make that pair add exactly zero
        float dx, dy, dz;
        const float rr = geom<DIM>(pi, pj, dx, dy, dz);
        return rr <= cut;
      },
      ExtraB{VP, DV},
      [&](int, const float4& pj, const NbrB& nb) {
        float dx, dy, dz;
        const float rr = geom<DIM>(pi, pj, dx, dy, dz);
        const float4 vj = nb.v;
        const bool fj = pj.w > 0.0f;
        const float rho_j = fabsf(pj.w);
        const float t = fmaf(fast_sqrt(rr), tq, 1.0f);
        const float w = (t * t * t) * fast_rcp(rho_j);  // kernel shape over rho_j
        const float inv_re = fast_rcp(rr + eta2);
        const float dvdx = DIM == 3 ? fmaf(vi.x - vj.x, dx, fmaf(vi.y - vj.y, dy, (vi.z - vj.z) * dz))
                                    : fmaf(vi.x - vj.x, dx, (vi.z - vj.z) * dz);
        dr1 = fmaf(w, dvdx, dr1);
        dr2 = fmaf(w, (rho_i - rho_j) * rr * inv_re, dr2);
        float vt = fmaf(-pmc, vj.w, pic);
        if (fi & fj) vt = fmaf(avca, dvdx * inv_re, fmaf(avcc, fmaf(damper, nb.dv, dofv_i), vt));
        vt *= w;
        ax = fmaf(vt, dx, ax);
        if (DIM == 3) ay = fmaf(vt, dy, ay);
        az = fmaf(vt, dz, az);
      });
  if (store) RATE[i] = make_float4(cm * fmaf(rho_i, dr1, dhc * dr2), ax, ay, az + P.gravity);
}

// ------------------------------------------------------------ staging
__device__ __forceinline__ int wave_min(int v) {
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o));
  return __builtin_amdgcn_readfirstlane(v);
}
__device__ __forceinline__ int wave_max(int v) {
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
  return __builtin_amdgcn_readfirstlane(v);
}


}  // namespace fast
}  // namespace sph
