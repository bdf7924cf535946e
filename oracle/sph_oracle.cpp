// CPU ORACLE of the delta-SPH dam-break step -- TEST INFRASTRUCTURE ONLY.
//
// This file restates the reference algorithm in plain serial/OpenMP C++ so the
// GPU library can be checked against it.  Only the test suite, the smoke entry
// point and the CPU-baseline leg of the benchmark load it; the product library
// never links or calls it.
//
// What it follows (reference = ZZPHUB/tmp_sph):
//   cell coordinate / hash      src/lib/mesh.cu:6-23, calcHashD :26-43
//   sort_by_key                 src/lib/mesh.cu:73-86 (stable ascending int sort)
//   reorder + permutation table opt_src_v0/lib/mesh.cu:106-177
//   kernel A                    opt_src_v1/equation/ns.cu:30-134
//   kernel B                    opt_src_v1/equation/ns.cu:137-230
//   predictor / corrector + EOS src/sph/sph_run.cu:425-519 with the ping-pong
//                               slot rule of opt_src_v0/sph/sph_run.cu:398,446
//   step order                  opt_src_v0/sph/sph_run.cu:269-309
// Deviations, all identical in effect for particles inside the padded grid:
//   * neighbour cells are bounds-checked per axis instead of 0 <= hash <= hash_max
//     (the reference reads past cellStart there, src/sph/sph_run.cu:92);
//   * cell coordinates are clamped into the grid;
//   * 2-D mode (plane x-z, 3 stencil rows), Tait EOS, FP64 positions and the
//     FP64 mode have no reference counterpart (they serve the 2-D and Tait
//     dam-break cases).
//
// Precision modes of orc_run / orc_eval:
//   0  FP32 everywhere, float positions, reference increments  (reference step)
//   1  FP32 fields, FP64 positions in the integrator          (GPU fast path policy)
//   2  FP64 everywhere                                          (host FP64 run)
// Build with -ffp-contract=off so every float operation rounds once, as written.
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../include/sph_capi.h"

namespace {

template <class R>
inline int cell_coord(R p, R mn, R size, int dim) {
  R f = std::floor((p - mn) / size);
  if (!(f >= (R)0)) f = (R)0;  // also maps NaN to 0
  if (f > (R)(dim - 1)) f = (R)(dim - 1);
  return (int)f;
}

template <class R>
inline int cell_hash(const sph_params& P, R x, R y, R z) {
  int cx = cell_coord<R>(x, (R)P.gridxmin, (R)P.gridSize, P.gridxdim);
  int cy = cell_coord<R>(y, (R)P.gridymin, (R)P.gridSize, P.gridydim);
  int cz = cell_coord<R>(z, (R)P.gridzmin, (R)P.gridSize, P.gridzdim);
  return cz * P.gridxdim * P.gridydim + cy * P.gridxdim + cx;
}

// ---------------------------------------------------------------- EOS
inline double pow_gamma(double x, double gamma) {
  if (gamma == 7.0) {
    double x2 = x * x;
    double x4 = x2 * x2;
    return x4 * x2 * x;
  }
  return std::pow(x, gamma);
}

// pow(par.cs, 2.0) * (rho - rho0) in double, stored in R (src/sph/sph_run.cu:467)
template <class R>
inline R eos_pressure(const sph_params& P, R rho) {
  if (P.eos == SPH_EOS_TAIT) {
    double x = (double)rho / (double)P.restDensity;
    return (R)((double)P.tait_B * (pow_gamma(x, (double)P.gamma) - 1.0));
  }
  double cs2 = (double)P.cs * (double)P.cs;
  return (R)(cs2 * (double)(rho - (R)P.restDensity));
}

// wall density from the extrapolated pressure (opt_src_v1/equation/ns.cu:131)
template <class R>
inline R eos_density(const sph_params& P, R p) {
  if (P.eos == SPH_EOS_TAIT)
    return (R)((double)P.restDensity *
               std::pow(1.0 + (double)p / (double)P.tait_B, 1.0 / (double)P.gamma));
  return p / ((R)P.cs * (R)P.cs) + (R)P.restDensity;
}

// ------------------------------------------------------ neighbour search
// Stencil reach of a grid: the smallest K >= 1 with K * gridSize >= 2h.  The
// reference's grid (gridSize = 2h) has K = 1, its 27-cell search
// (src/equation/ns.cu:37-55); a finer grid searches K cells on each side, as
// the reference's last version does with cells of one spacing
// (opt_src_v4/equation/ns.cu:62-93).
inline int stencil_reach(const sph_params& P) {
  int k = 1;
  while (k < 8 && (float)k * P.gridSize < P.kh * 0.999999f) ++k;
  return k;
}

// Calls f(j) for every candidate j of target cell `hash` in the reference
// order: z outer, y, x inner over cells; ascending sorted index in a cell.
template <int DIM, class F>
inline void for_candidates(const sph_params& P, int hash, const int* cellStart,
                           const int* cellEnd, F&& f) {
  const int X = P.gridxdim, Y = P.gridydim, Z = P.gridzdim, XY = X * Y;
  const int cz = hash / XY, cy = (hash % XY) / X, cx = hash % X;
  const int K = stencil_reach(P);
  for (int dz = -K; dz <= K; ++dz)
    for (int dy = (DIM == 3 ? -K : 0); dy <= (DIM == 3 ? K : 0); ++dy)
      for (int dx = -K; dx <= K; ++dx) {
        const int x = cx + dx, y = cy + dy, z = cz + dz;
        if (x < 0 || x >= X || y < 0 || y >= Y || z < 0 || z >= Z) continue;
        const int c = z * XY + y * X + x;
        const int s = cellStart[c];
        if (s == -1) continue;
        const int e = cellEnd[c];
        for (int j = s; j < e; ++j) f(j);
      }
}

// Magnitude sums used as the tolerance scale of near-cancelling results.
struct Mags {
  double* drho;       // sum |continuity terms|
  double* acc;        // per component sum |vtmp * dx_k|  (3N)
  double* dofv;       // sum |divergence terms| (filled by kernel A, read by kernel B)
  double* sens_drho;  // absolute change of drho that 2-ulp input densities can cause
};

// Kernel A (opt_src_v1/equation/ns.cu:30-134).
template <class R, int DIM>
void kernelA(const sph_params& P, const R* sPos, const R* sVel, R* sRho, R* sP, const int* sType,
             const int* cellStart, const int* cellEnd, const int* hash, R* dofv, int n, Mags mags) {
  const R h = (R)P.h, adh = (R)P.adh, g = (R)P.gravity, m = (R)P.particleMass;
  std::vector<R> wall_p(n), wall_rho(n);
  std::vector<char> is_wall(n, 0);
#pragma omp parallel for schedule(dynamic, 256)
  for (int i = 0; i < n; ++i) {
    R rhop_sum = 0, w_sum = 0, dofv_tmp = 0;
    double mdv = 0;
    const R px = sPos[3 * i], py = sPos[3 * i + 1], pz = sPos[3 * i + 2];
    const R vx = sVel[3 * i], vy = sVel[3 * i + 1], vz = sVel[3 * i + 2];
    const int ti = sType[i];
    for_candidates<DIM>(P, hash[i], cellStart, cellEnd, [&](int j) {
      if (j == i) return;
      const R dx = px - sPos[3 * j], dy = py - sPos[3 * j + 1], dz = pz - sPos[3 * j + 2];
      const R pres1 = sP[j], dens1 = sRho[j];
      const R rr = dx * dx + dy * dy + dz * dz;
      const R dvdx = (vx - sVel[3 * j]) * dx + (vy - sVel[3 * j + 1]) * dy + (vz - sVel[3 * j + 2]) * dz;
      const R q = std::sqrt(rr) / h;
      if (q <= (R)2) {
        R fr = ((R)1 - q / (R)2) * ((R)1 - q / (R)2) * ((R)1 - q / (R)2);
        R w = fr * ((R)1 - q / (R)2);
        w *= ((R)2 * q + (R)1) * adh;
        fr *= (R)-5 * adh / (h * h);
        const int tj = sType[j];
        if (ti != 1 && tj == 1) {
          rhop_sum += (pres1 - dens1 * ((R)0 * dx + (R)0 * dy + ((R)0 - g) * dz)) * w;
          w_sum += w;
        } else if (ti == 1 && tj == 1) {
          dofv_tmp -= fr * dvdx * m / dens1;
          mdv += std::fabs((double)fr * dvdx * m / dens1);
        }
      }
    });
    dofv[i] = dofv_tmp;
    if (mags.dofv) mags.dofv[i] = mdv;
    if (ti != 1) {
      R pw = std::fabs(w_sum) > (R)1.0E-8f ? rhop_sum / w_sum : (R)0;
      if (pw < (R)0) pw = (R)0;
      wall_p[i] = pw;
      wall_rho[i] = eos_density<R>(P, pw);
      is_wall[i] = 1;
    }
  }
  // the reference writes wall values in place while other threads read only
  // fluid values, so deferring the writes changes nothing
  for (int i = 0; i < n; ++i)
    if (is_wall[i]) {
      sP[i] = wall_p[i];
      sRho[i] = wall_rho[i];
    }
}

// Kernel B (opt_src_v1/equation/ns.cu:137-230).
template <class R, int DIM>
void kernelB(const sph_params& P, const R* sPos, const R* sVel, const R* sRho, const R* sP,
             const int* sType, const R* dofv, const int* cellStart, const int* cellEnd,
             const int* hash, R* drho, R* acc, int n, Mags mags) {
  const R h = (R)P.h, adh = (R)P.adh, m = (R)P.particleMass, delta = (R)P.delta, cs = (R)P.cs;
  const R eta = (R)P.eta, rho0 = (R)P.restDensity, afa = (R)P.afa, damper = (R)P.damper;
#pragma omp parallel for schedule(dynamic, 256)
  for (int i = 0; i < n; ++i) {
    const R px = sPos[3 * i], py = sPos[3 * i + 1], pz = sPos[3 * i + 2];
    const R vx = sVel[3 * i], vy = sVel[3 * i + 1], vz = sVel[3 * i + 2];
    const R pres0 = sP[i], dens0 = sRho[i], dofv0 = dofv[i];
    const int ti = sType[i];
    R dd = 0, ax = 0, ay = 0, az = 0;
    double md = 0, mx = 0, my = 0, mz = 0, sd = 0;
    for_candidates<DIM>(P, hash[i], cellStart, cellEnd, [&](int j) {
      if (j == i) return;
      const R dx = px - sPos[3 * j], dy = py - sPos[3 * j + 1], dz = pz - sPos[3 * j + 2];
      const R dvdx = (vx - sVel[3 * j]) * dx + (vy - sVel[3 * j + 1]) * dy + (vz - sVel[3 * j + 2]) * dz;
      const R pres1 = sP[j], dens1 = sRho[j], dofv1 = dofv[j];
      const R rr = dx * dx + dy * dy + dz * dz;
      const R q = std::sqrt(rr) / h;
      if (q <= (R)2) {
        const R fr = (R)-5 * adh * ((R)1 - q / (R)2) * ((R)1 - q / (R)2) * ((R)1 - q / (R)2) / (h * h);
        const R t1 = dens0 * dvdx * fr;
        const R t2 = (dens0 - dens1) * rr * fr * delta * h * cs / (rr + eta * eta);
        dd += (t1 + t2) * m / dens1;
        // the diffusion term takes the difference of two FP32 densities near
        // rho0, so a one-ulp change of either input (e.g. a wall density from
        // kernel A) moves it by |d t2 / d rho| * ulp: part of the scale
        // (an absolute error budget, not a magnitude: near rho0 one ulp of a
        // wall density moves drho by far more than 1e-5 of the term sums).
        // Only wall densities are computed here; fluid densities are inputs.
        if (ti != 1 || sType[j] != 1) {
          const double ulp_rho = 2.0 * std::ldexp(std::max(std::fabs((double)dens0), std::fabs((double)dens1)), -23);
          sd += std::fabs((double)rr * fr * delta * h * cs / (rr + eta * eta)) * ulp_rho * (double)m / (double)dens1;
        }
        md += (std::fabs((double)t1) + std::fabs((double)t2)) * (double)m / (double)dens1;
        R vtmp = -(pres0 + pres1) * m * fr / (dens0 * dens1);
        double vmag = std::fabs((double)vtmp);
        if (ti == 1 && sType[j] == 1) {
          const R av = h * cs * rho0 * m * fr * (damper * (dofv0 + dofv1) + afa * dvdx / (rr + eta * eta)) /
                       (dens0 * dens1);
          vtmp += av;
          const double s0 = mags.dofv ? std::max(std::fabs((double)dofv0), mags.dofv[i]) : std::fabs((double)dofv0);
          const double s1 = mags.dofv ? std::max(std::fabs((double)dofv1), mags.dofv[j]) : std::fabs((double)dofv1);
          vmag += std::fabs((double)h * cs * rho0 * m * fr / (dens0 * dens1)) *
                  (std::fabs((double)damper) * (s0 + s1) + std::fabs((double)afa * dvdx / (rr + eta * eta)));
        }
        ax += vtmp * dx;
        ay += vtmp * dy;
        az += vtmp * dz;
        mx += vmag * std::fabs((double)dx);
        my += vmag * std::fabs((double)dy);
        mz += vmag * std::fabs((double)dz);
      }
    });
    drho[i] = dd;
    acc[3 * i] = ax;
    acc[3 * i + 1] = ay;
    acc[3 * i + 2] = az + (R)P.gravity;
    if (mags.drho) mags.drho[i] = md;
    if (mags.sens_drho) mags.sens_drho[i] = sd;
    if (mags.acc) {
      mags.acc[3 * i] = mx;
      mags.acc[3 * i + 1] = my;
      mags.acc[3 * i + 2] = mz + std::fabs((double)P.gravity);
    }
  }
}

template <class R>
void run_ns(const sph_params& P, const R* sPos, const R* sVel, R* sRho, R* sP, const int* sType,
            R* drho, R* acc, const int* cellStart, const int* cellEnd, const int* hash, R* dofv, int n,
            Mags mags) {
  if (P.dim == 2) {
    kernelA<R, 2>(P, sPos, sVel, sRho, sP, sType, cellStart, cellEnd, hash, dofv, n, mags);
    kernelB<R, 2>(P, sPos, sVel, sRho, sP, sType, dofv, cellStart, cellEnd, hash, drho, acc, n, mags);
  } else {
    kernelA<R, 3>(P, sPos, sVel, sRho, sP, sType, cellStart, cellEnd, hash, dofv, n, mags);
    kernelB<R, 3>(P, sPos, sVel, sRho, sP, sType, dofv, cellStart, cellEnd, hash, drho, acc, n, mags);
  }
}

// Stable ascending sort of (key, index) pairs.  A counting sort over the cell
// range: the output is the unique stable order, the same as any stable sort.
void stable_sort_pairs(const sph_params& P, int* key, int* val, int n) {
  const int C = P.numCells;
  std::vector<int> cnt((size_t)C + 1, 0);
  for (int i = 0; i < n; ++i) cnt[(size_t)key[i] + 1]++;
  for (int c = 0; c < C; ++c) cnt[(size_t)c + 1] += cnt[c];
  std::vector<int> k2(n), v2(n);
  for (int i = 0; i < n; ++i) {
    const int d = cnt[key[i]]++;
    k2[d] = key[i];
    v2[d] = val[i];
  }
  std::memcpy(key, k2.data(), sizeof(int) * (size_t)n);
  std::memcpy(val, v2.data(), sizeof(int) * (size_t)n);
}

// reorder (opt_src_v0/lib/mesh.cu:106-177): cellStart = -1 everywhere, then
// cell ranges at key changes; gather into sorted order.
void cell_ranges(const sph_params& P, const int* hash, int* cellStart, int* cellEnd, int n) {
  std::fill(cellStart, cellStart + P.numCells, -1);
  for (int i = 0; i < n; ++i) {
    const int h = hash[i];
    if (i == 0 || h != hash[i - 1]) {
      cellStart[h] = i;
      if (i > 0) cellEnd[hash[i - 1]] = i;
    }
    if (i == n - 1) cellEnd[h] = i + 1;
  }
}

inline bool outside(const sph_params& P, double x, double y, double z) {
  return x > P.domain_max[0] || x < P.domain_min[0] || y > P.domain_max[1] || y < P.domain_min[1] ||
         z > P.domain_max[2] || z < P.domain_min[2];
}

// --------------------------------------------------------------- full state
template <class R, class X>
struct State {
  std::vector<X> pos;   // integrator positions (3N)
  std::vector<R> posR;  // positions used by hashing and the neighbour loops
  std::vector<R> vel, rho, p;
  std::vector<int> type, id;
  void resize(int n) {
    pos.resize(3 * (size_t)n);
    posR.resize(3 * (size_t)n);
    vel.resize(3 * (size_t)n);
    rho.resize(n);
    p.resize(n);
    type.resize(n);
    id.resize(n);
  }
};

template <class R, class X>
struct Sim {
  sph_params P;
  int n;
  State<R, X> s[2];
  std::vector<X> tpos;
  std::vector<R> tvel, trho;
  std::vector<int> key, idx, cellStart, cellEnd;
  std::vector<R> drho, acc, dofv;

  void init(const sph_params& prm, int nn, const float* pos, const float* vel, const float* rho,
            const float* p, const int* type) {
    P = prm;
    n = nn;
    for (auto& st : s) st.resize(n);
    for (int i = 0; i < n; ++i) {
      for (int k = 0; k < 3; ++k) {
        s[0].pos[3 * i + k] = (X)pos[3 * i + k];
        s[0].posR[3 * i + k] = (R)pos[3 * i + k];
        s[0].vel[3 * i + k] = (R)vel[3 * i + k];
      }
      s[0].rho[i] = (R)rho[i];
      s[0].p[i] = (R)p[i];
      s[0].type[i] = type[i];
      s[0].id[i] = i;
    }
    tpos.resize(3 * (size_t)n);
    tvel.resize(3 * (size_t)n);
    trho.resize(n);
    key.resize(n);
    idx.resize(n);
    cellStart.resize(P.numCells);
    cellEnd.assign(P.numCells, 0);
    drho.resize(n);
    acc.resize(3 * (size_t)n);
    dofv.resize(n);
  }

  // hash -> stable sort -> cell ranges -> gather (from set a into set b)
  void sort_into(int a, int b) {
    State<R, X>& in = s[a];
    State<R, X>& out = s[b];
#pragma omp parallel for
    for (int i = 0; i < n; ++i) {
      key[i] = cell_hash<R>(P, in.posR[3 * i], in.posR[3 * i + 1], in.posR[3 * i + 2]);
      idx[i] = i;
    }
    stable_sort_pairs(P, key.data(), idx.data(), n);
    cell_ranges(P, key.data(), cellStart.data(), cellEnd.data(), n);
#pragma omp parallel for
    for (int i = 0; i < n; ++i) {
      const int src = idx[i];
      for (int k = 0; k < 3; ++k) {
        out.pos[3 * i + k] = in.pos[3 * src + k];
        out.posR[3 * i + k] = in.posR[3 * src + k];
        out.vel[3 * i + k] = in.vel[3 * src + k];
      }
      out.rho[i] = in.rho[src];
      out.p[i] = in.p[src];
      out.type[i] = in.type[src];
      out.id[i] = in.id[src];
    }
  }

  void rates(int b, Mags mags) {
    State<R, X>& st = s[b];
    run_ns<R>(P, st.posR.data(), st.vel.data(), st.rho.data(), st.p.data(), st.type.data(), drho.data(),
              acc.data(), cellStart.data(), cellEnd.data(), key.data(), dofv.data(), n, mags);
  }

  // PC_predictionD / PC_correctionD with the precision policy of (R, X).
  void integrate(int b, bool pred) {
    State<R, X>& st = s[b];
    const R dt = (R)P.timeStep;
#pragma omp parallel for
    for (int i = 0; i < n; ++i) {
      if (st.type[i] != 1) continue;
      if (outside(P, (double)st.pos[3 * i], (double)st.pos[3 * i + 1], (double)st.pos[3 * i + 2])) {
        for (int k = 0; k < 3; ++k) {
          st.pos[3 * i + k] = (X)P.park_pos[k];
          st.posR[3 * i + k] = (R)P.park_pos[k];
          st.vel[3 * i + k] = 0;
        }
        if (P.dim == 2) {
          st.pos[3 * i + 1] = 0;
          st.posR[3 * i + 1] = 0;
        }
        st.p[i] = 0;
        st.rho[i] = (R)P.restDensity;
        st.type[i] = P.park_type;
        continue;
      }
      const int r = pred ? i : idx[i];
      if (pred) {
        for (int k = 0; k < 3; ++k) {
          tpos[3 * r + k] = st.pos[3 * i + k];
          tvel[3 * r + k] = st.vel[3 * i + k];
        }
        trho[r] = st.rho[i];
      }
      R rho;
      X x[3];
      R v[3];
      if constexpr (sizeof(R) == 4) {
        // float fields: the rate times dt is a float product, halved and added in double
        rho = (R)((double)trho[r] + (double)(drho[i] * dt) / 2.0);
        for (int k = 0; k < 3; ++k) v[k] = (R)((double)tvel[3 * r + k] + (double)(acc[3 * i + k] * dt) / 2.0);
        if constexpr (sizeof(X) == 4) {
          for (int k = 0; k < 3; ++k) x[k] = (X)((double)tpos[3 * r + k] + (double)(st.vel[3 * i + k] * dt) / 2.0);
        } else {
          const double hdt = 0.5 * (double)dt;
          for (int k = 0; k < 3; ++k) x[k] = tpos[3 * r + k] + (double)st.vel[3 * i + k] * hdt;
        }
      } else {
        rho = trho[r] + drho[i] * dt / 2.0;
        for (int k = 0; k < 3; ++k) v[k] = tvel[3 * r + k] + acc[3 * i + k] * dt / 2.0;
        for (int k = 0; k < 3; ++k) x[k] = tpos[3 * r + k] + st.vel[3 * i + k] * dt / 2.0;
      }
      if (rho < (R)P.restDensity) rho = (R)P.restDensity;
      if (!pred) {
        rho = 2 * rho - trho[r];
        if (rho < (R)P.restDensity) rho = (R)P.restDensity;
        for (int k = 0; k < 3; ++k) {
          x[k] = 2 * x[k] - tpos[3 * r + k];
          v[k] = 2 * v[k] - tvel[3 * r + k];
        }
      }
      if (P.dim == 2) {
        x[1] = 0;
        v[1] = 0;
      }
      st.rho[i] = rho;
      for (int k = 0; k < 3; ++k) {
        st.pos[3 * i + k] = x[k];
        st.posR[3 * i + k] = (R)x[k];
        st.vel[3 * i + k] = v[k];
      }
      st.p[i] = eos_pressure<R>(P, rho);
    }
  }

  void step() {
    Mags none{nullptr, nullptr, nullptr, nullptr};
    sort_into(0, 1);
    rates(1, none);
    integrate(1, true);
    sort_into(1, 0);
    rates(0, none);
    integrate(0, false);
  }

  void probe(double band, double* front, double* height) const {
    double f = 0, h = 0;
    const State<R, X>& st = s[0];
    for (int i = 0; i < n; ++i) {
      if (st.type[i] != 1) continue;
      const double x = (double)st.pos[3 * i], y = (double)st.pos[3 * i + 1], z = (double)st.pos[3 * i + 2];
      if (outside(P, x, y, z)) continue;
      f = std::max(f, x);
      if (x <= band) h = std::max(h, z);
    }
    *front = f;
    *height = h;
  }

  void download(float* pos, double* pos64, float* vel, float* rho, float* p, int* type) const {
    const State<R, X>& st = s[0];
    for (int i = 0; i < n; ++i) {
      const int o = st.id[i];
      for (int k = 0; k < 3; ++k) {
        if (pos) pos[3 * o + k] = (float)st.pos[3 * i + k];
        if (pos64) pos64[3 * o + k] = (double)st.pos[3 * i + k];
        if (vel) vel[3 * o + k] = (float)st.vel[3 * i + k];
      }
      if (rho) rho[o] = (float)st.rho[i];
      if (p) p[o] = (float)st.p[i];
      if (type) type[o] = st.type[i];
    }
  }
};

template <class R, class X>
int run_mode(const sph_params* P, int n, float* pos, double* pos64, float* vel, float* rho, float* p,
             int* type, int nsteps, double band, int probe_every, double* front, double* height,
             double* seconds) {
  Sim<R, X> sim;
  sim.init(*P, n, pos, vel, rho, p, type);
  int nprobe = 0;
  double t0 = omp_get_wtime();
  for (int k = 0; k < nsteps; ++k) {
    sim.step();
    if (front && probe_every > 0 && (k + 1) % probe_every == 0) {
      sim.probe(band, &front[nprobe], &height[nprobe]);
      ++nprobe;
    }
  }
  double t1 = omp_get_wtime();
  if (seconds) *seconds = t1 - t0;
  sim.download(pos, pos64, vel, rho, p, type);
  return nprobe;
}

template <class R, class X>
void eval_mode(const sph_params* P, int n, const float* pos, const float* vel, const float* rho,
               const float* p, const int* type, float* drho, float* acc, float* dofv, float* p_out,
               float* rho_out, int* keys, int* perm, double* mag_drho, double* mag_acc, double* mag_dofv,
               double* sens_drho) {
  Sim<R, X> sim;
  sim.init(*P, n, pos, vel, rho, p, type);
  sim.sort_into(0, 1);
  std::vector<double> md(n), ma(3 * (size_t)n), mdv(n), sd(n);
  sim.rates(1, Mags{md.data(), ma.data(), mdv.data(), sd.data()});
  const State<R, X>& st = sim.s[1];
  for (int i = 0; i < n; ++i) {
    const int o = st.id[i];
    if (drho) drho[o] = (float)sim.drho[i];
    if (dofv) dofv[o] = (float)sim.dofv[i];
    if (p_out) p_out[o] = (float)st.p[i];
    if (rho_out) rho_out[o] = (float)st.rho[i];
    for (int k = 0; k < 3; ++k) {
      if (acc) acc[3 * o + k] = (float)sim.acc[3 * i + k];
      if (mag_acc) mag_acc[3 * o + k] = ma[3 * (size_t)i + k];
    }
    if (mag_drho) mag_drho[o] = md[i];
    if (mag_dofv) mag_dofv[o] = mdv[i];
    if (sens_drho) sens_drho[o] = sd[i];
    if (keys) keys[i] = sim.key[i];
    if (perm) perm[i] = sim.idx[i];
  }
}

}  // namespace

extern "C" {

int orc_num_threads(int t) {
  if (t > 0) omp_set_num_threads(t);
  return omp_get_max_threads();
}

// ---- per-call restatements on the reference layouts (FP32) -----------------
void orc_calcHash(const sph_params* P, const float* pos, int* hash, int* idx, int n) {
  for (int i = 0; i < n; ++i) {
    hash[i] = cell_hash<float>(*P, pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
    idx[i] = i;
  }
}

void orc_sortParticles(const sph_params* P, int* hash, int* idx, int n) {
  stable_sort_pairs(*P, hash, idx, n);
}

void orc_reorder(const sph_params* P, const float* oPos, const float* oVel, const float* oRho,
                 const float* oP, const int* oType, float* sPos, float* sVel, float* sRho, float* sP,
                 int* sType, const int* hash, const int* idx, int* cellStart, int* cellEnd, int n,
                 int* table_l, const int* table_r) {
  cell_ranges(*P, hash, cellStart, cellEnd, n);
  for (int i = 0; i < n; ++i) {
    const int src = idx[i];
    for (int k = 0; k < 3; ++k) {
      sPos[3 * i + k] = oPos[3 * src + k];
      sVel[3 * i + k] = oVel[3 * src + k];
    }
    sRho[i] = oRho[src];
    sP[i] = oP[src];
    sType[i] = oType[src];
    if (table_l && table_r) table_l[i] = table_r[src] + (src - i);
  }
}

void orc_SPH_NS(const sph_params* P, const float* sPos, const float* sVel, float* sRho, float* sP,
                const int* sType, float* drho, float* acc, const int* cellStart, const int* cellEnd,
                const int* hash, float* dofv, int n, double* mag_drho, double* mag_acc,
                double* mag_dofv, double* sens_drho) {
  std::vector<double> tmp;
  if (!mag_dofv && (mag_drho || mag_acc)) {
    tmp.resize(n);
    mag_dofv = tmp.data();
  }
  run_ns<float>(*P, sPos, sVel, sRho, sP, sType, drho, acc, cellStart, cellEnd, hash, dofv, n,
                Mags{mag_drho, mag_acc, mag_dofv, sens_drho});
}

// PC_prediction (pred=1) / PC_correction (pred=0) on the reference layout.
void orc_PC(const sph_params* Pp, int pred, float* sPos, float* sVel, float* sRho, float* sP,
            const float* drho, const float* acc, const int* idx, float* tPos, float* tVel, float* tRho,
            int n, int* sType) {
  const sph_params& P = *Pp;
  const float dt = P.timeStep;
  for (int i = 0; i < n; ++i) {
    if (sType[i] != 1) continue;
    if (outside(P, (double)sPos[3 * i], (double)sPos[3 * i + 1], (double)sPos[3 * i + 2])) {
      for (int k = 0; k < 3; ++k) {
        sPos[3 * i + k] = P.park_pos[k];
        sVel[3 * i + k] = 0.0f;
      }
      sP[i] = 0.0f;
      sRho[i] = P.restDensity;
      sType[i] = P.park_type;
      continue;
    }
    const int r = pred ? i : idx[i];
    if (pred) {
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
    if (!pred) {
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
    sP[i] = eos_pressure<float>(P, rho);
  }
}

float orc_eos_pressure(const sph_params* P, float rho) { return eos_pressure<float>(*P, rho); }
float orc_eos_density(const sph_params* P, float p) { return eos_density<float>(*P, p); }
double orc_eos_pressure64(const sph_params* P, double rho) { return eos_pressure<double>(*P, rho); }
double orc_eos_density64(const sph_params* P, double p) { return eos_density<double>(*P, p); }

// ---- whole steps -------------------------------------------------------------
// State arrays in original particle order, updated in place.  front/height
// receive a probe every probe_every steps.  Returns the number of probes.
int orc_run(const sph_params* P, int mode, int n, float* pos, double* pos64, float* vel, float* rho,
            float* p, int* type, int nsteps, double band, int probe_every, double* front, double* height,
            double* seconds) {
  switch (mode) {
    case 0:
      return run_mode<float, float>(P, n, pos, pos64, vel, rho, p, type, nsteps, band, probe_every, front,
                                    height, seconds);
    case 1:
      return run_mode<float, double>(P, n, pos, pos64, vel, rho, p, type, nsteps, band, probe_every, front,
                                     height, seconds);
    default:
      return run_mode<double, double>(P, n, pos, pos64, vel, rho, p, type, nsteps, band, probe_every, front,
                                      height, seconds);
  }
}

// Rates of the state (original order in and out): hash, sort, gather, kernels
// A and B.  keys/perm come out in sorted order.
void orc_eval(const sph_params* P, int mode, int n, const float* pos, const float* vel, const float* rho,
              const float* p, const int* type, float* drho, float* acc, float* dofv, float* p_out,
              float* rho_out, int* keys, int* perm, double* mag_drho, double* mag_acc,
              double* mag_dofv, double* sens_drho) {
  if (mode == 2)
    eval_mode<double, double>(P, n, pos, vel, rho, p, type, drho, acc, dofv, p_out, rho_out, keys, perm,
                              mag_drho, mag_acc, mag_dofv, sens_drho);
  else
    eval_mode<float, double>(P, n, pos, vel, rho, p, type, drho, acc, dofv, p_out, rho_out, keys, perm,
                             mag_drho, mag_acc, mag_dofv, sens_drho);
}

}  // extern "C"
