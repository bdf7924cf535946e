// Host-side parameter builders and the dam-break case generator.
//
// sph_params_reference follows the reference host recipe
// (src/sph/sph_run.cu:11-22 constants, :60-101 derived values) including its
// float/double evaluation order, so the derived floats (h, kh, cell counts,
// particle mass, kernel constant) are the reference's values.
//
// sph_params_dambreak and sph_case_dambreak build the lattice dam-break cases
// (2-D and 3-D, Tait or linear EOS, CFL time step); they have no reference
// counterpart because the reference reads its initial state from .dat files
// that it does not ship (src/io/io.cu:6-111).
#include "sph_internal.h"

#include <cmath>
#include <cstring>

namespace sph {

// Largest float rr for which sqrtf(rr)/h <= 2 holds in IEEE FP32 (correctly
// rounded sqrt and division).  The predicate is monotone in rr, so the
// neighbour test `rr <= rr_cut` keeps exactly the pairs the reference keeps
// with `q = sqrtf(rr)/par.h; if (q <= 2.0f)` (opt_src_v1/equation/ns.cu:87-90).
float cutoff_rr(float h) {
  auto ok = [h](float rr) {
    volatile float s = std::sqrt(rr);
    volatile float q = s / h;
    return q <= 2.0f;
  };
  float rr = (2.0f * h) * (2.0f * h);
  if (ok(rr)) {
    for (;;) {
      float nx = std::nextafter(rr, INFINITY);
      if (!ok(nx)) break;
      rr = nx;
    }
  } else {
    while (!ok(rr)) rr = std::nextafter(rr, 0.0f);
  }
  return rr;
}

}  // namespace sph

using namespace sph;

extern "C" int sph_params_reference(sph_params* out, float dx_in, int numParticles, int numFluid) {
  if (!out || !(dx_in > 0.0f)) return set_error(SPH_E_INVAL, "sph_params_reference: bad argument");
  std::memset(out, 0, sizeof(*out));
  // src/sph/sph_run.cu:11-16 -- float locals
  float vlx = 3.221f, vly = 1.0f, vlz = 1.0f;
  float dx = dx_in;
  float rho0 = 1000.0f;
  float g = -9.81f, cs = 25.0f;
  float dt = 2e-4f;
  const double PI = 3.141592654;  // src/header/sph.cuh:4
  sph_params* t = out;
  t->dim = 3;
  t->timeStep = dt;
  t->cs = cs;
  t->numParticles = numParticles;
  t->numFluid = numFluid;
  t->gravity = g;
  t->particleR = dx;
  t->h = (float)(1.5 * dx);              // :67 (double product stored as float)
  t->kh = (float)(2.0 * t->h);           // :68
  t->eta = (float)(0.1 * t->h);          // :69
  t->delta = (float)0.2;                 // :70
  t->gridSize = t->kh;                   // :71
  t->restDensity = rho0;                 // :72
  t->particleMass = (float)(std::pow((double)t->particleR, 3) * t->restDensity);  // :73
  t->adh = (float)(21.0 / (16.0 * PI * t->h * t->h * t->h));                      // :74
  float xmin = (float)-0.1, ymin = (float)-0.1, zmin = (float)-0.1;               // :76-78
  float xmax = (float)(vlx + 0.1), ymax = (float)(vly + 0.1), zmax = (float)(vlz + 0.1);  // :79-81
  t->gridxmin = xmin;
  t->gridymin = ymin;
  t->gridzmin = zmin;
  t->gridxdim = (int)std::ceil((xmax - xmin) / t->gridSize);  // :88-90 (float arithmetic)
  t->gridydim = (int)std::ceil((ymax - ymin) / t->gridSize);
  t->gridzdim = (int)std::ceil((zmax - zmin) / t->gridSize);
  t->hash_max = t->gridxdim * t->gridydim * t->gridzdim + t->gridxdim * t->gridydim + t->gridxdim;  // :92
  t->numCells = t->gridxdim * t->gridydim * t->gridzdim;  // :93
  t->afa = (float)0.05;                                    // :100
  t->eos = SPH_EOS_LINEAR;
  t->gamma = 1.0f;
  t->tait_B = 0.0f;
  t->damper = 1.0f;
  t->rr_cut = cutoff_rr(t->h);
  // escape box and parking position, src/sph/sph_run.cu:437-447
  t->domain_min[0] = 0.0; t->domain_max[0] = 3.23;
  t->domain_min[1] = 0.0; t->domain_max[1] = 1.0;
  t->domain_min[2] = 0.0; t->domain_max[2] = 1.0;
  t->park_pos[0] = 0.0f; t->park_pos[1] = 0.5f; t->park_pos[2] = -0.08f;
  t->park_type = 1;
  return SPH_OK;
}

namespace {

struct Lattice {
  long long nx, ny, nz;     // tank cells per axis
  long long wx, wy, wz;     // column cells per axis
  int L;
};

int lattice_of(const sph_dambreak_cfg* c, Lattice* out) {
  if (!c || (c->dim != 2 && c->dim != 3) || !(c->dx > 0) || c->wall_layers < 0)
    return set_error(SPH_E_INVAL, "dambreak cfg: dim must be 2 or 3, dx > 0, wall_layers >= 0");
  Lattice l;
  l.nx = std::llround(c->tank[0] / c->dx);
  l.ny = c->dim == 3 ? std::llround(c->tank[1] / c->dx) : 1;
  l.nz = std::llround(c->tank[2] / c->dx);
  l.wx = std::llround(c->column[0] / c->dx);
  l.wy = c->dim == 3 ? std::llround(c->column[1] / c->dx) : 1;
  l.wz = std::llround(c->column[2] / c->dx);
  l.L = c->wall_layers;
  if (l.nx <= 0 || l.ny <= 0 || l.nz <= 0 || l.wx <= 0 || l.wy <= 0 || l.wz <= 0 ||
      l.wx > l.nx || l.wy > l.ny || l.wz > l.nz)
    return set_error(SPH_E_INVAL, "dambreak cfg: column must fit in the tank and both be non-empty");
  *out = l;
  return SPH_OK;
}

long long wall_count(const Lattice& l, int dim) {
  const long long L = l.L;
  if (dim == 2) return (l.nx + 2 * L) * L + 2 * l.nz * L;
  // floor over the extended footprint, y-faces over the extended x range,
  // x-faces between the y-faces
  return (l.nx + 2 * L) * (l.ny + 2 * L) * L + 2 * (l.nx + 2 * L) * l.nz * L + 2 * l.ny * l.nz * L;
}

double column_height(const sph_dambreak_cfg* c, const Lattice& l) { return (double)l.wz * c->dx; }

double sound_speed(const sph_dambreak_cfg* c, const Lattice& l) {
  if (c->c0 > 0) return c->c0;
  return 10.0 * std::sqrt(2.0 * std::fabs(c->g) * column_height(c, l));
}

}  // namespace

extern "C" int sph_case_dambreak_count(const sph_dambreak_cfg* cfg, int* n_total, int* n_fluid) {
  Lattice l;
  int rc = lattice_of(cfg, &l);
  if (rc) return rc;
  long long nf = l.wx * l.wy * l.wz;
  long long nt = nf + wall_count(l, cfg->dim);
  if (nt > 0x7fffffffLL) return set_error(SPH_E_INVAL, "dambreak cfg: more than 2^31-1 particles");
  if (n_total) *n_total = (int)nt;
  if (n_fluid) *n_fluid = (int)nf;
  return SPH_OK;
}

extern "C" int sph_params_dambreak(const sph_dambreak_cfg* c, sph_params* t) {
  if (!t) return set_error(SPH_E_INVAL, "sph_params_dambreak: null output");
  Lattice l;
  int rc = lattice_of(c, &l);
  if (rc) return rc;
  int nt = 0, nf = 0;
  rc = sph_case_dambreak_count(c, &nt, &nf);
  if (rc) return rc;
  std::memset(t, 0, sizeof(*t));
  const double dx = c->dx;
  const double h = c->h_over_dx * dx;
  const double c0 = sound_speed(c, l);
  t->dim = c->dim;
  t->numParticles = nt;
  t->numFluid = nf;
  t->particleR = (float)dx;
  t->h = (float)h;
  t->kh = (float)(2.0 * h);
  t->eta = (float)(0.1 * h);
  t->delta = (float)c->delta;
  t->afa = (float)c->alpha;
  t->restDensity = (float)c->rho0;
  t->particleMass = (float)(std::pow(dx, c->dim) * c->rho0);
  t->adh = c->dim == 3 ? (float)(21.0 / (16.0 * M_PI * h * h * h)) : (float)(7.0 / (4.0 * M_PI * h * h));
  t->cs = (float)c0;
  t->gravity = (float)c->g;
  t->timeStep = (float)(c->cfl * h / c0);
  t->eos = c->eos;
  t->gamma = (float)c->gamma;
  t->tait_B = (float)(c->rho0 * c0 * c0 / c->gamma);
  t->damper = 1.0f;
  t->rr_cut = cutoff_rr(t->h);
  double cell = 2.0 * h * (c->cell_over_kh > 0 ? c->cell_over_kh : 1.0);
  t->gridSize = (float)cell;
  const double pad = c->grid_pad;
  t->gridxmin = (float)-pad;
  t->gridzmin = (float)-pad;
  t->gridxdim = (int)std::ceil((c->tank[0] + 2 * pad) / (double)t->gridSize);
  t->gridzdim = (int)std::ceil((c->tank[2] + 2 * pad) / (double)t->gridSize);
  if (c->dim == 3) {
    t->gridymin = (float)-pad;
    t->gridydim = (int)std::ceil((c->tank[1] + 2 * pad) / (double)t->gridSize);
  } else {
    t->gridymin = 0.0f;
    t->gridydim = 1;
  }
  long long cells = (long long)t->gridxdim * t->gridydim * t->gridzdim;
  if (cells > (1LL << 30)) return set_error(SPH_E_INVAL, "dambreak cfg: cell grid too large");
  t->numCells = (int)cells;
  t->hash_max = t->gridxdim * t->gridydim * t->gridzdim + t->gridxdim * t->gridydim + t->gridxdim;
  t->domain_min[0] = 0.0; t->domain_max[0] = c->tank[0];
  if (c->dim == 3) {
    t->domain_min[1] = 0.0; t->domain_max[1] = c->tank[1];
  } else {
    t->domain_min[1] = -1.0; t->domain_max[1] = 1.0;
  }
  t->domain_min[2] = 0.0; t->domain_max[2] = c->tank[2];
  // parked below the floor inside the wall layers, like the reference's z = -0.08
  t->park_pos[0] = 0.0f;
  t->park_pos[1] = c->dim == 3 ? (float)(0.5 * c->tank[1]) : 0.0f;
  t->park_pos[2] = (float)(-0.5 * pad);
  t->park_type = 1;
  return SPH_OK;
}

extern "C" int sph_case_dambreak(const sph_dambreak_cfg* c, int cap, float* pos, float* vel, float* rho,
                                 float* p, int* type, sph_params* p_out, int* n_total, int* n_fluid) {
  Lattice l;
  int rc = lattice_of(c, &l);
  if (rc) return rc;
  int nt = 0, nf = 0;
  rc = sph_case_dambreak_count(c, &nt, &nf);
  if (rc) return rc;
  if (cap < nt || !pos || !vel || !rho || !p || !type)
    return set_error(SPH_E_INVAL, "sph_case_dambreak: output capacity too small or null array");
  sph_params prm;
  rc = sph_params_dambreak(c, &prm);
  if (rc) return rc;
  const double dx = c->dx;
  const double rho0 = c->rho0;
  const double c0 = prm.cs;
  const double H = column_height(c, l);
  const double gabs = std::fabs(c->g);
  const double B = rho0 * (double)c0 * (double)c0 / c->gamma;
  long long k = 0;
  auto put = [&](double x, double y, double z, int ty, double pr) {
    pos[3 * k] = (float)x;
    pos[3 * k + 1] = (float)y;
    pos[3 * k + 2] = (float)z;
    vel[3 * k] = vel[3 * k + 1] = vel[3 * k + 2] = 0.0f;
    double r;
    if (c->eos == SPH_EOS_TAIT)
      r = rho0 * std::pow(1.0 + pr / B, 1.0 / c->gamma);
    else
      r = rho0 + pr / ((double)c0 * (double)c0);
    p[k] = (float)pr;
    rho[k] = (float)r;
    type[k] = ty;
    ++k;
  };
  const long long L = l.L;
  auto at = [dx](long long i) { return ((double)i + 0.5) * dx; };
  // walls first (the reference numbers boundary particles before fluid)
  if (c->dim == 2) {
    for (long long layer = 0; layer < L; ++layer)
      for (long long i = -L; i < l.nx + L; ++i) put(at(i), 0.0, -at(layer), 0, 0.0);
    for (long long layer = 0; layer < L; ++layer)
      for (long long j = 0; j < l.nz; ++j) {
        put(-at(layer), 0.0, at(j), 0, 0.0);
        put((double)l.nx * dx + at(layer), 0.0, at(j), 0, 0.0);
      }
  } else {
    for (long long layer = 0; layer < L; ++layer)
      for (long long jy = -L; jy < l.ny + L; ++jy)
        for (long long i = -L; i < l.nx + L; ++i) put(at(i), at(jy), -at(layer), 0, 0.0);
    for (long long layer = 0; layer < L; ++layer)
      for (long long jz = 0; jz < l.nz; ++jz)
        for (long long i = -L; i < l.nx + L; ++i) {
          put(at(i), -at(layer), at(jz), 0, 0.0);
          put(at(i), (double)l.ny * dx + at(layer), at(jz), 0, 0.0);
        }
    for (long long layer = 0; layer < L; ++layer)
      for (long long jz = 0; jz < l.nz; ++jz)
        for (long long jy = 0; jy < l.ny; ++jy) {
          put(-at(layer), at(jy), at(jz), 0, 0.0);
          put((double)l.nx * dx + at(layer), at(jy), at(jz), 0, 0.0);
        }
  }
  // fluid column at rest, hydrostatic pressure rho0 |g| (H - z)
  for (long long jz = 0; jz < l.wz; ++jz)
    for (long long jy = 0; jy < l.wy; ++jy)
      for (long long i = 0; i < l.wx; ++i) {
        double z = at(jz);
        double y = c->dim == 3 ? at(jy) : 0.0;
        put(at(i), y, z, 1, rho0 * gabs * (H - z));
      }
  if (k != nt) return set_error(SPH_E_STATE, "sph_case_dambreak: internal count mismatch");
  if (p_out) *p_out = prm;
  if (n_total) *n_total = nt;
  if (n_fluid) *n_fluid = nf;
  return SPH_OK;
}

// Finer cell grid over the same box (no reference counterpart; the
// reference's last version also searches a grid finer than 2h,
// opt_src_v4/sph/sph_run.cu:73): the cell edge is divided by div (a power of
// two, so gridSize / div is exact and every fine cell lies inside one cell of
// the input grid), the cell counts are multiplied by div (y stays 1 in 2-D).
extern "C" int sph_params_subdivide(const sph_params* in, int div, sph_params* out) {
  if (!in || !out) return set_error(SPH_E_INVAL, "sph_params_subdivide: null argument");
  if (div != 1 && div != 2 && div != 4) return set_error(SPH_E_INVAL, "sph_params_subdivide: div must be 1, 2 or 4");
  sph_params t = *in;
  t.gridSize = in->gridSize / (float)div;
  t.gridxdim = in->gridxdim * div;
  t.gridydim = in->dim == 3 ? in->gridydim * div : in->gridydim;
  t.gridzdim = in->gridzdim * div;
  const long long cells = (long long)t.gridxdim * t.gridydim * t.gridzdim;
  if (cells > (1LL << 30)) return set_error(SPH_E_INVAL, "sph_params_subdivide: cell grid too large");
  t.numCells = (int)cells;
  const long long hm = cells + (long long)t.gridxdim * t.gridydim + t.gridxdim;
  t.hash_max = hm > 0x7fffffffLL ? 0x7fffffff : (int)hm;
  *out = t;
  return SPH_OK;
}
