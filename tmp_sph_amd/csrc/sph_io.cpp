// Host-side file formats of the reference's I/O layer:
//  * the initial-state loader `ini_fluid` (src/io/io.cu:6-111): one text
//    file per field in a directory, whitespace-separated values --
//    h_pos.dat, h_vel.dat, h_ac.dat (3 per particle), rhop.dat, p.dat (1 per
//    particle), particle_type.dat, particle_zone.dat (integers);
//  * the legacy VTK frames of `output_fluid_file` (ASCII: src/io/io.cu:164-196;
//    big-endian binary: opt_src_v3/io/io.cu:121-200): an unstructured grid of
//    points with point data "p" (scalar) and "velocity" (vector).  With no
//    optional fields the files are byte-identical to the reference's; the
//    optional density and type arrays are written as extra SCALARS sections
//    between them.
#include <algorithm>
#include <cerrno>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "sph_internal.h"

namespace {

int set_error(int code, const std::string& msg) { return sph::set_error(code, msg.c_str()); }

std::string join(const char* dir, const char* name) {
  std::string d(dir ? dir : ".");
  if (!d.empty() && d.back() != '/') d += '/';
  return d + name;
}

// Reads exactly `count` values from a whitespace-separated text file.
template <class T>
int read_values(const std::string& path, T* out, long long count) {
  FILE* f = std::fopen(path.c_str(), "r");
  if (!f) return set_error(SPH_E_INVAL, "cannot open " + path + ": " + std::strerror(errno));
  long long k = 0;
  for (; k < count; ++k) {
    int got;
    if constexpr (sizeof(T) == sizeof(float) && std::is_floating_point<T>::value)
      got = std::fscanf(f, "%f", reinterpret_cast<float*>(&out[k]));
    else
      got = std::fscanf(f, "%d", reinterpret_cast<int*>(&out[k]));
    if (got != 1) break;
  }
  std::fclose(f);
  if (k != count)
    return set_error(SPH_E_INVAL, path + ": expected " + std::to_string(count) + " values, read " + std::to_string(k));
  return SPH_OK;
}

int count_values(const std::string& path, long long* count) {
  FILE* f = std::fopen(path.c_str(), "r");
  if (!f) return set_error(SPH_E_INVAL, "cannot open " + path + ": " + std::strerror(errno));
  long long k = 0;
  char buf[128];
  while (std::fscanf(f, "%127s", buf) == 1) ++k;
  std::fclose(f);
  *count = k;
  return SPH_OK;
}

template <class T>
int write_values(const std::string& path, const T* v, long long count, int per_line) {
  FILE* f = std::fopen(path.c_str(), "w");
  if (!f) return set_error(SPH_E_INVAL, "cannot create " + path + ": " + std::strerror(errno));
  for (long long k = 0; k < count; ++k) {
    if constexpr (std::is_floating_point<T>::value)
      std::fprintf(f, "%.9g", (double)v[k]);
    else
      std::fprintf(f, "%d", (int)v[k]);
    std::fputc((k + 1) % per_line == 0 ? '\n' : ' ', f);
  }
  const bool ok = std::fclose(f) == 0;
  return ok ? SPH_OK : set_error(SPH_E_INVAL, "write failed: " + path);
}

// Writes `count` 32-bit words as big-endian bytes, in chunks.
void put_be32(FILE* f, const void* src, size_t count) {
  const unsigned char* p = (const unsigned char*)src;
  unsigned char buf[1 << 16];
  while (count) {
    const size_t k = std::min(count, sizeof(buf) / 4);
    for (size_t i = 0; i < k; ++i) {
      buf[4 * i + 0] = p[4 * i + 3];
      buf[4 * i + 1] = p[4 * i + 2];
      buf[4 * i + 2] = p[4 * i + 1];
      buf[4 * i + 3] = p[4 * i + 0];
    }
    std::fwrite(buf, 1, 4 * k, f);
    p += 4 * k;
    count -= k;
  }
}

}  // namespace

extern "C" int sph_count_dat(const char* dir, int* n) {
  if (!n) return set_error(SPH_E_INVAL, "null argument");
  long long k = 0;
  int rc = count_values(join(dir, "rhop.dat"), &k);
  if (rc) return rc;
  if (k > 0x7fffffffLL) return set_error(SPH_E_INVAL, "rhop.dat: too many particles");
  *n = (int)k;
  return SPH_OK;
}

extern "C" int sph_read_dat(const char* dir, int n, float* pos, float* vel, float* acc, float* rho, float* p,
                            int* type, int* zone) {
  if (n <= 0 || !pos || !vel || !rho || !p || !type) return set_error(SPH_E_INVAL, "sph_read_dat: bad argument");
  const long long n3 = 3LL * n;
  int rc = read_values(join(dir, "h_pos.dat"), pos, n3);
  if (!rc) rc = read_values(join(dir, "h_vel.dat"), vel, n3);
  if (!rc && acc) rc = read_values(join(dir, "h_ac.dat"), acc, n3);
  if (!rc) rc = read_values(join(dir, "rhop.dat"), rho, n);
  if (!rc) rc = read_values(join(dir, "p.dat"), p, n);
  if (!rc) rc = read_values(join(dir, "particle_type.dat"), type, n);
  if (!rc && zone) rc = read_values(join(dir, "particle_zone.dat"), zone, n);
  return rc;
}

extern "C" int sph_write_dat(const char* dir, int n, const float* pos, const float* vel, const float* acc,
                             const float* rho, const float* p, const int* type, const int* zone) {
  if (n <= 0 || !pos || !vel || !rho || !p || !type) return set_error(SPH_E_INVAL, "sph_write_dat: bad argument");
  const long long n3 = 3LL * n;
  std::vector<float> zeros3;
  std::vector<int> zeros;
  if (!acc) zeros3.assign((size_t)n3, 0.0f);
  if (!zone) zeros.assign((size_t)n, 0);
  int rc = write_values(join(dir, "h_pos.dat"), pos, n3, 3);
  if (!rc) rc = write_values(join(dir, "h_vel.dat"), vel, n3, 3);
  if (!rc) rc = write_values(join(dir, "h_ac.dat"), acc ? acc : zeros3.data(), n3, 3);
  if (!rc) rc = write_values(join(dir, "rhop.dat"), rho, n, 1);
  if (!rc) rc = write_values(join(dir, "p.dat"), p, n, 1);
  if (!rc) rc = write_values(join(dir, "particle_type.dat"), type, n, 1);
  if (!rc) rc = write_values(join(dir, "particle_zone.dat"), zone ? zone : zeros.data(), n, 1);
  return rc;
}

extern "C" int sph_write_vtk(const char* path, int n, const float* pos, const float* vel, const float* p,
                             const float* rho, const int* type, int binary) {
  if (!path || n < 0 || (n > 0 && (!pos || !vel || !p))) return set_error(SPH_E_INVAL, "sph_write_vtk: bad argument");
  FILE* f = std::fopen(path, binary ? "wb" : "w");
  if (!f) return set_error(SPH_E_INVAL, std::string("cannot create ") + path + ": " + std::strerror(errno));
  const char* ty = binary ? "float" : "double";
  std::fprintf(f, "# vtk DataFile Version 3.0\nsph data\n%s\nDATASET UNSTRUCTURED_GRID\n", binary ? "BINARY" : "ASCII");
  std::fprintf(f, "POINTS %d %s\n", n, ty);
  // ASCII values as the reference's stream prints them (scientific,
  // precision 6); binary values as big-endian 32-bit words with no separator
  // before the next keyword line, as the reference writes them.
  auto vec3 = [&](const float* v) {
    if (binary)
      put_be32(f, v, 3 * (size_t)n);
    else
      for (int i = 0; i < n; ++i)
        std::fprintf(f, "%.6e %.6e %.6e\n", (double)v[3 * i], (double)v[3 * i + 1], (double)v[3 * i + 2]);
  };
  auto scal = [&](const float* v) {
    if (binary)
      put_be32(f, v, (size_t)n);
    else
      for (int i = 0; i < n; ++i) std::fprintf(f, "%.6e\n", (double)v[i]);
  };
  vec3(pos);
  std::fprintf(f, "POINT_DATA %d\n", n);
  std::fprintf(f, "SCALARS p %s 1\nLOOKUP_TABLE DEFAULT\n", ty);
  scal(p);
  if (rho) {
    std::fprintf(f, "SCALARS rho %s 1\nLOOKUP_TABLE DEFAULT\n", ty);
    scal(rho);
  }
  if (type) {
    std::fprintf(f, "SCALARS type int 1\nLOOKUP_TABLE DEFAULT\n");
    if (binary)
      put_be32(f, type, (size_t)n);
    else
      for (int i = 0; i < n; ++i) std::fprintf(f, "%d\n", type[i]);
  }
  std::fprintf(f, "VECTORS velocity %s\n", ty);
  vec3(vel);
  const bool ok = std::fclose(f) == 0;
  return ok ? SPH_OK : set_error(SPH_E_INVAL, std::string("write failed: ") + path);
}
