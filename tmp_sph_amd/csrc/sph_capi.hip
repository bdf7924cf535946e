// C ABI of the MI355X delta-SPH step: runtime helpers, the reference call
// surface and the simulation context (fast path).
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "sph_kernels.h"

namespace sph {

namespace {
thread_local std::string g_last_error;
std::mutex g_mu;
sph_params g_params;
bool g_params_set = false;
hipStream_t g_stream = nullptr;

// Scratch of the reference-surface sortParticles, grown on demand like the
// temporary storage thrust::sort_by_key allocates.
struct SortPool {
  int cap = 0;
  int* keys_alt = nullptr;
  int* vals_alt = nullptr;
  unsigned* hist = nullptr;
  unsigned* totals = nullptr;
} g_pool;

// SPH_NS_simpleversion: 0 runs the fast path's row-staged loops on the
// caller's arrays (packed into NsPool), 1 the reference-order kernels.
int g_ns_mode = 0;

// Scratch of SPH_NS_simpleversion on the row-staged loops, grown on demand:
// the sorted state in the fast layout, the per-cell fluid flags, dofv and the
// rates, and a device copy of the parameters.
struct NsPool {
  int cap = 0;
  int cells = 0;
  FastSet st{};
  float* DV = nullptr;
  float4* RATE = nullptr;
  int* cellFluid = nullptr;
  sph_params* dP = nullptr;
  int* dcoff = nullptr;
  sph_params uploaded{};
  bool have_params = false;
} g_ns;
}  // namespace

int set_error(int code, const char* msg) {
  g_last_error = msg ? msg : "";
  return code;
}

int hip_status(int hip_err, const char* where) {
  if (hip_err == hipSuccess) return SPH_OK;
  (void)hipGetLastError();  // clear it, so a later launch check does not report it again
  std::string m = std::string(where) + ": " + hipGetErrorString((hipError_t)hip_err);
  g_last_error = m;
  return hip_err > 0 ? hip_err : 1000 - hip_err;
}

}  // namespace sph

using namespace sph;

#define SPH_HIP(call, where)                         \
  do {                                               \
    hipError_t e_ = (call);                          \
    if (e_ != hipSuccess) return hip_status(e_, where); \
  } while (0)

namespace {

int current_params(sph_params* out, hipStream_t* s) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_params_set) return set_error(SPH_E_NOPARAMS, "sph_set_params was not called");
  *out = g_params;
  *s = g_stream;
  return SPH_OK;
}

}  // namespace

namespace sph {

int check_params(const sph_params* p) {
  if (!p) return set_error(SPH_E_INVAL, "null params");
  if (p->dim != 2 && p->dim != 3) return set_error(SPH_E_INVAL, "params.dim must be 2 or 3");
  if (p->gridxdim <= 0 || p->gridydim <= 0 || p->gridzdim <= 0)
    return set_error(SPH_E_INVAL, "params: grid dimensions must be positive");
  if ((long long)p->gridxdim * p->gridydim * p->gridzdim != (long long)p->numCells)
    return set_error(SPH_E_INVAL, "params: numCells != gridxdim*gridydim*gridzdim");
  if (p->dim == 2 && p->gridydim != 1) return set_error(SPH_E_INVAL, "params: 2-D needs gridydim == 1");
  if (!(p->h > 0.0f) || !(p->gridSize > 0.0f) || !(p->restDensity > 0.0f))
    return set_error(SPH_E_INVAL, "params: h, gridSize and restDensity must be positive");
  if (!(p->gridSize >= p->kh * 0.999999f))
    return set_error(SPH_E_INVAL, "params: gridSize must be at least 2h for the 3-cell stencil");
  if (p->eos != SPH_EOS_LINEAR && p->eos != SPH_EOS_TAIT) return set_error(SPH_E_INVAL, "params: unknown eos");
  if (p->eos == SPH_EOS_TAIT && !(p->tait_B > 0.0f && p->gamma > 0.0f))
    return set_error(SPH_E_INVAL, "params: Tait EOS needs tait_B > 0 and gamma > 0");
  return SPH_OK;
}

}  // namespace sph

namespace {

int ensure_pool(int n) {
  if (g_pool.cap >= n && g_pool.keys_alt) return SPH_OK;
  if (g_pool.keys_alt) {
    hipFree(g_pool.keys_alt);
    hipFree(g_pool.vals_alt);
    hipFree(g_pool.hist);
    hipFree(g_pool.totals);
    g_pool = SortPool();
  }
  const size_t nn = (size_t)(n > 1 ? n : 1);
  SPH_HIP(hipMalloc(&g_pool.keys_alt, nn * sizeof(int)), "sort scratch");
  SPH_HIP(hipMalloc(&g_pool.vals_alt, nn * sizeof(int)), "sort scratch");
  SPH_HIP(hipMalloc(&g_pool.hist, (size_t)radix_hist_words((int)nn) * sizeof(unsigned)), "sort scratch");
  SPH_HIP(hipMalloc(&g_pool.totals, 256 * sizeof(unsigned)), "sort scratch");
  g_pool.cap = (int)nn;
  return SPH_OK;
}

// Grows the NS scratch to n particles and `cells` cells.
int ensure_ns_pool(int n, int cells) {
  NsPool& q = g_ns;
  if (!q.dP) {
    SPH_HIP(hipMalloc(&q.dP, sizeof(sph_params)), "SPH_NS scratch");
    SPH_HIP(hipMalloc(&q.dcoff, sizeof(int)), "SPH_NS scratch");
    SPH_HIP(hipMemset(q.dcoff, 0, sizeof(int)), "SPH_NS scratch");
  }
  if (q.cap < n || !q.st.PR) {
    hipFree(q.st.PR);
    hipFree(q.st.VP);
    hipFree(q.st.TY);
    hipFree(q.DV);
    hipFree(q.RATE);
    q.st = FastSet{};
    q.DV = nullptr;
    q.RATE = nullptr;
    q.cap = 0;
    const size_t nn = (size_t)(n > 1 ? n : 1);
    SPH_HIP(hipMalloc(&q.st.PR, (nn + CAND_PAD) * sizeof(float4)), "SPH_NS scratch");
    SPH_HIP(hipMalloc(&q.st.VP, (nn + CAND_PAD) * sizeof(float4)), "SPH_NS scratch");
    SPH_HIP(hipMalloc(&q.st.TY, nn * sizeof(int)), "SPH_NS scratch");
    SPH_HIP(hipMalloc(&q.DV, (nn + CAND_PAD) * sizeof(float)), "SPH_NS scratch");
    SPH_HIP(hipMalloc(&q.RATE, nn * sizeof(float4)), "SPH_NS scratch");
    q.cap = (int)nn;
  }
  if (q.cells < cells || !q.cellFluid) {
    hipFree(q.cellFluid);
    q.cellFluid = nullptr;
    q.cells = 0;
    SPH_HIP(hipMalloc(&q.cellFluid, (size_t)(cells > 1 ? cells : 1) * sizeof(int)), "SPH_NS scratch");
    q.cells = cells;
  }
  return SPH_OK;
}

// Kernels A and B of the row-staged fast path on the caller's sorted AoS
// state: pack, A (dofv, walls' pressure and density), B (rates of every
// particle, as the reference computes them), unpack.
int ns_fast(const sph_params& P, hipStream_t s, const float* sPos, const float* sVel, float* sRho, float* sP,
            const int* sType, float* drho, float* acc, const int* cellStart, const int* cellEnd, int n,
            const int* hash, float* dofv) {
  int rc = ensure_ns_pool(n, P.numCells);
  if (rc) return rc;
  NsPool& q = g_ns;
  if (!q.have_params || std::memcmp(&q.uploaded, &P, sizeof(P)) != 0) {
    SPH_HIP(hipMemcpyAsync(q.dP, &P, sizeof(P), hipMemcpyHostToDevice, s), "SPH_NS params");
    SPH_HIP(hipStreamSynchronize(s), "SPH_NS params");  // the host copy is a local
    q.uploaded = P;
    q.have_params = true;
  }
  InstTable T{q.dP, q.dcoff, 1, P.dim, P.numCells};
  FastWork w{};
  w.keys = const_cast<int*>(hash);  // read only
  w.cellStart = const_cast<int*>(cellStart);
  w.cellEnd = const_cast<int*>(cellEnd);
  w.cellFluid = q.cellFluid;
  w.DV = q.DV;
  w.RATE = q.RATE;
  sph_tuning tu{};
  tu.nbr_block = 64;
  tu.nbr_unroll = 4;  // the contexts' default shapes
  tu.xcd_remap = P.dim == 3 ? 32 : 1;
  tu.nbr_lds = 3;
  tu.nbr_prefetch = 1;
  tu.cell_div = 1;
  SPH_HIP(launch_ref_pack(sPos, sVel, sRho, sP, sType, hash, q.st, q.cellFluid, P.numCells, n, s), "SPH_NS pack");
  SPH_HIP(launch_nsA_fast(T, q.st, w, n, tu, s), "SPH_NS kernel A");
  SPH_HIP(launch_nsB_fast(T, q.st, w, n, tu, s, true), "SPH_NS kernel B");
  SPH_HIP(launch_ref_unpack(q.st, w, sType, drho, acc, dofv, sRho, sP, n, s), "SPH_NS unpack");
  return SPH_OK;
}

}  // namespace

// ---------------------------------------------------------------- runtime
extern "C" int sph_abi_sizes(int* params_size, int* cfg_size, int* tuning_size) {
  if (params_size) *params_size = (int)sizeof(sph_params);
  if (cfg_size) *cfg_size = (int)sizeof(sph_dambreak_cfg);
  if (tuning_size) *tuning_size = (int)sizeof(sph_tuning);
  return SPH_OK;
}

extern "C" const char* sph_last_error(void) { return g_last_error.c_str(); }

extern "C" int sph_device_count(int* count) {
  if (!count) return set_error(SPH_E_INVAL, "null count");
  int c = 0;
  hipError_t e = hipGetDeviceCount(&c);
  if (e != hipSuccess) c = 0;
  *count = c;
  return SPH_OK;
}

extern "C" int sph_set_device(int device) {
  SPH_HIP(hipSetDevice(device), "hipSetDevice");
  return SPH_OK;
}

extern "C" int sph_set_stream(void* s) {
  std::lock_guard<std::mutex> lk(g_mu);
  g_stream = (hipStream_t)s;
  return SPH_OK;
}

extern "C" int sph_device_synchronize(void) {
  SPH_HIP(hipDeviceSynchronize(), "hipDeviceSynchronize");
  return SPH_OK;
}

extern "C" int sph_malloc(void** p, unsigned long long bytes) {
  if (!p) return set_error(SPH_E_INVAL, "null pointer");
  *p = nullptr;
  SPH_HIP(hipMalloc(p, bytes ? bytes : 1), "hipMalloc");
  return SPH_OK;
}

extern "C" int sph_free(void* p) {
  if (p) SPH_HIP(hipFree(p), "hipFree");
  return SPH_OK;
}

extern "C" int sph_memcpy_htod(void* dst, const void* src, unsigned long long bytes) {
  if (bytes) SPH_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice), "hipMemcpy H2D");
  return SPH_OK;
}

extern "C" int sph_memcpy_dtoh(void* dst, const void* src, unsigned long long bytes) {
  if (bytes) SPH_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost), "hipMemcpy D2H");
  return SPH_OK;
}

extern "C" int sph_memcpy_dtod(void* dst, const void* src, unsigned long long bytes) {
  if (bytes) SPH_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice), "hipMemcpy D2D");
  return SPH_OK;
}

extern "C" int sph_memset(void* p, int value, unsigned long long bytes) {
  if (bytes) SPH_HIP(hipMemset(p, value, bytes), "hipMemset");
  return SPH_OK;
}

// ------------------------------------------------------------- parameters
extern "C" int sph_set_params(const sph_params* p) {
  int rc = check_params(p);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(g_mu);
  g_params = *p;
  g_params_set = true;
  return SPH_OK;
}

extern "C" int sph_get_params(sph_params* p) {
  if (!p) return set_error(SPH_E_INVAL, "null params");
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_params_set) return set_error(SPH_E_NOPARAMS, "sph_set_params was not called");
  *p = g_params;
  return SPH_OK;
}

// ------------------------------------------------- reference call surface
extern "C" int sph_calcHash(const float* pos, int* hash, int* idx, int n) {
  sph_params P;
  hipStream_t s;
  int rc = current_params(&P, &s);
  if (rc) return rc;
  if (n < 0 || (n > 0 && (!pos || !hash || !idx))) return set_error(SPH_E_INVAL, "sph_calcHash: bad argument");
  SPH_HIP(launch_hash_aos(P, pos, hash, idx, n, s), "calcHash");
  return SPH_OK;
}

extern "C" int sph_sortParticles(int* hash, int* idx, int n) {
  sph_params P;
  hipStream_t s;
  int rc = current_params(&P, &s);
  if (rc) return rc;
  if (n < 0 || (n > 0 && (!hash || !idx))) return set_error(SPH_E_INVAL, "sph_sortParticles: bad argument");
  if (n <= 1) return SPH_OK;
  std::lock_guard<std::mutex> lk(g_mu);
  rc = ensure_pool(n);
  if (rc) return rc;
  bool alt = false;
  SPH_HIP(radix_sort_pairs(hash, idx, g_pool.keys_alt, g_pool.vals_alt, g_pool.hist, g_pool.totals, n,
                           key_bits(P.numCells), s, &alt),
          "sortParticles");
  if (alt) {
    SPH_HIP(hipMemcpyAsync(hash, g_pool.keys_alt, (size_t)n * sizeof(int), hipMemcpyDeviceToDevice, s),
            "sortParticles copy");
    SPH_HIP(hipMemcpyAsync(idx, g_pool.vals_alt, (size_t)n * sizeof(int), hipMemcpyDeviceToDevice, s),
            "sortParticles copy");
  }
  // the pool is reused by the next call: keep the stream ordered behind it
  SPH_HIP(hipStreamSynchronize(s), "sortParticles sync");
  return SPH_OK;
}

// ------------------------------------------- sortParticles without allocation
// Scratch of sph_sortParticles_noalloc: a file-scope device array (one copy
// per GPU, resident for the life of the process) holding the alternate keys,
// the alternate indices, the per-tile digit histogram and the digit totals of
// a sort of up to SPH_NOALLOC_MAX_PARTICLES pairs.  Host code never uses the
// array's name as an address: each device's copy is found once with
// hipGetSymbolAddress and that pointer is what the kernels receive.
namespace {
constexpr long long kNoAllocWords = 3LL * SPH_NOALLOC_MAX_PARTICLES + 256;
static __device__ __align__(256) int scratch[kNoAllocWords];

constexpr int kMaxDevices = 64;
int* g_scratch_ptr[kMaxDevices] = {};
std::mutex g_scratch_mu;

// The calling thread's current device's copy of `scratch`.
int scratch_of_current_device(int** out) {
  int dev = 0;
  SPH_HIP(hipGetDevice(&dev), "sph_sortParticles_noalloc: hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices) return set_error(SPH_E_INVAL, "sph_sortParticles_noalloc: device index too large");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    SPH_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "sph_sortParticles_noalloc: hipGetSymbolAddress");
    if (!p) return set_error(SPH_E_STATE, "sph_sortParticles_noalloc: no device address for the scratch array");
    g_scratch_ptr[dev] = static_cast<int*>(p);
  }
  *out = g_scratch_ptr[dev];
  return SPH_OK;
}
}  // namespace

extern "C" int sph_sortParticles_noalloc(int* hash, int* idx, int n) {
  sph_params P;
  hipStream_t s;
  int rc = current_params(&P, &s);
  if (rc) return rc;
  if (n < 0 || (n > 0 && (!hash || !idx))) return set_error(SPH_E_INVAL, "sph_sortParticles_noalloc: bad argument");
  if (n > SPH_NOALLOC_MAX_PARTICLES)
    return set_error(SPH_E_INVAL, "sph_sortParticles_noalloc: more than SPH_NOALLOC_MAX_PARTICLES particles");
  if (n <= 1) return SPH_OK;
  int* base = nullptr;
  rc = scratch_of_current_device(&base);
  if (rc) return rc;
  // every word of the histogram and the totals is written by the sort before it is read
  int* keys_alt = base;
  int* vals_alt = base + SPH_NOALLOC_MAX_PARTICLES;
  unsigned* hist = reinterpret_cast<unsigned*>(base + 2LL * SPH_NOALLOC_MAX_PARTICLES);
  unsigned* totals = reinterpret_cast<unsigned*>(base + 3LL * SPH_NOALLOC_MAX_PARTICLES);
  static_assert(3LL * SPH_NOALLOC_MAX_PARTICLES + 256 == kNoAllocWords, "scratch layout");
  if ((long long)radix_hist_words(n) > (long long)SPH_NOALLOC_MAX_PARTICLES)
    return set_error(SPH_E_INVAL, "sph_sortParticles_noalloc: histogram exceeds the scratch array");
  bool alt = false;
  SPH_HIP(radix_sort_pairs(hash, idx, keys_alt, vals_alt, hist, totals, n, key_bits(P.numCells), s, &alt),
          "sortParticles_noalloc");
  if (alt) {
    SPH_HIP(hipMemcpyAsync(hash, keys_alt, (size_t)n * sizeof(int), hipMemcpyDeviceToDevice, s),
            "sortParticles_noalloc copy");
    SPH_HIP(hipMemcpyAsync(idx, vals_alt, (size_t)n * sizeof(int), hipMemcpyDeviceToDevice, s),
            "sortParticles_noalloc copy");
  }
  return SPH_OK;
}

extern "C" int sph_reorder(const float* oPos, const float* oVel, const float* oRho, const float* oP,
                           const int* oType, const int* oZone, float* sPos, float* sVel, float* sRho,
                           float* sP, int* sType, int* sZone, const int* hash, const int* idx,
                           int* cellStart, int* cellEnd, int n, int* table_l, const int* table_r) {
  sph_params P;
  hipStream_t s;
  int rc = current_params(&P, &s);
  if (rc) return rc;
  if (n < 0 || !cellStart || !cellEnd ||
      (n > 0 && (!oPos || !oVel || !oRho || !oP || !oType || !sPos || !sVel || !sRho || !sP || !sType ||
                 !hash || !idx)))
    return set_error(SPH_E_INVAL, "sph_reorder: bad argument");
  SPH_HIP(launch_reorder_aos(P, oPos, oVel, oRho, oP, oType, oZone, sPos, sVel, sRho, sP, sType, sZone,
                             hash, idx, cellStart, cellEnd, n, table_l, table_r, s),
          "reorder");
  return SPH_OK;
}

extern "C" int sph_recover(float* pos, float* vel, float* rho, float* p, int n, const int* idx, int* type,
                           int* zone, const float* sPos, const float* sVel, const float* sRho, const float* sP,
                           const int* sType, const int* sZone) {
  if (n < 0 || (n > 0 && (!pos || !vel || !rho || !p || !idx || !type || !sPos || !sVel || !sRho || !sP ||
                          !sType)) ||
      ((zone == nullptr) != (sZone == nullptr)))
    return set_error(SPH_E_INVAL, "sph_recover: bad argument");
  hipStream_t s;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    s = g_stream;
  }
  SPH_HIP(launch_recover_aos(pos, vel, rho, p, type, zone, idx, sPos, sVel, sRho, sP, sType, sZone, n, s),
          "recover");
  return SPH_OK;
}

extern "C" int sph_SPH_NS_simpleversion(const float* sPos, const float* sVel, float* sRho, float* sP,
                                        const int* sType, float* densitydt, float* Veldt,
                                        const int* cellStart, const int* cellEnd, int n,
                                        const int* hash, int timestep, float* dofv, float* rhop_sum,
                                        float* w_sum) {
  (void)timestep;
  (void)rhop_sum;
  (void)w_sum;
  sph_params P;
  hipStream_t s;
  int rc = current_params(&P, &s);
  if (rc) return rc;
  if (n < 0 || (n > 0 && (!sPos || !sVel || !sRho || !sP || !sType || !densitydt || !Veldt || !cellStart ||
                          !cellEnd || !hash || !dofv)))
    return set_error(SPH_E_INVAL, "sph_SPH_NS_simpleversion: bad argument");
  int mode;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    mode = g_ns_mode;
  }
  if (mode == 0) {
    std::lock_guard<std::mutex> lk(g_mu);  // one user of the shared scratch at a time
    return ns_fast(P, s, sPos, sVel, sRho, sP, sType, densitydt, Veldt, cellStart, cellEnd, n, hash, dofv);
  }
  SPH_HIP(launch_ns_aos(P, sPos, sVel, sRho, sP, sType, densitydt, Veldt, cellStart, cellEnd, hash, dofv,
                        n, s),
          "SPH_NS_simpleversion");
  return SPH_OK;
}

extern "C" int sph_reserve_scratch(int n) {
  if (n < 0) return set_error(SPH_E_INVAL, "sph_reserve_scratch: n must be >= 0");
  sph_params P;
  hipStream_t s;
  int rc = current_params(&P, &s);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(g_mu);
  rc = ensure_pool(n);
  if (!rc) rc = ensure_ns_pool(n, P.numCells);
  return rc;
}

extern "C" int sph_set_ns_mode(int mode) {
  if (mode != 0 && mode != 1) return set_error(SPH_E_INVAL, "sph_set_ns_mode: mode must be 0 or 1");
  std::lock_guard<std::mutex> lk(g_mu);
  g_ns_mode = mode;
  return SPH_OK;
}

extern "C" int sph_get_ns_mode(int* mode) {
  if (!mode) return set_error(SPH_E_INVAL, "sph_get_ns_mode: null argument");
  std::lock_guard<std::mutex> lk(g_mu);
  *mode = g_ns_mode;
  return SPH_OK;
}

static int pc_common(bool pred, float* sPos, float* sVel, float* sRho, float* sP, const float* drho,
                     const float* acc, const int* idx, float* tPos, float* tVel, float* tRho, int n,
                     int* sType) {
  sph_params P;
  hipStream_t s;
  int rc = current_params(&P, &s);
  if (rc) return rc;
  if (n < 0 || (n > 0 && (!sPos || !sVel || !sRho || !sP || !drho || !acc || !tPos || !tVel || !tRho ||
                          !sType || (!pred && !idx))))
    return set_error(SPH_E_INVAL, "sph_PC: bad argument");
  SPH_HIP(launch_pc_aos(P, pred, sPos, sVel, sRho, sP, drho, acc, idx, tPos, tVel, tRho, n, sType, s),
          pred ? "PC_prediction" : "PC_correction");
  return SPH_OK;
}

extern "C" int sph_PC_prediction(float* sPos, float* sVel, float* sRho, float* sP, const float* densitydt,
                                 const float* Veldt, const int* hash, const int* idx, float* Pos_tmp,
                                 float* Vel_tmp, float* density_tmp, int n, int* sType) {
  (void)hash;
  return pc_common(true, sPos, sVel, sRho, sP, densitydt, Veldt, idx, Pos_tmp, Vel_tmp, density_tmp, n,
                   sType);
}

extern "C" int sph_PC_correction(float* sPos, float* sVel, float* sRho, float* sP, const float* densitydt,
                                 const float* Veldt, const int* hash, const int* idx,
                                 const float* Pos_tmp, const float* Vel_tmp, const float* density_tmp,
                                 int n, int* sType) {
  (void)hash;
  return pc_common(false, sPos, sVel, sRho, sP, densitydt, Veldt, idx, const_cast<float*>(Pos_tmp),
                   const_cast<float*>(Vel_tmp), const_cast<float*>(density_tmp), n, sType);
}
