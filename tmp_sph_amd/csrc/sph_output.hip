// Checkpoint / restart of a simulation context and the asynchronous frame
// writer.  The writer follows the reference's output thread
// (opt_src_v3/io/io.cu:98-119, handshake at opt_src_v3/sph/sph_run.cu:268-300),
// whose busy-waiting compare-exchange loop is replaced by a queue, a mutex and
// a condition variable, and whose single host buffer becomes two pinned ones
// filled by a copy stream, so a frame's transfer and its file write both
// overlap the steps that follow.
#include <hip/hip_runtime.h>

#include <condition_variable>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "sph_internal.h"
#include "sph_sim_ctx.h"

using namespace sph;

namespace {

#define SPH_HIP(call, where)                            \
  do {                                                  \
    hipError_t e_ = (call);                             \
    if (e_ != hipSuccess) return hip_status(e_, where); \
  } while (0)

// ------------------------------------------------------------------ checkpoint

constexpr char kMagic[8] = {'S', 'P', 'H', 'C', 'K', 'P', 'T', '2'};

struct CkptHeader {
  char magic[8];
  uint32_t params_bytes;  // sizeof(sph_params) of the writer
  int32_t K, n, dim;
  int64_t steps;
  int32_t cell_div;       // cell grid the sorted order follows (tuning.cell_div)
  int32_t reserved;
};

// FNV-1a, 64-bit, over every byte after the header's magic.
struct Fnv {
  uint64_t h = 1469598103934665603ull;
  void add(const void* p, size_t len) {
    const unsigned char* b = (const unsigned char*)p;
    for (size_t i = 0; i < len; ++i) h = (h ^ b[i]) * 1099511628211ull;
  }
};

// The particle arrays of the current set, in sorted-slot order (2-D
// contexts keep no y coordinates).
struct Arrays {
  std::vector<float> PR, VP;  // 4 per particle
  std::vector<double> X[3];
  std::vector<int> ID, TY;
  void resize(size_t n, int dim) {
    PR.resize(4 * n);
    VP.resize(4 * n);
    for (int c = 0; c < 3; ++c) X[c].resize(c == 1 && dim == 2 ? 0 : n);
    ID.resize(n);
    TY.resize(n);
  }
  template <class F>
  void each(F&& f) {
    f(PR.data(), PR.size() * sizeof(float));
    f(VP.data(), VP.size() * sizeof(float));
    for (auto& x : X) f(x.data(), x.size() * sizeof(double));
    f(ID.data(), ID.size() * sizeof(int));
    f(TY.data(), TY.size() * sizeof(int));
  }
};

int copy_set(const FastSet& st, Arrays& a, size_t n, hipMemcpyKind kind) {
  const bool down = kind == hipMemcpyDeviceToHost;
  auto cp = [&](void* dev, void* host, size_t bytes) {
    return down ? hipMemcpy(host, dev, bytes, kind) : hipMemcpy(dev, host, bytes, kind);
  };
  hipError_t e = cp(st.PR, a.PR.data(), 16 * n);
  if (e == hipSuccess) e = cp(st.VP, a.VP.data(), 16 * n);
  for (int c = 0; c < 3 && e == hipSuccess; ++c)
    if (!a.X[c].empty()) e = cp(st.X[c], a.X[c].data(), 8 * n);
  if (e == hipSuccess) e = cp(st.ID, a.ID.data(), 4 * n);
  if (e == hipSuccess) e = cp(st.TY, a.TY.data(), 4 * n);
  return e == hipSuccess ? SPH_OK : hip_status(e, "checkpoint copy");
}

// ------------------------------------------------------------------ async output

struct Slot {
  float *dpos = nullptr, *dvel = nullptr, *dp = nullptr, *drho = nullptr;
  int* dtype = nullptr;
  float *hpos = nullptr, *hvel = nullptr, *hp = nullptr, *hrho = nullptr;
  int* htype = nullptr;
  hipEvent_t staged = nullptr, copied = nullptr;
  int frame = 0;
  bool busy = false;
};

}  // namespace

namespace sph {

struct OutputPipe {
  std::string dir;
  int binary = 0, fields = 0, n = 0;
  hipStream_t copy = nullptr;
  Slot slot[2];
  int next = 0;
  std::thread writer;
  std::mutex m;
  std::condition_variable cv;
  std::deque<int> queue;
  bool stop = false;
  int err = SPH_OK;
  std::string errmsg;
  long long written = 0;

  void run() {
    for (;;) {
      int k;
      {
        std::unique_lock<std::mutex> lk(m);
        cv.wait(lk, [&] { return stop || !queue.empty(); });
        if (queue.empty()) return;  // stop requested and nothing left
        k = queue.front();
        queue.pop_front();
      }
      Slot& sl = slot[k];
      int rc = SPH_OK;
      std::string msg;
      hipError_t e = hipEventSynchronize(sl.copied);
      if (e != hipSuccess) {
        rc = hip_status(e, "output: copy");
        msg = hipGetErrorString(e);
      } else {
        const std::string path = dir + "/fluid_t=" + std::to_string(sl.frame) + ".vtk";
        rc = sph_write_vtk(path.c_str(), n, sl.hpos, sl.hvel, sl.hp, (fields & 1) ? sl.hrho : nullptr,
                           (fields & 2) ? sl.htype : nullptr, binary);
        if (rc) msg = sph_last_error();
      }
      std::lock_guard<std::mutex> lk(m);
      if (rc == SPH_OK)
        ++written;
      else if (err == SPH_OK) {
        err = rc;
        errmsg = msg;
      }
      sl.busy = false;
      cv.notify_all();
    }
  }
};

void output_release(sph_sim* s) {
  OutputPipe* o = s->out;
  if (!o) return;
  if (o->writer.joinable()) {
    {
      std::lock_guard<std::mutex> lk(o->m);
      o->stop = true;
    }
    o->cv.notify_all();
    o->writer.join();
  }
  // teardown: release errors are not actionable
  for (Slot& sl : o->slot) {
    for (void* d : {(void*)sl.dpos, (void*)sl.dvel, (void*)sl.dp, (void*)sl.drho, (void*)sl.dtype}) (void)hipFree(d);
    for (void* h : {(void*)sl.hpos, (void*)sl.hvel, (void*)sl.hp, (void*)sl.hrho, (void*)sl.htype})
      (void)hipHostFree(h);
    if (sl.staged) (void)hipEventDestroy(sl.staged);
    if (sl.copied) (void)hipEventDestroy(sl.copied);
  }
  if (o->copy) (void)hipStreamDestroy(o->copy);
  delete o;
  s->out = nullptr;
}

}  // namespace sph

extern "C" int sph_sim_checkpoint(sph_sim* s, const char* path) {
  if (!s || !path) return set_error(SPH_E_INVAL, "sph_sim_checkpoint: null argument");
  const size_t n = (size_t)s->n;
  Arrays a;
  a.resize(n, s->dim);
  SPH_HIP(hipStreamSynchronize(s->stream), "sph_sim_checkpoint");
  int rc = copy_set(s->set[0], a, n, hipMemcpyDeviceToHost);
  if (rc) return rc;
  CkptHeader h{};
  std::memcpy(h.magic, kMagic, 8);
  h.params_bytes = sizeof(sph_params);
  h.K = s->K;
  h.n = s->n;
  h.dim = s->dim;
  h.steps = s->steps;
  h.cell_div = s->tun.cell_div;
  Fnv sum;
  sum.add((const char*)&h + 8, sizeof(h) - 8);
  sum.add(s->P.data(), sizeof(sph_params) * s->K);
  a.each([&](const void* p, size_t len) { sum.add(p, len); });
  FILE* f = std::fopen(path, "wb");
  if (!f) return set_error(SPH_E_INVAL, (std::string("sph_sim_checkpoint: cannot create ") + path).c_str());
  bool ok = std::fwrite(&h, sizeof(h), 1, f) == 1;
  ok = ok && std::fwrite(s->P.data(), sizeof(sph_params), s->K, f) == (size_t)s->K;
  a.each([&](const void* p, size_t len) { ok = ok && std::fwrite(p, 1, len, f) == len; });
  ok = ok && std::fwrite(&sum.h, sizeof(sum.h), 1, f) == 1;
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) return set_error(SPH_E_INVAL, (std::string("sph_sim_checkpoint: write failed: ") + path).c_str());
  return SPH_OK;
}

extern "C" int sph_sim_restore(const char* path, sph_sim** out) {
  if (!path || !out) return set_error(SPH_E_INVAL, "sph_sim_restore: null argument");
  *out = nullptr;
  FILE* f = std::fopen(path, "rb");
  if (!f) return set_error(SPH_E_INVAL, (std::string("sph_sim_restore: cannot open ") + path).c_str());
  auto fail = [&](const char* why) {
    std::fclose(f);
    return set_error(SPH_E_INVAL, (std::string("sph_sim_restore: ") + path + ": " + why).c_str());
  };
  CkptHeader h{};
  if (std::fread(&h, sizeof(h), 1, f) != 1 || std::memcmp(h.magic, kMagic, 8) != 0)
    return fail("not a checkpoint");
  if (h.params_bytes != sizeof(sph_params)) return fail("written by a build with another parameter layout");
  if (h.K <= 0 || h.n <= 0 || h.K > h.n || (h.dim != 2 && h.dim != 3) || (h.cell_div != 1 && h.cell_div != 2))
    return fail("corrupt header");
  std::vector<sph_params> P(h.K);
  if (std::fread(P.data(), sizeof(sph_params), h.K, f) != (size_t)h.K) return fail("truncated");
  long long total = 0;
  std::vector<int> counts(h.K);
  for (int k = 0; k < h.K; ++k) {
    counts[k] = P[k].numParticles;
    total += counts[k];
  }
  if (total != h.n) return fail("instance sizes do not add up");
  const size_t n = (size_t)h.n;
  Arrays a;
  a.resize(n, h.dim);
  bool ok = true;
  a.each([&](void* p, size_t len) { ok = ok && std::fread(p, 1, len, f) == len; });
  uint64_t want = 0;
  ok = ok && std::fread(&want, sizeof(want), 1, f) == 1;
  if (!ok) return fail("truncated");
  if (std::fgetc(f) != EOF) return fail("trailing bytes");
  std::fclose(f);
  Fnv sum;
  sum.add((const char*)&h + 8, sizeof(h) - 8);
  sum.add(P.data(), sizeof(sph_params) * h.K);
  a.each([&](const void* p, size_t len) { sum.add(p, len); });
  if (sum.h != want) return set_error(SPH_E_INVAL, (std::string("sph_sim_restore: ") + path + ": checksum mismatch").c_str());
  for (size_t i = 0; i < n; ++i)
    if (a.ID[i] < 0 || a.ID[i] >= h.n || a.TY[i] < 0 || (a.TY[i] >> 8) >= h.K)
      return set_error(SPH_E_INVAL, "sph_sim_restore: corrupt particle arrays");
  // Build a context of the right shape (its upload is overwritten below),
  // then install the saved sorted state and step count.
  std::vector<float> pos(3 * n), vel(3 * n), rho(n), prs(n);
  std::vector<int> type(n);
  for (size_t i = 0; i < n; ++i) {
    for (int c = 0; c < 3; ++c) pos[3 * i + c] = a.X[c].empty() ? 0.0f : (float)a.X[c][i];
    type[i] = a.TY[i] & 255;
    rho[i] = 1.0f;
  }
  sph_sim* s = nullptr;
  int rc = sph_sim_create_batch(P.data(), h.K, counts.data(), pos.data(), vel.data(), rho.data(), prs.data(),
                                type.data(), &s);
  if (rc) return rc;
  if (s->dim != h.dim) {
    sph_sim_destroy(s);
    return set_error(SPH_E_INVAL, "sph_sim_restore: dimension mismatch");
  }
  rc = copy_set(s->set[0], a, n, hipMemcpyHostToDevice);
  if (rc == SPH_OK && h.cell_div != 1) {  // the grid changes the order of the sums: keep the saved one
    sph_tuning t = s->tun;
    t.cell_div = h.cell_div;
    t.nbr_lds = 4;
    rc = sph_sim_set_tuning(s, &t);
  }
  if (rc) {
    sph_sim_destroy(s);
    return rc;
  }
  s->steps = h.steps;
  *out = s;
  return SPH_OK;
}

extern "C" int sph_sim_output_open(sph_sim* s, const char* dir, int binary, int fields) {
  if (!s || !dir || (binary != 0 && binary != 1) || fields < 0 || fields > 3)
    return set_error(SPH_E_INVAL, "sph_sim_output_open: bad argument");
  if (s->out) return set_error(SPH_E_STATE, "sph_sim_output_open: already open");
  OutputPipe* o = new OutputPipe();
  s->out = o;
  o->dir = dir;
  o->binary = binary;
  o->fields = fields;
  o->n = s->n;
  const size_t n = (size_t)s->n;
  hipError_t e = hipStreamCreateWithFlags(&o->copy, hipStreamNonBlocking);
  for (Slot& sl : o->slot) {
    if (e == hipSuccess) e = hipMalloc(&sl.dpos, 12 * n);
    if (e == hipSuccess) e = hipMalloc(&sl.dvel, 12 * n);
    if (e == hipSuccess) e = hipMalloc(&sl.dp, 4 * n);
    if (e == hipSuccess && (fields & 1)) e = hipMalloc(&sl.drho, 4 * n);
    if (e == hipSuccess && (fields & 2)) e = hipMalloc(&sl.dtype, 4 * n);
    if (e == hipSuccess) e = hipHostMalloc((void**)&sl.hpos, 12 * n, hipHostMallocDefault);
    if (e == hipSuccess) e = hipHostMalloc((void**)&sl.hvel, 12 * n, hipHostMallocDefault);
    if (e == hipSuccess) e = hipHostMalloc((void**)&sl.hp, 4 * n, hipHostMallocDefault);
    if (e == hipSuccess && (fields & 1)) e = hipHostMalloc((void**)&sl.hrho, 4 * n, hipHostMallocDefault);
    if (e == hipSuccess && (fields & 2)) e = hipHostMalloc((void**)&sl.htype, 4 * n, hipHostMallocDefault);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&sl.staged, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&sl.copied, hipEventDisableTiming);
  }
  if (e != hipSuccess) {
    const int rc = hip_status(e, "sph_sim_output_open");
    output_release(s);
    return rc;
  }
  o->writer = std::thread([o] { o->run(); });
  return SPH_OK;
}

extern "C" int sph_sim_output_frame(sph_sim* s, int frame) {
  if (!s) return set_error(SPH_E_INVAL, "null sim");
  OutputPipe* o = s->out;
  if (!o) return set_error(SPH_E_STATE, "sph_sim_output_frame: output not open");
  Slot& sl = o->slot[o->next];
  {
    std::unique_lock<std::mutex> lk(o->m);
    o->cv.wait(lk, [&] { return !sl.busy; });
    if (o->err) return set_error(o->err, o->errmsg.c_str());
  }
  // the gather runs on the step stream: the next step's writes into the
  // particle sets are ordered after it
  SPH_HIP(launch_download_fast(s->set[0], sl.dpos, sl.dvel, sl.drho, sl.dp, sl.dtype, nullptr, s->n, s->stream),
          "output: gather");
  SPH_HIP(hipEventRecord(sl.staged, s->stream), "output: record");
  SPH_HIP(hipStreamWaitEvent(o->copy, sl.staged, 0), "output: wait");
  const size_t n = (size_t)s->n;
  SPH_HIP(hipMemcpyAsync(sl.hpos, sl.dpos, 12 * n, hipMemcpyDeviceToHost, o->copy), "output: copy");
  SPH_HIP(hipMemcpyAsync(sl.hvel, sl.dvel, 12 * n, hipMemcpyDeviceToHost, o->copy), "output: copy");
  SPH_HIP(hipMemcpyAsync(sl.hp, sl.dp, 4 * n, hipMemcpyDeviceToHost, o->copy), "output: copy");
  if (sl.drho) SPH_HIP(hipMemcpyAsync(sl.hrho, sl.drho, 4 * n, hipMemcpyDeviceToHost, o->copy), "output: copy");
  if (sl.dtype) SPH_HIP(hipMemcpyAsync(sl.htype, sl.dtype, 4 * n, hipMemcpyDeviceToHost, o->copy), "output: copy");
  SPH_HIP(hipEventRecord(sl.copied, o->copy), "output: record");
  {
    std::lock_guard<std::mutex> lk(o->m);
    sl.frame = frame;
    sl.busy = true;
    o->queue.push_back(o->next);
  }
  o->cv.notify_all();
  o->next ^= 1;
  return SPH_OK;
}

extern "C" int sph_sim_output_flush(sph_sim* s, long long* frames_written) {
  if (!s) return set_error(SPH_E_INVAL, "null sim");
  OutputPipe* o = s->out;
  if (!o) return set_error(SPH_E_STATE, "sph_sim_output_flush: output not open");
  std::unique_lock<std::mutex> lk(o->m);
  o->cv.wait(lk, [&] { return o->queue.empty() && !o->slot[0].busy && !o->slot[1].busy; });
  if (frames_written) *frames_written = o->written;
  if (o->err) return set_error(o->err, o->errmsg.c_str());
  return SPH_OK;
}

extern "C" int sph_sim_output_close(sph_sim* s) {
  if (!s) return set_error(SPH_E_INVAL, "null sim");
  if (!s->out) return SPH_OK;
  const int rc = sph_sim_output_flush(s, nullptr);
  output_release(s);
  return rc;
}
