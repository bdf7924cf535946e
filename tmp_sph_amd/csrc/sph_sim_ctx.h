// The simulation context shared by the translation units that implement the
// sph_sim_* entry points (stepping, checkpoints, the asynchronous writer).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "sph_kernels.h"

namespace sph {
struct OutputPipe;
}

struct sph_sim {
  std::vector<sph_params> P;  // per instance (host copy)
  std::vector<int> cell_off;  // per instance
  sph::InstTable T{};         // grid in use (T1 or T2, per tuning.cell_div)
  sph::InstTable T1{};        // the instances' own grids
  sph::InstTable T2{};        // the same grids subdivided by 2 (K = 0 when too large)
  int K = 1;
  int dim = 3;
  int n = 0;
  int bits = 1;               // key bits of the grid in use
  int bits1 = 1, bits2 = 1;
  hipStream_t stream = nullptr;
  sph::FastSet set[2]{};
  sph::FastWork w{};
  sph_tuning tun{};
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  long long steps = 0;
  unsigned long long* probe = nullptr;  // 2K words (also 4 for pair statistics)
  unsigned* gbuf = nullptr;             // gauges / point probes: 5 x SPH_MAX_PROBES words
  std::vector<void*> allocs;
  sph::OutputPipe* out = nullptr;  // asynchronous frame writer (sph_sim_output_open)
  // kernel timing (eager steps with tuning.flags bit 0)
  std::vector<std::string> tnames;
  std::vector<double> tsum_ms;
  std::vector<long long> tcount;
  struct Mark {
    int slot;
    hipEvent_t a, b;
  };
  std::vector<Mark> marks;
  std::vector<hipEvent_t> event_pool;
  size_t event_next = 0;
};

namespace sph {
// Drains and frees the asynchronous writer of `s`, if one is open.
void output_release(sph_sim* s);
}
