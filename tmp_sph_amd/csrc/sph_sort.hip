// Stable LSD radix sort of (cell hash, particle index) pairs for gfx950.
//
// Replaces thrust::sort_by_key in sortParticles (src/lib/mesh.cu:73-86),
// which for int keys is a stable LSD radix sort.  Each 8-bit digit pass is
// three launches with no inter-block waiting:
//   upsweep   per-tile digit counts, written digit-major (hist[d][tile])
//   scan      one block per digit turns its row into exclusive tile offsets
//   downsweep per-tile stable ranks from wave64 ballots, scatter to
//             digit base + tile offset + rank
// Stability: a wave handles a contiguous run of its tile in 64-element
// rounds; lanes with equal digits are ranked by lane order (ballot match),
// rounds in order, waves in order, tiles in order.
#include "sph_kernels.h"

namespace sph {

namespace {

constexpr int RS_THREADS = 256;
constexpr int RS_WAVES = RS_THREADS / 64;
constexpr int RS_IPT_MAX = 16;  // keys per thread and tile of the largest tile
constexpr int RS_MIN_TILES = 256;  // small sorts use smaller tiles to fill the GPU

static_assert(RS_THREADS == 256, "one thread per digit in the per-digit phases");

// Lanes of the wave whose digit equals this lane's digit (among `active`,
// a prefix of the wave).
//
// The cell keys arrive nearly sorted (the previous step's order), and every
// LSD pass keeps each bucket of the previous digit in key order, so in almost
// every wave the digits do not decrease from lane to lane.  Then equal digits
// form one run of consecutive lanes, found with one shuffle and one ballot.
// Otherwise (a bucket boundary, or unsorted keys) the digit is matched bit by
// bit with one ballot per bit.
__device__ __forceinline__ unsigned long long match_digit(unsigned d, int nbits,
                                                          unsigned long long active) {
  const int lane = __lane_id();
  const bool on = (active >> lane) & 1ull;
  const unsigned dp = __shfl_up(d, 1);
  const bool start = lane == 0 || d != dp;
  if (__ballot(on && lane > 0 && d < dp) == 0ull) {  // wave-uniform: non-decreasing digits
    const unsigned long long starts = __ballot(on && start);
    const unsigned long long below = starts & (lane == 63 ? ~0ull : ((2ull << lane) - 1ull));
    const int first = below ? 63 - __builtin_clzll(below) : 0;  // start of this lane's run
    const unsigned long long above = starts & ~(lane == 63 ? ~0ull : ((2ull << lane) - 1ull));
    const unsigned long long upto = above ? ((1ull << __builtin_ctzll(above)) - 1ull) : ~0ull;
    return active & upto & ~((1ull << first) - 1ull);
  }
  unsigned long long m = active;
  for (int b = 0; b < nbits; ++b) {
    const int bit = (d >> b) & 1;
    const unsigned long long bal = __ballot(bit);
    m &= bit ? bal : ~bal;
  }
  return m;
}

// Per-tile digit counts.  Only counts are needed, so each thread takes RS_IPT
// consecutive keys (16-byte loads when the keys are 16-byte aligned) and adds
// each run of equal digits with one LDS atomic: nearly sorted keys give one or
// two runs per thread.
template <int RS_IPT>
__global__ __launch_bounds__(RS_THREADS) void rs_upsweep(const int* __restrict__ keys, int n,
                                                         int shift, int nbits,
                                                         unsigned* __restrict__ hist,
                                                         int numTiles, int vec) {
  __shared__ unsigned cnt[256];
  const int tid = threadIdx.x;
  cnt[tid] = 0;
  __syncthreads();
  const unsigned mask = (1u << nbits) - 1u;
  constexpr unsigned NONE = 0xffffffffu;  // past the end of the keys
  const long long b0 = (long long)blockIdx.x * (RS_THREADS * RS_IPT) + (long long)tid * RS_IPT;
  unsigned d[RS_IPT];
  if constexpr (RS_IPT % 4 == 0) {
    if (vec && b0 + RS_IPT <= n) {  // b0 is a multiple of 4
      const int4* p = reinterpret_cast<const int4*>(keys + b0);
#pragma unroll
      for (int q = 0; q < RS_IPT / 4; ++q) {
        const int4 v = p[q];
        d[4 * q] = ((unsigned)v.x >> shift) & mask;
        d[4 * q + 1] = ((unsigned)v.y >> shift) & mask;
        d[4 * q + 2] = ((unsigned)v.z >> shift) & mask;
        d[4 * q + 3] = ((unsigned)v.w >> shift) & mask;
      }
    } else {
#pragma unroll
      for (int k = 0; k < RS_IPT; ++k) d[k] = b0 + k < n ? (((unsigned)keys[b0 + k]) >> shift) & mask : NONE;
    }
  } else {
#pragma unroll
    for (int k = 0; k < RS_IPT; ++k) d[k] = b0 + k < n ? (((unsigned)keys[b0 + k]) >> shift) & mask : NONE;
  }
  unsigned cur = d[0], run = 0u;
#pragma unroll
  for (int k = 0; k < RS_IPT; ++k) {
    if (d[k] != cur) {
      if (cur != NONE) atomicAdd(&cnt[cur], run);
      cur = d[k];
      run = 0u;
    }
    ++run;
  }
  if (cur != NONE) atomicAdd(&cnt[cur], run);
  __syncthreads();
  hist[(size_t)tid * numTiles + blockIdx.x] = cnt[tid];
}

__global__ __launch_bounds__(256) void rs_scan_rows(unsigned* __restrict__ hist, int numTiles,
                                                    unsigned* __restrict__ totals) {
  __shared__ unsigned s[256];
  unsigned* row = hist + (size_t)blockIdx.x * numTiles;
  const int per = (numTiles + 255) / 256;
  const int b = threadIdx.x * per;
  const int e = min(b + per, numTiles);
  unsigned sum = 0;
  for (int t = b; t < e; ++t) sum += row[t];
  s[threadIdx.x] = sum;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    unsigned v = threadIdx.x >= off ? s[threadIdx.x - off] : 0u;
    __syncthreads();
    s[threadIdx.x] += v;
    __syncthreads();
  }
  unsigned run = threadIdx.x ? s[threadIdx.x - 1] : 0u;
  for (int t = b; t < e; ++t) {
    unsigned c = row[t];
    row[t] = run;
    run += c;
  }
  if (threadIdx.x == 255) totals[blockIdx.x] = s[255];
}

template <int RS_IPT>
__global__ __launch_bounds__(RS_THREADS) void rs_downsweep(
    const int* __restrict__ keys_in, const int* __restrict__ vals_in, int* __restrict__ keys_out,
    int* __restrict__ vals_out, int n, int shift, int nbits, const unsigned* __restrict__ hist,
    const unsigned* __restrict__ totals, int numTiles) {
  __shared__ unsigned wcnt[RS_WAVES][256];
  __shared__ unsigned dbase[256];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  for (int k = tid; k < RS_WAVES * 256; k += RS_THREADS) (&wcnt[0][0])[k] = 0u;
  // exclusive scan of the digit totals -> first output slot of each digit
  const unsigned own = totals[tid];
  dbase[tid] = own;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    unsigned v = tid >= off ? dbase[tid - off] : 0u;
    __syncthreads();
    dbase[tid] += v;
    __syncthreads();
  }
  const unsigned myBase = dbase[tid] - own + hist[(size_t)tid * numTiles + blockIdx.x];
  __syncthreads();
  dbase[tid] = myBase;

  const unsigned mask = (1u << nbits) - 1u;
  const unsigned long long ltmask = (1ull << lane) - 1ull;
  const long long base = (long long)blockIdx.x * (RS_THREADS * RS_IPT) + (long long)w * (64 * RS_IPT);
  int kk[RS_IPT], vv[RS_IPT];
  unsigned rk[RS_IPT];
#pragma unroll
  for (int r = 0; r < RS_IPT; ++r) {
    const long long i = base + r * 64 + lane;
    const bool valid = i < n;
    const int key = valid ? keys_in[i] : 0;
    const int val = valid ? vals_in[i] : 0;
    const unsigned d = (((unsigned)key) >> shift) & mask;
    const unsigned long long active = __ballot(valid);
    const unsigned long long peers = match_digit(d, nbits, active);
    const int leader = __ffsll((long long)peers) - 1;
    const unsigned before = wcnt[w][d];
    rk[r] = before + (unsigned)__popcll(peers & ltmask);
    if (valid && lane == leader) wcnt[w][d] = before + (unsigned)__popcll(peers);
    kk[r] = key;
    vv[r] = val;
  }
  __syncthreads();
  {
    // per digit: tile offset of each wave's run
    unsigned run = dbase[tid];
    for (int ww = 0; ww < RS_WAVES; ++ww) {
      unsigned c = wcnt[ww][tid];
      wcnt[ww][tid] = run;
      run += c;
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < RS_IPT; ++r) {
    const long long i = base + r * 64 + lane;
    if (i < n) {
      const unsigned d = (((unsigned)kk[r]) >> shift) & mask;
      const unsigned dst = wcnt[w][d] + rk[r];
      keys_out[dst] = kk[r];
      vals_out[dst] = vv[r];
    }
  }
}

__global__ void fill_int_kernel(int* __restrict__ p, int value, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = value;
}

}  // namespace

namespace {

// Keys per thread: the largest of 16, 8, 4, 2, 1 that still gives
// RS_MIN_TILES tiles (small sorts would otherwise run on a few dozen CUs).
int radix_ipt(int n) {
  int ipt = RS_IPT_MAX;
  while (ipt > 1 && (n + RS_THREADS * ipt - 1) / (RS_THREADS * ipt) < RS_MIN_TILES) ipt >>= 1;
  return ipt;
}

}  // namespace

int radix_hist_words(int n) {
  int tiles = (n + RS_THREADS - 1) / RS_THREADS;  // the smallest tile
  if (tiles < 1) tiles = 1;
  return 256 * tiles;
}

hipError_t fill_int(int* p, int value, long long n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  long long blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(fill_int_kernel, dim3((unsigned)blocks), dim3(256), 0, s, p, value, n);
  return hipGetLastError();
}

namespace {

template <int IPT>
hipError_t radix_passes(int* keys, int* vals, int* keys_alt, int* vals_alt, unsigned* hist, unsigned* totals, int n,
                        int bits, hipStream_t s, bool* result_in_alt) {
  const int numTiles = (n + RS_THREADS * IPT - 1) / (RS_THREADS * IPT);
  int* kin = keys;
  int* vin = vals;
  int* kout = keys_alt;
  int* vout = vals_alt;
  bool in_alt = false;
  for (int shift = 0; shift < bits; shift += 8) {
    const int nb = bits - shift < 8 ? bits - shift : 8;
    const int vec = ((uintptr_t)kin & 15u) == 0u;
    hipLaunchKernelGGL(rs_upsweep<IPT>, dim3(numTiles), dim3(RS_THREADS), 0, s, kin, n, shift, nb, hist, numTiles,
                       vec);
    hipLaunchKernelGGL(rs_scan_rows, dim3(256), dim3(256), 0, s, hist, numTiles, totals);
    hipLaunchKernelGGL(rs_downsweep<IPT>, dim3(numTiles), dim3(RS_THREADS), 0, s, kin, vin, kout, vout, n, shift, nb,
                       hist, totals, numTiles);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    int* t = kin; kin = kout; kout = t;
    t = vin; vin = vout; vout = t;
    in_alt = !in_alt;
  }
  *result_in_alt = in_alt;
  return hipSuccess;
}

}  // namespace

hipError_t radix_sort_pairs(int* keys, int* vals, int* keys_alt, int* vals_alt, unsigned* hist,
                            unsigned* totals, int n, int bits, hipStream_t s,
                            bool* result_in_alt) {
  *result_in_alt = false;
  if (n <= 1) return hipSuccess;
  switch (radix_ipt(n)) {
    case 16: return radix_passes<16>(keys, vals, keys_alt, vals_alt, hist, totals, n, bits, s, result_in_alt);
    case 8: return radix_passes<8>(keys, vals, keys_alt, vals_alt, hist, totals, n, bits, s, result_in_alt);
    case 4: return radix_passes<4>(keys, vals, keys_alt, vals_alt, hist, totals, n, bits, s, result_in_alt);
    case 2: return radix_passes<2>(keys, vals, keys_alt, vals_alt, hist, totals, n, bits, s, result_in_alt);
    default: return radix_passes<1>(keys, vals, keys_alt, vals_alt, hist, totals, n, bits, s, result_in_alt);
  }
}

}  // namespace sph
