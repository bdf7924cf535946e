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

// The digit of the previous lane: DPP wave_shr:1 (lane l reads lane l-1 within
// the same instruction; lane 0 keeps its own value, which is never used).
__device__ __forceinline__ unsigned prev_lane(unsigned d) {
  return (unsigned)__builtin_amdgcn_update_dpp((int)d, (int)d, 0x138, 0xf, 0xf, false);
}

// Lanes of the wave whose digit equals this lane's digit (among `active`,
// a prefix of the wave).
//
// The cell keys arrive nearly sorted (the previous step's order), and every
// LSD pass keeps each bucket of the previous digit in key order, so in almost
// every wave the digits do not decrease from lane to lane.  Then equal digits
// form one run of consecutive lanes, found with one DPP shift and one ballot.
// Otherwise (a bucket boundary, or unsorted keys) the digit is matched bit by
// bit with one ballot per bit.
__device__ __forceinline__ unsigned long long match_digit(unsigned d, int nbits,
                                                          unsigned long long active) {
  const int lane = __lane_id();
  const bool on = (active >> lane) & 1ull;
  const unsigned dp = prev_lane(d);
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

// Exclusive scan of one value per thread over a 256-thread block: a shuffle
// scan per wave, then the wave totals through `wsum` (one barrier).  `all`
// receives the block total.  `wsum` may be reused only after another barrier.
__device__ __forceinline__ unsigned block_exclusive_scan(unsigned v, unsigned* wsum, unsigned* all) {
  const int lane = __lane_id(), w = threadIdx.x >> 6;
  unsigned x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const unsigned y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  if (lane == 63) wsum[w] = x;
  __syncthreads();
  unsigned off = 0u, tot = 0u;
#pragma unroll
  for (int k = 0; k < RS_WAVES; ++k) {
    const unsigned s = wsum[k];
    off += k < w ? s : 0u;
    tot += s;
  }
  *all = tot;
  return off + x - v;
}

// One block per digit turns its row of per-tile counts into exclusive tile
// offsets.  Up to SCAN_REG entries per thread are read once into registers
// (every load in flight together); longer rows take a loop.
constexpr int SCAN_REG = 8;

__global__ __launch_bounds__(256) void rs_scan_rows(unsigned* __restrict__ hist, int numTiles,
                                                    unsigned* __restrict__ totals) {
  __shared__ unsigned wsum[RS_WAVES];
  unsigned* row = hist + (size_t)blockIdx.x * numTiles;
  const int per = (numTiles + 255) / 256;
  const int b = threadIdx.x * per;
  const int e = min(b + per, numTiles);
  unsigned all;
  if (per <= SCAN_REG) {
    unsigned c[SCAN_REG];
    unsigned sum = 0u;
#pragma unroll
    for (int k = 0; k < SCAN_REG; ++k) {
      c[k] = b + k < e ? row[b + k] : 0u;
      sum += c[k];
    }
    unsigned run = block_exclusive_scan(sum, wsum, &all);
#pragma unroll
    for (int k = 0; k < SCAN_REG; ++k) {
      if (b + k < e) row[b + k] = run;
      run += c[k];
    }
  } else {
    unsigned sum = 0u;
    for (int t = b; t < e; ++t) sum += row[t];
    unsigned run = block_exclusive_scan(sum, wsum, &all);
    for (int t = b; t < e; ++t) {
      const unsigned c = row[t];
      row[t] = run;
      run += c;
    }
  }
  if (threadIdx.x == 0) totals[blockIdx.x] = all;
}

template <int RS_IPT>
__global__ __launch_bounds__(RS_THREADS) void rs_downsweep(
    const int* __restrict__ keys_in, const int* __restrict__ vals_in, int* __restrict__ keys_out,
    int* __restrict__ vals_out, int n, int shift, int nbits, const unsigned* __restrict__ hist,
    const unsigned* __restrict__ totals, int numTiles) {
  __shared__ unsigned wcnt[RS_WAVES][256];
  __shared__ unsigned dbase[256];
  __shared__ unsigned wsum[RS_WAVES];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const long long base = (long long)blockIdx.x * (RS_THREADS * RS_IPT) + (long long)w * (64 * RS_IPT);
  // The digit totals and this tile's offsets are loaded first, then every key
  // and value of the wave, all before the first barrier: the wave waits for
  // HBM once (loaded inside the ranking rounds, each round waited for its own
  // loads), and the scan of the totals waits only for the first two loads.
  const unsigned own = totals[tid];
  const unsigned tileOff = hist[(size_t)tid * numTiles + blockIdx.x];
  __builtin_amdgcn_sched_barrier(0);  // keep these two loads ahead of the rest
  // Lanes past the end load the last element (n >= 2 here): one straight-line
  // path keeps the compiler's wait counts exact.  Those lanes stay out of the
  // ballots and are never stored.
  int kk[RS_IPT], vv[RS_IPT];
#pragma unroll
  for (int r = 0; r < RS_IPT; ++r) kk[r] = keys_in[min(base + r * 64 + lane, (long long)n - 1)];
#pragma unroll
  for (int r = 0; r < RS_IPT; ++r) vv[r] = vals_in[min(base + r * 64 + lane, (long long)n - 1)];
  for (int k = tid; k < RS_WAVES * 256; k += RS_THREADS) (&wcnt[0][0])[k] = 0u;
  // first output slot of each digit in this tile: exclusive scan of the digit
  // totals plus the tile's offset within the digit (the scan's barrier also
  // orders the clearing of wcnt before the ranking rounds)
  unsigned all;
  dbase[tid] = block_exclusive_scan(own, wsum, &all) + tileOff;

  const unsigned mask = (1u << nbits) - 1u;
  const unsigned long long ltmask = (1ull << lane) - 1ull;
  unsigned rk[RS_IPT];
#pragma unroll
  for (int r = 0; r < RS_IPT; ++r) {
    const bool valid = base + r * 64 + lane < n;
    const unsigned d = (((unsigned)kk[r]) >> shift) & mask;
    const unsigned long long active = __ballot(valid);
    const unsigned long long peers = match_digit(d, nbits, active);
    const int leader = __ffsll((long long)peers) - 1;
    const unsigned before = wcnt[w][d];
    rk[r] = before + (unsigned)__popcll(peers & ltmask);
    if (valid && lane == leader) wcnt[w][d] = before + (unsigned)__popcll(peers);
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
