// gfx950 kernels of the message-passing surrogate: a BF16 MFMA GEMM with
// FP32 accumulation and fused gather / concat / bias / ReLU / residual /
// ReLU-mask epilogues, plus the transposes, reductions, graph aggregations and
// the Adam update that training needs.
//
// MFMA: v_mfma_f32_16x16x32_bf16 (one wave, D[16x16] += A[16x32] * B[32x16]).
// Operand layout used here (lane l, element q):
//   A[l % 16][8 * (l / 16) + q],  B[8 * (l / 16) + q][l % 16],
//   D[4 * (l / 16) + q][l % 16]   (q < 4)
// The tests check it against numpy through mfma_tile().
#include "gnn_kernels.h"

#include <algorithm>

namespace sph {
namespace gnn {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int BM = 128;       // rows per block (4 waves x 32)
constexpr int BN = 128;       // columns per block
constexpr int BK = 32;        // reduction step (one MFMA K)
constexpr int LDB = BK + 8;   // LDS row stride of the staged B tile (bf16 elements)

inline unsigned cdiv(long long a, long long b) { return (unsigned)((a + b - 1) / b); }

__device__ __forceinline__ bf16x8 cvt8(float4 x, float4 y) {
  bf16x8 v;
  v[0] = (bf16)x.x; v[1] = (bf16)x.y; v[2] = (bf16)x.z; v[3] = (bf16)x.w;
  v[4] = (bf16)y.x; v[5] = (bf16)y.y; v[6] = (bf16)y.z; v[7] = (bf16)y.w;
  return v;
}

// Raw (unconverted) 8-element A fragment: two 16-byte words holding 8 FP32
// values, or 8 BF16 values in the first word.  Loading raw and converting one
// k-step later keeps the loads in flight across the MFMAs of the current step.
struct ARaw {
  uint4 a, b;
  int f32;
};

// The row of each segment that operand row `row` reads (its gathered index
// for a gathered segment), computed once per kernel.  Rows past M read row
// M - 1: their products land only in output rows that are never stored.
struct SegRows {
  int r[3];
};

__device__ __forceinline__ SegRows seg_rows(const AOperand& A, int row, int M) {
  const int rw = min(row, M - 1);
  SegRows s;
  s.r[0] = A.seg[0].rows ? A.seg[0].rows[rw] : rw;
  s.r[1] = A.nseg > 1 ? (A.seg[1].rows ? A.seg[1].rows[rw] : rw) : 0;
  s.r[2] = A.nseg > 2 ? (A.seg[2].rows ? A.seg[2].rows[rw] : rw) : 0;
  return s;
}

// The segment holding the k-step that starts at k0.  k0 is the same for all
// lanes (segment widths are multiples of the 32-wide k-step), so the choice
// stays in scalar registers; fields are picked with constant indices (a
// pointer into the kernel-argument struct would move it to memory).
struct SegSel {
  const void* ptr;
  int ld, is_f32, i, kk;
};

__device__ __forceinline__ SegSel seg_at(const AOperand& A, int k0) {
  const int w0 = A.seg[0].width, w1 = A.seg[1].width;
  SegSel s;
  s.i = (A.nseg == 1 || k0 < w0) ? 0 : (A.nseg == 2 || k0 < w0 + w1) ? 1 : 2;
  s.kk = s.i == 0 ? k0 : s.i == 1 ? k0 - w0 : k0 - w0 - w1;
  s.ptr = s.i == 0 ? A.seg[0].ptr : s.i == 1 ? A.seg[1].ptr : A.seg[2].ptr;
  s.ld = s.i == 0 ? A.seg[0].ld : s.i == 1 ? A.seg[1].ld : A.seg[2].ld;
  s.is_f32 = s.i == 0 ? A.seg[0].is_f32 : s.i == 1 ? A.seg[1].is_f32 : A.seg[2].is_f32;
  return s;
}

// K order inside a k-step.  BF16 operands: lane group lq (= lane / 16) holds
// k = 8 lq .. 8 lq + 7 (one 16-byte load per lane, 64 contiguous bytes per
// row).  FP32 operands: lane group lq holds k = 4 lq .. 4 lq + 3 and
// 16 + 4 lq .. 16 + 4 lq + 3, so each of its two 16-byte loads reads 64
// contiguous bytes of the row across the four lane groups (whole cache
// lines); the B tile is staged in LDS in the same order (store_b).  A
// product's k order only permutes the 32 terms of one MFMA.
__device__ __forceinline__ ARaw load_raw(const SegSel& sel, const SegRows& rows, int lq) {
  const int rr = sel.i == 0 ? rows.r[0] : sel.i == 1 ? rows.r[1] : rows.r[2];
  ARaw r;
  if (sel.is_f32) {
    const float* p = static_cast<const float*>(sel.ptr) + (size_t)rr * sel.ld + sel.kk + 4 * lq;
    r.a = *reinterpret_cast<const uint4*>(p);
    r.b = *reinterpret_cast<const uint4*>(p + 16);
    r.f32 = 1;
  } else {
    r.a = *reinterpret_cast<const uint4*>(static_cast<const bf16*>(sel.ptr) + (size_t)rr * sel.ld + sel.kk + 8 * lq);
    r.b = make_uint4(0u, 0u, 0u, 0u);
    r.f32 = 0;
  }
  return r;
}

// The 128 x 32 piece [k0, k0 + 32) of B (row n contiguous in k, row stride
// ldb): thread tid loads the 16-byte parts tid and tid + 256 (row part / 4,
// quarter part % 4), four lanes per row.
__device__ __forceinline__ void load_b(const bf16* __restrict__ B, int ldb, int k0, int tid, uint4& q0, uint4& q1) {
  q0 = *reinterpret_cast<const uint4*>(B + (size_t)(tid >> 2) * ldb + k0 + 8 * (tid & 3));
  q1 = *reinterpret_cast<const uint4*>(B + (size_t)((tid + 256) >> 2) * ldb + k0 + 8 * (tid & 3));
}

// Stores the two parts of load_b into the LDS tile [n][LDB] in the k order
// of the A operand: natural (perm = 0, BF16 A) or the FP32 order (perm = 1),
// where k = 8 p .. 8 p + 3 of quarter p go to slots lo(p) .. lo(p) + 3 and
// k = 8 p + 4 .. 8 p + 7 to lo(p) + 8 .., lo(p) = 16 p (p < 2), 16 (p - 2) + 4.
__device__ __forceinline__ void store_b(bf16* dst, uint4 q0, uint4 q1, int tid, int perm) {
  const int part = tid & 3;
  bf16* r0 = dst + (tid >> 2) * LDB;
  bf16* r1 = dst + ((tid + 256) >> 2) * LDB;
  if (!perm) {
    *reinterpret_cast<uint4*>(r0 + 8 * part) = q0;
    *reinterpret_cast<uint4*>(r1 + 8 * part) = q1;
  } else {
    const int lo = part < 2 ? 16 * part : 16 * (part - 2) + 4;
    *reinterpret_cast<uint2*>(r0 + lo) = make_uint2(q0.x, q0.y);
    *reinterpret_cast<uint2*>(r0 + lo + 8) = make_uint2(q0.z, q0.w);
    *reinterpret_cast<uint2*>(r1 + lo) = make_uint2(q1.x, q1.y);
    *reinterpret_cast<uint2*>(r1 + lo + 8) = make_uint2(q1.z, q1.w);
  }
}

__device__ __forceinline__ bf16x8 convert8(const ARaw& r) {
  if (r.f32) {
    float4 x, y;
    __builtin_memcpy(&x, &r.a, 16);
    __builtin_memcpy(&y, &r.b, 16);
    return cvt8(x, y);
  }
  bf16x8 v;
  __builtin_memcpy(&v, &r.a, 16);
  return v;
}

__device__ __forceinline__ float load_a1(const AOperand& A, int row, int k) {
  const ASeg* g = &A.seg[0];
  int kk = k;
  if (A.nseg > 1 && kk >= A.seg[0].width) {
    kk -= A.seg[0].width;
    g = &A.seg[1];
    if (A.nseg > 2 && kk >= A.seg[1].width) {
      kk -= A.seg[1].width;
      g = &A.seg[2];
    }
  }
  const int r = g->rows ? g->rows[row] : row;
  if (g->is_f32) return static_cast<const float*>(g->ptr)[(size_t)r * g->ld + kk];
  return (float)static_cast<const bf16*>(g->ptr)[(size_t)r * g->ld + kk];
}

// Main loop, software-pipelined: the A fragments and the B piece of step
// k + 1 are loaded while the MFMAs of step k run; the B tile is double
// buffered in LDS, so one barrier per step suffices (a buffer is rewritten two
// steps after it was read, and every wave has passed the barrier in between).
__global__ __launch_bounds__(256) void gemm_kernel(GemmArgs g) {
  __shared__ bf16 ldsB[2][BN * LDB];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int li = lane & 15, lq = lane >> 4;
  const int m0 = blockIdx.x * BM + w * 32;
  const int n0 = blockIdx.y * BN;
  const int z = blockIdx.z;
  int kbeg = 0, kend = g.K;
  if (g.ksplit > 1) {
    const int chunk = ((g.K + g.ksplit - 1) / g.ksplit + BK - 1) / BK * BK;
    kbeg = z * chunk;
    kend = min(g.K, kbeg + chunk);
  }
  f32x4 acc[2][8];
  for (int t = 0; t < 2; ++t)
    for (int c = 0; c < 8; ++c) acc[t][c] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  const int ldb = g.ldb > 0 ? g.ldb : g.K;
  const bf16* Bn = g.B + (size_t)n0 * ldb;
  uint4 bq0 = make_uint4(0u, 0u, 0u, 0u), bq1 = bq0;
  ARaw ar0{}, ar1{};
  int perm = 0;  // k order of the step whose operands are in flight (FP32 A: 1)
  const SegRows rows0 = seg_rows(g.A, m0 + li, g.M), rows1 = seg_rows(g.A, m0 + 16 + li, g.M);
  if (kbeg < kend) {
    load_b(Bn, ldb, kbeg, tid, bq0, bq1);
    const SegSel sel = seg_at(g.A, kbeg);
    perm = sel.is_f32;
    ar0 = load_raw(sel, rows0, lq);
    ar1 = load_raw(sel, rows1, lq);
  }
  int buf = 0;
  for (int k0 = kbeg; k0 < kend; k0 += BK) {
    store_b(ldsB[buf], bq0, bq1, tid, perm);
    bf16x8 af[2];
    af[0] = convert8(ar0);
    af[1] = convert8(ar1);
    __syncthreads();
    if (k0 + BK < kend) {  // next step's operands, in flight during the MFMAs
      load_b(Bn, ldb, k0 + BK, tid, bq0, bq1);
      const SegSel sel = seg_at(g.A, k0 + BK);
      perm = sel.is_f32;
      ar0 = load_raw(sel, rows0, lq);
      ar1 = load_raw(sel, rows1, lq);
    }
    const bf16* tileB = ldsB[buf];
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      const bf16x8 bf = *reinterpret_cast<const bf16x8*>(&tileB[(16 * c + li) * LDB + 8 * lq]);
      acc[0][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[0], bf, acc[0][c], 0, 0, 0);
      acc[1][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[1], bf, acc[1][c], 0, 0, 0);
    }
    buf ^= 1;
  }
  // Epilogue, one 16-row half at a time: the operands it reads (bias,
  // residual, mask) are all loaded before any is used -- rows past M read
  // row M - 1 -- so their latencies overlap; stores are predicated per row.
  const bool fin = g.ksplit <= 1;
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    float rs[8][4], bs[8];
    bool keep[8][4];
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      const int n = n0 + 16 * c + li;
      bs[c] = (fin && g.bias) ? g.bias[n] : 0.0f;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const size_t o = (size_t)min(m0 + 16 * t + 4 * lq + r, g.M - 1) * g.ldo + n;
        rs[c][r] = (fin && g.resid) ? g.resid[o] : 0.0f;
        keep[c][r] = (fin && g.mask) ? ((float)g.mask[o] > 0.0f) : true;
      }
    }
#pragma unroll
    for (int c = 0; c < 8; ++c)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + 16 * t + 4 * lq + r;
        const size_t o = (size_t)m * g.ldo + n0 + 16 * c + li;
        float v = acc[t][c][r];
        if (!fin) {
          if (m < g.M) g.out_f32[(size_t)z * g.M * g.ldo + o] = v;
          continue;
        }
        if (g.bias) v += bs[c];
        if (g.resid) v += rs[c][r];
        if (g.relu) v = fmaxf(v, 0.0f);
        if (!keep[c][r]) v = 0.0f;
        if (m < g.M) {
          if (g.out_f32) g.out_f32[o] = v;
          if (g.out_bf16) g.out_bf16[o] = (bf16)v;
        }
      }
  }
}

// ---------------------------------------------------------------- fused MLP
constexpr int LDH = 128 + 8;  // LDS row stride of a wave's hidden activations (bf16 elements)

__device__ __forceinline__ void mfma_step(f32x4 (&acc)[2][8], const bf16x8 (&af)[2], const bf16* tileB, int li,
                                          int lq) {
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const bf16x8 bf = *reinterpret_cast<const bf16x8*>(&tileB[(16 * c + li) * LDB + 8 * lq]);
    acc[0][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[0], bf, acc[0][c], 0, 0, 0);
    acc[1][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[1], bf, acc[1][c], 0, 0, 0);
  }
}

// relu(acc + bias) as BF16 into the wave's LDS rows (and to global h if given)
__device__ __forceinline__ void hidden_out(const f32x4 (&acc)[2][8], const float* __restrict__ bias, bf16* H,
                                           bf16* __restrict__ h, int m0, int M, int li, int lq) {
#pragma unroll
  for (int t = 0; t < 2; ++t)
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      const int n = 16 * c + li;
      const float b = bias[n];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = 16 * t + 4 * lq + r;
        const bf16 v = (bf16)fmaxf(acc[t][c][r] + b, 0.0f);
        H[row * LDH + n] = v;
        if (h && m0 + row < M) h[(size_t)(m0 + row) * 128 + n] = v;
      }
    }
}

// Operand loads run D k-steps ahead of the MFMAs (a ring of D register
// sets, indexed with compile-time constants): with one step of 16 MFMAs per
// wave far shorter than a gather's latency, deeper rings keep more bytes in
// flight per wave.

// One 128 x 128 layer whose A operand is the wave's LDS rows; continues the
// block-wide double-buffered B staging (buf toggles once per k-step).
template <int D>
__device__ __forceinline__ void lds_layer(f32x4 (&acc)[2][8], const bf16* H, const bf16* __restrict__ Wt,
                                          bf16 (*ldsB)[BN * LDB], int& buf, int tid, int li, int lq) {
  constexpr int steps = 128 / BK;
  for (int t = 0; t < 2; ++t)
    for (int c = 0; c < 8; ++c) acc[t][c] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  uint4 q[D][2];
#pragma unroll
  for (int d = 0; d < D; ++d) load_b(Wt, 128, d * BK, tid, q[d][0], q[d][1]);
  static_assert(steps % D == 0, "ring depth must divide the k-steps");
#pragma unroll 1
  for (int st0 = 0; st0 < steps; st0 += D)
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const int st = st0 + d;
      store_b(ldsB[buf], q[d][0], q[d][1], tid, 0);
      __syncthreads();
      load_b(Wt, 128, min(st + D, steps - 1) * BK, tid, q[d][0], q[d][1]);
      bf16x8 af[2];
      af[0] = *reinterpret_cast<const bf16x8*>(&H[li * LDH + st * BK + 8 * lq]);
      af[1] = *reinterpret_cast<const bf16x8*>(&H[(16 + li) * LDH + st * BK + 8 * lq]);
      mfma_step(acc, af, ldsB[buf], li, lq);
      buf ^= 1;
    }
}

// The FP32 A fragment of lane group lq in the FP32 k order (see load_raw)
__device__ __forceinline__ void load_f32x8(const SegSel& sel, const SegRows& rows, int lq, uint4& a, uint4& b) {
  const int rr = sel.i == 0 ? rows.r[0] : sel.i == 1 ? rows.r[1] : rows.r[2];
  const float* p = static_cast<const float*>(sel.ptr) + (size_t)rr * sel.ld + sel.kk + 4 * lq;
  a = *reinterpret_cast<const uint4*>(p);
  b = *reinterpret_cast<const uint4*>(p + 16);
}

__device__ __forceinline__ bf16x8 cvt_raw(uint4 a, uint4 b) {
  float4 x, y;
  __builtin_memcpy(&x, &a, 16);
  __builtin_memcpy(&y, &b, 16);
  return cvt8(x, y);
}

template <int D>
__global__ __launch_bounds__(256) void mlp3_kernel(Mlp3Args g) {
  __shared__ bf16 ldsB[2][BN * LDB];
  __shared__ bf16 ldsH[4][32 * LDH];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int li = lane & 15, lq = lane >> 4;
  const int m0 = blockIdx.x * BM + w * 32;
  bf16* H = ldsH[w];
  f32x4 acc[2][8];
  for (int t = 0; t < 2; ++t)
    for (int c = 0; c < 8; ++c) acc[t][c] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  int buf = 0;
  // layer 1: A gathered from global
  {
    const int steps = g.K1 / BK;  // uniform over the block: the barrier below is reached by all threads

    const SegRows rows0 = seg_rows(g.A, m0 + li, g.M), rows1 = seg_rows(g.A, m0 + 16 + li, g.M);
    uint4 q[D][2], a0[D][2], a1[D][2];
#pragma unroll
    for (int d = 0; d < D; ++d) {  // D <= steps (checked by the launcher)
        load_b(g.W1, g.K1, d * BK, tid, q[d][0], q[d][1]);
        const SegSel sel = seg_at(g.A, d * BK);
        load_f32x8(sel, rows0, lq, a0[d][0], a0[d][1]);
        load_f32x8(sel, rows1, lq, a1[d][0], a1[d][1]);
      }
#pragma unroll 1
    for (int st0 = 0; st0 < steps; st0 += D)  // steps % D == 0 (checked by the launcher)
#pragma unroll
      for (int d = 0; d < D; ++d) {
        const int st = st0 + d;
        store_b(ldsB[buf], q[d][0], q[d][1], tid, 1);
        bf16x8 af[2];
        af[0] = cvt_raw(a0[d][0], a0[d][1]);
        af[1] = cvt_raw(a1[d][0], a1[d][1]);
        __syncthreads();
        {  // unconditional (the last steps reload the last k-step): no register copies
          const int k = min(st + D, steps - 1) * BK;
          load_b(g.W1, g.K1, k, tid, q[d][0], q[d][1]);
          const SegSel sel = seg_at(g.A, k);
          load_f32x8(sel, rows0, lq, a0[d][0], a0[d][1]);
          load_f32x8(sel, rows1, lq, a1[d][0], a1[d][1]);
        }
        mfma_step(acc, af, ldsB[buf], li, lq);
        buf ^= 1;
      }
  }
  hidden_out(acc, g.b1, H, g.h1, m0, g.M, li, lq);
  lds_layer<D>(acc, H, g.W2, ldsB, buf, tid, li, lq);
  hidden_out(acc, g.b2, H, g.h2, m0, g.M, li, lq);
  lds_layer<D>(acc, H, g.W3, ldsB, buf, tid, li, lq);
  // residuals loaded in a batch per row half (rows past M read row M - 1)
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    float rs[8][4];
#pragma unroll
    for (int c = 0; c < 8; ++c)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        rs[c][r] = g.resid ? g.resid[(size_t)min(m0 + 16 * t + 4 * lq + r, g.M - 1) * g.ldo + 16 * c + li] : 0.0f;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      const int n = 16 * c + li;
      const float b = g.b3[n];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + 16 * t + 4 * lq + r;
        float v = acc[t][c][r] + b;
        if (g.resid) v += rs[c][r];
        if (m < g.M) {
          if (g.out_bf16)
            g.out_bf16[(size_t)m * g.ldo + n] = (bf16)v;
          else
            g.out[(size_t)m * g.ldo + n] = v;
        }
      }
    }
  }
}

// ------------------------------------------------- persistent edge MLP (inference)
// One block of EW waves per CU holds the three 128 x 128 weight matrices in
// LDS (96 KB, 16-byte chunks XOR-swizzled by row: the 16 rows of an MFMA
// fragment read hit 16 different chunks, free of bank conflicts) and one
// padded 32 x 128 hidden-activation tile per wave (row stride LDH); the LDS
// holds 7 such tiles.  Waves walk their tiles independently.
constexpr int EW = 7;
constexpr int LDF = 128 + 4;  // FP32 row stride of the epilogue's staging rows (16 x LDF x 4 B <= 32 x LDH x 2 B)
static_assert(16 * LDF * 4 <= 32 * LDH * 2, "staging rows must fit in the hidden tile");

__device__ __forceinline__ int swz(int row, int col) { return row * 128 + ((((col >> 3) ^ (row & 15))) << 3) + (col & 7); }

template <bool BE>  // BE: edge latents stored as BF16 (g.Eb), else FP32 (g.E)
__global__ __launch_bounds__(64 * EW) void edge_mlp_kernel(EdgeMlpArgs g) {
  __shared__ __attribute__((aligned(16))) bf16 Wl[3][128 * 128];
  __shared__ __attribute__((aligned(16))) bf16 Hs[EW][32 * LDH];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int li = lane & 15, lq = lane >> 4;
  // weights: 3 x 2048 chunks of 8; W1e in the FP32 k order of the A operand
  // (chunk 4 s + j of row n holds k = 32 s + 4 j .. + 3 and 32 s + 16 + 4 j .. + 3)
  for (int c = tid; c < 3 * 2048; c += 64 * EW) {
    const int mtx = c >> 11, n = (c >> 4) & 127, ch = c & 15;
    uint4 v;
    if (mtx == 0 && !BE) {  // FP32 A operand: the FP32 k order
      const bf16* src = g.W1 + (size_t)n * g.ldw1 + 32 * (ch >> 2) + 4 * (ch & 3);
      const uint2 lo = *reinterpret_cast<const uint2*>(src);
      const uint2 hi = *reinterpret_cast<const uint2*>(src + 16);
      v = make_uint4(lo.x, lo.y, hi.x, hi.y);
    } else if (mtx == 0) {  // BF16 A operand: natural k order
      v = *reinterpret_cast<const uint4*>(g.W1 + (size_t)n * g.ldw1 + 8 * ch);
    } else {
      v = *reinterpret_cast<const uint4*>((mtx == 1 ? g.W2 : g.W3) + (size_t)n * 128 + 8 * ch);
    }
    *reinterpret_cast<uint4*>(&Wl[mtx][swz(n, 8 * ch)]) = v;
  }
  __syncthreads();
  bf16* H = Hs[w];
  // B fragment of lane (li, lq), output block c, k-step s: row 16 c + li,
  // chunk (4 s + lq) ^ li = 4 (s ^ (li >> 2)) + (lq ^ (li & 3)); the
  // 16 c rows are an immediate offset
  const int wlane = li * 128 + 8 * (lq ^ (li & 3));
  const int sx = li >> 2;
  const int ntiles = (g.ne + 31) / 32;
  // each wave walks a contiguous chunk of tiles (g.chunk of them), so the
  // aggregation below can carry per-receiver sums from tile to tile
  const int t_begin = (blockIdx.x * EW + w) * g.chunk;
  const int t_end = min(t_begin + g.chunk, ntiles);
  if (t_begin >= t_end) return;  // wave-uniform; no barrier follows
  const int row0 = t_begin * 32, row1 = min(t_end * 32, g.ne);
  // receivers cut by this chunk's ends are summed by agg_fixup_kernel
  const int cut_first = (row0 > 0 && g.recv[row0 - 1] == g.recv[row0]) ? g.recv[row0] : -1;
  const int cut_last = (row1 < g.ne && g.recv[row1] == g.recv[row1 - 1]) ? g.recv[row1 - 1] : -1;
  int cur = g.recv[row0];  // receiver of the running sum
  float2 run = make_float2(0.0f, 0.0f);
  // receiver / sender ids of this lane's rows 16 t + 4 lq + r of a tile
  auto load_ids = [&](int tl, int (&ri)[2][4], int (&si)[2][4]) {
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = min(tl * 32 + 16 * t + 4 * lq + r, g.ne - 1);
        ri[t][r] = g.recv[m];
        si[t][r] = g.send[m];
      }
  };
  int ri[2][4], si[2][4];
  load_ids(t_begin, ri, si);
  for (int tile = t_begin; tile < t_end; ++tile) {  // wave-uniform: every wave exits
    const int m0 = tile * 32;
    const int rvec = g.recv[min(m0 + (lane & 31), g.ne - 1)];  // lane k < 32: receiver of row m0 + k
    // one memory round trip per tile: the edge latents (layer-1 operand) and
    // the node projections, whose sum P[recv] + P[send] seeds the
    // accumulators (ids loaded during the previous tile); projection columns
    // 8 li .. 8 li + 7 hold this lane's features 16 c + li
    uint4 ea[4][2][2];  // FP32: two words per fragment (FP32 k order); BF16: word 0 (natural order)
    if (BE) {
      const bf16* pa = g.Eb + (size_t)min(m0 + li, g.ne - 1) * 128 + 8 * lq;
      const bf16* pb = g.Eb + (size_t)min(m0 + 16 + li, g.ne - 1) * 128 + 8 * lq;
#pragma unroll
      for (int s4 = 0; s4 < 4; ++s4) {
        ea[s4][0][0] = *reinterpret_cast<const uint4*>(pa + 32 * s4);
        ea[s4][1][0] = *reinterpret_cast<const uint4*>(pb + 32 * s4);
        ea[s4][0][1] = ea[s4][1][1] = make_uint4(0u, 0u, 0u, 0u);
      }
    } else {
      const float* pa = g.E + (size_t)min(m0 + li, g.ne - 1) * 128 + 4 * lq;
      const float* pb = g.E + (size_t)min(m0 + 16 + li, g.ne - 1) * 128 + 4 * lq;
#pragma unroll
      for (int s4 = 0; s4 < 4; ++s4) {
        ea[s4][0][0] = *reinterpret_cast<const uint4*>(pa + 32 * s4);
        ea[s4][0][1] = *reinterpret_cast<const uint4*>(pa + 32 * s4 + 16);
        ea[s4][1][0] = *reinterpret_cast<const uint4*>(pb + 32 * s4);
        ea[s4][1][1] = *reinterpret_cast<const uint4*>(pb + 32 * s4 + 16);
      }
    }
    f32x4 acc[2][8];
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const float4* pr = reinterpret_cast<const float4*>(g.P + (size_t)ri[t][r] * 256 + 8 * li);
        const float4* ps = reinterpret_cast<const float4*>(g.P + (size_t)si[t][r] * 256 + 128 + 8 * li);
        const float4 r0 = pr[0], r1 = pr[1], s0 = ps[0], s1 = ps[1];
        acc[t][0][r] = r0.x + s0.x;
        acc[t][1][r] = r0.y + s0.y;
        acc[t][2][r] = r0.z + s0.z;
        acc[t][3][r] = r0.w + s0.w;
        acc[t][4][r] = r1.x + s1.x;
        acc[t][5][r] = r1.y + s1.y;
        acc[t][6][r] = r1.z + s1.z;
        acc[t][7][r] = r1.w + s1.w;
      }
    // layer 1 (E block of W1); A = E rows in the FP32 k order
#pragma unroll
    for (int s4 = 0; s4 < 4; ++s4) {
      bf16x8 af[2];
      if (BE) {
        __builtin_memcpy(&af[0], &ea[s4][0][0], 16);
        __builtin_memcpy(&af[1], &ea[s4][1][0], 16);
      } else {
        af[0] = cvt_raw(ea[s4][0][0], ea[s4][0][1]);
        af[1] = cvt_raw(ea[s4][1][0], ea[s4][1][1]);
      }
      const bf16* wb = &Wl[0][wlane + 32 * (s4 ^ sx)];
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        const bf16x8 bf = *reinterpret_cast<const bf16x8*>(wb + 2048 * c);
        acc[0][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[0], bf, acc[0][c], 0, 0, 0);
        acc[1][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[1], bf, acc[1][c], 0, 0, 0);
      }
    }
    // + bias, ReLU, into H
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        const float b = g.b1[16 * c + li];
#pragma unroll
        for (int r = 0; r < 4; ++r)
          H[(16 * t + 4 * lq + r) * LDH + 16 * c + li] = (bf16)fmaxf(acc[t][c][r] + b, 0.0f);
      }
    // the residual rows (row-contiguous, 32 lanes per row) and the next
    // tile's ids load during layers 2 and 3
    const int c4 = lane & 31;
    float4 rr[2][8];
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const size_t row = (size_t)min(m0 + 16 * t + 2 * j + (lane >> 5), g.ne - 1);
        if (BE) {
          const uint2 q = *reinterpret_cast<const uint2*>(g.Eb + row * 128 + 4 * c4);
          rr[t][j] = make_float4(__uint_as_float(q.x << 16), __uint_as_float(q.x & 0xffff0000u),
                                 __uint_as_float(q.y << 16), __uint_as_float(q.y & 0xffff0000u));
        } else {
          rr[t][j] = *reinterpret_cast<const float4*>(g.E + row * 128 + 4 * c4);
        }
      }
    load_ids(min(tile + 1, ntiles - 1), ri, si);
    // layers 2 and 3 (A = H, rows 16 t + li at k 32 s + 8 lq)
#pragma unroll
    for (int L = 1; L < 3; ++L) {
#pragma unroll
      for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int c = 0; c < 8; ++c) acc[t][c] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int s4 = 0; s4 < 4; ++s4) {
        bf16x8 af[2];
        af[0] = *reinterpret_cast<const bf16x8*>(&H[li * LDH + 32 * s4 + 8 * lq]);
        af[1] = *reinterpret_cast<const bf16x8*>(&H[(16 + li) * LDH + 32 * s4 + 8 * lq]);
        const bf16* wb = &Wl[L][wlane + 32 * (s4 ^ sx)];
#pragma unroll
        for (int c = 0; c < 8; ++c) {
          const bf16x8 bf = *reinterpret_cast<const bf16x8*>(wb + 2048 * c);
          acc[0][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[0], bf, acc[0][c], 0, 0, 0);
          acc[1][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[1], bf, acc[1][c], 0, 0, 0);
        }
      }
      if (L == 1) {
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
          for (int c = 0; c < 8; ++c) {
            const float b = g.b2[16 * c + li];
#pragma unroll
            for (int r = 0; r < 4; ++r)
              H[(16 * t + 4 * lq + r) * LDH + 16 * c + li] = (bf16)fmaxf(acc[t][c][r] + b, 0.0f);
          }
      }
    }
    // + bias + residual, in place: each 16-row half goes through the wave's
    // (now free) hidden-tile LDS as FP32 rows, so the residual is read and
    // the result written with row-contiguous 16-byte accesses
    float* F = reinterpret_cast<float*>(H);  // 16 rows x LDF floats fit in the tile
#pragma unroll
    for (int t = 0; t < 2; ++t) {
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        const float b = g.b3[16 * c + li];
#pragma unroll
        for (int r = 0; r < 4; ++r) F[(4 * lq + r) * LDF + 16 * c + li] = acc[t][c][r] + b;
      }
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int row = 2 * j + (lane >> 5), m = m0 + 16 * t + row;
        const float4 v = *reinterpret_cast<const float4*>(&F[row * LDF + 4 * c4]);
        const float4 e = rr[t][j];
        float4 o = make_float4(v.x + e.x, v.y + e.y, v.z + e.z, v.w + e.w);
        if (BE) {  // stored as BF16; the aggregation sums the stored values
          const bf16 b0 = (bf16)o.x, b1 = (bf16)o.y, b2 = (bf16)o.z, b3 = (bf16)o.w;
          o = make_float4((float)b0, (float)b1, (float)b2, (float)b3);
          unsigned short u[4];
          __builtin_memcpy(&u[0], &b0, 2);
          __builtin_memcpy(&u[1], &b1, 2);
          __builtin_memcpy(&u[2], &b2, 2);
          __builtin_memcpy(&u[3], &b3, 2);
          if (m < g.ne)
            *reinterpret_cast<uint2*>(g.Eb + (size_t)m * 128 + 4 * c4) =
                make_uint2(u[0] | ((unsigned)u[1] << 16), u[2] | ((unsigned)u[3] << 16));
        } else if (m < g.ne) {
          *reinterpret_cast<float4*>(g.E + (size_t)m * 128 + 4 * c4) = o;
        }
        *reinterpret_cast<float4*>(&F[row * LDF + 4 * c4]) = o;  // kept for the aggregation
      }
      // aggregation A[recv] = sum of the new latents of the receiver's edges,
      // in edge order as segment_sum_kernel adds them: lane l owns columns
      // 2 l, 2 l + 1; a receiver's sum is written when its last edge is seen
#pragma unroll
      for (int row = 0; row < 16; ++row) {
        const int m = m0 + 16 * t + row;
        if (m >= row1) break;  // wave-uniform
        const int rid = __builtin_amdgcn_readlane(rvec, 16 * t + row);
        if (rid != cur) {
          if (cur != cut_first && cur != cut_last)
            *reinterpret_cast<float2*>(g.agg + (size_t)cur * 128 + 2 * lane) = run;
          cur = rid;
          run = make_float2(0.0f, 0.0f);
        }
        const float2 v = *reinterpret_cast<const float2*>(&F[row * LDF + 2 * lane]);
        run.x += v.x;
        run.y += v.y;
      }
    }
  }
  if (cur != cut_first && cur != cut_last) *reinterpret_cast<float2*>(g.agg + (size_t)cur * 128 + 2 * lane) = run;
}

// Completes the aggregation of edge_mlp_kernel: receivers without edges get
// zero, receivers whose edges straddle two waves' chunks (chunk_rows edges
// each) are summed here in edge order; the others were written by the kernel.
__global__ void agg_fixup_kernel(const float* __restrict__ E, const bf16* __restrict__ Eb, const int* __restrict__ off,
                                 int nodes, int chunk_rows, float* __restrict__ agg) {
  const int node = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (node >= nodes) return;
  const int b = off[node], e = off[node + 1];
  if (e > b && b / chunk_rows == (e - 1) / chunk_rows) return;
  float s0 = 0.0f, s1 = 0.0f;
  for (int k = b; k < e; ++k) {
    float2 v;
    if (Eb) {
      const unsigned q = *reinterpret_cast<const unsigned*>(Eb + (size_t)k * 128 + 2 * lane);
      v = make_float2(__uint_as_float(q << 16), __uint_as_float(q & 0xffff0000u));
    } else {
      v = *reinterpret_cast<const float2*>(E + (size_t)k * 128 + 2 * lane);
    }
    s0 += v.x;
    s1 += v.y;
  }
  *reinterpret_cast<float2*>(agg + (size_t)node * 128 + 2 * lane) = make_float2(s0, s1);
}

__global__ void transpose_kernel(AOperand A, int R, int C, bf16* __restrict__ out, int ldo) {
  __shared__ float tile[32][33];
  const int r0 = blockIdx.x * 32, c0 = blockIdx.y * 32;
  for (int y = threadIdx.y; y < 32; y += blockDim.y) {
    const int r = r0 + y, c = c0 + threadIdx.x;
    tile[y][threadIdx.x] = (r < R && c < C) ? load_a1(A, r, c) : 0.0f;
  }
  __syncthreads();
  for (int y = threadIdx.y; y < 32; y += blockDim.y) {
    const int c = c0 + y, r = r0 + threadIdx.x;
    if (c < C && r < ldo) out[(size_t)c * ldo + r] = (bf16)tile[threadIdx.x][y];
  }
}

__global__ void reduce_partials_kernel(const float* __restrict__ part, int nparts, int count, float* __restrict__ out,
                                       int accumulate) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  float s = accumulate ? out[i] : 0.0f;
  for (int z = 0; z < nparts; ++z) s += part[(size_t)z * count + i];
  out[i] = s;
}

// partial column sums: block b sums rows b, b + G, b + 2G, ... for all N columns
// Partial column sums (N <= 128, N % 4 == 0): 256 threads = 8 row lanes x 32
// float4 column groups; each thread keeps four independent accumulators so
// several row loads are in flight; the 8 row lanes are combined in LDS.
__global__ __launch_bounds__(256) void column_partial_kernel(const float* __restrict__ X, int M, int N, int ld,
                                                             float* __restrict__ part) {
  __shared__ float4 red[8][32];
  const int cg = threadIdx.x & 31, rl = threadIdx.x >> 5;
  const bool active = 4 * cg < N;
  float4 a[4];
  for (int u = 0; u < 4; ++u) a[u] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  const long long stride = 8LL * gridDim.x;
  long long m = (long long)blockIdx.x * 8 + rl;
  if (active) {
    for (; m + 3 * stride < M; m += 4 * stride) {
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const float4 v = *reinterpret_cast<const float4*>(X + (size_t)(m + u * stride) * ld + 4 * cg);
        a[u].x += v.x;
        a[u].y += v.y;
        a[u].z += v.z;
        a[u].w += v.w;
      }
    }
    for (; m < M; m += stride) {
      const float4 v = *reinterpret_cast<const float4*>(X + (size_t)m * ld + 4 * cg);
      a[0].x += v.x;
      a[0].y += v.y;
      a[0].z += v.z;
      a[0].w += v.w;
    }
  }
  red[rl][cg] = make_float4(a[0].x + a[1].x + a[2].x + a[3].x, a[0].y + a[1].y + a[2].y + a[3].y,
                            a[0].z + a[1].z + a[2].z + a[3].z, a[0].w + a[1].w + a[2].w + a[3].w);
  __syncthreads();
  if (rl == 0 && active) {
    float4 t = red[0][cg];
    for (int r = 1; r < 8; ++r) {
      t.x += red[r][cg].x;
      t.y += red[r][cg].y;
      t.z += red[r][cg].z;
      t.w += red[r][cg].w;
    }
    *reinterpret_cast<float4*>(part + (size_t)blockIdx.x * N + 4 * cg) = t;
  }
}

// one wave per node, two channels per lane (width 128)
__global__ void segment_sum_kernel(const float* __restrict__ E, int ldE, const int* __restrict__ off,
                                   const int* __restrict__ rev, int nodes, float* __restrict__ agg, int ldA,
                                   int accumulate) {
  const int node = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (node >= nodes) return;
  float s0 = 0.0f, s1 = 0.0f;
  const int b = off[node], e = off[node + 1];
  for (int k = b; k < e; ++k) {
    const int row = rev ? rev[k] : k;
    const float2 v = *reinterpret_cast<const float2*>(E + (size_t)row * ldE + 2 * lane);
    s0 += v.x;
    s1 += v.y;
  }
  float2* o = reinterpret_cast<float2*>(agg + (size_t)node * ldA + 2 * lane);
  if (accumulate) {
    const float2 old = *o;
    s0 += old.x;
    s1 += old.y;
  }
  *o = make_float2(s0, s1);
}

__global__ void broadcast_kernel(const float* __restrict__ X, int ldX, const int* __restrict__ recv, int edges,
                                 float* __restrict__ E, int ldE, int accumulate) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int e = (int)(t >> 5), q = (int)(t & 31);
  if (e >= edges) return;
  const float4 v = *reinterpret_cast<const float4*>(X + (size_t)recv[e] * ldX + 4 * q);
  float4* o = reinterpret_cast<float4*>(E + (size_t)e * ldE + 4 * q);
  if (accumulate) {
    const float4 a = *o;
    *o = make_float4(a.x + v.x, a.y + v.y, a.z + v.z, a.w + v.w);
  } else {
    *o = v;
  }
}

__global__ void axpy_kernel(const float* __restrict__ x, float* __restrict__ y, long long count, float a) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (long long)gridDim.x * blockDim.x)
    y[i] += a * x[i];
}

__global__ void fill_kernel(float* __restrict__ p, float v, long long count) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (long long)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void to_bf16_kernel(const float* __restrict__ x, bf16* __restrict__ y, long long count) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (long long)gridDim.x * blockDim.x)
    y[i] = (bf16)x[i];
}

__global__ void pack_kernel(const float* __restrict__ W, int N, int K, bf16* __restrict__ Wb, bf16* __restrict__ Wt) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)N * K) return;
  const int n = (int)(i / K), k = (int)(i % K);
  const bf16 v = (bf16)W[i];
  Wb[i] = v;
  Wt[(size_t)k * N + n] = v;
}

__global__ void adam_kernel(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                            float* __restrict__ v, long long count, float lr, float b1, float b2, float eps, float c1,
                            float c2, float clip_scale) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (long long)gridDim.x * blockDim.x) {
    const float gi = g[i] * clip_scale;
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  }
}

// Sum of squares in a fixed order (block partials, then one block), so the
// gradient-clipping factor -- and with it training -- is reproducible.
constexpr int SQ_BLOCKS = 1024;

__global__ __launch_bounds__(256) void sum_squares_partial_kernel(const float* __restrict__ g, long long count,
                                                                  float* __restrict__ part) {
  __shared__ float red[4];
  float s = 0.0f;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (long long)gridDim.x * blockDim.x)
    s += g[i] * g[i];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

__global__ __launch_bounds__(256) void sum_squares_final_kernel(const float* __restrict__ part, int nparts,
                                                                float* __restrict__ out) {
  __shared__ float red[4];
  float s = 0.0f;
  for (int i = threadIdx.x; i < nparts; i += blockDim.x) s += part[i];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) *out = (red[0] + red[1]) + (red[2] + red[3]);
}

__global__ void mfma_tile_kernel(const bf16* __restrict__ A, const bf16* __restrict__ B, float* __restrict__ D) {
  const int l = threadIdx.x;
  bf16x8 a, b;
  for (int q = 0; q < 8; ++q) {
    a[q] = A[(l % 16) * 32 + 8 * (l / 16) + q];
    b[q] = B[(8 * (l / 16) + q) * 16 + (l % 16)];
  }
  f32x4 d = {0.0f, 0.0f, 0.0f, 0.0f};
  d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, d, 0, 0, 0);
  for (int q = 0; q < 4; ++q) D[(4 * (l / 16) + q) * 16 + (l % 16)] = d[q];
}

unsigned grid_for(long long count, int block = 256, unsigned cap = 8192) {
  long long b = (count + block - 1) / block;
  if (b < 1) b = 1;
  return (unsigned)(b > cap ? cap : b);
}

}  // namespace

hipError_t gemm(const GemmArgs& g, hipStream_t s) {
  if (g.M <= 0) return hipSuccess;
  if (g.N % BN != 0 || g.K % BK != 0 || g.K <= 0) return hipErrorInvalidValue;
  const int ks = g.ksplit > 1 ? g.ksplit : 1;
  hipLaunchKernelGGL(gemm_kernel, dim3(cdiv(g.M, BM), g.N / BN, ks), dim3(256), 0, s, g);
  return hipGetLastError();
}

hipError_t mlp3(const Mlp3Args& g, hipStream_t s) {
  if (g.M <= 0) return hipSuccess;
  if (g.K1 % BK != 0 || g.K1 <= 0 || !g.W1 || !g.W2 || !g.W3 || (!g.out && !g.out_bf16)) return hipErrorInvalidValue;
  if (g.out_bf16 && g.resid) return hipErrorInvalidValue;  // BF16 output only for the encoders
  int width = 0;
  for (int i = 0; i < g.A.nseg; ++i) {
    if (!g.A.seg[i].is_f32 || g.A.seg[i].width % BK != 0) return hipErrorInvalidValue;  // FP32 operands only
    width += g.A.seg[i].width;
  }
  if (width != g.K1) return hipErrorInvalidValue;
  const dim3 grid(cdiv(g.M, BM));
  if ((g.K1 / BK) % g.depth != 0) return hipErrorInvalidValue;
  switch (g.depth) {
    case 1: hipLaunchKernelGGL(mlp3_kernel<1>, grid, dim3(256), 0, s, g); break;
    case 2: hipLaunchKernelGGL(mlp3_kernel<2>, grid, dim3(256), 0, s, g); break;
    case 4: hipLaunchKernelGGL(mlp3_kernel<4>, grid, dim3(256), 0, s, g); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t edge_mlp(const EdgeMlpArgs& g, hipStream_t s) {
  if (g.ne <= 0) {  // no edges: every aggregate is zero
    if (g.nodes > 0 && g.agg && g.off)
      hipLaunchKernelGGL(agg_fixup_kernel, dim3(cdiv(g.nodes, 4)), dim3(256), 0, s, g.E, g.Eb, g.off, g.nodes, 1,
                         g.agg);
    return hipGetLastError();
  }
  if ((!g.E && !g.Eb) || !g.P || !g.recv || !g.send || !g.W1 || !g.W2 || !g.W3 || g.ldw1 < 128 || !g.agg || !g.off ||
      g.nodes <= 0)
    return hipErrorInvalidValue;
  int dev = 0, cus = 0;  // one persistent block per CU of the current device
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
    cus = 256;
  const int tiles = (g.ne + 31) / 32;
  const int blocks = std::min(cus, (tiles + EW - 1) / EW);
  EdgeMlpArgs a = g;
  a.chunk = (tiles + blocks * EW - 1) / (blocks * EW);
  if (g.Eb)
    hipLaunchKernelGGL(edge_mlp_kernel<true>, dim3(blocks), dim3(64 * EW), 0, s, a);
  else
    hipLaunchKernelGGL(edge_mlp_kernel<false>, dim3(blocks), dim3(64 * EW), 0, s, a);
  hipLaunchKernelGGL(agg_fixup_kernel, dim3(cdiv(g.nodes, 4)), dim3(256), 0, s, g.E, g.Eb, g.off, g.nodes,
                     32 * a.chunk, g.agg);
  return hipGetLastError();
}

hipError_t transpose_to_bf16(const AOperand& A, int R, int C, bf16* out, int ldo, hipStream_t s) {
  if (ldo <= 0 || C <= 0 || ldo < R) return ldo < R ? hipErrorInvalidValue : hipSuccess;
  hipLaunchKernelGGL(transpose_kernel, dim3(cdiv(ldo, 32), cdiv(C, 32)), dim3(32, 8), 0, s, A, R, C, out, ldo);
  return hipGetLastError();
}

hipError_t reduce_partials(const float* part, int nparts, int count, float* out, int accumulate, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(reduce_partials_kernel, dim3(cdiv(count, 256)), dim3(256), 0, s, part, nparts, count, out,
                     accumulate);
  return hipGetLastError();
}

hipError_t column_sum(const float* X, int M, int N, int ld, float* colsum, float* scratch, int accumulate,
                      hipStream_t s) {
  if (M <= 0) return hipSuccess;
  const int G = 256;
  if (N > 128 || N % 4 != 0 || ld % 4 != 0) return hipErrorInvalidValue;
  hipLaunchKernelGGL(column_partial_kernel, dim3(G), dim3(256), 0, s, X, M, N, ld, scratch);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return reduce_partials(scratch, G, N, colsum, accumulate, s);
}

hipError_t segment_sum(const float* E, int ldE, const int* off, const int* rev, int nodes, float* agg, int ldA,
                       int accumulate, hipStream_t s) {
  if (nodes <= 0) return hipSuccess;
  hipLaunchKernelGGL(segment_sum_kernel, dim3(cdiv(nodes, 4)), dim3(256), 0, s, E, ldE, off, rev, nodes, agg, ldA,
                     accumulate);
  return hipGetLastError();
}

hipError_t broadcast_to_edges(const float* X, int ldX, const int* recv, int edges, float* E, int ldE, int accumulate,
                              hipStream_t s) {
  if (edges <= 0) return hipSuccess;
  hipLaunchKernelGGL(broadcast_kernel, dim3(cdiv((long long)edges * 32, 256)), dim3(256), 0, s, X, ldX, recv, edges,
                     E, ldE, accumulate);
  return hipGetLastError();
}

hipError_t axpy(const float* x, float* y, long long count, float a, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(axpy_kernel, dim3(grid_for(count)), dim3(256), 0, s, x, y, count, a);
  return hipGetLastError();
}

hipError_t fill_f32(float* p, float v, long long count, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(fill_kernel, dim3(grid_for(count)), dim3(256), 0, s, p, v, count);
  return hipGetLastError();
}

hipError_t f32_to_bf16(const float* x, bf16* y, long long count, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(to_bf16_kernel, dim3(grid_for(count)), dim3(256), 0, s, x, y, count);
  return hipGetLastError();
}

__global__ void pack_projection_kernel(const bf16* __restrict__ W1b, bf16* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;  // < 2 * 128 * 128
  const int h = i >> 14, j = (i >> 7) & 127, k = i & 127;
  out[i] = W1b[(size_t)(16 * (j & 7) + (j >> 3)) * 384 + 128 * (1 + h) + k];
}

hipError_t pack_projection(const bf16* W1b, bf16* out, hipStream_t s) {
  hipLaunchKernelGGL(pack_projection_kernel, dim3(2 * 128 * 128 / 256), dim3(256), 0, s, W1b, out);
  return hipGetLastError();
}

hipError_t pack_weights(const float* W, int N, int K, bf16* Wb, bf16* Wt, hipStream_t s) {
  hipLaunchKernelGGL(pack_kernel, dim3(cdiv((long long)N * K, 256)), dim3(256), 0, s, W, N, K, Wb, Wt);
  return hipGetLastError();
}

hipError_t adam(float* w, const float* g, float* m, float* v, long long count, float lr, float b1, float b2, float eps,
                int t, float clip_scale, hipStream_t s) {
  const float c1 = 1.0f - powf(b1, (float)t), c2 = 1.0f - powf(b2, (float)t);
  hipLaunchKernelGGL(adam_kernel, dim3(grid_for(count)), dim3(256), 0, s, w, g, m, v, count, lr, b1, b2, eps, c1, c2,
                     clip_scale);
  return hipGetLastError();
}

hipError_t sum_squares(const float* g, long long count, float* out, float* scratch, hipStream_t s) {
  hipLaunchKernelGGL(sum_squares_partial_kernel, dim3(SQ_BLOCKS), dim3(256), 0, s, g, count, scratch);
  hipLaunchKernelGGL(sum_squares_final_kernel, dim3(1), dim3(256), 0, s, scratch, SQ_BLOCKS, out);
  return hipGetLastError();
}

hipError_t mfma_tile(const bf16* A, const bf16* B, float* D, hipStream_t s) {
  hipLaunchKernelGGL(mfma_tile_kernel, dim3(1), dim3(64), 0, s, A, B, D);
  return hipGetLastError();
}

}  // namespace gnn
}  // namespace sph
