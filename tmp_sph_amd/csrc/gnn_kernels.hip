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

__device__ __forceinline__ bf16x8 load_seg8(const ASeg& g, int row, int kk) {
  const int r = g.rows ? g.rows[row] : row;
  if (g.is_f32) {
    const float4* p = reinterpret_cast<const float4*>(static_cast<const float*>(g.ptr) + (size_t)r * g.ld + kk);
    return cvt8(p[0], p[1]);
  }
  return *reinterpret_cast<const bf16x8*>(static_cast<const bf16*>(g.ptr) + (size_t)r * g.ld + kk);
}

// 8 consecutive K elements of row `row` of the concatenated operand.
__device__ __forceinline__ bf16x8 load_a8(const AOperand& A, int row, int k, int M) {
  if (row >= M) {
    bf16x8 z;
    for (int q = 0; q < 8; ++q) z[q] = (bf16)0.0f;
    return z;
  }
  const int w0 = A.seg[0].width;
  if (A.nseg == 1 || k < w0) return load_seg8(A.seg[0], row, k);
  const int w1 = A.seg[1].width;
  if (A.nseg == 2 || k < w0 + w1) return load_seg8(A.seg[1], row, k - w0);
  return load_seg8(A.seg[2], row, k - w0 - w1);
}

// Raw (unconverted) 8-element A fragment: two 16-byte words holding 8 FP32
// values, or 8 BF16 values in the first word.  Loading raw and converting one
// k-step later keeps the loads in flight across the MFMAs of the current step.
struct ARaw {
  uint4 a, b;
  int f32;
};

__device__ __forceinline__ ARaw load_raw_seg(const ASeg& g, int row, int kk) {
  ARaw r;
  const int rr = g.rows ? g.rows[row] : row;
  if (g.is_f32) {
    const uint4* p = reinterpret_cast<const uint4*>(static_cast<const float*>(g.ptr) + (size_t)rr * g.ld + kk);
    r.a = p[0];
    r.b = p[1];
    r.f32 = 1;
  } else {
    r.a = *reinterpret_cast<const uint4*>(static_cast<const bf16*>(g.ptr) + (size_t)rr * g.ld + kk);
    r.b = make_uint4(0u, 0u, 0u, 0u);
    r.f32 = 0;
  }
  return r;
}

// (segments picked with constant indices: a pointer into the kernel-argument
// struct would move it to scratch)
__device__ __forceinline__ ARaw load_raw8(const AOperand& A, int row, int k, int M) {
  if (row >= M) {
    ARaw r;
    r.a = make_uint4(0u, 0u, 0u, 0u);
    r.b = r.a;
    r.f32 = 0;
    return r;
  }
  const int w0 = A.seg[0].width;
  if (A.nseg == 1 || k < w0) return load_raw_seg(A.seg[0], row, k);
  const int w1 = A.seg[1].width;
  if (A.nseg == 2 || k < w0 + w1) return load_raw_seg(A.seg[1], row, k - w0);
  return load_raw_seg(A.seg[2], row, k - w0 - w1);
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
  // this thread's share of the B tile: row n of B, 16 elements at 16 * half
  const int bn = tid >> 1, bhalf = tid & 1;
  const bf16* bsrc = g.B + (size_t)(n0 + bn) * g.K + 16 * bhalf;
  uint4 bq0 = make_uint4(0u, 0u, 0u, 0u), bq1 = bq0;
  ARaw ar0{}, ar1{};
  if (kbeg < kend) {
    bq0 = *reinterpret_cast<const uint4*>(bsrc + kbeg);
    bq1 = *reinterpret_cast<const uint4*>(bsrc + kbeg + 8);
    ar0 = load_raw8(g.A, m0 + li, kbeg + 8 * lq, g.M);
    ar1 = load_raw8(g.A, m0 + 16 + li, kbeg + 8 * lq, g.M);
  }
  int buf = 0;
  for (int k0 = kbeg; k0 < kend; k0 += BK) {
    *reinterpret_cast<uint4*>(&ldsB[buf][bn * LDB + 16 * bhalf]) = bq0;
    *reinterpret_cast<uint4*>(&ldsB[buf][bn * LDB + 16 * bhalf + 8]) = bq1;
    bf16x8 af[2];
    af[0] = convert8(ar0);
    af[1] = convert8(ar1);
    __syncthreads();
    if (k0 + BK < kend) {  // next step's operands, in flight during the MFMAs
      bq0 = *reinterpret_cast<const uint4*>(bsrc + k0 + BK);
      bq1 = *reinterpret_cast<const uint4*>(bsrc + k0 + BK + 8);
      ar0 = load_raw8(g.A, m0 + li, k0 + BK + 8 * lq, g.M);
      ar1 = load_raw8(g.A, m0 + 16 + li, k0 + BK + 8 * lq, g.M);
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
#pragma unroll
  for (int t = 0; t < 2; ++t)
#pragma unroll
    for (int c = 0; c < 8; ++c)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + 16 * t + 4 * lq + r;
        const int n = n0 + 16 * c + li;
        if (m >= g.M) continue;
        float v = acc[t][c][r];
        const size_t o = (size_t)m * g.ldo + n;
        if (g.ksplit > 1) {
          g.out_f32[(size_t)z * g.M * g.ldo + o] = v;
          continue;
        }
        if (g.bias) v += g.bias[n];
        if (g.resid) v += g.resid[o];
        if (g.relu) v = fmaxf(v, 0.0f);
        if (g.mask && !((float)g.mask[o] > 0.0f)) v = 0.0f;
        if (g.out_f32) g.out_f32[o] = v;
        if (g.out_bf16) g.out_bf16[o] = (bf16)v;
      }
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
