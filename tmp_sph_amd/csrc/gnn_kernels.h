// Internal interface of the message-passing surrogate's gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>

#include "sph_internal.h"

namespace sph {
namespace gnn {

typedef __bf16 bf16;

// One column block of a GEMM A operand: rows of `width` elements (a multiple
// of 32) read from `ptr` (row stride `ld` elements, FP32 or BF16), optionally
// through a row index (gather).  Up to three blocks are concatenated along K.
struct ASeg {
  const void* ptr;
  const int* rows;
  int ld;
  int width;
  int is_f32;
};

struct AOperand {
  ASeg seg[3];
  int nseg;
};

// C[M x N] = A[M x K] * B[N x K]^T with FP32 accumulation on BF16 MFMA.
// Epilogue, in order: + bias[n], + resid[m][n], ReLU, * (mask[m][n] > 0);
// stored as FP32 and/or BF16 (row stride ldo).  ksplit > 1 splits K into
// chunks (grid z) and writes raw FP32 partials at out_f32 + z * M * ldo.
struct GemmArgs {
  AOperand A;
  const bf16* B;
  const float* bias;
  const float* resid;
  const bf16* mask;
  float* out_f32;
  bf16* out_bf16;
  int M, N, K;
  int ldo;
  int relu;
  int ksplit;
  int ldb;  // row stride of B in elements (0: K)
};

hipError_t gemm(const GemmArgs& g, hipStream_t s);

// Inference form of an edge-update MLP call and the aggregation that follows
// it, persistent (one block of 7 waves per CU, the three weight matrices
// resident in LDS, no barriers after staging them): E[e] += L3(relu(L2(relu(E[e] W1e^T + P[recv[e]][0:128] +
// P[send[e]][128:256] + b1)))), where P (nodes x 256) holds the node
// projections V W1r^T | V W1s^T of the first layer's receiver and sender
// blocks (W1 = [W1e | W1r | W1s], 128 x 384, row stride ldw1) with columns in
// the lane order of pack_projection.  The same products as the concatenated
// first layer, summed in a different order.
struct EdgeMlpArgs {
  float* E;   // FP32 edge latents, or null when Eb holds them
  bf16* Eb;   // BF16 edge latents (inference with BF16 latents)
  int ne;
  const float* P;
  const int* recv;
  const int* send;
  const bf16* W1;
  int ldw1;
  const bf16* W2;
  const bf16* W3;
  const float* b1;
  const float* b2;
  const float* b3;
  // the message aggregation agg[i] = sum over edges [off[i], off[i+1]) of
  // the new E rows (edges ordered by receiver), in edge order, bit-identical
  // to segment_sum
  float* agg;
  const int* off;
  int nodes;
  int chunk;  // tiles per wave (set by edge_mlp)
};

hipError_t edge_mlp(const EdgeMlpArgs& g, hipStream_t s);

// One whole MLP call (three linear layers of width 128, ReLU after the first
// two) in one kernel: out[M x 128] = L3(relu(L2(relu(L1(A))))) + bias3 (+
// resid), the hidden activations kept in LDS as BF16 (and also stored to
// h1 / h2 when those are given, for the backward pass).  Same operand
// rounding and accumulation order as three gemm() calls, so the results are
// bit-identical to them.  W1 is 128 x K1, W2 and W3 are 128 x 128 (BF16, row
// n contiguous in k); resid may alias out (row m is read before it is written).
struct Mlp3Args {
  AOperand A;
  int K1;
  const bf16* W1;
  const bf16* W2;
  const bf16* W3;
  const float* b1;
  const float* b2;
  const float* b3;
  const float* resid;
  float* out;
  int ldo;
  bf16* h1;
  bf16* h2;
  int M;
  int depth;  // k-steps of operand loads in flight (1, 2 or 4; divides K1 / 32)
  bf16* out_bf16;  // when set, the output is stored here as BF16 (row stride ldo) instead of to `out`
};

hipError_t mlp3(const Mlp3Args& g, hipStream_t s);

// out[c * ldo + r] (BF16) = element (r, c) of the R x C matrix given as an A
// operand (the gathered / concatenated input), for r < ldo; columns
// R <= r < ldo are zero (padding of the reduction to a multiple of 32).
hipError_t transpose_to_bf16(const AOperand& A, int R, int C, bf16* out, int ldo, hipStream_t s);
// out[i] = sum_z part[z * count + i] (+ out[i] if accumulate)
hipError_t reduce_partials(const float* part, int nparts, int count, float* out, int accumulate, hipStream_t s);
// colsum[n] (+)= sum_m X[m][n], X is M x N FP32 with row stride ld
hipError_t column_sum(const float* X, int M, int N, int ld, float* colsum, float* scratch, int accumulate,
                      hipStream_t s);
// agg[i][:] = sum over e in [off[i], off[i+1]) of E[e][:] (width 128, FP32); rev != null reads E[rev[e]]
hipError_t segment_sum(const float* E, int ldE, const int* off, const int* rev, int nodes, float* agg, int ldA,
                       int accumulate, hipStream_t s);
// E[e][:] (+)= X[recv[e]][:] for every edge (width 128)
hipError_t broadcast_to_edges(const float* X, int ldX, const int* recv, int edges, float* E, int ldE, int accumulate,
                              hipStream_t s);
// y (+)= x, elementwise over count
hipError_t axpy(const float* x, float* y, long long count, float a, hipStream_t s);
hipError_t fill_f32(float* p, float v, long long count, hipStream_t s);
hipError_t f32_to_bf16(const float* x, bf16* y, long long count, hipStream_t s);
// W[N][K] (FP32) -> Wb[N][K] and Wt[K][N] (BF16)
hipError_t pack_weights(const float* W, int N, int K, bf16* Wb, bf16* Wt, hipStream_t s);
// Receiver and sender blocks of an edge MLP's first layer (W1b: BF16, 128 x
// 384) for the node projections, rows in lane order: out[h][j][k] =
// W1b[16 (j % 8) + j / 8][128 (1 + h) + k], so that projection column
// 8 i + c holds output feature 16 c + i and lane i of edge_mlp_kernel reads
// its eight features as two 16-byte words.
hipError_t pack_projection(const bf16* W1b, bf16* out, hipStream_t s);
// Adam step on count parameters (bias-corrected, step t >= 1)
hipError_t adam(float* w, const float* g, float* m, float* v, long long count, float lr, float b1, float b2,
                float eps, int t, float clip_scale, hipStream_t s);
// sum of squares of g into *out (FP32 atomics; used for gradient-norm clipping)
// *out = sum of g[i]^2 in a fixed order; scratch holds 1024 floats.
hipError_t sum_squares(const float* g, long long count, float* out, float* scratch, hipStream_t s);

// MFMA operand-layout probe: D(16x16) = A(16x32) * B(32x16), one wave.
hipError_t mfma_tile(const bf16* A, const bf16* B, float* D, hipStream_t s);

}  // namespace gnn
}  // namespace sph
