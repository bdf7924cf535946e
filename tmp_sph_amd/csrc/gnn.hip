// Message-passing particle surrogate: model, graph construction, forward,
// backward and Adam, all on the device with the kernels of gnn_kernels.hip.
//
// Graph: nodes are particles in their original order; edges join every pair
// closer than the SPH support (the pair set of the reference's kernels, built
// from the same cell hash + radix sort); edges are stored in CSR by receiver
// with both directions present, and rev[e] is the reverse edge of e, so the
// sender-side gradient is a segment sum too (no atomics, deterministic).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "gnn_kernels.h"
#include "sph_device.h"
#include "sph_kernels.h"

using namespace sph;
using namespace sph::gnn;

namespace {

constexpr int W = 128;  // latent / hidden width

struct Lin {
  int N, K;
  size_t w, b;  // offsets into the FP32 parameter vector
  size_t wb;    // offset into the packed BF16 weight buffers
};

struct Mlp {
  Lin l[3];
  int kin;
};

inline size_t round_up(size_t a, size_t b) { return (a + b - 1) / b * b; }

}  // namespace

struct sph_gnn {
  sph_gnn_cfg cfg{};
  hipStream_t stream = nullptr;
  int L = 10;
  std::vector<void*> allocs;
  // parameters
  std::vector<Lin> lins;
  Mlp enc_n{}, enc_e{}, dec{};
  std::vector<Mlp> pe, pn;
  size_t nparam = 0, nbf = 0;
  float *P = nullptr, *G = nullptr, *M1 = nullptr, *M2 = nullptr;
  bf16 *Wb = nullptr, *Wt = nullptr;
  int adam_t = 0;
  bool packed = false;
  // loaded state / graph
  sph_params prm{};
  int n = 0, ne = 0, nfluid = 0;
  bool has_target = false;
  int *off = nullptr, *recv = nullptr, *send = nullptr, *rev = nullptr;
  int *keys = nullptr, *idx = nullptr, *keys2 = nullptr, *idx2 = nullptr, *inv = nullptr, *deg = nullptr;
  unsigned *hist_sort = nullptr, *totals = nullptr;
  int *cellStart = nullptr, *cellEnd = nullptr;
  int cells_cap = 0;
  int* scan_blocks = nullptr;
  float *pos = nullptr, *vhist = nullptr, *rho = nullptr, *target = nullptr;
  int* type = nullptr;
  // activations
  float *Xn = nullptr, *Xe = nullptr;
  float *V = nullptr, *E = nullptr, *Agg = nullptr, *Y = nullptr;
  bf16* H = nullptr;  // hidden activations of every MLP call (training) or one call (inference)
  // backward scratch
  float *dY = nullptr, *dV = nullptr, *dE = nullptr, *dX = nullptr, *dZf[2] = {nullptr, nullptr};
  bf16 *dZb[2] = {nullptr, nullptr}, *Tz = nullptr, *Tx = nullptr;
  float *part = nullptr, *colpart = nullptr, *scal = nullptr;
  double* lpart = nullptr;  // per-block loss partials
  float* Pn = nullptr;      // inference: node projections of the edge MLPs' first layer (nodes x 256)
  bf16* Wp = nullptr;       // inference: their weights in lane order, per block (pack_projection)
  bf16* Eb = nullptr;       // inference with BF16 edge latents (fused = 9)
  int split_cap = 512;  // K chunks of the weight-gradient GEMMs (enough blocks to fill the GPU)
  // forward MLP calls: fused kernel with this load depth (1, 2, 4), three GEMMs (0); 8: edge MLPs by the
  // persistent kernel with node projections (inference contexts), other calls fused with depth 2;
  // 9: as 8 with the edge latents stored as BF16
  int fused = 2;
};

#define SPH_HIP(call, where)                            \
  do {                                                  \
    hipError_t e_ = (call);                             \
    if (e_ != hipSuccess) return hip_status(e_, where); \
  } while (0)
#define HIPCHK(call)                      \
  do {                                    \
    hipError_t e_ = (call);               \
    if (e_ != hipSuccess) return e_;      \
  } while (0)

namespace {

template <class T>
int galloc(sph_gnn* g, T** p, size_t count) {
  void* q = nullptr;
  SPH_HIP(hipMalloc(&q, (count ? count : 1) * sizeof(T)), "sph_gnn_create: hipMalloc");
  g->allocs.push_back(q);
  *p = (T*)q;
  return SPH_OK;
}

// ------------------------------------------------------------ graph kernels
__global__ void cells_only_kernel(const int* __restrict__ keys, int* __restrict__ cs, int* __restrict__ ce, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int h = keys[i];
  if (i == 0 || h != keys[i - 1]) {
    cs[h] = i;
    if (i > 0) ce[keys[i - 1]] = i;
  }
  if (i == n - 1) ce[h] = n;
}

__global__ void invert_kernel(const int* __restrict__ idx, int* __restrict__ inv, int n) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < n) inv[idx[s]] = s;
}

// Outside the escape box in the x-z plane: the solver parks such a fluid
// particle (below the floor, inside the wall layers) and leaves it there.
__device__ __forceinline__ bool outside_xz(const sph_params& P, float x, float z) {
  return (double)x > P.domain_max[0] || (double)x < P.domain_min[0] || (double)z > P.domain_max[2] ||
         (double)z < P.domain_min[2];
}
// A parked fluid particle: out of the graph (no edges in or out; the parked
// particles share one position next to the wall layers, where they would
// otherwise form one large clique) and out of the loss.
__device__ __forceinline__ bool parked(const sph_params& P, const float* pos, const int* type, int p) {
  return type[p] == 1 && outside_xz(P, pos[3 * p], pos[3 * p + 2]);
}

// Visits the candidate sorted slots j of sorted slot s (2-D: 3 stencil rows).
template <class F>
__device__ __forceinline__ void for_pairs2d(const sph_params& P, int key, const int* __restrict__ cs,
                                            const int* __restrict__ ce, F&& f) {
  const int X = P.gridxdim, Z = P.gridzdim;
  const int cz = key / X, cx = key - cz * X;
  const int x0 = cx > 0 ? cx - 1 : 0, x1 = cx < X - 1 ? cx + 1 : X - 1;
  for (int dz = -1; dz <= 1; ++dz) {
    const int z = cz + dz;
    if (z < 0 || z >= Z) continue;
    int s = -1, e = -1;
    for (int c = z * X + x0; c <= z * X + x1; ++c) {
      const int st = cs[c];
      if (st != -1) {
        if (s < 0) s = st;
        e = ce[c];
      }
    }
    for (int j = (s < 0 ? 0 : s); j < (s < 0 ? 0 : e); ++j) f(j);
  }
}

__global__ void degree_kernel(sph_params P, const float* __restrict__ pos, const int* __restrict__ type,
                              const int* __restrict__ keys, const int* __restrict__ idx, const int* __restrict__ cs,
                              const int* __restrict__ ce, int* __restrict__ deg, int n) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int p = idx[s];
  const float px = pos[3 * p], pz = pos[3 * p + 2];
  int d = 0;
  if (!parked(P, pos, type, p))
  for_pairs2d(P, keys[s], cs, ce, [&](int j) {
    if (j == s) return;
    const int q = idx[j];
    if (parked(P, pos, type, q)) return;
    const float dx = px - pos[3 * q], dz = pz - pos[3 * q + 2];
    if (dx * dx + dz * dz <= P.rr_cut) ++d;
  });
  deg[p] = d;
}

__global__ void fill_edges_kernel(sph_params P, const float* __restrict__ pos, const int* __restrict__ type,
                                  const int* __restrict__ keys, const int* __restrict__ idx,
                                  const int* __restrict__ cs, const int* __restrict__ ce, const int* __restrict__ off,
                                  int* __restrict__ send, int* __restrict__ recv, int n) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int p = idx[s];
  if (parked(P, pos, type, p)) return;
  const float px = pos[3 * p], pz = pos[3 * p + 2];
  int e = off[p];
  for_pairs2d(P, keys[s], cs, ce, [&](int j) {
    if (j == s) return;
    const int q = idx[j];
    if (parked(P, pos, type, q)) return;
    const float dx = px - pos[3 * q], dz = pz - pos[3 * q + 2];
    if (dx * dx + dz * dz <= P.rr_cut) {
      send[e] = q;
      recv[e] = p;
      ++e;
    }
  });
}

// rev[e]: the edge q <- p for e = p <- q; senders of a segment ascend in
// sorted slot, so a binary search on inv[] finds it.
__global__ void reverse_kernel(const int* __restrict__ off, const int* __restrict__ send, const int* __restrict__ recv,
                               const int* __restrict__ inv, int* __restrict__ rev, int ne) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  const int p = recv[e], q = send[e];
  const int target = inv[p];
  const int end = off[q + 1];
  int lo = off[q], hi = end;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (inv[send[mid]] < target)
      lo = mid + 1;
    else
      hi = mid;
  }
  // the pair test is symmetric, so the reverse edge exists; never point outside
  rev[e] = (lo < end && send[lo] == p) ? lo : e;
}

// exclusive scan of deg[0..n) into off[0..n], 1024 elements per block
__global__ void scan_block_kernel(const int* __restrict__ in, int* __restrict__ out, int* __restrict__ bsum, int n) {
  __shared__ int s[1024];
  const int i = blockIdx.x * 1024 + threadIdx.x;
  const int v = i < n ? in[i] : 0;
  s[threadIdx.x] = v;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {
    const int t = threadIdx.x >= o ? s[threadIdx.x - o] : 0;
    __syncthreads();
    s[threadIdx.x] += t;
    __syncthreads();
  }
  if (i < n) out[i] = s[threadIdx.x] - v;
  if (threadIdx.x == 1023) bsum[blockIdx.x] = s[1023];
}

__global__ void scan_sums_kernel(int* __restrict__ bsum, int nb, int* __restrict__ total_out) {
  // one thread: nb is small (n / 1024)
  if (threadIdx.x != 0) return;
  int run = 0;
  for (int b = 0; b < nb; ++b) {
    const int v = bsum[b];
    bsum[b] = run;
    run += v;
  }
  *total_out = run;
}

__global__ void scan_add_kernel(int* __restrict__ out, const int* __restrict__ bsum, int n) {
  const int i = blockIdx.x * 1024 + threadIdx.x;
  if (i < n) out[i] += bsum[blockIdx.x];
}

// ------------------------------------------------------------ features
__global__ void node_features_kernel(sph_params P, sph_gnn_cfg cfg, const float* __restrict__ pos,
                                     const float* __restrict__ vhist, const float* __restrict__ rho,
                                     const int* __restrict__ type, float* __restrict__ Xn, int n) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  float* x = Xn + (size_t)p * W;
  const float inv_v = 1.0f / cfg.v_scale;
  for (int t = 0; t < cfg.history; ++t) {
    x[2 * t] = vhist[((size_t)t * n + p) * 3] * inv_v;
    x[2 * t + 1] = vhist[((size_t)t * n + p) * 3 + 2] * inv_v;
  }
  const int b = 2 * cfg.history;
  const float R = P.kh;
  const float px = pos[3 * p], pz = pos[3 * p + 2];
  x[b] = fminf(fmaxf((px - (float)P.domain_min[0]) / R, 0.0f), 1.0f);
  x[b + 1] = fminf(fmaxf(((float)P.domain_max[0] - px) / R, 0.0f), 1.0f);
  x[b + 2] = fminf(fmaxf((pz - (float)P.domain_min[2]) / R, 0.0f), 1.0f);
  x[b + 3] = type[p] == 1 ? 1.0f : 0.0f;
  x[b + 4] = (rho[p] / P.restDensity - 1.0f) * cfg.rho_scale;
}

__global__ void edge_features_kernel(sph_params P, const float* __restrict__ pos, const int* __restrict__ send,
                                     const int* __restrict__ recv, float* __restrict__ Xe, int ne) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  const int p = recv[e], q = send[e];
  const float R = P.kh;
  const float dx = (pos[3 * p] - pos[3 * q]) / R, dz = (pos[3 * p + 2] - pos[3 * q + 2]) / R;
  float* x = Xe + (size_t)e * W;
  x[0] = dx;
  x[1] = dz;
  x[2] = sqrtf(dx * dx + dz * dz);
}

// loss over fluid nodes: mean of (y - t)^2 over 3 channels; dY = gradient
__global__ void loss_kernel(sph_params P, const float* __restrict__ pos, const float* __restrict__ Y,
                            const float* __restrict__ T, const int* __restrict__ type, float inv_count,
                            float* __restrict__ dY, double* __restrict__ loss, int n) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  double l = 0.0;
  if (p < n) {
    float* d = dY + (size_t)p * W;
    const bool fl = type[p] == 1 && !parked(P, pos, type, p);
    for (int c = 0; c < 3; ++c) {
      const float r = fl ? Y[(size_t)p * W + c] - T[(size_t)p * 3 + c] : 0.0f;
      d[c] = 2.0f * r * inv_count;
      l += (double)r * r;
    }
  }
  // block partial in a fixed order (the sum over blocks is loss_final_kernel's)
  for (int o = 32; o > 0; o >>= 1) l += __shfl_xor(l, o);
  __shared__ double wsum[4];
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = l;
  __syncthreads();
  if (threadIdx.x == 0) loss[blockIdx.x] = ((wsum[0] + wsum[1]) + (wsum[2] + wsum[3])) * inv_count;
}

// *loss = sum of the nb block partials, in a fixed order
__global__ __launch_bounds__(256) void loss_final_kernel(const double* __restrict__ part, int nb,
                                                         double* __restrict__ loss) {
  double l = 0.0;
  for (int i = threadIdx.x; i < nb; i += 256) l += part[i];
  for (int o = 32; o > 0; o >>= 1) l += __shfl_xor(l, o);
  __shared__ double wsum[4];
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = l;
  __syncthreads();
  if (threadIdx.x == 0) *loss = (wsum[0] + wsum[1]) + (wsum[2] + wsum[3]);
}

__global__ void gather_out_kernel(const float* __restrict__ Y, float* __restrict__ out, int n) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  for (int c = 0; c < 3; ++c) out[(size_t)p * 3 + c] = Y[(size_t)p * W + c];
}

inline unsigned nblk(long long n, int b = 256) { return (unsigned)((n + b - 1) / b); }

// ------------------------------------------------------------ model helpers
ASeg segf(const float* p, int ld, int width, const int* rows = nullptr) { return ASeg{p, rows, ld, width, 1}; }
ASeg segb(const bf16* p, int ld, int width, const int* rows = nullptr) { return ASeg{p, rows, ld, width, 0}; }
AOperand op1(ASeg a) {
  AOperand o{};
  o.seg[0] = a;
  o.nseg = 1;
  return o;
}
AOperand op2(ASeg a, ASeg b) {
  AOperand o = op1(a);
  o.seg[1] = b;
  o.nseg = 2;
  return o;
}
AOperand op3(ASeg a, ASeg b, ASeg c) {
  AOperand o = op2(a, b);
  o.seg[2] = c;
  o.nseg = 3;
  return o;
}

Mlp make_mlp(sph_gnn* g, int kin) {
  Mlp m{};
  m.kin = kin;
  const int ks[3] = {kin, W, W};
  for (int i = 0; i < 3; ++i) {
    Lin l{};
    l.N = W;
    l.K = ks[i];
    l.w = g->nparam;
    g->nparam += (size_t)l.N * l.K;
    l.b = g->nparam;
    g->nparam += (size_t)l.N;
    l.wb = g->nbf;
    g->nbf += (size_t)l.N * l.K;
    m.l[i] = l;
    g->lins.push_back(l);
  }
  return m;
}

hipError_t pack_all(sph_gnn* g) {
  for (const Lin& l : g->lins) HIPCHK(pack_weights(g->P + l.w, l.N, l.K, g->Wb + l.wb, g->Wt + l.wb, g->stream));
  if (g->Wp)
    for (int l = 0; l < g->L; ++l)
      HIPCHK(pack_projection(g->Wb + g->pe[l].l[0].wb, g->Wp + (size_t)l * 2 * W * W, g->stream));
  g->packed = true;
  return hipSuccess;
}

// out (f32, M x 128) = MLP(in) (+ resid); h1, h2 receive the hidden activations
hipError_t mlp_forward(sph_gnn* g, const Mlp& m, const AOperand& in, int M, bf16* h1, bf16* h2, float* out,
                       const float* resid, bf16* out_b = nullptr) {
  if (g->fused || out_b) {
    Mlp3Args f{};
    f.A = in;
    f.K1 = m.kin;
    f.W1 = g->Wb + m.l[0].wb;
    f.W2 = g->Wb + m.l[1].wb;
    f.W3 = g->Wb + m.l[2].wb;
    f.b1 = g->P + m.l[0].b;
    f.b2 = g->P + m.l[1].b;
    f.b3 = g->P + m.l[2].b;
    f.resid = resid;
    f.out = out;
    f.ldo = W;
    // the hidden activations are only read back by the backward pass
    f.h1 = g->cfg.train ? h1 : nullptr;
    f.h2 = g->cfg.train ? h2 : nullptr;
    f.M = M;
    f.depth = (g->fused == 0 || g->fused >= 8) ? 2 : g->fused;
    if (out_b) {  // BF16 output (edge encoder of an inference context with BF16 latents)
      f.out = nullptr;
      f.out_bf16 = out_b;
    }
    return mlp3(f, g->stream);
  }
  GemmArgs a{};
  a.A = in;
  a.B = g->Wb + m.l[0].wb;
  a.bias = g->P + m.l[0].b;
  a.out_bf16 = h1;
  a.M = M;
  a.N = W;
  a.K = m.kin;
  a.ldo = W;
  a.relu = 1;
  a.ksplit = 1;
  HIPCHK(gemm(a, g->stream));
  a.A = op1(segb(h1, W, W));
  a.B = g->Wb + m.l[1].wb;
  a.bias = g->P + m.l[1].b;
  a.out_bf16 = h2;
  a.K = W;
  HIPCHK(gemm(a, g->stream));
  a.A = op1(segb(h2, W, W));
  a.B = g->Wb + m.l[2].wb;
  a.bias = g->P + m.l[2].b;
  a.out_bf16 = nullptr;
  a.out_f32 = out;
  a.resid = resid;
  a.relu = 0;
  return gemm(a, g->stream);
}

// dW += dZ^T X and db += colsum(dZ) for one linear layer; dZ is M x 128
// (FP32 for the bias sum, BF16 for the GEMM), X the layer input (width K).
hipError_t weight_grad(sph_gnn* g, const float* dZf, const bf16* dZb, const AOperand& X, int M, const Lin& l) {
  const int Mp = (int)round_up((size_t)M, 32);
  // transposes with the reduction (row) dimension padded with zeros to Mp
  HIPCHK(transpose_to_bf16(op1(segb(dZb, W, W)), M, W, g->Tz, Mp, g->stream));
  HIPCHK(transpose_to_bf16(X, M, l.K, g->Tx, Mp, g->stream));
  int ks = Mp / 1024;
  if (ks < 1) ks = 1;
  if (ks > g->split_cap) ks = g->split_cap;
  GemmArgs a{};
  a.A = op1(segb(g->Tz, Mp, Mp));
  a.B = g->Tx;
  a.out_f32 = g->part;
  a.M = W;
  a.N = l.K;
  a.K = Mp;
  a.ldo = l.K;
  a.ksplit = ks;
  HIPCHK(gemm(a, g->stream));
  if (ks == 1) {
    HIPCHK(axpy(g->part, g->G + l.w, (long long)W * l.K, 1.0f, g->stream));
  } else {
    HIPCHK(reduce_partials(g->part, ks, W * l.K, g->G + l.w, 1, g->stream));
  }
  return column_sum(dZf, M, W, W, g->G + l.b, g->colpart, 1, g->stream);
}

// Backward of one MLP call: dOut (f32, M x 128) -> parameter gradients and,
// when dX is not null, dX (f32, M x kin) = gradient w.r.t. the input.
hipError_t mlp_backward(sph_gnn* g, const Mlp& m, const AOperand& in, int M, const bf16* h1, const bf16* h2,
                        const float* dOut, float* dX) {
  // the FP32 upstream gradient, rounded to BF16 for the transpose
  HIPCHK(f32_to_bf16(dOut, g->dZb[1], (long long)M * W, g->stream));
  HIPCHK(weight_grad(g, dOut, g->dZb[1], op1(segb(h2, W, W)), M, m.l[2]));
  GemmArgs a{};
  a.A = op1(segf(dOut, W, W));
  a.B = g->Wt + m.l[2].wb;
  a.mask = h2;
  a.out_f32 = g->dZf[0];
  a.out_bf16 = g->dZb[0];
  a.M = M;
  a.N = W;
  a.K = W;
  a.ldo = W;
  a.ksplit = 1;
  HIPCHK(gemm(a, g->stream));  // dZ2
  HIPCHK(weight_grad(g, g->dZf[0], g->dZb[0], op1(segb(h1, W, W)), M, m.l[1]));
  a.A = op1(segb(g->dZb[0], W, W));
  a.B = g->Wt + m.l[1].wb;
  a.mask = h1;
  a.out_f32 = g->dZf[1];
  a.out_bf16 = g->dZb[1];
  HIPCHK(gemm(a, g->stream));  // dZ1
  HIPCHK(weight_grad(g, g->dZf[1], g->dZb[1], in, M, m.l[0]));
  if (!dX) return hipSuccess;
  GemmArgs b{};
  b.A = op1(segb(g->dZb[1], W, W));
  b.B = g->Wt + m.l[0].wb;
  b.out_f32 = dX;
  b.M = M;
  b.N = m.kin;
  b.K = W;
  b.ldo = m.kin;
  b.ksplit = 1;
  return gemm(b, g->stream);
}

// Edge update of an inference context in place, E += MLP_e([E, V[recv], V[send]]),
// and the aggregation agg = segment sums of the new E by receiver: the first
// layer's receiver and sender blocks applied to the nodes first (P = V W1r^T |
// V W1s^T), then one persistent kernel per edge set.
hipError_t edge_forward_split(sph_gnn* g, int l, const float* V, float* E, bf16* Eb, float* agg) {
  const Mlp& m = g->pe[l];
  for (int half = 0; half < 2; ++half) {
    GemmArgs a{};
    a.A = op1(segf(V, W, W));
    a.B = g->Wp + ((size_t)2 * l + half) * W * W;  // columns [128, 256) / [256, 384) of W1, rows in lane order
    a.ldb = W;
    a.out_f32 = g->Pn + (size_t)W * half;
    a.M = g->n;
    a.N = W;
    a.K = W;
    a.ldo = 2 * W;
    a.ksplit = 1;
    HIPCHK(gemm(a, g->stream));
  }
  EdgeMlpArgs e{};
  e.E = E;
  e.Eb = Eb;
  e.ne = g->ne;
  e.P = g->Pn;
  e.recv = g->recv;
  e.send = g->send;
  e.W1 = g->Wb + m.l[0].wb;
  e.ldw1 = 3 * W;
  e.W2 = g->Wb + m.l[1].wb;
  e.W3 = g->Wb + m.l[2].wb;
  e.b1 = g->P + m.l[0].b;
  e.b2 = g->P + m.l[1].b;
  e.b3 = g->P + m.l[2].b;
  e.agg = agg;
  e.off = g->off;
  e.nodes = g->n;
  return edge_mlp(e, g->stream);
}

// hidden-activation slots: call c of the forward pass uses H + c * slot
size_t h_slot(const sph_gnn* g) { return (size_t)2 * W * (size_t)std::max(g->cfg.max_nodes, g->cfg.max_edges); }

bf16* h1_of(sph_gnn* g, int call) { return g->H + (g->cfg.train ? (size_t)call : 0) * h_slot(g); }
bf16* h2_of(sph_gnn* g, int call) { return h1_of(g, call) + (size_t)W * std::max(g->cfg.max_nodes, g->cfg.max_edges); }

float* V_of(sph_gnn* g, int l) { return g->V + (g->cfg.train ? (size_t)l : 0) * (size_t)g->cfg.max_nodes * W; }
float* E_of(sph_gnn* g, int l) { return g->E + (g->cfg.train ? (size_t)l : 0) * (size_t)g->cfg.max_edges * W; }
float* A_of(sph_gnn* g, int l) { return g->Agg + (g->cfg.train ? (size_t)l : 0) * (size_t)g->cfg.max_nodes * W; }

// Forward pass over the loaded graph.  With train = 0 the latents are updated
// in place and only one slot of hidden activations is used.
hipError_t forward(sph_gnn* g) {
  const int n = g->n, ne = g->ne, L = g->L;
  const bool tr = g->cfg.train != 0;
  int call = 0;
  HIPCHK(mlp_forward(g, g->enc_n, op1(segf(g->Xn, W, W)), n, h1_of(g, call), h2_of(g, call), V_of(g, 0), nullptr));
  ++call;
  const bool be = g->fused == 9 && !tr;  // BF16 edge latents
  HIPCHK(mlp_forward(g, g->enc_e, op1(segf(g->Xe, W, W)), ne, h1_of(g, call), h2_of(g, call), E_of(g, 0), nullptr,
                     be ? g->Eb : nullptr));
  ++call;
  for (int l = 0; l < L; ++l) {
    float* Vl = V_of(g, l);
    float* El = E_of(g, l);
    float* El1 = E_of(g, tr ? l + 1 : 0);
    float* Vl1 = V_of(g, tr ? l + 1 : 0);
    float* Al = A_of(g, l);
    if (g->fused >= 8 && !tr) {
      HIPCHK(edge_forward_split(g, l, Vl, be ? nullptr : El, be ? g->Eb : nullptr, Al));  // also aggregates
    } else {
      AOperand ein = op3(segf(El, W, W), segf(Vl, W, W, g->recv), segf(Vl, W, W, g->send));
      HIPCHK(mlp_forward(g, g->pe[l], ein, ne, h1_of(g, call), h2_of(g, call), El1, El));
      HIPCHK(segment_sum(El1, W, g->off, nullptr, n, Al, W, 0, g->stream));
    }
    ++call;
    AOperand nin = op2(segf(Vl, W, W), segf(Al, W, W));
    HIPCHK(mlp_forward(g, g->pn[l], nin, n, h1_of(g, call), h2_of(g, call), Vl1, Vl));
    ++call;
  }
  return mlp_forward(g, g->dec, op1(segf(V_of(g, tr ? L : 0), W, W)), n, h1_of(g, call), h2_of(g, call), g->Y,
                     nullptr);
}

hipError_t backward(sph_gnn* g) {
  const int n = g->n, ne = g->ne, L = g->L;
  int call = 2 + 2 * L;  // decoder call index
  // decoder: dV_L = dMLP(dY)
  HIPCHK(mlp_backward(g, g->dec, op1(segf(V_of(g, L), W, W)), n, h1_of(g, call), h2_of(g, call), g->dY, g->dV));
  HIPCHK(fill_f32(g->dE, 0.0f, (long long)ne * W, g->stream));
  for (int l = L - 1; l >= 0; --l) {
    float* Vl = V_of(g, l);
    float* El = E_of(g, l);
    float* Al = A_of(g, l);
    // node update: V_{l+1} = V_l + MLP_n([V_l, A_l]); dV holds dV_{l+1}
    --call;
    HIPCHK(mlp_backward(g, g->pn[l], op2(segf(Vl, W, W), segf(Al, W, W)), n, h1_of(g, call), h2_of(g, call), g->dV,
                        g->dX));
    // dV_l = dV_{l+1} + dX[:, 0:128];  dA_l = dX[:, 128:256] -> dE_{l+1} += dA_l[recv]
    HIPCHK(broadcast_to_edges(g->dX + W, 2 * W, g->recv, ne, g->dE, W, 1, g->stream));
    HIPCHK(hipMemcpy2DAsync(g->dZf[0], W * sizeof(float), g->dX, 2 * W * sizeof(float), W * sizeof(float), n,
                            hipMemcpyDeviceToDevice, g->stream));
    HIPCHK(axpy(g->dZf[0], g->dV, (long long)n * W, 1.0f, g->stream));
    // edge update: E_{l+1} = E_l + MLP_e([E_l, V_l[recv], V_l[send]]); dE holds dE_{l+1}
    --call;
    AOperand ein = op3(segf(El, W, W), segf(Vl, W, W, g->recv), segf(Vl, W, W, g->send));
    HIPCHK(mlp_backward(g, g->pe[l], ein, ne, h1_of(g, call), h2_of(g, call), g->dE, g->dX));
    // dV_l += segsum(dX[:, 128:256]) + segsum over the reverse edges of dX[:, 256:384]
    HIPCHK(segment_sum(g->dX + W, 3 * W, g->off, nullptr, n, g->dV, W, 1, g->stream));
    HIPCHK(segment_sum(g->dX + 2 * W, 3 * W, g->off, g->rev, n, g->dV, W, 1, g->stream));
    // dE_l = dE_{l+1} + dX[:, 0:128]
    HIPCHK(hipMemcpy2DAsync(g->dZf[0], W * sizeof(float), g->dX, 3 * W * sizeof(float), W * sizeof(float), ne,
                            hipMemcpyDeviceToDevice, g->stream));
    HIPCHK(axpy(g->dZf[0], g->dE, (long long)ne * W, 1.0f, g->stream));
  }
  // encoders (inputs are features: no input gradient)
  --call;
  HIPCHK(mlp_backward(g, g->enc_e, op1(segf(g->Xe, W, W)), ne, h1_of(g, call), h2_of(g, call), g->dE, nullptr));
  --call;
  return mlp_backward(g, g->enc_n, op1(segf(g->Xn, W, W)), n, h1_of(g, call), h2_of(g, call), g->dV, nullptr);
}

void release(sph_gnn* g) {
  for (void* p : g->allocs) hipFree(p);
  if (g->stream) hipStreamDestroy(g->stream);
  delete g;
}

// deterministic uniform in [-1, 1) from a 64-bit LCG
struct Lcg {
  unsigned long long s;
  float next() {
    s = s * 6364136223846793005ULL + 1442695040888963407ULL;
    return (float)((double)(s >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0);
  }
};

}  // namespace

extern "C" int sph_gnn_create(const sph_gnn_cfg* cfg, sph_gnn** out) {
  if (!out) return set_error(SPH_E_INVAL, "sph_gnn_create: null output");
  *out = nullptr;
  if (!cfg || cfg->layers < 1 || cfg->layers > 64 || cfg->history < 1 || 2 * cfg->history + 5 > W ||
      cfg->max_nodes <= 0 || cfg->max_edges <= 0 || !(cfg->v_scale > 0) || !(cfg->dv_scale > 0) ||
      !(cfg->drho_scale > 0))
    return set_error(SPH_E_INVAL, "sph_gnn_create: bad configuration");
  sph_gnn* g = new sph_gnn();
  g->cfg = *cfg;
  g->L = cfg->layers;
  g->fused = cfg->train ? 2 : 9;
  g->enc_n = make_mlp(g, W);
  g->enc_e = make_mlp(g, W);
  for (int l = 0; l < g->L; ++l) {
    g->pe.push_back(make_mlp(g, 3 * W));
    g->pn.push_back(make_mlp(g, 2 * W));
  }
  g->dec = make_mlp(g, W);
  int rc = SPH_OK;
  hipError_t he = hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking);
  if (he != hipSuccess) {
    rc = hip_status(he, "hipStreamCreate");
    release(g);
    return rc;
  }
  const size_t Nn = cfg->max_nodes, Ne = cfg->max_edges;
  const size_t Mx = std::max(Nn, Ne);
  const size_t Mp = round_up(Mx, 32);
  const bool tr = cfg->train != 0;
  const int calls = tr ? 3 + 2 * g->L : 1;
  rc = rc ? rc : galloc(g, &g->P, g->nparam);
  rc = rc ? rc : galloc(g, &g->G, g->nparam);
  rc = rc ? rc : galloc(g, &g->M1, g->nparam);
  rc = rc ? rc : galloc(g, &g->M2, g->nparam);
  rc = rc ? rc : galloc(g, &g->Wb, g->nbf);
  rc = rc ? rc : galloc(g, &g->Wt, g->nbf);
  rc = rc ? rc : galloc(g, &g->off, Nn + 1);
  rc = rc ? rc : galloc(g, &g->recv, Ne);
  rc = rc ? rc : galloc(g, &g->send, Ne);
  rc = rc ? rc : galloc(g, &g->rev, Ne);
  rc = rc ? rc : galloc(g, &g->keys, Nn);
  rc = rc ? rc : galloc(g, &g->idx, Nn);
  rc = rc ? rc : galloc(g, &g->keys2, Nn);
  rc = rc ? rc : galloc(g, &g->idx2, Nn);
  rc = rc ? rc : galloc(g, &g->inv, Nn);
  rc = rc ? rc : galloc(g, &g->deg, Nn + 1);
  rc = rc ? rc : galloc(g, &g->hist_sort, (size_t)radix_hist_words((int)Nn));
  rc = rc ? rc : galloc(g, &g->totals, 256);
  rc = rc ? rc : galloc(g, &g->scan_blocks, Nn / 1024 + 2);
  rc = rc ? rc : galloc(g, &g->pos, 3 * Nn);
  rc = rc ? rc : galloc(g, &g->vhist, (size_t)cfg->history * 3 * Nn);
  rc = rc ? rc : galloc(g, &g->rho, Nn);
  rc = rc ? rc : galloc(g, &g->type, Nn);
  rc = rc ? rc : galloc(g, &g->target, 3 * Nn);
  rc = rc ? rc : galloc(g, &g->Xn, Nn * W);
  rc = rc ? rc : galloc(g, &g->Xe, Ne * W);
  rc = rc ? rc : galloc(g, &g->V, (tr ? (size_t)g->L + 1 : 1) * Nn * W);
  rc = rc ? rc : galloc(g, &g->E, (tr ? (size_t)g->L + 1 : 1) * Ne * W);
  rc = rc ? rc : galloc(g, &g->Agg, (tr ? (size_t)g->L : 1) * Nn * W);
  if (!tr) rc = rc ? rc : galloc(g, &g->Pn, Nn * 2 * W);
  if (!tr) rc = rc ? rc : galloc(g, &g->Wp, (size_t)g->L * 2 * W * W);
  if (!tr) rc = rc ? rc : galloc(g, &g->Eb, Ne * W);
  rc = rc ? rc : galloc(g, &g->Y, Nn * W);
  rc = rc ? rc : galloc(g, &g->H, (size_t)calls * h_slot(g));
  rc = rc ? rc : galloc(g, &g->scal, 64);
  rc = rc ? rc : galloc(g, &g->lpart, (size_t)nblk(cfg->max_nodes));
  if (tr) {
    rc = rc ? rc : galloc(g, &g->dY, Nn * W);
    rc = rc ? rc : galloc(g, &g->dV, Nn * W);
    rc = rc ? rc : galloc(g, &g->dE, Ne * W);
    rc = rc ? rc : galloc(g, &g->dX, Mx * 3 * W);
    rc = rc ? rc : galloc(g, &g->dZf[0], Mx * W);
    rc = rc ? rc : galloc(g, &g->dZf[1], Mx * W);
    rc = rc ? rc : galloc(g, &g->dZb[0], Mx * W);
    rc = rc ? rc : galloc(g, &g->dZb[1], Mx * W);
    rc = rc ? rc : galloc(g, &g->Tz, (size_t)W * Mp);
    rc = rc ? rc : galloc(g, &g->Tx, (size_t)3 * W * Mp);
    rc = rc ? rc : galloc(g, &g->part, (size_t)g->split_cap * W * 3 * W);
    rc = rc ? rc : galloc(g, &g->colpart, (size_t)256 * W);
  }
  if (rc) {
    release(g);
    return rc;
  }
  // initialisation: U(-a, a) with a = sqrt(6 / K) for ReLU layers, a tenth of
  // that for the last layer of each MLP; zero biases; decoder rows >= 3 zero
  std::vector<float> host(g->nparam, 0.0f);
  Lcg rng{0x9E3779B97F4A7C15ULL ^ (unsigned long long)cfg->seed};
  for (size_t li = 0; li < g->lins.size(); ++li) {
    const Lin& l = g->lins[li];
    const bool last = (li % 3) == 2;
    const float a = std::sqrt(6.0f / (float)l.K) * (last ? 0.1f : 1.0f);
    for (int r = 0; r < l.N; ++r)
      for (int k = 0; k < l.K; ++k) host[l.w + (size_t)r * l.K + k] = a * rng.next();
  }
  {
    const Lin& l = g->dec.l[2];
    for (int r = 3; r < l.N; ++r)
      for (int k = 0; k < l.K; ++k) host[l.w + (size_t)r * l.K + k] = 0.0f;
  }
  he = hipMemcpy(g->P, host.data(), g->nparam * sizeof(float), hipMemcpyHostToDevice);
  if (he == hipSuccess) he = hipMemset(g->G, 0, g->nparam * sizeof(float));
  if (he == hipSuccess) he = hipMemset(g->M1, 0, g->nparam * sizeof(float));
  if (he == hipSuccess) he = hipMemset(g->M2, 0, g->nparam * sizeof(float));
  if (he == hipSuccess) he = pack_all(g);
  if (he == hipSuccess) he = hipStreamSynchronize(g->stream);
  if (he != hipSuccess) {
    rc = hip_status(he, "sph_gnn_create: init");
    release(g);
    return rc;
  }
  *out = g;
  return SPH_OK;
}

extern "C" int sph_gnn_destroy(sph_gnn* g) {
  if (!g) return SPH_OK;
  hipStreamSynchronize(g->stream);
  release(g);
  return SPH_OK;
}

extern "C" int sph_gnn_num_params(sph_gnn* g, long long* count) {
  if (!g || !count) return set_error(SPH_E_INVAL, "null argument");
  *count = (long long)g->nparam;
  return SPH_OK;
}

extern "C" int sph_gnn_param_layout(sph_gnn* g, int* layout5, int cap, int* nlayers) {
  if (!g || !nlayers) return set_error(SPH_E_INVAL, "null argument");
  *nlayers = (int)g->lins.size();
  if (layout5)
    for (int i = 0; i < *nlayers && i < cap; ++i) {
      const Lin& l = g->lins[i];
      layout5[5 * i] = (int)l.w;
      layout5[5 * i + 1] = l.N;
      layout5[5 * i + 2] = l.K;
      layout5[5 * i + 3] = (int)l.b;
      layout5[5 * i + 4] = 0;
    }
  return SPH_OK;
}

extern "C" int sph_gnn_get_params(sph_gnn* g, float* host) {
  if (!g || !host) return set_error(SPH_E_INVAL, "null argument");
  SPH_HIP(hipStreamSynchronize(g->stream), "sync");
  SPH_HIP(hipMemcpy(host, g->P, g->nparam * sizeof(float), hipMemcpyDeviceToHost), "get params");
  return SPH_OK;
}

extern "C" int sph_gnn_set_params(sph_gnn* g, const float* host) {
  if (!g || !host) return set_error(SPH_E_INVAL, "null argument");
  SPH_HIP(hipStreamSynchronize(g->stream), "sync");
  SPH_HIP(hipMemcpy(g->P, host, g->nparam * sizeof(float), hipMemcpyHostToDevice), "set params");
  SPH_HIP(pack_all(g), "pack");
  SPH_HIP(hipStreamSynchronize(g->stream), "sync");
  return SPH_OK;
}

extern "C" int sph_gnn_get_grads(sph_gnn* g, float* host) {
  if (!g || !host) return set_error(SPH_E_INVAL, "null argument");
  SPH_HIP(hipStreamSynchronize(g->stream), "sync");
  SPH_HIP(hipMemcpy(host, g->G, g->nparam * sizeof(float), hipMemcpyDeviceToHost), "get grads");
  return SPH_OK;
}

// Radius graph and input features of the state held in g->pos / vhist / rho /
// type (device, original order): hash, radix sort, cell ranges, degrees,
// scan, edges, reverse edges.  One host synchronisation (the edge count).
// Graph and features of the loaded state.  *overflow (when given) reports a
// graph with more edges than max_edges (nothing built) instead of an error.
static int build_graph_features(sph_gnn* g, const sph_params& p, int n, bool* overflow = nullptr) {
  hipStream_t s = g->stream;
  g->prm = p;
  g->n = n;
  const size_t N = (size_t)n;
  if (p.numCells > g->cells_cap) {
    if (g->cellStart) {
      hipFree(g->cellStart);
      hipFree(g->cellEnd);
      g->allocs.erase(std::remove(g->allocs.begin(), g->allocs.end(), (void*)g->cellStart), g->allocs.end());
      g->allocs.erase(std::remove(g->allocs.begin(), g->allocs.end(), (void*)g->cellEnd), g->allocs.end());
    }
    int rc = galloc(g, &g->cellStart, (size_t)p.numCells);
    if (rc) return rc;
    rc = galloc(g, &g->cellEnd, (size_t)p.numCells);
    if (rc) return rc;
    g->cells_cap = p.numCells;
  }
  SPH_HIP(launch_hash_aos(p, g->pos, g->keys, g->idx, n, s), "graph hash");
  bool alt = false;
  SPH_HIP(radix_sort_pairs(g->keys, g->idx, g->keys2, g->idx2, g->hist_sort, g->totals, n, key_bits(p.numCells), s,
                           &alt),
          "graph sort");
  int* keys = alt ? g->keys2 : g->keys;
  int* idx = alt ? g->idx2 : g->idx;
  SPH_HIP(fill_int(g->cellStart, -1, p.numCells, s), "graph cells");
  hipLaunchKernelGGL(cells_only_kernel, dim3(nblk(n)), dim3(256), 0, s, keys, g->cellStart, g->cellEnd, n);
  hipLaunchKernelGGL(invert_kernel, dim3(nblk(n)), dim3(256), 0, s, idx, g->inv, n);
  hipLaunchKernelGGL(degree_kernel, dim3(nblk(n)), dim3(256), 0, s, p, g->pos, g->type, keys, idx, g->cellStart,
                     g->cellEnd, g->deg, n);
  const int nb = (n + 1023) / 1024;
  hipLaunchKernelGGL(scan_block_kernel, dim3(nb), dim3(1024), 0, s, g->deg, g->off, g->scan_blocks, n);
  hipLaunchKernelGGL(scan_sums_kernel, dim3(1), dim3(64), 0, s, g->scan_blocks, nb, g->off + n);
  hipLaunchKernelGGL(scan_add_kernel, dim3(nb), dim3(1024), 0, s, g->off, g->scan_blocks, n);
  SPH_HIP(hipGetLastError(), "graph scan");
  int ne = 0;
  SPH_HIP(hipMemcpyAsync(&ne, g->off + n, sizeof(int), hipMemcpyDeviceToHost, s), "graph edges");
  SPH_HIP(hipStreamSynchronize(s), "graph sync");
  if (overflow) *overflow = ne > g->cfg.max_edges;
  if (ne > g->cfg.max_edges)
    return overflow ? SPH_OK : set_error(SPH_E_INVAL, "surrogate graph: more edges than max_edges");
  g->ne = ne;
  hipLaunchKernelGGL(fill_edges_kernel, dim3(nblk(n)), dim3(256), 0, s, p, g->pos, g->type, keys, idx,
                     g->cellStart, g->cellEnd, g->off, g->send, g->recv, n);
  if (ne > 0)
    hipLaunchKernelGGL(reverse_kernel, dim3(nblk(ne)), dim3(256), 0, s, g->off, g->send, g->recv, g->inv, g->rev, ne);
  SPH_HIP(fill_f32(g->Xn, 0.0f, (long long)N * W, s), "features");
  SPH_HIP(fill_f32(g->Xe, 0.0f, (long long)ne * W, s), "features");
  hipLaunchKernelGGL(node_features_kernel, dim3(nblk(n)), dim3(256), 0, s, p, g->cfg, g->pos, g->vhist, g->rho,
                     g->type, g->Xn, n);
  if (ne > 0)
    hipLaunchKernelGGL(edge_features_kernel, dim3(nblk(ne)), dim3(256), 0, s, p, g->pos, g->send, g->recv, g->Xe, ne);
  SPH_HIP(hipGetLastError(), "features");
  return SPH_OK;
}

// ---------------------------------------------------------------- rollout
namespace {

// Applies the prediction (fluid only): with T = k dt and the predicted rates,
// v' = v + T dv/dt, x' = x + T (v + v')/2, rho' = max(rho0, rho + T drho/dt),
// p' = EOS(rho'); walls keep their state.

__global__ void apply_kernel(sph_params P, sph_gnn_cfg cfg, int k, const float* __restrict__ Y,
                             const float* __restrict__ pos, const float* __restrict__ vel,
                             const float* __restrict__ rho, const int* __restrict__ type, float* __restrict__ pos2,
                             float* __restrict__ vel2, float* __restrict__ rho2, float* __restrict__ p2, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float x = pos[3 * i], z = pos[3 * i + 2];
  const float vx = vel[3 * i], vz = vel[3 * i + 2];
  float nx = x, nz = z, nvx = vx, nvz = vz, nr = rho[i];
  if (type[i] == 1 && !outside_xz(P, x, z)) {  // parked particles stay where they are
    const float* yp = Y + (size_t)i * W;
    float y[3] = {yp[0], yp[1], yp[2]};
    if (cfg.out_clip > 0.0f)
      for (int c = 0; c < 3; ++c) y[c] = fminf(fmaxf(y[c], -cfg.out_clip), cfg.out_clip);
    const float T = (float)k * P.timeStep;
    nvx = vx + (y[0] * cfg.dv_scale + cfg.out_mean[0]) * T;
    nvz = vz + (y[1] * cfg.dv_scale + cfg.out_mean[1]) * T;
    nr = fmaxf(rho[i] + (y[2] * cfg.drho_scale + cfg.out_mean[2]) * T, P.restDensity);
    nx = x + 0.5f * T * (vx + nvx);
    nz = z + 0.5f * T * (vz + nvz);
  }
  pos2[3 * i] = nx;
  pos2[3 * i + 1] = 0.0f;
  pos2[3 * i + 2] = nz;
  vel2[3 * i] = nvx;
  vel2[3 * i + 1] = 0.0f;
  vel2[3 * i + 2] = nvz;
  rho2[i] = nr;
  p2[i] = eos_pressure(P, nr);
}

// Judge of one interval S -> S' of length T, by the plain method's rates at
// both ends (kernel B's acceleration with gravity, a0 at S and a1 at S', and
// the continuity rate d0, d1; original particle order).  Over the fluid
// particles inside the escape box at the start (parked ones are left out):
//   density: |rho' - max(rho0, rho + T (d0 + d1) / 2)| / rho0, the continuity
//     integral by the trapezoid rule, clamped at rest density like the solver;
//   velocity: r = v' - v - T (a0 + a1) / 2, summed over all fluid (a bias of
//     the whole flow) and over the front band x >= xband (the surge front);
//   penetration: fluid predicted outside the escape box, and its depth.
// Per block (JW floats): [0] sum of density deviations, [1] count, [2] max
// deviation, [3] escaped, [4] deepest position beyond the box, [5] band count,
// [6] [7] band sums of r_x, r_z, [8] [9] sums of r_x, r_z.  Sums run in a
// fixed order (butterfly in the wave, then the block's waves), so the verdict
// does not depend on scheduling.
constexpr int JW = 10;
__device__ __forceinline__ bool judge_is_max(int c) { return c == 2 || c == 4; }

__global__ void __launch_bounds__(256) judge_kernel(sph_params P, float T, float xband,
                                                    const float* __restrict__ pos, const float* __restrict__ vel,
                                                    const float* __restrict__ rho, const float* __restrict__ pos2,
                                                    const float* __restrict__ vel2, const float* __restrict__ rho2,
                                                    const int* __restrict__ type, const float* __restrict__ a0,
                                                    const float* __restrict__ d0, const float* __restrict__ a1,
                                                    const float* __restrict__ d1, float* __restrict__ part, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  float v[JW];
  for (int c = 0; c < JW; ++c) v[c] = 0.0f;
  if (i < n && type[i] == 1 && !outside_xz(P, pos[3 * i], pos[3 * i + 2])) {
    const float x2 = pos2[3 * i], z2 = pos2[3 * i + 2];
    const double d = fmax(fmax(P.domain_min[0] - (double)x2, (double)x2 - P.domain_max[0]),
                          fmax(fmax(P.domain_min[2] - (double)z2, (double)z2 - P.domain_max[2]), 0.0));
    const float re = fmaxf(rho[i] + 0.5f * T * (d0[i] + d1[i]), P.restDensity);
    const float dev = fabsf(rho2[i] - re) / P.restDensity;
    const float rx = vel2[3 * i] - vel[3 * i] - 0.5f * T * (a0[3 * i] + a1[3 * i]);
    const float rz = vel2[3 * i + 2] - vel[3 * i + 2] - 0.5f * T * (a0[3 * i + 2] + a1[3 * i + 2]);
    v[0] = dev;
    v[1] = 1.0f;
    v[2] = dev != dev ? __builtin_huge_valf() : dev;  // a NaN prediction fails the density test
    v[3] = (outside_xz(P, x2, z2) || d != d) ? 1.0f : 0.0f;
    v[4] = (float)d;
    if (pos[3 * i] >= xband) {
      v[5] = 1.0f;
      v[6] = rx;
      v[7] = rz;
    }
    v[8] = rx;
    v[9] = rz;
  }
  for (int c = 0; c < JW; ++c) {
    float x = v[c];
    for (int o = 32; o > 0; o >>= 1) {
      const float y = __shfl_xor(x, o);
      x = judge_is_max(c) ? fmaxf(x, y) : x + y;
    }
    v[c] = x;
  }
  __shared__ float red[4][JW];
  const int wv = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
    for (int c = 0; c < JW; ++c) red[wv][c] = v[c];
  __syncthreads();
  if (threadIdx.x < JW) {
    const int c = threadIdx.x;
    float x = red[0][c];
    for (int w = 1; w < 4; ++w) x = judge_is_max(c) ? fmaxf(x, red[w][c]) : x + red[w][c];
    part[(size_t)blockIdx.x * JW + c] = x;
  }
}

// Sums (doubles) and maxima of the judge's per-block partials, in a fixed
// order: thread t takes blocks t, t + 256, ..., then a tree over the threads.
__global__ void __launch_bounds__(256) judge_reduce_kernel(const float* __restrict__ part, int nb,
                                                           double* __restrict__ out) {
  __shared__ double red[256][JW];
  const int t = threadIdx.x;
  double v[JW];
  for (int c = 0; c < JW; ++c) v[c] = 0.0;
  for (int b = t; b < nb; b += 256)
    for (int c = 0; c < JW; ++c) {
      const double x = part[(size_t)b * JW + c];
      v[c] = judge_is_max(c) ? fmax(v[c], x) : v[c] + x;
    }
  for (int c = 0; c < JW; ++c) red[t][c] = v[c];
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s)
      for (int c = 0; c < JW; ++c)
        red[t][c] = judge_is_max(c) ? fmax(red[t][c], red[t + s][c]) : red[t][c] + red[t + s][c];
    __syncthreads();
  }
  if (t < JW) out[t] = red[0][t];
}

// Launches the judge over one interval and reads its totals into J (waits).
hipError_t judge_read(const sph_params& P, float T, float xband, const float* pos, const float* vel, const float* rho,
                      const float* pos2, const float* vel2, const float* rho2, const int* type, const float* a0,
                      const float* d0, const float* a1, const float* d1, float* part, double* jout, int n,
                      hipStream_t s, double* J) {
  const unsigned nb = nblk(n);
  hipLaunchKernelGGL(judge_kernel, dim3(nb), dim3(256), 0, s, P, T, xband, pos, vel, rho, pos2, vel2, rho2, type,
                     a0, d0, a1, d1, part, n);
  hipLaunchKernelGGL(judge_reduce_kernel, dim3(1), dim3(256), 0, s, part, (int)nb, jout);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(J, jout, JW * sizeof(double), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  return e;
}

// Velocity residuals in units of g T: the front band's mean and the whole
// flow's mean, the larger component of each.
double judge_band_dev(const double* J, double gT) {
  return J[5] > 0 ? std::max(std::fabs(J[6]), std::fabs(J[7])) / J[5] / gT : 0.0;
}
double judge_bias_dev(const double* J, double gT) {
  return J[1] > 0 ? std::max(std::fabs(J[8]), std::fabs(J[9])) / J[1] / gT : 0.0;
}

}  // namespace

extern "C" int sph_gnn_rollout(sph_gnn* g, sph_sim* sim, sph_sim* judge_sim, int k, int intervals, float dev_tol,
                               float acc_tol, float drift_tol, double front_band, double band, int max_skip,
                               double* front, double* height, sph_rollout_stats* stats, sph_progress_fn progress,
                               void* user) {
  if (!g || !sim || !judge_sim || k <= 0 || intervals < 0 || !(front_band >= 0.0) || max_skip < 0)
    return set_error(SPH_E_INVAL, "sph_gnn_rollout: bad argument");
  int K = 0, KJ = 0;
  int rc = sph_sim_num_instances(sim, &K);
  if (!rc) rc = sph_sim_num_instances(judge_sim, &KJ);
  if (rc) return rc;
  if (K != 1 || KJ != 1) return set_error(SPH_E_INVAL, "sph_gnn_rollout: single-instance contexts only");
  sph_params P, PJ;
  rc = sph_sim_instance_params(sim, 0, &P);
  if (!rc) rc = sph_sim_instance_params(judge_sim, 0, &PJ);
  if (rc) return rc;
  if (P.dim != 2) return set_error(SPH_E_INVAL, "sph_gnn_rollout: the surrogate is 2-D");
  if (PJ.numParticles != P.numParticles || PJ.timeStep != P.timeStep || PJ.h != P.h || PJ.dim != P.dim)
    return set_error(SPH_E_INVAL, "sph_gnn_rollout: the judge context must hold the same case");
  const int n = P.numParticles;
  if (n > g->cfg.max_nodes) return set_error(SPH_E_INVAL, "sph_gnn_rollout: more particles than max_nodes");
  hipStream_t s = g->stream;
  const size_t N = (size_t)n;
  const int H = g->cfg.history;
  const unsigned nb = nblk(n);
  // state (pos/rho live in g), the next state (prediction or plain), plain
  // rates at both, judge partials and totals
  float *vel = nullptr, *prs = nullptr, *pos2 = nullptr, *vel2 = nullptr, *rho2 = nullptr, *p2 = nullptr;
  float *a0 = nullptr, *d0 = nullptr, *a1 = nullptr, *d1 = nullptr, *part = nullptr;
  double* jout = nullptr;
  hipError_t he = hipMalloc(&vel, 3 * N * sizeof(float));
  if (he == hipSuccess) he = hipMalloc(&prs, N * sizeof(float));
  if (he == hipSuccess) he = hipMalloc(&pos2, 3 * N * sizeof(float));
  if (he == hipSuccess) he = hipMalloc(&vel2, 3 * N * sizeof(float));
  if (he == hipSuccess) he = hipMalloc(&rho2, N * sizeof(float));
  if (he == hipSuccess) he = hipMalloc(&p2, N * sizeof(float));
  if (he == hipSuccess) he = hipMalloc(&a0, 3 * N * sizeof(float));
  if (he == hipSuccess) he = hipMalloc(&a1, 3 * N * sizeof(float));
  if (he == hipSuccess) he = hipMalloc(&d0, N * sizeof(float));
  if (he == hipSuccess) he = hipMalloc(&d1, N * sizeof(float));
  if (he == hipSuccess) he = hipMalloc(&part, (size_t)nb * JW * sizeof(float));
  if (he == hipSuccess) he = hipMalloc(&jout, JW * sizeof(double));
  auto release = [&]() {
    hipFree(vel); hipFree(prs); hipFree(pos2); hipFree(vel2); hipFree(rho2); hipFree(p2);
    hipFree(a0); hipFree(a1); hipFree(d0); hipFree(d1); hipFree(part); hipFree(jout);
  };
  if (he != hipSuccess) {
    release();
    return hip_status(he, "sph_gnn_rollout: scratch");
  }
  const float T = (float)k * P.timeStep;
  const double gT = std::fabs((double)P.gravity) * (double)T;
  sph_rollout_stats st{};
  double dev_sum = 0.0;
  int judged = 0, filled = 0;
  auto t0 = std::chrono::steady_clock::now();
  auto push_history = [&](const float* v) -> hipError_t {
    for (int t = H - 1; t > 0; --t)
      HIPCHK(hipMemcpyAsync(g->vhist + (size_t)t * 3 * N, g->vhist + (size_t)(t - 1) * 3 * N, 3 * N * sizeof(float),
                            hipMemcpyDeviceToDevice, s));
    HIPCHK(hipMemcpyAsync(g->vhist, v, 3 * N * sizeof(float), hipMemcpyDeviceToDevice, s));
    return hipStreamSynchronize(s);  // `v` is rewritten by the context's stream next
  };
  // plain rates of a state (device buffers, written on g's stream) from the
  // judge context; positions enter it rounded to FP32, as the network sees them
  auto plain_rates = [&](const float* p, const float* v, const float* r, const float* pr, float* acc,
                         float* drho) -> int {
    hipError_t e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_status(e, "rollout judge state");
    int r2 = sph_sim_set_state_device(judge_sim, p, v, r, pr, g->type);
    if (!r2) r2 = sph_sim_eval_device(judge_sim, drho, acc);
    if (!r2) r2 = sph_sim_synchronize(judge_sim);
    return r2;
  };
  // the judge's reading of the interval from the current state to (p, v, r)
  double J[JW];
  auto read_judge = [&](float xband, const float* p, const float* v, const float* r) -> int {
    const hipError_t e = judge_read(P, T, xband, g->pos, vel, g->rho, p, v, r, g->type, a0, d0, a1, d1, part, jout,
                                    n, s, J);
    return e == hipSuccess ? SPH_OK : hip_status(e, "rollout judge");
  };
  auto band_dev = [&]() { return judge_band_dev(J, gT); };
  auto bias_dev = [&]() { return judge_bias_dev(J, gT); };
  // velocity drift the accepted network steps have left in the front band
  // (sums of its signed mean residuals, x and z, units of g T); plain
  // intervals add none
  double drift[2] = {0.0, 0.0};
  // plain intervals replay the captured step (bit-identical to eager steps)
  const bool captured = sph_sim_capture(sim) == SPH_OK;
  double f_now = 0.0, h_now = 0.0;
  rc = sph_sim_get_state_device(sim, g->pos, vel, g->rho, prs, g->type);
  if (!rc) rc = sph_sim_probe(sim, band, &f_now, &h_now);
  if (!rc) rc = plain_rates(g->pos, vel, g->rho, prs, a0, d0);
  if (!rc && (he = push_history(vel)) != hipSuccess) rc = hip_status(he, "rollout history");
  filled = 1;
  // back-off after rejections: `wait` intervals go to the plain step without
  // a network attempt (and without the judge's reading); the next rejection
  // doubles the wait up to max_skip, an acceptance clears it
  int wait = 0, next_wait = 1;
  bool rates_now = true;  // a0/d0 hold the plain rates of the current state
  for (int it = 0; it < intervals && !rc; ++it) {
    const float xband = (float)(f_now - front_band);
    bool plain = filled < H;
    bool skipped = false;
    if (!plain && wait > 0) {
      --wait;
      plain = skipped = true;
      ++st.skipped;
    }
    if (!plain && !rates_now) {
      if ((rc = plain_rates(g->pos, vel, g->rho, prs, a0, d0))) break;
      rates_now = true;
    }
    if (!plain) {
      bool overflow = false;
      rc = build_graph_features(g, P, n, &overflow);
      if (rc) break;
      if (overflow) {  // a state whose graph does not fit: the plain step takes the interval
        plain = true;
        ++st.fallbacks;
        ++st.rejected_graph;
      }
    }
    if (!plain) {
      if ((he = forward(g)) != hipSuccess) { rc = hip_status(he, "rollout forward"); break; }
      hipLaunchKernelGGL(apply_kernel, dim3(nb), dim3(256), 0, s, P, g->cfg, k, g->Y, g->pos, vel, g->rho,
                         g->type, pos2, vel2, rho2, p2, n);
      if ((rc = plain_rates(pos2, vel2, rho2, p2, a1, d1))) break;
      if ((rc = read_judge(xband, pos2, vel2, rho2))) break;
      // the largest per-particle density deviation decides (a mean would hide
      // a local compression); the mean is kept for the statistics
      const double dev = J[1] > 0 ? J[2] : 0.0;
      const double bd = band_dev(), bs = bias_dev();
      dev_sum += J[1] > 0 ? J[0] / J[1] : 0.0;
      ++judged;
      st.max_dev = std::max(st.max_dev, dev);
      st.max_band_dev = std::max(st.max_band_dev, bd);
      st.max_bias_dev = std::max(st.max_bias_dev, bs);
      st.max_penetration = std::max(st.max_penetration, J[4]);
      const double rb[2] = {J[5] > 0 ? J[6] / J[5] / gT : 0.0, J[5] > 0 ? J[7] / J[5] / gT : 0.0};
      const bool ok_drift = std::fabs(drift[0] + rb[0]) <= drift_tol && std::fabs(drift[1] + rb[1]) <= drift_tol;
      const bool ok_pen = J[3] == 0.0, ok_rho = dev <= dev_tol, ok_vel = bd <= acc_tol && bs <= acc_tol && ok_drift;
      if (!ok_pen) ++st.rejected_penetration;
      if (!ok_rho) ++st.rejected_density;
      if (!ok_vel) ++st.rejected_velocity;
      if (ok_pen && ok_rho && ok_vel) {
        rc = sph_sim_set_state_device(sim, pos2, vel2, rho2, p2, g->type);
        if (rc) break;
        ++st.accepted;
        drift[0] += rb[0];
        drift[1] += rb[1];
        st.max_drift = std::max({st.max_drift, std::fabs(drift[0]), std::fabs(drift[1])});
        next_wait = 1;
      } else {
        plain = true;
        ++st.fallbacks;
        if (max_skip > 0) {
          wait = std::min(next_wait, max_skip);
          next_wait = std::min(2 * next_wait, max_skip);
        }
      }
    }
    if (plain) {
      rc = captured ? sph_sim_replay(sim, k) : sph_sim_step(sim, k);
      if (!rc) rc = sph_sim_get_state_device(sim, pos2, vel2, rho2, p2, nullptr);
      if (!rc && !skipped) {
        // the judge's reading of a plain interval; its end-state rates are the
        // next interval's start-state rates
        rc = plain_rates(pos2, vel2, rho2, p2, a1, d1);
        if (!rc) rc = read_judge(xband, pos2, vel2, rho2);
        if (!rc) {
          st.plain_max_dev = std::max(st.plain_max_dev, J[1] > 0 ? J[2] : 0.0);
          st.plain_max_band_dev = std::max(st.plain_max_band_dev, band_dev());
          st.plain_max_bias_dev = std::max(st.plain_max_bias_dev, bias_dev());
        }
      }
      if (rc) break;
      st.plain_steps += k;
    }
    // the new state (accepted prediction or plain result) becomes the current one
    he = hipMemcpyAsync(g->pos, pos2, 3 * N * sizeof(float), hipMemcpyDeviceToDevice, s);
    if (he == hipSuccess) he = hipMemcpyAsync(vel, vel2, 3 * N * sizeof(float), hipMemcpyDeviceToDevice, s);
    if (he == hipSuccess) he = hipMemcpyAsync(g->rho, rho2, N * sizeof(float), hipMemcpyDeviceToDevice, s);
    if (he == hipSuccess) he = hipMemcpyAsync(prs, p2, N * sizeof(float), hipMemcpyDeviceToDevice, s);
    if (he != hipSuccess) { rc = hip_status(he, "rollout state"); break; }
    std::swap(a0, a1);
    std::swap(d0, d1);
    rates_now = !skipped;  // a skipped interval computed no end-state rates
    if ((he = push_history(vel)) != hipSuccess) { rc = hip_status(he, "rollout history"); break; }
    filled = std::min(filled + 1, H);
    rc = sph_sim_probe(sim, band, &f_now, &h_now);
    if (rc) break;
    if (front) front[it] = f_now;
    if (height) height[it] = h_now;
    if (progress && (it + 1) % 1000 == 0) progress(it + 1, st.accepted, user);
  }
  hipStreamSynchronize(s);
  st.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  st.mean_dev = judged ? dev_sum / judged : 0.0;
  if (stats) *stats = st;
  release();
  return rc;
}

extern "C" int sph_gnn_judge(sph_sim* judge_sim, int k, double front_band, const float* pos, const float* vel,
                             const float* rho, const float* p, const int* type, const float* pos2,
                             const float* vel2, const float* rho2, const float* p2, double* out) {
  if (!judge_sim || k <= 0 || !pos || !vel || !rho || !p || !type || !pos2 || !vel2 || !rho2 || !p2 || !out ||
      !(front_band >= 0.0))
    return set_error(SPH_E_INVAL, "sph_gnn_judge: bad argument");
  int K = 0;
  int rc = sph_sim_num_instances(judge_sim, &K);
  if (rc) return rc;
  if (K != 1) return set_error(SPH_E_INVAL, "sph_gnn_judge: single-instance contexts only");
  sph_params P;
  rc = sph_sim_instance_params(judge_sim, 0, &P);
  if (rc) return rc;
  if (P.dim != 2) return set_error(SPH_E_INVAL, "sph_gnn_judge: the surrogate is 2-D");
  const int n = P.numParticles;
  const size_t N = (size_t)n;
  // front at the start: max x of fluid inside the escape box
  double front = 0.0;
  bool any = false;
  for (size_t i = 0; i < N; ++i) {
    const double x = pos[3 * i], z = pos[3 * i + 2];
    if (type[i] != 1 || x > P.domain_max[0] || x < P.domain_min[0] || z > P.domain_max[2] || z < P.domain_min[2])
      continue;
    front = any ? std::max(front, x) : x;
    any = true;
  }
  const size_t f3 = 3 * N * sizeof(float), f1 = N * sizeof(float);
  std::vector<void*> bufs;
  auto dev = [&](size_t bytes) -> void* {
    void* q = nullptr;
    if (hipMalloc(&q, bytes) != hipSuccess) return nullptr;
    bufs.push_back(q);
    return q;
  };
  float *dpos = (float*)dev(f3), *dvel = (float*)dev(f3), *drho = (float*)dev(f1), *dp = (float*)dev(f1);
  float *dpos2 = (float*)dev(f3), *dvel2 = (float*)dev(f3), *drho2 = (float*)dev(f1), *dp2 = (float*)dev(f1);
  float *a0 = (float*)dev(f3), *a1 = (float*)dev(f3), *d0 = (float*)dev(f1), *d1 = (float*)dev(f1);
  int* dtype = (int*)dev(N * sizeof(int));
  float* part = (float*)dev((size_t)nblk(n) * JW * sizeof(float));
  double* jout = (double*)dev(JW * sizeof(double));
  hipError_t e = bufs.size() == 15 ? hipSuccess : hipErrorOutOfMemory;
  if (e == hipSuccess) e = hipMemcpy(dpos, pos, f3, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dvel, vel, f3, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(drho, rho, f1, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dp, p, f1, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dtype, type, N * sizeof(int), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dpos2, pos2, f3, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dvel2, vel2, f3, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(drho2, rho2, f1, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dp2, p2, f1, hipMemcpyHostToDevice);
  if (e != hipSuccess) rc = hip_status(e, "sph_gnn_judge: buffers");
  if (!rc) rc = sph_sim_set_state_device(judge_sim, dpos, dvel, drho, dp, dtype);
  if (!rc) rc = sph_sim_eval_device(judge_sim, d0, a0);
  if (!rc) rc = sph_sim_set_state_device(judge_sim, dpos2, dvel2, drho2, dp2, dtype);
  if (!rc) rc = sph_sim_eval_device(judge_sim, d1, a1);
  if (!rc) rc = sph_sim_synchronize(judge_sim);
  double J[JW];
  const float T = (float)k * P.timeStep;
  if (!rc) {
    e = judge_read(P, T, (float)(front - front_band), dpos, dvel, drho, dpos2, dvel2, drho2, dtype, a0, d0, a1, d1,
                   part, jout, n, nullptr, J);
    if (e != hipSuccess) rc = hip_status(e, "sph_gnn_judge");
  }
  for (void* q : bufs) hipFree(q);
  if (rc) return rc;
  const double gT = std::fabs((double)P.gravity) * (double)T;
  out[0] = J[1] > 0 ? J[2] : 0.0;
  out[1] = J[1] > 0 ? J[0] / J[1] : 0.0;
  out[2] = judge_band_dev(J, gT);
  out[3] = judge_bias_dev(J, gT);
  out[4] = J[3];
  out[5] = J[4];
  out[6] = J[5];
  return SPH_OK;
}

extern "C" int sph_gnn_load_state(sph_gnn* g, const sph_params* p, int n, const float* pos, const float* hist,
                                  const float* rho, const int* type, const float* target) {
  if (!g || !p || n <= 0 || !pos || !hist || !rho || !type) return set_error(SPH_E_INVAL, "sph_gnn_load_state: bad argument");
  if (p->dim != 2) return set_error(SPH_E_INVAL, "sph_gnn_load_state: the surrogate is 2-D");
  if (n > g->cfg.max_nodes) return set_error(SPH_E_INVAL, "sph_gnn_load_state: more nodes than max_nodes");
  hipStream_t s = g->stream;
  g->prm = *p;
  g->n = n;
  const size_t N = (size_t)n;
  SPH_HIP(hipMemcpyAsync(g->pos, pos, 3 * N * sizeof(float), hipMemcpyHostToDevice, s), "upload");
  SPH_HIP(hipMemcpyAsync(g->vhist, hist, (size_t)g->cfg.history * 3 * N * sizeof(float), hipMemcpyHostToDevice, s),
          "upload");
  SPH_HIP(hipMemcpyAsync(g->rho, rho, N * sizeof(float), hipMemcpyHostToDevice, s), "upload");
  SPH_HIP(hipMemcpyAsync(g->type, type, N * sizeof(int), hipMemcpyHostToDevice, s), "upload");
  g->has_target = target != nullptr;
  if (target) SPH_HIP(hipMemcpyAsync(g->target, target, 3 * N * sizeof(float), hipMemcpyHostToDevice, s), "upload");
  int rc = build_graph_features(g, *p, n);
  if (rc) return rc;
  // fluid nodes in the loss: inside the escape box (the kernels' test)
  g->nfluid = 0;
  for (int i = 0; i < n; ++i) {
    const double x = pos[3 * i], z = pos[3 * i + 2];
    g->nfluid += type[i] == 1 &&
                 !(x > p->domain_max[0] || x < p->domain_min[0] || z > p->domain_max[2] || z < p->domain_min[2]);
  }
  SPH_HIP(hipStreamSynchronize(s), "load sync");
  return SPH_OK;
}

extern "C" int sph_gnn_graph(sph_gnn* g, int* edges, int* offsets, int* senders, int* receivers) {
  if (!g || !edges) return set_error(SPH_E_INVAL, "null argument");
  *edges = g->ne;
  if (offsets) SPH_HIP(hipMemcpy(offsets, g->off, ((size_t)g->n + 1) * sizeof(int), hipMemcpyDeviceToHost), "graph");
  if (senders) SPH_HIP(hipMemcpy(senders, g->send, (size_t)g->ne * sizeof(int), hipMemcpyDeviceToHost), "graph");
  if (receivers) SPH_HIP(hipMemcpy(receivers, g->recv, (size_t)g->ne * sizeof(int), hipMemcpyDeviceToHost), "graph");
  return SPH_OK;
}

extern "C" int sph_gnn_features(sph_gnn* g, float* node_feat, float* edge_feat) {
  if (!g) return set_error(SPH_E_INVAL, "null argument");
  if (node_feat) SPH_HIP(hipMemcpy(node_feat, g->Xn, (size_t)g->n * W * sizeof(float), hipMemcpyDeviceToHost), "features");
  if (edge_feat) SPH_HIP(hipMemcpy(edge_feat, g->Xe, (size_t)g->ne * W * sizeof(float), hipMemcpyDeviceToHost), "features");
  return SPH_OK;
}

extern "C" int sph_gnn_get_tensor(sph_gnn* g, int kind, int index, void* host, long long count) {
  if (!g || !host) return set_error(SPH_E_INVAL, "null argument");
  if (!g->cfg.train) return set_error(SPH_E_STATE, "sph_gnn_get_tensor: activations are kept only with train=1");
  if (g->n <= 0) return set_error(SPH_E_STATE, "sph_gnn_get_tensor: no state loaded");
  const int L = g->L, ncalls = 3 + 2 * L;
  const void* src = nullptr;
  long long want = 0;
  size_t esz = sizeof(float);
  switch (kind) {
    case 0:
    case 1: {
      if (index < 0 || index >= ncalls) return set_error(SPH_E_INVAL, "sph_gnn_get_tensor: bad MLP call index");
      // edge-set calls: the edge encoder (1) and every edge update (2 + 2l)
      const bool edges = index == 1 || (index >= 2 && index < 2 + 2 * L && index % 2 == 0);
      want = (long long)(edges ? g->ne : g->n) * W;
      src = kind == 0 ? h1_of(g, index) : h2_of(g, index);
      esz = sizeof(bf16);
      break;
    }
    case 2:
      if (index < 0 || index > L) return set_error(SPH_E_INVAL, "sph_gnn_get_tensor: bad layer");
      src = V_of(g, index);
      want = (long long)g->n * W;
      break;
    case 3:
      if (index < 0 || index > L) return set_error(SPH_E_INVAL, "sph_gnn_get_tensor: bad layer");
      src = E_of(g, index);
      want = (long long)g->ne * W;
      break;
    case 4:
      if (index < 0 || index >= L) return set_error(SPH_E_INVAL, "sph_gnn_get_tensor: bad layer");
      src = A_of(g, index);
      want = (long long)g->n * W;
      break;
    case 5:
      src = g->Y;
      want = (long long)g->n * W;
      break;
    default:
      return set_error(SPH_E_INVAL, "sph_gnn_get_tensor: unknown kind");
  }
  if (count != want) return set_error(SPH_E_INVAL, "sph_gnn_get_tensor: count does not match the tensor");
  SPH_HIP(hipStreamSynchronize(g->stream), "get tensor sync");
  SPH_HIP(hipMemcpy(host, src, (size_t)want * esz, hipMemcpyDeviceToHost), "get tensor");
  return SPH_OK;
}

extern "C" int sph_gnn_forward(sph_gnn* g, float* out, double* loss) {
  if (!g || g->n <= 0) return set_error(SPH_E_STATE, "sph_gnn_forward: no state loaded");
  hipStream_t s = g->stream;
  SPH_HIP(forward(g), "gnn forward");
  if (g->has_target && g->cfg.train) {
    SPH_HIP(fill_f32(g->dY, 0.0f, (long long)g->n * W, s), "loss");
    const float inv = g->nfluid > 0 ? 1.0f / (3.0f * (float)g->nfluid) : 0.0f;
    hipLaunchKernelGGL(loss_kernel, dim3(nblk(g->n)), dim3(256), 0, s, g->prm, g->pos, g->Y, g->target, g->type,
                       inv, g->dY, g->lpart, g->n);
    hipLaunchKernelGGL(loss_final_kernel, dim3(1), dim3(256), 0, s, g->lpart, (int)nblk(g->n), (double*)g->scal);
    SPH_HIP(hipGetLastError(), "loss");
  }
  if (out) {
    float* d = nullptr;
    SPH_HIP(hipMalloc(&d, (size_t)g->n * 3 * sizeof(float)), "out");
    hipLaunchKernelGGL(gather_out_kernel, dim3(nblk(g->n)), dim3(256), 0, s, g->Y, d, g->n);
    hipError_t e = hipStreamSynchronize(s);
    if (e == hipSuccess) e = hipMemcpy(out, d, (size_t)g->n * 3 * sizeof(float), hipMemcpyDeviceToHost);
    hipFree(d);
    SPH_HIP(e, "out");
  }
  if (loss) {
    double l = 0.0;
    if (g->has_target && g->cfg.train)
      SPH_HIP(hipMemcpyAsync(&l, g->scal, sizeof(double), hipMemcpyDeviceToHost, s), "loss");
    SPH_HIP(hipStreamSynchronize(s), "loss sync");
    *loss = l;
  }
  SPH_HIP(hipStreamSynchronize(s), "forward sync");
  return SPH_OK;
}

extern "C" int sph_gnn_backward(sph_gnn* g) {
  if (!g || !g->cfg.train || !g->has_target) return set_error(SPH_E_STATE, "sph_gnn_backward: needs train=1 and a target");
  SPH_HIP(backward(g), "gnn backward");
  SPH_HIP(hipStreamSynchronize(g->stream), "backward sync");
  return SPH_OK;
}

extern "C" int sph_gnn_zero_grad(sph_gnn* g) {
  if (!g) return set_error(SPH_E_INVAL, "null argument");
  SPH_HIP(fill_f32(g->G, 0.0f, (long long)g->nparam, g->stream), "zero grad");
  return SPH_OK;
}

extern "C" int sph_gnn_adam(sph_gnn* g, float lr, float clip) {
  if (!g) return set_error(SPH_E_INVAL, "null argument");
  if (!g->cfg.train) return set_error(SPH_E_STATE, "sph_gnn_adam: needs train = 1");
  hipStream_t s = g->stream;
  float scale = 1.0f;
  if (clip > 0.0f) {
    SPH_HIP(sum_squares(g->G, (long long)g->nparam, g->scal + 2, g->colpart, s), "clip");
    float sq = 0.0f;
    SPH_HIP(hipMemcpyAsync(&sq, g->scal + 2, sizeof(float), hipMemcpyDeviceToHost, s), "clip");
    SPH_HIP(hipStreamSynchronize(s), "clip");
    const float norm = std::sqrt(sq);
    if (norm > clip) scale = clip / norm;
  }
  ++g->adam_t;
  SPH_HIP(adam(g->P, g->G, g->M1, g->M2, (long long)g->nparam, lr, 0.9f, 0.999f, 1e-8f, g->adam_t, scale, s), "adam");
  SPH_HIP(pack_all(g), "pack");
  SPH_HIP(hipStreamSynchronize(s), "adam sync");
  return SPH_OK;
}

extern "C" int sph_gnn_set_fused(sph_gnn* g, int fused) {
  if (!g || !(fused == 0 || fused == 1 || fused == 2 || fused == 4 || fused == 8 || fused == 9))
    return set_error(SPH_E_INVAL, "sph_gnn_set_fused: bad argument");
  if (fused >= 8 && g->cfg.train) return set_error(SPH_E_STATE, "sph_gnn_set_fused: modes 8 and 9 need train = 0");
  g->fused = fused;
  return SPH_OK;
}

extern "C" int sph_gnn_mfma_probe(const unsigned short* A, const unsigned short* B, float* D) {
  if (!A || !B || !D) return set_error(SPH_E_INVAL, "null argument");
  bf16 *dA = nullptr, *dB = nullptr;
  float* dD = nullptr;
  hipError_t e = hipMalloc(&dA, 16 * 32 * sizeof(bf16));
  if (e == hipSuccess) e = hipMalloc(&dB, 32 * 16 * sizeof(bf16));
  if (e == hipSuccess) e = hipMalloc(&dD, 16 * 16 * sizeof(float));
  if (e == hipSuccess) e = hipMemcpy(dA, A, 16 * 32 * sizeof(bf16), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dB, B, 32 * 16 * sizeof(bf16), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = mfma_tile(dA, dB, dD, nullptr);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e == hipSuccess) e = hipMemcpy(D, dD, 16 * 16 * sizeof(float), hipMemcpyDeviceToHost);
  hipFree(dA);
  hipFree(dB);
  hipFree(dD);
  SPH_HIP(e, "sph_gnn_mfma_probe");
  return SPH_OK;
}
