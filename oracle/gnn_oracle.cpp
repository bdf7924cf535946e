// CPU ORACLE of the message-passing surrogate -- TEST INFRASTRUCTURE ONLY.
//
// A plain FP64 restatement of the network the GPU library runs (encoders, L
// processor blocks of residual edge and node MLPs with sum aggregation by
// receiver, decoder; MLPs = Linear-ReLU-Linear-ReLU-Linear, width 128) and of
// its loss (mean squared error over fluid nodes, 3 outputs) with a
// hand-written backward pass.  Inputs are the features and graph the GPU
// built, so only the network is restated.  The network has no reference
// counterpart; the backward pass is pinned by central finite differences.
//
// With bf16 = 1 the operands of every matrix product and the stored hidden
// activations are rounded to BF16 exactly where the GPU rounds them, so the
// GPU's FP32-accumulated gradients can be compared tightly.
#include <omp.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

constexpr int W = 128;

inline double bfr(double x) {
  // round to nearest even BF16 through FP32
  float f = (float)x;
  uint32_t u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7f800000u) == 0x7f800000u) return f;
  const uint32_t lsb = (u >> 16) & 1u;
  u += 0x7fffu + lsb;
  u &= 0xffff0000u;
  std::memcpy(&f, &u, 4);
  return (double)f;
}

struct Net {
  int L;
  const float* P;  // flat parameters
  std::vector<int> lay;  // 5 ints per linear layer
  bool bf;
  double q(double x) const { return bf ? bfr(x) : x; }
  const float* Wp(int li) const { return P + lay[5 * li]; }
  const float* Bp(int li) const { return P + lay[5 * li + 3]; }
  int K(int li) const { return lay[5 * li + 2]; }
};

typedef std::vector<double> Vec;

// y[m][n] = sum_k q(x[m][k]) q(w[n][k]) + b[n]  (x is M x K, row stride K)
void linear(const Net& net, int li, const Vec& x, int M, Vec& y) {
  const int Kd = net.K(li);
  const float* w = net.Wp(li);
  const float* b = net.Bp(li);
  std::vector<double> wq((size_t)W * Kd);
  for (size_t i = 0; i < wq.size(); ++i) wq[i] = net.q(w[i]);
  y.assign((size_t)M * W, 0.0);
#pragma omp parallel for schedule(static)
  for (int m = 0; m < M; ++m) {
    std::vector<double> xr(Kd);
    for (int k = 0; k < Kd; ++k) xr[k] = net.q(x[(size_t)m * Kd + k]);
    for (int n = 0; n < W; ++n) {
      double s = 0.0;
      const double* wr = &wq[(size_t)n * Kd];
      for (int k = 0; k < Kd; ++k) s += xr[k] * wr[k];
      y[(size_t)m * W + n] = s + (double)b[n];
    }
  }
}

struct MlpAct {
  Vec x, h1, h2;  // input (M x K), hidden activations (stored as the GPU stores them)
};

// out = MLP(x); records activations
void mlp_fwd(const Net& net, int first, const Vec& x, int M, MlpAct& a, Vec& out) {
  a.x = x;
  linear(net, first, x, M, a.h1);
  for (double& v : a.h1) v = net.q(v > 0 ? v : 0);
  linear(net, first + 1, a.h1, M, a.h2);
  for (double& v : a.h2) v = net.q(v > 0 ? v : 0);
  linear(net, first + 2, a.h2, M, out);
}

// gradients of one linear layer: dW[n][k] += sum_m q(dz[m][n]) q(x[m][k]); db[n] += sum_m dz[m][n]
void lin_grad(const Net& net, int li, const Vec& dz, const Vec& x, int M, double* G) {
  const int Kd = net.K(li);
  double* gw = G + net.lay[5 * li];
  double* gb = G + net.lay[5 * li + 3];
#pragma omp parallel for schedule(static)
  for (int n = 0; n < W; ++n) {
    double sb = 0.0;
    std::vector<double> acc(Kd, 0.0);
    for (int m = 0; m < M; ++m) {
      const double d = dz[(size_t)m * W + n];
      sb += d;
      const double dq = net.q(d);
      for (int k = 0; k < Kd; ++k) acc[k] += dq * net.q(x[(size_t)m * Kd + k]);
    }
    for (int k = 0; k < Kd; ++k) gw[(size_t)n * Kd + k] += acc[k];
    gb[n] += sb;
  }
}

// dx[m][k] = sum_n q(dz[m][n]) q(w[n][k]) (* mask)
void lin_back(const Net& net, int li, const Vec& dz, int M, const Vec* mask, Vec& dx) {
  const int Kd = net.K(li);
  const float* w = net.Wp(li);
  dx.assign((size_t)M * Kd, 0.0);
#pragma omp parallel for schedule(static)
  for (int m = 0; m < M; ++m) {
    for (int k = 0; k < Kd; ++k) {
      double s = 0.0;
      for (int n = 0; n < W; ++n) s += net.q(dz[(size_t)m * W + n]) * net.q(w[(size_t)n * Kd + k]);
      if (mask && !((*mask)[(size_t)m * Kd + k] > 0)) s = 0.0;
      dx[(size_t)m * Kd + k] = s;
    }
  }
}

void mlp_bwd(const Net& net, int first, const MlpAct& a, int M, const Vec& dout, double* G, Vec* dx) {
  lin_grad(net, first + 2, dout, a.h2, M, G);
  Vec dz2, dz1;
  lin_back(net, first + 2, dout, M, &a.h2, dz2);
  lin_grad(net, first + 1, dz2, a.h1, M, G);
  lin_back(net, first + 1, dz2, M, &a.h1, dz1);
  lin_grad(net, first, dz1, a.x, M, G);
  if (dx) lin_back(net, first, dz1, M, nullptr, *dx);
}

}  // namespace

extern "C" {

// Loss and (if G != null) its gradient for the loaded graph.  layout5 as in
// sph_gnn_param_layout; Xn n x 128, Xe ne x 128; off[n+1], send/recv[ne];
// target n x 3; type n.  Returns the loss (mean over fluid nodes and 3 outputs).
double orc_gnn_loss(int L, const float* params, const int* layout5, int nlin, int n, int ne, const float* Xn,
                    const float* Xe, const int* off, const int* send, const int* recv, const float* target,
                    const int* type, int bf16, double* G, double* Yout) {
  Net net{L, params, std::vector<int>(layout5, layout5 + 5 * nlin), bf16 != 0};
  const int enc_n = 0, enc_e = 3, dec = 6 + 6 * L;
  auto pe = [](int l) { return 6 + 6 * l; };
  auto pn = [](int l) { return 9 + 6 * l; };
  Vec xn(Xn, Xn + (size_t)n * W), xe(Xe, Xe + (size_t)ne * W);
  std::vector<Vec> V(L + 1), E(L + 1), A(L);
  std::vector<MlpAct> ae(L), an(L);
  MlpAct aen, aee, adec;
  mlp_fwd(net, enc_n, xn, n, aen, V[0]);
  mlp_fwd(net, enc_e, xe, ne, aee, E[0]);
  for (int l = 0; l < L; ++l) {
    Vec in((size_t)ne * 3 * W);
    for (int e = 0; e < ne; ++e)
      for (int c = 0; c < W; ++c) {
        in[(size_t)e * 3 * W + c] = E[l][(size_t)e * W + c];
        in[(size_t)e * 3 * W + W + c] = V[l][(size_t)recv[e] * W + c];
        in[(size_t)e * 3 * W + 2 * W + c] = V[l][(size_t)send[e] * W + c];
      }
    Vec d;
    mlp_fwd(net, pe(l), in, ne, ae[l], d);
    E[l + 1] = E[l];
    for (size_t i = 0; i < d.size(); ++i) E[l + 1][i] += d[i];
    A[l].assign((size_t)n * W, 0.0);
    for (int i = 0; i < n; ++i)
      for (int e = off[i]; e < off[i + 1]; ++e)
        for (int c = 0; c < W; ++c) A[l][(size_t)i * W + c] += E[l + 1][(size_t)e * W + c];
    Vec nin((size_t)n * 2 * W);
    for (int i = 0; i < n; ++i)
      for (int c = 0; c < W; ++c) {
        nin[(size_t)i * 2 * W + c] = V[l][(size_t)i * W + c];
        nin[(size_t)i * 2 * W + W + c] = A[l][(size_t)i * W + c];
      }
    mlp_fwd(net, pn(l), nin, n, an[l], d);
    V[l + 1] = V[l];
    for (size_t i = 0; i < d.size(); ++i) V[l + 1][i] += d[i];
  }
  Vec Y;
  mlp_fwd(net, dec, V[L], n, adec, Y);
  if (Yout)
    for (int i = 0; i < n; ++i)
      for (int c = 0; c < 3; ++c) Yout[(size_t)i * 3 + c] = Y[(size_t)i * W + c];
  int nf = 0;
  for (int i = 0; i < n; ++i) nf += type[i] == 1;
  const double inv = nf ? 1.0 / (3.0 * nf) : 0.0;
  double loss = 0.0;
  Vec dY((size_t)n * W, 0.0);
  for (int i = 0; i < n; ++i) {
    if (type[i] != 1) continue;
    for (int c = 0; c < 3; ++c) {
      const double r = Y[(size_t)i * W + c] - target[(size_t)i * 3 + c];
      loss += r * r * inv;
      dY[(size_t)i * W + c] = 2.0 * r * inv;
    }
  }
  if (!G) return loss;
  Vec dV, dE((size_t)ne * W, 0.0), dx;
  mlp_bwd(net, dec, adec, n, dY, G, &dV);
  for (int l = L - 1; l >= 0; --l) {
    mlp_bwd(net, pn(l), an[l], n, dV, G, &dx);
    for (int i = 0; i < n; ++i)
      for (int c = 0; c < W; ++c) dV[(size_t)i * W + c] += dx[(size_t)i * 2 * W + c];
    for (int e = 0; e < ne; ++e)
      for (int c = 0; c < W; ++c) dE[(size_t)e * W + c] += dx[(size_t)recv[e] * 2 * W + W + c];
    mlp_bwd(net, pe(l), ae[l], ne, dE, G, &dx);
    for (int e = 0; e < ne; ++e)
      for (int c = 0; c < W; ++c) {
        dE[(size_t)e * W + c] += dx[(size_t)e * 3 * W + c];
        dV[(size_t)recv[e] * W + c] += dx[(size_t)e * 3 * W + W + c];
        dV[(size_t)send[e] * W + c] += dx[(size_t)e * 3 * W + 2 * W + c];
      }
  }
  mlp_bwd(net, enc_e, aee, ne, dE, G, nullptr);
  mlp_bwd(net, enc_n, aen, n, dV, G, nullptr);
  return loss;
}

}  // extern "C"
