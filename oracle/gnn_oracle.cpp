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

struct Acts {
  MlpAct en, ee, dec;
  std::vector<MlpAct> e, v;  // per processor block
};

// Backward pass from dY (n x 128) given every MLP's recorded activations;
// accumulates parameter gradients into G.
void backward(const Net& net, int L, int n, int ne, const int* send, const int* recv, const Acts& a,
              const Vec& dY, double* G) {
  const int enc_n = 0, enc_e = 3, dec = 6 + 6 * L;
  auto pe = [](int l) { return 6 + 6 * l; };
  auto pn = [](int l) { return 9 + 6 * l; };
  Vec dV, dE((size_t)ne * W, 0.0), dx;
  mlp_bwd(net, dec, a.dec, n, dY, G, &dV);
  for (int l = L - 1; l >= 0; --l) {
    // node update V_{l+1} = V_l + MLP([V_l, A_l]); A_l = sum of E_{l+1} over received edges
    mlp_bwd(net, pn(l), a.v[l], n, dV, G, &dx);
    for (int i = 0; i < n; ++i)
      for (int c = 0; c < W; ++c) dV[(size_t)i * W + c] += dx[(size_t)i * 2 * W + c];
    for (int e = 0; e < ne; ++e)
      for (int c = 0; c < W; ++c) dE[(size_t)e * W + c] += dx[(size_t)recv[e] * 2 * W + W + c];
    // edge update E_{l+1} = E_l + MLP([E_l, V_l[recv], V_l[send]])
    mlp_bwd(net, pe(l), a.e[l], ne, dE, G, &dx);
    for (int e = 0; e < ne; ++e)
      for (int c = 0; c < W; ++c) {
        dE[(size_t)e * W + c] += dx[(size_t)e * 3 * W + c];
        dV[(size_t)recv[e] * W + c] += dx[(size_t)e * 3 * W + W + c];
        dV[(size_t)send[e] * W + c] += dx[(size_t)e * 3 * W + 2 * W + c];
      }
  }
  mlp_bwd(net, enc_e, a.ee, ne, dE, G, nullptr);
  mlp_bwd(net, enc_n, a.en, n, dV, G, nullptr);
}

Vec edge_input(const Vec& E, const Vec& V, int ne, const int* send, const int* recv) {
  Vec in((size_t)ne * 3 * W);
  for (int e = 0; e < ne; ++e)
    for (int c = 0; c < W; ++c) {
      in[(size_t)e * 3 * W + c] = E[(size_t)e * W + c];
      in[(size_t)e * 3 * W + W + c] = V[(size_t)recv[e] * W + c];
      in[(size_t)e * 3 * W + 2 * W + c] = V[(size_t)send[e] * W + c];
    }
  return in;
}

Vec node_input(const Vec& V, const Vec& A, int n) {
  Vec in((size_t)n * 2 * W);
  for (int i = 0; i < n; ++i)
    for (int c = 0; c < W; ++c) {
      in[(size_t)i * 2 * W + c] = V[(size_t)i * W + c];
      in[(size_t)i * 2 * W + W + c] = A[(size_t)i * W + c];
    }
  return in;
}

// mean squared error over fluid nodes and 3 outputs, and its gradient
double loss_grad(int n, const Vec& Y, const float* target, const int* type, Vec& dY) {
  int nf = 0;
  for (int i = 0; i < n; ++i) nf += type[i] == 1;
  const double inv = nf ? 1.0 / (3.0 * nf) : 0.0;
  double loss = 0.0;
  dY.assign((size_t)n * W, 0.0);
  for (int i = 0; i < n; ++i) {
    if (type[i] != 1) continue;
    for (int c = 0; c < 3; ++c) {
      const double r = Y[(size_t)i * W + c] - target[(size_t)i * 3 + c];
      loss += r * r * inv;
      dY[(size_t)i * W + c] = 2.0 * r * inv;
    }
  }
  return loss;
}

double bf16_to_double(uint16_t b) {
  const uint32_t u = (uint32_t)b << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return (double)f;
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
  Acts act;
  act.e.resize(L);
  act.v.resize(L);
  mlp_fwd(net, enc_n, xn, n, act.en, V[0]);
  mlp_fwd(net, enc_e, xe, ne, act.ee, E[0]);
  for (int l = 0; l < L; ++l) {
    Vec d;
    mlp_fwd(net, pe(l), edge_input(E[l], V[l], ne, send, recv), ne, act.e[l], d);
    E[l + 1] = E[l];
    for (size_t i = 0; i < d.size(); ++i) E[l + 1][i] += d[i];
    A[l].assign((size_t)n * W, 0.0);
    for (int i = 0; i < n; ++i)
      for (int e = off[i]; e < off[i + 1]; ++e)
        for (int c = 0; c < W; ++c) A[l][(size_t)i * W + c] += E[l + 1][(size_t)e * W + c];
    mlp_fwd(net, pn(l), node_input(V[l], A[l], n), n, act.v[l], d);
    V[l + 1] = V[l];
    for (size_t i = 0; i < d.size(); ++i) V[l + 1][i] += d[i];
  }
  Vec Y;
  mlp_fwd(net, dec, V[L], n, act.dec, Y);
  if (Yout)
    for (int i = 0; i < n; ++i)
      for (int c = 0; c < 3; ++c) Yout[(size_t)i * 3 + c] = Y[(size_t)i * W + c];
  Vec dY;
  const double loss = loss_grad(n, Y, target, type, dY);
  if (G) backward(net, L, n, ne, send, recv, act, dY, G);
  return loss;
}

// Gradient of the BF16 network evaluated at activations recorded by another
// implementation: Vs (L+1) x n x 128 and Es (L+1) x ne x 128 node and edge
// latents, As L x n x 128 aggregated messages, H the hidden activations h1, h2
// (BF16 bit patterns) of MLP calls 0 .. 2+2L in order (rows = ne for calls 1
// and 2+2l, n otherwise), Y n x 128 decoder output.  The backward pass is the
// one orc_gnn_loss runs (operands rounded to BF16 where the device rounds
// them), but its ReLU masks and layer inputs come from the recorded forward,
// so it does not inherit the forward's sensitivity to rounding-boundary
// crossings.  Returns the loss of Y.
double orc_gnn_backward_at(int L, const float* params, const int* layout5, int nlin, int n, int ne,
                           const float* Xn, const float* Xe, const int* send, const int* recv, const float* target,
                           const int* type, const float* Vs, const float* Es, const float* As, const uint16_t* H,
                           const float* Yin, double* G) {
  Net net{L, params, std::vector<int>(layout5, layout5 + 5 * nlin), true};
  auto vec = [](const float* p, size_t count) { return Vec(p, p + count); };
  const size_t nW = (size_t)n * W, eW = (size_t)ne * W;
  const uint16_t* h = H;
  auto take = [&](MlpAct& a, Vec x, int rows) {
    a.x = std::move(x);
    a.h1.resize((size_t)rows * W);
    a.h2.resize((size_t)rows * W);
    for (size_t i = 0; i < a.h1.size(); ++i) a.h1[i] = bf16_to_double(h[i]);
    h += (size_t)rows * W;
    for (size_t i = 0; i < a.h2.size(); ++i) a.h2[i] = bf16_to_double(h[i]);
    h += (size_t)rows * W;
  };
  Acts act;
  act.e.resize(L);
  act.v.resize(L);
  take(act.en, vec(Xn, nW), n);
  take(act.ee, vec(Xe, eW), ne);
  for (int l = 0; l < L; ++l) {
    const Vec Vl = vec(Vs + l * nW, nW), El = vec(Es + l * eW, eW), Al = vec(As + l * nW, nW);
    take(act.e[l], edge_input(El, Vl, ne, send, recv), ne);
    take(act.v[l], node_input(Vl, Al, n), n);
  }
  take(act.dec, vec(Vs + L * nW, nW), n);
  const Vec Y = vec(Yin, nW);
  Vec dY;
  const double loss = loss_grad(n, Y, target, type, dY);
  backward(net, L, n, ne, send, recv, act, dY, G);
  return loss;
}

}  // extern "C"
