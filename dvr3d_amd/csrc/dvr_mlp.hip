// Training of the learned Davidson preconditioner on the GPU.
//
// Data: the eigenpairs (lambda_k, z_k) that the dense FP64 path produced for
// one Hamiltonian H (N x N).  Each training pair imitates a Davidson Ritz
// pair: x = normalise(z_k + sum_l c_l z_l) over the 40 levels nearest to
// lambda_k, c_l ~ 0.05 N(0,1) / (1 + |lambda_l - lambda_k| / 5e-3 Eh);
// theta = x^T H x, r = H x - theta x, and the true error e = z_k - x projected
// out of x.  The correction of the learned preconditioner is
// t_j = w_j r_j / (D_j - theta) (w = network output per basis function); the
// loss is 1 - cos^2(P t, e), P = I - x x^T, averaged over the batch.
//
// One step: sample generation (MFMA GEMMs), FP32 forward with stored
// activations, loss and dL/dt per pair, dL/dw, FP32 backward, Adam.
#include "dvr_internal.h"

#include <cmath>
#include <numeric>
#include <random>

namespace dvr {

int mlp_train(hipStream_t s, const double* H, int N, const double* Z, int nz,
              const std::vector<double>& lam, const double* D, const double* st, float* P,
              int steps, int S, unsigned long long seed, float lr, std::vector<double>& hist,
              MlpWork& w) {
  if (nz < 2 || S < 1 || steps < 1) return 1;
  const long R = (long)N * S;
  w.X.ensure(sizeof(double) * R);
  w.AX.ensure(sizeof(double) * R);
  w.Rr.ensure(sizeof(double) * R);
  w.E.ensure(sizeof(double) * R);
  w.T.ensure(sizeof(double) * R);
  w.G.ensure(sizeof(double) * R);
  w.Cm.ensure(sizeof(double) * (size_t)nz * S);
  w.th.ensure(sizeof(double) * S);
  w.c.ensure(sizeof(double) * S);
  w.loss.ensure(sizeof(double) * S);
  w.F.ensure(sizeof(float) * R * MLP_NF);
  w.W.ensure(sizeof(float) * R);
  w.GW.ensure(sizeof(float) * R);
  w.H1.ensure(sizeof(float) * R * MLP_WIDTH);
  w.H2.ensure(sizeof(float) * R * MLP_WIDTH);
  w.GA1.ensure(sizeof(float) * R * MLP_WIDTH);
  w.GA2.ensure(sizeof(float) * R * MLP_WIDTH);
  const long pcap = 64L * MLP_WIDTH * MLP_WIDTH;
  w.partial.ensure(sizeof(float) * pcap);
  w.grad.ensure(sizeof(float) * MLP_NPARAM);
  w.m.ensure(sizeof(float) * MLP_NPARAM);
  w.v.ensure(sizeof(float) * MLP_NPARAM);
  DVR_HIP(hipMemsetAsync(w.m.p, 0, sizeof(float) * MLP_NPARAM, s));
  DVR_HIP(hipMemsetAsync(w.v.p, 0, sizeof(float) * MLP_NPARAM, s));
  {
    std::vector<float> ones(R, 1.0f);
    w.ones.ensure(sizeof(float) * R);
    DVR_HIP(hipMemcpyAsync(w.ones.p, ones.data(), sizeof(float) * R, hipMemcpyHostToDevice, s));
  }
  std::mt19937_64 rng(seed);
  std::normal_distribution<double> gauss(0.0, 1.0);
  std::uniform_int_distribution<int> pick(0, nz - 1);
  double* X = w.X.as<double>();
  double* AX = w.AX.as<double>();
  double* Rr = w.Rr.as<double>();
  double* E = w.E.as<double>();
  double* th = w.th.as<double>();
  hist.clear();
  std::vector<double> cm((size_t)nz * S), hl(S), nrm(S);
  std::vector<int> ks(S);
  std::vector<int> order(nz);
  for (int step = 1; step <= steps; ++step) {
    // ---- training pairs
    std::fill(cm.begin(), cm.end(), 0.0);
    for (int q = 0; q < S; ++q) {
      const int k = pick(rng);
      ks[q] = k;
      std::iota(order.begin(), order.end(), 0);
      const int nn = std::min(nz, 41);
      std::partial_sort(order.begin(), order.begin() + nn, order.end(), [&](int a, int b) {
        return std::fabs(lam[a] - lam[k]) < std::fabs(lam[b] - lam[k]);
      });
      cm[(size_t)q * nz + k] = 1.0;
      for (int t = 0; t < nn; ++t) {
        const int l = order[t];
        if (l == k) continue;
        cm[(size_t)q * nz + l] = 0.05 * gauss(rng) / (1.0 + std::fabs(lam[l] - lam[k]) / 5e-3);
      }
    }
    DVR_HIP(hipMemcpyAsync(w.Cm.p, cm.data(), sizeof(double) * cm.size(), hipMemcpyHostToDevice, s));
    launch_gemm(s, "mlp_pairs", false, false, N, S, nz, 1.0, Z, N, w.Cm.as<double>(), nz, 0.0, X, N,
                1, nullptr, false);
    launch_coldot(s, X, X, N, S, w.c.as<double>());
    DVR_HIP(hipMemcpyAsync(nrm.data(), w.c.p, sizeof(double) * S, hipMemcpyDeviceToHost, s));
    DVR_HIP(hipStreamSynchronize(s));
    for (int q = 0; q < S; ++q) nrm[q] = 1.0 / std::sqrt(nrm[q]);
    DVR_HIP(hipMemcpyAsync(w.c.p, nrm.data(), sizeof(double) * S, hipMemcpyHostToDevice, s));
    launch_scale_cols(s, X, N, N, S, w.c.as<double>());
    launch_gemm(s, "mlp_pairs", false, false, N, S, N, 1.0, H, N, X, N, 0.0, AX, N, 1, nullptr,
                false);
    launch_coldot(s, X, AX, N, S, th);
    DVR_HIP(hipMemcpyAsync(Rr, AX, sizeof(double) * R, hipMemcpyDeviceToDevice, s));
    launch_col_axpy(s, X, th, -1.0, N, S, Rr);
    // e = z_k - x, projected out of x
    launch_gather_cols(s, Z, N, upload(w.kidx, ks, s), S, E, N, N);
    {
      std::vector<double> one(S, 1.0);
      DVR_HIP(hipMemcpyAsync(w.c.p, one.data(), sizeof(double) * S, hipMemcpyHostToDevice, s));
      launch_col_axpy(s, X, w.c.as<double>(), -1.0, N, S, E);
      launch_coldot(s, X, E, N, S, w.c.as<double>());
      launch_col_axpy(s, X, w.c.as<double>(), -1.0, N, S, E);
    }
    // ---- forward, loss, backward, update
    launch_mlp_features(s, D, st, N, S, th, th, w.F.as<float>());
    launch_mlp_fwd(s, P, R, w.F.as<float>(), w.W.as<float>(), w.H1.as<float>(), w.H2.as<float>());
    launch_mlp_apply(s, w.W.as<float>(), D, th, Rr, N, S, w.T.as<double>());
    launch_mlp_loss(s, w.T.as<double>(), X, E, N, S, w.G.as<double>(), w.loss.as<double>());
    launch_mlp_gw(s, w.G.as<double>(), D, th, Rr, N, S, 1.0 / S, w.GW.as<float>());
    launch_mlp_bwd(s, P, R, w.F.as<float>(), w.H1.as<float>(), w.H2.as<float>(), w.GW.as<float>(),
                   w.GA1.as<float>(), w.GA2.as<float>(), w.grad.as<float>(), w.partial.as<float>(),
                   pcap);
    launch_mlp_bias_grads(s, R, w.ones.as<float>(), w.GA1.as<float>(), w.GA2.as<float>(),
                          w.GW.as<float>(), w.grad.as<float>(), w.partial.as<float>(), pcap);
    launch_adam(s, P, w.grad.as<float>(), w.m.as<float>(), w.v.as<float>(), MLP_NPARAM, lr, step);
    DVR_HIP(hipMemcpyAsync(hl.data(), w.loss.p, sizeof(double) * S, hipMemcpyDeviceToHost, s));
    DVR_HIP(hipStreamSynchronize(s));
    double mean = 0.0;
    for (double v : hl) mean += v;
    hist.push_back(mean / S);
  }
  return 0;
}

}  // namespace dvr
