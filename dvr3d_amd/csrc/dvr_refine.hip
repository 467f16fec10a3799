// FP64 residual-correction refinement of eigenpairs computed on an
// approximate Hamiltonian (the split-precision assembly).
//
// The exact FP64 operator is applied without forming H3, from the
// contraction coefficients of the 3D basis (mkham3 loop 81 + loadh,
// source/dvr3drjz_segmented.f90:2459-2463, 2554-2609):
//
//   (H3 X)_c = C_c^T  sum_c' xk1(c, c') C_c' X_c'  +  E2_c X_c
//
// (C_c: the G x iv2_c expansion of the chosen 2D functions of r1 block c).
// Then one Rayleigh-Ritz step on span(X) with the exact operator, and the
// residual norms of the refined pairs, all in FP64 on the device.
#include "dvr_internal.h"

namespace dvr {

void apply_h_exact(hipStream_t s, const double* cint, long G, const std::vector<int>& col0,
                   const std::vector<int>& iv2, const double* xk, const double* e2, int N3,
                   const double* X, int nv, double* Y, RefineWork& w) {
  const int nc = (int)iv2.size();
  const int chunk = 128;
  w.U.ensure(sizeof(double) * (size_t)G * chunk * nc);
  w.V.ensure(sizeof(double) * (size_t)G * chunk * nc);
  for (int v0 = 0; v0 < nv; v0 += chunk) {
    const int q = std::min(chunk, nv - v0);
    const long gq = G * q;
    double* U = w.U.as<double>();
    double* V = w.V.as<double>();
    for (int c = 0; c < nc; ++c) {
      double* Uc = U + (long)c * gq;
      if (iv2[c] == 0) {
        DVR_HIP(hipMemsetAsync(Uc, 0, sizeof(double) * gq, s));
        continue;
      }
      launch_gemm(s, "hx_expand", false, false, (int)G, q, iv2[c], 1.0, cint + (long)col0[c] * G, G,
                  X + (long)v0 * N3 + col0[c], N3, 0.0, Uc, G, 1, nullptr, false);
    }
    // V(:, c) = sum_c' U(:, c') xk1(c', c)   over the flattened G*q rows
    launch_gemm(s, "hx_mix", false, false, (int)gq, nc, nc, 1.0, U, gq, xk, nc, 0.0, V, gq, 1,
                nullptr, false);
    for (int c = 0; c < nc; ++c) {
      if (iv2[c] == 0) continue;
      const int ks = gemm_auto_ksplit(iv2[c], q, (int)G);
      if (ks > 1) w.work.ensure(sizeof(double) * gemm_work_doubles(iv2[c], q, ks));
      launch_gemm(s, "hx_project", true, false, iv2[c], q, (int)G, 1.0, cint + (long)col0[c] * G, G,
                  V + (long)c * gq, G, 0.0, Y + (long)v0 * N3 + col0[c], N3, ks,
                  ks > 1 ? w.work.as<double>() : nullptr, false);
    }
  }
  launch_add_diag(s, Y, X, e2, N3, nv);
}

double refine_pairs(hipStream_t s, const double* cint, long G, const std::vector<int>& col0,
                    const std::vector<int>& iv2, const double* xk, const double* e2, int N3,
                    double* X, int nv, std::vector<double>& eval, RefineWork& w) {
  w.Y.ensure(sizeof(double) * (size_t)N3 * nv);
  w.T.ensure(sizeof(double) * (size_t)nv * nv);
  w.Z.ensure(sizeof(double) * (size_t)nv * nv);
  w.X2.ensure(sizeof(double) * (size_t)N3 * nv);
  w.Y2.ensure(sizeof(double) * (size_t)N3 * nv);
  w.th.ensure(sizeof(double) * nv);
  w.nrm.ensure(sizeof(double) * nv);
  double* Y = w.Y.as<double>();
  apply_h_exact(s, cint, G, col0, iv2, xk, e2, N3, X, nv, Y, w);
  // Rayleigh-Ritz on span(X) with the exact operator
  double* T = w.T.as<double>();
  const int ks = gemm_auto_ksplit(nv, nv, N3);
  if (ks > 1) w.work.ensure(sizeof(double) * gemm_work_doubles(nv, nv, ks));
  launch_gemm(s, "refine_proj", true, false, nv, nv, N3, 1.0, X, N3, Y, N3, 0.0, T, nv, ks,
              ks > 1 ? w.work.as<double>() : nullptr, false);
  launch_symmetrize(s, T, nv);
  w.eig.tridiag(std::vector<int>{nv}, T, std::vector<long>{0}, s);
  w.eig.values(std::vector<int>{nv}, s);
  eval.assign(w.eig.hval.begin(), w.eig.hval.begin() + nv);
  w.eig.vectors(std::vector<int>{nv}, w.Z.as<double>(), std::vector<long>{0}, s);
  launch_gemm(s, "refine_rot", false, false, N3, nv, nv, 1.0, X, N3, w.Z.as<double>(), nv, 0.0,
              w.X2.as<double>(), N3, 1, nullptr, false);
  launch_gemm(s, "refine_rot", false, false, N3, nv, nv, 1.0, Y, N3, w.Z.as<double>(), nv, 0.0,
              w.Y2.as<double>(), N3, 1, nullptr, false);
  DVR_HIP(hipMemcpyAsync(w.th.p, eval.data(), sizeof(double) * nv, hipMemcpyHostToDevice, s));
  launch_resid_norms(s, w.X2.as<double>(), w.Y2.as<double>(), N3, w.th.as<double>(), nv, N3,
                     w.nrm.as<double>());
  std::vector<double> nr(nv);
  DVR_HIP(hipMemcpyAsync(nr.data(), w.nrm.p, sizeof(double) * nv, hipMemcpyDeviceToHost, s));
  DVR_HIP(hipMemcpyAsync(X, w.X2.p, sizeof(double) * (size_t)N3 * nv, hipMemcpyDeviceToDevice, s));
  DVR_HIP(hipStreamSynchronize(s));
  double mx = 0.0;
  for (double r : nr) mx = std::max(mx, r);
  return mx;
}

// Rayleigh quotients with the exact operator, one per pair, for eigenvectors
// X (N3 x nv, orthonormal) of an approximate H3 when every level is wanted
// (a Rayleigh-Ritz over their span would be a second full eigensolve):
// eval_i = x_i^T H x_i, whose error is of the order of the squared residual
// over the gap to the next level; X is left as it is.  Returns the largest
// residual norm ||H x_i - eval_i x_i||.
double refine_rq(hipStream_t s, const double* cint, long G, const std::vector<int>& col0,
                 const std::vector<int>& iv2, const double* xk, const double* e2, int N3,
                 const double* X, int nv, std::vector<double>& eval, RefineWork& w) {
  w.Y.ensure(sizeof(double) * (size_t)N3 * nv);
  w.th.ensure(sizeof(double) * nv);
  w.nrm.ensure(sizeof(double) * nv);
  double* Y = w.Y.as<double>();
  apply_h_exact(s, cint, G, col0, iv2, xk, e2, N3, X, nv, Y, w);
  launch_coldot(s, X, Y, N3, nv, w.th.as<double>());
  launch_resid_norms(s, X, Y, N3, w.th.as<double>(), nv, N3, w.nrm.as<double>());
  eval.assign(nv, 0.0);
  std::vector<double> nr(nv);
  DVR_HIP(hipMemcpyAsync(eval.data(), w.th.p, sizeof(double) * nv, hipMemcpyDeviceToHost, s));
  DVR_HIP(hipMemcpyAsync(nr.data(), w.nrm.p, sizeof(double) * nv, hipMemcpyDeviceToHost, s));
  DVR_HIP(hipStreamSynchronize(s));
  double mx = 0.0;
  for (double r : nr) mx = std::max(mx, r);
  return mx;
}

}  // namespace dvr
