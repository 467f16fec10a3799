// The expectation values of XPECT3 on one GPU (source/xpect3.f90, cited as
// X:<line>): <psi_i | P_n | psi_i> of every level over every property.
//
//   1. DSRD/REST (X:763-925): a k block of FBR coefficients is evaluated at
//      the quadrature points with sqrt(w) folded in, d_b = ang (npot x nin)
//      X_b (nin x nrad ne): one MFMA GEMM over all of its levels; the DVR
//      amplitudes of a J = 0 file are already values at the grid points and
//      are used as they are;
//   2. TRANS (X:1048-1078): the reference adds P_n(g) |d(i, g)|^2 point by
//      point; here the values are squared in one pass and the sums over the
//      grid are one TN GEMM out += (d o d)^T P (M = levels, N = properties,
//      K = grid points, split over work-groups).
#include "../../include/dvr3d_amd.h"
#include "dvr_internal.h"

#include <algorithm>
#include <vector>

namespace dvr {

namespace {

// S[g + K i] = D[g + ld i]^2 for g < K, i < ne
__global__ void __launch_bounds__(256)
k_xp_square(const double* __restrict__ D, long ld, long K, long ne, double* __restrict__ S) {
  const long tot = K * ne;
  for (long x = blockIdx.x * 256L + threadIdx.x; x < tot; x += (long)gridDim.x * 256) {
    const long i = x / K, g = x - i * K;
    const double v = D[g + ld * i];
    S[x] = v * v;
  }
}

int grid_rows(long n) { return (int)std::min<long>((n + 255) / 256, 16384); }

}  // namespace

}  // namespace dvr

using namespace dvr;

extern "C" int dvr_xpect3(const dvr_xp_dims_t* D, const dvr_xp_block_t* blk, const double* prop,
                          double* out) {
  if (!D || !blk || !prop || !out || D->npts < 1 || D->ne < 1 || D->nblk < 1 || D->nprop < 1 ||
      D->npot < 0)
    return DVR_ERR_ARG;
  const long K = D->npts, ne = D->ne, np = D->nprop;
  for (int b = 0; b < D->nblk; ++b) {
    const dvr_xp_block_t& B = blk[b];
    if (!B.vec) continue;                                    // empty block
    if (B.nin < 0) return DVR_ERR_ARG;
    if (B.nin == 0) {
      if (B.ldv < K) return DVR_ERR_ARG;                     // values at the K points as given
    } else {
      if (!B.ang || D->npot < 1 || K % D->npot != 0) return DVR_ERR_ARG;
      if (B.ldv != (long)B.nin * (K / D->npot)) return DVR_ERR_ARG;
    }
  }
  hipStream_t s = nullptr;
  try {
    DVR_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    DevBuf dp, dout, X, A, V, S, work;
    const double* P = prop;
    if (!on_device(prop)) {
      dp.ensure(sizeof(double) * K * np);
      DVR_HIP(hipMemcpyAsync(dp.p, prop, sizeof(double) * K * np, hipMemcpyHostToDevice, s));
      P = dp.as<double>();
    }
    dout.ensure(sizeof(double) * ne * np);
    DVR_HIP(hipMemcpyAsync(dout.p, out, sizeof(double) * ne * np, hipMemcpyHostToDevice, s));
    S.ensure(sizeof(double) * K * ne);
    for (int b = 0; b < D->nblk; ++b) {
      const dvr_xp_block_t& B = blk[b];
      if (!B.vec) continue;
      const double* x = B.vec;
      const long nv = B.ldv * ne;
      if (!on_device(x)) {
        X.ensure(sizeof(double) * nv);
        DVR_HIP(hipMemcpyAsync(X.p, B.vec, sizeof(double) * nv, hipMemcpyHostToDevice, s));
        x = X.as<double>();
      }
      const double* vals = x;
      long ld = B.ldv;
      if (B.nin > 0) {
        // values at the points: (npot x nin) times (nin x nrad*ne), in level
        // chunks that keep the 64-column tile count under 65536
        const long nin = B.nin, nrad = K / D->npot;
        const double* ang = B.ang;
        if (!on_device(ang)) {
          A.ensure(sizeof(double) * D->npot * nin);
          DVR_HIP(hipMemcpyAsync(A.p, B.ang, sizeof(double) * D->npot * nin,
                                 hipMemcpyHostToDevice, s));
          ang = A.as<double>();
        }
        V.ensure(sizeof(double) * K * ne);
        const long chunk = std::max<long>(1, std::min<long>(ne, (65535L * 64) / nrad));
        for (long l0 = 0; l0 < ne; l0 += chunk) {
          const long m = std::min(chunk, ne - l0);
          launch_gemm_var(s, "xp:rest", false, false, D->npot, (int)(nrad * m), (int)nin, 1.0, ang,
                          D->npot, x + nin * nrad * l0, nin, 0.0, V.as<double>() + K * l0,
                          D->npot, false, 16, false, false);
        }
        vals = V.as<double>();
        ld = K;
      }
      DVR_LAUNCH("xp:square", 16.0 * K * ne, 0.0, k_xp_square, dim3(grid_rows(K * ne)),
                 dim3(256), 0, s, vals, ld, K, ne, S.as<double>());
      const int ks = gemm_auto_ksplit((int)ne, (int)np, (int)std::min<long>(K, 1L << 30));
      work.ensure(sizeof(double) * std::max<size_t>(1, gemm_work_doubles((int)ne, (int)np, ks)));
      launch_gemm(s, "xp:trans", true, false, (int)ne, (int)np, (int)K, 1.0, S.as<double>(), K,
                  P, K, 1.0, dout.as<double>(), ne, ks, work.as<double>(), false);
    }
    DVR_HIP(hipMemcpyAsync(out, dout.p, sizeof(double) * ne * np, hipMemcpyDeviceToHost, s));
    DVR_HIP(hipStreamSynchronize(s));
    (void)hipStreamDestroy(s);
  } catch (const std::exception& e) {
    fprintf(stderr, "dvr_xpect3: %s\n", e.what());
    if (s) {
      (void)hipStreamSynchronize(s);
      (void)hipStreamDestroy(s);
    }
    return DVR_ERR_HIP;
  }
  return DVR_OK;
}
