// The transition moments of DIPOLE3 on one GPU (source/dipole3_segmented.f90,
// cited as P:<line>): given the wavefunctions of every k block of a ket and a
// bra file, the z and x moments of every ket/bra level pair.
//
//   1. DSRD (P:1521-1545, with JTRAN P:1559-1640 for a J = 0 IWAVE file):
//      each block's vectors at the angular integration points.  The input
//      of block b is nin functions per radial point (associated Legendre
//      FBR functions, or the DVR points), and its values at the points are
//      d_b = ang_b (ipot x nin) X_b (nin x nrad ne): one MFMA GEMM per
//      block over all of its levels at once, written straight into the
//      layout the moment products read (K = nrad ipot rows, one column per
//      level);
//   2. TRANS (P:1642-1777): the reference adds nrad*ipot rank-one DGER
//      updates per block pair; here the bra side is scaled by the dipole
//      values on the grid (LAGPT P:1187-1217, weights included) and the
//      pair's whole sum is one TN GEMM t += factor d1^T (dip .* d2) with
//      K = nrad ipot, split over work-groups when the output is small.
// Transformed blocks stay resident until their last pair has used them.
#include "../../include/dvr3d_amd.h"
#include "dvr_internal.h"

#include <algorithm>
#include <cstdlib>
#include <memory>
#include <vector>

namespace dvr {

namespace {

// E[i + K ie] = w[i] D[i + K ie]
__global__ void __launch_bounds__(256)
k_dip_scale(const double* __restrict__ D, const double* __restrict__ w, long K, long ne,
            double* __restrict__ E) {
  const long tot = K * ne;
  for (long x = blockIdx.x * 256L + threadIdx.x; x < tot; x += (long)gridDim.x * 256) {
    const long i = x % K;
    E[x] = w[i] * D[x];
  }
}

int grid_rows(long n) { return (int)std::min<long>((n + 255) / 256, 16384); }

}  // namespace

}  // namespace dvr

using namespace dvr;

extern "C" int dvr_dipole3(const dvr_dip_dims_t* D, const dvr_dip_block_t* ket,
                           const dvr_dip_block_t* bra, const double* dipol,
                           const dvr_dip_pair_t* pairs, double* tz, double* tx) {
  if (!D || !dipol || !tz || !tx || D->nrad < 1 || D->ipot < 1 || D->ne1 < 1 || D->ne2 < 1 ||
      D->nblk1 < 1 || D->nblk2 < 1 || D->npair < 0 || (D->npair > 0 && (!pairs || !ket || !bra)))
    return DVR_ERR_ARG;
  const long nrad = D->nrad, ipot = D->ipot, K = nrad * ipot;
  const int ne[2] = {D->ne1, D->ne2}, nblk[2] = {D->nblk1, D->nblk2};
  const dvr_dip_block_t* side[2] = {ket, bra};
  // last pair that reads each block (-1: never read)
  std::vector<int> last[2] = {std::vector<int>(nblk[0], -1), std::vector<int>(nblk[1], -1)};
  for (int q = 0; q < D->npair; ++q) {
    const dvr_dip_pair_t& p = pairs[q];
    if (p.b1 < 0 || p.b1 >= nblk[0] || p.b2 < 0 || p.b2 >= nblk[1] || (p.comp != 0 && p.comp != 1))
      return DVR_ERR_ARG;
    if (p.factor == 0.0) continue;
    last[0][p.b1] = q;
    last[1][p.b2] = q;
  }
  for (int sd = 0; sd < 2; ++sd)
    for (int b = 0; b < nblk[sd]; ++b)
      if (last[sd][b] >= 0) {
        const dvr_dip_block_t& B = side[sd][b];
        if (B.nin < 1 || !B.vec || !B.ang) return DVR_ERR_ARG;
      }

  hipStream_t s = nullptr;
  const char* tv = getenv("DVR_DIP_TIMES");
  const bool times = tv && atoi(tv) != 0;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  try {
    DVR_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    if (times) {
      DVR_HIP(hipEventCreate(&e0));
      DVR_HIP(hipEventCreate(&e1));
      DVR_HIP(hipEventRecord(e0, s));
    }
    DevBuf ddip, dt[2], E, X, A, work;
    const long nt = (long)ne[0] * ne[1];
    ddip.ensure(sizeof(double) * 2 * K);
    DVR_HIP(hipMemcpyAsync(ddip.p, dipol, sizeof(double) * 2 * K, hipMemcpyHostToDevice, s));
    double* hts[2] = {tz, tx};
    for (int c = 0; c < 2; ++c) {
      dt[c].ensure(sizeof(double) * nt);
      DVR_HIP(hipMemcpyAsync(dt[c].p, hts[c], sizeof(double) * nt, hipMemcpyHostToDevice, s));
    }
    std::vector<std::unique_ptr<DevBuf>> val[2];
    val[0].resize(nblk[0]);
    val[1].resize(nblk[1]);
    // DSRD of one block: values at the integration points, K x ne
    auto transform = [&](int sd, int b) {
      const dvr_dip_block_t& B = side[sd][b];
      const long nin = B.nin, n = ne[sd];
      val[sd][b].reset(new DevBuf());
      DevBuf& out = *val[sd][b];
      out.ensure(sizeof(double) * K * n);
      A.ensure(sizeof(double) * ipot * nin);
      DVR_HIP(hipMemcpyAsync(A.p, B.ang, sizeof(double) * ipot * nin, hipMemcpyHostToDevice, s));
      // the product's second grid dimension counts 64-column tiles of
      // nrad * levels: levels in chunks that keep it under 65536
      const long lev_max = std::max<long>(1, (65535L * 64) / nrad);
      const long chunk = std::min<long>(n, lev_max);
      X.ensure(sizeof(double) * nin * nrad * chunk);
      for (long l0 = 0; l0 < n; l0 += chunk) {
        const long m = std::min(chunk, n - l0);
        // the previous chunk's product reads X: the copy is ordered after it
        DVR_HIP(hipMemcpyAsync(X.p, B.vec + nin * nrad * l0, sizeof(double) * nin * nrad * m,
                               hipMemcpyHostToDevice, s));
        launch_gemm_var(s, "dip:dsrd", false, false, (int)ipot, (int)(nrad * m), (int)nin, 1.0,
                        A.as<double>(), ipot, X.as<double>(), nin, 0.0,
                        out.as<double>() + K * l0, ipot, false, 16, false, false);
      }
      // A and X are rewritten by the next block's copies, which the stream
      // orders after these products
    };
    for (int q = 0; q < D->npair; ++q) {
      const dvr_dip_pair_t& p = pairs[q];
      if (p.factor == 0.0) continue;
      if (!val[0][p.b1]) transform(0, p.b1);
      if (!val[1][p.b2]) transform(1, p.b2);
      E.ensure(sizeof(double) * K * ne[1]);
      DVR_LAUNCH("dip:scale", 16.0 * K * ne[1] + 8.0 * K, 0.0, k_dip_scale,
                 dim3(grid_rows(K * ne[1])), dim3(256), 0, s, val[1][p.b2]->as<double>(),
                 ddip.as<double>() + K * p.comp, K, (long)ne[1], E.as<double>());
      const int ks = gemm_auto_ksplit(ne[0], ne[1], (int)std::min<long>(K, 1L << 30));
      work.ensure(sizeof(double) * std::max<size_t>(1, gemm_work_doubles(ne[0], ne[1], ks)));
      launch_gemm(s, "dip:trans", true, false, ne[0], ne[1], (int)K, p.factor,
                  val[0][p.b1]->as<double>(), K, E.as<double>(), K, 1.0, dt[p.comp].as<double>(),
                  ne[0], ks, work.as<double>(), false);
      // release what no later pair reads (the stream orders the free after
      // the products that read it only once it is synchronised)
      for (int sd = 0; sd < 2; ++sd) {
        const int b = sd == 0 ? p.b1 : p.b2;
        if (last[sd][b] == q) {
          DVR_HIP(hipStreamSynchronize(s));
          val[sd][b].reset();
        }
      }
    }
    for (int c = 0; c < 2; ++c)
      DVR_HIP(hipMemcpyAsync(hts[c], dt[c].p, sizeof(double) * nt, hipMemcpyDeviceToHost, s));
    DVR_HIP(hipStreamSynchronize(s));
    if (times) {
      float ms = 0.0f;
      DVR_HIP(hipEventRecord(e1, s));
      DVR_HIP(hipEventSynchronize(e1));
      DVR_HIP(hipEventElapsedTime(&ms, e0, e1));
      fprintf(stderr, "dvr_dipole3: %10.3f ms\n", ms);
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
    }
    (void)hipStreamDestroy(s);
  } catch (const std::exception& e) {
    fprintf(stderr, "dvr_dipole3: %s\n", e.what());
    if (s) {
      (void)hipStreamSynchronize(s);
      (void)hipStreamDestroy(s);
    }
    return DVR_ERR_HIP;
  }
  return DVR_OK;
}
