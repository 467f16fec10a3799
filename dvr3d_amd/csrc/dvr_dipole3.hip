// The transition moments of DIPOLE3 on one GPU (source/dipole3_segmented.f90,
// cited as P:<line>): given the wavefunctions of every k block of a ket and a
// bra file, the z and x moments of every ket/bra level pair.
//
//   1. DSRD (P:1511-1548, with JTRAN P:1564-1645 for a J = 0 IWAVE file):
//      each block's vectors at the angular integration points.  The input
//      of block b is nin functions per radial point (associated Legendre
//      FBR functions, or the DVR points), and its values at the points are
//      d_b = ang_b (ipot x nin) X_b (nin x nrad ne): one MFMA GEMM per
//      block over all of its levels at once, written straight into the
//      layout the moment products read (K = nrad ipot rows, one column per
//      level);
//   2. TRANS (P:1648-1776): the reference adds nrad*ipot rank-one DGER
//      updates per block pair; here the bra side is scaled by the dipole
//      values on the grid (LAGPT P:1206-1230, weights included) and the
//      pair's whole sum is one TN GEMM t += factor d1^T (dip .* d2) with
//      K = nrad ipot, split over work-groups when the output is small.
// Transformed blocks stay resident until their last pair has used them.
#include "../../include/dvr3d_amd.h"
#include "dvr_internal.h"

#include <algorithm>
#include <cstdlib>
#include <memory>
#include <utility>
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
                           const dvr_dip_pair_t* pairs, double* tz, double* tx,
                           dvr_dip_stats_t* stats) {
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
  // with stats: an event pair around every phase launch, summed at the end
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev[3];
  auto tick = [&](int ph) -> hipEvent_t {
    if (!stats) return nullptr;
    hipEvent_t a;
    DVR_HIP(hipEventCreate(&a));
    DVR_HIP(hipEventRecord(a, s));
    ev[ph].push_back({a, nullptr});
    return a;
  };
  auto tock = [&](int ph) {
    if (!stats) return;
    hipEvent_t b;
    DVR_HIP(hipEventCreate(&b));
    DVR_HIP(hipEventRecord(b, s));
    ev[ph].back().second = b;
  };
  auto drop_events = [&]() {
    for (auto& v : ev)
      for (auto& e : v) {
        if (e.first) (void)hipEventDestroy(e.first);
        if (e.second) (void)hipEventDestroy(e.second);
      }
  };
  try {
    DVR_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    DevBuf ddip, dt[2], E, X, A, work;
    const long nt = (long)ne[0] * ne[1];
    const double* dip = dipol;
    if (!on_device(dipol)) {
      ddip.ensure(sizeof(double) * 2 * K);
      DVR_HIP(hipMemcpyAsync(ddip.p, dipol, sizeof(double) * 2 * K, hipMemcpyHostToDevice, s));
      dip = ddip.as<double>();
    }
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
      const double* ang = B.ang;
      if (!on_device(ang)) {
        A.ensure(sizeof(double) * ipot * nin);
        DVR_HIP(hipMemcpyAsync(A.p, B.ang, sizeof(double) * ipot * nin, hipMemcpyHostToDevice,
                               s));
        ang = A.as<double>();
      }
      const bool vdev = on_device(B.vec);
      // the product's second grid dimension counts 64-column tiles of
      // nrad * levels: levels in chunks that keep it under 65536
      const long lev_max = std::max<long>(1, (65535L * 64) / nrad);
      const long chunk = std::min<long>(n, lev_max);
      if (!vdev) X.ensure(sizeof(double) * nin * nrad * chunk);
      for (long l0 = 0; l0 < n; l0 += chunk) {
        const long m = std::min(chunk, n - l0);
        const double* x = B.vec + nin * nrad * l0;
        if (!vdev) {
          // the previous chunk's product reads X: the copy is ordered after it
          DVR_HIP(hipMemcpyAsync(X.p, x, sizeof(double) * nin * nrad * m, hipMemcpyHostToDevice,
                                 s));
          x = X.as<double>();
        }
        tick(0);
        launch_gemm_var(s, "dip:dsrd", false, false, (int)ipot, (int)(nrad * m), (int)nin, 1.0,
                        ang, ipot, x, nin, 0.0, out.as<double>() + K * l0, ipot, false, 16,
                        false, false);
        tock(0);
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
      tick(1);
      DVR_LAUNCH("dip:scale", 16.0 * K * ne[1] + 8.0 * K, 0.0, k_dip_scale,
                 dim3(grid_rows(K * ne[1])), dim3(256), 0, s, val[1][p.b2]->as<double>(),
                 dip + K * p.comp, K, (long)ne[1], E.as<double>());
      tock(1);
      const int ks = gemm_auto_ksplit(ne[0], ne[1], (int)std::min<long>(K, 1L << 30));
      work.ensure(sizeof(double) * std::max<size_t>(1, gemm_work_doubles(ne[0], ne[1], ks)));
      tick(2);
      launch_gemm(s, "dip:trans", true, false, ne[0], ne[1], (int)K, p.factor,
                  val[0][p.b1]->as<double>(), K, E.as<double>(), K, 1.0, dt[p.comp].as<double>(),
                  ne[0], ks, work.as<double>(), false);
      tock(2);
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
    if (stats) {
      double ms3[3] = {0.0, 0.0, 0.0};
      for (int ph = 0; ph < 3; ++ph)
        for (auto& e : ev[ph]) {
          float ms = 0.0f;
          DVR_HIP(hipEventElapsedTime(&ms, e.first, e.second));
          ms3[ph] += ms;
        }
      stats->transform_ms = ms3[0];
      stats->scale_ms = ms3[1];
      stats->trans_ms = ms3[2];
      stats->ntransform = (int32_t)ev[0].size();
      stats->ntrans = (int32_t)ev[2].size();
    }
    drop_events();
    (void)hipStreamDestroy(s);
  } catch (const std::exception& e) {
    fprintf(stderr, "dvr_dipole3: %s\n", e.what());
    if (s) {
      (void)hipStreamSynchronize(s);
      (void)hipStreamDestroy(s);
    }
    drop_events();
    return DVR_ERR_HIP;
  }
  return DVR_OK;
}
