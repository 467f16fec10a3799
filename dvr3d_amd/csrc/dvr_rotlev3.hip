// The rotational step ROTLEV3 on one GPU (source/rotlev3.f90, cited as
// R:<line>): given the vibrational functions of every k block of a
// DVR3DRJZ zrot run (IWAVE), build the Coriolis-coupled Hamiltonian over
// the selected functions and solve it for both parity sets.
//
//   1. JTRAN (R:839-904): every block's vibrational vectors from the DVR in
//      theta to the associated Legendre basis, coef = P_sel X with P_sel the
//      block's rows of the stored pleg*sqrt(w) array: one GEMM per block over
//      all of its vectors at once (N = npnt1 npnt2 mvib);
//   2. SOLRT (R:671-836): the Coriolis block between k-1 and k is the sum of
//      rank-one updates cor * coef_k[j, :] coef_{k-1}[i, :]^T over the radial
//      points and the shared l (R:776-810).  The (j, i, cor) triples are
//      listed on the host in the reference's loop order; the rows are
//      gathered (scaled by cor) into two K x m panels and the block is one
//      MFMA GEMM A^T B written straight into the lower triangle of H, then
//      mirrored;
//   3. LOADH (R:907-972) / DGROT (R:1046-1197): the diagonal holds the
//      vibrational levels; the lowest NEVAL eigenpairs of H (set 1) and, for
//      KMIN > 1, of its trailing block without k = 0 (set 2, f parity) come
//      from the batched dense solver;
//   4. DSTORE (R:1297-1323): the eigenvectors back in the FBR basis per k
//      block, d_k = coef_k C_k, one GEMM per block (the vectors DIPOLE3
//      reads).
#include "../../include/dvr3d_amd.h"
#include "dvr_internal.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace dvr {

namespace {

int grid_rows(long n) { return (int)std::min<long>((n + 255) / 256, 16384); }

// P_sel of one block: P[j + nang * a] = pleg[jj_j + (maxleg + 1) * a]
__global__ void __launch_bounds__(256)
k_rot_psel(const double* __restrict__ pleg, int maxleg, int idvr, int jj0, int jdia, int nang,
           double* __restrict__ P) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= nang * idvr) return;
  const int j = t % nang, a = t / nang;
  P[t] = pleg[(long)(jj0 + (j + 1) * jdia) + (long)(maxleg + 1) * a];
}

// A2[t + K a] = w[t] coef2[J[t] + ld2 a] (a < m2); B1[t + K b] = coef1[I[t] + ld1 b] (b < m1)
__global__ void __launch_bounds__(256)
k_rot_gather(const double* __restrict__ coef2, long ld2, int m2, const double* __restrict__ coef1,
             long ld1, int m1, const int* __restrict__ J, const int* __restrict__ I,
             const double* __restrict__ w, long K, double* __restrict__ A2,
             double* __restrict__ B1) {
  const long tot = K * (long)(m2 + m1);
  for (long x = blockIdx.x * 256L + threadIdx.x; x < tot; x += (long)gridDim.x * 256) {
    const long col = x / K, t = x - col * K;
    if (col < m2)
      A2[x] = w[t] * coef2[(long)J[t] + ld2 * col];
    else
      B1[x - K * m2] = coef1[(long)I[t] + ld1 * (col - m2)];
  }
}

__global__ void __launch_bounds__(256)
k_rot_diag(double* __restrict__ H, long ld, const double* __restrict__ e, int n) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) H[(long)i * ld + i] = e[i];
}

struct Terms {
  std::vector<int> J, I;
  std::vector<double> w;
};

// The rank-one terms of the Coriolis block between blocks k-1 (set 1) and
// k (set 2), in the order of the loops R:776-810 (0-based rows).
Terms coriolis_terms(const dvr_rot_dims_t* D, const double* rm2, int k, int jst1, int maxlg1,
                     int jst2, int maxlg2, int nblk1) {
  const double sqrt2 = 1.4142135623731;   // R:693, the reference's constant
  const int jdia = std::max(D->idia, 1);
  const long jjp1 = (long)D->jrot * (D->jrot + 1), kkp1 = (long)k * (k - 1);
  const int n2max = D->zquad2 ? 1 : D->npnt2;
  Terms T;
  long ii = 0, j = 0;
  const long ijump = (jst2 - jst1) / jdia;                 // truncation, as Fortran
  const long jjump = maxlg2 / jdia - maxlg1 / jdia;
  for (int n1 = 1; n1 <= D->npnt2; ++n1)
    for (int m1 = 1; m1 <= D->npnt1; ++m1) {
      ii += ijump;
      for (int l1 = jst2; l1 <= maxlg1; l1 += jdia) {
        const long llp1 = (long)l1 * (l1 + 1);
        double cjlp = -std::sqrt((double)((llp1 - kkp1) * (jjp1 - kkp1)));
        if (k == 1) cjlp = sqrt2 * cjlp;
        ++ii;
        ++j;
        long i = D->zquad2 ? ii : (ii - 1) % nblk1 + 1;
        for (int n2 = 1; n2 <= n2max; ++n2) {
          double cor;
          if (D->zquad2) {
            i = ii;
            cor = cjlp * (D->zembed ? rm2[n1 - 1] : rm2[m1 - 1]);
          } else {
            cor = cjlp * rm2[(long)(n1 - 1) * D->npnt2 + n2 - 1];
          }
          T.J.push_back((int)(j - 1));
          T.I.push_back((int)(i - 1));
          T.w.push_back(cor);
          i += nblk1;
        }
      }
      j += jjump;
    }
  return T;
}

}  // namespace

// First associated Legendre function and count of a block (JTRAN R:879-885).
void rot_angular(int k, int jdia, int ipar, int maxleg, int lincr, int* jstart, int* jj0,
                 int* nang) {
  int js = k, j0 = -jdia;
  if (js % jdia != ipar) {
    ++j0;
    ++js;
  }
  *jstart = js;
  *jj0 = j0;
  *nang = (maxleg + lincr - js) / jdia + 1;
}

}  // namespace dvr

using namespace dvr;

extern "C" int dvr_rotlev3(const dvr_rot_dims_t* D, const dvr_rot_block_t* blk,
                           const double* pleg, const double* dvrvec, const double* evib,
                           const double* rm2, int32_t* jst_nang, double* eval, double* vec,
                           double* dvec) {
  if (!D || !blk || !pleg || !evib || !rm2 || !eval) return DVR_ERR_ARG;
  const int nblk = D->nblk, jdia = std::max(D->idia, 1);
  const long nrad = (long)D->npnt1 * D->npnt2;
  if (nblk < 1 || nblk > 4096 || D->npnt1 < 1 || D->npnt2 < 1 || D->jrot < 0 ||
      (D->kfirst != 0 && D->kfirst != 1) || D->nsets < 1 || D->nsets > 2 ||
      (D->nsets == 2 && nblk < 2) || D->ipar < 0 || D->ipar >= jdia ||
      D->nlim < (D->zquad2 ? (D->zembed ? D->npnt2 : D->npnt1) : D->npnt2 * D->npnt2))
    return DVR_ERR_ARG;
  if (D->nvec && (!vec || !dvec || !dvrvec)) return DVR_ERR_ARG;
  // blocks, offsets, angular ranges
  std::vector<long> off(nblk + 1, 0), poff(nblk + 1, 0), xoff(nblk + 1, 0);
  std::vector<int> jst(nblk), jj0(nblk), nang(nblk), maxlg(nblk);
  for (int b = 0; b < nblk; ++b) {
    const dvr_rot_block_t& B = blk[b];
    if (B.mvib < 0 || B.idvr < 1 || B.maxleg < 0 || B.lincr < 0) return DVR_ERR_ARG;
    const int k = D->kfirst + b;
    rot_angular(k, jdia, D->ipar, B.maxleg, B.lincr, &jst[b], &jj0[b], &nang[b]);
    if (nang[b] < 1 || jj0[b] + jdia < 0 || jj0[b] + nang[b] * jdia > B.maxleg)
      return DVR_ERR_ARG;   // rows outside the stored pleg array
    int m2 = B.maxleg + B.lincr;
    if (m2 % jdia != D->ipar) --m2;                      // R:741-742
    maxlg[b] = m2;
    off[b + 1] = off[b] + B.mvib;
    poff[b + 1] = poff[b] + (long)(B.maxleg + 1) * B.idvr;
    xoff[b + 1] = xoff[b] + (long)B.mvib * B.idvr * nrad;
    if (jst_nang) {
      jst_nang[2 * b] = jst[b];
      jst_nang[2 * b + 1] = nang[b];
    }
  }
  if (xoff[nblk] > 0 && !dvrvec) return DVR_ERR_ARG;
  const int nbass = (int)off[nblk];
  if (nbass < 1) return DVR_ERR_ARG;
  const int ibass1 = nbass - blk[0].mvib;
  const int neval0 = D->neval[0], neval1 = D->nsets == 2 ? D->neval[1] : 0;
  if (neval0 < 1 || neval0 > nbass || (D->nsets == 2 && (neval1 < 1 || neval1 > ibass1)))
    return DVR_ERR_ARG;

  hipStream_t s = nullptr;
  // DVR_ROT_TIMES=1: device time of each phase (events on the stream),
  // printed to stderr at the end of the call
  const char* tv = getenv("DVR_ROT_TIMES");
  const bool times = tv && atoi(tv) != 0;
  std::vector<hipEvent_t> ev_t;
  std::vector<const char*> ev_n;
  auto mark = [&](const char* name) {
    if (!times) return;
    hipEvent_t e;
    DVR_HIP(hipEventCreate(&e));
    DVR_HIP(hipEventRecord(e, s));
    ev_t.push_back(e);
    ev_n.push_back(name);
  };
  try {
    DVR_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    mark("start");
    DevBuf dpl, dx, de, dP, dH, dH1, dA, dB, dJ, dI, dw, work;
    std::vector<DevBuf> coef(nblk);
    // host copies stay alive until the stream has been synchronised
    const std::vector<double> plv(pleg, pleg + poff[nblk]), ev(evib, evib + nbass),
        rmv(rm2, rm2 + D->nlim);
    std::vector<Terms> terms(nblk);
    upload(dpl, plv, s);
    upload(de, ev, s);
    if (xoff[nblk] > 0) {
      dx.ensure(sizeof(double) * xoff[nblk]);
      DVR_HIP(hipMemcpyAsync(dx.p, dvrvec, sizeof(double) * xoff[nblk], hipMemcpyHostToDevice,
                             s));
    }
    mark("upload");
    // ---- JTRAN: coef_b (nang nrad x mvib) = P_sel (nang x idvr) X_b (idvr x nrad mvib)
    for (int b = 0; b < nblk; ++b) {
      const dvr_rot_block_t& B = blk[b];
      if (B.mvib == 0) continue;
      const long nk = (long)nang[b] * nrad;
      dP.ensure(sizeof(double) * nang[b] * B.idvr);
      DVR_LAUNCH("rot:psel", 16.0 * nang[b] * B.idvr, 0.0, k_rot_psel,
                 dim3(grid_rows((long)nang[b] * B.idvr)), dim3(256), 0, s,
                 dpl.as<double>() + poff[b], B.maxleg, B.idvr, jj0[b], jdia, nang[b],
                 dP.as<double>());
      coef[b].ensure(sizeof(double) * nk * B.mvib);
      launch_gemm_var(s, "rot:jtran", false, false, nang[b], (int)(nrad * B.mvib), B.idvr, 1.0,
                      dP.as<double>(), nang[b], dx.as<double>() + xoff[b], B.idvr, 0.0,
                      coef[b].as<double>(), nang[b], false, 16, false, false);
    }
    mark("jtran");
    // ---- H: levels on the diagonal, Coriolis blocks below it
    dH.ensure(sizeof(double) * (size_t)nbass * nbass);
    DVR_HIP(hipMemsetAsync(dH.p, 0, sizeof(double) * (size_t)nbass * nbass, s));
    DVR_LAUNCH("rot:diag", 16.0 * nbass, 0.0, k_rot_diag, dim3((nbass + 255) / 256), dim3(256), 0,
               s, dH.as<double>(), (long)nbass, de.as<double>(), nbass);
    for (int b = 1; b < nblk; ++b) {
      const int m2 = blk[b].mvib, m1 = blk[b - 1].mvib;
      if (m2 == 0 || m1 == 0) continue;
      const int k = D->kfirst + b;
      const int nblk1 = (int)((long)blk[b - 1].idvr * nrad / D->npnt2);
      terms[b] = coriolis_terms(D, rmv.data(), k, jst[b - 1], maxlg[b - 1], jst[b], maxlg[b],
                                nblk1);
      const Terms& T = terms[b];
      const long K = (long)T.J.size();
      if (K == 0) continue;
      const long nk2 = (long)nang[b] * nrad, nk1 = (long)nang[b - 1] * nrad;
      for (long t = 0; t < K; ++t)
        if (T.J[t] < 0 || T.J[t] >= nk2 || T.I[t] < 0 || T.I[t] >= nk1)
          throw HipError("Coriolis term outside a block's FBR rows");
      // the previous block's gather may still read these buffers: the
      // uploads are ordered after it on the stream
      upload(dJ, T.J, s);
      upload(dI, T.I, s);
      upload(dw, T.w, s);
      dA.ensure(sizeof(double) * K * m2);
      dB.ensure(sizeof(double) * K * m1);
      DVR_LAUNCH("rot:gather", 16.0 * K * (m1 + m2), 0.0, k_rot_gather,
                 dim3(grid_rows(K * (m1 + m2))), dim3(256), 0, s, coef[b].as<double>(), nk2, m2,
                 coef[b - 1].as<double>(), nk1, m1, dJ.as<int>(), dI.as<int>(), dw.as<double>(),
                 K, dA.as<double>(), dB.as<double>());
      const int ks = gemm_auto_ksplit(m2, m1, (int)std::min<long>(K, 1L << 30));
      work.ensure(sizeof(double) * std::max<size_t>(1, gemm_work_doubles(m2, m1, ks)));
      launch_gemm(s, "rot:coriolis", true, false, m2, m1, (int)K, 1.0, dA.as<double>(), K,
                  dB.as<double>(), K, 0.0, dH.as<double>() + off[b] + (long)nbass * off[b - 1],
                  nbass, ks, work.as<double>(), false);
    }
    // the dense solver reads both triangles
    launch_mirror_lower(s, dH.as<double>(), nbass);
    mark("coriolis");
    // the f set (KMIN > 1) is the trailing block without k = 0; copied
    // before the solve of set 1 overwrites H
    if (D->nsets == 2) {
      dH1.ensure(sizeof(double) * (size_t)ibass1 * ibass1);
      const long m0 = off[1];
      DVR_HIP(hipMemcpy2DAsync(dH1.p, sizeof(double) * ibass1,
                               dH.as<double>() + m0 + (long)nbass * m0, sizeof(double) * nbass,
                               sizeof(double) * ibass1, ibass1, hipMemcpyDeviceToDevice, s));
    }
    // ---- DGROT and DSTORE per set
    long eoff = 0, voff = 0, doff = 0;
    for (int set = 0; set < D->nsets; ++set) {
      const int n = set == 0 ? nbass : ibass1;
      const int nev = set == 0 ? neval0 : neval1;
      const int b0 = set;                       // first block of the set
      double* Hs = set == 0 ? dH.as<double>() : dH1.as<double>();
      EigWork w;
      TwoStage ts;
      DevBuf cp, Z, dd;
      const int want = std::min(n, nev + 16);
      bool two = false;
      if (n >= 4 * twostage_bandwidth() && (D->nvec ? env_twostage_vec() : env_twostage())) {
        ts.keep = D->nvec != 0;
        two = twostage_values(Hs, n, want, s, ts, w, cp);
      } else {
        w.tridiag(std::vector<int>{n}, Hs, std::vector<long>{0}, s);
        w.values(std::vector<int>{want}, s);
      }
      std::memcpy(eval + eoff, w.hval.data(), sizeof(double) * nev);
      eoff += nev;
      if (!D->nvec) {
        mark(set == 0 ? "solve1" : "solve2");
        continue;
      }
      const int mx = w.extend(0, nev);
      Z.ensure(sizeof(double) * (size_t)n * mx);
      if (two)
        twostage_vectors(n, mx, Z.as<double>(), s, ts, w);
      else
        w.vectors(std::vector<int>{mx}, Z.as<double>(), std::vector<long>{0}, s);
      mark(set == 0 ? "solve1" : "solve2");
      DVR_HIP(hipMemcpy2DAsync(vec + voff, sizeof(double) * n, Z.p, sizeof(double) * n,
                               sizeof(double) * n, nev, hipMemcpyDeviceToHost, s));
      voff += (long)n * nev;
      long row = 0;
      for (int b = b0; b < nblk; ++b) {
        const long nk = (long)nang[b] * nrad;
        dd.ensure(sizeof(double) * nk * nev);
        if (blk[b].mvib > 0) {
          launch_gemm_var(s, "rot:dstore", false, false, (int)nk, nev, blk[b].mvib, 1.0,
                          coef[b].as<double>(), nk, Z.as<double>() + row, n, 0.0,
                          dd.as<double>(), nk, false, 16, false, false);
        } else {
          DVR_HIP(hipMemsetAsync(dd.p, 0, sizeof(double) * nk * nev, s));
        }
        DVR_HIP(hipMemcpyAsync(dvec + doff, dd.p, sizeof(double) * nk * nev,
                               hipMemcpyDeviceToHost, s));
        // the buffer is reused by the next block: wait for the copy
        DVR_HIP(hipStreamSynchronize(s));
        doff += nk * nev;
        row += blk[b].mvib;
      }
      DVR_HIP(hipStreamSynchronize(s));
      mark(set == 0 ? "dstore1" : "dstore2");
    }
    DVR_HIP(hipStreamSynchronize(s));
    for (size_t i = 1; i < ev_t.size(); ++i) {
      float ms = 0.0f;
      DVR_HIP(hipEventElapsedTime(&ms, ev_t[i - 1], ev_t[i]));
      fprintf(stderr, "dvr_rotlev3: %-9s %10.3f ms\n", ev_n[i], ms);
    }
    for (auto e : ev_t) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(s);
  } catch (const std::exception& e) {
    fprintf(stderr, "dvr_rotlev3: %s\n", e.what());
    if (s) {
      (void)hipStreamSynchronize(s);
      (void)hipStreamDestroy(s);
    }
    return DVR_ERR_HIP;
  }
  return DVR_OK;
}
