// Symmetrised-Radau driver (IDIA = -2) on one GPU: the body of
// SUBROUTINE nfmain (source/dvr3drjz_segmented.f90:3006-3122, cited as
// D:<line>) for one k block.
//
//   1. every angle's symmetrised radial Hamiltonian (blc2d1, D:3125-3220),
//      all nalf of them built by one kernel;
//   2. all nalf 2D eigenproblems in one batched eigensolve (diag, D:3060);
//   3. selection on the host from the 2D levels (choosr D:3223 / cut2dr
//      D:3323, testiv D:3380);
//   4. the 3D Hamiltonian (bloc3d / blc2d2 / vecmul, D:3455-3595): blc2d2 is
//      diagonal in the symmetrised radial functions, w_p L(g, g') with
//      w_p = (1/2)(1/r_a^2 + 1/r_b^2)/mu, so the block (g, g') of the 3D
//      Hamiltonian is L(g, g') C_g^T C_g' with C_g = diag(sqrt w) V_g: one
//      lower-tile GEMM of C^T C on MFMA, then the L-matrix scale, the 2D
//      levels on the diagonal and the mirror (the kernel of the theta-first
//      path's mkham3);
//   5. the final eigensolve (diag3d, D:2661) and the wavefunction records of
//      transr (D:3599-3646).
#include "../../include/dvr3d_amd.h"
#include "dvr_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

namespace dvr {

namespace {

int grid_n(long n) { return (int)std::min<long>((n + 255) / 256, 16384); }

// H[g] (np x np, column-major, both triangles) of every angle g, element
// (pp, p) as blc2d1 accumulates it: the rotational diagonal (J > 0; the
// r1 = r2 function of each beta receives it twice, D:3159-3175), the kinetic
// terms of both bonds with the parity sign q, the potential at (r_a, r_b)
// and q times at (r_b, r_a), then 1/sqrt(2) per r1 = r2 function.
// Pair p = (pa[p], pb[p]); V[(g*npnt + b)*npnt + a] = V(r_a, r_b, x_g).
__global__ void __launch_bounds__(256)
k_build_h2r(double* __restrict__ H, int np, int nalf, const int* __restrict__ pa,
            const int* __restrict__ pb, const double* __restrict__ hr, int npnt,
            const double* __restrict__ V, const double* __restrict__ xalf,
            const double* __restrict__ r, double q, int jrot, int kz, int ipar, double term,
            double term2, double term3, double urr1) {
#pragma clang fp contract(off)
  const long nn = (long)np * np;
  const long tot = nn * nalf;
  const double f2 = sqrt(0.5);
  for (long t = blockIdx.x * 256L + threadIdx.x; t < tot; t += (long)gridDim.x * 256) {
    const int g = (int)(t / nn);
    const long e = t - (long)g * nn;
    const int pp = (int)(e % np), p = (int)(e / np);
    const int a = pa[p], b = pb[p], ap = pa[pp], bp = pb[pp];
    const double* Vg = V + (long)g * npnt * npnt;
    double h = 0.0;
    if (jrot != 0 && pp == p) {
      const double x = xalf[g];
      double fact = term + term2 / (1.0 - x);
      if (kz == 1) fact = fact + term3 * (1.0 + x) / (1.0 - x);
      const double wsum = (0.5 / (r[a] * r[a]) + 0.5 / (r[b] * r[b])) * fact / urr1;
      h = h + wsum;
      if (ipar == 0 && a == b) h = h + wsum;
    }
    if (b == bp) h = h + hr[ap + (long)a * npnt];
    if (b == ap) h = h + q * hr[bp + (long)a * npnt];
    if (a == bp) h = h + q * hr[ap + (long)b * npnt];
    if (a == ap) h = h + hr[bp + (long)b * npnt];
    if (a == ap && b == bp) h = h + Vg[a + (long)b * npnt];
    if (a == bp && b == ap) h = h + q * Vg[b + (long)a * npnt];
    if (ap == bp) h = h * f2;
    if (a == b) h = h * f2;
    H[(long)g * nn + e] = h;
  }
}

// C[:, col0_g + l] = sqrt(w) .* Z_g[:, l] for the chosen functions, and the
// angle and 2D level of every 3D basis function.
__global__ void __launch_bounds__(256)
k_nf_cmat(const double* __restrict__ Z, const long* __restrict__ zoff,
          const int* __restrict__ colg, const int* __restrict__ coll,
          const double* __restrict__ sw, int np, int n3, double* __restrict__ C) {
  const long tot = (long)np * n3;
  for (long t = blockIdx.x * 256L + threadIdx.x; t < tot; t += (long)gridDim.x * 256) {
    const int p = (int)(t % np), col = (int)(t / np);
    const int g = colg[col], l = coll[col];
    C[t] = sw[p] * Z[zoff[g] + p + (long)l * np];
  }
}

// wvfunc[lev][irs * iang + gi] = sum_l Z_g(irs, l) Z3(col0_g + l, lev) over
// the angles g with functions (gi their position among them), as transr.
__global__ void __launch_bounds__(256)
k_nf_trans(const double* __restrict__ Z, const long* __restrict__ zoff,
           const int* __restrict__ gsel, const int* __restrict__ col0,
           const int* __restrict__ nv, int iang, int np, const double* __restrict__ Z3,
           int n3, int nvec, double* __restrict__ out) {
  const long nb = (long)np * iang;
  const long tot = nb * nvec;
  for (long t = blockIdx.x * 256L + threadIdx.x; t < tot; t += (long)gridDim.x * 256) {
    const int lev = (int)(t / nb);
    const long idx = t - (long)lev * nb;
    const int irs = (int)(idx / iang), gi = (int)(idx % iang);
    const int g = gsel[gi];
    const double* zg = Z + zoff[g] + irs;
    const double* z3 = Z3 + col0[g] + (long)lev * n3;
    double s = 0.0;
    for (int l = 0; l < nv[g]; ++l) s += zg[(long)l * np] * z3[l];
    out[t] = s;
  }
}

}  // namespace
}  // namespace dvr

using namespace dvr;

extern "C" int dvr_nfmain(const dvr_nf_dims_t* D, const double* hr, const double* htheta,
                          const double* r, const double* xalf, double ur1, double urr1,
                          const double* vgrid, double emax2, int32_t* iv2, int32_t* nham3,
                          int32_t* iang_out, double* eval, double* wvfunc) {
  if (!D || !hr || !htheta || !r || !xalf || !vgrid || !iv2 || !nham3 || !eval)
    return DVR_ERR_ARG;
  const int npnt = D->npnt, nalf = D->nalf, ipar = D->ipar;
  if (npnt < 1 || nalf < 1 || (ipar != 0 && ipar != 1) || D->max3d < 1 || D->neval < 1 ||
      D->nvec < 0 || D->nvec > D->neval || (D->nvec > 0 && !wvfunc) || !(ur1 > 0.0))
    return DVR_ERR_ARG;
  const int np = npnt * (npnt + 1 - 2 * ipar) / 2;
  if (np < 1) return DVR_ERR_ARG;
  // J > 0 (rotational terms of blc2d1, testiv) is built below but has no
  // reference run to be checked against: refused until it has
  if (D->jrot != 0) return DVR_ERR_ARG;
  hipStream_t s = nullptr;
  try {
    DVR_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    std::vector<int> pa, pb;
    for (int b = 0; b < npnt; ++b)
      for (int a = 0; a <= b - ipar; ++a) {
        pa.push_back(a);
        pb.push_back(b);
      }
    DevBuf dpa, dpb, dhr, dht, dr, dx, dv, H2;
    // host copies stay alive until the stream has been synchronised
    const std::vector<double> hrv(hr, hr + (size_t)npnt * npnt),
        htv(htheta, htheta + (size_t)nalf * nalf), rv(r, r + npnt), xv(xalf, xalf + nalf),
        vv(vgrid, vgrid + (size_t)npnt * npnt * nalf);
    upload(dpa, pa, s);
    upload(dpb, pb, s);
    upload(dhr, hrv, s);
    upload(dht, htv, s);
    upload(dr, rv, s);
    upload(dx, xv, s);
    upload(dv, vv, s);
    // rotational factors of nfmain (D:3056-3064)
    const int jrot = D->jrot, kz = D->kz;
    double term = 0.0, term2 = 0.0, term3 = 0.0;
    if (jrot != 0) {
      term = (double)(jrot * jrot + jrot - kz * kz) / 8.0;
      term2 = (double)(jrot * jrot + jrot - 3 * kz * kz) / 4.0;
      if (std::abs(kz) == 1) {
        term3 = (double)(jrot * jrot + jrot) / 16.0;
        if (D->kmin >= 1 && D->zrot) term3 = -term3;
      }
    }
    const long nn = (long)np * np;
    H2.ensure(sizeof(double) * nn * nalf);
    DVR_LAUNCH("nf_h2", 8.0 * nn * nalf, 0.0, k_build_h2r, dim3(grid_n(nn * nalf)), dim3(256), 0,
               s, H2.as<double>(), np, nalf, dpa.as<int>(), dpb.as<int>(), dhr.as<double>(), npnt,
               dv.as<double>(), dx.as<double>(), dr.as<double>(), ipar ? -1.0 : 1.0, jrot, kz,
               ipar, term, term2, term3, urr1);
    // ---- all angles' 2D levels (batched eigensolve)
    EigWork w2;
    std::vector<int> nvec(nalf, np);
    std::vector<long> aoff(nalf);
    for (int g = 0; g < nalf; ++g) aoff[g] = (long)g * nn;
    w2.tridiag(nvec, H2.as<double>(), aoff, s);
    w2.values(nvec, s);
    // ---- selection (choosr D:3223-3320 / cut2dr D:3323-3377)
    std::vector<int> nv(nalf, 0);
    if (D->zcut) {
      int tot = 0;
      for (int g = 0; g < nalf; ++g) {
        const double* e = w2.hval.data() + w2.heoff[g];
        int k = 0;
        while (k < np && e[k] < emax2) ++k;
        nv[g] = k;
        tot += k;
      }
      if (tot > D->max3d) {
        (void)hipStreamDestroy(s);
        return DVR_ERR_CORE_3D;
      }
    } else {
      std::vector<double> all(w2.hval.begin(), w2.hval.begin() + (size_t)np * nalf);
      std::sort(all.begin(), all.end());
      const double emax = (long)all.size() <= D->max3d ? all.back() : all[D->max3d - 1];
      for (int g = 0; g < nalf; ++g) {
        const double* e = w2.hval.data() + w2.heoff[g];
        int k = 0;
        while (k < np && e[k] <= emax) ++k;
        nv[g] = k;
      }
    }
    // ---- testiv (D:3380-3452): J > 0 bisector embedding
    if (jrot > 0) {
      int iend = nalf;
      for (int i = nalf / 2; i <= nalf; ++i)
        if (i >= 1 && nv[i - 1] == 0) {
          iend = i;
          break;
        }
      for (int i = iend + 1; i <= nalf; ++i) nv[i - 1] = 0;
      if (nv[nalf - 1] > 0) {
        if (!D->zlin) {
          (void)hipStreamDestroy(s);
          return DVR_ERR_ARG;   // "wavefunction has amplitude for theta = 0"
        }
        nv[nalf - 1] = 0;
      }
    }
    int n3 = 0, iang = 0;
    std::vector<int> col0(nalf, 0), gsel;
    for (int g = 0; g < nalf; ++g) {
      col0[g] = n3;
      n3 += nv[g];
      if (nv[g] > 0) gsel.push_back(g);
    }
    iang = (int)gsel.size();
    for (int g = 0; g < nalf; ++g) iv2[g] = nv[g];
    *nham3 = n3;
    if (iang_out) *iang_out = iang;
    if (n3 <= 0) {
      (void)hipStreamDestroy(s);
      return DVR_ERR_ARG;
    }
    // ---- chosen 2D vectors, C, 3D Hamiltonian
    std::vector<int> mext(nalf);
    std::vector<long> zoff(nalf);
    long zt = 0;
    for (int g = 0; g < nalf; ++g) {
      mext[g] = w2.extend(g, nv[g]);
      zoff[g] = zt;
      zt += (long)np * mext[g];
    }
    DevBuf Z2, dzoff, dcolg, dcoll, dsw, C, H3, de2, dblk;
    Z2.ensure(sizeof(double) * (zt + 1));
    w2.vectors(mext, Z2.as<double>(), zoff, s);
    std::vector<int> colg(n3), coll(n3);
    std::vector<double> e2sel(n3);
    for (int g = 0; g < nalf; ++g)
      for (int l = 0; l < nv[g]; ++l) {
        colg[col0[g] + l] = g;
        coll[col0[g] + l] = l;
        e2sel[col0[g] + l] = w2.hval[w2.heoff[g] + l];
      }
    std::vector<double> sw(np);
    for (int p = 0; p < np; ++p) {
      const double ra = r[pa[p]], rb = r[pb[p]];
      sw[p] = std::sqrt(0.5 / (ra * ra * ur1) + 0.5 / (rb * rb * ur1));
    }
    C.ensure(sizeof(double) * (size_t)np * n3);
    DVR_LAUNCH("nf_cmat", 16.0 * np * (double)n3, 0.0, k_nf_cmat, dim3(grid_n((long)np * n3)),
               dim3(256), 0, s, Z2.as<double>(), upload(dzoff, zoff, s), upload(dcolg, colg, s),
               upload(dcoll, coll, s), upload(dsw, sw, s), np, n3, C.as<double>());
    H3.ensure(sizeof(double) * (size_t)n3 * n3);
    launch_gemm_var(s, "nf_syrk", true, false, n3, n3, np, 1.0, C.as<double>(), np,
                    C.as<double>(), np, 0.0, H3.as<double>(), n3, true, 16, false, true);
    launch_scale_mirror(s, H3.as<double>(), n3, dht.as<double>(), nalf, dcolg.as<int>(),
                        upload(de2, e2sel, s));
    // ---- final eigensolve (diag3d)
    const int meval = std::min(D->neval, n3);
    const int nv3 = std::min(D->nvec, meval);
    EigWork w3;
    TwoStage ts;
    DevBuf cp, Z3;
    const int want = std::min(n3, meval + 16);
    bool two = false;
    if (n3 >= 4 * twostage_bandwidth() && (nv3 > 0 ? env_twostage_vec() : env_twostage())) {
      ts.keep = nv3 > 0;
      two = twostage_values(H3.as<double>(), n3, want, s, ts, w3, cp);
    } else {
      w3.tridiag(std::vector<int>{n3}, H3.as<double>(), std::vector<long>{0}, s);
      w3.values(std::vector<int>{want}, s);
    }
    std::memcpy(eval, w3.hval.data(), sizeof(double) * meval);
    if (nv3 > 0) {
      const int mx = w3.extend(0, nv3);
      Z3.ensure(sizeof(double) * (size_t)n3 * mx);
      if (two)
        twostage_vectors(n3, mx, Z3.as<double>(), s, ts, w3);
      else
        w3.vectors(std::vector<int>{mx}, Z3.as<double>(), std::vector<long>{0}, s);
      DevBuf out, dgs, dc0, dnv;
      const long nb = (long)np * iang;
      out.ensure(sizeof(double) * nb * nv3);
      DVR_LAUNCH("nf_trans", 8.0 * nb * nv3, 0.0, k_nf_trans, dim3(grid_n(nb * nv3)), dim3(256),
                 0, s, Z2.as<double>(), dzoff.as<long>(), upload(dgs, gsel, s),
                 upload(dc0, col0, s), upload(dnv, nv, s), iang, np, Z3.as<double>(), n3, nv3,
                 out.as<double>());
      DVR_HIP(hipMemcpyAsync(wvfunc, out.p, sizeof(double) * nb * nv3, hipMemcpyDeviceToHost, s));
    }
    DVR_HIP(hipStreamSynchronize(s));
    (void)hipStreamDestroy(s);
  } catch (const std::exception& e) {
    fprintf(stderr, "dvr_nfmain: %s\n", e.what());
    if (s) (void)hipStreamDestroy(s);
    return DVR_ERR_HIP;
  }
  return DVR_OK;
}
