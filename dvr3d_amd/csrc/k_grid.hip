// Streaming kernels over the DVR grid: potential evaluation, packing of the
// kept 1D eigenpairs, expansion of 2D vectors on the grid (mkham3's cint,
// source/dvr3drjz_segmented.f90:2362-2411) and the wavefunction transform
// (trans, source/dvr3drjz_segmented.f90:3650-3845).
#include "dvr_internal.h"

namespace dvr {

__global__ void __launch_bounds__(256)
k_pes_ab2(double* __restrict__ v, const double* __restrict__ rc,
          const double* __restrict__ rb, const double* __restrict__ xa, int nc,
          int nb, int na, int zr2r1, double de, double aa, double re,
          double kb, double cte, double cc) {
  // no fused multiply-add in this kernel: products and sums are rounded one
  // by one, as in the reference's Fortran evaluation, and V comes out
  // symmetric in r1 <-> r2 bit for bit
#pragma clang fp contract(off)
  const long total = (long)nc * nb * na;
  for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (long)gridDim.x * blockDim.x) {
    const int a = (int)(idx % na);
    const long cb = idx / na;
    const int b = (int)(cb % nb);
    const int c = (int)(cb / nb);
    const double r1 = zr2r1 ? rc[c] : rb[b];
    const double r2 = zr2r1 ? rb[b] : rc[c];
    const double x = xa[a];
    const double t1 = 1.0 - exp(-aa * (r1 - re));
    const double t2 = 1.0 - exp(-aa * (r2 - re));
    const double msum = t1 * t1 + t2 * t2;
    const double dx = x - cte;
    v[idx] = de * msum + kb * (dx * dx) + cc * ((r1 - re) * (r2 - re));
  }
}

void launch_pes_ab2(hipStream_t s, double* v, const double* rc,
                    const double* rb, const double* xa, int nc, int nb, int na,
                    int zr2r1, const double par[6]) {
  const long total = (long)nc * nb * na;
  int grid = (int)((total + 255) / 256);
  if (grid > 65535 * 8) grid = 65535 * 8;
  DVR_LAUNCH("pes", 8.0 * total, 0.0, k_pes_ab2, dim3(grid), dim3(256), 0, s, v, rc, rb, xa, nc,
                   nb, na, zr2r1, par[0], par[1], par[2], par[3], par[4],
                   par[5]);
  DVR_LAUNCH_CHECK();
}

__global__ void __launch_bounds__(256)
k_pack_1d(const double* __restrict__ v1full, const double* __restrict__ e1full,
          const int* __restrict__ src_row, int rows, int n,
          double* __restrict__ v1, double* __restrict__ e1p) {
  const long total = (long)rows * n;
  for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (long)gridDim.x * blockDim.x) {
    const int r = (int)(idx / n);
    const int a = (int)(idx % n);
    const long src = src_row[r];
    v1[idx] = v1full[src * n + a];
    if (a == 0) e1p[r] = e1full[src];
  }
}

void launch_pack_1d(hipStream_t s, const double* v1full, const double* e1full,
                    const int* src_row, int rows, int n, double* v1,
                    double* e1p) {
  if (rows <= 0) return;
  const long total = (long)rows * n;
  int grid = (int)((total + 255) / 256);
  if (grid > 65535 * 8) grid = 65535 * 8;
  DVR_LAUNCH("pack1d", 16.0 * total, 0.0, k_pack_1d, dim3(grid), dim3(256), 0, s, v1full, e1full,
                   src_row, rows, n, v1, e1p);
  DVR_LAUNCH_CHECK();
}

// cint (mkham3, D:2362-2411): the chosen 2D vectors of every c block on its
// (a, b) grid, row g = b na + a.  For each (c, b) the rows b na .. b na + na - 1
// of the block's columns are the product of the kept 1D vectors of (c, b)
// (na x cnt, rows of v1) and rows off .. off + cnt - 1 of the block's 2D
// vectors (cnt x iv2_c): one work-group per (64 columns, b, c), both factors
// staged through LDS in chunks of CI_K 1D levels, each thread one a and 16
// columns per 64-row slice of a, every element written once, coalesced along
// a.  The row order is free: the assembly and the refinement only contract
// over g, and the wavefunction scatter reads it back in this order.
constexpr int CI_K = 32;

__global__ void __launch_bounds__(256)
k_cint(const double* __restrict__ v1, const double* __restrict__ c2,
       double* __restrict__ cint, const CBlock* __restrict__ cb,
       const int* __restrict__ iv1off, int na, int nb) {
  __shared__ double As[CI_K][64];
  __shared__ double Bs[CI_K][65];
  const int c = blockIdx.z, b = blockIdx.y;
  const CBlock B = cb[c];
  const int j0 = blockIdx.x * 64;
  if (j0 >= B.iv2) return;                       // uniform over the work-group
  const int jn = min(64, B.iv2 - j0);
  const int off = iv1off[c * (nb + 1) + b];
  const int cnt = iv1off[c * (nb + 1) + b + 1] - off;
  const long G = (long)na * nb;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const double* vrow = v1 + (long)(B.rowbase + off) * na;        // 1D vector k at vrow + k na
  const double* c2b = c2 + B.z2off + off + (long)j0 * B.nham2;   // (k, j) at c2b + k + j nham2
  double* out = cint + (long)b * na + G * (B.col0 + j0);         // (a, j) at out + a + j G
  for (int a0 = 0; a0 < na; a0 += 64) {
    double acc[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) acc[q] = 0.0;
    for (int k0 = 0; k0 < cnt; k0 += CI_K) {
      const int kc = min(CI_K, cnt - k0);
      __syncthreads();
      for (int x = t; x < CI_K * 64; x += 256) {
        const int kk = x >> 6, aa = x & 63;
        As[kk][aa] = (kk < kc && a0 + aa < na) ? vrow[(long)(k0 + kk) * na + a0 + aa] : 0.0;
        const int kb = x % CI_K, jb = x / CI_K;
        Bs[kb][jb] = (kb < kc && jb < jn) ? c2b[k0 + kb + (long)jb * B.nham2] : 0.0;
      }
      __syncthreads();
      for (int kk = 0; kk < kc; ++kk) {
        const double av = As[kk][lane];
#pragma unroll
        for (int q = 0; q < 16; ++q) acc[q] = fma(av, Bs[kk][w * 16 + q], acc[q]);
      }
    }
    if (a0 + lane < na) {
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int j = w * 16 + q;
        if (j < jn) out[a0 + lane + (long)j * G] = acc[q];
      }
    }
  }
}

void launch_cint(hipStream_t s, const double* v1, const double* c2, double* cint,
                 const CBlock* cb, const std::vector<CBlock>& hcb, const int* iv1off, int na,
                 int nb) {
  int jt = 0, nc = (int)hcb.size();
  double fl = 0.0;
  for (const CBlock& B : hcb) {
    jt = std::max(jt, (B.iv2 + 63) / 64);
    fl += 2.0 * na * (double)B.nham2 * B.iv2;
  }
  if (jt == 0 || nb <= 0 || nc <= 0) return;
  long n3 = 0;
  for (const CBlock& B : hcb) n3 += B.iv2;
  DVR_LAUNCH("cint", 8.0 * (double)na * nb * n3, fl, k_cint, dim3(jt, nb, nc), dim3(256), 0, s,
             v1, c2, cint, cb, iv1off, na, nb);
  DVR_LAUNCH_CHECK();
}

// Wavefunction amplitudes on the DVR grid (trans, D:3650-3845): for every
// c block, the grid values of the levels are cint_c (G x iv2_c, g = b na + a)
// times the block's rows of the 3D eigenvectors, one FP64 MFMA GEMM per block
// into T[g + G (c + nc l)]; the scatter then copies each run of na values
// (contiguous in a on both sides) into the IWAVE record layout
// psi(idvr, npnt1, npnt2) per level: theta-first with zr2r1 -> (c = r1,
// b = r2), else (c = r2, b = r1).  Blocks without chosen functions stay zero.
__global__ void __launch_bounds__(256)
k_trans_scatter(const double* __restrict__ T, int na, int nb, int nc, long nlev,
                const CBlock* __restrict__ cb, double* __restrict__ psi, long npsi, long sb,
                long sc) {
  const long G = (long)na * nb;
  const long total = G * nc * nlev;
  for (long x = blockIdx.x * 256L + threadIdx.x; x < total; x += (long)gridDim.x * 256) {
    const long y = x / G;                 // c + nc l
    const long g = x - y * G;             // b na + a
    const int c = (int)(y % nc);
    if (cb[c].iv2 <= 0) continue;
    const long l = y / nc;
    const int b = (int)(g / na), a = (int)(g - (long)b * na);
    psi[npsi * l + sc * c + sb * b + a] = T[x];
  }
}

void launch_trans(hipStream_t s, const double* cint, const double* c3, double* psi,
                  const CBlock* cb, const std::vector<CBlock>& hcb, int nc, int na, int nb,
                  int n3, int nvec, int zr2r1, int idvr, int npnt1, int npnt2, DevBuf& tmp) {
  const long G = (long)na * nb;
  if (G <= 0 || nvec <= 0 || nc <= 0) return;
  tmp.ensure(sizeof(double) * (size_t)G * nc * nvec);
  double* T = tmp.as<double>();
  for (int c = 0; c < nc; ++c) {
    const CBlock& B = hcb[c];
    if (B.iv2 <= 0) continue;
    launch_gemm(s, "trans", false, false, (int)G, nvec, B.iv2, 1.0, cint + G * B.col0, G,
                c3 + B.col0, n3, 0.0, T + G * c, G * nc, 1, nullptr, false);
  }
  const long npsi = (long)idvr * npnt1 * npnt2;
  const long sb = zr2r1 ? (long)idvr * npnt1 : (long)idvr;
  const long sc = zr2r1 ? (long)idvr : (long)idvr * npnt1;
  const long tot = G * nc * nvec;
  DVR_LAUNCH("trans_scatter", 16.0 * tot, 0.0, k_trans_scatter,
             dim3((unsigned)std::min<long>((tot + 255) / 256, 65535L * 16)), dim3(256), 0, s,
             (const double*)T, na, nb, nc, (long)nvec, cb, psi, npsi, sb, sc);
  DVR_LAUNCH_CHECK();
}

}  // namespace dvr
