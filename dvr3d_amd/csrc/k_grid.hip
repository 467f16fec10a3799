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

// One thread per (g, column) of cint.  Consecutive threads walk g with b
// fastest, matching the column-major cint layout, so the stores coalesce.
__global__ void __launch_bounds__(256)
k_cint(const double* __restrict__ v1, const double* __restrict__ c2,
       double* __restrict__ cint, const CBlock* __restrict__ cb,
       const int* __restrict__ blk3, const int* __restrict__ iv1off, int na,
       int nb, int n3) {
  const long G = (long)na * nb;
  const long total = G * n3;
  for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (long)gridDim.x * blockDim.x) {
    const int col = (int)(idx / G);
    const int g = (int)(idx % G);
    const int a = g / nb;
    const int b = g % nb;
    const int c = blk3[col];
    const CBlock B = cb[c];
    const int j = col - B.col0;
    const int off = iv1off[c * (nb + 1) + b];
    const int cnt = iv1off[c * (nb + 1) + b + 1] - off;
    const double* vr = v1 + (long)(B.rowbase + off) * na + a;
    const double* cr = c2 + B.z2off + off + (long)j * B.nham2;
    double sum = 0.0;
    for (int k = 0; k < cnt; ++k) sum += cr[k] * vr[(long)k * na];
    cint[idx] = sum;
  }
}

void launch_cint(hipStream_t s, const double* v1, const double* c2,
                 double* cint, const CBlock* cb, const int* blk3,
                 const int* iv1off, int na, int nb, int n3) {
  const long total = (long)na * nb * n3;
  if (total <= 0) return;
  int grid = (int)((total + 255) / 256);
  if (grid > 65535 * 16) grid = 65535 * 16;
  DVR_LAUNCH("cint", 8.0 * total, 0.0, k_cint, dim3(grid), dim3(256), 0, s, v1, c2, cint, cb,
                   blk3, iv1off, na, nb, n3);
  DVR_LAUNCH_CHECK();
}

// Wavefunction amplitudes on the DVR grid (trans, D:3650-3845): for every
// c block, the (a, b) grid values of the levels are cint_c (G x iv2_c,
// g = a nb + b) times the block's rows of the 3D eigenvectors, one FP64 MFMA
// GEMM per block into T[g + G (c + nc l)]; a tiled scatter then writes them
// in the IWAVE record layout psi(idvr, npnt1, npnt2) per level:
// theta-first with zr2r1 -> (c = r1, b = r2), else (c = r2, b = r1).
// T (G x nc) x l: 64 x 64 tiles of (a, b) through LDS, read along b and
// written along a; blocks without chosen functions stay zero.
__global__ void __launch_bounds__(256)
k_trans_scatter(const double* __restrict__ T, int na, int nb, int nc, long nlev,
                const CBlock* __restrict__ cb, double* __restrict__ psi, long npsi, long sb,
                long sc) {
  __shared__ double tl[64][65];
  const long G = (long)na * nb;
  const int nta = (na + 63) / 64;
  const int a0 = (blockIdx.x % nta) * 64, b0 = (blockIdx.x / nta) * 64;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  for (long y = blockIdx.y; y < (long)nc * nlev; y += gridDim.y) {
    const int c = (int)(y % nc);
    const long l = y / nc;
    if (cb[c].iv2 <= 0) continue;       // uniform over the work-group
    const double* src = T + G * y;
    for (int i = ty; i < 64; i += 4) {
      const int a = a0 + i, b = b0 + tx;
      if (a < na && b < nb) tl[i][tx] = src[(long)a * nb + b];
    }
    __syncthreads();
    double* dst = psi + npsi * l + sc * c;
    for (int j = ty; j < 64; j += 4) {
      const int a = a0 + tx, b = b0 + j;
      if (a < na && b < nb) dst[a + sb * b] = tl[tx][j];
    }
    __syncthreads();
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
  const unsigned gx = (unsigned)(((na + 63) / 64) * ((nb + 63) / 64));
  const unsigned gy = (unsigned)std::min<long>((long)nc * nvec, 65535);
  DVR_LAUNCH("trans_scatter", 16.0 * G * nc * nvec, 0.0, k_trans_scatter, dim3(gx, gy),
             dim3(256), 0, s, (const double*)T, na, nb, nc, (long)nvec, cb, psi, npsi, sb, sc);
  DVR_LAUNCH_CHECK();
}

}  // namespace dvr
