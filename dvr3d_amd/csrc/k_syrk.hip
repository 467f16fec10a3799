// Grouped symmetric rank-K products with a block-constant scale, on FP64 MFMA.
//
//   C = (X^T X) .* S + diag(dg),   S(i,j) = xk(blk(i), blk(j))
//
// This is the shape of both contraction stages of the theta-first path:
//  * mkham2 (source/dvr3drjz_segmented.f90:2228-2298): X = the kept 1D
//    eigenvectors of one r1 block (K = npnta angles, N = nham2 functions),
//    blk = the r2 point of each function, xk = xk2, dg = 1D eigenvalues;
//  * mkham3 loop 81 + loadh (D:2459-2463, D:2554-2609): X = cint (K = npnta *
//    npntb grid points, N = nham3), blk = the r1 block, xk = xk1,
//    dg = the chosen 2D eigenvalues.
// Only lower-triangular 64x64 tiles are computed; each writes its block and
// the mirrored block.  Work-group = 4 wavefronts, each a 32x32 sub-tile of
// 2x2 v_mfma_f64_16x16x4f64 accumulators, operands staged through LDS.
// Lane layout of v_mfma_f64_16x16x4f64 (checked on gfx950): A(i,k) comes from
// lane i + 16k, B(k,j) from lane j + 16k, and result register r of lane l is
// D(4r + l/16, l%16).
#include "dvr_internal.h"

namespace dvr {

namespace {
constexpr int TM = 64;   // tile edge
constexpr int KT = 16;   // k depth per LDS stage
constexpr int LDT = TM + 1;
typedef double v4d __attribute__((ext_vector_type(4)));
}  // namespace

int syrk_tile() { return TM; }

__global__ void __launch_bounds__(256)
k_syrk_scaled(const SyrkDesc* __restrict__ desc, int ngroup,
              const double* __restrict__ xk, int ldxk) {
  __shared__ double As[KT * LDT];
  __shared__ double Bs[KT * LDT];
  const int bid = blockIdx.x;
  int g = 0;
  while (g + 1 < ngroup && desc[g + 1].tile0 <= bid) ++g;
  const SyrkDesc D = desc[g];
  const int t = bid - D.tile0;
  if (t >= D.ntile) return;
  // t -> (ti, tj), ti >= tj
  int ti = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
  while ((ti + 1) * (ti + 2) / 2 <= t) ++ti;
  while (ti * (ti + 1) / 2 > t) --ti;
  const int tj = t - ti * (ti + 1) / 2;
  const int row0 = ti * TM, col0 = tj * TM;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wy = wave >> 1, wx = wave & 1;
  const int l16 = lane & 15, lk = lane >> 4;

  v4d acc[2][2];
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) acc[a][b] = v4d{0.0, 0.0, 0.0, 0.0};

  const double* X = D.X;
  const long ldx = D.ldx;
  const int K = D.K, N = D.N;
  for (int k0 = 0; k0 < K; k0 += KT) {
    for (int idx = tid; idx < KT * TM; idx += 256) {
      const int kk = idx % KT;
      const int cc = idx / KT;
      const int k = k0 + kk;
      const int ri = row0 + cc, cj = col0 + cc;
      As[kk * LDT + cc] = (k < K && ri < N) ? X[k + (long)ri * ldx] : 0.0;
      Bs[kk * LDT + cc] = (k < K && cj < N) ? X[k + (long)cj * ldx] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int k4 = 0; k4 < KT; k4 += 4) {
      const int kr = (k4 + lk) * LDT;
      double a0 = As[kr + wy * 32 + l16];
      double a1 = As[kr + wy * 32 + 16 + l16];
      double b0 = Bs[kr + wx * 32 + l16];
      double b1 = Bs[kr + wx * 32 + 16 + l16];
      acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
      acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
      acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
      acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
    }
    __syncthreads();
  }

  double* C = D.C;
  const int* blk = D.blk;
  const double* dg = D.diag;
  for (int bi = 0; bi < 2; ++bi)
    for (int bj = 0; bj < 2; ++bj)
      for (int r = 0; r < 4; ++r) {
        const int i = row0 + wy * 32 + bi * 16 + 4 * r + lk;
        const int j = col0 + wx * 32 + bj * 16 + l16;
        if (i >= N || j >= N) continue;
        {
#pragma clang fp contract(off)
          // (sum * xk) + diag, rounded one by one as in mkham2/loadh
          double v = acc[bi][bj][r] * xk[blk[i] + (long)ldxk * blk[j]];
          if (i == j && dg) v = v + dg[i];
          C[i + (long)N * j] = v;
          if (ti != tj) C[j + (long)N * i] = v;
        }
      }
}

// ---------------------------------------------------------------------------
// Split-precision variant: each FP64 operand x is split into hi = fl32(x) and
// lo = fl32(x - hi) while it is staged to LDS; the tile product is
// hi.hi + hi.lo + lo.hi on v_mfma_f32_16x16x4f32 with FP32 accumulation
// inside each 16-deep k-slice, and the slices are summed in FP64.  Result
// register r of lane l of the FP32 MFMA is D(4 (l/16) + r, l%16).
// ---------------------------------------------------------------------------
typedef float v4f __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(256)
k_syrk_split32(const SyrkDesc* __restrict__ desc, int ngroup, const double* __restrict__ xk,
               int ldxk) {
  __shared__ float Ah[KT * LDT], Al[KT * LDT], Bh[KT * LDT], Bl[KT * LDT];
  const int bid = blockIdx.x;
  int g = 0;
  while (g + 1 < ngroup && desc[g + 1].tile0 <= bid) ++g;
  const SyrkDesc D = desc[g];
  const int t = bid - D.tile0;
  if (t >= D.ntile) return;
  int ti = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
  while ((ti + 1) * (ti + 2) / 2 <= t) ++ti;
  while (ti * (ti + 1) / 2 > t) --ti;
  const int tj = t - ti * (ti + 1) / 2;
  const int row0 = ti * TM, col0 = tj * TM;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wy = wave >> 1, wx = wave & 1;
  const int l16 = lane & 15, lk = lane >> 4;
  double acc[2][2][4];
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      for (int r = 0; r < 4; ++r) acc[a][b][r] = 0.0;
  const double* X = D.X;
  const long ldx = D.ldx;
  const int K = D.K, N = D.N;
  for (int k0 = 0; k0 < K; k0 += KT) {
    for (int idx = tid; idx < KT * TM; idx += 256) {
      const int kk = idx % KT, cc = idx / KT;
      const int k = k0 + kk;
      const int ri = row0 + cc, cj = col0 + cc;
      const double xa = (k < K && ri < N) ? X[k + (long)ri * ldx] : 0.0;
      const double xb = (k < K && cj < N) ? X[k + (long)cj * ldx] : 0.0;
      const float ah = (float)xa, bh = (float)xb;
      Ah[kk * LDT + cc] = ah;
      Al[kk * LDT + cc] = (float)(xa - (double)ah);
      Bh[kk * LDT + cc] = bh;
      Bl[kk * LDT + cc] = (float)(xb - (double)bh);
    }
    __syncthreads();
    v4f f[2][2];
    for (int a = 0; a < 2; ++a)
      for (int b = 0; b < 2; ++b) f[a][b] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int k4 = 0; k4 < KT; k4 += 4) {
      const int kr = (k4 + lk) * LDT;
      float ah[2], al[2], bh[2], bl[2];
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        ah[u] = Ah[kr + wy * 32 + u * 16 + l16];
        al[u] = Al[kr + wy * 32 + u * 16 + l16];
        bh[u] = Bh[kr + wx * 32 + u * 16 + l16];
        bl[u] = Bl[kr + wx * 32 + u * 16 + l16];
      }
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
          f[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(ah[a], bh[b], f[a][b], 0, 0, 0);
          f[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(ah[a], bl[b], f[a][b], 0, 0, 0);
          f[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(al[a], bh[b], f[a][b], 0, 0, 0);
        }
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int r = 0; r < 4; ++r) acc[a][b][r] += (double)f[a][b][r];
    __syncthreads();
  }
  double* C = D.C;
  const int* blk = D.blk;
  const double* dg = D.diag;
  for (int bi = 0; bi < 2; ++bi)
    for (int bj = 0; bj < 2; ++bj)
      for (int r = 0; r < 4; ++r) {
        const int i = row0 + wy * 32 + bi * 16 + 4 * lk + r;
        const int j = col0 + wx * 32 + bj * 16 + l16;
        if (i >= N || j >= N) continue;
        double v = acc[bi][bj][r] * xk[blk[i] + (long)ldxk * blk[j]];
        if (i == j && dg) v = v + dg[i];
        C[i + (long)N * j] = v;
        if (ti != tj) C[j + (long)N * i] = v;
      }
}

void launch_syrk_split32(hipStream_t s, const SyrkDesc* d_desc, int ngroup, int total_tiles,
                         const double* xk, int ldxk, double bytes, double flops) {
  if (total_tiles <= 0) return;
  DVR_LAUNCH("syrk_split32", bytes, flops, k_syrk_split32, dim3(total_tiles), dim3(256), 0, s,
             d_desc, ngroup, xk, ldxk);
  DVR_LAUNCH_CHECK();
}

// Scale and mirror after a lower-tile product C = X^T X (one group): for
// i >= j, C(i,j) <- C(i,j) * xk(blk(i), blk(j)) (+ dg(i) on the diagonal),
// rounded one by one as in mkham2/loadh, and C(j,i) <- C(i,j).  One
// work-group per lower 64x64 tile; the mirrored tile is stored through LDS
// so that both stores run along columns.
__global__ void __launch_bounds__(256)
k_scale_mirror(double* __restrict__ C, int n, const double* __restrict__ xk, int ldxk,
               const int* __restrict__ blk, const double* __restrict__ dg) {
  __shared__ double T[TM * LDT];
  const int t = blockIdx.x;
  int ti = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
  while ((ti + 1) * (ti + 2) / 2 <= t) ++ti;
  while (ti * (ti + 1) / 2 > t) --ti;
  const int tj = t - ti * (ti + 1) / 2;
  const int row0 = ti * TM, col0 = tj * TM;
  const int tid = threadIdx.x;
  for (int e = tid; e < TM * TM; e += 256) {
    const int r = e & (TM - 1), c = e / TM;
    const int i = row0 + r, j = col0 + c;
    if (i < n && j < n && i >= j) {
#pragma clang fp contract(off)
      double v = C[i + (long)n * j] * xk[blk[i] + (long)ldxk * blk[j]];
      if (i == j && dg) v = v + dg[i];
      C[i + (long)n * j] = v;
      T[r * LDT + c] = v;
    }
  }
  __syncthreads();
  for (int e = tid; e < TM * TM; e += 256) {
    const int a = e & (TM - 1), b = e / TM;
    const int row = col0 + a, col = row0 + b;   // C(row, col) = C(col, row)
    if (row < n && col < n && row < col) C[row + (long)n * col] = T[b * LDT + a];
  }
}

void launch_scale_mirror(hipStream_t s, double* C, int n, const double* xk, int ldxk,
                         const int* blk, const double* dg) {
  const int nt = (n + TM - 1) / TM;
  const int tiles = nt * (nt + 1) / 2;
  if (tiles <= 0) return;
  DVR_LAUNCH("scale_mirror", 8.0 * 2.0 * (double)n * n, 0.0, k_scale_mirror, dim3(tiles),
             dim3(256), 0, s, C, n, xk, ldxk, blk, dg);
  DVR_LAUNCH_CHECK();
}

void launch_syrk_scaled(hipStream_t s, const SyrkDesc* d_desc, int ngroup,
                        int total_tiles, const double* xk, int ldxk,
                        double bytes, double flops) {
  if (total_tiles <= 0) return;
  DVR_LAUNCH("syrk", bytes, flops, k_syrk_scaled, dim3(total_tiles), dim3(256), 0, s, d_desc,
                   ngroup, xk, ldxk);
  DVR_LAUNCH_CHECK();
}

}  // namespace dvr
