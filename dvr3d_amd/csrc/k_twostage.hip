// Two-stage reduction of a dense symmetric matrix to tridiagonal form, FP64,
// for the plain final solve (diag3d -> diag -> dsyev, source/
// dvr3drjz_segmented.f90:2661-2764, 2638-2659; dsyev reduces with DSYTRD,
// source/dsyev.f:6209, one stage, BLAS-2 at the reference's LWORK = 3N).
//
// Stage 1 (band reduction, bandwidth BW = 64), per panel of BW columns:
//   * the panel below the band, P (m x BW), is factored P = Q R by shifted
//     Cholesky QR three times (Gram matrix on MFMA, LDL^T of the 64 x 64
//     Gram block, Q = P M^T), which needs no column-by-column sweep;
//   * Householder reconstruction (modified LU of Q1 - S, Q1 the top block,
//     S = diag(+-1)) turns Q into the compact WY form H = I - V T V^T with
//     H(:, 1:BW) = Q S, so that H^T P = [S R; 0];
//   * the trailing matrix A22 <- H^T A22 H as X = A22 V T, M = T^T V^T X,
//     Y = X - V M / 2, A22 -= V Y^T + Y V^T: one GEMM of width BW for X and
//     one lower-tile rank-2BW GEMM, both on FP64 MFMA; the upper tiles are
//     then mirrored for the next panel's product.
//   All of it is GEMM-shaped: 4/3 n^3 flops on MFMA instead of the one-stage
//   reduction's 4/3 n^3 flops of matrix-vector products.
// Stage 2 (band to tridiagonal) chases the bulges with Householder
// reflections of length BW: sweep s annihilates column s below the
// subdiagonal and pushes the fill-in down the band in tasks of one BW block
// each.  Task k of sweep s runs in launch 2 s + k: tasks in one launch touch
// disjoint blocks (sweep s - 1 is two blocks ahead), so every launch runs all
// of its tasks at once, one work-group of 256 threads each, with the diagonal
// block in registers and the block below it in LDS.
// A panel whose Cholesky QR breaks down (a rank-deficient panel) is reported,
// and the caller then uses the one-stage reduction.
//
// Eigenvectors (dsyev with JOBZ = 'V', the reference's call): inverse
// iteration gives those of the tridiagonal T for the wanted levels only,
// and they are carried back through both stages:
//   * stage 2: the chase keeps every reflector H(s, k) (sweep s, task k,
//     rows s + 1 + k BW .. s + k BW + BW).  A later sweep's reflector that
//     overlaps an earlier one has a task index no larger, so the reflectors
//     of BT consecutive sweeps at one task form a block (g, k), applied in
//     descending sweep order, and blocks on one wavefront L = (G-1-g) + k
//     touch disjoint rows: one launch per wavefront.  The vectors are held
//     transposed (one lane per vector, rows contiguous across lanes), and a
//     lane keeps the BW rows of the current reflector in a register ring
//     that slides up one row per reflector;
//   * stage 1: per panel, in reverse, Y <- Y - V T^T V^T Y in the transposed
//     layout as three FP64 MFMA GEMMs, after the reflectors of the last
//     short panel.
#include "dvr_internal.h"

#include <cmath>
#include <utility>

namespace dvr {

namespace {

constexpr int BW = 64;          // bandwidth of the first stage
constexpr int LDB = 2 * BW;     // rows of the band storage (diagonals 0 .. 2 BW - 1)
constexpr int LAG = 2;          // launches between consecutive sweeps of stage 2
constexpr int LD = BW + 1;      // LDS leading dimension of 64 x 64 blocks

__device__ inline double wsum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return __shfl(v, 0, 64);
}

// block sum over 256 threads (red: 4 doubles of LDS)
__device__ inline double bsum(double v, double* red) {
  v = wsum(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  return red[0] + red[1] + red[2] + red[3];
}

// ---------------------------------------------------------------------------
// Stage 1 kernels
// ---------------------------------------------------------------------------

// After gram_prep + LDL^T of the Gram block (d, Linv, column scales sc):
// M = diag(d)^-1/2 Linv diag(sc), so that Q = Z M^T has orthonormal columns.
// flags[0] |= 1 when a pivot is below 1e-12 on a non-zero column or when
// only some columns are zero (a rank-deficient panel); on the first pass an
// all-zero panel sets zero[0] = 1 (nothing to reduce) instead.
__global__ void __launch_bounds__(256)
k_cholqr_m(const double* __restrict__ d, const double* __restrict__ Linv,
           const double* __restrict__ sc, double* __restrict__ M, int* __restrict__ flags,
           int* __restrict__ zero, int first) {
  __shared__ double r[BW];
  __shared__ int nz, bad;
  const int t = threadIdx.x;
  if (t == 0) {
    nz = 0;
    bad = 0;
  }
  __syncthreads();
  if (t < BW) {
    const bool z = !(sc[t] > 0.0);
    if (z) atomicAdd(&nz, 1);
    if (!z && !(d[t] > 1e-12)) atomicAdd(&bad, 1);
    r[t] = d[t] > 0.0 ? 1.0 / sqrt(d[t]) : 0.0;
  }
  __syncthreads();
  const bool allzero = nz == BW;
  if (t == 0) {
    if (first) zero[0] = allzero ? 1 : 0;
    if (!allzero && (nz > 0 || bad > 0)) flags[0] |= 1;
  }
  for (int idx = t; idx < BW * BW; idx += 256) {
    const int i = idx % BW, j = idx / BW;
    M[idx] = allzero ? 0.0 : r[i] * Linv[idx] * sc[j];
  }
}

// Householder reconstruction from the top block Q1 (ld ldq) of an
// orthonormal Q: modified LU Q1 - S = L1 U without pivoting, s_i = -sign of
// the pivot (so every pivot has modulus >= 1), then
//   V(1:BW, :) = L1 (unit lower), Uinv = U^-1, T = -U S L1^-T, Sv = s.
// The rows below (V2 = Q2 U^-1) are one GEMM by the caller.  A zero panel
// gets T = 0 (H = I), Sv = 1.
__global__ void __launch_bounds__(256)
k_house_rec(const double* __restrict__ Q, long ldq, double* __restrict__ V, long ldv,
            double* __restrict__ Uinv, double* __restrict__ T, double* __restrict__ Sv,
            const int* __restrict__ zero) {
  extern __shared__ double sm[];
  double* W = sm;                 // LU in place
  double* X = W + BW * LD;        // U^-1
  double* Y = X + BW * LD;        // L1^-1
  double* s = Y + BW * LD;        // signs
  const int t = threadIdx.x;
  const bool zp = zero[0] != 0;
  for (int idx = t; idx < BW * BW; idx += 256) {
    const int i = idx % BW, j = idx / BW;
    W[i + j * LD] = Q[i + (long)j * ldq];
  }
  __syncthreads();
  if (zp) {
    for (int idx = t; idx < BW * BW; idx += 256) {
      const int i = idx % BW, j = idx / BW;
      T[idx] = 0.0;
      Uinv[idx] = i == j ? 1.0 : 0.0;
      V[i + (long)j * ldv] = i == j ? 1.0 : 0.0;
    }
    if (t < BW) Sv[t] = 1.0;
    return;
  }
  for (int i = 0; i < BW; ++i) {
    if (t == 0) {
      const double p = W[i + i * LD];
      const double si = p >= 0.0 ? -1.0 : 1.0;
      s[i] = si;
      W[i + i * LD] = p - si;
    }
    __syncthreads();
    const double piv = W[i + i * LD];
    if (t > i && t < BW) W[t + i * LD] /= piv;
    __syncthreads();
    const int nr = BW - 1 - i;
    for (int idx = t; idx < nr * nr; idx += 256) {
      const int k = i + 1 + idx % nr, j = i + 1 + idx / nr;
      W[k + j * LD] -= W[k + i * LD] * W[i + j * LD];
    }
    __syncthreads();
  }
  // U^-1 (column j by thread j) and L1^-1 (column j by thread 64 + j)
  if (t < BW) {
    const int j = t;
    for (int i = BW - 1; i > j; --i) X[i + j * LD] = 0.0;
    X[j + j * LD] = 1.0 / W[j + j * LD];
    for (int i = j - 1; i >= 0; --i) {
      double a = 0.0;
      for (int k = i + 1; k <= j; ++k) a += W[i + k * LD] * X[k + j * LD];
      X[i + j * LD] = -a / W[i + i * LD];
    }
  } else if (t < 2 * BW) {
    const int j = t - BW;
    for (int i = 0; i < j; ++i) Y[i + j * LD] = 0.0;
    Y[j + j * LD] = 1.0;
    for (int i = j + 1; i < BW; ++i) {
      double a = 0.0;
      for (int k = j; k < i; ++k) a += W[i + k * LD] * Y[k + j * LD];
      Y[i + j * LD] = -a;
    }
  }
  __syncthreads();
  // T = -U S L1^-T: T(i, j) = -sum_{k >= i} U(i, k) s_k L1inv(j, k), upper
  for (int idx = t; idx < BW * BW; idx += 256) {
    const int i = idx % BW, j = idx / BW;
    double a = 0.0;
    if (i <= j)
      for (int k = i; k <= j; ++k) a += W[i + k * LD] * s[k] * Y[j + k * LD];
    T[idx] = -a;
    Uinv[idx] = X[i + j * LD];
    V[i + (long)j * ldv] = i > j ? W[i + j * LD] : (i == j ? 1.0 : 0.0);
  }
  if (t < BW) Sv[t] = s[t];
}

// The reduced panel into A: rows j0+BW .. j0+BW+m of columns j0 .. j0+BW:
// the band block S R (upper triangle of the top BW rows) and, for the
// back-transformation, the Householder vectors V below it (strict lower
// part of the top block and all rows below).
__global__ void __launch_bounds__(256)
k_panel_store(double* __restrict__ A, long lda, int j0, int m, const double* __restrict__ R,
              const double* __restrict__ Sv, const double* __restrict__ V) {
  const long n = (long)m * BW;
  for (long t = blockIdx.x * 256L + threadIdx.x; t < n; t += (long)gridDim.x * 256) {
    const int i = (int)(t % m), c = (int)(t / m);
    double v;
    if (i < BW && i <= c) v = Sv[i] * R[i + c * BW];
    else v = V[i + (long)m * c];
    A[(j0 + BW + i) + (long)(j0 + c) * lda] = v;
  }
}

// Upper tiles (64 x 64, strictly above the diagonal) of the m x m block at A
// from the transposed lower tiles, through LDS.  Tile pair q enumerates
// (I, J), I < J, column-wise.
__global__ void __launch_bounds__(256)
k_mirror_tiles(double* __restrict__ A, long lda, int m) {
  __shared__ double tl[BW][BW + 1];
  // decode q -> (I, J), J >= 1, I < J: q = J (J - 1) / 2 + I
  const long q = blockIdx.x;
  int J = (int)((1.0 + sqrt(1.0 + 8.0 * (double)q)) * 0.5);
  while ((long)J * (J - 1) / 2 > q) --J;
  while ((long)(J + 1) * J / 2 <= q) ++J;
  const int I = (int)(q - (long)J * (J - 1) / 2);
  const int r0 = J * BW, c0 = I * BW;       // source tile (J, I): rows r0.., cols c0..
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  for (int c = ty; c < BW; c += 4) {
    const int gi = r0 + tx, gj = c0 + c;
    tl[tx][c] = (gi < m && gj < m) ? A[gi + (long)gj * lda] : 0.0;
  }
  __syncthreads();
  // destination (I, J): element (c0 + a, r0 + b) = source (r0 + b, c0 + a)
  for (int bcol = ty; bcol < BW; bcol += 4) {
    const int gi = c0 + tx, gj = r0 + bcol;
    if (gi < m && gj < m) A[gi + (long)gj * lda] = tl[bcol][tx];
  }
}

// Last panel when fewer than BW rows remain below it (0 < m < BW): one
// work-group, Householder reflections in LDS column by column (the panel P,
// m x BW, becomes upper trapezoidal) and the m x m trailing block updated
// from both sides with each reflection.  The reflectors stay below the
// diagonal of P, their factors in tau.
__global__ void __launch_bounds__(256)
k_last_panel(double* __restrict__ A, long lda, int j0, int m, double* __restrict__ tau) {
  extern __shared__ double lp[];
  double (*P)[BW + 1] = reinterpret_cast<double (*)[BW + 1]>(lp);            // P[row][col]
  double (*C)[BW + 1] = reinterpret_cast<double (*)[BW + 1]>(lp + BW * LD);  // trailing, full
  double* v = lp + 2 * BW * LD;
  double* w = v + BW;
  double* red = w + BW;
  double* sc = red + 4;
  const int t = threadIdx.x;
  const int r0 = j0 + BW;
  for (int idx = t; idx < m * BW; idx += 256) {
    const int i = idx % m, c = idx / m;
    P[i][c] = A[(r0 + i) + (long)(j0 + c) * lda];
  }
  for (int idx = t; idx < m * m; idx += 256) {
    const int i = idx % m, j = idx / m;
    const int a = i >= j ? i : j, b = i >= j ? j : i;
    C[i][j] = A[(r0 + a) + (long)(r0 + b) * lda];
  }
  __syncthreads();
  for (int c = 0; c + 1 < m; ++c) {
    double part = 0.0;
    for (int i = c + 1 + t; i < m; i += 256) part += P[i][c] * P[i][c];
    const double sigma = bsum(part, red);
    if (t == 0) {
      const double alpha = P[c][c];
      double tv = 0.0, beta = alpha, scale = 0.0;
      if (sigma > 0.0) {
        beta = -copysign(sqrt(alpha * alpha + sigma), alpha);
        tv = (beta - alpha) / beta;
        scale = 1.0 / (alpha - beta);
      }
      sc[0] = tv;
      sc[1] = scale;
      P[c][c] = beta;
      tau[c] = tv;
    }
    __syncthreads();
    const double tv = sc[0];
    // v over all m rows of the trailing block, zero above row c
    if (t < m) v[t] = t < c ? 0.0 : (t == c ? 1.0 : P[t][c] * sc[1]);
    __syncthreads();
    if (t > c && t < m) P[t][c] = v[t];
    // left application to the remaining panel columns
    for (int cc = c + 1 + t; cc < BW; cc += 256) {
      double a = 0.0;
      for (int i = c; i < m; ++i) a += v[i] * P[i][cc];
      a *= tv;
      for (int i = c; i < m; ++i) P[i][cc] -= a * v[i];
    }
    // both sides of the whole trailing block (its rows above c are coupled
    // to the rows the reflection mixes)
    if (t < m) {
      double a = 0.0;
      for (int j = c; j < m; ++j) a += C[t][j] * v[j];
      w[t] = tv * a;
    }
    __syncthreads();
    double kp = t < m ? w[t] * v[t] : 0.0;
    const double K = bsum(kp, red);
    const double al = -0.5 * tv * K;
    if (t < m) w[t] += al * v[t];
    __syncthreads();
    for (int idx = t; idx < m * m; idx += 256) {
      const int i = idx % m, j = idx / m;
      C[i][j] -= v[i] * w[j] + w[i] * v[j];
    }
    __syncthreads();
  }
  for (int idx = t; idx < m * BW; idx += 256) {
    const int i = idx % m, c = idx / m;
    A[(r0 + i) + (long)(j0 + c) * lda] = P[i][c];
  }
  for (int idx = t; idx < m * m; idx += 256) {
    const int i = idx % m, j = idx / m;
    if (i >= j) A[(r0 + i) + (long)(r0 + j) * lda] = C[i][j];
  }
}

// ---------------------------------------------------------------------------
// Stage 2 kernels
// ---------------------------------------------------------------------------

// Band storage AB[d + j LDB] = A(j + d, j), d < LDB (zero beyond the band).
__global__ void __launch_bounds__(256)
k_band_extract(const double* __restrict__ A, long lda, int n, double* __restrict__ AB) {
  const long tot = (long)n * LDB;
  for (long t = blockIdx.x * 256L + threadIdx.x; t < tot; t += (long)gridDim.x * 256) {
    const int dd = (int)(t % LDB), j = (int)(t / LDB);
    AB[t] = (dd <= BW && j + dd < n) ? A[(j + dd) + (long)j * lda] : 0.0;
  }
}

// One launch of the bulge chase: work-group i runs task k = t - LAG s of
// sweep s = s0 + i.  Vs[s BW ...], taus[s]: the reflector of the sweep's
// previous task.
//   k = 0: reflector of column s below the subdiagonal;
//   k > 0: the off-diagonal block O (rows r0 .., columns r0 - BW ..) gets
//          the previous reflector from the right, then a reflector of its
//          column 0 from the left;
//   both: the reflector from both sides of the diagonal block at r0.
// Each block is updated in one pass: with y = taup O vp, x = O(:, 0) - y vp0
// (the new column 0), v, tau, beta from x, a = v.y and u = v^T O,
//   O <- O - y vp^T - v w^T,  w_c = tau (u_c - a vp_c)   (c >= 1),
//   D <- D - v z^T - z v^T,   z = tau D v - tau^2 (v.D v) v / 2.
// 256 threads: thread (row r = tid / 4, quarter q = tid % 4) handles the
// 16 columns q, q + 4, ... of row r (or the 16 rows of column r); the
// quarter sums meet through lane shuffles.  All global loads are issued at
// the start.  With V2 set, the reflector is also kept for the
// back-transformation at V2[(k nsw + s) BW ...], its factor at tau2[k nsw + s].
__global__ void __launch_bounds__(256)
k_sb2st_step(double* __restrict__ AB, int n, int t, int s0, double* __restrict__ Vs,
             double* __restrict__ taus, double* __restrict__ V2, double* __restrict__ tau2) {
  extern __shared__ double sm[];
  double* O = sm;                // off-diagonal block (ld LD)
  double* vb = O + BW * LD;      // v
  double* pb = vb + BW;          // vp
  double* wb = pb + BW;          // w, then z
  double* red = wb + BW;         // 10
  const int tid = threadIdx.x, r = tid >> 2, q = tid & 3;
  const int s = s0 + blockIdx.x;
  const int k = t - LAG * s;
  if (s > n - 3 || k < 0) return;
  const int r0 = s + 1 + k * BW;
  if (r0 > n - 1) return;
  const int bb = min(BW, n - r0);
  const int c0 = r0 - BW;
  const bool act = r < bb;
  // ---- loads into registers: diagonal block (row r, both triangles from
  // the lower band), O, vp
  double dreg[16], oreg[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    const int j = q + 4 * i;
    const int a = r >= j ? r : j, b = r >= j ? j : r;   // (a, b) lower
    dreg[i] = (a < bb) ? AB[(a - b) + (long)(r0 + b) * LDB] : 0.0;
    oreg[i] = (k > 0 && act) ? AB[(BW + r - j) + (long)(c0 + j) * LDB] : 0.0;
  }
  double xcol = 0.0;
  if (k == 0 && q == 0 && act) xcol = AB[(1 + r) + (long)s * LDB];
  if (tid < BW) pb[tid] = k > 0 ? Vs[(long)s * BW + tid] : 0.0;
  const double taup = k > 0 ? taus[s] : 0.0;
#pragma unroll
  for (int i = 0; i < 16; ++i) O[r + (q + 4 * i) * LD] = oreg[i];
  __syncthreads();
  // ---- y = taup O vp (row r), x = new column 0
  double y = 0.0;
#pragma unroll
  for (int i = 0; i < 16; ++i) y += oreg[i] * pb[q + 4 * i];
  y += __shfl_xor(y, 1, 64);
  y += __shfl_xor(y, 2, 64);
  y *= taup;
  const int lane = tid & 63, wv = tid >> 6;
  const double x = k == 0 ? __shfl(xcol, lane & ~3, 64) : O[r] - y * pb[0];
  // sigma = sum_{r >= 1} x_r^2, xi = sum_{r >= 1} x_r y_r (one thread per row)
  double s1 = (q == 0 && act && r >= 1) ? x * x : 0.0;
  double s2 = (q == 0 && act && r >= 1) ? x * y : 0.0;
  s1 = wsum(s1);
  s2 = wsum(s2);
  if (lane == 0) {
    red[wv] = s1;
    red[4 + wv] = s2;
  }
  if (tid == 0) {
    red[8] = y;      // y_0
    red[9] = x;      // x_0, the leading entry of the reflected column
  }
  __syncthreads();
  const double sigma = red[0] + red[1] + red[2] + red[3];
  const double xi = red[4] + red[5] + red[6] + red[7];
  const double y0 = red[8], al0 = red[9];
  double tau = 0.0, beta = al0, scale = 0.0;
  if (sigma > 0.0) {
    beta = -copysign(sqrt(al0 * al0 + sigma), al0);
    tau = (beta - al0) / beta;
    scale = 1.0 / (al0 - beta);
  }
  const double v = r == 0 ? 1.0 : (act ? x * scale : 0.0);
  if (q == 0) vb[r] = v;
  if (k == 0 && q == 0 && act) AB[(1 + r) + (long)s * LDB] = r == 0 ? beta : 0.0;
  __syncthreads();
  if (k > 0) {
    const double av = y0 + scale * xi;           // v . y
    // u_c = v^T O(:, c) for column c = r (16 rows per quarter)
    double u = 0.0;
#pragma unroll
    for (int i = 0; i < 16; ++i) u += vb[q + 4 * i] * O[(q + 4 * i) + r * LD];
    u += __shfl_xor(u, 1, 64);
    u += __shfl_xor(u, 2, 64);
    if (q == 0) wb[r] = tau * (u - av * pb[r]);
    __syncthreads();
    if (act) {
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        const int c = q + 4 * i;
        const double val = c == 0 ? (r == 0 ? beta : 0.0) : oreg[i] - y * pb[c] - v * wb[c];
        AB[(BW + r - c) + (long)(c0 + c) * LDB] = val;
      }
    }
    __syncthreads();
  }
  // ---- diagonal block from both sides
  double z = 0.0;
#pragma unroll
  for (int i = 0; i < 16; ++i) z += dreg[i] * vb[q + 4 * i];
  z += __shfl_xor(z, 1, 64);
  z += __shfl_xor(z, 2, 64);
  z *= tau;
  double kv = (q == 0) ? v * z : 0.0;
  kv = wsum(kv);
  if (lane == 0) red[wv] = kv;
  __syncthreads();
  const double K = red[0] + red[1] + red[2] + red[3];
  z -= 0.5 * tau * K * v;
  if (q == 0) wb[r] = z;
  __syncthreads();
  if (act) {
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const int j = q + 4 * i;
      if (j <= r) AB[(r - j) + (long)(r0 + j) * LDB] = dreg[i] - v * wb[j] - z * vb[j];
    }
  }
  if (q == 0) Vs[(long)s * BW + r] = v;
  if (tid == 0) taus[s] = tau;
  if (V2) {
    const long ri = (long)k * (n - 2) + s;
    if (q == 0) V2[ri * BW + r] = v;
    if (tid == 0) tau2[ri] = tau;
  }
}

__global__ void __launch_bounds__(256)
k_band_to_tri(const double* __restrict__ AB, int n, double* __restrict__ d, double* __restrict__ e) {
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    d[i] = AB[(long)i * LDB];
    if (i + 1 < n) e[i] = AB[1 + (long)i * LDB];
  }
}

// ---------------------------------------------------------------------------
// Back-transformation kernels
// ---------------------------------------------------------------------------

constexpr int BT = 64;   // sweeps per stage-2 back-transformation block
constexpr int PF = 4;    // window rows loaded ahead of the ring

static_assert(BW == 64, "the register ring of the back-transformation holds BW = 64 rows");

// Explicit address spaces: row pointers that pass through an empty asm stay
// global (vector memory instructions), and reflector reads through the
// constant space stay scalar loads.
typedef __attribute__((address_space(1))) double gdouble;
typedef const __attribute__((address_space(4))) double kdouble;

// Reflector of sweep s0 + T on window rows T .. T + BW - 1 (ring slots
// (T + j) & 63).  Before it, the ring's bottom row T + BW leaves (stored at
// sp) and row T enters from the prefetch buffer, which then loads row
// T - PF (at lp).  sp and lp walk up one row per step; they pass through an
// empty asm so that the compiler keeps one pointer each in VGPRs instead of
// materialising every row's address.  The reflector is read twice (scalar
// loads, once for the dot product, once for the update) so that its 64
// values need not stay in SGPRs across the reduction.
template <int T>
__device__ __forceinline__ void bt2_step(double (&z)[BW], double (&pre)[PF], gdouble*& sp,
                                         gdouble*& lp, long ldt, int s0, int nsw, long krow,
                                         kdouble* V2, kdouble* tau2) {
  if (T < BT - 1) {
    *sp = z[T & (BW - 1)];
    sp -= ldt;
    asm volatile("" : "+v"(sp));
    z[T & (BW - 1)] = pre[T % PF];
    if (T - PF >= 0) {
      pre[T % PF] = *lp;
      lp -= ldt;
      asm volatile("" : "+v"(lp));
    }
  }
  // slots of sweeps without task k hold v = 0, tau = 0 (zero-filled); past
  // the last sweep the factor is 0 on a valid slot
  const int s = s0 + T;
  const long ri = krow + (s < nsw ? s : nsw - 1);
  const double tau = tau2[ri] * (s < nsw ? 1.0 : 0.0);
  kdouble* v = V2 + ri * BW;
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
#pragma unroll
  for (int j = 0; j < BW; j += 4) {
    a0 = fma(v[j], z[(T + j) & (BW - 1)], a0);
    a1 = fma(v[j + 1], z[(T + j + 1) & (BW - 1)], a1);
    a2 = fma(v[j + 2], z[(T + j + 2) & (BW - 1)], a2);
    a3 = fma(v[j + 3], z[(T + j + 3) & (BW - 1)], a3);
  }
  const double f = -tau * ((a0 + a1) + (a2 + a3));
  asm volatile("" : "+s"(v));
#pragma unroll
  for (int j = 0; j < BW; ++j) z[(T + j) & (BW - 1)] = fma(f, v[j], z[(T + j) & (BW - 1)]);
}

template <int... I>
__device__ __forceinline__ void bt2_steps(std::integer_sequence<int, I...>, double (&z)[BW],
                                          double (&pre)[PF], gdouble*& sp, gdouble*& lp, long ldt,
                                          int s0, int nsw, long krow, kdouble* V2, kdouble* tau2) {
  (bt2_step<BT - 1 - I>(z, pre, sp, lp, ldt, s0, nsw, krow, V2, tau2), ...);
}

// One wavefront L of the stage-2 back-transformation on Zt (vector c's row r
// at Zt[c + r ldt]; ldt a multiple of 64 and at least BT + BW - 1 rows below
// row n - 1, all zero, as are the columns nv .. ldt - 1).  Work-group
// (x, y): block (g, k) with k = klo + x, g = G - 1 - (L - k), vectors
// 64 y .. 64 y + 63.  The block's window is rows R0 = g BT + 1 + k BW ..
// R0 + BT + BW - 2; sweep s0 + T of the block acts on window rows
// T .. T + BW - 1.  Window rows past n - 1 meet only zero reflector entries
// (a reflector ends at row n - 1), so the padding stays zero and no lane or
// row needs a guard.
__global__ void __launch_bounds__(64)
k_bt2(double* __restrict__ Zt, long ldt, int n, int L, int G, int klo,
      const double* __restrict__ V2, const double* __restrict__ tau2) {
  const int nsw = n - 2;
  const int k = klo + blockIdx.x;
  const int g = G - 1 - (L - k);
  const int s0 = g * BT;
  // the group's first sweep has the most tasks
  if (g < 0 || s0 >= nsw || k > (n - 2 - s0) / BW) return;
  const long R0 = (long)s0 + 1 + (long)k * BW;
  double* zb = Zt + (long)(blockIdx.y * 64 + threadIdx.x) + R0 * ldt;
  double z[BW], pre[PF];
#pragma unroll
  for (int r = 0; r < BW; ++r) z[(BT - 1 + r) & (BW - 1)] = zb[(BT - 1 + r) * ldt];
#pragma unroll
  for (int q = 1; q <= PF; ++q) pre[(BT - 1 - q) % PF] = zb[(BT - 1 - q) * ldt];
  gdouble* sp = (gdouble*)(zb + (BT - 2 + BW) * ldt);   // row leaving at the second step
  gdouble* lp = (gdouble*)(zb + (BT - 2 - PF) * ldt);   // row loaded at the second step
  bt2_steps(std::make_integer_sequence<int, BT>{}, z, pre, sp, lp, ldt, s0, nsw, (long)k * nsw,
            (kdouble*)V2, (kdouble*)tau2);
#pragma unroll
  for (int r = 0; r < BW; ++r) zb[r * ldt] = z[r & (BW - 1)];
}

// V of the panel at column j0 (m x BW, ld m) from the reduced A: below the
// diagonal of the panel's top block and all rows under it as stored, unit
// diagonal, zero above.
__global__ void __launch_bounds__(256)
k_panel_v(const double* __restrict__ A, long lda, int j0, int m, double* __restrict__ V) {
  const long tot = (long)m * BW;
  for (long t = blockIdx.x * 256L + threadIdx.x; t < tot; t += (long)gridDim.x * 256) {
    const int i = (int)(t % m), c = (int)(t / m);
    V[t] = i < c ? 0.0 : (i == c ? 1.0 : A[(j0 + BW + i) + (long)(j0 + c) * lda]);
  }
}

// The reflectors of the last short panel (m < BW rows below it, column j0),
// last one first, on Zt: thread = vector.  Reflector c acts on rows
// j0 + BW + c .. j0 + BW + m - 1, v = (1, A(below)), factor tau[c].
__global__ void __launch_bounds__(256)
k_bt1_last(double* __restrict__ Zt, long ldt, int nv, const double* __restrict__ A, long lda,
           int j0, int m, const double* __restrict__ tau) {
  const int col = blockIdx.x * 256 + threadIdx.x;
  if (col >= nv) return;
  const long r0 = j0 + BW;
  double* zc = Zt + col;
  for (int c = m - 2; c >= 0; --c) {
    const double tv = tau[c];
    if (tv == 0.0) continue;
    const double* v = A + r0 + (long)(j0 + c) * lda;
    double d = zc[(r0 + c) * ldt];
    for (int t = c + 1; t < m; ++t) d = fma(v[t], zc[(r0 + t) * ldt], d);
    d *= tv;
    zc[(r0 + c) * ldt] -= d;
    for (int t = c + 1; t < m; ++t) zc[(r0 + t) * ldt] = fma(-d, v[t], zc[(r0 + t) * ldt]);
  }
}

// B (cols x rows, ld ldb) = A^T, A rows x cols (ld lda), in 64 x 64 tiles
// through LDS.
__global__ void __launch_bounds__(256)
k_transpose(const double* __restrict__ A, long lda, int rows, int cols, double* __restrict__ B,
            long ldb) {
  __shared__ double tl[64][65];
  const int r0 = blockIdx.x * 64, c0 = blockIdx.y * 64;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  for (int c = ty; c < 64; c += 4)
    if (r0 + tx < rows && c0 + c < cols) tl[c][tx] = A[(r0 + tx) + (long)(c0 + c) * lda];
  __syncthreads();
  for (int r = ty; r < 64; r += 4)
    if (c0 + tx < cols && r0 + r < rows) B[(c0 + tx) + (long)(r0 + r) * ldb] = tl[tx][r];
}

__global__ void __launch_bounds__(256) k_zero(double* __restrict__ x, long n) {
  for (long t = blockIdx.x * 256L + threadIdx.x; t < n; t += (long)gridDim.x * 256) x[t] = 0.0;
}

int grid_of(long n) { return (int)std::min<long>((n + 255) / 256, 8192); }

size_t house_rec_lds() { return sizeof(double) * (3 * BW * LD + BW); }
size_t last_panel_lds() { return sizeof(double) * (2 * BW * LD + 2 * BW + 6); }
size_t sb2st_lds() { return sizeof(double) * (BW * LD + 3 * BW + 12); }

}  // namespace

int twostage_bandwidth() { return BW; }

// Reduce A (n x n, ld lda, both triangles valid; destroyed) to tridiagonal
// form (d[n], e[n-1] on the device).  Returns 0, or 1 when a panel was rank
// deficient (the caller then reduces a copy of A in one stage).
int TwoStage::reduce(double* A, long lda, int n, double* d, double* e, hipStream_t s) {
  if (n < 4 * BW) throw HipError("two-stage reduction: n < 256");
  const int m0 = n - BW;
  // workspace of the split-K GEMMs: the largest over every panel's shapes
  // (the K split of the product A22 W grows as m shrinks, so the first
  // panel is not always the largest)
  size_t wneed = 1;
  for (int jj = 0; n - jj - BW >= BW; jj += BW) {
    const int m = n - jj - BW;
    const int shp[4][3] = {{BW, BW, m}, {m, BW, m}, {m, BW, BW}, {m - BW, BW, BW}};
    for (const auto& x : shp)
      if (x[0] > 0)
        wneed = std::max(wneed, gemm_work_doubles(x[0], x[1], gemm_auto_ksplit(x[0], x[1], x[2])));
  }
  work.ensure(sizeof(double) * wneed);
  vyv.ensure(sizeof(double) * (size_t)m0 * 3 * BW);
  wb.ensure(sizeof(double) * (size_t)m0 * BW);
  qa.ensure(sizeof(double) * (size_t)m0 * BW);
  qb.ensure(sizeof(double) * (size_t)m0 * BW);
  const int npan = (n + BW - 1) / BW;
  for (DevBuf* x : {&G, &dl, &Linv, &sc, &M, &Uinv, &Rb, &Zs, &Mm})
    x->ensure(sizeof(double) * BW * BW);
  Sv.ensure(sizeof(double) * BW);
  Tst.ensure(sizeof(double) * (size_t)npan * BW * BW);
  zero.ensure(sizeof(int) * npan);
  flags.ensure(sizeof(int));
  taul.ensure(sizeof(double) * BW);
  DVR_HIP(hipMemsetAsync(flags.p, 0, sizeof(int), s));
  DVR_HIP(hipFuncSetAttribute((const void*)k_house_rec,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)house_rec_lds()));
  DVR_HIP(hipFuncSetAttribute((const void*)k_last_panel,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)last_panel_lds()));
  DVR_HIP(hipFuncSetAttribute((const void*)k_sb2st_step,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)sb2st_lds()));
  double* W = work.as<double>();
  auto gemm = [&](const char* name, bool ta, bool tb, int M_, int N_, int K_, double alpha,
                  const double* A_, long lda_, const double* B_, long ldb_, double beta, double* C_,
                  long ldc_) {
    const int ks = gemm_auto_ksplit(M_, N_, K_);
    if (gemm_work_doubles(M_, N_, ks) > wneed)
      throw HipError("two-stage reduction: split-K workspace too small");
    launch_gemm(s, name, ta, tb, M_, N_, K_, alpha, A_, lda_, B_, ldb_, beta, C_, ldc_, ks,
                ks > 1 ? W : nullptr, false);
  };
  const double eps = 2.220446049250313e-16;
  int p = 0;
  int j0 = 0;
  for (; n - j0 - BW >= BW; j0 += BW, ++p) {
    const int m = n - j0 - BW;
    double* P = A + (j0 + BW) + (long)j0 * lda;
    double* A22 = A + (j0 + BW) * (1 + lda);
    // P = Q R: shifted Cholesky QR, then two plain passes
    const double* Z = P;
    long ldz = lda;
    double* qout[3] = {qa.as<double>(), qb.as<double>(), qa.as<double>()};
    for (int pass = 0; pass < 3; ++pass) {
      gemm("s1_gram", true, false, BW, BW, m, 1.0, Z, ldz, Z, ldz, 0.0, G.as<double>(), BW);
      const double shift = pass == 0 ? 11.0 * ((double)m * BW + BW * (BW + 1.0)) * eps * BW : 0.0;
      launch_gram_prep(s, G.as<double>(), BW, shift, sc.as<double>());
      launch_ldl_block(s, G.as<double>(), BW, BW, dl.as<double>(), Linv.as<double>());
      DVR_LAUNCH("s1_cholqr", 8.0 * 3 * BW * BW, 0.0, k_cholqr_m, dim3(1), dim3(256), 0, s,
                 dl.as<double>(), Linv.as<double>(), sc.as<double>(), M.as<double>(),
                 flags.as<int>(), zero.as<int>() + p, pass == 0 ? 1 : 0);
      gemm("s1_q", false, true, m, BW, BW, 1.0, Z, ldz, M.as<double>(), BW, 0.0, qout[pass], m);
      Z = qout[pass];
      ldz = m;
    }
    const double* Q = qa.as<double>();
    double* V = vyv.as<double>();
    double* X = V + (size_t)m * BW;
    double* T = Tst.as<double>() + (size_t)p * BW * BW;
    DVR_LAUNCH("s1_house_rec", 8.0 * 4 * BW * BW, 0.0, k_house_rec, dim3(1), dim3(256),
               house_rec_lds(), s, Q, (long)m, V, (long)m, Uinv.as<double>(), T, Sv.as<double>(),
               zero.as<int>() + p);
    gemm("s1_r", true, false, BW, BW, m, 1.0, Q, m, P, lda, 0.0, Rb.as<double>(), BW);
    gemm("s1_v2", false, false, m - BW, BW, BW, 1.0, Q + BW, m, Uinv.as<double>(), BW, 0.0, V + BW,
         m);
    DVR_HIP(hipMemcpyAsync(V + (size_t)2 * m * BW, V, sizeof(double) * (size_t)m * BW,
                           hipMemcpyDeviceToDevice, s));
    // X = A22 V T, M = T^T V^T X, Y = X - V M / 2 (in place of X)
    gemm("s1_vt", false, false, m, BW, BW, 1.0, V, m, T, BW, 0.0, wb.as<double>(), m);
    gemm("s1_symm", false, false, m, BW, m, 1.0, A22, lda, wb.as<double>(), m, 0.0, X, m);
    gemm("s1_vtx", true, false, BW, BW, m, 1.0, V, m, X, m, 0.0, Zs.as<double>(), BW);
    gemm("s1_m", true, false, BW, BW, BW, 1.0, T, BW, Zs.as<double>(), BW, 0.0, Mm.as<double>(),
         BW);
    gemm("s1_y", false, false, m, BW, BW, -0.5, V, m, Mm.as<double>(), BW, 1.0, X, m);
    // A22 -= [V Y] [Y V]^T on the lower tiles, then the upper tiles mirrored
    launch_gemm_var(s, "s1_syr2k", false, true, m, m, 2 * BW, -1.0, V, m, X, m, 1.0, A22, lda,
                    true, 16, true, true);
    const long nt = (m + BW - 1) / BW;
    if (nt > 1)
      DVR_LAUNCH("s1_mirror", 8.0 * (double)m * m, 0.0, k_mirror_tiles,
                 dim3((unsigned)(nt * (nt - 1) / 2)), dim3(256), 0, s, A22, lda, m);
    DVR_LAUNCH("s1_store", 16.0 * m * BW, 0.0, k_panel_store, dim3(grid_of((long)m * BW)),
               dim3(256), 0, s, A, lda, j0, m, Rb.as<double>(), Sv.as<double>(), V);
  }
  const int mlast = n - j0 - BW;
  if (mlast > 0)
    DVR_LAUNCH("s1_last", 8.0 * 2 * (double)(mlast + BW) * (mlast + BW), 0.0, k_last_panel, dim3(1),
               dim3(256), last_panel_lds(), s, A, lda, j0, mlast, taul.as<double>());
  Ared = A;
  lda_red = lda;
  n_red = n;
  npanel = p;
  jlast = j0;
  this->mlast = mlast;
  // ---- stage 2
  AB.ensure(sizeof(double) * (size_t)n * LDB);
  vs.ensure(sizeof(double) * (size_t)n * BW);
  taus.ensure(sizeof(double) * n);
  DVR_LAUNCH("s2_band", 8.0 * 2 * (double)n * LDB, 0.0, k_band_extract, dim3(grid_of((long)n * LDB)),
             dim3(256), 0, s, A, lda, n, AB.as<double>());
  const int nsweep = n - 2;   // sweeps 0 .. n - 3
  auto kmax = [&](int sw) { return (n - 2 - sw) / BW; };   // last task with r0 <= n - 1
  double* v2 = nullptr;
  double* t2 = nullptr;
  if (keep) {
    // one slot per (task, sweep); slots without a task stay zero
    const size_t nslot = (size_t)(kmax(0) + 1) * nsweep;
    V2.ensure(sizeof(double) * nslot * BW);
    tau2.ensure(sizeof(double) * nslot);
    v2 = V2.as<double>();
    t2 = tau2.as<double>();
    DVR_LAUNCH("s2_zero", 8.0 * nslot * BW, 0.0, k_zero, dim3(grid_of((long)nslot * BW)),
               dim3(256), 0, s, v2, (long)(nslot * BW));
    DVR_LAUNCH("s2_zero", 8.0 * nslot, 0.0, k_zero, dim3(grid_of((long)nslot)), dim3(256), 0, s,
               t2, (long)nslot);
  }
  const int tmax = LAG * (nsweep - 1) + kmax(0);
  int slo = 0;
  for (int t = 0; t <= tmax; ++t) {
    const int shi = std::min(nsweep - 1, t / LAG);
    while (slo <= shi && t - LAG * slo > kmax(slo)) ++slo;
    if (slo > shi) continue;
    DVR_LAUNCH("s2_chase", 0.0, 0.0, k_sb2st_step, dim3(shi - slo + 1), dim3(256), sb2st_lds(), s,
               AB.as<double>(), n, t, slo, vs.as<double>(), taus.as<double>(), v2, t2);
  }
  DVR_LAUNCH("s2_tri", 16.0 * n, 0.0, k_band_to_tri, dim3(grid_of(n)), dim3(256), 0, s,
             AB.as<double>(), n, d, e);
  DVR_LAUNCH_CHECK();
  int hf = 0;
  DVR_HIP(hipMemcpyAsync(&hf, flags.p, sizeof(int), hipMemcpyDeviceToHost, s));
  DVR_HIP(hipStreamSynchronize(s));
  return hf ? 1 : 0;
}

void TwoStage::back(double* Z, long ldz, int nv, hipStream_t s) {
  if (!keep || !Ared || !V2.p) throw HipError("two-stage back-transformation: reflectors not kept");
  const int n = n_red;
  if (nv <= 0) return;
  // transposed vectors, columns padded to a multiple of 64 and BT + BW rows
  // of zeros below the last row (the wavefront kernel's windows run into them)
  const long ldt = (nv + 63) / 64 * 64;
  const long zrows = (long)n + BT + BW;
  Zt.ensure(sizeof(double) * (size_t)(ldt * zrows));
  double* zt = Zt.as<double>();
  DVR_LAUNCH("bt_zero", 8.0 * ldt * zrows, 0.0, k_zero, dim3(grid_of(ldt * zrows)), dim3(256), 0, s,
             zt, ldt * zrows);
  const dim3 tg((n + 63) / 64, (nv + 63) / 64);
  DVR_LAUNCH("bt_transpose", 16.0 * (double)n * nv, 0.0, k_transpose, tg, dim3(256), 0, s,
             (const double*)Z, ldz, n, nv, zt, ldt);
  // ---- stage 2: wavefronts L = 0 .. G - 1 + K, blocks (g, k) with
  // g = G - 1 - (L - k) and k <= kmax(g BT)
  const int nsw = n - 2;
  const int K = (n - 2) / BW;
  const int G = (nsw + BT - 1) / BT;
  const unsigned ny = (unsigned)(ldt / 64);
  for (int L = 0; L <= G - 1 + K; ++L) {
    const int klo = std::max(0, L - (G - 1));
    int khi = std::min(K, L);
    // last k whose block is not empty (kmax falls as g, and so k, grows)
    while (khi >= klo && khi > (n - 2 - (G - 1 - (L - khi)) * BT) / BW) --khi;
    if (khi < klo) continue;
    const double nblk = (double)(khi - klo + 1);
    DVR_LAUNCH("bt2", nblk * ldt * 16.0 * (BT + BW - 1), nblk * ldt * BT * 4.0 * BW, k_bt2,
               dim3((unsigned)(khi - klo + 1), ny), dim3(64), 0, s, zt, ldt, n, L, G, klo,
               (const double*)V2.as<double>(), (const double*)tau2.as<double>());
  }
  // ---- stage 1: the last short panel, then the panels in reverse
  if (mlast > 1)
    DVR_LAUNCH("bt1_last", 16.0 * mlast * nv, 2.0 * mlast * mlast * nv, k_bt1_last, dim3((nv + 255) / 256),
               dim3(256), 0, s, zt, ldt, nv, (const double*)Ared, lda_red, jlast, mlast,
               (const double*)taul.as<double>());
  if (npanel > 0) {
    const int m0 = n - BW;
    size_t wneed = 1;
    for (int p = 0; p < npanel; ++p) {
      const int m = n - p * BW - BW;
      const int shp[3][3] = {{nv, BW, m}, {nv, BW, BW}, {nv, m, BW}};
      for (const auto& x : shp)
        wneed = std::max(wneed, gemm_work_doubles(x[0], x[1], gemm_auto_ksplit(x[0], x[1], x[2])));
    }
    work.ensure(sizeof(double) * wneed);
    Vp.ensure(sizeof(double) * (size_t)m0 * BW);
    bw1.ensure(sizeof(double) * (size_t)nv * BW);
    bw2.ensure(sizeof(double) * (size_t)nv * BW);
    double* W = work.as<double>();
    auto gemm = [&](const char* name, bool ta, bool tb, int M_, int N_, int K_, double alpha,
                    const double* A_, long lda_, const double* B_, long ldb_, double beta,
                    double* C_, long ldc_) {
      const int ks = gemm_auto_ksplit(M_, N_, K_);
      if (gemm_work_doubles(M_, N_, ks) > wneed)
        throw HipError("two-stage back-transformation: split-K workspace too small");
      launch_gemm(s, name, ta, tb, M_, N_, K_, alpha, A_, lda_, B_, ldb_, beta, C_, ldc_, ks,
                  ks > 1 ? W : nullptr, false);
    };
    for (int p = npanel - 1; p >= 0; --p) {
      const int j0 = p * BW;
      const int m = n - j0 - BW;
      double* V = Vp.as<double>();
      DVR_LAUNCH("bt1_v", 16.0 * m * BW, 0.0, k_panel_v, dim3(grid_of((long)m * BW)), dim3(256), 0,
                 s, (const double*)Ared, lda_red, j0, m, V);
      double* Y = zt + (long)(j0 + BW) * ldt;
      const double* T = Tst.as<double>() + (size_t)p * BW * BW;
      // Y <- Y - (Y V) T^T V^T   (Yt = rows of the transposed vectors)
      gemm("bt1_yv", false, false, nv, BW, m, 1.0, Y, ldt, V, m, 0.0, bw1.as<double>(), nv);
      gemm("bt1_t", false, true, nv, BW, BW, 1.0, bw1.as<double>(), nv, T, BW, 0.0,
           bw2.as<double>(), nv);
      gemm("bt1_upd", false, true, nv, m, BW, -1.0, bw2.as<double>(), nv, V, m, 1.0, Y, ldt);
    }
  }
  const dim3 tb((nv + 63) / 64, (n + 63) / 64);
  DVR_LAUNCH("bt_transpose", 16.0 * (double)n * nv, 0.0, k_transpose, tb, dim3(256), 0, s,
             (const double*)zt, ldt, nv, n, Z, ldz);
  DVR_LAUNCH_CHECK();
}

void twostage_vectors(int n, int m, double* Z, hipStream_t s, TwoStage& ts, EigWork& W) {
  W.vectors(std::vector<int>{m}, Z, std::vector<long>{0}, s, false);
  ts.back(Z, n, m, s);
  DVR_HIP(hipStreamSynchronize(s));
}

bool twostage_values(const double* A, int n, int want, hipStream_t s, TwoStage& ts, EigWork& W,
                     DevBuf& copy) {
  copy.ensure(sizeof(double) * (size_t)n * n);
  DVR_HIP(hipMemcpyAsync(copy.p, A, sizeof(double) * (size_t)n * n, hipMemcpyDeviceToDevice, s));
  W.tridiag(std::vector<int>{n}, copy.as<double>(), std::vector<long>{0}, s, false, false);
  if (ts.reduce(copy.as<double>(), n, n, W.b.d, W.b.e, s) == 0) {
    W.bounds(s);
    W.values(std::vector<int>{want}, s);
    return true;
  }
  // a rank-deficient panel: the one-stage reduction of the intact A
  DVR_HIP(hipMemcpyAsync(copy.p, A, sizeof(double) * (size_t)n * n, hipMemcpyDeviceToDevice, s));
  W.tridiag(std::vector<int>{n}, copy.as<double>(), std::vector<long>{0}, s);
  W.values(std::vector<int>{want}, s);
  return false;
}

}  // namespace dvr
