// Small device kernels of the block-Davidson final solve: column gathers,
// Ritz residuals and their norms, symmetrisation of projected matrices.
#include "dvr_internal.h"

namespace dvr {

// dst[:, c] = src[:, idx[c]]   (nrows x ncol, column-major)
__global__ void __launch_bounds__(256)
k_gather_cols(const double* __restrict__ src, long lds, const int* __restrict__ idx, int ncol,
              double* __restrict__ dst, long ldd, int nrows) {
  const long n = (long)nrows * ncol;
  for (long t = blockIdx.x * 256L + threadIdx.x; t < n; t += (long)gridDim.x * 256) {
    const int r = (int)(t % nrows), c = (int)(t / nrows);
    dst[r + (long)c * ldd] = src[r + (long)idx[c] * lds];
  }
}

// norms[j] = || AX[:, j] - theta[j] X[:, j] ||, one work-group per column
__global__ void __launch_bounds__(256)
k_resid_norms(const double* __restrict__ X, const double* __restrict__ AX, long ld,
              const double* __restrict__ theta, int nrows, double* __restrict__ norms) {
  __shared__ double red[4];
  const int j = blockIdx.x;
  const double th = theta[j];
  const double* x = X + (long)j * ld;
  const double* ax = AX + (long)j * ld;
  double s = 0.0;
  for (int i = threadIdx.x; i < nrows; i += 256) {
    const double r = ax[i] - th * x[i];
    s += r * r;
  }
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) norms[j] = sqrt(red[0] + red[1] + red[2] + red[3]);
}

// R[:, c] = AX[:, idx[c]] - theta[idx[c]] X[:, idx[c]]
__global__ void __launch_bounds__(256)
k_resid_gather(const double* __restrict__ X, const double* __restrict__ AX, long ld,
               const double* __restrict__ theta, const int* __restrict__ idx, int ncol,
               double* __restrict__ R, long ldr, int nrows) {
  const long n = (long)nrows * ncol;
  for (long t = blockIdx.x * 256L + threadIdx.x; t < n; t += (long)gridDim.x * 256) {
    const int r = (int)(t % nrows), c = (int)(t / nrows);
    const int j = idx[c];
    R[r + (long)c * ldr] = AX[r + (long)j * ld] - theta[j] * X[r + (long)j * ld];
  }
}

// T <- (T + T^T) / 2 (m x m, ld m)
__global__ void __launch_bounds__(256) k_symmetrize(double* __restrict__ T, int m) {
  const long n = (long)m * m;
  for (long t = blockIdx.x * 256L + threadIdx.x; t < n; t += (long)gridDim.x * 256) {
    const int r = (int)(t % m), c = (int)(t / m);
    if (r > c) {
      const double v = 0.5 * (T[r + (long)c * m] + T[c + (long)r * m]);
      T[r + (long)c * m] = v;
      T[c + (long)r * m] = v;
    }
  }
}

// U[:, c] *= scal[c] (m x ncol, ld m)
__global__ void __launch_bounds__(256)
k_scale_cols(double* __restrict__ U, long ld, int m, int ncol, const double* __restrict__ scal) {
  const long n = (long)m * ncol;
  for (long t = blockIdx.x * 256L + threadIdx.x; t < n; t += (long)gridDim.x * 256) {
    const int r = (int)(t % m), c = (int)(t / m);
    U[r + (long)c * ld] *= scal[c];
  }
}

// G <- S G S + eps I with S = diag(1/sqrt(G_kk)) (q x q, one work-group);
// sc[k] = S_kk (0 for an empty column)
__global__ void __launch_bounds__(256)
k_gram_prep(double* __restrict__ G, int q, double eps, double* __restrict__ sc) {
  __shared__ double s[64];
  const int t = threadIdx.x;
  if (t < q) {
    const double g = G[t + (long)t * q];
    s[t] = g > 0.0 ? 1.0 / sqrt(g) : 0.0;
    sc[t] = s[t];
  }
  __syncthreads();
  for (int idx = t; idx < q * q; idx += 256) {
    const int i = idx % q, j = idx / q;
    double v = G[idx] * s[i] * s[j];
    if (i == j) v = (s[i] > 0.0 ? v : 1.0) + eps;
    G[idx] = v;
  }
}

// M(i,j) = r[i] * Linv(i,j) * sc[j]  (q x q)
__global__ void __launch_bounds__(256)
k_form_m(const double* __restrict__ Linv, int q, const double* __restrict__ r,
         const double* __restrict__ sc, double* __restrict__ M) {
  for (int idx = threadIdx.x; idx < q * q; idx += 256) {
    const int i = idx % q, j = idx / q;
    M[idx] = r[i] * Linv[idx] * sc[j];
  }
}

static int grid_for(long n) { return (int)std::min<long>((n + 255) / 256, 8192); }

void launch_gather_cols(hipStream_t s, const double* src, long lds, const int* idx, int ncol,
                        double* dst, long ldd, int nrows) {
  if (ncol <= 0 || nrows <= 0) return;
  DVR_LAUNCH("gather", 16.0 * nrows * ncol, 0.0, k_gather_cols, dim3(grid_for((long)nrows * ncol)),
             dim3(256), 0, s, src, lds, idx, ncol, dst, ldd, nrows);
}

void launch_resid_norms(hipStream_t s, const double* X, const double* AX, long ld,
                        const double* theta, int m, int nrows, double* norms) {
  if (m <= 0) return;
  DVR_LAUNCH("resid_norms", 16.0 * nrows * m, 3.0 * nrows * m, k_resid_norms, dim3(m), dim3(256),
             0, s, X, AX, ld, theta, nrows, norms);
}

void launch_resid_gather(hipStream_t s, const double* X, const double* AX, long ld,
                         const double* theta, const int* idx, int ncol, double* R, long ldr,
                         int nrows) {
  if (ncol <= 0) return;
  DVR_LAUNCH("resid", 24.0 * nrows * ncol, 2.0 * nrows * ncol, k_resid_gather,
             dim3(grid_for((long)nrows * ncol)), dim3(256), 0, s, X, AX, ld, theta, idx, ncol, R,
             ldr, nrows);
}

// T(r, c) -> T(c, r) for r > c (m x m, ld m)
__global__ void __launch_bounds__(256) k_mirror_lower(double* __restrict__ T, int m) {
  const long n = (long)m * m;
  for (long t = blockIdx.x * 256L + threadIdx.x; t < n; t += (long)gridDim.x * 256) {
    const int r = (int)(t % m), c = (int)(t / m);
    if (r > c) T[c + (long)r * m] = T[r + (long)c * m];
  }
}

void launch_mirror_lower(hipStream_t s, double* T, int m) {
  if (m <= 1) return;
  DVR_LAUNCH("mirror", 8.0 * m * m, 0.0, k_mirror_lower, dim3(grid_for((long)m * m)), dim3(256),
             0, s, T, m);
}

void launch_symmetrize(hipStream_t s, double* T, int m) {
  if (m <= 1) return;
  DVR_LAUNCH("symm", 16.0 * m * m, 0.0, k_symmetrize, dim3(grid_for((long)m * m)), dim3(256), 0,
             s, T, m);
}

void launch_gram_prep(hipStream_t s, double* G, int q, double eps, double* sc) {
  if (q <= 0 || q > 64) throw HipError("gram_prep: size out of range");
  DVR_LAUNCH("gram_prep", 16.0 * q * q, 0.0, k_gram_prep, dim3(1), dim3(256), 0, s, G, q, eps, sc);
}

void launch_form_m(hipStream_t s, const double* Linv, int q, const double* r, const double* sc,
                   double* M) {
  DVR_LAUNCH("form_m", 16.0 * q * q, 0.0, k_form_m, dim3(1), dim3(256), 0, s, Linv, q, r, sc, M);
}

void launch_scale_cols(hipStream_t s, double* U, long ld, int m, int ncol, const double* scal) {
  if (ncol <= 0) return;
  DVR_LAUNCH("scale_cols", 16.0 * m * ncol, 0.0, k_scale_cols, dim3(grid_for((long)m * ncol)),
             dim3(256), 0, s, U, ld, m, ncol, scal);
}

}  // namespace dvr

namespace dvr {

// Y[i, v] += e[i] * X[i, v]  (N x nv, ld N)
__global__ void __launch_bounds__(256)
k_add_diag(double* __restrict__ Y, const double* __restrict__ X, const double* __restrict__ e,
           int N, int nv) {
  const long n = (long)N * nv;
  for (long t = blockIdx.x * 256L + threadIdx.x; t < n; t += (long)gridDim.x * 256)
    Y[t] += e[t % N] * X[t];
}

// out[i] = H[i, i]  (H: N x N, ld N)
__global__ void __launch_bounds__(256)
k_get_diag(const double* __restrict__ H, int N, double* __restrict__ out) {
  for (long i = blockIdx.x * 256L + threadIdx.x; i < N; i += (long)gridDim.x * 256)
    out[i] = H[i * (N + 1L)];
}

// column c of V (N x b, ld N) <- the unit vector e_{idx[c]}
__global__ void __launch_bounds__(256)
k_unit_cols(double* __restrict__ V, int N, const int* __restrict__ idx, int b) {
  const long n = (long)N * b;
  for (long t = blockIdx.x * 256L + threadIdx.x; t < n; t += (long)gridDim.x * 256) {
    const int r = (int)(t % N), c = (int)(t / N);
    V[t] = r == idx[c] ? 1.0 : 0.0;
  }
}

void launch_unit_cols(hipStream_t s, double* V, int N, const int* idx, int b) {
  const long n = (long)N * b;
  DVR_LAUNCH("unit_cols", 8.0 * n, 0.0, k_unit_cols,
             dim3((int)std::min<long>((n + 255) / 256, 8192)), dim3(256), 0, s, V, N, idx, b);
}

void launch_get_diag(hipStream_t s, const double* H, int N, double* out) {
  DVR_LAUNCH("get_diag", 16.0 * N, 0.0, k_get_diag, dim3((N + 255) / 256), dim3(256), 0, s, H, N,
             out);
}

void launch_add_diag(hipStream_t s, double* Y, const double* X, const double* e, int N, int nv) {
  const long n = (long)N * nv;
  DVR_LAUNCH("add_diag", 24.0 * n, 2.0 * n, k_add_diag,
             dim3((int)std::min<long>((n + 255) / 256, 8192)), dim3(256), 0, s, Y, X, e, N, nv);
}

}  // namespace dvr
