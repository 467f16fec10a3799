// LDS-resident single-work-group symmetric eigensolver for n <= 128.
//
// One 256-thread work-group per problem: the matrix is copied into LDS,
// reduced to tridiagonal form by Householder reflections (same convention as
// k_syev.hip: v(j+1) = 1, H = I - tau v v^T, reflectors kept below the
// subdiagonal), the wanted eigenvalues are found by Sturm multisection (one
// wavefront per eigenvalue), eigenvectors of T by inverse iteration (one
// thread per cluster of close eigenvalues, modified Gram-Schmidt inside a
// cluster) and back-transformed in registers: 16 threads share one vector,
// 8 rows each, and reduce their dot products with shuffles, so the
// back-transformation needs no work-group barrier.
//
// Uses in the path:
//  * the 1D angular problems (mkham1 -> diag -> cut1d, source/dvr3drjz_
//    segmented.f90:2121-2225, 2638-2659, 2870-2938): all n eigenvalues and
//    only the eigenvectors below EMAX1 (the ones cut1d keeps);
//  * the Rayleigh-Ritz problems of the block-Davidson final solve.
#include "dvr_internal.h"
#include "inviter.h"
#include <cfloat>

namespace dvr {

namespace {
constexpr int NT = 256;
constexpr int NWV = NT / 64;
constexpr int VG = 16;          // threads per vector in the back-transform
constexpr int VR = 8;           // rows per thread (n <= VG * VR = 128)
constexpr double kEpsL = 2.220446049250313e-16;

__device__ inline double wred(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return __shfl(v, 0, 64);
}

__device__ inline double bred(double v, double* red) {
  v = wred(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  double s = 0.0;
  for (int i = 0; i < NWV; ++i) s += red[i];
  return s;
}

}  // namespace

size_t syev_lds_bytes(int n) {
  return (size_t)n * (n + 1) * 8 + (size_t)8 * n * 8 + NWV * 8 + (size_t)(n + 2) * 4;
}

size_t syev_lds_scratch_doubles(int n, int nsolve) {
  return (size_t)6 * n * n * (size_t)nsolve;
}

// A_p at Ain + p*astride (n x n, column-major, both triangles), p =
// src[blockIdx.x] (or blockIdx.x).  Writes the lowest kcnt eigenvalues to
// eval[p*estride + k] and the lowest m eigenvectors to Z[p*zstride + k*n + i],
// where m = mfixed, or (mfixed < 0) the number of eigenvalues <= emax, which
// is also stored in cnt[p].
__global__ void __launch_bounds__(NT)
k_syev_lds(const double* __restrict__ Ain, long astride, int n, int kcnt,
           const int* __restrict__ src, double* __restrict__ eval, long estride,
           double* __restrict__ Z, long zstride, int* __restrict__ cnt,
           int mfixed, double emax, double* __restrict__ scratch) {
  extern __shared__ double sm[];
  const int p = src ? src[blockIdx.x] : (int)blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lda = n + 1;
  double* A = sm;
  double* d = A + (long)n * lda;
  double* e = d + n;
  double* e2 = e + n;
  double* tau = e2 + n;
  double* va = tau + n;
  double* vb = va + n;
  double* w = vb + n;
  double* sv = w + n;
  double* red = sv + n;
  int* cls = reinterpret_cast<int*>(red + NWV);   // n + 2 ints

  const double* src_a = Ain + (long)p * astride;
  for (int idx = tid; idx < n * n; idx += NT) {
    const int j = idx / n, i = idx - j * n;
    A[i + j * lda] = src_a[idx];
  }
  __syncthreads();

  // ---- tridiagonalisation
  if (n == 1 && tid == 0) d[0] = A[0];
  for (int j = 0; j + 1 < n; ++j) {
    double* vprev = (j & 1) ? va : vb;
    double* vcur = (j & 1) ? vb : va;
    double* col = A + j * lda;
    double part = 0.0;
    if (j > 0) {
      const double taup = tau[j - 1];
      double kp = 0.0;
      for (int i = j + tid; i < n; i += NT) kp += sv[i] * vprev[i];
      const double K = taup * bred(kp, red);
      const double alpha = -0.5 * taup * K;
      const double vj = vprev[j];
      const double wj = taup * sv[j] + alpha * vj;
      for (int i = j + tid; i < n; i += NT) {
        const double wi = taup * sv[i] + alpha * vprev[i];
        w[i] = wi;
        const double ci = col[i] - (vprev[i] * wj + wi * vj);
        col[i] = ci;
        if (i >= j + 2) part += ci * ci;
      }
    } else {
      for (int i = j + 2 + tid; i < n; i += NT) part += col[i] * col[i];
    }
    const double sigma = bred(part, red);
    if (j == n - 2) {
      if (tid == 0) {
        d[j] = col[j];
        e[j] = col[n - 1];
        tau[j] = 0.0;
        double dl = A[(n - 1) + (n - 1) * lda];
        if (j > 0) dl -= (vprev[n - 1] * w[n - 1] + w[n - 1] * vprev[n - 1]);
        d[n - 1] = dl;
      }
      __syncthreads();
      break;
    }
    const double alpha = col[j + 1];
    double tj, beta, scal;
    if (sigma == 0.0) {
      tj = 0.0;
      beta = alpha;
      scal = 0.0;
    } else {
      const double mu = sqrt(alpha * alpha + sigma);
      beta = (alpha <= 0.0) ? mu : -mu;
      tj = (beta - alpha) / beta;
      scal = 1.0 / (alpha - beta);
    }
    for (int i = j + 1 + tid; i < n; i += NT) {
      const double vi = (i == j + 1) ? 1.0 : col[i] * scal;
      vcur[i] = vi;
      if (i > j + 1) col[i] = vi;
    }
    __syncthreads();
    if (tid == 0) {
      d[j] = col[j];
      e[j] = beta;
      tau[j] = tj;
      col[j + 1] = beta;
    }
    for (int k = j + 1 + wave; k < n; k += NWV) {
      double* colk = A + k * lda;
      double acc = 0.0;
      if (j > 0) {
        const double wk = w[k], vk = vprev[k];
        for (int i = j + 1 + lane; i < n; i += 64) {
          const double u = colk[i] - (vprev[i] * wk + w[i] * vk);
          colk[i] = u;
          acc += u * vcur[i];
        }
      } else {
        for (int i = j + 1 + lane; i < n; i += 64) acc += colk[i] * vcur[i];
      }
      acc = wred(acc);
      if (lane == 0) sv[k] = acc;
    }
    __syncthreads();
  }

  // ---- bounds, e^2
  double lo = DBL_MAX, hi = -DBL_MAX, mx2 = 0.0;
  for (int i = tid; i < n; i += NT) {
    const double em = i > 0 ? fabs(e[i - 1]) : 0.0;
    const double ep = i < n - 1 ? fabs(e[i]) : 0.0;
    lo = fmin(lo, d[i] - em - ep);
    hi = fmax(hi, d[i] + em + ep);
    if (i < n - 1) {
      e2[i] = e[i] * e[i];
      mx2 = fmax(mx2, e2[i]);
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    lo = fmin(lo, __shfl_xor(lo, off, 64));
    hi = fmax(hi, __shfl_xor(hi, off, 64));
    mx2 = fmax(mx2, __shfl_xor(mx2, off, 64));
  }
  __shared__ double rl[NWV], rh[NWV], rm[NWV];
  if (lane == 0) {
    rl[wave] = lo;
    rh[wave] = hi;
    rm[wave] = mx2;
  }
  __syncthreads();
  lo = rl[0];
  hi = rh[0];
  mx2 = rm[0];
  for (int i = 1; i < NWV; ++i) {
    lo = fmin(lo, rl[i]);
    hi = fmax(hi, rh[i]);
    mx2 = fmax(mx2, rm[i]);
  }
  const double tn = fmax(fabs(lo), fabs(hi));
  const double piv = DBL_MIN * fmax(1.0, mx2);
  const double pad = 2.1 * kEpsL * tn * n + 4.0 * piv;
  lo -= pad;
  hi += pad;

  // ---- eigenvalues: one thread per eigenvalue, bisection on the Sturm
  // count (all lanes read the same d, e2 entry at each step: LDS broadcast).
  // A wavefront per eigenvalue with 64 shifts per step cut the steps 8-fold
  // but ran the n/4 eigenvalues of each wavefront one after another.
  double* ev = eval + (long)p * estride;
  for (int k = tid; k < kcnt; k += NT) {
    double l = lo, h = hi;
    for (int it = 0; it < 80; ++it) {
      const double width = h - l;
      if (width <= 2.0 * kEpsL * fmax(fabs(l), fabs(h)) || width <= 4.0 * piv) break;
      const double x = 0.5 * (l + h);
      if (x <= l || x >= h) break;
      double q = d[0] - x;
      if (fabs(q) < piv) q = -piv;
      int c = q <= 0.0;
      for (int i = 1; i < n; ++i) {
        q = (d[i] - x) - e2[i - 1] * rcp_refined(q);
        if (fabs(q) < piv) q = -piv;
        c += q <= 0.0;
      }
      if (c <= k) l = x; else h = x;
    }
    const double lam = 0.5 * (l + h);
    ev[k] = lam;
    sv[k] = lam;     // LDS copy for the vector phase
  }
  __syncthreads();

  // ---- number of vectors, clusters
  if (tid == 0) {
    int m = mfixed;
    if (m < 0) {
      m = 0;
      while (m < kcnt && !(sv[m] > emax)) ++m;
    }
    if (cnt) cnt[p] = m;
    const double ctol = 1.0e-6 * fmax(tn, 1e-300);
    int nc = 0;
    cls[0] = 0;
    for (int k = 1; k < m; ++k)
      if (sv[k] - sv[k - 1] >= ctol) cls[++nc] = k;
    if (m > 0) ++nc;
    cls[nc] = m;
    cls[n + 1] = nc;
  }
  __syncthreads();
  const int ncl = cls[n + 1];
  const int m = ncl > 0 ? cls[ncl] : 0;
  if (m == 0) return;
  double* Zp = Z + (long)p * zstride;

  // ---- inverse iteration, one thread per cluster, scratch interleaved
  // across the threads (array a, element i of thread c at (a*n + i)*n + c)
  {
    const int c = tid;
    if (c < ncl)
      inverse_iteration_cluster<8>(n, d, e, sv, cls[c], cls[c + 1], tn,
                                scratch + (long)blockIdx.x * 6 * n * n + c, (long)n * n, n, Zp,
                                (unsigned long long)p);
  }
  __syncthreads();

  // ---- back-transformation Z <- Q Z in registers, 16 threads per vector
  const int g = tid % VG;              // row group: rows g, g+16, ..., g+112
  for (int q0 = 0; q0 < m; q0 += NT / VG) {
    const int q = q0 + tid / VG;
    const bool act = q < m;
    double z[VR];
    double* zq = Zp + (long)(act ? q : 0) * n;
#pragma unroll
    for (int r = 0; r < VR; ++r) {
      const int i = g + VG * r;
      z[r] = (act && i < n) ? zq[i] : 0.0;
    }
    for (int j = n - 3; j >= 0; --j) {
      const double tj = tau[j];
      if (tj == 0.0) continue;
      const double* v = A + j * lda;
      double s = 0.0;
#pragma unroll
      for (int r = 0; r < VR; ++r) {
        const int i = g + VG * r;
        const double vi = (i == j + 1) ? 1.0 : ((i > j + 1 && i < n) ? v[i] : 0.0);
        s += vi * z[r];
      }
      s += __shfl_xor(s, 1, 64);
      s += __shfl_xor(s, 2, 64);
      s += __shfl_xor(s, 4, 64);
      s += __shfl_xor(s, 8, 64);
      s *= tj;
#pragma unroll
      for (int r = 0; r < VR; ++r) {
        const int i = g + VG * r;
        const double vi = (i == j + 1) ? 1.0 : ((i > j + 1 && i < n) ? v[i] : 0.0);
        z[r] -= s * vi;
      }
    }
    if (act) {
#pragma unroll
      for (int r = 0; r < VR; ++r) {
        const int i = g + VG * r;
        if (i < n) zq[i] = z[r];
      }
    }
  }
}

void launch_syev_lds(hipStream_t s, const double* A, long astride, int n, int kcnt,
                     int nsolve, const int* src, double* eval, long estride, double* Z,
                     long zstride, int* cnt, int mfixed, double emax, double* scratch) {
  if (nsolve <= 0) return;
  if (n > VG * VR) throw HipError("syev_lds: n > 128");
  const size_t lds = syev_lds_bytes(n);
  DVR_HIP(hipFuncSetAttribute((const void*)k_syev_lds,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  DVR_LAUNCH("syev_lds", 8.0 * nsolve * (double)n * n, 0.0, k_syev_lds, dim3(nsolve),
             dim3(NT), lds, s, A, astride, n, kcnt, src, eval, estride, Z, zstride, cnt,
             mfixed, emax, scratch);
  DVR_LAUNCH_CHECK();
}

// H1 of every (c,b) point into a dense n x n buffer (mkham1 with ztheta=T,
// D:2180-2193, rounded as the reference: (V + wterm) + L*wsum, unfused).
__global__ void __launch_bounds__(256)
k_build_h1(const double* __restrict__ vgrid, const double* __restrict__ xl,
           const double* __restrict__ wsum, const double* __restrict__ wterm,
           int n, int nsolve, const int* __restrict__ src, double* __restrict__ H) {
#pragma clang fp contract(off)
  const long per = (long)n * n;
  const long total = per * nsolve;
  for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (long)gridDim.x * blockDim.x) {
    const int slot = (int)(idx / per);
    const int r = (int)(idx % per);
    const int j = r / n, i = r - j * n;
    const int p = src ? src[slot] : slot;
    double a = xl[i * n + j] * wsum[p];
    if (i == j) a = (vgrid[(long)p * n + i] + (wterm ? wterm[p] : 0.0)) + a;
    H[(long)p * per + r] = a;
  }
}

void launch_build_h1(hipStream_t s, const double* vgrid, const double* xl,
                     const double* wsum, const double* wterm, int n, int nsolve,
                     const int* src, double* H) {
  const long total = (long)n * n * nsolve;
  int grid = (int)((total + 255) / 256);
  if (grid > 65535 * 16) grid = 65535 * 16;
  DVR_LAUNCH("build_h1", 8.0 * (double)total, 0.0, k_build_h1, dim3(grid), dim3(256), 0, s,
             vgrid, xl, wsum, wterm, n, nsolve, src, H);
  DVR_LAUNCH_CHECK();
}

}  // namespace dvr
