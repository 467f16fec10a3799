// Batched 1D angular eigenproblems of the theta-first DVR3DRJZ path.
//
// Reference: mkham1 (source/dvr3drjz_segmented.f90:2121-2225) builds, at each
// (r1, r2) grid point, H1(k,k') = delta_kk' (V(r1,r2,x_k) + wterm)
//                                + (w1 + w2) L(k,k'),
// then diag -> dsyev (D:1987, D:2638-2659) and cut1d (D:2870-2938) keeps the
// ascending eigenpairs with E <= EMAX1.
//
// Here one workgroup owns one problem.  The matrix (and, for n <= 96, the
// eigenvector accumulator) lives in LDS; the eigensolver is a parallel cyclic
// Jacobi method with the round-robin pair ordering: each step rotates n/2
// disjoint (p,q) pairs at once, and every 2x2 block (pair u rows, pair v
// columns) is updated by a single work item, which writes the block and its
// transpose so the matrix stays exactly symmetric.
#include "dvr_internal.h"
#include <cfloat>

namespace dvr {

namespace {
constexpr int kThreads = 512;
constexpr int kMaxSweeps = 40;
constexpr double kPad = 1.0e300;  // diagonal of the padding row (odd n)

// Round-robin ("chess tournament") pairing of np players in step s < np-1:
// player np-1 meets s, and (s+k, s-k) mod (np-1) meet for k = 1..np/2-1.
__device__ inline void pair_of(int s, int k, int np, int& p, int& q) {
  const int m1 = np - 1;
  int pa, pb;
  if (k == 0) {
    pa = m1;
    pb = s;
  } else {
    pa = s + k;
    if (pa >= m1) pa -= m1;
    pb = s - k;
    if (pb < 0) pb += m1;
  }
  p = pa < pb ? pa : pb;
  q = pa < pb ? pb : pa;
}
}  // namespace

__global__ void __launch_bounds__(kThreads)
k_jacobi1d(const double* __restrict__ vgrid, const double* __restrict__ xl,
           const double* __restrict__ wsum, const double* __restrict__ wterm,
           int n, double emaxau, double* __restrict__ e1,
           double* __restrict__ v1, int* __restrict__ cnt,
           double* __restrict__ vscratch, int vlds,
           const int* __restrict__ src_prob, int* __restrict__ sweeps_out) {
  extern __shared__ double smem[];
  const int p = src_prob ? src_prob[blockIdx.x] : (int)blockIdx.x;
  const int tid = threadIdx.x;
  const int np = n + (n & 1);
  const int m = np / 2;
  const int lda = np + 1;
  double* A = smem;
  double* Vt = vlds ? (A + (long)np * lda) : (vscratch + (long)blockIdx.x * np * np);
  double* rest = vlds ? (Vt + (long)np * np) : (A + (long)np * lda);
  double* Cs = rest;
  double* Sn = Cs + m;
  double* Tt = Sn + m;
  int* Pp = reinterpret_cast<int*>(Tt + m);
  int* Qq = Pp + m;
  int* Mode = Qq + m;   // 0 none, 1 rotate, 2 zero a negligible apq
  int* flag = Mode + m;
  int* rank = flag + 1;
  unsigned* tri = reinterpret_cast<unsigned*>(rank + n);  // (u<<16)|v, u <= v
  const int ntri = m * (m + 1) / 2;
  for (int u = tid; u < m; u += kThreads) {
    const int base = u * m - u * (u - 1) / 2;  // first slot of row u
    for (int v = u; v < m; ++v) tri[base + (v - u)] = ((unsigned)u << 16) | (unsigned)v;
  }

  // ---- H1 in LDS (mkham1 with ztheta=T, D:2180-2193), rounded like the
  // reference: (V + wterm) + L*wsum with an unfused product
  const double ws = wsum[p];
  const double wt = wterm ? wterm[p] : 0.0;
  for (int idx = tid; idx < np * np; idx += kThreads) {
    const int i = idx / np;
    const int j = idx - i * np;
    double a;
    if (i < n && j < n) {
#pragma clang fp contract(off)
      a = xl[i * n + j] * ws;
      if (i == j) a = (vgrid[(long)p * n + i] + wt) + a;
    } else {
      a = (i == j) ? kPad : 0.0;
    }
    A[i * lda + j] = a;
    Vt[(long)i * np + j] = (i == j) ? 1.0 : 0.0;
  }
  __syncthreads();

  int sweep = 0;
  for (; sweep < kMaxSweeps; ++sweep) {
    if (tid == 0) *flag = 0;
    __syncthreads();
    for (int s = 0; s < np - 1; ++s) {
      // rotation parameters for the m disjoint pairs of this step
      for (int k = tid; k < m; k += kThreads) {
        int pp, qq;
        pair_of(s, k, np, pp, qq);
        const double apq = A[pp * lda + qq];
        const double app = A[pp * lda + pp];
        const double aqq = A[qq * lda + qq];
        int mode = 0;
        double c = 1.0, sn = 0.0, t = 0.0;
        if (apq != 0.0) {
          const double g = 100.0 * fabs(apq);
          if (fabs(app) + g == fabs(app) && fabs(aqq) + g == fabs(aqq)) {
            mode = 2;
          } else {
            const double theta = (aqq - app) / (2.0 * apq);
            if (fabs(theta) > 1.0e150) {
              t = 0.5 / theta;
            } else {
              t = 1.0 / (fabs(theta) + sqrt(theta * theta + 1.0));
              if (theta < 0.0) t = -t;
            }
            c = 1.0 / sqrt(t * t + 1.0);
            sn = t * c;
            mode = 1;
            *flag = 1;
          }
        }
        Pp[k] = pp;
        Qq[k] = qq;
        Cs[k] = c;
        Sn[k] = sn;
        Tt[k] = t;
        Mode[k] = mode;
      }
      __syncthreads();
      // A <- J^T A J, one 2x2 block per work item (u <= v)
      for (int w = tid; w < ntri; w += kThreads) {
        const unsigned uv = tri[w];
        const int u = (int)(uv >> 16);
        const int v = (int)(uv & 0xffffu);
        const int p1 = Pp[u], q1 = Qq[u];
        if (u == v) {
          const int mode = Mode[u];
          if (mode == 1) {
            const double apq = A[p1 * lda + q1];
            const double t = Tt[u];
            A[p1 * lda + p1] -= t * apq;
            A[q1 * lda + q1] += t * apq;
          }
          if (mode != 0) {
            A[p1 * lda + q1] = 0.0;
            A[q1 * lda + p1] = 0.0;
          }
          continue;
        }
        const int mu = Mode[u], mv = Mode[v];
        if (mu != 1 && mv != 1) continue;
        const int p2 = Pp[v], q2 = Qq[v];
        const double c1 = Cs[u], s1 = Sn[u], c2 = Cs[v], s2 = Sn[v];
        const double x = A[p1 * lda + p2], y = A[p1 * lda + q2];
        const double z = A[q1 * lda + p2], ww = A[q1 * lda + q2];
        const double x1 = c2 * x - s2 * y, y1 = s2 * x + c2 * y;
        const double z1 = c2 * z - s2 * ww, w1 = s2 * z + c2 * ww;
        const double X = c1 * x1 - s1 * z1, Z = s1 * x1 + c1 * z1;
        const double Y = c1 * y1 - s1 * w1, W = s1 * y1 + c1 * w1;
        A[p1 * lda + p2] = X;
        A[p1 * lda + q2] = Y;
        A[q1 * lda + p2] = Z;
        A[q1 * lda + q2] = W;
        A[p2 * lda + p1] = X;
        A[q2 * lda + p1] = Y;
        A[p2 * lda + q1] = Z;
        A[q2 * lda + q1] = W;
      }
      // V <- V J  (rows of V^T)
      const int lane = tid & 63, wv = tid >> 6;
      for (int v = wv; v < m; v += kThreads / 64) {
        if (Mode[v] != 1) continue;
        for (int i = lane; i < np; i += 64) {
        const double c = Cs[v], sn = Sn[v];
        double* vp = Vt + (long)Pp[v] * np + i;
        double* vq = Vt + (long)Qq[v] * np + i;
        const double a = *vp, b = *vq;
        *vp = c * a - sn * b;
        *vq = sn * a + c * b;
        }
      }
      __syncthreads();
    }
    const int f = *flag;
    __syncthreads();
    if (f == 0) break;
  }

  // ---- ascending order (rank by value, index breaks ties) and cut1d
  for (int i = tid; i < n; i += kThreads) {
    const double di = A[i * lda + i];
    int r = 0;
    for (int j = 0; j < n; ++j) {
      const double dj = A[j * lda + j];
      r += (dj < di) || (dj == di && j < i);
    }
    rank[i] = r;
  }
  __syncthreads();
  for (int idx = tid; idx < n * n; idx += kThreads) {
    const int i = idx / n;
    const int a = idx - i * n;
    v1[((long)p * n + rank[i]) * n + a] = Vt[(long)i * np + a];
  }
  if (tid < 64) {
    int kept = 0;
    for (int i = tid; i < n; i += 64) {
      const double di = A[i * lda + i];
      e1[(long)p * n + rank[i]] = di;
      kept += !(di > emaxau);
    }
    for (int off = 32; off > 0; off >>= 1) kept += __shfl_down(kept, off, 64);
    if (tid == 0) {
      cnt[p] = kept;
      if (sweeps_out) sweeps_out[p] = sweep;
    }
  }
}

static size_t jacobi_lds_bytes(int n, bool vlds) {
  const int np = n + (n & 1);
  const int m = np / 2;
  size_t b = (size_t)np * (np + 1) * 8;
  if (vlds) b += (size_t)np * np * 8;
  b += (size_t)m * 3 * 8 + (size_t)m * 3 * 4 + 4 + (size_t)n * 4;
  b += (size_t)m * (m + 1) / 2 * 4;
  return b;
}

static bool jacobi_vlds(int n) { return jacobi_lds_bytes(n, true) <= 160 * 1024; }

// Mirror test of the 1D inputs: problem (c,b) equals problem (b,c) bit for
// bit when V(c,b,:) == V(b,c,:), wsum and wterm are symmetric (square grid).
__global__ void __launch_bounds__(256)
k_mirror_check(const double* __restrict__ v, const double* __restrict__ wsum,
               const double* __restrict__ wterm, int nb, int na,
               int* __restrict__ flag) {
  const long total = (long)nb * nb * na;
  for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (long)gridDim.x * blockDim.x) {
    const int a = (int)(idx % na);
    const long cb = idx / na;
    const int b = (int)(cb % nb);
    const int c = (int)(cb / nb);
    if (c <= b) continue;
    bool diff = v[idx] != v[((long)b * nb + c) * na + a];
    if (a == 0) {
      diff |= wsum[c * nb + b] != wsum[b * nb + c];
      if (wterm) diff |= wterm[c * nb + b] != wterm[b * nb + c];
    }
    if (diff) *flag = 1;
  }
}

// Copy the solution of problem (b,c), b < c, to its mirror (c,b).
__global__ void __launch_bounds__(256)
k_mirror_copy(int nb, int n, double* __restrict__ e1, double* __restrict__ v1,
              int* __restrict__ cnt) {
  const long per = (long)n * n + n + 1;
  const long total = (long)nb * nb * per;
  for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (long)gridDim.x * blockDim.x) {
    const long pp = idx / per;
    const long r = idx % per;
    const int c = (int)(pp / nb), b = (int)(pp % nb);
    if (c >= b) continue;          // destination (c,b) with c < b
    const long dst = (long)c * nb + b, src = (long)b * nb + c;
    if (r < (long)n * n)
      v1[dst * n * n + r] = v1[src * n * n + r];
    else if (r < (long)n * n + n)
      e1[dst * n + (r - (long)n * n)] = e1[src * n + (r - (long)n * n)];
    else
      cnt[dst] = cnt[src];
  }
}

void launch_mirror_check(hipStream_t s, const double* v, const double* wsum,
                         const double* wterm, int nb, int na, int* flag) {
  const long total = (long)nb * nb * na;
  int grid = (int)((total + 255) / 256);
  if (grid > 65535 * 8) grid = 65535 * 8;
  DVR_HIP(hipMemsetAsync(flag, 0, sizeof(int), s));
  hipLaunchKernelGGL(k_mirror_check, dim3(grid), dim3(256), 0, s, v, wsum, wterm,
                     nb, na, flag);
  DVR_LAUNCH_CHECK();
}

void launch_mirror_copy(hipStream_t s, int nb, int n, double* e1, double* v1,
                        int* cnt) {
  const long total = (long)nb * nb * ((long)n * n + n + 1);
  int grid = (int)((total + 255) / 256);
  if (grid > 65535 * 16) grid = 65535 * 16;
  DVR_LAUNCH("mirror_copy", 16.0 * (double)nb * (nb - 1) / 2 * ((double)n * n + n), 0.0, k_mirror_copy, dim3(grid), dim3(256), 0, s, nb, n, e1, v1, cnt);
  DVR_LAUNCH_CHECK();
}

size_t jacobi1d_scratch_doubles(int n, int nprob) {
  if (jacobi_vlds(n)) return 0;
  const size_t np = n + (n & 1);
  return np * np * (size_t)nprob;
}

void launch_jacobi1d(hipStream_t s, const double* vgrid, const double* xl,
                     const double* wsum, const double* wterm, int n, int nprob,
                     double emaxau, double* e1, double* v1, int* cnt,
                     double* vscratch, const int* src_prob, int* sweeps) {
  if (nprob <= 0) return;
  const bool vlds = jacobi_vlds(n);
  const size_t lds = jacobi_lds_bytes(n, vlds);
  if (lds > 160 * 1024)
    throw HipError("angular DVR size too large for the LDS-resident solver");
  if (!vlds && !vscratch) throw HipError("jacobi1d: scratch required");
  DVR_HIP(hipFuncSetAttribute((const void*)k_jacobi1d,
                              hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)lds));
  (void)0;
  DVR_LAUNCH("jacobi1d", 8.0 * nprob * ((double)n * n + 3.0 * n), 0.0, k_jacobi1d, dim3(nprob), dim3(kThreads), lds, s, vgrid, xl,
                   wsum, wterm, n, emaxau, e1, v1, cnt, vscratch, vlds ? 1 : 0,
                   src_prob, sweeps);
  DVR_LAUNCH_CHECK();
}

}  // namespace dvr
