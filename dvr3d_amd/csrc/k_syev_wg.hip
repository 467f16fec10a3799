// Single-work-group dense symmetric eigensolver: one 1024-thread work-group
// solves one problem from start to finish inside one launch -- Householder
// tridiagonalisation, Sturm multisection for the lowest cnt eigenvalues,
// inverse iteration for the lowest m eigenvectors, and back-transformation.
//
// It replaces dsyev('V','L') (source/dvr3drjz_segmented.f90:2638-2659) for
// the many small and medium problems of the path: the 2D blocks (n of a few
// hundred) and the Rayleigh-Ritz problems of the iterative final solve.  The
// multi-launch solver in k_syev.hip keeps the large single problems, where
// a whole-GPU pass per column pays off; here the per-column cost is a few
// work-group barriers instead of two kernel launches.
//
// Numerics are those of k_syev.hip: same reflector convention (v(j+1) = 1,
// H = I - tau v v^T), same Sturm count and pivot guard, same inverse
// iteration with modified Gram-Schmidt inside clusters of close eigenvalues.
#include "dvr_internal.h"
#include <cfloat>

namespace dvr {

namespace {
constexpr int WG = 1024;
constexpr int NW = WG / 64;
constexpr double kEpsW = 2.220446049250313e-16;

__device__ inline double wsum64(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return __shfl(v, 0, 64);
}

__device__ inline double bsum(double v, double* red) {
  v = wsum64(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  double s = 0.0;
  for (int i = 0; i < NW; ++i) s += red[i];
  return s;
}

__device__ inline double umix(unsigned long long x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  x = x ^ (x >> 31);
  return (double)(x >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}
}  // namespace

size_t syev_wg_scratch_doubles(int n, int m) {
  return (size_t)6 * (size_t)n * (size_t)(m > 0 ? m : 1);
}

__global__ void __launch_bounds__(WG)
k_syev_wg(EigBatch b, const int* __restrict__ cnt, const long* __restrict__ eoff,
          double* __restrict__ eval, const int* __restrict__ mvec,
          const long* __restrict__ zoff, double* __restrict__ Z,
          double* __restrict__ scratch, const long* __restrict__ soff,
          int phase) {
  __shared__ double red[NW];
  __shared__ int cl_start[WG + 1];
  __shared__ int ncl_s;
  const int p = blockIdx.x;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int n = b.n[p];
  const long vo = b.voff[p];
  double* A = b.A + b.aoff[p];
  double* d = b.d + vo;
  double* e = b.e + vo;
  double* e2 = b.e2 + vo;
  double* tau = b.tau + vo;
  double* w = b.wbuf + vo;
  double* sv = b.pbuf + vo;
  const int kcnt = cnt[p];
  const int m = mvec ? mvec[p] : 0;
  double* ev = eval + eoff[p];

  double tn;
  if (phase & 1) {
  // ---------------- 1. tridiagonalisation
  if (n == 1) {
    if (tid == 0) d[0] = A[0];
  }
  for (int j = 0; j + 1 < n; ++j) {
    double* vprev = ((j - 1) & 1 ? b.vbuf1 : b.vbuf0) + vo;
    double* vcur = (j & 1 ? b.vbuf1 : b.vbuf0) + vo;
    double* col = A + (long)j * n;
    double part = 0.0;
    double wj = 0.0, vj = 0.0;
    if (j > 0) {
      const double taup = tau[j - 1];
      double kp = 0.0;
      for (int i = j + tid; i < n; i += WG) kp += sv[i] * vprev[i];
      const double K = taup * bsum(kp, red);
      const double alpha = -0.5 * taup * K;
      vj = vprev[j];
      wj = taup * sv[j] + alpha * vj;
      for (int i = j + tid; i < n; i += WG) {
        const double wi = taup * sv[i] + alpha * vprev[i];
        w[i] = wi;
        const double ci = col[i] - (vprev[i] * wj + wi * vj);
        col[i] = ci;
        if (i >= j + 2) part += ci * ci;
      }
    } else {
      for (int i = j + 2 + tid; i < n; i += WG) part += col[i] * col[i];
    }
    const double sigma = bsum(part, red);   // also orders the column writes
    if (j == n - 2) {
      if (tid == 0) {
        d[j] = col[j];
        e[j] = col[n - 1];
        tau[j] = 0.0;
        double dl = A[(n - 1) + (long)(n - 1) * n];
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
    for (int i = j + 1 + tid; i < n; i += WG) {
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
    // trailing square: pending update (j-1) and s = A v_j, one wave per column
    for (int k = j + 1 + wave; k < n; k += NW) {
      double* colk = A + (long)k * n;
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
      acc = wsum64(acc);
      if (lane == 0) sv[k] = acc;
    }
    __syncthreads();
  }

  // ---------------- 2. Gershgorin bounds and e^2
  double lo = DBL_MAX, hi = -DBL_MAX, mx2 = 0.0;
  for (int i = tid; i < n; i += WG) {
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
  __shared__ double rlo[NW], rhi[NW], rm2[NW];
  if (lane == 0) {
    rlo[wave] = lo;
    rhi[wave] = hi;
    rm2[wave] = mx2;
  }
  __syncthreads();
  lo = rlo[0];
  hi = rhi[0];
  mx2 = rm2[0];
  for (int i = 1; i < NW; ++i) {
    lo = fmin(lo, rlo[i]);
    hi = fmax(hi, rhi[i]);
    mx2 = fmax(mx2, rm2[i]);
  }
  tn = fmax(fabs(lo), fabs(hi));
  const double piv = DBL_MIN * fmax(1.0, mx2);
  const double pad = 2.1 * kEpsW * tn * n + 4.0 * piv;
  lo -= pad;
  hi += pad;
  if (tid == 0) {
    b.gl[p] = lo;
    b.gu[p] = hi;
    b.tnrm[p] = tn;
    b.pivmin[p] = piv;
  }

  // ---------------- 3. multisection, one wave per eigenvalue
  for (int k = wave; k < kcnt; k += NW) {
    double l = lo, h = hi;
    for (int it = 0; it < 48; ++it) {
      const double width = h - l;
      if (width <= 2.0 * kEpsW * fmax(fabs(l), fabs(h)) || width <= 4.0 * piv) break;
      const double step = width / 65.0;
      const double x = l + (double)(lane + 1) * step;
      double q = d[0] - x;
      if (fabs(q) < piv) q = -piv;
      int c = q <= 0.0;
      for (int i = 1; i < n; ++i) {
        q = (d[i] - x) - e2[i - 1] / q;
        if (fabs(q) < piv) q = -piv;
        c += q <= 0.0;
      }
      const unsigned long long mask = __ballot(c <= k);
      double nl, nh;
      if (mask == 0ull) {
        nl = l;
        nh = l + step;
      } else {
        const int L = 63 - __clzll((long long)mask);
        nl = l + (double)(L + 1) * step;
        nh = (L == 63) ? h : l + (double)(L + 2) * step;
      }
      if (nl == l && nh == h) break;
      l = nl;
      h = nh;
    }
    if (lane == 0) ev[k] = 0.5 * (l + h);
  }
  __syncthreads();
  } else {
    tn = b.tnrm[p];
  }
  if (!(phase & 2) || m <= 0) return;

  // ---------------- 4. clusters of close eigenvalues among the first m
  if (tid == 0) {
    const double ctol = 1.0e-6 * fmax(tn, 1e-300);
    int nc = 0;
    cl_start[nc++] = 0;
    for (int k = 1; k < m; ++k) {
      if (ev[k] - ev[k - 1] >= ctol && nc < WG) cl_start[nc++] = k;
    }
    cl_start[nc] = m;
    ncl_s = nc;
  }
  __syncthreads();
  const int ncl = ncl_s;
  double* Zp = Z + zoff[p];

  // ---------------- 5. inverse iteration, one thread per cluster
  for (int c = tid; c < ncl; c += WG) {
    const int kb = cl_start[c], ke = cl_start[c + 1];
    double* DL = scratch + soff[p] + (long)c * 6 * n;
    double* DD = DL + n;
    double* DU = DD + n;
    double* DU2 = DU + n;
    double* PV = DU2 + n;
    double* X = PV + n;
    const double tnn = tn > 0.0 ? tn : 1.0;
    const double pertol = 10.0 * kEpsW * tnn, tiny = kEpsW * tnn;
    double lprev = 0.0;
    for (int k = kb; k < ke; ++k) {
      double lam = ev[k];
      if (k > kb && lam - lprev < pertol) lam = lprev + pertol;
      lprev = lam;
      for (int i = 0; i < n; ++i) {
        DD[i] = d[i] - lam;
        if (i < n - 1) {
          DL[i] = e[i];
          DU[i] = e[i];
        }
        DU2[i] = 0.0;
        PV[i] = 0.0;
      }
      for (int i = 0; i < n - 1; ++i) {
        const double di = DD[i], li = DL[i];
        if (fabs(di) >= fabs(li)) {
          if (di != 0.0) {
            const double f = li / di;
            DL[i] = f;
            DD[i + 1] -= f * DU[i];
          } else {
            DD[i] = tiny;
            DL[i] = 0.0;
          }
        } else {
          const double f = di / li;
          DD[i] = li;
          DL[i] = f;
          const double t = DU[i];
          DU[i] = DD[i + 1];
          DD[i + 1] = t - f * DD[i + 1];
          if (i < n - 2) {
            DU2[i] = DU[i + 1];
            DU[i + 1] = -f * DU[i + 1];
          }
          PV[i] = 1.0;
        }
      }
      for (int i = 0; i < n; ++i)
        if (DD[i] == 0.0) DD[i] = tiny;
      for (int i = 0; i < n; ++i)
        X[i] = umix(((unsigned long long)p << 40) ^ ((unsigned long long)k << 20) ^ (unsigned long long)i);
      for (int it = 0; it < 3; ++it) {
        for (int i = 0; i < n - 1; ++i) {
          if (PV[i] == 0.0) {
            X[i + 1] -= DL[i] * X[i];
          } else {
            const double t = X[i];
            X[i] = X[i + 1];
            X[i + 1] = t - DL[i] * X[i];
          }
        }
        X[n - 1] /= DD[n - 1];
        if (n > 1) X[n - 2] = (X[n - 2] - DU[n - 2] * X[n - 1]) / DD[n - 2];
        for (int i = n - 3; i >= 0; --i) X[i] = (X[i] - DU[i] * X[i + 1] - DU2[i] * X[i + 2]) / DD[i];
        double mxv = 0.0;
        for (int i = 0; i < n; ++i) mxv = fmax(mxv, fabs(X[i]));
        const double sc = mxv > 0.0 ? 1.0 / mxv : 1.0;
        for (int i = 0; i < n; ++i) X[i] *= sc;
        for (int q = kb; q < k; ++q) {
          const double* zq = Zp + (long)q * n;
          double dot = 0.0;
          for (int i = 0; i < n; ++i) dot += X[i] * zq[i];
          for (int i = 0; i < n; ++i) X[i] -= dot * zq[i];
        }
        double nrm = 0.0;
        for (int i = 0; i < n; ++i) nrm += X[i] * X[i];
        nrm = nrm > 0.0 ? 1.0 / sqrt(nrm) : 0.0;
        for (int i = 0; i < n; ++i) X[i] *= nrm;
      }
      double* zk = Zp + (long)k * n;
      for (int i = 0; i < n; ++i) zk[i] = X[i];
    }
  }
  __syncthreads();

  // ---------------- 6. back-transformation, one wave per vector
  for (int q = wave; q < m; q += NW) {
    double* z = Zp + (long)q * n;
    for (int j = n - 3; j >= 0; --j) {
      const double tj = tau[j];
      if (tj == 0.0) continue;
      const double* v = A + (long)j * n;   // v(j+1) = 1, v(i) = A(i,j) below
      double s = (lane == 0) ? z[j + 1] : 0.0;
      for (int i = j + 2 + lane; i < n; i += 64) s += v[i] * z[i];
      s = tj * wsum64(s);
      if (lane == 0) z[j + 1] -= s;
      for (int i = j + 2 + lane; i < n; i += 64) z[i] -= s * v[i];
      __threadfence_block();   // the next reflector reads other lanes' z
    }
  }
}

void launch_syev_wg(hipStream_t s, const EigBatch& b, const int* cnt,
                    const long* eoff, double* eval, const int* mvec,
                    const long* zoff, double* Z, double* scratch,
                    const long* soff, int phase) {
  if (b.nprob <= 0) return;
  DVR_LAUNCH(phase & 1 ? "syev_wg_values" : "syev_wg_vectors", 0.0, 0.0, k_syev_wg, dim3(b.nprob), dim3(WG), 0, s, b, cnt, eoff, eval,
                   mvec, zoff, Z, scratch, soff, phase);
  DVR_LAUNCH_CHECK();
}

}  // namespace dvr
