// Inverse iteration for one cluster of eigenvalues of a symmetric tridiagonal
// matrix, executed by a single thread (device code shared by k_syev.hip and
// k_syev_lds.hip).
//
// For each eigenvalue lam_k of the cluster [kb, ke): T - lam I = P L U with
// partial pivoting (the dgttrf recurrence), three solves from a hashed start
// vector, modified Gram-Schmidt against the earlier vectors of the cluster
// (eigenvalues closer than 10 eps ||T|| are separated first, as in dstein).
// The scratch arrays (multipliers, reciprocal pivots, the two superdiagonals
// of U, the iterate) are strided so that the threads of a wavefront that
// handle neighbouring clusters touch one contiguous line per step; the values
// carried along the recurrences stay in registers, so the only dependent
// chain is arithmetic.
#pragma once

#include <type_traits>

namespace dvr {

__device__ inline double ii_hash(unsigned long long x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  x = x ^ (x >> 31);
  return (double)(x >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

// s0: this thread's scratch; array a (0..5) starts at s0 + a*astride and its
// element i is at + i*stride.  Vector k is written to Zp + k*n.
//
// The three recurrences (factorisation, forward and backward sweeps) move
// CH elements between memory and registers per burst.  Full bursts carry no
// per-element guards, and the pivoting choice is a select rather than a
// branch, so a burst is straight-line code whose only dependent chain is the
// arithmetic of the recurrence.
template <int CH>
__device__ __forceinline__ void inverse_iteration_cluster(int n, const double* __restrict__ d,
                                                          const double* __restrict__ e,
                                                          const double* __restrict__ lams,
                                                          int kb, int ke, double tn, double* s0,
                                                          long astride, long stride,
                                                          double* Zp, unsigned long long seed) {
  double* FL = s0;
  double* FD = s0 + astride;
  double* FU = s0 + 2 * astride;
  double* FU2 = s0 + 3 * astride;
  double* X = s0 + 4 * astride;
  double* PV = s0 + 5 * astride;   // 1 where rows i and i+1 were interchanged
#define II(arr, i) (arr)[(long)(i) * stride]
  constexpr double eps = 2.220446049250313e-16;
  const double tnn = tn > 0.0 ? tn : 1.0;
  const double pertol = 10.0 * eps * tnn, tiny = eps * tnn;
  const int nm1 = n - 1;                 // elimination steps 0 .. n-2
  const int nfull = nm1 / CH * CH;       // steps covered by full bursts
  double lprev = 0.0;
  for (int k = kb; k < ke; ++k) {
    double lam = lams[k];
    if (k > kb && lam - lprev < pertol) lam = lprev + pertol;
    lprev = lam;
    // factorisation T - lam I = P L U (the dgttrf recurrence): step i
    // eliminates row i+1 with row i or, when |T(i+1,i)| > |pivot|, after
    // interchanging them
    double dc = d[0] - lam, uc = n > 1 ? e[0] : 0.0;
    auto fact_burst = [&](int i0, auto full) {
      constexpr bool F = decltype(full)::value;
      double ev[CH], dv[CH];
#pragma unroll
      for (int q = 0; q < CH; ++q) {
        const int i = i0 + q;
        ev[q] = (F || i < nm1) ? e[i] : 0.0;
        dv[q] = (F || i < nm1) ? d[i + 1] : 0.0;
      }
      const double etail = (i0 + CH < nm1) ? e[i0 + CH] : 0.0;
#pragma unroll
      for (int q = 0; q < CH; ++q) {
        const int i = i0 + q;
        if (F || i < nm1) {
          const double li = ev[q];
          const double dn = dv[q] - lam;
          const double enext = (q + 1 < CH) ? ev[q + 1] : etail;
          const double un = (i < n - 2) ? enext : 0.0;
          const bool sw = fabs(dc) < fabs(li);
          const double den = sw ? li : dc, num = sw ? dc : li;
          const double rden = rcp_refined(den == 0.0 ? tiny : den);
          const double fl = num * rden;   // den == 0 only with li == 0
          II(FL, i) = fl;
          II(FD, i) = rden;
          II(FU, i) = sw ? dn : uc;
          II(FU2, i) = sw ? un : 0.0;
          II(PV, i) = sw ? 1.0 : 0.0;
          dc = sw ? uc - fl * dn : dn - fl * uc;
          uc = sw ? -fl * un : un;
        }
      }
    };
    for (int i0 = 0; i0 < nfull; i0 += CH) fact_burst(i0, std::true_type{});
    if (nfull < nm1) fact_burst(nfull, std::false_type{});
    II(FD, n - 1) = 1.0 / (dc == 0.0 ? tiny : dc);
    double scale = 1.0;
    for (int it = 0; it < 3; ++it) {
      const unsigned long long base = (seed << 40) ^ ((unsigned long long)k << 20);
      // forward sweep: L^-1 with the row interchanges
      double xc = it == 0 ? ii_hash(base) : II(X, 0) * scale;
      auto fwd_burst = [&](int i0, auto full) {
        constexpr bool F = decltype(full)::value;
        double fl[CH], xn[CH], sw[CH];
#pragma unroll
        for (int q = 0; q < CH; ++q) {
          const int i = i0 + q;
          if (F || i < nm1) {
            fl[q] = II(FL, i);
            sw[q] = II(PV, i);
            xn[q] = it == 0 ? ii_hash(base ^ (unsigned long long)(i + 1)) : II(X, i + 1) * scale;
          }
        }
#pragma unroll
        for (int q = 0; q < CH; ++q) {
          const int i = i0 + q;
          if (F || i < nm1) {
            const bool s = sw[q] != 0.0;
            const double a = s ? xn[q] : xc, bq = s ? xc : xn[q];
            xc = bq - fl[q] * a;
            II(X, i) = a;
          }
        }
      };
      for (int i0 = 0; i0 < nfull; i0 += CH) fwd_burst(i0, std::true_type{});
      if (nfull < nm1) fwd_burst(nfull, std::false_type{});
      // backward sweep: U^-1, tracking max |x|
      double x1 = xc * II(FD, n - 1);
      II(X, n - 1) = x1;
      double mxv = fabs(x1);
      double x0 = 0.0;
      if (n > 1) {
        x0 = (II(X, n - 2) - II(FU, n - 2) * x1) * II(FD, n - 2);
        II(X, n - 2) = x0;
        mxv = fmax(mxv, fabs(x0));
      }
      // rows n-3 .. 0, CH at a time from the top end
      auto bwd_burst = [&](int i1, auto full) {
        constexpr bool F = decltype(full)::value;
        double xv[CH], fu[CH], fu2[CH], fd[CH];
#pragma unroll
        for (int q = 0; q < CH; ++q) {
          const int i = i1 - q;
          if (F || i >= 0) {
            xv[q] = II(X, i);
            fu[q] = II(FU, i);
            fu2[q] = II(FU2, i);
            fd[q] = II(FD, i);
          }
        }
#pragma unroll
        for (int q = 0; q < CH; ++q) {
          if (F || i1 - q >= 0) {
            const double xi = (xv[q] - fu[q] * x0 - fu2[q] * x1) * fd[q];
            mxv = fmax(mxv, fabs(xi));
            x1 = x0;
            x0 = xi;
            II(X, i1 - q) = xi;
          }
        }
      };
      int i1 = n - 3;
      for (; i1 - (CH - 1) >= 0; i1 -= CH) bwd_burst(i1, std::true_type{});
      if (i1 >= 0) bwd_burst(i1, std::false_type{});
      const double sc = mxv > 0.0 ? 1.0 / mxv : 1.0;
      if (k > kb) {
        for (int i = 0; i < n; ++i) II(X, i) *= sc;
        for (int q = kb; q < k; ++q) {
          const double* zq = Zp + (long)q * n;
          double dot = 0.0;
          for (int i = 0; i < n; ++i) dot += II(X, i) * zq[i];
          for (int i = 0; i < n; ++i) II(X, i) -= dot * zq[i];
        }
        double nrm = 0.0;
        for (int i = 0; i < n; ++i) nrm += II(X, i) * II(X, i);
        scale = nrm > 0.0 ? 1.0 / sqrt(nrm) : 0.0;
      } else {
        double n0 = 0.0, n1 = 0.0, n2 = 0.0, n3 = 0.0;
        int i = 0;
        for (; i + 3 < n; i += 4) {
          const double t0 = II(X, i) * sc, t1 = II(X, i + 1) * sc;
          const double t2 = II(X, i + 2) * sc, t3 = II(X, i + 3) * sc;
          n0 += t0 * t0;
          n1 += t1 * t1;
          n2 += t2 * t2;
          n3 += t3 * t3;
        }
        for (; i < n; ++i) {
          const double t = II(X, i) * sc;
          n0 += t * t;
        }
        const double nrm = (n0 + n1) + (n2 + n3);
        scale = nrm > 0.0 ? sc / sqrt(nrm) : 0.0;
      }
    }
    double* zk = Zp + (long)k * n;
    for (int i = 0; i < n; ++i) zk[i] = II(X, i) * scale;
  }
#undef II
}

}  // namespace dvr
