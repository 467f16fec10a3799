// Batched dense symmetric eigensolver in FP64: the replacement for the
// reference's  diag -> dsyev('V','L', ...)  calls
// (source/dvr3drjz_segmented.f90:2638-2659; call sites D:1987, 2030, 2699).
//
// Stages (each batched over independent problems of unequal size):
//   1. Householder reduction to tridiagonal form, A = Q T Q^T, one column at
//      a time.  Per column j two launches:
//        k_tri_col : finishes w_{j-1} = p - (tau/2)(p.v) v, applies the
//                    pending rank-2 update to column j only, and forms the
//                    reflector (v_j, tau_j) that annihilates A(j+2:n, j);
//        k_tri_upd : applies the pending rank-2 update (v_{j-1}, w_{j-1}) to
//                    the trailing columns and, in the same pass over the
//                    matrix, forms p_j = tau_j A v_j.
//      The reflectors stay in the lower part of A (v(j+1) = 1 implied).
//   2. Gershgorin bounds; eigenvalues by Sturm-count multisection: one
//      wavefront per wanted eigenvalue, 64 shifts per step.
//   3. Eigenvectors of T by inverse iteration (tridiagonal LU with partial
//      pivoting), with modified Gram-Schmidt inside clusters of close
//      eigenvalues; one thread per cluster.
//   4. Back-transformation Z <- Q Z in blocks of NB reflectors in WY form,
//      Q_blk = I - V T V^T, one work-group per (problem, 8-column tile).
#include "dvr_internal.h"
#include "inviter.h"
#include <cfloat>

namespace dvr {

namespace {
constexpr int TB = 256;
constexpr int NB = 32;   // reflectors per WY block
constexpr int CT = 8;    // columns per back-transform tile
constexpr int RC = 64;   // rows per staged chunk in the back-transform
constexpr double kEps = 2.220446049250313e-16;

__device__ inline double block_sum(double v, double* red) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  double s = 0.0;
  for (int i = 0; i < TB / 64; ++i) s += red[i];
  return s;
}

__device__ inline double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}
}  // namespace

int backtransform_nb() { return NB; }

// ---------------------------------------------------------------------------
// 1. tridiagonalisation
// ---------------------------------------------------------------------------
// k_tri_col uses a larger work-group: it is latency bound (three dependent
// reductions over one column per step).
constexpr int TC = 1024;

__device__ inline double block_sum_n(double v, double* red, int nthreads) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  double s = 0.0;
  for (int i = 0; i < nthreads / 64; ++i) s += red[i];
  return s;
}

// The last TAIL columns of every problem (all of a problem of n <= TAIL) are
// reduced by one work-group with the trailing TAIL x TAIL block in LDS
// (k_tri_tail): one launch instead of two per column where the columns are
// short and every launch is latency.
constexpr int TAIL = 128;
__device__ __host__ inline int tail_start(int n) { return n > TAIL ? n - TAIL : 0; }
bool tri_fused_enabled();

// pbuf holds s = A v (unscaled); p = tau s.
__global__ void __launch_bounds__(TC) k_tri_col(EigBatch b, int j) {
  __shared__ double red[TC / 64];
  const int p = blockIdx.x;
  const int tid = threadIdx.x;
  const int n = b.n[p];
  const long vo = b.voff[p];
  double* A = b.A + b.aoff[p];
  if (j >= tail_start(n)) return;   // k_tri_tail's part
  double* vprev = ((j - 1) & 1 ? b.vbuf1 : b.vbuf0) + vo;
  double* vcur = (j & 1 ? b.vbuf1 : b.vbuf0) + vo;
  double* w = b.wbuf + vo;
  double* sv = b.pbuf + vo;
  double* col = A + (long)j * n;
  // pending update of step j-1 on column j, fused with the norm of the part
  // of the column the new reflector annihilates
  double part = 0.0;
  if (j > 0) {
    const double taup = b.tau[vo + j - 1];
    double kp = 0.0;
    for (int i = j + tid; i < n; i += TC) kp += sv[i] * vprev[i];
    const double K = taup * block_sum_n(kp, red, TC);
    const double alpha = -0.5 * taup * K;
    const double vj = vprev[j];
    const double wj = taup * sv[j] + alpha * vj;
    for (int i = j + tid; i < n; i += TC) {
      const double wi = taup * sv[i] + alpha * vprev[i];
      w[i] = wi;
      const double ci = col[i] - (vprev[i] * wj + wi * vj);
      col[i] = ci;
      if (i >= j + 2) part += ci * ci;
    }
    if (j == n - 2) __syncthreads();  // w(n-1) is read below by thread 0
  } else {
    for (int i = j + 2 + tid; i < n; i += TC) part += col[i] * col[i];
  }
  if (j == n - 2) {
    if (tid == 0) {
      b.d[vo + j] = col[j];
      b.e[vo + j] = col[n - 1];
      b.tau[vo + j] = 0.0;
      double dl = A[(n - 1) + (long)(n - 1) * n];
      if (j > 0) dl -= (vprev[n - 1] * w[n - 1] + w[n - 1] * vprev[n - 1]);
      b.d[vo + n - 1] = dl;
    }
    return;
  }
  const double sigma = block_sum_n(part, red, TC);
  const double alpha = col[j + 1];
  double tau, beta, scal;
  if (sigma == 0.0) {
    tau = 0.0;
    beta = alpha;
    scal = 0.0;
  } else {
    const double mu = sqrt(alpha * alpha + sigma);
    beta = (alpha <= 0.0) ? mu : -mu;
    tau = (beta - alpha) / beta;
    scal = 1.0 / (alpha - beta);
  }
  for (int i = j + 1 + tid; i < n; i += TC) {
    const double vi = (i == j + 1) ? 1.0 : col[i] * scal;
    vcur[i] = vi;
    if (i > j + 1) col[i] = vi;
  }
  __syncthreads();
  if (tid == 0) {
    b.d[vo + j] = col[j];
    b.e[vo + j] = beta;
    b.tau[vo + j] = tau;
    col[j + 1] = beta;
  }
}

// Trailing update + matrix-vector product over the full trailing square
// A(j+1:n, j+1:n), both triangles kept current so every column is contiguous:
//   A(i,k) -= v(i) w(k) + w(i) v(k)      (pending update of step j-1)
//   s(k)    = sum_i A(i,k) v_j(i)
// Each wavefront owns two columns and walks their rows together, so the three
// row vectors (v_{j-1}, w, v_j) are loaded once per row for both columns; each
// lane keeps four rows of both columns in flight.  s(k) is complete within the
// wave, so no atomics are needed.
constexpr int UC = 8;   // columns per work-group (2 per wave)

__global__ void __launch_bounds__(TB) k_tri_upd(EigBatch b, int j, int pending) {
  const int bid = blockIdx.x;
  int p = 0, base = 0;
  for (; p < b.nprob; ++p) {
    const int n = b.n[p];
    const int cnt = (j <= n - 3 && j < tail_start(n)) ? (n - j - 1 + UC - 1) / UC : 0;
    if (bid < base + cnt) break;
    base += cnt;
  }
  if (p >= b.nprob) return;
  const int n = b.n[p];
  const long vo = b.voff[p];
  double* A = b.A + b.aoff[p];
  const double* vprev = ((j - 1) & 1 ? b.vbuf1 : b.vbuf0) + vo;
  const double* vcur = (j & 1 ? b.vbuf1 : b.vbuf0) + vo;
  const double* w = b.wbuf + vo;
  double* sv = b.pbuf + vo;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int k0 = j + 1 + (bid - base) * UC + 2 * wave;
  if (k0 >= n) return;
  const bool two = k0 + 1 < n;
  double* ca = A + (long)k0 * n;
  double* cb = two ? A + (long)(k0 + 1) * n : ca;
  const int i0 = j + 1;
  double sa0 = 0.0, sa1 = 0.0, sb0 = 0.0, sb1 = 0.0;
  if (j > 0 && pending) {
    const double wa = w[k0], pa = vprev[k0];
    const double wb = two ? w[k0 + 1] : 0.0, pb = two ? vprev[k0 + 1] : 0.0;
    int i = i0 + lane;
    for (; i + 192 < n; i += 256) {
      double x[4], y[4], vp[4], ww[4], vc[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        x[q] = ca[i + 64 * q];
        y[q] = cb[i + 64 * q];
        vp[q] = vprev[i + 64 * q];
        ww[q] = w[i + 64 * q];
        vc[q] = vcur[i + 64 * q];
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const double ua = x[q] - (vp[q] * wa + ww[q] * pa);
        const double ub = y[q] - (vp[q] * wb + ww[q] * pb);
        ca[i + 64 * q] = ua;
        if (two) cb[i + 64 * q] = ub;
        if (q & 1) {
          sa1 += ua * vc[q];
          sb1 += ub * vc[q];
        } else {
          sa0 += ua * vc[q];
          sb0 += ub * vc[q];
        }
      }
    }
    for (; i < n; i += 64) {
      const double vpi = vprev[i], wi = w[i], vci = vcur[i];
      const double ua = ca[i] - (vpi * wa + wi * pa);
      const double ub = cb[i] - (vpi * wb + wi * pb);
      ca[i] = ua;
      if (two) cb[i] = ub;
      sa0 += ua * vci;
      sb0 += ub * vci;
    }
  } else {
    int i = i0 + lane;
    for (; i + 192 < n; i += 256) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const double vc = vcur[i + 64 * q];
        if (q & 1) {
          sa1 += ca[i + 64 * q] * vc;
          sb1 += cb[i + 64 * q] * vc;
        } else {
          sa0 += ca[i + 64 * q] * vc;
          sb0 += cb[i + 64 * q] * vc;
        }
      }
    }
    for (; i < n; i += 64) {
      const double vc = vcur[i];
      sa0 += ca[i] * vc;
      sb0 += cb[i] * vc;
    }
  }
  const double ra = wave_sum(sa0 + sa1);
  const double rb = wave_sum(sb0 + sb1);
  if (lane == 0) {
    sv[k0] = ra;
    if (two) sv[k0 + 1] = rb;
  }
}

// Blocked reduction of the large problems (the form of LAPACK's dsytrd /
// dlatrd): the trailing square stays as it was at the start of a panel of
// NBP columns, whose reflectors v and vectors w are kept in the panel buffer
// [V | W | V] (n_p rows, 3 NBP columns per problem).  Per column j (slot
// c = j - j0 of the panel starting at j0) two launches:
//   k_tri_pcol : w of column j-1 from s = A_panel v_{j-1} (k_tri_upd without
//                its update) with the panel's corrections,
//                  w = tau p + alpha v,  p = s - V (W^T v) - W (V^T v),
//                  alpha = -(tau^2 / 2) p.v;
//                column j updated by the panel's reflectors so far, and its
//                reflector (v_j, tau_j) formed as in k_tri_col;
//   k_tri_upd  : s_j = A_panel v_j over the trailing square (read only).
// After the panel, k_tri_pcol<true> finishes the last w, and one NT GEMM per
// problem applies A -= [V W] [W V]^T to the trailing square (both
// triangles).  The trailing square is read once per column and written once
// per panel, where the column-by-column form reads and writes it per column.
// Problems whose columns end inside the panel (the LDS tail's start) get zero
// slots, so their GEMM adds nothing for them.
constexpr int NBP = 32;
int tri_panel_width() { return NBP; }

template <bool FIN>
__global__ void __launch_bounds__(TC) k_tri_pcol(EigBatch b, int j, int j0) {
  __shared__ double red[TC / 64];
  __shared__ double part2[2 * NBP][TC / 64];
  __shared__ double ys[2 * NBP];   // W^T v, V^T v of the previous column
  __shared__ double cj[2 * NBP];   // W(j, c'), V(j, c') of this column
  const int p = blockIdx.x, tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int n = b.n[p];
  const long vo = b.voff[p];
  const int ts = tail_start(n);
  double* A = b.A + b.aoff[p];
  double* V = b.pan + 3L * NBP * vo;
  double* W = V + (long)NBP * n;
  double* V2 = V + 2L * NBP * n;
  const int c = j - j0;
  // the last real column of this problem in the panel is min(j0 + NBP, ts) - 1;
  // its w is finished at column min(j0 + NBP, ts) (or by the FIN launch)
  const int jreal = min(j0 + NBP, ts);
  const bool finish = c > 0 && j - 1 < jreal && j <= jreal;
  if (finish) {
    // w of column j - 1 (slot c - 1); v_{j-1} lives in rows >= j
    const int cp = c - 1;
    const double* v = V + (long)cp * n;
    const double* sv = b.pbuf + vo;
    const double taup = b.tau[vo + j - 1];
    // W(:, c')^T v and V(:, c')^T v for c' < cp
    for (int q = 0; q < cp; ++q) {
      double a1 = 0.0, a2 = 0.0;
      for (int i = j + tid; i < n; i += TC) {
        const double vi = v[i];
        a1 += W[i + (long)q * n] * vi;
        a2 += V[i + (long)q * n] * vi;
      }
      a1 = wave_sum(a1);
      a2 = wave_sum(a2);
      if (lane == 0) {
        part2[q][wave] = a1;
        part2[NBP + q][wave] = a2;
      }
    }
    __syncthreads();
    if (tid < NBP && tid < cp) {
      double y1 = 0.0, y2 = 0.0;
      for (int w = 0; w < TC / 64; ++w) {
        y1 += part2[tid][w];
        y2 += part2[NBP + tid][w];
      }
      ys[tid] = y1;
      ys[NBP + tid] = y2;
    }
    __syncthreads();
    // p = s - V y1 - W y2 (kept in the W column), K = p.v
    double* wcol = W + (long)cp * n;
    double kp = 0.0;
    for (int i = j + tid; i < n; i += TC) {
      double pi = sv[i];
      for (int q = 0; q < cp; ++q)
        pi -= V[i + (long)q * n] * ys[q] + W[i + (long)q * n] * ys[NBP + q];
      wcol[i] = pi;
      kp += pi * v[i];
    }
    const double K = taup * block_sum_n(kp, red, TC);
    const double alpha = -0.5 * taup * K;
    for (int i = tid; i < n; i += TC)
      wcol[i] = i < j ? 0.0 : taup * wcol[i] + alpha * v[i];
    __syncthreads();
  }
  if (FIN) return;
  if (j >= ts || j > n - 2) {
    // no column here for this problem: an empty slot
    if (c < NBP)
      for (int i = tid; i < n; i += TC) {
        V[i + (long)c * n] = 0.0;
        V2[i + (long)c * n] = 0.0;
        W[i + (long)c * n] = 0.0;
      }
    return;
  }
  double* col = A + (long)j * n;
  // column j with the panel's updates: col -= V W(j, :)^T + W V(j, :)^T
  if (tid < c) {
    cj[tid] = W[j + (long)tid * n];
    cj[NBP + tid] = V[j + (long)tid * n];
  }
  __syncthreads();
  double part = 0.0;
  for (int i = j + tid; i < n; i += TC) {
    double ci = col[i];
    for (int q = 0; q < c; ++q)
      ci -= V[i + (long)q * n] * cj[q] + W[i + (long)q * n] * cj[NBP + q];
    col[i] = ci;
    if (i >= j + 2) part += ci * ci;
  }
  const double sigma = block_sum_n(part, red, TC);   // column j visible
  const double alpha = col[j + 1];
  double tau, beta, scal;
  if (sigma == 0.0) {
    tau = 0.0;
    beta = alpha;
    scal = 0.0;
  } else {
    const double mu = sqrt(alpha * alpha + sigma);
    beta = (alpha <= 0.0) ? mu : -mu;
    tau = (beta - alpha) / beta;
    scal = 1.0 / (alpha - beta);
  }
  double* vcur = (j & 1 ? b.vbuf1 : b.vbuf0) + vo;
  for (int i = tid; i < n; i += TC) {
    const double vi = i <= j ? 0.0 : (i == j + 1 ? 1.0 : col[i] * scal);
    V[i + (long)c * n] = vi;
    V2[i + (long)c * n] = vi;
    if (i > j) vcur[i] = vi;
    if (i > j + 1) col[i] = vi;
  }
  __syncthreads();
  if (tid == 0) {
    b.d[vo + j] = col[j];
    b.e[vo + j] = beta;
    b.tau[vo + j] = tau;
    col[j + 1] = beta;
  }
}

// One launch per column for problems of at most FMAX rows (the
// Rayleigh-Ritz problems of the iterative solver), where every launch is
// latency: each work-group of the column-j launch repeats the column phase
// of k_tri_col itself -- K = tau s.v, w, the pending update of column j, its
// norm and the reflector v_j, all on at most FMAX rows held in LDS -- and then
// runs the trailing phase of k_tri_upd on its own UC columns.  Storage
// differs from the two-launch form: s = A v alternates between pbuf (even j)
// and wbuf (odd j), since the launch reads s_{j-1} while it writes s_j; w is
// never stored; the reflector v_j goes to vbuf[j&1] and reaches the lower
// part of column j one launch later (written by the next launch's first
// work-group, when no work-group reads column j any more; the tail kernel
// does it for the last one).
constexpr int FMAX = 512;

__global__ void __launch_bounds__(TB) k_tri_fused(EigBatch b, int j) {
  // v_{j-1}, s_{j-1} and the updated column j; w_{j-1} and v_j are formed
  // from them where they are needed, so that the two reductions are the
  // launch's only barriers
  __shared__ double vp[FMAX], sp[FMAX], cl[FMAX];
  __shared__ double red[TB / 64];
  const int bid = blockIdx.x;
  int p = 0, base = 0;
  for (; p < b.nprob; ++p) {
    const int n = b.n[p];
    const int cnt = (j < tail_start(n)) ? (n - j - 1 + UC - 1) / UC : 0;
    if (bid < base + cnt) break;
    base += cnt;
  }
  if (p >= b.nprob) return;
  const int n = b.n[p];
  const long vo = b.voff[p];
  double* A = b.A + b.aoff[p];
  const int tid = threadIdx.x;
  const bool first = bid == base;   // the problem's first work-group
  const double* sprev = ((j - 1) & 1 ? b.wbuf : b.pbuf) + vo;
  double* scur = (j & 1 ? b.wbuf : b.pbuf) + vo;
  const double* vprev_g = ((j - 1) & 1 ? b.vbuf1 : b.vbuf0) + vo;
  double* vcur_g = (j & 1 ? b.vbuf1 : b.vbuf0) + vo;
  const int lane = tid & 63, wave = tid >> 6;
  const double* col = A + (long)j * n;
  // every global read of the launch is issued here, so that they share one
  // memory round trip: v_{j-1}, s_{j-1} and column j (rows j.., two per
  // thread) and this wavefront's two trailing columns (rows j+1.., eight per
  // lane)
  constexpr int RQ = FMAX / TB, RL = FMAX / 64;
  double vr[RQ], sr[RQ], cr[RQ];
#pragma unroll
  for (int q = 0; q < RQ; ++q) {
    const int i = j + tid + q * TB;
    const bool in = i < n;
    vr[q] = (in && j > 0) ? vprev_g[i] : 0.0;
    sr[q] = (in && j > 0) ? sprev[i] : 0.0;
    cr[q] = in ? col[i] : 0.0;
  }
  const int k0 = j + 1 + (bid - base) * UC + 2 * wave;
  const bool have = k0 < n, two = k0 + 1 < n;
  double* ca = A + (long)(have ? k0 : 0) * n;
  double* cb = two ? A + (long)(k0 + 1) * n : ca;
  double xa[RL], xb[RL];
#pragma unroll
  for (int r = 0; r < RL; ++r) {
    const int i = j + 1 + lane + 64 * r;
    xa[r] = (have && i < n) ? ca[i] : 0.0;
    xb[r] = (two && i < n) ? cb[i] : 0.0;
  }
  const double taup = j > 0 ? b.tau[vo + j - 1] : 0.0;
  // the previous reflector into the lower part of its column
  if (first && j > 0)
    for (int i = j + 1 + tid; i < n; i += TB) A[i + (long)(j - 1) * n] = vprev_g[i];
  // pending update of step j-1: w = tau (s - (tau/2)(s.v) v)
  double kp = 0.0;
#pragma unroll
  for (int q = 0; q < RQ; ++q) {
    const int i = j + tid + q * TB;
    if (i < n) {
      vp[i] = vr[q];
      sp[i] = sr[q];
      kp += sr[q] * vr[q];
    }
  }
  const double K = taup * block_sum(kp, red);   // vp, sp visible
  const double alpha = -0.5 * taup * K;
  auto wof = [&](int i) { return taup * sp[i] + alpha * vp[i]; };   // w_{j-1}(i)
  // column j with the pending update, and the norm below its subdiagonal
  const double wj = wof(j), vpj = vp[j];
  double c[RQ];
  double part = 0.0;
#pragma unroll
  for (int q = 0; q < RQ; ++q) {
    const int i = j + tid + q * TB;
    c[q] = 0.0;
    if (i < n) {
      c[q] = cr[q] - (vp[i] * wj + wof(i) * vpj);
      cl[i] = c[q];
      if (i >= j + 2) part += c[q] * c[q];
    }
  }
  const double sigma = block_sum(part, red);   // cl visible
  const double alph = cl[j + 1];
  double tj, beta, scal;
  if (sigma == 0.0) {
    tj = 0.0;
    beta = alph;
    scal = 0.0;
  } else {
    const double mu = sqrt(alph * alph + sigma);
    beta = (alph <= 0.0) ? mu : -mu;
    tj = (beta - alph) / beta;
    scal = 1.0 / (alph - beta);
  }
  if (first) {
#pragma unroll
    for (int q = 0; q < RQ; ++q) {
      const int i = j + tid + q * TB;
      if (i > j && i < n) vcur_g[i] = (i == j + 1) ? 1.0 : c[q] * scal;
      if (i == j) {
        b.d[vo + j] = c[q];
        b.e[vo + j] = beta;
        b.tau[vo + j] = tj;
      }
    }
  }
  // trailing phase on this work-group's columns (two per wavefront)
  if (!have) return;
  const double wa = wof(k0), pa = vp[k0];
  const double wb = two ? wof(k0 + 1) : 0.0, pb = two ? vp[k0 + 1] : 0.0;
  double sa = 0.0, sb = 0.0;
#pragma unroll
  for (int r = 0; r < RL; ++r) {
    const int i = j + 1 + lane + 64 * r;
    if (i < n) {
      const double vpi = vp[i], wi = wof(i);
      const double vji = (i == j + 1) ? 1.0 : cl[i] * scal;
      const double ua = xa[r] - (vpi * wa + wi * pa);
      const double ub = xb[r] - (vpi * wb + wi * pb);
      ca[i] = ua;
      if (two) cb[i] = ub;
      sa += ua * vji;
      sb += ub * vji;
    }
  }
  sa = wave_sum(sa);
  sb = wave_sum(sb);
  if (lane == 0) {
    scur[k0] = sa;
    if (two) scur[k0 + 1] = sb;
  }
}

// Columns tail_start(n) .. n-2 of problem blockIdx.x: the trailing block
// A(j0:n, j0:n) (both triangles) is copied to LDS together with the state the
// column launches leave behind -- the reflector v_{j0-1}, s_{j0-1} = A v_{j0-1}
// and tau_{j0-1}, the rank-2 update of step j0-1 still pending -- and the
// column and trailing phases of every remaining step run on it exactly as in
// k_tri_col / k_tri_upd.  The lower triangle (reflectors) is written back.
// work-group size of the tail kernel: DVR_TAIL_THREADS (256, 512, 1024).
// On the 448-row Rayleigh-Ritz problems of c3 one tail took 0.55 / 0.42 /
// 0.48 ms; 512 is the default.
int tail_threads() {
  const char* v = getenv("DVR_TAIL_THREADS");
  const int t = v ? atoi(v) : 512;
  return t == 256 || t == 1024 ? t : 512;
}
constexpr int LDB = TAIL + 1;
constexpr size_t TAIL_LDS = sizeof(double) * ((size_t)TAIL * LDB + 4 * TAIL);

template <int TT>
__global__ void __launch_bounds__(TT) k_tri_tail(EigBatch b, int fused) {
  extern __shared__ double sh[];
  double* B = sh;
  double* vp = B + TAIL * LDB;   // v_{j-1}
  double* vc = vp + TAIL;        // v_j (the two swap roles every step)
  double* w = vc + TAIL;         // w_{j-1}
  double* ss = w + TAIL;         // s_{j-1}, then s_j
  __shared__ double red[TT / 64];
  const int p = blockIdx.x, tid = threadIdx.x;
  const int n = b.n[p];
  const long vo = b.voff[p];
  double* A = b.A + b.aoff[p];
  if (n == 1) {
    if (tid == 0) b.d[vo] = A[0];
    return;
  }
  const int j0 = tail_start(n), t = n - j0;
  for (int e = tid; e < t * t; e += TT) {
    const int i = e % t, k = e / t;
    B[i + k * LDB] = A[(j0 + i) + (long)(j0 + k) * n];
  }
  double taup = 0.0;
  if (j0 > 0 && fused == 2) {
    // blocked reduction: every update up to column j0 already applied
    for (int i = tid; i < t; i += TT) {
      vp[i] = 0.0;
      ss[i] = 0.0;
    }
  } else if (j0 > 0) {
    const double* vprev = ((j0 - 1) & 1 ? b.vbuf1 : b.vbuf0) + vo;
    // s_{j0-1}: pbuf, or (one-launch columns) the buffer of that parity
    const double* sprev = (fused && ((j0 - 1) & 1) ? b.wbuf : b.pbuf) + vo;
    for (int i = tid; i < t; i += TT) {
      vp[i] = vprev[j0 + i];
      ss[i] = sprev[j0 + i];
    }
    // the one-launch columns leave the last reflector to be stored
    if (fused)
      for (int i = j0 + 1 + tid; i < n; i += TT) A[i + (long)(j0 - 1) * n] = vprev[i];
    taup = b.tau[vo + j0 - 1];
  }
  __syncthreads();
  for (int j = j0; j <= n - 2; ++j) {
    const int jl = j - j0;
    double part = 0.0;
    if (j > 0) {
      double kp = 0.0;
      for (int i = jl + tid; i < t; i += TT) kp += ss[i] * vp[i];
      const double K = taup * block_sum_n(kp, red, TT);
      const double alpha = -0.5 * taup * K;
      const double vj = vp[jl];
      const double wj = taup * ss[jl] + alpha * vj;
      for (int i = jl + tid; i < t; i += TT) {
        const double wi = taup * ss[i] + alpha * vp[i];
        w[i] = wi;
        const double ci = B[i + jl * LDB] - (vp[i] * wj + wi * vj);
        B[i + jl * LDB] = ci;
        if (i >= jl + 2) part += ci * ci;
      }
    } else {
      for (int i = jl + 2 + tid; i < t; i += TT) part += B[i + jl * LDB] * B[i + jl * LDB];
    }
    const double sigma = block_sum_n(part, red, TT);   // w and column jl now visible
    if (j == n - 2) {
      if (tid == 0) {
        b.d[vo + j] = B[jl + jl * LDB];
        b.e[vo + j] = B[(t - 1) + jl * LDB];
        b.tau[vo + j] = 0.0;
        double dl = B[(t - 1) + (t - 1) * LDB];
        if (j > 0) dl -= (vp[t - 1] * w[t - 1] + w[t - 1] * vp[t - 1]);
        b.d[vo + n - 1] = dl;
      }
      break;
    }
    const double alpha = B[(jl + 1) + jl * LDB];
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
    for (int i = jl + 1 + tid; i < t; i += TT) {
      const double vi = (i == jl + 1) ? 1.0 : B[i + jl * LDB] * scal;
      vc[i] = vi;
      if (i > jl + 1) B[i + jl * LDB] = vi;
    }
    __syncthreads();
    if (tid == 0) {
      b.d[vo + j] = B[jl + jl * LDB];
      b.e[vo + j] = beta;
      b.tau[vo + j] = tj;
      B[(jl + 1) + jl * LDB] = beta;
    }
    // trailing phase: pending update of step j-1 on columns jl+1.. (rows
    // jl+1..) and s_j = A v_j, eight threads per column
    {
      const int grp = tid >> 3, sub = tid & 7;
      for (int k = jl + 1 + grp; k < t; k += TT / 8) {
        const double wk = j > 0 ? w[k] : 0.0, pk = j > 0 ? vp[k] : 0.0;
        double acc = 0.0;
        for (int i = jl + 1 + sub; i < t; i += 8) {
          double u = B[i + k * LDB];
          if (j > 0) {
            u -= vp[i] * wk + w[i] * pk;
            B[i + k * LDB] = u;
          }
          acc += u * vc[i];
        }
        for (int off = 4; off > 0; off >>= 1) acc += __shfl_down(acc, off, 8);
        if (sub == 0) ss[k] = acc;
      }
    }
    __syncthreads();
    double* const vt = vp;
    vp = vc;
    vc = vt;
    taup = tj;
  }
  __syncthreads();
  for (int e = tid; e < t * t; e += TT) {
    const int i = e % t, k = e / t;
    if (i >= k) A[(j0 + i) + (long)(j0 + k) * n] = B[i + k * LDB];
  }
}

void launch_tridiag(hipStream_t s, const EigBatch& b, const int* h_n, const long* h_aoff,
                    const long* h_voff) {
  int jend = 0, nmax = 0;
  for (int p = 0; p < b.nprob; ++p) {
    jend = std::max(jend, tail_start(h_n[p]));
    nmax = std::max(nmax, h_n[p]);
  }
  if (nmax <= 0) return;
  const bool fused = nmax <= FMAX && tri_fused_enabled();
  for (int j = 0; j < jend && fused; ++j) {
    double bytes = 0.0, flops = 0.0;
    int grid = 0;
    for (int p = 0; p < b.nprob; ++p) {
      if (j >= tail_start(h_n[p])) continue;
      const double t = (double)(h_n[p] - j - 1);
      grid += (h_n[p] - j - 1 + UC - 1) / UC;
      bytes += (j > 0 ? 16.0 : 8.0) * t * t + 8.0 * 6.0 * t;
      flops += (j > 0 ? 6.0 : 2.0) * t * t;
    }
    DVR_LAUNCH("tri_fused", bytes, flops, k_tri_fused, dim3(grid), dim3(TB), 0, s, b, j);
  }
  const bool blocked = !fused && tri_blocked_enabled(h_n, b.nprob);
  for (int j0 = 0; j0 < jend && blocked; j0 += NBP) {
    for (int j = j0; j < j0 + NBP; ++j) {
      double cbytes = 0.0, ubytes = 0.0, uflops = 0.0;
      int grid = 0;
      for (int p = 0; p < b.nprob; ++p) {
        const double m = (double)(h_n[p] - j);
        cbytes += 8.0 * 6.0 * NBP * std::max(m, 0.0);
        if (j >= tail_start(h_n[p]) || j > h_n[p] - 3) continue;
        grid += (h_n[p] - j - 1 + UC - 1) / UC;
        const double t = m - 1.0;
        ubytes += 8.0 * t * t + 8.0 * 2.0 * t;
        uflops += 2.0 * t * t;
      }
      DVR_LAUNCH("tri_pcol", cbytes, 0.0, k_tri_pcol<false>, dim3(b.nprob), dim3(TC), 0, s, b, j,
                 j0);
      if (grid > 0)
        DVR_LAUNCH("tri_symv", ubytes, uflops, k_tri_upd, dim3(grid), dim3(TB), 0, s, b, j, 0);
    }
    DVR_LAUNCH("tri_pcol", 0.0, 0.0, k_tri_pcol<true>, dim3(b.nprob), dim3(TC), 0, s, b, j0 + NBP,
               j0);
    // the panel's rank-2 NBP update of every problem's trailing square
    for (int p = 0; p < b.nprob; ++p) {
      const int np = h_n[p], ts = tail_start(np);
      if (j0 >= ts) continue;
      const int js = std::min(j0 + NBP, ts), m = np - js;
      if (m <= 0) continue;
      double* V = b.pan + 3L * NBP * h_voff[p];
      launch_gemm(s, "tri_pupd", false, true, m, m, 2 * NBP, -1.0, V + js, np,
                  V + (long)NBP * np + js, np, 1.0, b.A + h_aoff[p] + js + (long)js * np, np,
                  1, nullptr, false);
    }
  }
  for (int j = 0; j < jend && !fused && !blocked; ++j) {
    double cbytes = 0.0, ubytes = 0.0, uflops = 0.0;
    int grid = 0;
    for (int p = 0; p < b.nprob; ++p) {
      if (j >= tail_start(h_n[p])) continue;
      const double m = (double)(h_n[p] - j);
      cbytes += 8.0 * 6.0 * m;
      grid += (h_n[p] - j - 1 + UC - 1) / UC;
      const double t = m - 1.0;
      // trailing square read (+ written back when an update is pending)
      ubytes += (j > 0 ? 16.0 : 8.0) * t * t + 8.0 * 4.0 * t;
      uflops += (j > 0 ? 6.0 : 2.0) * t * t;
    }
    DVR_LAUNCH("tri_col", cbytes, 0.0, k_tri_col, dim3(b.nprob), dim3(TC), 0, s, b, j);
    if (grid > 0) {
      DVR_LAUNCH("tri_upd", ubytes, uflops, k_tri_upd, dim3(grid), dim3(TB), 0, s, b, j, 1);
    }
  }
  for (const void* f : {(const void*)k_tri_tail<256>, (const void*)k_tri_tail<512>,
                        (const void*)k_tri_tail<1024>})
    DVR_HIP(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)TAIL_LDS));
  double tbytes = 0.0;
  for (int p = 0; p < b.nprob; ++p) {
    const double t = (double)(h_n[p] - tail_start(h_n[p]));
    tbytes += 8.0 * 2.0 * t * t;
  }
  const int tt = tail_threads();
  const int mode = fused ? 1 : (blocked ? 2 : 0);
  if (tt == 256)
    DVR_LAUNCH("tri_tail", tbytes, 0.0, k_tri_tail<256>, dim3(b.nprob), dim3(256), TAIL_LDS, s, b,
               mode);
  else if (tt == 512)
    DVR_LAUNCH("tri_tail", tbytes, 0.0, k_tri_tail<512>, dim3(b.nprob), dim3(512), TAIL_LDS, s, b,
               mode);
  else
    DVR_LAUNCH("tri_tail", tbytes, 0.0, k_tri_tail<1024>, dim3(b.nprob), dim3(1024), TAIL_LDS, s,
               b, mode);
  DVR_LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------
// 2. bounds and multisection
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(TB) k_tri_bounds(EigBatch b) {
  __shared__ double rmin[TB], rmax[TB];
  const int p = blockIdx.x;
  const int n = b.n[p];
  const long vo = b.voff[p];
  const double* d = b.d + vo;
  const double* e = b.e + vo;
  double* e2 = b.e2 + vo;
  double lo = DBL_MAX, hi = -DBL_MAX, m2 = 0.0;
  for (int i = threadIdx.x; i < n; i += TB) {
    const double em = i > 0 ? fabs(e[i - 1]) : 0.0;
    const double ep = i < n - 1 ? fabs(e[i]) : 0.0;
    lo = fmin(lo, d[i] - em - ep);
    hi = fmax(hi, d[i] + em + ep);
    if (i < n - 1) {
      const double q = e[i] * e[i];
      e2[i] = q;
      m2 = fmax(m2, q);
    }
  }
  rmin[threadIdx.x] = lo;
  rmax[threadIdx.x] = hi;
  __syncthreads();
  for (int st = TB / 2; st > 0; st >>= 1) {
    if (threadIdx.x < st) {
      rmin[threadIdx.x] = fmin(rmin[threadIdx.x], rmin[threadIdx.x + st]);
      rmax[threadIdx.x] = fmax(rmax[threadIdx.x], rmax[threadIdx.x + st]);
    }
    __syncthreads();
  }
  const double glo = rmin[0], ghi = rmax[0];
  __syncthreads();
  rmax[threadIdx.x] = m2;
  __syncthreads();
  for (int st = TB / 2; st > 0; st >>= 1) {
    if (threadIdx.x < st) rmax[threadIdx.x] = fmax(rmax[threadIdx.x], rmax[threadIdx.x + st]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double tn = fmax(fabs(glo), fabs(ghi));
    const double piv = DBL_MIN * fmax(1.0, rmax[0]);
    const double pad = 2.1 * kEps * tn * n + 4.0 * piv;
    b.gl[p] = glo - pad;
    b.gu[p] = ghi + pad;
    b.tnrm[p] = tn;
    b.pivmin[p] = piv;
  }
}

void launch_tridiag_bounds(hipStream_t s, const EigBatch& b) {
  hipLaunchKernelGGL(k_tri_bounds, dim3(b.nprob), dim3(TB), 0, s, b);
  DVR_LAUNCH_CHECK();
}

// Sturm count: the number of negative pivots of T - x I = L D L^T.  The
// quotient e2/q uses rcp_refined (the recurrence's critical path);
// |q| >= pivmin keeps it away from denormals.  d and e2 are read eight at a
// time.

__device__ inline int sturm_count(const double* __restrict__ d,
                                  const double* __restrict__ e2, int n,
                                  double x, double pivmin) {
  double q = d[0] - x;
  if (fabs(q) < pivmin) q = -pivmin;
  int c = q <= 0.0;
  int i = 1;
  for (; i + 7 < n; i += 8) {
    double dv[8], ev[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      dv[u] = d[i + u];
      ev[u] = e2[i + u - 1];
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      q = (dv[u] - x) - ev[u] * rcp_refined(q);
      if (fabs(q) < pivmin) q = -pivmin;
      c += q <= 0.0;
    }
  }
  for (; i < n; ++i) {
    q = (d[i] - x) - e2[i - 1] * rcp_refined(q);
    if (fabs(q) < pivmin) q = -pivmin;
    c += q <= 0.0;
  }
  return c;
}

__global__ void __launch_bounds__(TB)
k_bisect(EigBatch b, const int* __restrict__ il, const int* __restrict__ kofs,
         int total, double* __restrict__ eval, const long* __restrict__ eoff) {
  const int lane = threadIdx.x & 63;
  const int gid = __builtin_amdgcn_readfirstlane(blockIdx.x * (TB / 64) + (threadIdx.x >> 6));
  if (gid >= total) return;
  int lo_p = 0, hi_p = b.nprob - 1;
  while (lo_p < hi_p) {  // largest p with kofs[p] <= gid
    const int mid = (lo_p + hi_p + 1) >> 1;
    if (kofs[mid] <= gid) lo_p = mid; else hi_p = mid - 1;
  }
  const int p = __builtin_amdgcn_readfirstlane(lo_p);
  const int kloc = gid - kofs[p];
  const int k = il[p] + kloc;
  const int n = b.n[p];
  const long vo = b.voff[p];
  const double* d = b.d + vo;
  const double* e2 = b.e2 + vo;
  const double piv = b.pivmin[p];
  double lo = b.gl[p], hi = b.gu[p];
  for (int it = 0; it < 48; ++it) {
    const double width = hi - lo;
    if (width <= 2.0 * kEps * fmax(fabs(lo), fabs(hi)) || width <= 4.0 * piv) break;
    const double step = width / 65.0;
    const double x = lo + (double)(lane + 1) * step;
    const int c = sturm_count(d, e2, n, x, piv);
    const unsigned long long mask = __ballot(c <= k);
    double nlo, nhi;
    if (mask == 0ull) {
      nlo = lo;
      nhi = lo + 1.0 * step;
    } else {
      const int L = 63 - __clzll((long long)mask);
      nlo = lo + (double)(L + 1) * step;
      nhi = (L == 63) ? hi : lo + (double)(L + 2) * step;
    }
    if (nlo == lo && nhi == hi) break;
    lo = nlo;
    hi = nhi;
  }
  if (lane == 0) eval[eoff[p] + kloc] = 0.5 * (lo + hi);
}

void launch_bisect(hipStream_t s, const EigBatch& b, const int* il,
                   const int* kofs, int total, double* eval,
                   const long* eoff) {
  if (total <= 0) return;
  const int grid = (total + TB / 64 - 1) / (TB / 64);
  DVR_LAUNCH("bisect", 0.0, 0.0, k_bisect, dim3(grid), dim3(TB), 0, s, b, il, kofs, total,
                   eval, eoff);
  DVR_LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------
// 3. inverse iteration
// ---------------------------------------------------------------------------

size_t stein_scratch_doubles(int nmax, int nthreads) {
  return (size_t)6 * (size_t)nmax * (size_t)nthreads;
}

__global__ void __launch_bounds__(TB)
k_stein(EigBatch b, const Cluster* __restrict__ cl, int ncl,
        const double* __restrict__ eval, const long* __restrict__ eoff,
        double* __restrict__ Z, const long* __restrict__ zoff,
        double* __restrict__ scratch, int T) {
  const int t = blockIdx.x * TB + threadIdx.x;
  if (t >= ncl) return;
  const Cluster C = cl[t];
  const int p = C.p;
  const long vo = b.voff[p];
  const long nm = b.nmax;
  inverse_iteration_cluster<8>(b.n[p], b.d + vo, b.e + vo, eval + eoff[p], C.kb, C.ke, b.tnrm[p],
                            scratch + t, nm * T, T, Z + zoff[p], (unsigned long long)p);
}

// The same inverse iteration with the per-thread scratch in LDS, for
// problems small enough that several threads' scratch fits in one CU's LDS
// (the Rayleigh-Ritz problems of the iterative solver): with few clusters
// the global-memory version above runs a handful of wavefronts whose every
// burst of loads waits on memory.  One wavefront per work-group; its first TW
// lanes take one cluster each, their arrays interleaved so that a step
// touches TW consecutive doubles.  When all of a work-group's clusters belong
// to one problem its tridiagonal is staged in LDS as well.
constexpr int LDS_BYTES_MAX = 160 * 1024;

// DVR_TRI=twolaunch keeps two launches per column for small problems too
bool tri_fused_enabled() {
  const char* v = getenv("DVR_TRI");
  return !(v && std::string(v) == "twolaunch");
}

// The blocked (panel) reduction for batches of at least 8 problems over 1024
// rows with one over 2048: c2's 2D batch (30 problems up to 3629 rows) went
// 1090 -> 695 ms per block; where the batch is small the one-work-group
// column kernel of each panel step is the critical path -- one 4000-row
// problem with 300 vectors (c4's dense solves) 165 -> 277 ms, c3's 2D batch
// (64 problems up to 1116 rows, short per-problem panel GEMMs) 41 -> 61 ms.
// DVR_TRI=blocked takes it for every batch with a problem over FMAX rows,
// DVR_TRI=column (or twolaunch) never.
bool tri_blocked_enabled(const int* hn, int np) {
  const char* v = getenv("DVR_TRI");
  const std::string t = v ? v : "";
  if (t == "column" || t == "twolaunch") return false;
  int nmax = 0, big = 0;
  for (int p = 0; p < np; ++p) {
    nmax = std::max(nmax, hn[p]);
    big += hn[p] > 1024;
  }
  if (t == "blocked") return nmax > FMAX;
  return nmax > 2048 && big >= 8;
}

// DVR_BT=global keeps the back-transformation on the global-memory kernels
bool bt_use_lds() {
  const char* v = getenv("DVR_BT");
  return !(v && std::string(v) == "global");
}

// DVR_STEIN=global keeps every problem on the global-memory version
bool stein_use_lds() {
  const char* v = getenv("DVR_STEIN");
  return !(v && std::string(v) == "global");
}

int stein_lds_lanes(int nmax) {
  if (nmax <= 0) return 0;
  const long avail = LDS_BYTES_MAX / 8 - 2L * nmax;
  return (int)std::min<long>(64, std::max<long>(0, avail / (6L * nmax)));
}

__global__ void __launch_bounds__(64)
k_stein_lds(EigBatch b, const Cluster* __restrict__ cl, int ncl,
            const double* __restrict__ eval, const long* __restrict__ eoff,
            double* __restrict__ Z, const long* __restrict__ zoff, int TW) {
  extern __shared__ double sh[];
  const int nm = b.nmax;
  const int c0 = blockIdx.x * TW;
  const int cnt = min(TW, ncl - c0);
  const int lane = threadIdx.x;
  double* dl = sh + (long)6 * nm * TW;
  double* el = dl + nm;
  const int p0 = cl[c0].p;
  const bool staged = cl[c0 + cnt - 1].p == p0;
  if (staged) {
    const long vo = b.voff[p0];
    for (int i = lane; i < b.n[p0]; i += 64) {
      dl[i] = b.d[vo + i];
      el[i] = b.e[vo + i];
    }
  }
  __syncthreads();
  if (lane >= cnt) return;
  const Cluster C = cl[c0 + lane];
  const int p = C.p;
  const long vo = b.voff[p];
  inverse_iteration_cluster<8>(b.n[p], staged ? dl : b.d + vo, staged ? el : b.e + vo,
                                eval + eoff[p], C.kb, C.ke, b.tnrm[p], sh + lane,
                                (long)nm * TW, TW, Z + zoff[p], (unsigned long long)p);
}

// Inverse iteration with one wavefront per cluster, for problems whose
// scratch does not fit in LDS (large n): the same arithmetic as
// inverse_iteration_cluster, with the recurrences (factorisation and the two
// sweeps of each of the three iterations) on lane 0, but their operands
// moving through LDS in chunks of SW elements that the whole wavefront
// loads and stores with coalesced accesses, the next chunk's loads in flight
// while lane 0 works on the current one.  Scaling, Gram-Schmidt against the
// cluster's earlier vectors, norms and the final store are wavefront-
// parallel.  With one thread per cluster (k_stein) a c3-size problem (20000
// rows, 1000 vectors) runs 16 wavefronts whose every burst waits on memory.
// Per cluster the scratch arrays FL, FD, FU, FU2, X, PV (nmax doubles each)
// are contiguous.
constexpr int SW = 256;
constexpr int SWL = SW / 64;

__device__ inline double wave_bcast_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return __shfl(v, 0, 64);
}

__global__ void __launch_bounds__(64)
k_stein_wave(EigBatch b, const Cluster* __restrict__ cl, int ncl,
             const double* __restrict__ eval, const long* __restrict__ eoff,
             double* __restrict__ Z, const long* __restrict__ zoff,
             double* __restrict__ scratch) {
  __shared__ double in[4][SW + 1];
  __shared__ double out[5][SW];
  __shared__ double bc[2];
  const int ci = blockIdx.x;
  if (ci >= ncl) return;
  const int lane = threadIdx.x;
  const Cluster C = cl[ci];
  const int p = C.p;
  const int n = b.n[p];
  const long vo = b.voff[p];
  const double* __restrict__ d = b.d + vo;
  const double* __restrict__ e = b.e + vo;
  const double* __restrict__ lams = eval + eoff[p];
  double* __restrict__ Zp = Z + zoff[p];
  const long nm = b.nmax;
  double* FL = scratch + (long)ci * 6 * nm;
  double* FD = FL + nm;
  double* FU = FD + nm;
  double* FU2 = FU + nm;
  double* X = FU2 + nm;
  double* PV = X + nm;
  const double tnn = b.tnrm[p] > 0.0 ? b.tnrm[p] : 1.0;
  const double pertol = 10.0 * kEps * tnn, tiny = kEps * tnn;
  const int nm1 = n - 1;
  const unsigned long long seed = (unsigned long long)p;
  // chunk [lo, lo + cnt) of a into registers (lane-strided), zero past cnt
  auto ldc = [&](const double* a, long lo, int cnt, double(&r)[SWL]) {
#pragma unroll
    for (int t = 0; t < SWL; ++t) {
      const int q = lane + 64 * t;
      r[t] = q < cnt ? a[lo + q] : 0.0;
    }
  };
  auto stl = [&](double* sdst, const double(&r)[SWL]) {
#pragma unroll
    for (int t = 0; t < SWL; ++t) sdst[lane + 64 * t] = r[t];
  };
  double lprev = 0.0;
  for (int k = C.kb; k < C.ke; ++k) {
    double lam = lams[k];
    if (k > C.kb && lam - lprev < pertol) lam = lprev + pertol;
    lprev = lam;
    // ---- factorisation T - lam I = P L U (the dgttrf recurrence)
    double dc = d[0] - lam, uc = n > 1 ? e[0] : 0.0;
    {
      double re[SWL], rd[SWL], rx = 0.0;
      auto pre = [&](int j0) {
        const int cnt = min(SW, nm1 - j0);
        ldc(e, j0, cnt, re);
        ldc(d + 1, j0, cnt, rd);
        rx = (lane == 0 && j0 + SW < nm1) ? e[j0 + SW] : 0.0;
      };
      if (nm1 > 0) pre(0);
      for (int i0 = 0; i0 < nm1; i0 += SW) {
        const int cnt = min(SW, nm1 - i0);
        __syncthreads();
        stl(in[0], re);
        stl(in[1], rd);
        if (lane == 0) in[0][SW] = rx;
        __syncthreads();
        if (i0 + SW < nm1) pre(i0 + SW);
        if (lane == 0) {
          #pragma unroll 4
          for (int q = 0; q < cnt; ++q) {
            const int i = i0 + q;
            const double li = in[0][q];
            const double dn = in[1][q] - lam;
            const double un = (i < n - 2) ? in[0][q + 1] : 0.0;
            const bool sw = fabs(dc) < fabs(li);
            const double den = sw ? li : dc, num = sw ? dc : li;
            const double rden = rcp_refined(den == 0.0 ? tiny : den);
            const double fl = num * rden;
            out[0][q] = fl;
            out[1][q] = rden;
            out[2][q] = sw ? dn : uc;
            out[3][q] = sw ? un : 0.0;
            out[4][q] = sw ? 1.0 : 0.0;
            dc = sw ? uc - fl * dn : dn - fl * uc;
            uc = sw ? -fl * un : un;
          }
        }
        __syncthreads();
#pragma unroll
        for (int t = 0; t < SWL; ++t) {
          const int q = lane + 64 * t;
          if (q < cnt) {
            FL[i0 + q] = out[0][q];
            FD[i0 + q] = out[1][q];
            FU[i0 + q] = out[2][q];
            FU2[i0 + q] = out[3][q];
            PV[i0 + q] = out[4][q];
          }
        }
      }
      if (lane == 0) FD[n - 1] = 1.0 / (dc == 0.0 ? tiny : dc);
    }
    __threadfence();
    __syncthreads();
    double scale = 1.0;
    for (int it = 0; it < 3; ++it) {
      const unsigned long long base = (seed << 40) ^ ((unsigned long long)k << 20);
      // ---- forward sweep: L^-1 with the row interchanges
      double xc = it == 0 ? ii_hash(base) : X[0] * scale;
      {
        double rl[SWL], rp[SWL], rn[SWL];
        auto pre = [&](int j0) {
          const int cnt = min(SW, nm1 - j0);
          ldc(FL, j0, cnt, rl);
          ldc(PV, j0, cnt, rp);
          if (it == 0) {
#pragma unroll
            for (int t = 0; t < SWL; ++t)
              rn[t] = ii_hash(base ^ (unsigned long long)(j0 + lane + 64 * t + 1));
          } else {
            ldc(X + 1, j0, cnt, rn);
#pragma unroll
            for (int t = 0; t < SWL; ++t) rn[t] *= scale;
          }
        };
        if (nm1 > 0) pre(0);
        for (int i0 = 0; i0 < nm1; i0 += SW) {
          const int cnt = min(SW, nm1 - i0);
          __syncthreads();
          stl(in[0], rl);
          stl(in[1], rp);
          stl(in[2], rn);
          __syncthreads();
          if (i0 + SW < nm1) pre(i0 + SW);
          if (lane == 0) {
            #pragma unroll 4
            for (int q = 0; q < cnt; ++q) {
              const bool sflag = in[1][q] != 0.0;
              const double xn = in[2][q];
              const double a = sflag ? xn : xc, bq = sflag ? xc : xn;
              xc = bq - in[0][q] * a;
              out[0][q] = a;
            }
          }
          __syncthreads();
#pragma unroll
          for (int t = 0; t < SWL; ++t) {
            const int q = lane + 64 * t;
            if (q < cnt) X[i0 + q] = out[0][q];
          }
        }
      }
      __threadfence();
      __syncthreads();
      // ---- backward sweep: U^-1, tracking max |x| (lane 0)
      double x1 = 0.0, x0 = 0.0, mxv = 0.0;
      if (lane == 0) {
        x1 = xc * FD[n - 1];
        X[n - 1] = x1;
        mxv = fabs(x1);
        if (n > 1) {
          x0 = (X[n - 2] - FU[n - 2] * x1) * FD[n - 2];
          X[n - 2] = x0;
          mxv = fmax(mxv, fabs(x0));
        }
      }
      {
        double rx[SWL], ru[SWL], ru2[SWL], rf[SWL];
        auto pre = [&](int hi) {
          const int lo = max(0, hi - SW + 1), cnt = hi - lo + 1;
          ldc(X, lo, cnt, rx);
          ldc(FU, lo, cnt, ru);
          ldc(FU2, lo, cnt, ru2);
          ldc(FD, lo, cnt, rf);
        };
        if (n - 3 >= 0) pre(n - 3);
        for (int i1 = n - 3; i1 >= 0; i1 -= SW) {
          const int lo = max(0, i1 - SW + 1), cnt = i1 - lo + 1;
          __syncthreads();
          stl(in[0], rx);
          stl(in[1], ru);
          stl(in[2], ru2);
          stl(in[3], rf);
          __syncthreads();
          if (i1 - SW >= 0) pre(i1 - SW);
          if (lane == 0) {
            #pragma unroll 4
            for (int q = cnt - 1; q >= 0; --q) {
              const double xi = (in[0][q] - in[1][q] * x0 - in[2][q] * x1) * in[3][q];
              mxv = fmax(mxv, fabs(xi));
              x1 = x0;
              x0 = xi;
              out[0][q] = xi;
            }
          }
          __syncthreads();
#pragma unroll
          for (int t = 0; t < SWL; ++t) {
            const int q = lane + 64 * t;
            if (q < cnt) X[lo + q] = out[0][q];
          }
        }
      }
      if (lane == 0) bc[0] = mxv;
      __threadfence();
      __syncthreads();
      mxv = bc[0];
      const double sc = mxv > 0.0 ? 1.0 / mxv : 1.0;
      if (k > C.kb) {
        for (int i = lane; i < n; i += 64) X[i] *= sc;
        for (int q = C.kb; q < k; ++q) {
          const double* zq = Zp + (long)q * n;
          double dot = 0.0;
          for (int i = lane; i < n; i += 64) dot += X[i] * zq[i];
          dot = wave_bcast_sum(dot);
          for (int i = lane; i < n; i += 64) X[i] -= dot * zq[i];
        }
        double nrm = 0.0;
        for (int i = lane; i < n; i += 64) nrm += X[i] * X[i];
        nrm = wave_bcast_sum(nrm);
        scale = nrm > 0.0 ? 1.0 / sqrt(nrm) : 0.0;
      } else {
        double nrm = 0.0;
        for (int i = lane; i < n; i += 64) {
          const double t = X[i] * sc;
          nrm += t * t;
        }
        nrm = wave_bcast_sum(nrm);
        scale = nrm > 0.0 ? sc / sqrt(nrm) : 0.0;
      }
      __threadfence();
      __syncthreads();
    }
    double* zk = Zp + (long)k * n;
    for (int i = lane; i < n; i += 64) zk[i] = X[i] * scale;
    __threadfence();
    __syncthreads();
  }
}

// One wavefront per cluster where the clusters are too few for the
// one-thread-per-cluster kernel to fill the GPU (at most 4096, i.e. 64 of its
// wavefronts: c3's 3D solve, 980 clusters of 20000 rows, 234 -> 30 ms); a
// batch with more (c3's 2D stage, ~19000 clusters of <= 1116 rows) keeps
// one thread each (16.0 ms there against 18.8 ms).  DVR_STEIN=thread (or
// global) never takes the wavefront kernel, DVR_STEIN=wave always does.
bool stein_use_wave(int ncl) {
  const char* v = getenv("DVR_STEIN");
  const std::string t = v ? v : "";
  if (t == "thread" || t == "global") return false;
  return t == "wave" || ncl <= 4096;
}

void launch_stein(hipStream_t s, const EigBatch& b, const Cluster* cl,
                  int ncl, const double* eval, const long* eoff, double* Z,
                  const long* zoff, double* scratch, int nthreads) {
  if (ncl <= 0) return;
  const int tw = stein_lds_lanes(b.nmax);
  if (tw >= 4 && stein_use_lds()) {
    const size_t lds = sizeof(double) * ((size_t)6 * b.nmax * tw + 2 * (size_t)b.nmax);
    DVR_HIP(hipFuncSetAttribute((const void*)k_stein_lds,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    DVR_LAUNCH("stein", 0.0, 0.0, k_stein_lds, dim3((ncl + tw - 1) / tw), dim3(64), lds, s, b, cl,
               ncl, eval, eoff, Z, zoff, tw);
    DVR_LAUNCH_CHECK();
    return;
  }
  if (stein_use_wave(ncl) && nthreads >= ncl) {
    DVR_LAUNCH("stein", 0.0, 0.0, k_stein_wave, dim3(ncl), dim3(64), 0, s, b, cl, ncl, eval, eoff,
               Z, zoff, scratch);
    DVR_LAUNCH_CHECK();
    return;
  }
  const int grid = (ncl + TB - 1) / TB;
  DVR_LAUNCH("stein", 0.0, 0.0, k_stein, dim3(grid), dim3(TB), 0, s, b, cl, ncl, eval,
                   eoff, Z, zoff, scratch, nthreads);
  DVR_LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------
// 4. back-transformation
// ---------------------------------------------------------------------------
// Reflector r of block blk: index j = blk*NB + r, vector v(i) for i > j with
// v(j+1) = 1 and v(i) = A(i, j) below.  T is the upper-triangular factor of
// the forward column-wise WY form (H_j0 ... H_j1 = I - V T V^T).
__device__ inline double refl(const double* A, int n, int j, int i) {
  if (i <= j) return 0.0;
  if (i == j + 1) return 1.0;
  return A[i + (long)j * n];
}

__global__ void __launch_bounds__(TB)
k_larft(EigBatch b, const int* __restrict__ nblk_ofs, double* __restrict__ tbuf,
        const long* __restrict__ toff) {
  __shared__ double G[NB * NB];
  __shared__ double Ts[NB * NB];
  const int bid = blockIdx.x;
  int lo_p = 0, hi_p = b.nprob - 1;
  while (lo_p < hi_p) {
    const int mid = (lo_p + hi_p + 1) >> 1;
    if (nblk_ofs[mid] <= bid) lo_p = mid; else hi_p = mid - 1;
  }
  const int p = lo_p;
  const int blk = bid - nblk_ofs[p];
  const int n = b.n[p];
  const int nr = n - 2;
  const int j0 = blk * NB;
  if (nr <= 0 || j0 >= nr) return;
  const int nbk = min(NB, nr - j0);
  const double* A = b.A + b.aoff[p];
  const double* tau = b.tau + b.voff[p];
  for (int idx = threadIdx.x; idx < NB * NB; idx += TB) {
    G[idx] = 0.0;
    Ts[idx] = 0.0;
  }
  __syncthreads();
  // G(r,c) = v_r . v_c for r < c
  for (int idx = threadIdx.x; idx < nbk * nbk; idx += TB) {
    const int r = idx % nbk, c = idx / nbk;
    if (r >= c) continue;
    const int jr = j0 + r, jc = j0 + c;
    double s = refl(A, n, jr, jc + 1);
    for (int i = jc + 2; i < n; ++i) s += A[i + (long)jr * n] * A[i + (long)jc * n];
    G[r + NB * c] = s;
  }
  __syncthreads();
  for (int c = 0; c < nbk; ++c) {
    const double tc = tau[j0 + c];
    if (threadIdx.x < c) {
      const int r = threadIdx.x;
      double s = 0.0;
      for (int l = r; l < c; ++l) s += Ts[r + NB * l] * G[l + NB * c];
      Ts[r + NB * c] = -tc * s;
    }
    if (threadIdx.x == 0) Ts[c + NB * c] = tc;
    __syncthreads();
  }
  double* Tout = tbuf + toff[p] + (long)blk * NB * NB;
  for (int idx = threadIdx.x; idx < NB * NB; idx += TB) Tout[idx] = Ts[idx];
}

__global__ void __launch_bounds__(TB)
k_ormtr(EigBatch b, double* __restrict__ Z, const long* __restrict__ zoff,
        const int* __restrict__ m, const int* __restrict__ tile_ofs,
        const double* __restrict__ tbuf, const long* __restrict__ toff) {
  __shared__ double Vs[RC * (NB + 1)];
  __shared__ double Xs[RC * CT];
  __shared__ double W[NB * CT];
  __shared__ double W2[NB * CT];
  const int bid = blockIdx.x;
  int lo_p = 0, hi_p = b.nprob - 1;
  while (lo_p < hi_p) {
    const int mid = (lo_p + hi_p + 1) >> 1;
    if (tile_ofs[mid] <= bid) lo_p = mid; else hi_p = mid - 1;
  }
  const int p = lo_p;
  const int tile = bid - tile_ofs[p];
  const int n = b.n[p];
  const int mp = m[p];
  const int c0 = tile * CT;
  if (c0 >= mp) return;
  const int ct = min(CT, mp - c0);
  const int nr = n - 2;
  if (nr <= 0) return;
  const double* A = b.A + b.aoff[p];
  double* Zp = Z + zoff[p] + (long)c0 * n;
  const double* Tp = tbuf + toff[p];
  const int nblk = (nr + NB - 1) / NB;
  const int tid = threadIdx.x;
  const int tr = tid / CT, tc = tid % CT;  // (r, c) owner for W
  for (int blk = nblk - 1; blk >= 0; --blk) {
    const int j0 = blk * NB;
    const int nbk = min(NB, nr - j0);
    const int i0 = j0 + 1;
    // W = V^T X
    double acc = 0.0;
    for (int ib = i0; ib < n; ib += RC) {
      __syncthreads();
      for (int idx = tid; idx < RC * NB; idx += TB) {
        const int rr = idx % RC, r = idx / RC;
        const int i = ib + rr;
        Vs[rr * (NB + 1) + r] = (i < n && r < nbk) ? refl(A, n, j0 + r, i) : 0.0;
      }
      for (int idx = tid; idx < RC * CT; idx += TB) {
        const int rr = idx % RC, c = idx / RC;
        const int i = ib + rr;
        Xs[rr * CT + c] = (i < n && c < ct) ? Zp[i + (long)c * n] : 0.0;
      }
      __syncthreads();
      for (int rr = 0; rr < RC; ++rr) acc += Vs[rr * (NB + 1) + tr] * Xs[rr * CT + tc];
    }
    W[tr * CT + tc] = acc;
    __syncthreads();
    // W2 = T W  (T upper triangular)
    {
      double s = 0.0;
      const double* Tb = Tp + (long)blk * NB * NB;
      for (int l = tr; l < nbk; ++l) s += Tb[tr + NB * l] * W[l * CT + tc];
      W2[tr * CT + tc] = (tr < nbk) ? s : 0.0;
    }
    __syncthreads();
    // X -= V W2
    for (int ib = i0; ib < n; ib += RC) {
      __syncthreads();
      for (int idx = tid; idx < RC * NB; idx += TB) {
        const int rr = idx % RC, r = idx / RC;
        const int i = ib + rr;
        Vs[rr * (NB + 1) + r] = (i < n && r < nbk) ? refl(A, n, j0 + r, i) : 0.0;
      }
      __syncthreads();
      for (int idx = tid; idx < RC * CT; idx += TB) {
        const int rr = idx % RC, c = idx / RC;
        const int i = ib + rr;
        if (i >= n || c >= ct) continue;
        double s = 0.0;
        for (int r = 0; r < nbk; ++r) s += Vs[rr * (NB + 1) + r] * W2[r * CT + c];
        Zp[i + (long)c * n] -= s;
      }
    }
    __syncthreads();
  }
}

// LDS-resident versions for problems of at most BT_LDS rows (the
// Rayleigh-Ritz problems of the iterative solver): the global-memory kernels
// above reload the reflectors and the eigenvector tile from memory for every
// 64-row chunk of every block, a round trip per chunk.  Here a work-group
// stages the whole reflector block (and, in the back-transform, its
// eigenvector tile) in LDS once per block.
constexpr int BT_LDS = 448;
constexpr int VLD = NB + 1;   // row stride of a staged reflector block

// T factor of block blk: G = V^T V over the block's rows, then the forward
// recurrence T(0:c, c) = -tau_c T(0:c, 0:c) G(0:c, c), T(c, c) = tau_c.
__global__ void __launch_bounds__(TB)
k_larft_lds(EigBatch b, const int* __restrict__ nblk_ofs, double* __restrict__ tbuf,
            const long* __restrict__ toff) {
  extern __shared__ double sh[];
  const int bid = blockIdx.x;
  int lo_p = 0, hi_p = b.nprob - 1;
  while (lo_p < hi_p) {
    const int mid = (lo_p + hi_p + 1) >> 1;
    if (nblk_ofs[mid] <= bid) lo_p = mid; else hi_p = mid - 1;
  }
  const int p = lo_p;
  const int blk = bid - nblk_ofs[p];
  const int n = b.n[p];
  const int nr = n - 2;
  const int j0 = blk * NB;
  if (nr <= 0 || j0 >= nr) return;
  const int nbk = min(NB, nr - j0);
  const int i0 = j0 + 1, nrow = n - i0;
  const double* A = b.A + b.aoff[p];
  const double* tau = b.tau + b.voff[p];
  double* Vl = sh;                       // nrow x VLD
  double* G = Vl + (long)nrow * VLD;     // NB x NB
  double* Ts = G + NB * NB;              // NB x NB
  const int tid = threadIdx.x;
  for (int r = 0; r < NB; ++r)
    for (int il = tid; il < nrow; il += TB)
      Vl[il * VLD + r] = r < nbk ? refl(A, n, j0 + r, i0 + il) : 0.0;
  for (int e = tid; e < NB * NB; e += TB) Ts[e] = 0.0;
  __syncthreads();
  for (int e = tid; e < NB * NB; e += TB) {
    const int r = e % NB, c = e / NB;
    double g = 0.0;
    if (r < c && c < nbk) {
      double g0 = 0.0, g1 = 0.0;
      int il = 0;
      for (; il + 1 < nrow; il += 2) {
        g0 += Vl[il * VLD + r] * Vl[il * VLD + c];
        g1 += Vl[(il + 1) * VLD + r] * Vl[(il + 1) * VLD + c];
      }
      if (il < nrow) g0 += Vl[il * VLD + r] * Vl[il * VLD + c];
      g = g0 + g1;
    }
    G[e] = g;
  }
  __syncthreads();
  for (int c = 0; c < nbk; ++c) {
    const double tc = tau[j0 + c];
    if (tid < c) {
      const int r = tid;
      double sacc = 0.0;
      for (int l = r; l < c; ++l) sacc += Ts[r + NB * l] * G[l + NB * c];
      Ts[r + NB * c] = -tc * sacc;
    }
    if (tid == 0) Ts[c + NB * c] = tc;
    __syncthreads();
  }
  double* Tout = tbuf + toff[p] + (long)blk * NB * NB;
  for (int e = tid; e < NB * NB; e += TB) Tout[e] = Ts[e];
}

// Z tile (all n rows, CT columns) in LDS for the whole back-transform; per
// block: stage V and T, W = V^T Z, W2 = T W, Z -= V W2.
__global__ void __launch_bounds__(TB)
k_ormtr_lds(EigBatch b, double* __restrict__ Z, const long* __restrict__ zoff,
            const int* __restrict__ m, const int* __restrict__ tile_ofs,
            const double* __restrict__ tbuf, const long* __restrict__ toff) {
  extern __shared__ double sh[];
  const int bid = blockIdx.x;
  int lo_p = 0, hi_p = b.nprob - 1;
  while (lo_p < hi_p) {
    const int mid = (lo_p + hi_p + 1) >> 1;
    if (tile_ofs[mid] <= bid) lo_p = mid; else hi_p = mid - 1;
  }
  const int p = lo_p;
  const int tile = bid - tile_ofs[p];
  const int n = b.n[p];
  const int mp = m[p];
  const int c0 = tile * CT;
  if (c0 >= mp) return;
  const int ct = min(CT, mp - c0);
  const int nr = n - 2;
  if (nr <= 0) return;
  const double* A = b.A + b.aoff[p];
  double* Zp = Z + zoff[p] + (long)c0 * n;
  const double* Tp = tbuf + toff[p];
  const int nmx = b.nmax;
  double* Vl = sh;                        // (nmx - 1) x VLD
  double* Zl = Vl + (long)(nmx - 1) * VLD;  // n x CT, row-major
  double* W = Zl + (long)nmx * CT;        // NB x CT
  double* W2 = W + NB * CT;               // NB x CT
  double* Tl = W2 + NB * CT;              // NB x NB
  const int tid = threadIdx.x;
  for (int c = 0; c < CT; ++c)
    for (int i = tid; i < n; i += TB) Zl[i * CT + c] = c < ct ? Zp[i + (long)c * n] : 0.0;
  const int nblk = (nr + NB - 1) / NB;
  const int tr = tid / CT, tc = tid % CT;   // owner of W(tr, tc)
  for (int blk = nblk - 1; blk >= 0; --blk) {
    const int j0 = blk * NB;
    const int nbk = min(NB, nr - j0);
    const int i0 = j0 + 1, nrow = n - i0;
    __syncthreads();   // previous block's update of Zl done
    for (int r = 0; r < NB; ++r)
      for (int il = tid; il < nrow; il += TB)
        Vl[il * VLD + r] = r < nbk ? refl(A, n, j0 + r, i0 + il) : 0.0;
    const double* Tb = Tp + (long)blk * NB * NB;
    for (int e = tid; e < NB * NB; e += TB) Tl[e] = Tb[e];
    __syncthreads();
    {
      double w0 = 0.0, w1 = 0.0;
      const double* zc = Zl + (long)i0 * CT + tc;
      int il = 0;
      for (; il + 1 < nrow; il += 2) {
        w0 += Vl[il * VLD + tr] * zc[il * CT];
        w1 += Vl[(il + 1) * VLD + tr] * zc[(il + 1) * CT];
      }
      if (il < nrow) w0 += Vl[il * VLD + tr] * zc[il * CT];
      W[tr * CT + tc] = w0 + w1;
    }
    __syncthreads();
    {
      double sacc = 0.0;
      for (int l = tr; l < nbk; ++l) sacc += Tl[tr + NB * l] * W[l * CT + tc];
      W2[tr * CT + tc] = tr < nbk ? sacc : 0.0;
    }
    __syncthreads();
    for (int e = tid; e < nrow * CT; e += TB) {
      const int il = e / CT, c = e % CT;
      double sacc = 0.0;
#pragma unroll 8
      for (int r = 0; r < NB; ++r) sacc += Vl[il * VLD + r] * W2[r * CT + c];
      Zl[(i0 + il) * CT + c] -= sacc;
    }
  }
  __syncthreads();
  for (int c = 0; c < ct; ++c)
    for (int i = tid; i < n; i += TB) Zp[i + (long)c * n] = Zl[i * CT + c];
}

void launch_backtransform(hipStream_t s, const EigBatch& b, double* Z,
                          const long* zoff, const int* m, const int* h_n,
                          const int* h_m, double* tbuf, const long* toff,
                          long tbuf_total, DevBuf& ofs) {
  (void)tbuf_total;
  // host-side offsets of blocks and tiles, uploaded through a small buffer
  const int np = b.nprob;
  int* h = (int*)malloc(sizeof(int) * 2 * (np + 1));
  int nbt = 0, ntt = 0;
  for (int p = 0; p < np; ++p) {
    h[p] = nbt;
    h[np + 1 + p] = ntt;
    const int nr = h_n[p] - 2;
    nbt += nr > 0 ? (nr + NB - 1) / NB : 0;
    ntt += (nr > 0 && h_m[p] > 0) ? (h_m[p] + CT - 1) / CT : 0;
  }
  h[np] = nbt;
  h[np + 1 + np] = ntt;
  ofs.ensure(sizeof(int) * 2 * (np + 1));
  int* d = ofs.as<int>();
  DVR_HIP(hipMemcpyAsync(d, h, sizeof(int) * 2 * (np + 1), hipMemcpyHostToDevice, s));
  double obytes = 0.0, oflops = 0.0;
  for (int p = 0; p < np; ++p) {
    const double n = h_n[p], mm = h_m[p];
    obytes += 8.0 * (n * n * 0.5 + 2.0 * n * mm * ((n - 2.0) / NB > 1.0 ? (n - 2.0) / NB : 1.0));
    oflops += 4.0 * n * n * mm;
  }
  const bool lds = b.nmax <= BT_LDS && bt_use_lds();
  if (nbt > 0) {
    if (lds) {
      const size_t sh = sizeof(double) * ((size_t)(b.nmax - 1) * VLD + 2 * NB * NB);
      DVR_HIP(hipFuncSetAttribute((const void*)k_larft_lds,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh));
      DVR_LAUNCH("larft", 0.0, 0.0, k_larft_lds, dim3(nbt), dim3(TB), sh, s, b, d, tbuf, toff);
    } else {
      DVR_LAUNCH("larft", 0.0, 0.0, k_larft, dim3(nbt), dim3(TB), 0, s, b, d, tbuf, toff);
    }
  }
  if (ntt > 0) {
    if (lds) {
      const size_t sh = sizeof(double) * ((size_t)(b.nmax - 1) * VLD + (size_t)b.nmax * CT +
                                          2 * NB * CT + NB * NB);
      DVR_HIP(hipFuncSetAttribute((const void*)k_ormtr_lds,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh));
      DVR_LAUNCH("ormtr", obytes, oflops, k_ormtr_lds, dim3(ntt), dim3(TB), sh, s, b, Z, zoff, m,
                 d + np + 1, tbuf, toff);
    } else {
      DVR_LAUNCH("ormtr", obytes, oflops, k_ormtr, dim3(ntt), dim3(TB), 0, s, b, Z, zoff, m,
                 d + np + 1, tbuf, toff);
    }
  }
  DVR_LAUNCH_CHECK();
  DVR_HIP(hipStreamSynchronize(s));
  free(h);
}

}  // namespace dvr
