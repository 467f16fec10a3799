// General FP64 matrix product on MFMA, column-major:
//
//   C = alpha * op(A) * op(B) + beta * C,   op(X) = X or X^T
//
// The workhorse of the block-Davidson final solve (H V products, projections
// V^T W, Ritz-vector formation V Y, Gram-Schmidt) and of the blocked LDL^T
// factorisation (trailing update A22 -= L21 W21^T, lower tiles only).
//
// Work-group = 64x64 tile of C, 4 wavefronts in 2x2, each a 32x32 sub-tile
// of 2x2 v_mfma_f64_16x16x4f64 accumulators; operands staged through LDS in
// k-slices of 16 with the next slice prefetched into registers while the
// current one is multiplied.  Lane layout of v_mfma_f64_16x16x4f64 (checked
// on gfx950): A(i,k) from lane i + 16k, B(k,j) from lane j + 16k, result
// register r of lane l is D(4r + l/16, l%16); the kernel feeds the B slice as
// the first operand so that D is the transposed sub-tile.
//
// K can be split over blockIdx.z: each slice then writes its partial tile to
// a workspace and a second kernel sums the slices (deterministic, no atomics).
#include "dvr_internal.h"

#include <type_traits>

namespace dvr {

namespace {
constexpr int TM = 64;
constexpr int KT0 = 16;   // k-slice granularity of the split-K chunks
constexpr int LDT = TM + 1;
typedef double v4d __attribute__((ext_vector_type(4)));

template <bool TA>
__device__ inline double load_a(const GemmArgs& g, int i, int k) {
  if (i >= g.M || k >= g.K) return 0.0;
  return TA ? g.A[k + (long)i * g.lda] : g.A[i + (long)k * g.lda];
}
template <bool TB>
__device__ inline double load_b(const GemmArgs& g, int k, int j) {
  if (j >= g.N || k >= g.K) return 0.0;
  return TB ? g.B[j + (long)k * g.ldb] : g.B[k + (long)j * g.ldb];
}
}  // namespace

// Lower-tile enumeration shared by the tile kernels: block L of a one-
// dimensional grid over the nt(nt+1)/2 tiles on or below the diagonal, in
// groups of 8 row-tiles walked column by column.
__device__ inline void lower_tile(long L, int nt, int& ti, int& tj) {
  int r0 = (int)((sqrt(8.0 * (double)L + 1.0) - 1.0) * 0.5);
  r0 -= r0 & 7;
  while (r0 > 0 && (long)r0 * (r0 + 1) / 2 > L) r0 -= 8;
  while (r0 + 8 < nt && (long)(r0 + 8) * (r0 + 9) / 2 <= L) r0 += 8;
  const int r1 = min(r0 + 8, nt), gm = r1 - r0;
  long q = L - (long)r0 * (r0 + 1) / 2;
  if (q < (long)r0 * gm) {
    tj = (int)(q / gm);
    ti = r0 + (int)(q % gm);
  } else {
    q -= (long)r0 * gm;
    int c = r0;
    while (q >= r1 - c) {
      q -= r1 - c;
      ++c;
    }
    tj = c;
    ti = c + (int)q;
  }
}

// XCD-aware enumeration of the lower tiles (lower == 2).  Work-group L runs on
// XCD L % 8 (round-robin dispatch); the lower triangle is cut into super-tiles
// of 8 x 8 tiles, super-tile S goes to XCD S % 8, and the 64 work-groups of a
// super-tile follow one another on their XCD, so that its 8 row panels and 8
// column panels are read from HBM once into that XCD's L2 while the
// super-tiles stay spread evenly over the XCDs.  Returns false for the empty
// slots (above the diagonal of a diagonal super-tile, beyond the last tile).
__device__ inline bool lower_tile_xcd(long L, int nt, int& ti, int& tj) {
  const int xcd = (int)(L & 7);
  const long j = L >> 3;
  const long S = xcd + 8 * (j >> 6);
  const int slot = (int)(j & 63);
  const int ns = (nt + 7) >> 3;
  if (S >= (long)ns * (ns + 1) / 2) return false;
  int si = (int)((sqrt(8.0 * (double)S + 1.0) - 1.0) * 0.5);
  while ((long)si * (si + 1) / 2 > S) --si;
  while ((long)(si + 1) * (si + 2) / 2 <= S) ++si;
  const int sj = (int)(S - (long)si * (si + 1) / 2);
  ti = si * 8 + (slot & 7);
  tj = sj * 8 + (slot >> 3);
  return ti < nt && tj <= ti;
}

template <bool TA, bool TB, int KT, bool CP>
__global__ void __launch_bounds__(256) k_gemm(GemmArgs g0) {
  constexpr int NQ = KT / 4;   // elements of each operand per thread and slice
  __shared__ double As[KT * LDT];
  __shared__ double Bs[KT * LDT];
  GemmArgs g = g0;
  int z = blockIdx.z;
  if (g.nbatch > 1) {
    g.A += z * g.sa;
    g.B += z * g.sb;
    g.C += z * g.sc;
    z = 0;
  }
  int ti = blockIdx.x, tj = blockIdx.y;
  // 1D grid over the lower tiles: neighbouring work-groups (dealt round-robin
  // to the XCDs) share their column panel, and each XCD keeps one row panel
  // per group of 8 row-tiles in L2
  if (g.lower == 2) {
    if (!lower_tile_xcd(blockIdx.x, g.ntile, ti, tj)) return;
  } else if (g.lower) {
    lower_tile(blockIdx.x, g.ntile, ti, tj);
  }
  const int row0 = ti * TM, col0 = tj * TM;
  const int kbeg = z * g.kchunk;
  const int kend = min(g.K, kbeg + g.kchunk);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wy = wave >> 1, wx = wave & 1;
  const int l16 = lane & 15, lk = lane >> 4;

  // element (kk, cc) of the 16x64 slice handled by this thread in round q:
  // the fast index follows the contiguous direction of the operand in memory
  auto a_pos = [&](int q, int& kk, int& cc) {
    const int idx = tid + 256 * q;
    if (TA) { kk = idx % KT; cc = idx / KT; } else { cc = idx & 63; kk = idx >> 6; }
  };
  auto b_pos = [&](int q, int& kk, int& cc) {
    const int idx = tid + 256 * q;
    if (TB) { cc = idx & 63; kk = idx >> 6; } else { kk = idx % KT; cc = idx / KT; }
  };

  v4d acc[2][2];
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) acc[a][b] = v4d{0.0, 0.0, 0.0, 0.0};

  // C of this thread's 16 outputs, read now so that the load latency is
  // hidden behind the k loop (beta != 0 without a K split)
  const bool cpre = CP && g.beta != 0.0 && g.ksplit == 1;
  double cv[2][2][4];
  if (cpre) {
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
      for (int bj = 0; bj < 2; ++bj)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int i = row0 + wy * 32 + bi * 16 + l16;
          const int j = col0 + wx * 32 + bj * 16 + 4 * r + lk;
          cv[bi][bj][r] = (i < g.M && j < g.N) ? g.C[i + (long)g.ldc * j] : 0.0;
        }
  }

  // the next operand slice is staged through registers while the current
  // one is multiplied
  double ra[NQ], rb[NQ];
  // tiles entirely inside C with whole k slices load through running
  // pointers (no per-element bounds arithmetic); edge tiles go element-wise
  const bool full = row0 + TM <= g.M && col0 + TM <= g.N && (kend - kbeg) % KT == 0;
  const double* pa;
  const double* pb;
  long qa, qb, sa, sb;   // pointer steps per round q and per k slice
  if (TA) {
    pa = g.A + kbeg + tid % KT + (long)(row0 + tid / KT) * g.lda;
    qa = (long)(256 / KT) * g.lda;
    sa = KT;
  } else {
    pa = g.A + row0 + (tid & 63) + (long)(kbeg + (tid >> 6)) * g.lda;
    qa = 4 * g.lda;
    sa = (long)KT * g.lda;
  }
  if (TB) {
    pb = g.B + col0 + (tid & 63) + (long)(kbeg + (tid >> 6)) * g.ldb;
    qb = 4 * g.ldb;
    sb = (long)KT * g.ldb;
  } else {
    pb = g.B + kbeg + tid % KT + (long)(col0 + tid / KT) * g.ldb;
    qb = (long)(256 / KT) * g.ldb;
    sb = KT;
  }
  auto fetch = [&](int k0) {
    if (full) {
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        ra[q] = pa[q * qa];
        rb[q] = pb[q * qb];
      }
      pa += sa;
      pb += sb;
      return;
    }
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      int kk, cc;
      a_pos(q, kk, cc);
      ra[q] = (k0 + kk < kend) ? load_a<TA>(g, row0 + cc, k0 + kk) : 0.0;
      b_pos(q, kk, cc);
      rb[q] = (k0 + kk < kend) ? load_b<TB>(g, k0 + kk, col0 + cc) : 0.0;
    }
  };
  auto stage = [&]() {
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      int kk, cc;
      a_pos(q, kk, cc);
      As[kk * LDT + cc] = ra[q];
      b_pos(q, kk, cc);
      Bs[kk * LDT + cc] = rb[q];
    }
  };
  auto multiply = [&]() {
    // all fragments of the slice first (distinct registers), so that the LDS
    // reads of later k steps are in flight while the first MFMAs run
    double fa[KT / 4][2], fb[KT / 4][2];
#pragma unroll
    for (int t = 0; t < KT / 4; ++t) {
      const int kr = (4 * t + lk) * LDT;
      fa[t][0] = As[kr + wy * 32 + l16];
      fa[t][1] = As[kr + wy * 32 + 16 + l16];
      fb[t][0] = Bs[kr + wx * 32 + l16];
      fb[t][1] = Bs[kr + wx * 32 + 16 + l16];
    }
    // keep the scheduler from sinking the reads back between the MFMAs
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int t = 0; t < KT / 4; ++t) {
      // B operand first: the accumulator holds the transposed sub-tile, so
      // that lanes 0..15 of a result register own 16 consecutive rows of C
      // (coalesced stores in column-major C)
      acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(fb[t][0], fa[t][0], acc[0][0], 0, 0, 0);
      acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(fb[t][1], fa[t][0], acc[0][1], 0, 0, 0);
      acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(fb[t][0], fa[t][1], acc[1][0], 0, 0, 0);
      acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(fb[t][1], fa[t][1], acc[1][1], 0, 0, 0);
    }
  };
  if (kbeg < kend) fetch(kbeg);
  for (int k0 = kbeg; k0 < kend; k0 += KT) {
    stage();
    __syncthreads();
    if (k0 + KT < kend) fetch(k0 + KT);
    multiply();
    __syncthreads();
  }

  // register r of lane l holds C(i, j) with i = l%16, j = 4r + l/16 inside
  // the 16x16 block (bi, bj)
#pragma unroll
  for (int bi = 0; bi < 2; ++bi)
#pragma unroll
    for (int bj = 0; bj < 2; ++bj)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = row0 + wy * 32 + bi * 16 + l16;
        const int j = col0 + wx * 32 + bj * 16 + 4 * r + lk;
        if (i >= g.M || j >= g.N) continue;
        if (g.ksplit > 1) {
          g.work[(long)z * g.M * g.N + i + (long)g.M * j] = acc[bi][bj][r];
        } else {
          double* c = g.C + i + (long)g.ldc * j;
          if (cpre)
            *c = g.alpha * acc[bi][bj][r] + g.beta * cv[bi][bj][r];
          else
            *c = g.beta == 0.0 ? g.alpha * acc[bi][bj][r] : g.alpha * acc[bi][bj][r] + g.beta * *c;
        }
      }
}

// The lower 64 x 64 tiles (ti >= tj) of an nt x nt tile grid, enumerated in
// super-tiles of S x S tiles: super-rows top to bottom, their super-tiles
// left to right, tiles column by column inside (the diagonal super-tile's
// lower triangle likewise).  lower64_count gives their number, lower64_tile
// tile number idx.
__device__ inline long lower64_count(int nt, int S) {
  long n = 0;
  for (int si = 0; S * si < nt; ++si) {
    const int h = min(S, nt - S * si);
    n += (long)S * h * si + (long)h * (h + 1) / 2;
  }
  return n;
}
__device__ inline void lower64_tile(long idx, int nt, int S, int& ti, int& tj) {
  int si = 0;
  long base = 0;
  for (;;) {
    const int h = min(S, nt - S * si);
    const long cnt = (long)S * h * si + (long)h * (h + 1) / 2;
    if (idx < base + cnt) break;
    base += cnt;
    ++si;
  }
  const int h = min(S, nt - S * si);
  long rem = idx - base;
  if (rem < (long)S * h * si) {
    const int sj = (int)(rem / (S * h));
    const int loc = (int)(rem % (S * h));
    tj = S * sj + loc / h;
    ti = S * si + loc % h;
  } else {
    rem -= (long)S * h * si;
    int c = 0;
    while (rem >= h - c) {
      rem -= h - c;
      ++c;
    }
    tj = S * si + c;
    ti = S * si + c + (int)rem;
  }
}

// Lower-tile NT update C = alpha A B^T + beta C on the 64 x 64 blocks on or
// below the diagonal of an M x M output (M >= 64, K = 4 NSTEPS): the
// trailing update of the blocked LDL^T factorisation.
//
// Persistent: the grid is 8 x (work-groups per XCD) of four independent
// wavefronts (no LDS, no barrier), one wavefront per SIMD.  The valid tiles
// (lower64_tile order) are cut into eight contiguous ranges, range x for the
// work-groups b with b % 8 == x (the XCD that round-robin dispatch puts them
// on), and dealt round-robin to that XCD's wavefronts, so every wavefront
// gets within one tile of the mean and each XCD's L2 holds the panels of the
// super-tiles it works on.  A wavefront owns one 64 x 64 tile at a time as
// 4 x 4 MFMA accumulators and streams
//   - its operand fragments, A(r0 + 16f + p, 4t + q) and B(c0 + 16f + p,
//     4t + q) for lane p + 16q, through a ring of D register sets: the loads
//     of step t + D are in flight while the 16 MFMAs of step t run, and the
//     ring runs on into the next tile, so a tile starts without a load wait;
//   - its C tile, 64 / NSTEPS loads per k step into registers, so that C is
//     read while the MFMAs run and the epilogue only writes.  C is read and
//     written once, with non-temporal loads and stores, so that it streams
//     past L2 without evicting the panels the XCD's other tiles still read
//     (fetched bytes -21% on a 19456^2 update, same time).
// The k loop is unrolled (the C registers need compile-time indices).  The
// k order of the sums (MFMA groups of 4, t ascending) and the epilogue
// alpha acc + beta C are those of k_gemm: both kernels give the same bits.
template <int NSTEPS, int D>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1)))
k_gemm_lower_w64(GemmArgs g) {
  // the ring runs on into the next tile, whose step t must find its
  // fragments in slot t % D: D divides NSTEPS
  static_assert(NSTEPS % 16 == 0 && NSTEPS <= 64 && NSTEPS % D == 0, "k steps");
  constexpr int CPS = 64 / NSTEPS;   // C loads per k step
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int p = lane & 15, q = lane >> 4;
  const int M = g.M, nt = (M + 63) >> 6;
  const int S = g.st;
  const long ntiles = lower64_count(nt, S);
  const int x = blockIdx.x & 7;
  const long W = (long)(gridDim.x >> 3) * 4;   // wavefronts per XCD
  const long end = ntiles * (x + 1) / 8;
  long idx = ntiles * x / 8 + (long)(blockIdx.x >> 3) * 4 + wave;
  if (idx >= end) return;
  int ti, tj;
  lower64_tile(idx, nt, S, ti, tj);
  // a tile of the ragged last tile row (M % 64 != 0) computes the 64 rows
  // (and, on the diagonal, columns) ending at M, so that every load stays
  // inside the matrix, and stores only its own rows and columns
  auto base = [&](int t) { return min(t * 64, M - 64); };
  const long sa = 4 * g.lda, sb = 4 * g.ldb, ldc = g.ldc;
  const double* pa = g.A + base(ti) + p + (long)q * g.lda;
  const double* pb = g.B + base(tj) + p + (long)q * g.ldb;
  double fa[D][4], fb[D][4];
  auto load = [&](int slot, const double* a, const double* b) {
#pragma unroll
    for (int f = 0; f < 4; ++f) {
      fa[slot][f] = a[16 * f];
      fb[slot][f] = b[16 * f];
    }
  };
#pragma unroll
  for (int s = 0; s < D; ++s) load(s, pa + s * sa, pb + s * sb);
  v4d acc[4][4];
  double cv[4][4][4];
  const double alpha = g.alpha, beta = g.beta;
  for (;;) {
    // the wavefront's next tile: its first D steps load during this one
    const long in = idx + W;
    const bool more = in < end;
    int tin = ti, tjn = tj;
    if (more) lower64_tile(in, nt, S, tin, tjn);
    const double* pan = g.A + base(tin) + p + (long)q * g.lda;
    const double* pbn = g.B + base(tjn) + p + (long)q * g.ldb;
    const int lr = base(ti), lc = base(tj);
    const double* pc = g.C + (lr + p) + (long)(lc + q) * ldc;
#pragma unroll
    for (int f = 0; f < 4; ++f)
#pragma unroll
      for (int h = 0; h < 4; ++h) acc[f][h] = v4d{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int t = 0; t < NSTEPS; ++t) {
      const int s = t % D;
      // B fragment first: register r of lane p + 16q then holds
      // C(r0 + 16f + p, c0 + 16h + 4r + q)
#pragma unroll
      for (int f = 0; f < 4; ++f)
#pragma unroll
        for (int h = 0; h < 4; ++h)
          acc[f][h] = __builtin_amdgcn_mfma_f64_16x16x4f64(fb[s][h], fa[s][f], acc[f][h], 0, 0, 0);
      if (t + D < NSTEPS) load(s, pa + (t + D) * sa, pb + (t + D) * sb);
      else load(s, pan + (t + D - NSTEPS) * sa, pbn + (t + D - NSTEPS) * sb);
#pragma unroll
      for (int j = t * CPS; j < (t + 1) * CPS; ++j) {
        const int f = j >> 4, h = (j >> 2) & 3, r = j & 3;
        cv[f][h][r] = __builtin_nontemporal_load(pc + 16 * f + (long)(16 * h + 4 * r) * ldc);
      }
      // keep each step's loads where they are (no hoisting into a burst)
      __builtin_amdgcn_sched_barrier(0);
    }
    // two copies of the epilogue: straight stores for whole tiles, stores
    // under a lane mask for a ragged one (a wave-uniform branch)
    auto epilogue = [&](auto guarded) {
#pragma unroll
      for (int f = 0; f < 4; ++f) {
        double* c = g.C + (lr + 16 * f + p) + (long)(lc + q) * ldc;
#pragma unroll
        for (int h = 0; h < 4; ++h)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const double v =
                beta == 0.0 ? alpha * acc[f][h][r] : alpha * acc[f][h][r] + beta * cv[f][h][r];
            if (!decltype(guarded)::value ||
                (lr + 16 * f + p >= ti * 64 && lc + 16 * h + 4 * r + q >= tj * 64))
              __builtin_nontemporal_store(v, c + (long)(16 * h + 4 * r) * ldc);
          }
      }
    };
    if (lr != ti * 64 || lc != tj * 64) epilogue(std::true_type{});
    else epilogue(std::false_type{});
    if (!more) break;
    idx = in;
    ti = tin;
    tj = tjn;
    pa = pan;
    pb = pbn;
  }
}

__global__ void __launch_bounds__(256) k_gemm_reduce(GemmArgs g) {
  const long n = (long)g.M * g.N;
  for (long t = blockIdx.x * 256L + threadIdx.x; t < n; t += (long)gridDim.x * 256) {
    const int i = (int)(t % g.M), j = (int)(t / g.M);
    if (g.lower && (i / TM) < (j / TM)) continue;
    double s = 0.0;
    for (int z = 0; z < g.ksplit; ++z) s += g.work[(long)z * n + t];
    double* c = g.C + i + (long)g.ldc * j;
    *c = g.beta == 0.0 ? g.alpha * s : g.alpha * s + g.beta * *c;
  }
}

int gemm_auto_ksplit(int M, int N, int K) {
  const long tiles = (long)((M + TM - 1) / TM) * ((N + TM - 1) / TM);
  if (tiles >= 768 || K < 512) return 1;
  long ks = (1024 + tiles - 1) / tiles;
  ks = std::min<long>(ks, std::max(1, K / 256));
  return (int)std::max<long>(ks, 1);
}

size_t gemm_work_doubles(int M, int N, int ksplit) {
  return ksplit > 1 ? (size_t)M * N * ksplit : 0;
}

static void launch_gemm_kt(hipStream_t s, const char* name, bool ta, bool tb, int M, int N, int K,
                           double alpha, const double* A, long lda, const double* B, long ldb,
                           double beta, double* C, long ldc, int ksplit, double* work, bool lower,
                           int kt, bool cp = false, bool xcd_lower = false) {
  if (M <= 0 || N <= 0) return;
  if (K <= 0) ksplit = 1;
  if (ksplit < 1) ksplit = 1;
  GemmArgs g;
  g.M = M; g.N = N; g.K = K; g.A = A; g.lda = lda; g.B = B; g.ldb = ldb; g.C = C; g.ldc = ldc;
  g.alpha = alpha; g.beta = beta; g.lower = lower ? 1 : 0;
  // chunks of whole 32-deep slices (a multiple of both slice depths)
  int kchunk = (K + ksplit - 1) / ksplit;
  kchunk = ((kchunk + 31) / 32) * 32;
  if (kchunk <= 0) kchunk = 32;
  ksplit = K > 0 ? (K + kchunk - 1) / kchunk : 1;
  if (ksplit > 1 && !work) throw HipError("gemm: split-K needs a workspace");
  g.ksplit = ksplit; g.kchunk = kchunk; g.work = work;
  if (lower && M != N) throw HipError("gemm: lower-tile product needs a square output");
  g.ntile = (M + TM - 1) / TM;
  if (lower && xcd_lower && ksplit == 1) g.lower = 2;
  long lower_wgs = (long)g.ntile * (g.ntile + 1) / 2;
  if (g.lower == 2) {
    const long ns = (g.ntile + 7) / 8;
    lower_wgs = 8L * 64 * ((ns * (ns + 1) / 2 + 7) / 8);
  }
  const dim3 grid(lower ? (unsigned)lower_wgs : (M + TM - 1) / TM,
                  lower ? 1u : (unsigned)((N + TM - 1) / TM), ksplit);
  const double flops = 2.0 * M * (double)N * K * (lower ? 0.5 : 1.0);
  const double bytes = 8.0 * ((double)M * K + (double)K * N +
                              (beta != 0.0 ? 2.0 : 1.0) * M * N * (lower ? 0.5 : 1.0));
#define DVR_GEMM_LAUNCH(KT_, CP_)                                                          \
  do {                                                                                          \
    if (ta && tb) DVR_LAUNCH(name, bytes, flops, (k_gemm<true, true, KT_, CP_>), grid,      \
                             dim3(256), 0, s, g);                                               \
    else if (ta) DVR_LAUNCH(name, bytes, flops, (k_gemm<true, false, KT_, CP_>), grid,      \
                            dim3(256), 0, s, g);                                                \
    else if (tb) DVR_LAUNCH(name, bytes, flops, (k_gemm<false, true, KT_, CP_>), grid,      \
                            dim3(256), 0, s, g);                                                \
    else DVR_LAUNCH(name, bytes, flops, (k_gemm<false, false, KT_, CP_>), grid, dim3(256), \
                    0, s, g);                                                                   \
  } while (0)
  if (kt == 32) {
    if (cp) DVR_GEMM_LAUNCH(32, true);
    else DVR_GEMM_LAUNCH(32, false);
  } else {
    if (cp) DVR_GEMM_LAUNCH(16, true);
    else DVR_GEMM_LAUNCH(16, false);
  }
#undef DVR_GEMM_LAUNCH
  if (ksplit > 1) {
    const long n = (long)M * N;
    const int blocks = (int)std::min<long>((n + 255) / 256, 4096);
    DVR_LAUNCH("gemm_reduce", 8.0 * n * (ksplit + 1), 0.0, k_gemm_reduce, dim3(blocks), dim3(256),
               0, s, g);
  }
  DVR_LAUNCH_CHECK();
}

void launch_gemm_batched(hipStream_t s, const char* name, bool ta, bool tb, int M, int N, int K,
                         double alpha, const double* A, long lda, long sa, const double* B,
                         long ldb, long sb, double beta, double* C, long ldc, long sc,
                         int nbatch) {
  if (M <= 0 || N <= 0 || nbatch <= 0) return;
  GemmArgs g;
  g.M = M; g.N = N; g.K = K; g.A = A; g.lda = lda; g.B = B; g.ldb = ldb; g.C = C; g.ldc = ldc;
  g.alpha = alpha; g.beta = beta; g.lower = 0;
  g.kchunk = K > 0 ? ((K + 31) / 32) * 32 : 32;
  g.ksplit = 1; g.work = nullptr;
  g.ntile = (M + TM - 1) / TM;
  g.sa = sa; g.sb = sb; g.sc = sc; g.nbatch = nbatch;
  const dim3 grid((M + TM - 1) / TM, (N + TM - 1) / TM, nbatch);
  const double flops = 2.0 * M * (double)N * K * nbatch;
  const double bytes = 8.0 * nbatch * ((double)M * K + (double)K * N +
                                       (beta != 0.0 ? 2.0 : 1.0) * M * N);
  if (ta && tb) DVR_LAUNCH(name, bytes, flops, (k_gemm<true, true, 16, false>), grid, dim3(256), 0, s, g);
  else if (ta) DVR_LAUNCH(name, bytes, flops, (k_gemm<true, false, 16, false>), grid, dim3(256), 0, s, g);
  else if (tb) DVR_LAUNCH(name, bytes, flops, (k_gemm<false, true, 16, false>), grid, dim3(256), 0, s, g);
  else DVR_LAUNCH(name, bytes, flops, (k_gemm<false, false, 16, false>), grid, dim3(256), 0, s, g);
  DVR_LAUNCH_CHECK();
}

void launch_gemm_var(hipStream_t s, const char* name, bool ta, bool tb, int M, int N, int K,
                     double alpha, const double* A, long lda, const double* B, long ldb,
                     double beta, double* C, long ldc, bool lower, int kt, bool cp,
                     bool xcd_lower) {
  launch_gemm_kt(s, name, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, 1, nullptr, lower,
                 kt, cp, xcd_lower);
}

// super-tile width (in 64-row tiles) of the persistent lower-tile update;
// DVR_LDL_ST overrides it
int ldl_super_tile() {
  const char* v = getenv("DVR_LDL_ST");
  const int st = v ? atoi(v) : 16;
  return st >= 1 ? st : 16;
}

void launch_gemm_lower_nt(hipStream_t s, const char* name, int M, int K, double alpha,
                          const double* A, long lda, const double* B, long ldb, double beta,
                          double* C, long ldc, int wgs_per_xcd, bool xcd_lower) {
  if (M <= 0) return;
  if (M < 64 || (K != 64 && K != 128 && K != 256) || wgs_per_xcd < 1) {
    const std::string small = std::string(name) + "_small";
    launch_gemm_kt(s, small.c_str(), false, true, M, M, K, alpha, A, lda, B, ldb, beta, C, ldc, 1,
                   nullptr, true, 16, true, xcd_lower);
    return;
  }
  GemmArgs g;
  g.M = M; g.N = M; g.K = K; g.A = A; g.lda = lda; g.B = B; g.ldb = ldb;
  g.C = C; g.ldc = ldc; g.alpha = alpha; g.beta = beta;
  g.ksplit = 1; g.kchunk = K; g.work = nullptr; g.lower = 1; g.ntile = (M + 63) / 64;
  g.st = ldl_super_tile();
  const dim3 grid(8u * (unsigned)wgs_per_xcd);
  const double flops = (double)M * M * K;   // lower half: 2 M^2 K / 2
  const double bytes = 8.0 * (2.0 * M * K + (beta != 0.0 ? 2.0 : 1.0) * 0.5 * M * M);
  if (K == 256) DVR_LAUNCH(name, bytes, flops, (k_gemm_lower_w64<64, 4>), grid, dim3(256), 0, s, g);
  else if (K == 128) DVR_LAUNCH(name, bytes, flops, (k_gemm_lower_w64<32, 4>), grid, dim3(256), 0, s, g);
  else DVR_LAUNCH(name, bytes, flops, (k_gemm_lower_w64<16, 4>), grid, dim3(256), 0, s, g);
  DVR_LAUNCH_CHECK();
}

void launch_gemm(hipStream_t s, const char* name, bool ta, bool tb, int M, int N, int K,
                 double alpha, const double* A, long lda, const double* B, long ldb,
                 double beta, double* C, long ldc, int ksplit, double* work, bool lower) {
  launch_gemm_kt(s, name, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, ksplit, work, lower,
                 16);
}

// ---------------------------------------------------------------------------
// Autotuner: the first call of a shape times the candidate (k-slice depth,
// K split) pairs on a scratch output and keeps the fastest; later calls of
// the shape reuse the choice.
// ---------------------------------------------------------------------------
static bool capturing(hipStream_t s) {
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  DVR_HIP(hipStreamIsCapturing(s, &st));
  return st != hipStreamCaptureStatusNone;
}

GemmTuner::~GemmTuner() {
  if (ev0) (void)hipEventDestroy(ev0);
  if (ev1) (void)hipEventDestroy(ev1);
}

void GemmTuner::run(hipStream_t s, const char* name, bool ta, bool tb, int M, int N, int K,
                    double alpha, const double* A, long lda, const double* B, long ldb,
                    double beta, double* C, long ldc, bool lower) {
  if (M <= 0 || N <= 0) return;
  const Key key{ta, tb, lower, M, N, K};
  Choice ch{16, gemm_auto_ksplit(M, N, K)};
  auto it = cache.find(key);
  if (it != cache.end()) {
    ch = it->second;
  } else if (enabled && !g_prof && (double)M * N * K >= 1e7 && !capturing(s)) {
    if (!ev0) {
      DVR_HIP(hipEventCreate(&ev0));
      DVR_HIP(hipEventCreate(&ev1));
    }
    const int ka = gemm_auto_ksplit(M, N, K);
    std::vector<Choice> cand;
    for (int kt : {16, 32})
      for (int ks : {1, ka, 2 * ka, 4 * ka}) {
        if (ks > 1 && K / ks < 64) continue;
        bool dup = false;
        for (const Choice& c : cand) dup |= (c.kt == kt && c.ksplit == ks);
        if (!dup) cand.push_back(Choice{kt, ks});
      }
    scratch.ensure(sizeof(double) * (size_t)M * N);
    float best = 1e30f;
    for (const Choice& c : cand) {
      if (c.ksplit > 1) work.ensure(sizeof(double) * gemm_work_doubles(M, N, c.ksplit));
      float t = 1e30f;
      for (int rep = 0; rep < 2; ++rep) {
        DVR_HIP(hipEventRecord(ev0, s));
        launch_gemm_kt(s, "tune", ta, tb, M, N, K, 1.0, A, lda, B, ldb, 0.0, scratch.as<double>(), M,
                       c.ksplit, c.ksplit > 1 ? work.as<double>() : nullptr, lower, c.kt);
        DVR_HIP(hipEventRecord(ev1, s));
        DVR_HIP(hipEventSynchronize(ev1));
        float ms = 0.f;
        DVR_HIP(hipEventElapsedTime(&ms, ev0, ev1));
        t = std::min(t, ms);
      }
      if (t < best) {
        best = t;
        ch = c;
      }
    }
    cache[key] = ch;
  }
  if (ch.ksplit > 1) work.ensure(sizeof(double) * gemm_work_doubles(M, N, ch.ksplit));
  launch_gemm_kt(s, name, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, ch.ksplit,
                 ch.ksplit > 1 ? work.as<double>() : nullptr, lower, ch.kt);
}

}  // namespace dvr
