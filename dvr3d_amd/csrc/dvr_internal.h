// Internal declarations shared by the HIP translation units of libdvr3d_amd.
// Kernels live in k_*.hip; each file exposes host-side launch_* wrappers.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>
#include <cstdint>
#include <cstdio>
#include <string>
#include <stdexcept>
#include <vector>

namespace dvr {

struct HipError : std::runtime_error {
  explicit HipError(const std::string& s) : std::runtime_error(s) {}
};

#define DVR_HIP(call)                                                        \
  do {                                                                       \
    hipError_t err_ = (call);                                                \
    if (err_ != hipSuccess)                                                  \
      throw ::dvr::HipError(std::string(#call) + ": " +                      \
                            hipGetErrorString(err_) + " (" __FILE__ ":" +   \
                            std::to_string(__LINE__) + ")");                 \
  } while (0)

#define DVR_LAUNCH_CHECK() DVR_HIP(hipGetLastError())

// ---------------------------------------------------------------------------
// Optional per-launch timing: while a Profiler is installed (one profiled
// dvr_run), every launch wrapper brackets its kernel with two HIP events on
// the launch stream and records the algorithmic bytes and flops of the launch.
// ---------------------------------------------------------------------------
struct ProfRec {
  std::string name;
  hipEvent_t a, b;
  double bytes, flops;
};
struct Profiler {
  std::vector<hipEvent_t> pool;
  size_t used = 0;
  std::vector<ProfRec> recs;
  std::string phase;
  hipEvent_t event() {
    if (used == pool.size()) {
      hipEvent_t e;
      DVR_HIP(hipEventCreate(&e));
      pool.push_back(e);
    }
    return pool[used++];
  }
  void reset() {
    used = 0;
    recs.clear();
  }
  ~Profiler() {
    for (auto e : pool) (void)hipEventDestroy(e);
  }
};
extern thread_local Profiler* g_prof;
inline void prof_begin(hipStream_t s, const char* kernel, double bytes, double flops) {
  if (!g_prof) return;
  ProfRec r;
  r.name = g_prof->phase.empty() ? std::string(kernel) : g_prof->phase + ":" + kernel;
  r.a = g_prof->event();
  r.b = g_prof->event();
  r.bytes = bytes;
  r.flops = flops;
  DVR_HIP(hipEventRecord(r.a, s));
  g_prof->recs.push_back(r);
}
inline void prof_end(hipStream_t s) {
  if (!g_prof) return;
  DVR_HIP(hipEventRecord(g_prof->recs.back().b, s));
}
inline void prof_phase(const char* ph) {
  if (g_prof) g_prof->phase = ph;
}
// Launch a kernel; while profiling, through hipExtLaunchKernelGGL so the two
// events timestamp the dispatch itself (no extra queue packets around it).
#define DVR_LAUNCH(pname, pbytes, pflops, kern, grid, block, shm, strm, ...)      \
  do {                                                                           \
    if (::dvr::g_prof) {                                                         \
      ::dvr::ProfRec r_;                                                         \
      r_.name = ::dvr::g_prof->phase.empty()                                     \
                    ? std::string(pname)                                         \
                    : ::dvr::g_prof->phase + ":" + (pname);                      \
      r_.a = ::dvr::g_prof->event();                                             \
      r_.b = ::dvr::g_prof->event();                                             \
      r_.bytes = (pbytes);                                                       \
      r_.flops = (pflops);                                                       \
      ::dvr::g_prof->recs.push_back(r_);                                         \
      hipExtLaunchKernelGGL(kern, grid, block, shm, strm, r_.a, r_.b, 0,         \
                            __VA_ARGS__);                                        \
    } else {                                                                     \
      hipLaunchKernelGGL(kern, grid, block, shm, strm, __VA_ARGS__);             \
    }                                                                            \
  } while (0)

// ---------------------------------------------------------------------------
// grid-side kernels (k_grid.hip)
// ---------------------------------------------------------------------------
// Analytic AB2 surface on the (c, b, a) grid, a fastest:
//   V = De[(1-e^{-a(r1-re)})^2 + (1-e^{-a(r2-re)})^2] + kb (x - cte)^2
//       + c (r1-re)(r2-re)
// par = {De, a, re, kb, cte, c}.  zr2r1: c indexes r1 and b indexes r2.
void launch_pes_ab2(hipStream_t s, double* v, const double* rc,
                    const double* rb, const double* xa, int nc, int nb, int na,
                    int zr2r1, const double par[6]);

// Gather the kept 1D eigenpairs into packed per-block storage.
//   src_row[r]  : row of the (nprob*n) x n eigenvector array for packed row r
//   v1 (rows x n, row-major), e1p[r]
void launch_pack_1d(hipStream_t s, const double* v1full, const double* e1full,
                    const int* src_row, int rows, int n, double* v1,
                    double* e1p);

// Per-c-block bookkeeping of the contracted basis.
struct CBlock {
  int rowbase;  // first packed 1D row of this block
  int nham2;    // 2D dimension of the block
  int iv2;      // chosen 2D functions (columns of the 3D basis)
  int col0;     // first 3D basis index of the block
  long z2off;   // offset of the block's 2D vectors (nham2 x iv2, col-major)
};

// Expansion of the chosen 2D vectors on the (a, b) grid of each c block:
//   cint[g + G*(col0+j)] = sum_{k<iv1(c,b)} v1[(rowbase+off+k)*na + a]
//                                          * c2[z2off + off+k + j*nham2]
// with g = a*nb + b and off = iv1off[c*(nb+1)+b].
void launch_cint(hipStream_t s, const double* v1, const double* c2,
                 double* cint, const CBlock* cb, const int* blk3,
                 const int* iv1off, int na, int nb, int n3);

// Wavefunction amplitudes of nvec levels on the DVR grid, in the IWAVE
// record layout psi(idvr, npnt1, npnt2) per level.
void launch_trans(hipStream_t s, const double* cint, const double* c3,
                  double* psi, const CBlock* cb, int nc, int na, int nb,
                  int n3, int nvec, int zr2r1, int idvr, int npnt1,
                  int npnt2);

// ---------------------------------------------------------------------------
// 1D angular eigenproblems (k_jacobi.hip)
// ---------------------------------------------------------------------------
// One workgroup per (c,b) problem: H1 = diag(V + wterm) + wsum*L, parallel
// cyclic Jacobi in LDS, eigenpairs sorted ascending, count of E <= emax.
//   e1[p*n + k], v1[(p*n + k)*n + a] (row k = eigenvector k), cnt[p]
void launch_jacobi1d(hipStream_t s, const double* vgrid, const double* xl,
                     const double* wsum, const double* wterm, int n, int nprob,
                     double emaxau, double* e1, double* v1, int* cnt,
                     double* vscratch, const int* src_prob, int* sweeps);
size_t jacobi1d_scratch_doubles(int n, int nprob);
// Mirror symmetry of the 1D problems (square (c,b) grid): flag = 0 when
// every problem (c,b) has the same inputs as (b,c); then only c >= b need
// solving and k_mirror_copy fills the rest.
void launch_mirror_check(hipStream_t s, const double* v, const double* wsum,
                         const double* wterm, int nb, int na, int* flag);
void launch_mirror_copy(hipStream_t s, int nb, int n, double* e1, double* v1,
                        int* cnt);

// ---------------------------------------------------------------------------
// symmetric rank-K products on MFMA (k_syrk.hip)
// ---------------------------------------------------------------------------
// For each group g:  C_g = (X_g^T X_g) .* S_g + diag(dg_g), with
//   X_g : K_g x N_g column-major (ld = ldx_g),
//   S_g(i,j) = xk[blk_g[i] + ldxk*blk_g[j]]  (a block-constant scale),
//   C_g : N_g x N_g column-major, ld = N_g, both triangles written.
struct SyrkDesc {
  const double* X;
  long ldx;
  int K;
  int N;
  double* C;
  const int* blk;
  const double* diag;
  int tile0;  // first tile index of this group in the flat launch
  int ntile;  // number of lower-triangular tiles
};
// bytes/flops: algorithmic traffic and work of the launch (profiling only).
void launch_syrk_scaled(hipStream_t s, const SyrkDesc* d_desc, int ngroup,
                        int total_tiles, const double* xk, int ldxk,
                        double bytes, double flops);
int syrk_tile();

// ---------------------------------------------------------------------------
// batched dense symmetric eigensolver (k_syev.hip)
// ---------------------------------------------------------------------------
struct EigBatch {
  int nprob;
  int nmax;
  const int* n;       // [nprob]
  const long* aoff;   // [nprob] matrix offset, column-major, ld = n
  const long* voff;   // [nprob] offset of per-row vectors (sum of n)
  double* A;
  double* d;
  double* e;
  double* e2;
  double* tau;
  double* vbuf0;      // Householder vector ping-pong
  double* vbuf1;
  double* wbuf;
  double* pbuf;
  double* gl;         // [nprob] Gershgorin lower / upper bounds, norm, pivmin
  double* gu;
  double* tnrm;
  double* pivmin;
};
// Householder tridiagonalisation A = Q T Q^T (lower), column by column.
void launch_tridiag(hipStream_t s, const EigBatch& b, const int* h_n);
// Gershgorin bounds + e^2.
void launch_tridiag_bounds(hipStream_t s, const EigBatch& b);
// Eigenvalues k in [il_p, il_p + cnt_p) of each problem, by Sturm-count
// multisection (one wavefront per eigenvalue).  kofs = prefix of cnt.
void launch_bisect(hipStream_t s, const EigBatch& b, const int* il,
                   const int* kofs, int total, double* eval,
                   const long* eoff);
// Inverse iteration.  Cluster c covers eigenvalues [kb, ke) of problem p
// (local indices relative to il_p).  Z_p is n_p x m_p column-major.
struct Cluster {
  int p;
  int kb;
  int ke;
};
void launch_stein(hipStream_t s, const EigBatch& b, const Cluster* cl,
                  int ncl, const double* eval, const long* eoff,
                  double* Z, const long* zoff, double* scratch, int nthreads);
size_t stein_scratch_doubles(int nmax, int nthreads);
// Back-transformation Z <- Q Z with the stored reflectors (blocked WY).
void launch_backtransform(hipStream_t s, const EigBatch& b, double* Z,
                          const long* zoff, const int* m, const int* h_n,
                          const int* h_m, double* tbuf, const long* toff,
                          long tbuf_total);
int backtransform_nb();
// Single-work-group solver (k_syev_wg.hip): phase bit 0 = tridiagonalise and
// find the lowest cnt[p] eigenvalues; bit 1 = the lowest mvec[p] vectors into
// Z (from the reflectors, d, e and eigenvalues left by phase 0).
void launch_syev_wg(hipStream_t s, const EigBatch& b, const int* cnt,
                    const long* eoff, double* eval, const int* mvec,
                    const long* zoff, double* Z, double* scratch,
                    const long* soff, int phase);
size_t syev_wg_scratch_doubles(int n, int m);
// LDS-resident single-work-group solver for n <= 128 (k_syev_lds.hip).
// Problem p = src[slot] (or slot): A at A + p*astride (n x n, both
// triangles); lowest kcnt eigenvalues -> eval[p*estride + k]; lowest m
// eigenvectors -> Z[p*zstride + k*n + i], m = mfixed or, when mfixed < 0,
// the count of eigenvalues <= emax (also written to cnt[p]).
void launch_syev_lds(hipStream_t s, const double* A, long astride, int n, int kcnt,
                     int nsolve, const int* src, double* eval, long estride, double* Z,
                     long zstride, int* cnt, int mfixed, double emax, double* scratch);
size_t syev_lds_scratch_doubles(int n, int nsolve);
// Dense H1 of every solved (c,b) problem into H[p*n*n] (column-major).
void launch_build_h1(hipStream_t s, const double* vgrid, const double* xl,
                     const double* wsum, const double* wterm, int n, int nsolve,
                     const int* src, double* H);

}  // namespace dvr
