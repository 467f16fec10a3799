// Internal declarations shared by the HIP translation units of libdvr3d_amd.
// Kernels live in k_*.hip; each file exposes host-side launch_* wrappers.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>
#include <algorithm>
#include <map>
#include <tuple>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <stdexcept>
#include <vector>

namespace dvr {

// 1/x from the hardware reciprocal estimate and two Newton steps: within a
// few ulps for normal x, and about half the dependent instructions of the
// correctly rounded division (used on the critical path of recurrences)
__device__ inline double rcp_refined(double x) {
  double r = __builtin_amdgcn_rcp(x);
  r = fma(r, fma(-x, r, 1.0), r);
  return fma(r, fma(-x, r, 1.0), r);
}


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

// true for memory the device can read in place (hipMalloc'd); host memory,
// registered or not, is staged through device buffers by the callers
inline bool on_device(const void* p) {
  hipPointerAttribute_t a;
  if (hipPointerGetAttributes(&a, p) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return a.type == hipMemoryTypeDevice;
}

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
// with g = b*na + a and off = iv1off[c*(nb+1)+b]; hcb: the blocks on the host.
void launch_cint(hipStream_t s, const double* v1, const double* c2, double* cint,
                 const CBlock* cb, const std::vector<CBlock>& hcb, const int* iv1off, int na,
                 int nb);

struct DevBuf;
// Wavefunction amplitudes of nvec levels on the DVR grid, in the IWAVE
// record layout psi(idvr, npnt1, npnt2) per level.
// hcb: the blocks on the host; tmp: staging of the per-block products
// (G nc nvec doubles).
void launch_trans(hipStream_t s, const double* cint, const double* c3, double* psi,
                  const CBlock* cb, const std::vector<CBlock>& hcb, int nc, int na, int nb,
                  int n3, int nvec, int zr2r1, int idvr, int npnt1, int npnt2, DevBuf& tmp);

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
// After a lower-tile C = X^T X: C(i,j) *= xk(blk(i), blk(j)), + dg(i) on
// the diagonal, mirrored to the upper triangle (C: n x n, ld n).
void launch_scale_mirror(hipStream_t s, double* C, int n, const double* xk, int ldxk,
                         const int* blk, const double* dg);
// bytes/flops: algorithmic traffic and work of the launch (profiling only).
void launch_syrk_scaled(hipStream_t s, const SyrkDesc* d_desc, int ngroup,
                        int total_tiles, const double* xk, int ldxk,
                        double bytes, double flops);
int syrk_tile();
// same product with FP32 split (hi + lo) operands on FP32 MFMA
void launch_syrk_split32(hipStream_t s, const SyrkDesc* d_desc, int ngroup, int total_tiles,
                         const double* xk, int ldxk, double bytes, double flops);

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
  double* pan;        // panel buffer of the blocked reduction (3 tri_panel_width() x sum n)
};
// Householder tridiagonalisation A = Q T Q^T (lower): in panels for batches
// of at least 8 problems over 1024 rows with one over 2048 (needs b.pan;
// DVR_TRI=blocked: any batch with a problem over 512, DVR_TRI=column: never),
// column by column otherwise; h_aoff and h_voff are the host copies of aoff
// and voff.
void launch_tridiag(hipStream_t s, const EigBatch& b, const int* h_n, const long* h_aoff,
                    const long* h_voff);
bool tri_blocked_enabled(const int* hn, int np);
int tri_panel_width();
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
struct DevBuf;
void launch_backtransform(hipStream_t s, const EigBatch& b, double* Z,
                          const long* zoff, const int* m, const int* h_n,
                          const int* h_m, double* tbuf, const long* toff,
                          long tbuf_total, DevBuf& ofs);
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

// ---------------------------------------------------------------------------
// Device buffers and the batched dense eigensolver driver
// ---------------------------------------------------------------------------
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  void ensure(size_t b) {
    if (b <= bytes && p) return;
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    if (b == 0) b = 8;
    DVR_HIP(hipMalloc(&p, b));
    bytes = b;
  }
  template <class T>
  T* as() const { return static_cast<T*>(p); }
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
};

template <class T>
T* upload(DevBuf& d, const std::vector<T>& h, hipStream_t s) {
  d.ensure(sizeof(T) * std::max<size_t>(h.size(), 1));
  if (!h.empty())
    DVR_HIP(hipMemcpyAsync(d.p, h.data(), sizeof(T) * h.size(),
                           hipMemcpyHostToDevice, s));
  return d.as<T>();
}

// Working storage of one batched eigensolve.
struct EigWork {
  DevBuf n, aoff, voff, d, e, e2, tau, vb0, vb1, wb, pb, gl, gu, tn, piv;
  DevBuf il, kofs, eoff, eval;
  DevBuf zoff, m, clus, scr, tbuf, toff, btofs, pan;
  EigBatch b{};
  std::vector<int> hn;
  std::vector<long> hvoff;
  std::vector<int> hcnt;
  std::vector<long> heoff;
  std::vector<double> htn;
  std::vector<double> hval;  // concatenated eigenvalues (heoff offsets)

  bool wg = false;  // single-work-group solver (values and vectors split)
  // eigenvalues closer than cluster_rel * ||T|| share one inverse-iteration
  // thread and are Gram-Schmidt orthogonalised against each other
  double cluster_rel = 1.0e-6;
  DevBuf soffb;

  // Reduce A_p (at A + aoff[p], n_p x n_p) to tridiagonal form.  With the
  // single-work-group solver, or with reduce = false (d and e are then
  // filled by the caller, who calls bounds() afterwards), this only sets the
  // batch up.
  void tridiag(const std::vector<int>& n_, double* A,
               const std::vector<long>& aoff_, hipStream_t s, bool use_wg = false,
               bool reduce = true) {
    wg = use_wg;
    hn = n_;
    const int np = (int)hn.size();
    hvoff.assign(np, 0);
    long tot = 0;
    int nmax = 0;
    for (int p = 0; p < np; ++p) {
      hvoff[p] = tot;
      tot += hn[p];
      nmax = std::max(nmax, hn[p]);
    }
    const size_t vb = sizeof(double) * (tot + 1);
    for (DevBuf* x : {&d, &e, &e2, &tau, &vb0, &vb1, &wb, &pb}) x->ensure(vb);
    for (DevBuf* x : {&gl, &gu, &tn, &piv}) x->ensure(sizeof(double) * np);
    b.nprob = np;
    b.nmax = nmax;
    b.n = upload(n, hn, s);
    b.aoff = upload(aoff, aoff_, s);
    b.voff = upload(voff, hvoff, s);
    b.A = A;
    b.d = d.as<double>();
    b.e = e.as<double>();
    b.e2 = e2.as<double>();
    b.tau = tau.as<double>();
    b.vbuf0 = vb0.as<double>();
    b.vbuf1 = vb1.as<double>();
    b.wbuf = wb.as<double>();
    b.pbuf = pb.as<double>();
    b.gl = gl.as<double>();
    b.gu = gu.as<double>();
    b.tnrm = tn.as<double>();
    b.pivmin = piv.as<double>();
    if (!wg && reduce) {
      b.pan = nullptr;
      if (tri_blocked_enabled(hn.data(), np)) {
        pan.ensure(sizeof(double) * 3 * tri_panel_width() * (size_t)(tot + 1));
        b.pan = pan.as<double>();
      }
      launch_tridiag(s, b, hn.data(), aoff_.data(), hvoff.data());
      launch_tridiag_bounds(s, b);
    }
  }
  // Gershgorin bounds and e^2 of a tridiagonal filled by the caller
  void bounds(hipStream_t s) { launch_tridiag_bounds(s, b); }

  // Lowest cnt[p] eigenvalues of every problem -> hval (synchronises).
  void values(const std::vector<int>& cnt, hipStream_t s) {
    const int np = b.nprob;
    hcnt = cnt;
    std::vector<int> hil(np, 0), hk(np + 1, 0);
    heoff.assign(np, 0);
    long tot = 0;
    for (int p = 0; p < np; ++p) {
      hk[p + 1] = hk[p] + cnt[p];
      heoff[p] = tot;
      tot += cnt[p];
    }
    eval.ensure(sizeof(double) * (tot + 1));
    if (wg) {
      launch_syev_wg(s, b, upload(kofs, cnt, s), upload(eoff, heoff, s),
                     eval.as<double>(), nullptr, nullptr, nullptr, nullptr, nullptr, 1);
    } else {
      launch_bisect(s, b, upload(il, hil, s), upload(kofs, hk, s), hk[np],
                    eval.as<double>(), upload(eoff, heoff, s));
    }
    hval.assign(tot, 0.0);
    htn.assign(np, 0.0);
    if (tot)
      DVR_HIP(hipMemcpyAsync(hval.data(), eval.p, sizeof(double) * tot,
                             hipMemcpyDeviceToHost, s));
    DVR_HIP(hipMemcpyAsync(htn.data(), tn.p, sizeof(double) * np,
                           hipMemcpyDeviceToHost, s));
    DVR_HIP(hipStreamSynchronize(s));
  }

  // Number of vectors to compute so that the last wanted one does not split
  // a cluster of close eigenvalues.
  int extend(int p, int want) const {
    if (want <= 0) return 0;
    const double ctol = cluster_rel * std::max(htn[p], 1e-300);
    const double* v = hval.data() + heoff[p];
    int k = want;
    while (k < hcnt[p] && v[k] - v[k - 1] < ctol) ++k;
    return k;
  }

  // Eigenvectors 0..m[p]-1 of every problem into Z (n_p x m_p at zoff[p]).
  // back = false leaves the eigenvectors of the tridiagonal matrices (the
  // caller back-transforms them itself, as the two-stage reduction does).
  void vectors(const std::vector<int>& mv, double* Z,
               const std::vector<long>& zoff_, hipStream_t s, bool back = true) {
    const int np = b.nprob;
    if (wg) {
      std::vector<long> so(np, 0);
      long st = 0;
      for (int p = 0; p < np; ++p) {
        so[p] = st;
        st += (long)syev_wg_scratch_doubles(hn[p], mv[p]);
      }
      scr.ensure(sizeof(double) * (st + 1));
      launch_syev_wg(s, b, kofs.as<int>(), eoff.as<long>(), eval.as<double>(),
                     upload(m, mv, s), upload(zoff, zoff_, s), Z, scr.as<double>(),
                     upload(soffb, so, s), 2);
      DVR_HIP(hipStreamSynchronize(s));
      return;
    }
    std::vector<Cluster> cl;
    for (int p = 0; p < np; ++p) {
      const int mp = mv[p];
      if (mp <= 0) continue;
      const double ctol = cluster_rel * std::max(htn[p], 1e-300);
      const double* v = hval.data() + heoff[p];
      int kb = 0;
      for (int k = 1; k <= mp; ++k) {
        if (k == mp || v[k] - v[k - 1] >= ctol) {
          cl.push_back(Cluster{p, kb, k});
          kb = k;
        }
      }
    }
    const int ncl = (int)cl.size();
    if (ncl == 0) return;
    const int T = ncl;
    scr.ensure(sizeof(double) * stein_scratch_doubles(b.nmax, T));
    clus.ensure(sizeof(Cluster) * ncl);
    DVR_HIP(hipMemcpyAsync(clus.p, cl.data(), sizeof(Cluster) * ncl,
                           hipMemcpyHostToDevice, s));
    const long* dz = upload(zoff, zoff_, s);
    launch_stein(s, b, clus.as<Cluster>(), ncl, eval.as<double>(),
                 eoff.as<long>(), Z, dz, scr.as<double>(), T);
    if (!back) return;
    // WY factors
    const int NBk = backtransform_nb();
    std::vector<long> ht(np, 0);
    long tt = 0;
    for (int p = 0; p < np; ++p) {
      ht[p] = tt;
      const int nr = hn[p] - 2;
      tt += nr > 0 ? (long)((nr + NBk - 1) / NBk) * NBk * NBk : 0;
    }
    tbuf.ensure(sizeof(double) * (tt + 1));
    launch_backtransform(s, b, Z, dz, upload(m, mv, s), hn.data(), mv.data(),
                         tbuf.as<double>(), upload(toff, ht, s), tt, btofs);
    DVR_HIP(hipStreamSynchronize(s));
  }
};

// ---------------------------------------------------------------------------
// Two-stage tridiagonalisation (k_twostage.hip): band reduction on MFMA
// (bandwidth 64), then bulge chasing
// ---------------------------------------------------------------------------
struct TwoStage {
  DevBuf work, vyv, wb, qa, qb, G, dl, Linv, sc, M, Uinv, Rb, Zs, Mm, Sv, Tst, zero, flags, taul;
  DevBuf AB, vs, taus;
  // kept for the back-transformation when keep is set: every reflector of
  // the bulge chase (V2: BW doubles per (task, sweep), tau2), the reduced
  // matrix (its first-stage reflectors below the band) and the panel count
  bool keep = false;
  DevBuf V2, tau2, Zt, Vp, bw1, bw2;
  double* Ared = nullptr;
  long lda_red = 0;
  int n_red = 0, npanel = 0, jlast = 0, mlast = 0;
  // A (n x n, ld lda, both triangles) is destroyed; d[n], e[n-1] on the
  // device.  Returns 0, or 1 when a panel was rank deficient.
  int reduce(double* A, long lda, int n, double* d, double* e, hipStream_t s);
  // Z (n x nv, ld ldz): eigenvectors of the tridiagonal matrix of the last
  // reduce (with keep set) -> eigenvectors of its A, in place.
  void back(double* Z, long ldz, int nv, hipStream_t s);
};
int twostage_bandwidth();
// DVR_TWOSTAGE=0 switches the two-stage reduction off (one-stage instead)
inline bool env_twostage() {
  const char* v = getenv("DVR_TWOSTAGE");
  return !v || atoi(v) != 0;
}
// DVR_TWOSTAGE_VEC=0 keeps eigenvector solves on the one-stage reduction
inline bool env_twostage_vec() {
  const char* v = getenv("DVR_TWOSTAGE_VEC");
  return env_twostage() && (!v || atoi(v) != 0);
}
// Lowest `want` eigenvalues of the symmetric A (device, n x n, ld n, both
// triangles; left intact) through the two-stage reduction of a copy in
// `copy`, with the one-stage reduction of A as fallback; the values end in
// W.hval (W.heoff[0] = 0).  Returns true when the two-stage path was used.
bool twostage_values(const double* A, int n, int want, hipStream_t s, TwoStage& ts, EigWork& W,
                     DevBuf& copy);
// After twostage_values returned true with ts.keep set: eigenvectors 0..m-1
// of A into Z (n x m, ld n): inverse iteration on the tridiagonal matrix,
// then back through the bulge chase and the band reduction.
void twostage_vectors(int n, int m, double* Z, hipStream_t s, TwoStage& ts, EigWork& W);

// ---------------------------------------------------------------------------
// General FP64 GEMM on MFMA (k_gemm.hip)
// ---------------------------------------------------------------------------
struct GemmArgs {
  int M, N, K;
  const double* A;
  long lda;
  const double* B;
  long ldb;
  double* C;
  long ldc;
  double alpha, beta;
  int ksplit, kchunk, lower, ntile;
  double* work;
  // strided batch over blockIdx.z (without K split): product z uses
  // A + z sa, B + z sb, C + z sc
  long sa = 0, sb = 0, sc = 0;
  int nbatch = 1;
  int st = 16;   // super-tile width (tiles) of the persistent lower-tile update
};
// C = alpha op(A) op(B) + beta C (column-major).  ksplit > 1 splits K over
// work-groups; work then holds gemm_work_doubles(M, N, ksplit) doubles.
// lower: only 64x64 tiles on or below the diagonal of C are formed.
void launch_gemm(hipStream_t s, const char* name, bool ta, bool tb, int M, int N, int K,
                 double alpha, const double* A, long lda, const double* B, long ldb,
                 double beta, double* C, long ldc, int ksplit, double* work, bool lower);
size_t gemm_work_doubles(int M, int N, int ksplit);
// nbatch products of one shape at constant strides (no K split)
void launch_gemm_batched(hipStream_t s, const char* name, bool ta, bool tb, int M, int N, int K,
                         double alpha, const double* A, long lda, long sa, const double* B,
                         long ldb, long sb, double beta, double* C, long ldc, long sc,
                         int nbatch);
// Without K split, with the k-slice depth kt (16, 32), for beta != 0 C read
// at the start of each tile (cp) so that its latency hides behind the k loop,
// and for lower outputs optionally the XCD-aware tile order (xcd_lower:
// 8 x 8-tile super-tiles, each on one XCD, so that panels are read into each
// L2 once per super-tile).
void launch_gemm_var(hipStream_t s, const char* name, bool ta, bool tb, int M, int N, int K,
                     double alpha, const double* A, long lda, const double* B, long ldb,
                     double beta, double* C, long ldc, bool lower, int kt, bool cp,
                     bool xcd_lower = false);
// Lower-tile NT update C = alpha A B^T + beta C (M x M, lower 64 x 64 tiles)
// on the persistent wavefront-tile kernel with wgs_per_xcd work-groups per
// XCD (one wavefront per SIMD each); M < 64 or K other than 64, 128, 256 on
// the 64 x 64 kernel (launch name + "_small").  Same bits as launch_gemm_var
// with tb and lower.
void launch_gemm_lower_nt(hipStream_t s, const char* name, int M, int K, double alpha,
                          const double* A, long lda, const double* B, long ldb, double beta,
                          double* C, long ldc, int wgs_per_xcd, bool xcd_lower);
// split of K that gives about a thousand work-groups for small outputs
int gemm_auto_ksplit(int M, int N, int K);

// Per-shape choice of the GEMM k-slice depth (16 or 32) and K split, timed on
// first use (GemmTuner::run); enabled = false keeps the fixed rule.
struct GemmTuner {
  struct Key {
    bool ta, tb, lower;
    int M, N, K;
    bool operator<(const Key& o) const {
      return std::tie(ta, tb, lower, M, N, K) < std::tie(o.ta, o.tb, o.lower, o.M, o.N, o.K);
    }
  };
  struct Choice {
    int kt, ksplit;
  };
  std::map<Key, Choice> cache;
  bool enabled = true;
  DevBuf scratch, work;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  void run(hipStream_t s, const char* name, bool ta, bool tb, int M, int N, int K, double alpha,
           const double* A, long lda, const double* B, long ldb, double beta, double* C, long ldc,
           bool lower);
  ~GemmTuner();
};

// ---------------------------------------------------------------------------
// LDL^T of H - sigma I without pivoting (k_ldl.hip)
// ---------------------------------------------------------------------------
struct LdlFactor {
  int sb = 2048;  // super-block of the triangular solves (set before ensure)
  int n = 0, nsb = 0;
  DevBuf A, d, linv64, linvsb, w, work;
  std::vector<double> hd;
  void ensure(int n);
  // factors H - sigma I (H: n x n full square, ld n); returns the inertia
  // count (number of negative pivots = eigenvalues of H below sigma)
  // (-1 when a pivot is not finite or below n eps max|d|: the count is then
  // not trustworthy)
  int factor(const double* H, double sigma, hipStream_t s);
  // the same in two halves: the launches (the host does not wait), then the
  // wait for s and the inertia count
  void factor_begin(const double* H, double sigma, hipStream_t s);
  int factor_end(hipStream_t s);
  // T <- (H - sigma I)^-1 R for nrhs columns; R is destroyed
  void solve(double* R, double* T, int nrhs, hipStream_t s);
  // the same through a captured HIP graph per nrhs (the sequence of ~160
  // launches is replayed; recaptured when any buffer it uses moved)
  void solve_graph(double* R, double* T, int nrhs, hipStream_t s);
  bool use_graph = true;
  int panel_sub = 4;            // sub-panels of 64 columns per panel (trailing K = 64 panel_sub)
  bool lookahead = true;        // trailing updates overlapped with the next panel
  int upd_kt = 16;              // trailing-update GEMM: k-slice depth, C read at the start,
  bool upd_cp = true;           // XCD-aware tile order
  bool upd_xcd = true;
  int upd_wide = 32;            // trailing update on the persistent wavefront-tile kernel with
                                // this many work-groups per XCD; 0: the 64 x 64 kernel
  GemmTuner* tuner = nullptr;   // per-shape GEMM choices (nullptr: fixed rule)
  hipStream_t aux = nullptr;    // trailing updates overlapped with the next panel
  hipEvent_t ev_panel = nullptr, ev_trail = nullptr;
  ~LdlFactor();

 private:
  struct Captured {
    hipGraphExec_t exec = nullptr;
    const void* key[6] = {};
  };
  std::vector<std::pair<int, Captured>> graphs;
  int graph_sb = 0;   // super-block size the graphs were captured with
  void clear_graphs();
};

// ---------------------------------------------------------------------------
// Block-Davidson final solve (dvr_davidson.hip, kernels in k_dav.hip)
// ---------------------------------------------------------------------------
void launch_gather_cols(hipStream_t s, const double* src, long lds, const int* idx, int ncol,
                        double* dst, long ldd, int nrows);
void launch_resid_norms(hipStream_t s, const double* X, const double* AX, long ld,
                        const double* theta, int m, int nrows, double* norms);
void launch_resid_gather(hipStream_t s, const double* X, const double* AX, long ld,
                         const double* theta, const int* idx, int ncol, double* R, long ldr,
                         int nrows);
void launch_symmetrize(hipStream_t s, double* T, int m);
// T(c, r) <- T(r, c) for r > c: the upper triangle from the lower one
void launch_mirror_lower(hipStream_t s, double* T, int m);
void launch_gram_prep(hipStream_t s, double* G, int q, double eps, double* sc);
void launch_form_m(hipStream_t s, const double* Linv, int q, const double* r, const double* sc,
                   double* M);
// LDL^T of one nb x nb block (nb <= 64) in place: D to d, L^-1 to Linv (ld nb)
void launch_ldl_block(hipStream_t s, double* A, long lda, int nb, double* d, double* Linv);
void launch_scale_cols(hipStream_t s, double* U, long ld, int m, int ncol, const double* scal);

struct DavParams {
  int nev = 0;          // wanted levels
  int prec = 0;         // 0 shift-and-invert, 1 diagonal, 2 learned diagonal
  const float* mlp = nullptr;     // learned preconditioner parameters (device)
  const double* st = nullptr;     // static features of the basis (device, 3 x N)
  int block = 64;       // correction block
  int slice = 256;      // levels locked per certified slice
  int mmax = 448;       // largest search space
  int keep = 320;       // Ritz vectors kept when the space is full (0: mmax - block)
  double tol = 1e-7;    // residual norm (Eh) for locking
  int max_iter = 2000;
};
struct DavStats {
  int iterations = 0, matvecs = 0, solves = 0, factorizations = 0, slices = 0;
  int certified = 0, missing = 0;
  // device time per phase (HIP events on the solver's stream, ms):
  // Rayleigh-Ritz, shift-and-invert solves, basis expansion, factorisations
  // and the start (diagonal, start block, first projection)
  double phase_ms[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
};
// HIP timing events reused across solves (created on demand)
struct PhaseClock {
  std::vector<hipEvent_t> ev;
  std::vector<int> lab;
  int used = 0;
  PhaseClock() = default;
  PhaseClock(const PhaseClock&) = delete;
  PhaseClock& operator=(const PhaseClock&) = delete;
  ~PhaseClock() {
    for (hipEvent_t e : ev) (void)hipEventDestroy(e);
  }
  void mark(int label, hipStream_t s);
  // adds the time from each mark to the next one to out[label]
  void collect(double* out);
};
class DavidsonSolver {
 public:
  // 0: certified; 2: stalled; 3: a level below the last shift was missed;
  // 5: too small a problem for the iteration (n < nev + 1 + 3 block)
  int solve(const double* H, int n, const DavParams& P, hipStream_t s);
  int solve_impl(const double* H, int n, const DavParams& P, hipStream_t s);
  // the lowest nvec eigenvectors (ascending) into Z (n x nvec)
  void vectors(int nvec, double* Z, hipStream_t s);
  std::vector<double> eval;
  DavStats stats;
  double sigma = 0.0;
  GemmTuner tuner;      // per-shape GEMM choices shared with the factorisations
  bool rr_single_wg = false;   // Rayleigh-Ritz through the one-work-group solver
  bool use_graph = true;       // shift-and-invert solves replayed from HIP graphs
  int ldl_sub = 4;             // LDL^T panel width in 64-column sub-panels
  bool ldl_xcd = true;         // XCD-aware tile order of the trailing updates
  int ldl_sb = 2048;           // super-block of the shift-and-invert triangular solves
  bool ldl_cp = true;          // trailing updates read C before the k loop
  int ldl_wide = 32;           // trailing updates on the persistent kernel (work-groups per XCD; 0 off)
  // residual norm below which the Ritz pairs up to one past the top of a
  // slice fix the next shift: its factorisation then starts on a second
  // stream while the slice converges (0: after the slice, on the same stream)
  double early_tol = 1e-4;
  // iterations with diagonal corrections while the first factorisation runs
  // on the second stream (needs early_tol > 0; 0: factored before them)
  int first_overlap = 0;
  // k > 0: the k-th inertia count of a solve is reported ambiguous (tests of
  // the retries at nudged shifts; 0 in use)
  int force_ambiguous = 0;
  ~DavidsonSolver();

 private:
  void gemm(hipStream_t s, const char* name, bool ta, bool tb, int M, int N, int K, double alpha,
            const double* A, long lda, const double* B, long ldb, double beta, double* C, long ldc,
            bool lower = false);
  void small_eig(int m, double* S, double* U, std::vector<double>& w, hipStream_t s);
  int expand(const double* H, int q0, hipStream_t s);
  int expand_chunk(const double* H, double* W, int q0, hipStream_t s);
  int cholqr(const double* Z, int q, double* tmp, double* out, hipStream_t s,
             const double* raw2 = nullptr);
  int N = 0, m = 0, nlocked = 0;
  DevBuf V, AV, X, AX, XL, Rb, Tb, Tm, Y, Y2, G, G2, U, C1, thd, nrmd, idxd, scal, work;
  DevBuf gd, gsc, gr, gm, Dd, Fd, Wd, tha, rawd;
  EigWork rr;
  LdlFactor ldl, ldl2;         // current shift; the next one (early factorisation)
  hipStream_t s2 = nullptr;    // stream of the early factorisation
  hipEvent_t ev_s = nullptr;
  bool first_pending = false;  // the first factorisation still running on s2
  PhaseClock clock;
  std::vector<double> thL, gw;
  std::vector<int> perm;
};

// ---------------------------------------------------------------------------
// Learned preconditioner (k_mlp.hip, dvr_mlp.hip)
// ---------------------------------------------------------------------------
#define MLP_NF 6
#define MLP_WIDTH 128
#define MLP_NPARAM (MLP_NF * MLP_WIDTH + MLP_WIDTH + MLP_WIDTH * MLP_WIDTH + MLP_WIDTH + MLP_WIDTH + 1)
#define MLP_DELTA 1.0e-3   // energy scale of the features (Eh)
#define MLP_DMIN 1.0e-4    // smallest |D_j - theta| of the diagonal correction (Eh)
void launch_mlp_fwd(hipStream_t s, const float* P, long R, const float* F, float* W, float* H1,
                    float* H2);
void launch_mlp_infer_bf16(hipStream_t s, const float* P, long R, const float* F, float* W);
void launch_mlp_bwd(hipStream_t s, const float* P, long R, const float* F, const float* H1,
                    const float* H2, const float* GW, float* GA1, float* GA2, float* grad,
                    float* partial, long partial_floats);
void launch_mlp_bias_grads(hipStream_t s, long R, const float* ones, const float* GA1,
                           const float* GA2, const float* GW, float* grad, float* partial,
                           long partial_floats);
void launch_adam(hipStream_t s, float* P, const float* g, float* m, float* v, int n, float lr,
                 int step);
void launch_mlp_features(hipStream_t s, const double* D, const double* st, int N, int S,
                         const double* theta, const double* sigma, float* F);
void launch_mlp_apply(hipStream_t s, const float* w, const double* D, const double* theta,
                      const double* R, int N, int S, double* T);
void launch_mlp_gw(hipStream_t s, const double* gt, const double* D, const double* theta,
                   const double* R, int N, int S, double scale, float* gw);
void launch_mlp_loss(hipStream_t s, const double* T, const double* X, const double* E, int N,
                     int S, double* G, double* loss);
void launch_coldot(hipStream_t s, const double* X, const double* Y, int N, int S, double* out);
void launch_col_axpy(hipStream_t s, const double* X, const double* c, double alpha, int N, int S,
                     double* Y);

struct MlpWork {
  DevBuf X, AX, Rr, E, T, G, Cm, th, c, loss, F, W, GW, H1, H2, GA1, GA2, partial, grad, m, v,
      ones, kidx;
};
// Trains the parameters P (device, MLP_NPARAM floats) in place on the
// eigenpairs (lam, Z: N x nz) of H; D = diag H, st = static features
// (e2, r1 fraction, 2D index fraction; 3 x N).  hist: mean loss per step.
int mlp_train(hipStream_t s, const double* H, int N, const double* Z, int nz,
              const std::vector<double>& lam, const double* D, const double* st, float* P,
              int steps, int S, unsigned long long seed, float lr, std::vector<double>& hist,
              MlpWork& w);

// ---------------------------------------------------------------------------
// FP64 refinement with the exact operator (dvr_refine.hip)
// ---------------------------------------------------------------------------
void launch_add_diag(hipStream_t s, double* Y, const double* X, const double* e, int N, int nv);
// out[i] = H[i, i]
void launch_get_diag(hipStream_t s, const double* H, int N, double* out);
// column c of V (N x b, ld N) <- e_{idx[c]}
void launch_unit_cols(hipStream_t s, double* V, int N, const int* idx, int b);
struct RefineWork {
  DevBuf U, V, Y, T, Z, X2, Y2, th, nrm, work;
  EigWork eig;
};
// Y = H3 X from the contraction coefficients (no H3 needed)
void apply_h_exact(hipStream_t s, const double* cint, long G, const std::vector<int>& col0,
                   const std::vector<int>& iv2, const double* xk, const double* e2, int N3,
                   const double* X, int nv, double* Y, RefineWork& w);
// Rayleigh-Ritz of span(X) (N3 x nv, orthonormal) with the exact operator: X
// is replaced by the refined vectors, eval by the refined values; returns the
// largest exact residual norm.
double refine_pairs(hipStream_t s, const double* cint, long G, const std::vector<int>& col0,
                    const std::vector<int>& iv2, const double* xk, const double* e2, int N3,
                    double* X, int nv, std::vector<double>& eval, RefineWork& w);
// Rayleigh quotient of every column of X (orthonormal) with the exact
// operator (all levels wanted); returns the largest residual norm.
double refine_rq(hipStream_t s, const double* cint, long G, const std::vector<int>& col0,
                 const std::vector<int>& iv2, const double* xk, const double* e2, int N3,
                 const double* X, int nv, std::vector<double>& eval, RefineWork& w);

}  // namespace dvr
