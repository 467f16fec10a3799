// C ABI and stage orchestration of the theta-first DVR3DRJZ path on one GPU.
//
// The stage sequence is that of jhmain (source/dvr3drjz_segmented.f90:
// 1915-2118); every intermediate stays in device memory instead of the
// reference's scratch units (IVECS1/IEIGS1/INTVEC/IVECS2/IVINT/IBAND).  The
// host only sees the small integer truncation data (iv1, iv2), which it needs
// to size the following launches, and the 2D eigenvalues for the global
// selection (choose, D:2767-2867).
#include "../../include/dvr3d_amd.h"
#include "dvr_internal.h"

#include <algorithm>
#include <numeric>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

using namespace dvr;

namespace dvr {
thread_local Profiler* g_prof = nullptr;
}

namespace {


}  // namespace

struct dvr_ctx {
  dvr_dims_t dims{};
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  bool basis_set = false, pot_set = false, ran = false;
  double emax1 = 0.0, emax2 = 0.0;
  DevBuf xl, xkb, xkc, wsum, wterm, rc, rb, xa, v;
  DevBuf e1full, v1full, cnt, vscr, sweeps, mflag, uniq, h1, lscr;
  bool mirror_used = false, eig1d_householder = false;
  DevBuf srcrow, rowb, v1, e1p, iv1off;
  DevBuf H2, sdesc2, Z2;
  DevBuf cint, blk3, e2sel, cblk, sdesc3, H3, Z3, psi;
  EigWork w2, w3;
  TwoStage ts;
  DevBuf H3b;                 // copy of H3 reduced by the two-stage path
  bool twostage_used = false;
  DavidsonSolver dav;
  int dav_status = -1;
  std::string dav_error;   // message of an exception inside the Davidson solver
  bool keep_h = false, h_kept = false, split32 = false;
  double refine_resid = 0.0;
  DevBuf Zr, permd;
  RefineWork rw;
  DevBuf H3keep, st3, mlpP;
  DevBuf ptmp;                // per-block wavefunction products before the scatter
  int prec_kind = 0;
  bool mlp_set = false;
  MlpWork mw;
  bool has_wterm = false;
  std::vector<int> h_iv1;    // [c*npntb + b]
  std::vector<int> h_nham2, h_iv2;
  std::vector<double> h_e2all;
  int nham3 = 0, meval = 0, nvec_done = 0;
  std::vector<double> h_eval;
  double stage_ms[DVR_NSTAGE] = {0};
  hipEvent_t ev[DVR_NSTAGE + 1] = {};
  bool profiling = false;
  std::unique_ptr<Profiler> prof;
  struct KStat {
    std::string name;
    int launches;
    double seconds, bytes, flops;
  };
  std::vector<KStat> kstats;
};

namespace {

int fail(dvr_ctx_t* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

int env_int(const char* name, int def) {
  const char* v = getenv(name);
  return v ? atoi(v) : def;
}

struct DeviceGuard {
  int prev = -1;
  explicit DeviceGuard(int dev) {
    if (hipGetDevice(&prev) != hipSuccess) prev = -1;
    (void)hipSetDevice(dev);
  }
  ~DeviceGuard() {
    if (prev >= 0) (void)hipSetDevice(prev);
  }
};

void run_pipeline(dvr_ctx_t* C) {
  const dvr_dims_t& D = C->dims;
  hipStream_t s = C->stream;
  const int na = D.npnta, nb = D.npntb, nc = D.npntc;
  const int nprob = nb * nc;
  const double emax1 = C->emax1;
  for (int i = 0; i < DVR_NSTAGE; ++i) C->stage_ms[i] = 0.0;
  DVR_HIP(hipEventRecord(C->ev[0], s));

  // ---- 1D: mkham1 + diag + cut1d for every (c, b)
  prof_phase("1d");
  C->e1full.ensure(sizeof(double) * (size_t)nprob * na);
  C->v1full.ensure(sizeof(double) * (size_t)nprob * na * na);
  C->cnt.ensure(sizeof(int) * nprob);
  C->sweeps.ensure(sizeof(int) * nprob);
  // problems (c,b) and (b,c) with identical inputs are solved once
  bool mirror = false;
  if (nb == nc && nb > 1) {
    C->mflag.ensure(sizeof(int));
    launch_mirror_check(s, C->v.as<double>(), C->wsum.as<double>(),
                        C->has_wterm ? C->wterm.as<double>() : nullptr, nb, na,
                        C->mflag.as<int>());
    int flag = 1;
    DVR_HIP(hipMemcpyAsync(&flag, C->mflag.p, sizeof(int), hipMemcpyDeviceToHost, s));
    DVR_HIP(hipStreamSynchronize(s));
    mirror = flag == 0;
  }
  C->mirror_used = mirror;
  std::vector<int> uniq;
  if (mirror) {
    for (int c = 0; c < nc; ++c)
      for (int b = 0; b <= c; ++b) uniq.push_back(c * nb + b);
  }
  const int nsolve = mirror ? (int)uniq.size() : nprob;
  const int* d_src = mirror ? upload(C->uniq, uniq, s) : nullptr;
  // 1D eigensolver: LDS Householder (default for npnta <= 128) or the
  // LDS Jacobi sweeps; DVR_EIG1D=jacobi|householder overrides
  const char* e1env = getenv("DVR_EIG1D");
  const bool hh = e1env ? std::string(e1env) == "householder" : na <= 128;
  C->eig1d_householder = hh;
  if (hh) {
    C->h1.ensure(sizeof(double) * (size_t)nprob * na * na);
    launch_build_h1(s, C->v.as<double>(), C->xl.as<double>(), C->wsum.as<double>(),
                    C->has_wterm ? C->wterm.as<double>() : nullptr, na, nsolve, d_src,
                    C->h1.as<double>());
    C->lscr.ensure(sizeof(double) * syev_lds_scratch_doubles(na, nsolve));
    launch_syev_lds(s, C->h1.as<double>(), (long)na * na, na, na, nsolve, d_src,
                    C->e1full.as<double>(), na, C->v1full.as<double>(), (long)na * na,
                    C->cnt.as<int>(), D.zall ? na : -1, emax1, C->lscr.as<double>());
  } else {
    const size_t vs = jacobi1d_scratch_doubles(na, nsolve);
    if (vs) C->vscr.ensure(sizeof(double) * vs);
    launch_jacobi1d(s, C->v.as<double>(), C->xl.as<double>(),
                    C->wsum.as<double>(),
                    C->has_wterm ? C->wterm.as<double>() : nullptr, na, nsolve,
                    emax1, C->e1full.as<double>(), C->v1full.as<double>(),
                    C->cnt.as<int>(), vs ? C->vscr.as<double>() : nullptr,
                    d_src, C->sweeps.as<int>());
  }
  if (mirror)
    launch_mirror_copy(s, nb, na, C->e1full.as<double>(), C->v1full.as<double>(),
                       C->cnt.as<int>());
  std::vector<int> cnt(nprob);
  DVR_HIP(hipMemcpyAsync(cnt.data(), C->cnt.p, sizeof(int) * nprob,
                         hipMemcpyDeviceToHost, s));
  DVR_HIP(hipEventRecord(C->ev[1], s));
  DVR_HIP(hipStreamSynchronize(s));
  if (D.zall)
    for (int& x : cnt) x = na;
  C->h_iv1 = cnt;

  // ---- sizes of the 2D blocks (D:2021-2025), capacity check (D:2925)
  std::vector<int> nham2(nc, 0), rowbase(nc, 0), iv1off(nc * (nb + 1), 0);
  int rows = 0;
  for (int c = 0; c < nc; ++c) {
    rowbase[c] = rows;
    int s2 = 0;
    for (int b = 0; b < nb; ++b) {
      iv1off[c * (nb + 1) + b] = s2;
      s2 += cnt[c * nb + b];
    }
    iv1off[c * (nb + 1) + nb] = s2;
    nham2[c] = s2;
    if (s2 > D.max2d && !D.zall)
      throw std::runtime_error("core2d");
    rows += s2;
  }
  C->h_nham2 = nham2;
  std::vector<int> srcrow(rows), rowb(rows);
  for (int c = 0; c < nc; ++c)
    for (int b = 0; b < nb; ++b)
      for (int k = 0; k < cnt[c * nb + b]; ++k) {
        const int r = rowbase[c] + iv1off[c * (nb + 1) + b] + k;
        srcrow[r] = (c * nb + b) * na + k;
        rowb[r] = b;
      }
  C->v1.ensure(sizeof(double) * ((size_t)rows * na + 1));
  C->e1p.ensure(sizeof(double) * (rows + 1));
  launch_pack_1d(s, C->v1full.as<double>(), C->e1full.as<double>(),
                 upload(C->srcrow, srcrow, s), rows, na, C->v1.as<double>(),
                 C->e1p.as<double>());
  const int* d_rowb = upload(C->rowb, rowb, s);
  const int* d_iv1off = upload(C->iv1off, iv1off, s);

  // ---- mkham2: grouped SYRK over the c blocks
  prof_phase("2d");
  std::vector<int> pc;  // c of each nonempty block
  std::vector<int> pn;
  std::vector<long> h2off;
  long h2tot = 0;
  for (int c = 0; c < nc; ++c)
    if (nham2[c] > 0) {
      pc.push_back(c);
      pn.push_back(nham2[c]);
      h2off.push_back(h2tot);
      h2tot += (long)nham2[c] * nham2[c];
    }
  const int np2 = (int)pc.size();
  if (np2 == 0) throw std::runtime_error("empty 2D basis: no 1D level below EMAX1");
  C->H2.ensure(sizeof(double) * (h2tot + 1));
  {
    const int TM = syrk_tile();
    std::vector<SyrkDesc> desc(np2);
    int tiles = 0;
    for (int q = 0; q < np2; ++q) {
      const int c = pc[q];
      const int nt = (pn[q] + TM - 1) / TM;
      SyrkDesc& x = desc[q];
      x.X = C->v1.as<double>() + (size_t)rowbase[c] * na;
      x.ldx = na;
      x.K = na;
      x.N = pn[q];
      x.C = C->H2.as<double>() + h2off[q];
      x.blk = d_rowb + rowbase[c];
      x.diag = C->e1p.as<double>() + rowbase[c];
      x.tile0 = tiles;
      x.ntile = nt * (nt + 1) / 2;
      tiles += x.ntile;
    }
    double by = 0.0, fl = 0.0;
    for (int q = 0; q < np2; ++q) {
      by += 8.0 * ((double)na * pn[q] + (double)pn[q] * pn[q]);
      fl += (double)na * pn[q] * (pn[q] + 1.0);
    }
    launch_syrk_scaled(s, upload(C->sdesc2, desc, s), np2, tiles,
                       C->xkb.as<double>(), nb, by, fl);
  }
  DVR_HIP(hipEventRecord(C->ev[2], s));

  // ---- 2D eigenvalues (diag at D:2030), all of them
  int nmax2 = 0;
  for (int x : pn) nmax2 = std::max(nmax2, x);
  const char* e2env = getenv("DVR_EIG2D");
  // multi-launch by default: measured faster for the 2D blocks (the single
  // work-group solver is latency bound at one work-group per CU)
  const bool wg2 = e2env ? (std::string(e2env) == "wg") : false;
  C->w2.tridiag(pn, C->H2.as<double>(), h2off, s, wg2);
  C->w2.values(pn, s);
  DVR_HIP(hipEventRecord(C->ev[3], s));
  {
    std::vector<double> all;
    for (int q = 0; q < np2; ++q)
      all.insert(all.end(), C->w2.hval.begin() + C->w2.heoff[q],
                 C->w2.hval.begin() + C->w2.heoff[q] + pn[q]);
    C->h_e2all = all;
  }

  // ---- selection of the final basis: choose (D:2767-2867) or cut2d (D:2941)
  std::vector<int> iv2(nc, 0);
  int n3 = 0;
  if (D.zcut || D.zall) {
    const double emax2 = C->emax2;
    for (int q = 0; q < np2; ++q) {
      const double* e = C->w2.hval.data() + C->w2.heoff[q];
      int k = 0;
      if (D.zall)
        k = pn[q];
      else
        while (k < pn[q] && !(e[k] > emax2)) ++k;
      iv2[pc[q]] = k;
      n3 += k;
    }
    if (n3 > D.max3d && !D.zall) throw std::runtime_error("core3d");
  } else {
    long tot = 0;
    for (int q = 0; q < np2; ++q) tot += pn[q];
    const int low3d = (int)std::min<long>(D.max3d, tot);
    std::vector<std::tuple<double, int, int>> keys;
    keys.reserve(tot);
    for (int q = 0; q < np2; ++q) {
      const double* e = C->w2.hval.data() + C->w2.heoff[q];
      for (int k = 0; k < pn[q]; ++k) keys.emplace_back(e[k], pc[q], k);
    }
    // lowest energies first; on equal energies the lower c block first and
    // within a block the lower index, as the strict '<' merge of D:2808-2813
    std::partial_sort(keys.begin(), keys.begin() + low3d, keys.end());
    for (int i = 0; i < low3d; ++i) iv2[std::get<1>(keys[i])]++;
    n3 = low3d;
  }
  C->h_iv2 = iv2;
  C->nham3 = n3;
  if (n3 <= 0) throw std::runtime_error("empty final basis");

  // ---- chosen 2D eigenvectors
  std::vector<int> m2(np2), m2ext(np2);
  std::vector<long> z2off(np2);
  long z2tot = 0;
  for (int q = 0; q < np2; ++q) {
    m2[q] = iv2[pc[q]];
    m2ext[q] = C->w2.extend(q, m2[q]);
    z2off[q] = z2tot;
    z2tot += (long)pn[q] * m2ext[q];
  }
  C->Z2.ensure(sizeof(double) * (z2tot + 1));
  C->w2.vectors(m2ext, C->Z2.as<double>(), z2off, s);
  DVR_HIP(hipEventRecord(C->ev[4], s));

  // ---- 3D: cint expansion + mkham3 SYRK + loadh diagonal
  prof_phase("3d");
  std::vector<CBlock> cb(nc);
  std::vector<int> blk3(n3);
  std::vector<double> e2sel(n3);
  {
    int col = 0, q = 0;
    for (int c = 0; c < nc; ++c) {
      cb[c].rowbase = rowbase[c];
      cb[c].nham2 = nham2[c];
      cb[c].iv2 = iv2[c];
      cb[c].col0 = col;
      cb[c].z2off = 0;
      if (nham2[c] > 0) {
        cb[c].z2off = z2off[q];
        const double* e = C->w2.hval.data() + C->w2.heoff[q];
        for (int j = 0; j < iv2[c]; ++j) {
          blk3[col + j] = c;
          e2sel[col + j] = e[j];
        }
        ++q;
      }
      col += iv2[c];
    }
  }
  {
    // static features of the learned preconditioner: parent 2D eigenvalue,
    // r1 point, index of the function inside its r1 block (fractions)
    int nmax2 = 1;
    for (int c = 0; c < nc; ++c) nmax2 = std::max(nmax2, iv2[c]);
    std::vector<double> st(3 * (size_t)n3);
    int col = 0;
    for (int c = 0; c < nc; ++c)
      for (int j = 0; j < iv2[c]; ++j, ++col) {
        st[col] = e2sel[col];
        st[n3 + col] = nc > 1 ? (double)c / (nc - 1) : 0.0;
        st[2 * (size_t)n3 + col] = nmax2 > 1 ? (double)j / (nmax2 - 1) : 0.0;
      }
    upload(C->st3, st, s);
  }
  const long G = (long)na * nb;
  C->cint.ensure(sizeof(double) * (size_t)G * n3);
  const CBlock* d_cb = upload(C->cblk, cb, s);
  const int* d_blk3 = upload(C->blk3, blk3, s);
  launch_cint(s, C->v1.as<double>(), C->Z2.as<double>(), C->cint.as<double>(), d_cb, cb,
              d_iv1off, na, nb);
  C->H3.ensure(sizeof(double) * (size_t)n3 * n3);
  {
    const int TM = syrk_tile();
    const int nt = (n3 + TM - 1) / TM;
    std::vector<SyrkDesc> desc(1);
    desc[0].X = C->cint.as<double>();
    desc[0].ldx = G;
    desc[0].K = (int)G;
    desc[0].N = n3;
    desc[0].C = C->H3.as<double>();
    desc[0].blk = d_blk3;
    desc[0].diag = upload(C->e2sel, e2sel, s);
    desc[0].tile0 = 0;
    desc[0].ntile = nt * (nt + 1) / 2;
    // split-precision FP32 assembly when DVR_ASSEMBLY=split32, with FP64
    // refinement after the solve: a Rayleigh-Ritz step with the exact
    // operator after the Davidson solve, a Rayleigh quotient per level after
    // the dense one
    const char* asmenv = getenv("DVR_ASSEMBLY");
    C->split32 = asmenv && std::string(asmenv) == "split32";
    if (C->split32)
      launch_syrk_split32(s, upload(C->sdesc3, desc, s), 1, desc[0].ntile, C->xkc.as<double>(),
                          nc, 8.0 * ((double)G * n3 + (double)n3 * n3),
                          (double)G * n3 * (n3 + 1.0));
    else if (env_int("DVR_SYRK3_GEMM", 1)) {
      // the lower tiles of cint^T cint on the tuned GEMM (XCD-aware tile
      // order), then the block scale, the diagonal and the mirror
      launch_gemm_var(s, "syrk", true, false, n3, n3, (int)G, 1.0, C->cint.as<double>(), G,
                      C->cint.as<double>(), G, 0.0, C->H3.as<double>(), n3, true,
                      env_int("DVR_SYRK3_KT", 16) == 32 ? 32 : 16, false, true);
      launch_scale_mirror(s, C->H3.as<double>(), n3, C->xkc.as<double>(), nc, d_blk3,
                          desc[0].diag);
    } else
      launch_syrk_scaled(s, upload(C->sdesc3, desc, s), 1, desc[0].ntile,
                         C->xkc.as<double>(), nc,
                         8.0 * ((double)G * n3 + (double)n3 * n3),
                         (double)G * n3 * (n3 + 1.0));
  }
  DVR_HIP(hipEventRecord(C->ev[5], s));

  // ---- final solve (diag3d, D:2661-2764)
  const int meval = std::min(D.neval, n3);
  C->meval = meval;
  const int nvec = std::min(D.nvec, meval);
  C->nvec_done = nvec;
  C->dav_status = -1;
  bool dense = D.solver != DVR_SOLVER_DAVIDSON;
  if (!dense) {
    // block-Davidson with certified slices; the dense path is the fallback
    DavParams P;
    P.nev = meval;
    P.block = env_int("DVR_DAV_BLOCK", P.block);
    C->dav.rr_single_wg = env_int("DVR_DAV_RRWG", 0) != 0;
    C->dav.use_graph = env_int("DVR_GRAPH", 1) != 0;
    C->dav.ldl_sub = env_int("DVR_LDL_SUB", 4);
    C->dav.ldl_xcd = env_int("DVR_LDL_XCD", 1) != 0;
    C->dav.ldl_sb = std::max(64, env_int("DVR_LDL_SB", 2048) / 64 * 64);
    C->dav.ldl_cp = env_int("DVR_LDL_CP", 1) != 0;
    C->dav.ldl_wide = env_int("DVR_LDL_WIDE", 32);
    if (const char* t = getenv("DVR_DAV_EARLY")) C->dav.early_tol = atof(t);
    C->dav.force_ambiguous = std::max(0, env_int("DVR_DAV_FORCE_AMBIG", 0));
    // kind 3 needs iterations beside the first factorisation to be used in
    C->dav.first_overlap =
        std::max(0, env_int("DVR_DAV_FIRST", C->prec_kind == DVR_PREC_SI_LEARNED ? 6 : 0));
    P.slice = env_int("DVR_DAV_SLICE", P.slice);
    P.mmax = env_int("DVR_DAV_MMAX", P.mmax);
    P.keep = env_int("DVR_DAV_KEEP", P.keep);
    P.max_iter = env_int("DVR_DAV_MAXIT", 2000);
    P.prec = C->prec_kind;
    if (P.prec == DVR_PREC_LEARNED || P.prec == DVR_PREC_SI_LEARNED) {
      if (!C->mlp_set) throw std::runtime_error("learned preconditioner selected without parameters");
      P.mlp = C->mlpP.as<float>();
      P.st = C->st3.as<double>();
    }
    if (const char* t = getenv("DVR_DAV_TOL")) P.tol = atof(t);
    C->dav.tuner.enabled = env_int("DVR_TUNE", 1) != 0;
    C->dav_error.clear();
    try {
      C->dav_status = C->dav.solve(C->H3.as<double>(), n3, P, s);
    } catch (const std::exception& e) {
      // an error inside the iterative solver hands the problem to the dense
      // one (H3 is intact: the solver works on copies); a device fault
      // resurfaces there
      C->dav_status = 4;
      C->dav_error = e.what();
    }
    if (C->dav_status == 0) {
      C->h_eval.assign(C->dav.eval.begin(), C->dav.eval.begin() + meval);
      if (C->split32) {
        // the levels of the split-precision H3, refined with the exact FP64
        // operator: Rayleigh-Ritz on the span of the meval Davidson vectors
        prof_phase("refine");
        C->Zr.ensure(sizeof(double) * (size_t)n3 * meval);
        C->dav.vectors(meval, C->Zr.as<double>(), s);
        std::vector<int> col0(nc), iv2v(nc);
        for (int c = 0; c < nc; ++c) {
          col0[c] = cb[c].col0;
          iv2v[c] = cb[c].iv2;
        }
        std::vector<double> ev;
        C->refine_resid = refine_pairs(s, C->cint.as<double>(), G, col0, iv2v, C->xkc.as<double>(),
                                       C->e2sel.as<double>(), n3, C->Zr.as<double>(), meval, ev,
                                       C->rw);
        C->h_eval.assign(ev.begin(), ev.begin() + meval);
        prof_phase("3d");
      }
    } else {
      dense = true;   // H3 is intact: the Davidson solver works on copies
    }
  }
  C->h_kept = false;
  if (C->keep_h) {
    C->H3keep.ensure(sizeof(double) * (size_t)n3 * n3);
    DVR_HIP(hipMemcpyAsync(C->H3keep.p, C->H3.p, sizeof(double) * (size_t)n3 * n3,
                           hipMemcpyDeviceToDevice, s));
    C->h_kept = true;
  }
  C->twostage_used = false;
  std::vector<int> rq_perm;   // dense split32: level order after the refinement
  if (dense) {
    const int want = std::min(n3, meval + 16);
    // two-stage reduction (band reduction on MFMA, bulge chasing) of a copy
    // of H3, its reflectors kept when eigenvectors are wanted (also those the
    // split32 refinement needs); for small N3, or with DVR_TWOSTAGE_VEC=0
    // and eigenvectors, the one-stage reduction
    const bool wantvec = nvec > 0 || C->split32;
    if (n3 >= 4 * twostage_bandwidth() && (wantvec ? env_twostage_vec() : env_twostage())) {
      C->ts.keep = wantvec;
      C->twostage_used = twostage_values(C->H3.as<double>(), n3, want, s, C->ts, C->w3, C->H3b);
    } else {
      C->w3.tridiag(std::vector<int>{n3}, C->H3.as<double>(), std::vector<long>{0}, s);
      C->w3.values(std::vector<int>{want}, s);
    }
    C->h_eval.assign(C->w3.hval.begin(), C->w3.hval.begin() + meval);
    if (C->split32) {
      // every wanted level of the split-precision H3 refined by its Rayleigh
      // quotient with the exact FP64 operator
      prof_phase("refine");
      const int mext = C->w3.extend(0, meval);
      C->Zr.ensure(sizeof(double) * (size_t)n3 * mext);
      if (C->twostage_used)
        twostage_vectors(n3, mext, C->Zr.as<double>(), s, C->ts, C->w3);
      else
        C->w3.vectors(std::vector<int>{mext}, C->Zr.as<double>(), std::vector<long>{0}, s);
      std::vector<int> col0(nc), iv2v(nc);
      for (int c = 0; c < nc; ++c) {
        col0[c] = cb[c].col0;
        iv2v[c] = cb[c].iv2;
      }
      std::vector<double> ev;
      C->refine_resid = refine_rq(s, C->cint.as<double>(), G, col0, iv2v, C->xkc.as<double>(),
                                  C->e2sel.as<double>(), n3, C->Zr.as<double>(), meval, ev, C->rw);
      rq_perm.resize(meval);
      std::iota(rq_perm.begin(), rq_perm.end(), 0);
      std::stable_sort(rq_perm.begin(), rq_perm.end(), [&](int x, int y) { return ev[x] < ev[y]; });
      for (int k = 0; k < meval; ++k) C->h_eval[k] = ev[rq_perm[k]];
      prof_phase("3d");
    }
  }
  DVR_HIP(hipEventRecord(C->ev[6], s));

  // ---- eigenvectors and wavefunctions (trans, D:3650-3845)
  if (nvec > 0) {
    if (dense && C->split32) {
      // the refined levels' vectors, in the refined order
      std::vector<int> p(rq_perm.begin(), rq_perm.begin() + nvec);
      C->Z3.ensure(sizeof(double) * (size_t)n3 * nvec);
      launch_gather_cols(s, C->Zr.as<double>(), n3, upload(C->permd, p, s), nvec,
                         C->Z3.as<double>(), n3, n3);
    } else if (dense) {
      const int mext = C->w3.extend(0, nvec);
      C->Z3.ensure(sizeof(double) * (size_t)n3 * mext);
      if (C->twostage_used)
        twostage_vectors(n3, mext, C->Z3.as<double>(), s, C->ts, C->w3);
      else
        C->w3.vectors(std::vector<int>{mext}, C->Z3.as<double>(), std::vector<long>{0}, s);
    } else if (C->split32) {
      C->Z3.ensure(sizeof(double) * (size_t)n3 * nvec);
      DVR_HIP(hipMemcpyAsync(C->Z3.p, C->Zr.p, sizeof(double) * (size_t)n3 * nvec,
                             hipMemcpyDeviceToDevice, s));
    } else {
      C->Z3.ensure(sizeof(double) * (size_t)n3 * nvec);
      C->dav.vectors(nvec, C->Z3.as<double>(), s);
    }
    const size_t npsi = (size_t)D.idvr * D.npnt1 * D.npnt2;
    C->psi.ensure(sizeof(double) * npsi * nvec);
    DVR_HIP(hipMemsetAsync(C->psi.p, 0, sizeof(double) * npsi * nvec, s));
    launch_trans(s, C->cint.as<double>(), C->Z3.as<double>(), C->psi.as<double>(), d_cb, cb, nc,
                 na, nb, n3, nvec, D.zr2r1, D.idvr, D.npnt1, D.npnt2, C->ptmp);
  }
  DVR_HIP(hipEventRecord(C->ev[7], s));
  DVR_HIP(hipStreamSynchronize(s));
  for (int i = 0; i < DVR_NSTAGE; ++i) {
    float ms = 0.f;
    DVR_HIP(hipEventElapsedTime(&ms, C->ev[i], C->ev[i + 1]));
    C->stage_ms[i] = ms;
  }
}

}  // namespace

extern "C" {

int dvr_ctx_create(const dvr_dims_t* dims, int32_t device, dvr_ctx_t** out) {
  if (!dims || !out) return DVR_ERR_ARG;
  *out = nullptr;
  const dvr_dims_t& D = *dims;
  if (D.npnta <= 0 || D.npntb <= 0 || D.npntc <= 0 || D.max2d <= 0 ||
      D.max3d <= 0 || D.neval <= 0 || D.nvec < 0 || D.idvr < D.npnta ||
      (D.solver != DVR_SOLVER_DENSE && D.solver != DVR_SOLVER_DAVIDSON))
    return DVR_ERR_ARG;
  if (D.nvec > 0 && D.npnt1 * D.npnt2 != D.npntb * D.npntc) return DVR_ERR_ARG;
  std::unique_ptr<dvr_ctx_t> c(new dvr_ctx_t());
  c->dims = D;
  c->device = device;
  try {
    DeviceGuard g(device);
    DVR_HIP(hipSetDevice(device));
    DVR_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    for (auto& e : c->ev) DVR_HIP(hipEventCreate(&e));
  } catch (const std::exception& e) {
    return DVR_ERR_HIP;
  }
  *out = c.release();
  return DVR_OK;
}

void dvr_ctx_destroy(dvr_ctx_t* c) {
  if (!c) return;
  {
    DeviceGuard g(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto& e : c->ev)
      if (e) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    c->stream = nullptr;
    delete c;
  }
}

const char* dvr_last_error(const dvr_ctx_t* c) {
  return c ? c->err.c_str() : "null context";
}

int dvr_set_basis(dvr_ctx_t* c, const double* xlmatr, const double* xkb,
                  const double* xkc, const double* wsum, const double* wterm,
                  double emax1, double emax2) {
  if (!c || !xlmatr || !xkb || !xkc || !wsum) return fail(c, DVR_ERR_ARG, "null argument");
  const dvr_dims_t& D = c->dims;
  try {
    DeviceGuard g(c->device);
    auto put = [&](DevBuf& b, const double* h, size_t n) {
      b.ensure(sizeof(double) * n);
      DVR_HIP(hipMemcpyAsync(b.p, h, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    };
    put(c->xl, xlmatr, (size_t)D.npnta * D.npnta);
    put(c->xkb, xkb, (size_t)D.npntb * D.npntb);
    put(c->xkc, xkc, (size_t)D.npntc * D.npntc);
    put(c->wsum, wsum, (size_t)D.npntc * D.npntb);
    c->has_wterm = wterm != nullptr;
    if (wterm) put(c->wterm, wterm, (size_t)D.npntc * D.npntb);
    DVR_HIP(hipStreamSynchronize(c->stream));
  } catch (const std::exception& e) {
    return fail(c, DVR_ERR_HIP, e.what());
  }
  c->emax1 = emax1;
  c->emax2 = emax2;
  c->basis_set = true;
  return DVR_OK;
}

int dvr_set_potential(dvr_ctx_t* c, const double* v) {
  if (!c || !v) return fail(c, DVR_ERR_ARG, "null argument");
  const dvr_dims_t& D = c->dims;
  try {
    DeviceGuard g(c->device);
    const size_t n = (size_t)D.npnta * D.npntb * D.npntc;
    c->v.ensure(sizeof(double) * n);
    DVR_HIP(hipMemcpyAsync(c->v.p, v, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    DVR_HIP(hipStreamSynchronize(c->stream));
  } catch (const std::exception& e) {
    return fail(c, DVR_ERR_HIP, e.what());
  }
  c->pot_set = true;
  return DVR_OK;
}

int dvr_eval_potential_ab2(dvr_ctx_t* c, const double* rc, const double* rb,
                           const double* xa, const double* par) {
  if (!c || !rc || !rb || !xa || !par) return fail(c, DVR_ERR_ARG, "null argument");
  const dvr_dims_t& D = c->dims;
  try {
    DeviceGuard g(c->device);
    const size_t n = (size_t)D.npnta * D.npntb * D.npntc;
    c->v.ensure(sizeof(double) * n);
    auto put = [&](DevBuf& b, const double* h, size_t k) {
      b.ensure(sizeof(double) * k);
      DVR_HIP(hipMemcpyAsync(b.p, h, sizeof(double) * k, hipMemcpyHostToDevice, c->stream));
    };
    put(c->rc, rc, D.npntc);
    put(c->rb, rb, D.npntb);
    put(c->xa, xa, D.npnta);
    launch_pes_ab2(c->stream, c->v.as<double>(), c->rc.as<double>(),
                   c->rb.as<double>(), c->xa.as<double>(), D.npntc, D.npntb,
                   D.npnta, D.zr2r1, par);
    DVR_HIP(hipStreamSynchronize(c->stream));
  } catch (const std::exception& e) {
    return fail(c, DVR_ERR_HIP, e.what());
  }
  c->pot_set = true;
  return DVR_OK;
}

int dvr_run(dvr_ctx_t* c) {
  if (!c) return DVR_ERR_ARG;
  if (!c->basis_set || !c->pot_set)
    return fail(c, DVR_ERR_ARG, "dvr_run: basis and potential must be set first");
  c->ran = false;
  try {
    DeviceGuard g(c->device);
    if (c->profiling) {
      if (!c->prof) c->prof.reset(new Profiler());
      c->prof->reset();
      c->prof->phase.clear();
      g_prof = c->prof.get();
    }
    struct ProfOff {
      ~ProfOff() { g_prof = nullptr; }
    } prof_off;
    run_pipeline(c);
    if (c->profiling) {
      DVR_HIP(hipStreamSynchronize(c->stream));
      c->kstats.clear();
      for (const ProfRec& r : c->prof->recs) {
        float ms = 0.f;
        DVR_HIP(hipEventElapsedTime(&ms, r.a, r.b));
        auto it = std::find_if(c->kstats.begin(), c->kstats.end(),
                               [&](const dvr_ctx::KStat& k) { return k.name == r.name; });
        if (it == c->kstats.end()) {
          c->kstats.push_back(dvr_ctx::KStat{r.name, 0, 0.0, 0.0, 0.0});
          it = c->kstats.end() - 1;
        }
        it->launches += 1;
        it->seconds += 1e-3 * ms;
        it->bytes += r.bytes;
        it->flops += r.flops;
      }
    }
  } catch (const HipError& e) {
    return fail(c, DVR_ERR_HIP, e.what());
  } catch (const std::runtime_error& e) {
    const std::string w = e.what();
    if (w == "core2d")
      return fail(c, DVR_ERR_CORE_2D, "**** core exceeded: reduce cut-off emax1 ****");
    if (w == "core3d")
      return fail(c, DVR_ERR_CORE_3D, "**** core exceeded: reduce cut-off emax2 ****");
    return fail(c, DVR_ERR_ARG, w);
  } catch (const std::exception& e) {
    return fail(c, DVR_ERR_HIP, e.what());
  }
  c->ran = true;
  return DVR_OK;
}

// The k blocks of one rotational run are independent problems of different
// sizes (the ccmain loop, D:857-1007, solves them one after another).  Each
// context owns its stream and buffers, so one host thread per context lets
// the blocks share the GPU: a small block's launches fill the gaps of a
// larger one's.
int dvr_run_blocks(dvr_ctx_t* const* ctxs, int n) {
  if (!ctxs || n < 0) return DVR_ERR_ARG;
  for (int i = 0; i < n; ++i) {
    if (!ctxs[i]) return DVR_ERR_ARG;
    for (int j = 0; j < i; ++j)
      if (ctxs[j] == ctxs[i]) return DVR_ERR_ARG;   // one thread per context
  }
  if (n == 1) return dvr_run(ctxs[0]);
  std::vector<int> rc(n, DVR_OK);
  std::vector<std::thread> th;
  th.reserve(n);
  for (int i = 0; i < n; ++i) th.emplace_back([&, i] { rc[i] = dvr_run(ctxs[i]); });
  for (auto& t : th) t.join();
  for (int i = 0; i < n; ++i)
    if (rc[i] != DVR_OK) return rc[i];
  return DVR_OK;
}

int dvr_get_sizes(const dvr_ctx_t* c, int32_t* iv1, int32_t* iv2, int32_t* nham3) {
  if (!c || !c->ran) return DVR_ERR_ARG;
  const dvr_dims_t& D = c->dims;
  if (iv1)
    for (int cc = 0; cc < D.npntc; ++cc)
      for (int b = 0; b < D.npntb; ++b)
        iv1[cc + D.npntc * b] = c->h_iv1[cc * D.npntb + b];
  if (iv2)
    for (int cc = 0; cc < D.npntc; ++cc) iv2[cc] = c->h_iv2[cc];
  if (nham3) *nham3 = c->nham3;
  return DVR_OK;
}

int dvr_get_eigenvalues(const dvr_ctx_t* c, double* eval, int32_t* meval) {
  if (!c || !c->ran) return DVR_ERR_ARG;
  if (meval) *meval = c->meval;
  if (eval) std::memcpy(eval, c->h_eval.data(), sizeof(double) * c->meval);
  return DVR_OK;
}

int dvr_get_1d_eigenvalues(const dvr_ctx_t* c, double* e1) {
  if (!c || !c->ran || !e1) return DVR_ERR_ARG;
  const dvr_dims_t& D = c->dims;
  DeviceGuard g(c->device);
  const size_t n = (size_t)D.npnta * D.npntb * D.npntc;
  if (hipMemcpy(e1, c->e1full.p, sizeof(double) * n, hipMemcpyDeviceToHost) != hipSuccess)
    return DVR_ERR_HIP;
  return DVR_OK;
}

int dvr_get_2d_eigenvalues(const dvr_ctx_t* c, double* e2, int32_t* total) {
  if (!c || !c->ran) return DVR_ERR_ARG;
  if (total) *total = (int32_t)c->h_e2all.size();
  if (e2) std::memcpy(e2, c->h_e2all.data(), sizeof(double) * c->h_e2all.size());
  return DVR_OK;
}

int dvr_get_wavefunctions(const dvr_ctx_t* c, double* psi) {
  if (!c || !c->ran || !psi || c->nvec_done <= 0) return DVR_ERR_ARG;
  const dvr_dims_t& D = c->dims;
  DeviceGuard g(c->device);
  const size_t n = (size_t)D.idvr * D.npnt1 * D.npnt2 * c->nvec_done;
  if (hipMemcpy(psi, c->psi.p, sizeof(double) * n, hipMemcpyDeviceToHost) != hipSuccess)
    return DVR_ERR_HIP;
  return DVR_OK;
}

int dvr_get_vectors(const dvr_ctx_t* c, double* vec) {
  if (!c || !c->ran || !vec || c->nvec_done <= 0) return DVR_ERR_ARG;
  DeviceGuard g(c->device);
  const size_t n = (size_t)c->nham3 * c->nvec_done;
  if (hipMemcpy(vec, c->Z3.p, sizeof(double) * n, hipMemcpyDeviceToHost) != hipSuccess)
    return DVR_ERR_HIP;
  return DVR_OK;
}

int dvr_get_solver_stats(const dvr_ctx_t* c, int32_t* st, int32_t n) {
  if (!c || !c->ran || !st || n < 1) return DVR_ERR_ARG;
  const DavStats& d = c->dav.stats;
  const int32_t v[15] = {c->dav_status, d.iterations, d.matvecs, d.solves, d.factorizations,
                         d.slices, d.certified, d.missing,
                         c->dims.solver == DVR_SOLVER_DAVIDSON && c->dav_status != 0,
                         (int32_t)std::lround(d.phase_ms[0] * 1e3),
                         (int32_t)std::lround(d.phase_ms[1] * 1e3),
                         (int32_t)std::lround(d.phase_ms[2] * 1e3),
                         (int32_t)std::lround(d.phase_ms[3] * 1e3),
                         (int32_t)std::lround(d.phase_ms[4] * 1e3), c->twostage_used ? 1 : 0};
  for (int i = 0; i < n && i < 15; ++i) st[i] = v[i];
  return DVR_OK;
}

int dvr_keep_hamiltonian(dvr_ctx_t* c, int32_t on) {
  if (!c) return DVR_ERR_ARG;
  c->keep_h = on != 0;
  return DVR_OK;
}

int dvr_set_preconditioner(dvr_ctx_t* c, int32_t kind, const float* params) {
  const bool learned = kind == DVR_PREC_LEARNED || kind == DVR_PREC_SI_LEARNED;
  if (!c || kind < 0 || kind > 3 || (learned && !params)) return DVR_ERR_ARG;
  try {
    DeviceGuard g(c->device);
    c->prec_kind = kind;
    if (learned) {
      c->mlpP.ensure(sizeof(float) * MLP_NPARAM);
      DVR_HIP(hipMemcpy(c->mlpP.p, params, sizeof(float) * MLP_NPARAM, hipMemcpyHostToDevice));
      c->mlp_set = true;
    }
  } catch (const std::exception& e) {
    return fail(c, DVR_ERR_HIP, e.what());
  }
  return DVR_OK;
}

int dvr_mlp_train(dvr_ctx_t* c, int32_t steps, int32_t batch, uint64_t seed, float lr,
                  float* params, double* loss_hist) {
  if (!c || !params || steps < 1 || batch < 1) return DVR_ERR_ARG;
  if (!c->ran || !c->h_kept || c->nvec_done < 2)
    return fail(c, DVR_ERR_ARG,
                "dvr_mlp_train: needs a run with dvr_keep_hamiltonian and nvec >= 2");
  try {
    DeviceGuard g(c->device);
    hipStream_t s = c->stream;
    const int n3 = c->nham3, nz = c->nvec_done;
    DevBuf P, Dd;
    P.ensure(sizeof(float) * MLP_NPARAM);
    DVR_HIP(hipMemcpyAsync(P.p, params, sizeof(float) * MLP_NPARAM, hipMemcpyHostToDevice, s));
    Dd.ensure(sizeof(double) * n3);
    launch_get_diag(s, c->H3keep.as<double>(), n3, Dd.as<double>());
    std::vector<double> lam(c->h_eval.begin(), c->h_eval.begin() + nz);
    std::vector<double> hist;
    const int rc = mlp_train(s, c->H3keep.as<double>(), n3, c->Z3.as<double>(), nz, lam,
                             Dd.as<double>(), c->st3.as<double>(), P.as<float>(), steps, batch,
                             (unsigned long long)seed, lr, hist, c->mw);
    if (rc != 0) return fail(c, DVR_ERR_ARG, "dvr_mlp_train: bad sizes");
    DVR_HIP(hipMemcpyAsync(params, P.p, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToHost, s));
    DVR_HIP(hipStreamSynchronize(s));
    if (loss_hist) std::memcpy(loss_hist, hist.data(), sizeof(double) * hist.size());
  } catch (const std::exception& e) {
    return fail(c, DVR_ERR_HIP, e.what());
  }
  return DVR_OK;
}

int dvr_mlp_forward(const float* params, int64_t nrows, const float* feats, float* w,
                    int32_t bf16) {
  if (!params || nrows <= 0 || !feats || !w) return DVR_ERR_ARG;
  try {
    hipStream_t s;
    DVR_HIP(hipStreamCreate(&s));
    {
      DevBuf P, F, W;
      P.ensure(sizeof(float) * MLP_NPARAM);
      F.ensure(sizeof(float) * nrows * MLP_NF);
      W.ensure(sizeof(float) * nrows);
      DVR_HIP(hipMemcpyAsync(P.p, params, sizeof(float) * MLP_NPARAM, hipMemcpyHostToDevice, s));
      DVR_HIP(hipMemcpyAsync(F.p, feats, sizeof(float) * nrows * MLP_NF, hipMemcpyHostToDevice, s));
      if (bf16)
        launch_mlp_infer_bf16(s, P.as<float>(), nrows, F.as<float>(), W.as<float>());
      else
        launch_mlp_fwd(s, P.as<float>(), nrows, F.as<float>(), W.as<float>(), nullptr, nullptr);
      DVR_HIP(hipMemcpyAsync(w, W.p, sizeof(float) * nrows, hipMemcpyDeviceToHost, s));
      DVR_HIP(hipStreamSynchronize(s));
    }
    DVR_HIP(hipStreamDestroy(s));
  } catch (const std::exception& e) {
    fprintf(stderr, "dvr_mlp_forward: %s\n", e.what());
    return DVR_ERR_HIP;
  }
  return DVR_OK;
}

int dvr_mlp_backward(const float* params, int64_t nrows, const float* feats, const float* gw,
                     float* grad) {
  if (!params || nrows <= 0 || !feats || !gw || !grad) return DVR_ERR_ARG;
  try {
    hipStream_t s;
    DVR_HIP(hipStreamCreate(&s));
    {
      const long R = nrows;
      DevBuf P, F, W, GW, H1, H2, GA1, GA2, G, part, ones;
      P.ensure(sizeof(float) * MLP_NPARAM);
      F.ensure(sizeof(float) * R * MLP_NF);
      W.ensure(sizeof(float) * R);
      GW.ensure(sizeof(float) * R);
      H1.ensure(sizeof(float) * R * MLP_WIDTH);
      H2.ensure(sizeof(float) * R * MLP_WIDTH);
      GA1.ensure(sizeof(float) * R * MLP_WIDTH);
      GA2.ensure(sizeof(float) * R * MLP_WIDTH);
      G.ensure(sizeof(float) * MLP_NPARAM);
      const long pcap = 64L * MLP_WIDTH * MLP_WIDTH;
      part.ensure(sizeof(float) * pcap);
      std::vector<float> one(R, 1.0f);
      ones.ensure(sizeof(float) * R);
      DVR_HIP(hipMemcpyAsync(ones.p, one.data(), sizeof(float) * R, hipMemcpyHostToDevice, s));
      DVR_HIP(hipMemcpyAsync(P.p, params, sizeof(float) * MLP_NPARAM, hipMemcpyHostToDevice, s));
      DVR_HIP(hipMemcpyAsync(F.p, feats, sizeof(float) * R * MLP_NF, hipMemcpyHostToDevice, s));
      DVR_HIP(hipMemcpyAsync(GW.p, gw, sizeof(float) * R, hipMemcpyHostToDevice, s));
      launch_mlp_fwd(s, P.as<float>(), R, F.as<float>(), W.as<float>(), H1.as<float>(),
                     H2.as<float>());
      launch_mlp_bwd(s, P.as<float>(), R, F.as<float>(), H1.as<float>(), H2.as<float>(),
                     GW.as<float>(), GA1.as<float>(), GA2.as<float>(), G.as<float>(),
                     part.as<float>(), pcap);
      launch_mlp_bias_grads(s, R, ones.as<float>(), GA1.as<float>(), GA2.as<float>(),
                            GW.as<float>(), G.as<float>(), part.as<float>(), pcap);
      DVR_HIP(hipMemcpyAsync(grad, G.p, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToHost, s));
      DVR_HIP(hipStreamSynchronize(s));
    }
    DVR_HIP(hipStreamDestroy(s));
  } catch (const std::exception& e) {
    fprintf(stderr, "dvr_mlp_backward: %s\n", e.what());
    return DVR_ERR_HIP;
  }
  return DVR_OK;
}

int dvr_ldl_solve(const double* a, int32_t n, double sigma, const double* r, int32_t nrhs,
                  double* t, int32_t* inertia) {
  if (!a || n <= 0 || nrhs < 0 || (nrhs > 0 && (!r || !t))) return DVR_ERR_ARG;
  bool bad = false;
  try {
    hipStream_t s;
    DVR_HIP(hipStreamCreate(&s));
    {
      DevBuf H, R, T;
      H.ensure(sizeof(double) * (size_t)n * n);
      DVR_HIP(hipMemcpyAsync(H.p, a, sizeof(double) * (size_t)n * n, hipMemcpyHostToDevice, s));
      LdlFactor f;
      f.panel_sub = env_int("DVR_LDL_SUB", 4);
      f.upd_xcd = env_int("DVR_LDL_XCD", 1) != 0;
      f.lookahead = env_int("DVR_LDL_LOOKAHEAD", 1) != 0;
      f.sb = std::max(64, env_int("DVR_LDL_SB", 2048) / 64 * 64);
      f.upd_cp = env_int("DVR_LDL_CP", 1) != 0;
      f.upd_wide = env_int("DVR_LDL_WIDE", 32);
      f.ensure(n);
      const int nu = f.factor(H.as<double>(), sigma, s);
      if (inertia) *inertia = nu;
      if (nu < 0) {
        DVR_HIP(hipStreamSynchronize(s));
        bad = true;
      } else if (nrhs > 0) {
        R.ensure(sizeof(double) * (size_t)n * nrhs);
        T.ensure(sizeof(double) * (size_t)n * nrhs);
        DVR_HIP(hipMemcpyAsync(R.p, r, sizeof(double) * (size_t)n * nrhs, hipMemcpyHostToDevice,
                               s));
        f.solve(R.as<double>(), T.as<double>(), nrhs, s);
        DVR_HIP(hipMemcpyAsync(t, T.p, sizeof(double) * (size_t)n * nrhs, hipMemcpyDeviceToHost,
                               s));
      }
      DVR_HIP(hipStreamSynchronize(s));
    }
    DVR_HIP(hipStreamDestroy(s));
  } catch (const std::exception& e) {
    fprintf(stderr, "dvr_ldl_solve: %s\n", e.what());
    return DVR_ERR_HIP;
  }
  return bad ? DVR_ERR_NOCONV : DVR_OK;
}

int dvr_set_profiling(dvr_ctx_t* c, int32_t on) {
  if (!c) return DVR_ERR_ARG;
  c->profiling = on != 0;
  return DVR_OK;
}

int dvr_get_profile(const dvr_ctx_t* c, int32_t idx, char* name, int32_t name_len,
                    int32_t* launches, double* seconds, double* bytes, double* flops) {
  if (!c || idx < 0 || idx >= (int)c->kstats.size()) return DVR_ERR_ARG;
  const auto& k = c->kstats[idx];
  if (name && name_len > 0) {
    std::strncpy(name, k.name.c_str(), name_len - 1);
    name[name_len - 1] = 0;
  }
  if (launches) *launches = k.launches;
  if (seconds) *seconds = k.seconds;
  if (bytes) *bytes = k.bytes;
  if (flops) *flops = k.flops;
  return DVR_OK;
}

int dvr_get_stage_times(const dvr_ctx_t* c, double* ms) {
  if (!c || !ms) return DVR_ERR_ARG;
  for (int i = 0; i < DVR_NSTAGE; ++i) ms[i] = c->stage_ms[i];
  return DVR_OK;
}

int dvr_syev_batch(int32_t nprob, const int32_t* n, const double* a,
                   const int32_t* nvec, double* w, double* z) {
  if (nprob <= 0 || !n || !a || !nvec || !w) return DVR_ERR_ARG;
  for (int p = 0; p < nprob; ++p)
    if (n[p] <= 0 || nvec[p] < 0 || nvec[p] > n[p]) return DVR_ERR_ARG;
  hipStream_t s = nullptr;
  try {
    DVR_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    std::vector<int> hn(n, n + nprob), cntv(nvec, nvec + nprob);
    std::vector<long> aoff(nprob);
    long tot = 0;
    for (int p = 0; p < nprob; ++p) {
      aoff[p] = tot;
      tot += (long)hn[p] * hn[p];
    }
    DevBuf A;
    A.ensure(sizeof(double) * tot);
    // symmetrise from the lower triangle on the host while staging
    std::vector<double> stage(a, a + tot);
    for (int p = 0; p < nprob; ++p) {
      double* M = stage.data() + aoff[p];
      const int np = hn[p];
      for (int j = 0; j < np; ++j)
        for (int i = 0; i < j; ++i) M[i + (long)np * j] = M[j + (long)np * i];
    }
    DVR_HIP(hipMemcpyAsync(A.p, stage.data(), sizeof(double) * tot, hipMemcpyHostToDevice, s));
    EigWork W;
    int nmx = 0, nmn = 1 << 30;
    for (int x : hn) {
      nmx = std::max(nmx, x);
      nmn = std::min(nmn, x);
    }
    // every problem of order >= 256: two-stage reduction of each problem in
    // turn, eigenvectors back-transformed through both stages
    if (nmn >= 4 * twostage_bandwidth() && (z ? env_twostage_vec() : env_twostage())) {
      TwoStage ts;
      ts.keep = z != nullptr;
      DevBuf cp, Zp;
      long wo = 0, zo = 0;
      for (int p = 0; p < nprob; ++p) {
        const int want = std::min(hn[p], cntv[p] + (cntv[p] < hn[p] ? 16 : 0));
        EigWork Wp;
        const bool two = twostage_values(A.as<double>() + aoff[p], hn[p], want, s, ts, Wp, cp);
        std::memcpy(w + wo, Wp.hval.data(), sizeof(double) * cntv[p]);
        wo += cntv[p];
        if (z && cntv[p] > 0) {
          const int mext = Wp.extend(0, cntv[p]);
          Zp.ensure(sizeof(double) * (size_t)hn[p] * mext);
          if (two)
            twostage_vectors(hn[p], mext, Zp.as<double>(), s, ts, Wp);
          else
            Wp.vectors(std::vector<int>{mext}, Zp.as<double>(), std::vector<long>{0}, s);
          DVR_HIP(hipMemcpyAsync(z + zo, Zp.p, sizeof(double) * (size_t)hn[p] * cntv[p],
                                 hipMemcpyDeviceToHost, s));
          DVR_HIP(hipStreamSynchronize(s));
          zo += (long)hn[p] * cntv[p];
        }
      }
      DVR_HIP(hipStreamSynchronize(s));
      (void)hipStreamDestroy(s);
      return DVR_OK;
    }
    // DVR_SYEV=wg selects the single-work-group solver, =multi the
    // multi-launch one
    const char* senv = getenv("DVR_SYEV");
    const bool use_wg = senv ? std::string(senv) == "wg" : false;
    W.tridiag(hn, A.as<double>(), aoff, s, use_wg);
    std::vector<int> want(nprob);
    for (int p = 0; p < nprob; ++p) want[p] = std::min(hn[p], cntv[p] + (cntv[p] < hn[p] ? 16 : 0));
    W.values(want, s);
    long wo = 0;
    for (int p = 0; p < nprob; ++p) {
      std::memcpy(w + wo, W.hval.data() + W.heoff[p], sizeof(double) * cntv[p]);
      wo += cntv[p];
    }
    if (z) {
      std::vector<int> mext(nprob);
      std::vector<long> zoff(nprob);
      long zt = 0;
      for (int p = 0; p < nprob; ++p) {
        mext[p] = W.extend(p, cntv[p]);
        zoff[p] = zt;
        zt += (long)hn[p] * mext[p];
      }
      DevBuf Z;
      Z.ensure(sizeof(double) * (zt + 1));
      W.vectors(mext, Z.as<double>(), zoff, s);
      long zo = 0;
      for (int p = 0; p < nprob; ++p) {
        DVR_HIP(hipMemcpyAsync(z + zo, Z.as<double>() + zoff[p],
                               sizeof(double) * (size_t)hn[p] * cntv[p],
                               hipMemcpyDeviceToHost, s));
        zo += (long)hn[p] * cntv[p];
      }
      DVR_HIP(hipStreamSynchronize(s));
    }
    DVR_HIP(hipStreamSynchronize(s));
    (void)hipStreamDestroy(s);
  } catch (const std::exception& e) {
    fprintf(stderr, "dvr_syev_batch: %s\n", e.what());
    if (s) (void)hipStreamDestroy(s);
    return DVR_ERR_HIP;
  }
  return DVR_OK;
}

int dvr_diag(double* ham, int32_t maxham, int32_t nham, double* eig) {
  if (!ham || !eig || nham <= 0 || maxham < nham) return DVR_ERR_ARG;
  std::vector<double> a((size_t)nham * nham), z((size_t)nham * nham);
  for (int j = 0; j < nham; ++j)
    for (int i = 0; i < nham; ++i) a[i + (size_t)nham * j] = ham[i + (size_t)maxham * j];
  const int32_t nv = nham;
  const int rc = dvr_syev_batch(1, &nham, a.data(), &nv, eig, z.data());
  if (rc != DVR_OK) return rc;
  for (int j = 0; j < nham; ++j)
    for (int i = 0; i < nham; ++i) ham[i + (size_t)maxham * j] = z[i + (size_t)nham * j];
  return DVR_OK;
}

}  // extern "C"
