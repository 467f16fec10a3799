// Block-Davidson final solve with certified spectrum slicing.
//
// Replaces the dense diag3d eigensolve (source/dvr3drjz_segmented.f90:
// 2661-2764 -> diag 2638-2659) when only the lowest NEVAL of N3 levels are
// wanted.  The levels are found in slices from the bottom of the spectrum up:
//
//  * in slice s the search space is expanded with shift-and-invert
//    corrections t = (H - sigma_s I)^-1 r of the lowest unconverged Ritz
//    pairs (r = H x - theta x), sigma_s being the upper end of the previous
//    slice; Ritz pairs are locked in ascending order once ||r|| <= tol;
//  * when SLICE levels are locked, sigma_{s+1} is placed half-way between the
//    highest locked level and the next Ritz value, H - sigma_{s+1} I is
//    factored as L D L^T and Sylvester's law of inertia (number of negative
//    pivots) must equal the number of locked levels below sigma_{s+1}.  A
//    mismatch (a missed level) or a stalled iteration hands the problem back
//    to the dense solver.
//
// Everything runs on the device in FP64: MFMA GEMMs for H V, the projections
// and the Gram-Schmidt passes; the projected Rayleigh-Ritz eigenproblems
// through the dense eigensolver; each new block orthonormalised by Cholesky
// QR (twice, with classical Gram-Schmidt against the basis in between).  The
// host only makes the locking and restart decisions from the Ritz values and
// residual norms.
#include "dvr_internal.h"

#include <cmath>
#include <numeric>

namespace dvr {

void DavidsonSolver::gemm(hipStream_t s, const char* name, bool ta, bool tb, int M, int N, int K,
                          double alpha, const double* A, long lda, const double* B, long ldb,
                          double beta, double* C, long ldc, bool lower) {
  tuner.run(s, name, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, lower);
}

// eigen-decomposition of the symmetric m x m matrix whose lower triangle is
// at S (destroyed): ascending values to host w, vectors to U (m x m, ld m)
void DavidsonSolver::small_eig(int m, double* S, double* U, std::vector<double>& w,
                               hipStream_t s) {
  launch_mirror_lower(s, S, m);
  // only (near-)degenerate Ritz values share an inverse-iteration thread; the
  // small loss of orthogonality between the other vectors (eps ||T|| / gap)
  // is removed by two Newton-Schulz steps U <- U (3 I - U^T U) / 2, which mix
  // columns in proportion to their overlap and so keep the residuals at the
  // eps ||T|| level
  rr.cluster_rel = 1e-10;
  rr.tridiag(std::vector<int>{m}, S, std::vector<long>{0}, s, rr_single_wg);
  rr.values(std::vector<int>{m}, s);
  w.assign(rr.hval.begin(), rr.hval.begin() + m);
  if (U) {
    rr.vectors(std::vector<int>{m}, U, std::vector<long>{0}, s);
    double* g = G2.as<double>();
    double* u2 = Y2.as<double>();
    for (int it = 0; it < 2; ++it) {
      gemm(s, "ns_gram", true, false, m, m, m, 1.0, U, m, U, m, 0.0, g, m);
      DVR_HIP(hipMemcpyAsync(u2, U, sizeof(double) * (size_t)m * m, hipMemcpyDeviceToDevice, s));
      gemm(s, "ns_upd", false, false, m, m, m, -0.5, U, m, g, m, 1.5, u2, m);
      DVR_HIP(hipMemcpyAsync(U, u2, sizeof(double) * (size_t)m * m, hipMemcpyDeviceToDevice, s));
    }
  }
}

// Orthonormalise the q0 columns at Tb (N x q0, ld N, destroyed) against the
// locked vectors and the current basis, append the survivors to V (with
// their products H V); returns their number.  Blocks wider than one Gram
// factorisation (64) go through in chunks of 64, each orthogonalised
// against the basis grown by the chunks before it.
int DavidsonSolver::expand(const double* H, int q0, hipStream_t s) {
  if (q0 <= 0) return 0;
  if (q0 <= 64) return expand_chunk(H, Tb.as<double>(), q0, s);
  const int m0 = m;
  int qt = 0;
  for (int c0 = 0; c0 < q0; c0 += 64) {
    const int q = expand_chunk(H, Tb.as<double>() + (long)c0 * N, std::min(64, q0 - c0), s);
    m += q;
    qt += q;
  }
  m = m0;
  return qt;
}

int DavidsonSolver::expand_chunk(const double* H, double* Wt, int q0, hipStream_t s) {
  double* Vp = V.as<double>();
  double* XLp = XL.as<double>();
  double* Cp = C1.as<double>();
  auto cgs = [&](double* Z, int q) {
    if (nlocked > 0) {
      gemm(s, "orth_proj", true, false, nlocked, q, N, 1.0, XLp, N, Z, N, 0.0, Cp, nlocked);
      gemm(s, "orth_upd", false, false, N, q, nlocked, -1.0, XLp, N, Cp, nlocked, 1.0, Z, N);
    }
    if (m > 0) {
      gemm(s, "orth_proj", true, false, m, q, N, 1.0, Vp, N, Z, N, 0.0, Cp, m);
      gemm(s, "orth_upd", false, false, N, q, m, -1.0, Vp, N, Cp, m, 1.0, Z, N);
    }
  };
  // norms of the raw corrections: a column that the projections reduce
  // below 1e-10 of its norm lies in the span of the basis and is dropped
  rawd.ensure(sizeof(double) * 64);
  launch_coldot(s, Wt, Wt, N, q0, rawd.as<double>());
  std::vector<double> raw2(q0);
  DVR_HIP(hipMemcpyAsync(raw2.data(), rawd.p, sizeof(double) * q0, hipMemcpyDeviceToHost, s));
  cgs(Wt, q0);
  cgs(Wt, q0);
  double* Vn = Vp + (long)m * N;
  double* tmp = X.as<double>();   // free between the restart and the next RR
  int q = cholqr(Wt, q0, tmp, Vn, s, raw2.data());
  if (q == 0) return 0;
  cgs(Vn, q);
  q = cholqr(Vn, q, tmp, Wt, s);
  if (q == 0) return 0;
  DVR_HIP(hipMemcpyAsync(Vn, Wt, sizeof(double) * (size_t)N * q, hipMemcpyDeviceToDevice, s));
  // AV for the new columns
  gemm(s, "matvec", false, false, N, q, N, 1.0, H, N, Vn, N, 0.0, AV.as<double>() + (long)m * N, N);
  stats.matvecs += q;
  return q;
}

// Cholesky QR of the q columns at Z: G = Z^T Z is scaled to unit diagonal,
// regularised by 1e-14 and factored as L D L^T; out receives the columns of
// Z S L^-T D^-1/2 whose pivot exceeds 1e-12 (the others are numerically
// dependent and dropped) and, with raw2 (host, the squared norms of the
// columns before their projection), whose squared norm is above 1e-20
// raw2.  tmp: N x q workspace.  Returns the number kept.
int DavidsonSolver::cholqr(const double* Z, int q, double* tmp, double* out, hipStream_t s,
                           const double* raw2) {
  double* Gp = G.as<double>();
  gemm(s, "gram", true, false, q, q, N, 1.0, Z, N, Z, N, 0.0, Gp, q);
  gd.ensure(sizeof(double) * 64);
  gsc.ensure(sizeof(double) * 64);
  gr.ensure(sizeof(double) * 64);
  gm.ensure(sizeof(double) * 64 * 64);
  launch_gram_prep(s, Gp, q, 1e-14, gsc.as<double>());
  launch_ldl_block(s, Gp, q, q, gd.as<double>(), U.as<double>());
  std::vector<double> dh(q), sch(q);
  DVR_HIP(hipMemcpyAsync(dh.data(), gd.p, sizeof(double) * q, hipMemcpyDeviceToHost, s));
  DVR_HIP(hipMemcpyAsync(sch.data(), gsc.p, sizeof(double) * q, hipMemcpyDeviceToHost, s));
  DVR_HIP(hipStreamSynchronize(s));
  std::vector<double> r(q, 0.0);
  std::vector<int> keep;
  for (int k = 0; k < q; ++k)
    if (sch[k] > 0.0 && dh[k] > 1e-12 &&
        (!raw2 || 1.0 / (sch[k] * sch[k]) > 1e-20 * raw2[k])) {
      r[k] = 1.0 / std::sqrt(dh[k]);
      keep.push_back(k);
    }
  if (keep.empty()) return 0;
  launch_form_m(s, U.as<double>(), q, upload(gr, r, s), gsc.as<double>(), gm.as<double>());
  gemm(s, "cholqr", false, true, N, q, q, 1.0, Z, N, gm.as<double>(), q, 0.0, tmp, N);
  launch_gather_cols(s, tmp, N, upload(idxd, keep, s), (int)keep.size(), out, N, N);
  return (int)keep.size();
}

void PhaseClock::mark(int label, hipStream_t s) {
  if (used == (int)ev.size()) {
    hipEvent_t e;
    DVR_HIP(hipEventCreate(&e));
    ev.push_back(e);
    lab.push_back(-1);
  }
  lab[used] = label;
  DVR_HIP(hipEventRecord(ev[used], s));
  ++used;
}

void PhaseClock::collect(double* out) {
  for (int i = 0; i + 1 < used; ++i) {
    if (lab[i] < 0) continue;
    float ms = 0.0f;
    DVR_HIP(hipEventElapsedTime(&ms, ev[i], ev[i + 1]));
    out[lab[i]] += ms;
  }
  used = 0;
}

enum { PH_RR = 0, PH_SI = 1, PH_EXPAND = 2, PH_FACTOR = 3, PH_START = 4 };

DavidsonSolver::~DavidsonSolver() {
  if (s2) {
    (void)hipStreamSynchronize(s2);
    (void)hipStreamDestroy(s2);
  }
  if (ev_s) (void)hipEventDestroy(ev_s);
}

int DavidsonSolver::solve(const double* H, int n, const DavParams& P, hipStream_t s) {
  clock.used = 0;
  int rc = 0;
  try {
    rc = solve_impl(H, n, P, s);
  } catch (...) {
    // nothing may still read H or the factors when the caller takes over
    if (s2) (void)hipStreamSynchronize(s2);
    (void)hipStreamSynchronize(s);
    clock.used = 0;
    throw;
  }
  // an early factorisation still running when the solve ended (stall)
  if (s2) DVR_HIP(hipStreamSynchronize(s2));
  clock.mark(-1, s);
  DVR_HIP(hipStreamSynchronize(s));
  clock.collect(stats.phase_ms);
  return rc;
}

int DavidsonSolver::solve_impl(const double* H, int n, const DavParams& P, hipStream_t s) {
  stats = DavStats{};
  N = n;
  const int nev = std::min(P.nev, n);
  const int b = std::max(1, std::min(std::min(P.block, n), 128));   // Gram matrices of <= 64 columns
  const int K = std::max(b, P.slice);
  const int nlmax = std::min(n, nev + 1);
  // the locked levels plus the search space plus one correction block must
  // leave room in R^n: otherwise the corrections have nowhere to go and the
  // orthonormalisation only produces rounding noise.  Too small a problem
  // for that is left to the dense solver.
  if (n < nlmax + 3 * b) return 5;
  const int mmax = std::min(std::min(n, std::max(P.mmax, 2 * b)), n - nlmax - b);
  const int keep = P.keep > 0 ? std::min(std::max(P.keep, b), mmax - b) : 0;
  const double tol = P.tol;
  V.ensure(sizeof(double) * (size_t)N * mmax);
  AV.ensure(sizeof(double) * (size_t)N * mmax);
  X.ensure(sizeof(double) * (size_t)N * mmax);
  AX.ensure(sizeof(double) * (size_t)N * mmax);
  XL.ensure(sizeof(double) * (size_t)N * nlmax);
  Rb.ensure(sizeof(double) * (size_t)N * b);
  Tb.ensure(sizeof(double) * (size_t)N * b);
  Tm.ensure(sizeof(double) * (size_t)mmax * mmax);
  Y.ensure(sizeof(double) * (size_t)mmax * mmax);
  Y2.ensure(sizeof(double) * (size_t)mmax * mmax);
  G2.ensure(sizeof(double) * (size_t)mmax * mmax);
  G.ensure(sizeof(double) * (size_t)b * b);
  U.ensure(sizeof(double) * (size_t)b * b);
  C1.ensure(sizeof(double) * (size_t)std::max(nlmax, mmax) * b);
  thd.ensure(sizeof(double) * mmax);
  nrmd.ensure(sizeof(double) * mmax);
  auto setup = [&](LdlFactor& f) {
    f.sb = ldl_sb;
    f.ensure(N);
    f.use_graph = use_graph;
    f.panel_sub = ldl_sub;
    f.upd_xcd = ldl_xcd;
    f.upd_cp = ldl_cp;
    f.upd_wide = ldl_wide;
    f.tuner = &tuner;
  };
  setup(ldl);
  // with early_tol > 0 the factorisation of the next shift runs on s2 into
  // the second factor while the current slice converges on s
  const bool early_on = early_tol > 0.0;
  if (early_on) {
    setup(ldl2);
    if (!s2) {
      DVR_HIP(hipStreamCreate(&s2));
      DVR_HIP(hipEventCreateWithFlags(&ev_s, hipEventDisableTiming));
    }
  }
  LdlFactor* cur = &ldl;
  LdlFactor* nxt = &ldl2;
  thL.clear();
  nlocked = 0;
  first_pending = false;

  // start: unit vectors on the b smallest diagonal elements
  std::vector<double> dg(N);
  clock.mark(PH_START, s);
  Dd.ensure(sizeof(double) * N);
  launch_get_diag(s, H, N, Dd.as<double>());
  DVR_HIP(hipMemcpyAsync(dg.data(), Dd.p, sizeof(double) * N, hipMemcpyDeviceToHost, s));
  DVR_HIP(hipStreamSynchronize(s));
  std::vector<int> ord(N);
  std::iota(ord.begin(), ord.end(), 0);
  std::partial_sort(ord.begin(), ord.begin() + b, ord.end(),
                    [&](int x, int y) { return dg[x] < dg[y] || (dg[x] == dg[y] && x < y); });
  std::vector<int> idx0(ord.begin(), ord.begin() + b);
  {
    const int* di = upload(idxd, idx0, s);
    launch_unit_cols(s, V.as<double>(), N, di, b);
    launch_gather_cols(s, H, N, di, b, AV.as<double>(), N, N);
  }
  m = b;
  stats.matvecs += b;

  std::vector<double> th, nrm(mmax);
  auto rayleigh_ritz = [&]() {
    double* T = Tm.as<double>();
    gemm(s, "proj", true, false, m, m, N, 1.0, V.as<double>(), N, AV.as<double>(), N, 0.0, T, m,
         true);
    small_eig(m, T, Y.as<double>(), th, s);
    gemm(s, "ritz", false, false, N, m, m, 1.0, V.as<double>(), N, Y.as<double>(), m, 0.0,
         X.as<double>(), N);
    gemm(s, "ritz", false, false, N, m, m, 1.0, AV.as<double>(), N, Y.as<double>(), m, 0.0,
         AX.as<double>(), N);
    DVR_HIP(hipMemcpyAsync(thd.p, th.data(), sizeof(double) * m, hipMemcpyHostToDevice, s));
    launch_resid_norms(s, X.as<double>(), AX.as<double>(), N, thd.as<double>(), m, N,
                       nrmd.as<double>());
    DVR_HIP(hipMemcpyAsync(nrm.data(), nrmd.p, sizeof(double) * m, hipMemcpyDeviceToHost, s));
    DVR_HIP(hipStreamSynchronize(s));
  };

  // kind 3: shift-and-invert, with learned corrections in the iterations
  // that run while the first factorisation is on the second stream
  const bool si = P.prec == 0 || P.prec == 3;
  const bool learned = P.prec == 2 || P.prec == 3;
  if (learned) {
    // Dd already holds the diagonal
    if (!P.mlp || !P.st) throw HipError("davidson: learned preconditioner without parameters");
    Fd.ensure(sizeof(float) * (size_t)N * b * MLP_NF);
    Wd.ensure(sizeof(float) * (size_t)N * b);
  }
  // inertia of a factorisation as the solver sees it; with force_ambiguous
  // = k > 0 the k-th count of the solve is reported ambiguous (-1), to test
  // the retries at nudged shifts
  int ncount = 0;
  auto inertia = [&](int nu) {
    ++ncount;
    return (force_ambiguous > 0 && ncount == force_ambiguous) ? -1 : nu;
  };
  // first shift (shift-and-invert only): the lowest Ritz value of the start
  // block; the diagonal preconditioners factor only to certify slices
  if (si) {
    std::vector<double> w0;
    double* T = Tm.as<double>();
    gemm(s, "proj", true, false, m, m, N, 1.0, V.as<double>(), N, AV.as<double>(), N, 0.0, T, m,
         true);
    small_eig(m, T, nullptr, w0, s);
    sigma = w0[0];
    stats.factorizations++;
    if (first_overlap > 0 && early_on) {
      // factored on s2; until it is ready the corrections use the diagonal
      DVR_HIP(hipEventRecord(ev_s, s));
      DVR_HIP(hipStreamWaitEvent(s2, ev_s, 0));
      cur->factor_begin(H, sigma, s2);
      first_pending = true;
    } else {
      clock.mark(PH_FACTOR, s);
      // a zero or non-finite pivot (shift on an eigenvalue): the solves
      // would not be usable
      // a near-zero pivot (ambiguous inertia): the shift a little lower
      int nu0 = inertia(cur->factor(H, sigma, s));
      for (int t = 0; nu0 < 0 && t < 2; ++t) {
        sigma -= 1e-6 * (std::fabs(sigma) + 1.0);
        nu0 = inertia(cur->factor(H, sigma, s));
        stats.factorizations++;
      }
      if (nu0 < 0) return 3;
    }
  }
  // every later use of the first factor on s comes after this
  bool first_bad = false;
  auto first_done = [&]() {
    if (!first_pending) return;
    int nu0 = inertia(cur->factor_end(s2));
    // ambiguous inertia (near-zero pivot): the shift a little lower, on s
    for (int t = 0; nu0 < 0 && t < 2; ++t) {
      sigma -= 1e-6 * (std::fabs(sigma) + 1.0);
      nu0 = inertia(cur->factor(H, sigma, s));
      stats.factorizations++;
    }
    first_bad = nu0 < 0;
    first_pending = false;
  };

  for (;;) {
    int locked_slice = 0;
    double th_next = 0.0;
    bool have_next = false;
    // early mode: the next shift sig_e is fixed and factored on s2; the slice
    // then takes every level below it and ends when their number equals the
    // inertia nu_e
    bool early = false, launch_e = false;
    double sig_e = 0.0, gap_e = 0.0;
    int nu_e = -1;
    auto found_below = [&](double sg) {
      int f = 0;
      for (double t : thL) f += t < sg;
      return f;
    };
    stats.slices++;
    for (;;) {
      if (stats.iterations >= P.max_iter) return 2;   // stalled
      stats.iterations++;
      clock.mark(PH_RR, s);
      rayleigh_ritz();
      int nl = 0;
      if (early)
        while (nl < m && nrm[nl] <= tol && th[nl] < sig_e && nlocked + nl < nlmax) ++nl;
      else
        while (nl < m && nrm[nl] <= tol && locked_slice + nl < K && nlocked + nl < nlmax) ++nl;
      if (nl > 0) {
        DVR_HIP(hipMemcpyAsync(XL.as<double>() + (long)nlocked * N, X.p,
                               sizeof(double) * (size_t)N * nl, hipMemcpyDeviceToDevice, s));
        for (int j = 0; j < nl; ++j) thL.push_back(th[j]);
        nlocked += nl;
        locked_slice += nl;
      }
      bool end_slice = false;
      if (early) {
        // no Ritz value left below the next shift: compare with its inertia
        if (!(nl < m && th[nl] < sig_e) || nlocked >= nlmax) {
          if (nu_e < 0) {
            clock.mark(PH_FACTOR, s);
            nu_e = inertia(nxt->factor_end(s2));
            // a near-zero pivot makes the count ambiguous: the shift a
            // quarter of the gap higher, still between the same two Ritz
            // values as when it was placed, factored on s
            for (int t = 0; nu_e < 0 && t < 2; ++t) {
              sig_e += 0.25 * gap_e;
              nu_e = inertia(nxt->factor(H, sig_e, s));
              stats.factorizations++;
            }
            first_pending = false;   // s2 has drained
            // no trustworthy inertia at this shift: certification failed
            if (nu_e < 0) return 3;
          }
          // fewer levels than the inertia: one has not appeared yet, iterate on
          end_slice = found_below(sig_e) >= nu_e || nlocked >= nlmax;
        }
      } else if (locked_slice >= K || nlocked >= nlmax) {
        end_slice = true;
      } else if (early_on) {
        // the Ritz pairs up to one past the top of the slice near convergence:
        // the next shift goes half-way between the top level and the next
        const int r = std::min(K - locked_slice, nlmax - nlocked);
        if (nl + r < m) {
          bool ok = true;
          for (int j = nl; j <= nl + r; ++j) ok = ok && nrm[j] <= early_tol;
          if (ok) {
            sig_e = 0.5 * (th[nl + r - 1] + th[nl + r]);
            gap_e = th[nl + r] - th[nl + r - 1];
            early = launch_e = true;
          }
        }
      }
      if (end_slice) {
        const int rest = m - nl;
        if (rest > 0) {
          DVR_HIP(hipMemcpyAsync(V.p, X.as<double>() + (long)nl * N, sizeof(double) * (size_t)N * rest,
                                 hipMemcpyDeviceToDevice, s));
          DVR_HIP(hipMemcpyAsync(AV.p, AX.as<double>() + (long)nl * N,
                                 sizeof(double) * (size_t)N * rest, hipMemcpyDeviceToDevice, s));
          th_next = th[nl];
          have_next = true;
        }
        m = rest;
        break;
      }
      std::vector<int> act;
      for (int j = nl; j < m && (int)act.size() < b; ++j)
        if (!(nrm[j] <= tol)) act.push_back(j);
      // every Ritz pair left is converged but cannot be locked (the space is
      // exhausted, or a level below the next shift has not appeared): no
      // correction to expand with, the iteration has stalled
      if (act.empty()) return 2;
      int mk = m - nl;
      // full space: thick restart on the lowest Ritz vectors
      if (mk + (int)act.size() > mmax) mk = keep > 0 ? std::min(keep, mmax - (int)act.size())
                                                     : mmax - (int)act.size();
      const int na = (int)act.size();
      clock.mark(PH_SI, s);
      launch_resid_gather(s, X.as<double>(), AX.as<double>(), N, thd.as<double>(),
                          upload(idxd, act, s), na, Rb.as<double>(), N, N);
      // a fixed number of diagonal corrections while the first factorisation
      // runs (fixed, so that the iterates do not depend on timing)
      bool use_si = si;
      if (si && first_pending) {
        if (stats.iterations > first_overlap) first_done();
        else use_si = false;
      }
      if (use_si) {
        if (first_bad) return 3;
        cur->solve_graph(Rb.as<double>(), Tb.as<double>(), na, s);
      } else {
        std::vector<double> ta(na);
        for (int c = 0; c < na; ++c) ta[c] = th[act[c]];
        const double* thp = upload(tha, ta, s);
        const float* wp = nullptr;
        if (learned) {
          launch_mlp_features(s, Dd.as<double>(), P.st, N, na, thp, thp, Fd.as<float>());
          launch_mlp_infer_bf16(s, P.mlp, (long)N * na, Fd.as<float>(), Wd.as<float>());
          wp = Wd.as<float>();
        }
        launch_mlp_apply(s, wp, Dd.as<double>(), thp, Rb.as<double>(), N, na, Tb.as<double>());
      }
      stats.solves += na;
      // restart on the kept Ritz vectors (contiguous columns nl .. nl+mk)
      clock.mark(PH_EXPAND, s);
      if (mk > 0) {
        DVR_HIP(hipMemcpyAsync(V.p, X.as<double>() + (long)nl * N, sizeof(double) * (size_t)N * mk,
                               hipMemcpyDeviceToDevice, s));
        DVR_HIP(hipMemcpyAsync(AV.p, AX.as<double>() + (long)nl * N,
                               sizeof(double) * (size_t)N * mk, hipMemcpyDeviceToDevice, s));
      }
      m = mk;
      const int q = expand(H, na, s);
      if (q == 0) return 2;
      m += q;
      if (launch_e) {
        // the factorisation of the next shift, behind this iteration's work
        // on s (which also ends every earlier use of the second factor)
        DVR_HIP(hipEventRecord(ev_s, s));
        DVR_HIP(hipStreamWaitEvent(s2, ev_s, 0));
        nxt->factor_begin(H, sig_e, s2);
        stats.factorizations++;
        launch_e = false;
      }
    }
    // certify the levels found so far
    int nu = 0, found = 0;
    if (early) {
      found = found_below(sig_e);
      if (found == nu_e) {
        nu = nu_e;
        sigma = sig_e;
        std::swap(cur, nxt);
      } else {
        // the space of locked levels filled up before every level below
        // sig_e was found (a shift placed from rough Ritz values): certify
        // at the top of what was found instead, as without the early shift
        early = false;
      }
    }
    if (!early) {
      const double top = *std::max_element(thL.begin(), thL.end());
      double sig_hi = have_next ? 0.5 * (top + th_next)
                                      : top + 1e-6 * (std::fabs(top) + 1.0);
      clock.mark(PH_FACTOR, s);
      first_done();
      nu = inertia(cur->factor(H, sig_hi, s));
      // ambiguous inertia (near-zero pivot): a quarter of the gap higher
      for (int t = 0; nu < 0 && t < 2; ++t) {
        sig_hi += have_next ? 0.25 * (th_next - top) : 1e-6 * (std::fabs(top) + 1.0);
        nu = inertia(cur->factor(H, sig_hi, s));
        stats.factorizations++;
      }
      stats.factorizations++;
      if (nu < 0) return 3;
      sigma = sig_hi;
      found = found_below(sig_hi);
    }
    stats.certified = nu;
    if (found != nu) {
      stats.missing = nu - found;
      return 3;
    }
    if (nu >= nev) break;
    if (m == 0) {
      // basis exhausted: restart from fresh unit vectors is not attempted
      return 2;
    }
  }
  // ascending order of the locked levels
  perm.resize(nlocked);
  std::iota(perm.begin(), perm.end(), 0);
  std::sort(perm.begin(), perm.end(), [&](int x, int y) { return thL[x] < thL[y]; });
  eval.resize(nev);
  for (int k = 0; k < nev; ++k) eval[k] = thL[perm[k]];
  return 0;
}

void DavidsonSolver::vectors(int nvec, double* Z, hipStream_t s) {
  std::vector<int> p(perm.begin(), perm.begin() + nvec);
  launch_gather_cols(s, XL.as<double>(), N, upload(idxd, p, s), nvec, Z, N, N);
}

}  // namespace dvr
