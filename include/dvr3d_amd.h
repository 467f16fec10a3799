/*
 * dvr3d_amd.h -- C ABI of the MI355X (gfx950) implementation of the DVR3DRJZ
 * theta-first vibrational Hamiltonian path.
 *
 * The reference program (ExoMol DVR3D, source/dvr3drjz_segmented.f90, cited
 * below as D:<line>) has no foreign-function interface; its replaceable units
 * are Fortran subroutines that share COMMON blocks and scratch files.  This
 * ABI exposes the same units with plain pointers and sizes so that a Fortran
 * host binds them with ISO_C_BINDING:
 *
 *   dvr_diag            <- SUBROUTINE diag(ham,maxham,nham,eig)     D:2638-2659
 *   dvr_ctx_* + dvr_run <- the body of SUBROUTINE jhmain            D:1915-2118
 *                          (mkham1 D:2121, diag/cut1d D:1987/2870, mkham2
 *                          D:2228, choose/cut2d D:2767/2941, mkham3 D:2301,
 *                          loadh D:2554, diag3d D:2661, trans D:3650)
 *
 * Conventions: 32-bit ints, IEEE double; matrices column-major like the
 * Fortran caller; errors are returned as an int (0 = success) and the message
 * is available from dvr_last_error(); no C++ exception crosses the ABI.  One
 * context owns its device memory and one HIP stream; calls on one context must
 * not overlap.
 */
#ifndef DVR3D_AMD_H
#define DVR3D_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define DVR_OK 0
#define DVR_ERR_ARG 1       /* invalid argument or call order */
#define DVR_ERR_HIP 2       /* HIP runtime error (message has details) */
#define DVR_ERR_CORE_2D 3   /* "core exceeded: reduce cut-off emax1" (D:2923-2929) */
#define DVR_ERR_CORE_3D 4   /* "core exceeded: reduce cut-off emax2" (D:2981-2985) */
#define DVR_ERR_NOCONV 5    /* iterative solver did not converge */

/* final-solve methods */
#define DVR_SOLVER_DENSE 0     /* dense FP64 eigensolver (plain form) */
#define DVR_SOLVER_DAVIDSON 1  /* block-Davidson, dense fallback (chosen form) */

/* stage timers reported by dvr_get_stage_times (milliseconds) */
#define DVR_STAGE_1D 0       /* 1D build + Jacobi + cut            */
#define DVR_STAGE_2D_BUILD 1 /* packing + mkham2 SYRK             */
#define DVR_STAGE_2D_EIG 2   /* 2D eigenvalues (tridiag+bisect)   */
#define DVR_STAGE_2D_VEC 3   /* chosen 2D eigenvectors            */
#define DVR_STAGE_3D_BUILD 4 /* cint + mkham3 SYRK + loadh        */
#define DVR_STAGE_3D_EIG 5   /* final eigenvalues                 */
#define DVR_STAGE_3D_VEC 6   /* final eigenvectors + trans        */
#define DVR_NSTAGE 7

typedef struct dvr_ctx dvr_ctx_t;

/*
 * Problem dimensions after the reference's insize clamps (D:232-355).
 * theta-first order (ztheta=T): a = angle (npnta = nalf/max(1,idia)),
 * b = the second coordinate and c = the last one; zr2r1=1 means b = r2 and
 * c = r1 (D:310-318).
 */
typedef struct dvr_dims {
  int32_t npnta, npntb, npntc;
  int32_t max2d, max3d, neval;
  int32_t zcut;   /* 1: final basis by energy cut emax2 (cut2d), 0: lowest max3d (choose) */
  int32_t zall;   /* 1: no truncation of intermediate solutions */
  int32_t zr2r1;
  int32_t idvr, npnt1, npnt2; /* wavefunction record layout psi(idvr,npnt1,npnt2) */
  int32_t nvec;   /* number of lowest levels whose wavefunctions are wanted (0..neval) */
  int32_t solver; /* DVR_SOLVER_* */
} dvr_dims_t;

/* Create / destroy a context on HIP device `device`. */
int dvr_ctx_create(const dvr_dims_t* dims, int32_t device, dvr_ctx_t** out);
void dvr_ctx_destroy(dvr_ctx_t* ctx);
const char* dvr_last_error(const dvr_ctx_t* ctx);

/*
 * Host-computed basis arrays for one k block (copied; caller keeps ownership).
 *   xlmatr [npnta*npnta]  angular L matrix (lmatrx, D:1846-1911)
 *   xkb    [npntb*npntb]  DVR kinetic matrix of coordinate b (k1k2, D:1603)
 *   xkc    [npntc*npntc]  DVR kinetic matrix of coordinate c
 *   wsum   [npntc*npntb]  (w1+w2) at (c,b), index c*npntb + b (mkham1 D:2155-2157)
 *   wterm  [npntc*npntb]  rotational diagonal shift, or NULL for J=0 (D:2159-2173)
 *   emax1, emax2          cut-off energies in hartree (EMAX1/EMAX2 card / autocm)
 */
int dvr_set_basis(dvr_ctx_t* ctx, const double* xlmatr, const double* xkb,
                  const double* xkc, const double* wsum, const double* wterm,
                  double emax1, double emax2);

/* Potential on the grid, V[(c*npntb + b)*npnta + a] in hartree, as returned by
 * the user's POTV(V,R1,R2,XCOS) at the grid points (mkham1, D:2178). */
int dvr_set_potential(dvr_ctx_t* ctx, const double* v);

/*
 * Evaluate the analytic AB2 surface on the device instead:
 *   V = De[(1-exp(-a(r1-re)))^2 + (1-exp(-a(r2-re)))^2]
 *       + kb (cos(theta) - cte)^2 + c (r1-re)(r2-re)
 * rc[npntc], rb[npntb] are the radial points of coordinates c and b, xa[npnta]
 * the angular points; par = {De, a, re, kb, cte, c}.
 */
int dvr_eval_potential_ab2(dvr_ctx_t* ctx, const double* rc, const double* rb,
                           const double* xa, const double* par);

/* Run 1D -> 2D -> selection -> 3D -> final solve (-> wavefunctions). */
int dvr_run(dvr_ctx_t* ctx);

/* dvr_run on n contexts at once (the k blocks of a J > 0 run with zrot:
 * the ccmain k loop, source/dvr3drjz_segmented.f90:857-1007, which calls
 * jhmain once per k).  Every context must be distinct (DVR_ERR_ARG
 * otherwise, before anything runs) and ready to run; they execute
 * concurrently on their own streams.  Returns DVR_OK, or the status of the
 * first context (in array order) that failed; each context keeps its own
 * error text. */
int dvr_run_blocks(dvr_ctx_t* const* ctxs, int n);

/* Truncation results: iv1 as the reference's iv1(ndimc,ndimb) (column-major,
 * iv1[c + npntc*b]), iv2[npntc], nham3 (D:1997, D:2053, D:2086). */
int dvr_get_sizes(const dvr_ctx_t* ctx, int32_t* iv1, int32_t* iv2,
                  int32_t* nham3);

/* Lowest meval = min(neval, nham3) eigenvalues in hartree, ascending. */
int dvr_get_eigenvalues(const dvr_ctx_t* ctx, double* eval, int32_t* meval);

/* All 1D eigenvalues e1[(c*npntb + b)*npnta + k] (ascending in k). */
int dvr_get_1d_eigenvalues(const dvr_ctx_t* ctx, double* e1);

/* All 2D eigenvalues of every c block, concatenated in c order (sum of nham2). */
int dvr_get_2d_eigenvalues(const dvr_ctx_t* ctx, double* e2, int32_t* total);

/* Wavefunction amplitudes of the nvec lowest levels, each level a record
 * psi(idvr,npnt1,npnt2) as written to IWAVE by trans (D:3841). */
int dvr_get_wavefunctions(const dvr_ctx_t* ctx, double* psi);

/* Final eigenvectors in the contracted basis, nham3 x nvec column-major. */
int dvr_get_vectors(const dvr_ctx_t* ctx, double* vec);

/* Per-stage device times of the last dvr_run, DVR_NSTAGE doubles (ms). */
int dvr_get_stage_times(const dvr_ctx_t* ctx, double* ms);

/* Statistics of the last final solve (no counterpart in the reference, which
 * only has the dense dsyev call).  Fills up to n of: [0] Davidson status
 * (-1 not run, 0 certified, 2 stalled, 3 missed level or no trustworthy
 * inertia at a shift, 4 error inside the iterative solver, 5 too small a
 * problem for the iteration: N3 < NEVAL + 1 + 3 x 64), [1] iterations,
 * [2] H-vector products, [3] shift-and-invert solves, [4] LDL^T
 * factorisations, [5] slices, [6] inertia count at the last shift,
 * [7] levels missing at the failed certification, [8] 1 when the dense
 * solver was used as fallback; device time (microseconds, HIP events on the
 * solver's stream) of [9] the Rayleigh-Ritz steps, [10] the shift-and-invert
 * solves, [11] the basis expansions (orthogonalisation and H-vector
 * products), [12] the LDL^T factorisations (a factorisation run on the
 * second stream beside a slice counts only for the time the solver's stream
 * waits for it), [13] the start (diagonal, start block, first projection),
 * [14] 1 when the dense eigenvalues came from the two-stage reduction. */
int dvr_get_solver_stats(const dvr_ctx_t* ctx, int32_t* stats, int32_t n);

/* Shift-and-invert kernel of the Davidson solver, exposed for testing and
 * for callers that need eigenvalue counts (no counterpart in the reference).
 * Factors A - sigma I = L D L^T (A: n x n column-major, full square, host
 * memory), returns in *inertia the number of negative pivots (= eigenvalues
 * of A below sigma by Sylvester's law) and, for nrhs > 0, solves
 * (A - sigma I) t = r for the nrhs columns of r (n x nrhs) into t.  The
 * factorisation is unpivoted: when a pivot is not finite or below
 * n eps max|d| the count is ambiguous, *inertia is set to -1 and
 * DVR_ERR_NOCONV is returned (nothing is solved). */
int dvr_ldl_solve(const double* a, int32_t n, double sigma, const double* r, int32_t nrhs,
                  double* t, int32_t* inertia);

/* ---- learned Davidson preconditioner (no counterpart in the reference) ----
 * A 3-layer perceptron (6 features -> 128 -> 128 -> 1, tanh) gives per basis
 * function j of a Ritz pair (theta, x, r) the weight w_j of the correction
 * t_j = w_j r_j / (D_j - theta), D = diag H3.  Parameters: DVR_MLP_NPARAM
 * floats, layout W1[f*128+o], b1[128], W2[k*128+o], b2[128], W3[128], b3. */
#define DVR_MLP_NPARAM 17537
#define DVR_PREC_SHIFT_INVERT 0  /* (H - sigma I)^-1 r, LDL^T per slice (default) */
#define DVR_PREC_DIAGONAL 1      /* r / (D - theta)                             */
#define DVR_PREC_LEARNED 2       /* w(features) r / (D - theta), BF16 inference */
#define DVR_PREC_SI_LEARNED 3    /* shift-and-invert; the first DVR_DAV_FIRST (default
                                    6) iterations run the learned correction while the
                                    first shift is factored on a second stream */

/* Keep a copy of H3 after the next dvr_run (needed by dvr_mlp_train). */
int dvr_keep_hamiltonian(dvr_ctx_t* ctx, int32_t on);

/* Correction of the Davidson solver (solver DVR_SOLVER_DAVIDSON); params
 * (DVR_MLP_NPARAM floats, copied) for DVR_PREC_LEARNED and
 * DVR_PREC_SI_LEARNED. */
int dvr_set_preconditioner(dvr_ctx_t* ctx, int32_t kind, const float* params);

/* Trains params in place (Adam, steps x batch training pairs) on the
 * eigenpairs of the last dense run (nvec >= 2, dvr_keep_hamiltonian set);
 * loss_hist (optional) receives the mean loss of every step. */
int dvr_mlp_train(dvr_ctx_t* ctx, int32_t steps, int32_t batch, uint64_t seed, float lr,
                  float* params, double* loss_hist);

/* Test entry points: network output for nrows feature rows (6 floats each),
 * FP32 (bf16 = 0) or BF16-MFMA (bf16 = 1); gradients of sum(gw * w) with
 * respect to the parameters (FP32). */
int dvr_mlp_forward(const float* params, int64_t nrows, const float* feats, float* w,
                    int32_t bf16);
int dvr_mlp_backward(const float* params, int64_t nrows, const float* feats, const float* gw,
                     float* grad);

/*
 * Kernel-level timing (no counterpart in the reference, which only prints
 * the total run time, D:67/D:75-78).  With profiling on, the next dvr_run
 * brackets every kernel launch with HIP events on the context's stream and
 * accumulates, per "<stage>:<kernel>" name, the launch count, device
 * seconds and the launches' algorithmic bytes and flops.  Entry idx =
 * 0, 1, ... returns DVR_ERR_ARG past the last kernel name.
 */
int dvr_set_profiling(dvr_ctx_t* ctx, int32_t on);
int dvr_get_profile(const dvr_ctx_t* ctx, int32_t idx, char* name, int32_t name_len,
                    int32_t* launches, double* seconds, double* bytes, double* flops);

/*
 * Symmetrised-Radau driver (IDIA = -2, AB2 molecules): the body of
 * SUBROUTINE nfmain(hr,htheta,r,theta,kz) (D:3006-3122) for one k block in
 * one call -- every angle's symmetrised radial Hamiltonian (blc2d1,
 * D:3125), their eigenpairs (diag), the selection (choosr D:3223 or cut2dr
 * D:3323, then testiv D:3380), the 3D Hamiltonian (bloc3d/vecmul D:3455),
 * the final eigensolve (diag3d) and transr's wavefunction records (D:3599).
 *   hr     [npnt*npnt]  radial DVR kinetic matrix (k1k2 of r1, D:1603)
 *   htheta [nalf*nalf]  angular L matrix (lmatrx, D:1846), full range
 *   r [npnt], xalf [nalf]  radial points, angular points (cos theta)
 *   ur1, urr1           vibrational and rotational reduced masses (setcon)
 *   vgrid  [nalf*npnt*npnt]  V[(g*npnt + b)*npnt + a] = POTV(r_a, r_b, x_g)
 *   emax2               cut energy in hartree (zcut = 1)
 * Outputs: iv2[nalf] 2D functions kept per angle, *nham3, *iang (angles
 * with functions), eval[min(neval, nham3)] ascending (hartree) and, for
 * nvec > 0, wvfunc: nvec records of npair*iang amplitudes (radial function
 * outer, angle inner), npair = npnt(npnt+1-2 ipar)/2, as transr writes them.
 * Returns DVR_ERR_CORE_3D when cut2dr keeps more than max3d functions.
 * J = 0 only (jrot != 0 returns DVR_ERR_ARG): the J > 0 terms (blc2d1's
 * rotational diagonal, testiv) are not yet checked against a reference run.
 */
typedef struct dvr_nf_dims {
  int32_t npnt;   /* radial points of each bond (npnt1 = npnt2, D:248) */
  int32_t nalf;   /* angular points (full range) */
  int32_t ipar;   /* 0: even, 1: odd symmetrised radial functions */
  int32_t max3d, neval;
  int32_t zcut;   /* 1: 2D levels below emax2 (cut2dr); 0: the lowest max3d (choosr) */
  int32_t nvec;   /* levels whose wavefunction records are wanted (0..neval) */
  int32_t jrot, kz, kmin, zrot, zlin;
} dvr_nf_dims_t;
int dvr_nfmain(const dvr_nf_dims_t* dims, const double* hr, const double* htheta,
               const double* r, const double* xalf, double ur1, double urr1,
               const double* vgrid, double emax2, int32_t* iv2, int32_t* nham3,
               int32_t* iang, double* eval, double* wvfunc);

/*
 * The rotational step ROTLEV3 (source/rotlev3.f90, cited R:<line>): the body
 * of VRMAIN (R:482-629) for one run -- JTRAN of every k block's vibrational
 * vectors (R:839-904), the Coriolis blocks of SOLRT (R:671-836), the dense
 * solve of DGROT (R:1046-1197, DSYEVX 'I': lowest NEVAL levels) for the first
 * parity set over all blocks and, with nsets = 2 (KMIN > 1), for the second
 * over blocks 1..nblk-1, and the vectors of DSTORE (R:1297-1323).
 * The caller reads the IWAVE file and selects (SELECT, R:301-478); block b
 * (b = 0 .. nblk-1) has k = kfirst + b and keeps its mvib lowest functions.
 *   pleg    concatenated per block: (maxleg+1) x idvr, the pleg*w record
 *   dvrvec  concatenated per block: mvib vectors of idvr*npnt1*npnt2
 *           amplitudes (angle fastest), as in IWAVE
 *   evib    concatenated per block: the mvib levels (hartree)
 *   rm2     the r**-2 record of IWAVE (nlim values)
 * Outputs: jst_nang[2 b], [2 b + 1] = first l and count of the associated
 * Legendre functions of block b (JTRAN R:879-885; may be NULL); eval:
 * neval[0] then neval[1] levels; with nvec != 0, vec: the eigenvectors of
 * each set (ibass x neval, ibass = the set's function count) one set after
 * the other, and dvec: per set, per block of the set, nkbas x neval with
 * nkbas = nang * npnt1 * npnt2 (the DSTORE record of that block).
 */
typedef struct dvr_rot_dims {
  int32_t nblk;       /* k blocks */
  int32_t kfirst;     /* k of block 0: 0 (KMIN >= 1) or 1 (KMIN = 0) */
  int32_t jrot, idia, ipar;
  int32_t npnt1, npnt2;
  int32_t zembed, zquad2;
  int32_t nlim;       /* length of rm2 */
  int32_t nsets;      /* 1, or 2 for KMIN > 1 */
  int32_t neval[2];   /* levels per set */
  int32_t nvec;       /* 1: eigenvectors and DSTORE vectors wanted */
} dvr_rot_dims_t;
typedef struct dvr_rot_block {
  int32_t maxleg, idvr, lincr;   /* the block's (kz, maxleg, idvr, lincr) record */
  int32_t mvib;                  /* vibrational functions kept */
} dvr_rot_block_t;
int dvr_rotlev3(const dvr_rot_dims_t* dims, const dvr_rot_block_t* blocks, const double* pleg,
                const double* dvrvec, const double* evib, const double* rm2, int32_t* jst_nang,
                double* eval, double* vec, double* dvec);

/*
 * The transition moments of DIPOLE3 (source/dipole3_segmented.f90, cited as
 * P:<line>), replacing the body of DMAIN (P:708-1115): DSRD's evaluation of
 * every k block's wavefunctions at the angular integration points
 * (P:1511-1548; for a J = 0 IWAVE file also JTRAN, P:1564-1645) and TRANS's
 * sums over the radial and angular points (P:1648-1776, rank-one DGER
 * updates in the reference).
 *
 * Block b of each side (ket: nblk1 blocks, bra: nblk2) gives
 *   vec  ne vectors of nin*nrad values, vector ie at vec + ie*nin*nrad,
 *        index a + nin*mn (a = angular function or DVR point, mn = radial
 *        point, r1 fastest): the DSTORE record of a KVEC file, or the IWAVE
 *        wavefunction records of a J = 0 run
 *   ang  ipot x nin (column-major): input function a at integration point p,
 *        i.e. the normalised associated Legendre values of ASLEG for FBR
 *        input, or those times JTRAN's P_sel for DVR input
 * Its values at the points are d_b[p + ipot*mn, ie] = sum_a ang[p, a]
 * vec[a + nin*mn, ie].  dipol holds 2 * nrad*ipot values, the z component
 * (nu = 0) then the x component (nu = +-1) of LAGPT (P:1206-1230: the
 * surface times the quadrature weights, doubled for IDIA = 2), index
 * p + ipot*mn.  Each pair (b1, b2, comp, factor) adds
 *   t[ie1, ie2] += factor * sum_i dipol[comp][i] d1_b1[i, ie1] d2_b2[i, ie2]
 * to tz (comp 0) or tx (comp 1), both ne1 x ne2 column-major, accumulated
 * in place (zero them for a fresh run; a restart passes the stored sums).
 * factor = threej * xfac * (-1)^kk1 as TRANS forms it (P:1699-1700); pairs
 * with factor 0 are skipped.  Blocks no pair uses are never read.  vec, ang
 * and dipol may be host memory (copied in) or device memory of the current
 * device (read in place: a caller that already holds the vectors on the GPU,
 * e.g. from dvr_rotlev3's products, moves nothing over PCIe); tz and tx are
 * host memory.  Runs on a stream of its own and returns when done.
 */
typedef struct dvr_dip_dims {
  int32_t nrad;       /* radial points of a block, npnt1*npnt2 */
  int32_t ipot;       /* angular integration points */
  int32_t ne1, ne2;   /* ket and bra levels */
  int32_t nblk1, nblk2;
  int32_t npair;
} dvr_dip_dims_t;
typedef struct dvr_dip_block {
  int32_t nin;        /* input functions per radial point */
  const double* vec;
  const double* ang;
} dvr_dip_block_t;
typedef struct dvr_dip_pair {
  int32_t b1, b2;     /* ket block, bra block (0-based) */
  int32_t comp;       /* 0: z (tz), 1: x (tx) */
  double factor;
} dvr_dip_pair_t;
/* Device times of the call (HIP events on its stream; stats may be NULL):
 * the blocks' evaluation at the points, the dipole scaling and the moment
 * products (split-K reduction included), with their launch counts. */
typedef struct dvr_dip_stats {
  double transform_ms, scale_ms, trans_ms;
  int32_t ntransform, ntrans;
} dvr_dip_stats_t;
int dvr_dipole3(const dvr_dip_dims_t* dims, const dvr_dip_block_t* ket,
                const dvr_dip_block_t* bra, const double* dipol, const dvr_dip_pair_t* pairs,
                double* tz, double* tx, dvr_dip_stats_t* stats);

/*
 * The expectation values of XPECT3 (source/xpect3.f90, cited as X:<line>),
 * replacing the sums of X3MAIN (X:436-520): DSRD/REST's evaluation of each k
 * block at the quadrature points (X:763-925) and TRANS's grid sums
 * (X:998-1082):
 *   out[i + ne*n] += sum_g prop[g + npts*n] * d_b[g, i]^2  over every block b
 * (out: ne x nprop column-major, accumulated in place).  Block b gives
 *   nin = 0:  vec holds each level's values at the npts grid points already
 *             (the DVR amplitudes of a J = 0 IWAVE file: jhmain's
 *             psi(idvr,npnt1,npnt2), or nfmain's kept angles per radial
 *             pair), level i at vec + i*ldv, ldv >= npts;
 *   nin > 0:  vec holds ne vectors of nin*nrad FBR coefficients (ldv =
 *             nin*nrad, the DSTORE record of a KVEC block), ang the npot x nin
 *             column-major values sqrt(w_p) P_l(x_p) of REST, and
 *             d_b[p + npot*mn, i] = sum_a ang[p, a] vec[a + nin*mn + ldv*i]
 *             with npts = nrad*npot;
 *   vec = NULL: an empty block (skipped).
 * prop holds the npts x nprop property values on the same grid (PROP at the
 * points, averaged over the homonuclear exchange as X:470-500, 634-646 do).
 * vec, ang and prop may be host or device memory; out is host memory.
 */
typedef struct dvr_xp_dims {
  int32_t npts;       /* grid points of a block */
  int32_t npot;       /* quadrature points per radial point (nin > 0 blocks) */
  int32_t ne;         /* levels */
  int32_t nblk;
  int32_t nprop;
} dvr_xp_dims_t;
typedef struct dvr_xp_block {
  int32_t nin;
  int64_t ldv;
  const double* vec;
  const double* ang;
} dvr_xp_block_t;
int dvr_xpect3(const dvr_xp_dims_t* dims, const dvr_xp_block_t* blocks, const double* prop,
               double* out);

/*
 * Drop-in for the reference's SUBROUTINE diag(ham,maxham,nham,eig)
 * (D:2638-2659: dsyev('V','L',nham,ham,maxham,eig,...)).  On entry the lower
 * triangle of the column-major ham(maxham,nham) holds the symmetric matrix; on
 * return columns 1..nham hold orthonormal eigenvectors and eig(nham) the
 * eigenvalues in ascending order.  Runs on the current HIP device.
 */
int dvr_diag(double* ham, int32_t maxham, int32_t nham, double* eig);

/*
 * Batched form used by the pipeline, exposed for callers with several
 * independent problems: nprob symmetric matrices of orders n[p], stored one
 * after another (each n[p]*n[p], column-major, lower triangle read).
 * Returns the lowest nvec[p] eigenvalues of each (w, concatenated) and, if z
 * is not NULL, their eigenvectors (n[p] x nvec[p] each, concatenated).
 */
int dvr_syev_batch(int32_t nprob, const int32_t* n, const double* a,
                   const int32_t* nvec, double* w, double* z);

/*
 * dvr_syev_batch without any device allocation (no counterpart in the
 * reference): for callers that cannot allocate device memory -- code under
 * HIP graph capture, tight solve loops, harnesses that forbid hipMalloc.
 * Same arguments and results as dvr_syev_batch; every order n[p] must be
 * <= 128 (the problems are solved in LDS), and the device workspace of the
 * call, sum over runs of consecutive problems of one order n with up to m
 * wanted vectors of nb (7 n^2 + n + n m) doubles, must fit in a static array
 * of 4 Mi doubles (32 MiB) that each GPU holds until the process ends;
 * larger requests return DVR_ERR_ARG.  Runs on the null stream of the
 * current device.  Calls that overlap on one GPU share that array and race.
 */
int dvr_syev_batch_noalloc(int32_t nprob, const int32_t* n, const double* a,
                           const int32_t* nvec, double* w, double* z);

#ifdef __cplusplus
}
#endif

#endif /* DVR3D_AMD_H */
