/*
 * stda_hip.h -- C ABI of the MI355X excited-state core (sTDA / sTD-DFT /
 * linear-response polarizability).
 *
 * This is the drop-in boundary for the block of the reference's RKS driver
 * `stda` (stda.f) that runs from the Loewdin orbitals to the spectrum:
 * stda.f:454-1471, with the solvers of srpapack.f:32-197 and
 * linear_response.f:23-369.  The caller (the Fortran driver through an
 * ISO_C_BINDING module, or the Python host layer through ctypes) hands over
 * the FP64 quantities the reference holds at stda.f:452 and receives FP64
 * results.  Plain pointers and sizes only; all arrays are column-major
 * (Fortran order) and owned by the caller unless stated otherwise.
 *
 * Every function returns an int status: STDA_HIP_OK (0) or one of the error
 * codes below; stda_hip_error_string() gives the message the reference
 * prints with `stop` in the same situation.
 *
 * Threading: a context is used by one host thread at a time.  Calls on
 * different contexts may run concurrently; each context owns its own HIP
 * stream and device memory.
 */
#ifndef STDA_HIP_H
#define STDA_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Index type of the arrays the reference's driver holds as default Fortran
 * integers: iaoat, iconf and the -dual atom list.  32-bit in the default build
 * (libstda_hip.so); 64-bit in the build compiled with -DSTDA_HIP_ILP64
 * (libstda_hip_ilp64.so), which pairs with std2 built with 8-byte default
 * integers (meson_options.txt:37-43, meson.build:62-65; lin8 at
 * stda.f:2183-2190).  Scalars keep their fixed C types in both builds. */
#ifdef STDA_HIP_ILP64
typedef int64_t stda_hip_int;
#else
typedef int32_t stda_hip_int;
#endif

enum {
    STDA_HIP_OK = 0,
    STDA_HIP_ERR_ARG = 1,          /* invalid argument / dimension */
    STDA_HIP_ERR_NO_CSF = 2,       /* 'no CSF, increase energy threshold (-e option)' stda.f:316,644 */
    STDA_HIP_ERR_DEVICE = 3,       /* HIP runtime or library failure */
    STDA_HIP_ERR_NOMEM = 4,        /* device allocation failed, or would leave less than the
                                      runtime's reserve free (4 GiB, STDA_HIP_DEVICE_RESERVE_MB) */
    STDA_HIP_ERR_STATE = 5,        /* call out of order (e.g. solve before setup) */
    STDA_HIP_ERR_NOT_CONVERGED = 6,/* iterative solver hit its iteration limit */
    STDA_HIP_ERR_AMB_NOT_PD = 7,   /* 'matrix power impossible' srpapack.f:229 (A-B not positive definite) */
    STDA_HIP_ERR_NO_ROOTS = 8,     /* 'internal error in diag' stda.f:1152 / 'RPA diag failed' srpapack.f:92 */
    STDA_HIP_ERR_TOO_LARGE = 9     /* problem exceeds a documented size limit of an entry point */
};

/* Moments: order of the 9 packed arrays (each moci*(moci+1)/2 FP64,
 * element lin(i,j) = min + max*(max-1)/2 (1-based) holding M(min,max)):
 * xl yl zl (dipole length), xm ym zm (angular momentum), xv yv zv (velocity),
 * as produced by onetri+shrink at stda.f:346-412. */

typedef struct {
    int32_t ncent;          /* atoms */
    int32_t nao;            /* AOs */
    int32_t no, nv;         /* window occupied / virtual MOs (moci = no + nv) */
    const double *clow;     /* clow(nao, moci): Loewdin MOs S^1/2 C, stda.f:436-452 */
    const double *epsi;     /* epsi(moci): orbital energies (Hartree), ascending */
    const stda_hip_int *iaoat; /* iaoat(nao): 1-based atom of each AO (intslvm.f:66-77) */
    const double *xyz;      /* xyz(4, ncent): Bohr; xyz(4,:) = nuclear charge */
    const double *moments;  /* 9 * moci*(moci+1)/2, see above; may be NULL (no properties) */
} stda_hip_system;

typedef struct {
    double thr_ev;          /* Ethr: spectral range (eV), main.f:87 default 7 */
    double thrp;            /* PT2 selection threshold (Hartree), main.f:93 default 1e-4 */
    double fthr_ev;         /* max CSF selection range (eV); <= 0 means unlimited (main.f:86) */
    double ax;              /* Fock-exchange fraction a_x */
    double alphak, betaj;   /* gamma exponents; < -99 selects 1.42+0.48ax / 0.20+1.83ax (stda.f:490-500) */
    int32_t triplet;        /* ak = 0 instead of 2 (stda.f:482-483) */
    int32_t kshift;         /* K(ia,ia)-dependent diagonal shift (xTB mode, stda.f:1750-1768) */
    double kshift_max_ev, kshift_width_ev, kshift_steep;   /* defaults 0.5, 0.1, 4.0 */
} stda_hip_params;

typedef struct stda_hip_ctx stda_hip_ctx;

/* Bytes of stda_hip_int in this build (4 or 8): a caller checks that it was
 * linked against the library of its integer size.  No context needed. */
int stda_hip_index_size(void);

/* Create a context on HIP device `device` (-1: current device). */
int stda_hip_create(stda_hip_ctx **ctx, int device);
int stda_hip_destroy(stda_hip_ctx *ctx);
const char *stda_hip_error_string(int status);
/* Message of the last error raised on this context (more detail than the status). */
const char *stda_hip_last_error(const stda_hip_ctx *ctx);

/*
 * Stage 1 (replaces stda.f:454-790): gamma^J/gamma^K (stda.f:509-532),
 * Loewdin transition charges (lo12pop, stda.f:465-763), contracted
 * intermediates (stda.f:555-589,742), CSF screening and ordering
 * (stda.f:593-668) and PT2 selection (ptselect, stda.f:1773-1863).
 * Device-resident afterwards: charges, gamma, CSF list, diagonal.
 * Out: nci = primary + PT2-selected CSFs; nprim = primaries (nex);
 * nsec = secondaries considered in PT2 (nexpt).
 */
int stda_hip_setup(stda_hip_ctx *ctx, const stda_hip_system *sys, const stda_hip_params *par,
                   int64_t *nci, int64_t *nprim, int64_t *nsec);

/* stda_hip_setup in two halves: load copies the system to the device (host
 * -> HBM transfers, AO reordering on the host) and select runs the device
 * pipeline on the resident data.  select may be called again (same result). */
int stda_hip_load(stda_hip_ctx *ctx, const stda_hip_system *sys, const stda_hip_params *par);
int stda_hip_select(stda_hip_ctx *ctx, int64_t *nci, int64_t *nprim, int64_t *nsec);

/* Axis-resolved transition properties of TDA states (std2 -aniso: tdax.dat,
 * tday.dat, tdaz.dat; stda.f:1360-1376 and, velocity-corrected,
 * apbtrafo.f:245-262).  w (k) excitation energies, x (nci, k; leading
 * dimension ldx) their vectors.  out(4,3,k): for axis q = x, y, z the
 * fL, fV, RL and RV (Rv(corr) when velcorr) of that axis, scaled as the
 * reference prints them: f = (f_x + f_y + f_z)/3 and R = R_x + R_y + R_z. */
int stda_hip_props_xyz(stda_hip_ctx *ctx, int32_t k, const double *w, const double *x,
                       int64_t ldx, int32_t velcorr, double *out);

/* One-electron AO integrals over contracted Cartesian Gaussians, one thread
 * per AO pair (replaces the libcint calls of the reference's integral driver:
 * overlap; dipole length -<mu|r|nu>, libcint.f:252-352; velocity
 * -<mu|d/dr|nu>, libcint.f:653-756; angular momentum <mu|r x d/dr|nu>,
 * intpack.f90:1208-1283; origin 0).  Per AO m: primitives
 * [prim_ptr[m], prim_ptr[m+1]) with exponents alpha and contraction
 * coefficients coef (normalisation included), centre center[3m..3m+2] (Bohr),
 * Cartesian exponents lmn[3m..3m+2] (each <= 4).  Out (column-major, nao x
 * nao): S and ao9 = dipole x, y, z, angular momentum x, y, z, velocity
 * x, y, z (the order stda_hip_transform takes). */
int stda_hip_ao_ints(stda_hip_ctx *ctx, int32_t nao, const int64_t *prim_ptr, const double *alpha,
                     const double *coef, const double *center, const int32_t *lmn, double *S,
                     double *ao9);

/* The upstream transforms before the boundary, on the device: the nine MO
 * moment arrays moments(npk, 9) = shrink(ca^T A_q ca) (onetri.f:20-55,
 * stda.f:351-412, shrink stda.f:1695-1706; A_q = ao9[q]: dipole x, y, z,
 * angular momentum x, y, z, velocity x, y, z; packed as the reference's
 * lin(i,j), npk = moci (moci+1)/2) and the Loewdin orbitals clow = S^1/2 ca
 * (makel, stda.f:1615-1691, 430-452).  ca (nao, moci), S (nao, nao),
 * ao9 (nao, nao, 9), clow (nao, moci), all column-major host arrays.
 * A negative eigenvalue of S stops as the reference does unless bfw (-BFW:
 * set to zero); |<c'|c'> - 1| > 1e-2 is the reference's MO norm error.
 * Needs a context (its stream and BLAS handle), not a loaded system. */
int stda_hip_transform(stda_hip_ctx *ctx, int32_t nao, int32_t moci, const double *ca,
                       const double *S, const double *ao9, int32_t bfw, double *clow,
                       double *moments);

/* stda_hip_ao_ints followed by stda_hip_transform with the AO matrices kept on
 * the device: the basis arrays as for stda_hip_ao_ints, then moci, ca, bfw,
 * clow and moments as for stda_hip_transform. */
int stda_hip_prepare_ao(stda_hip_ctx *ctx, int32_t nao, const int64_t *prim_ptr,
                        const double *alpha, const double *coef, const double *center,
                        const int32_t *lmn, int32_t moci, const double *ca, int32_t bfw,
                        double *clow, double *moments);

/* Selection variants, applied by the next stda_hip_select / stda_hip_setup:
 *   natoms > 0: dual thresholds of std2 -dual (stda_rw_dual, stda-rw_dual.f:658-853;
 *     the 'dual' file holds thr2_ev and the atom list).  Occupied MOs whose
 *     squared Loewdin coefficients on the listed atoms (1-based) sum above 0.1
 *     keep Ethr and the PT2 window; the others use thr2_ev and the window
 *     2 (1 + 0.8 a_x) thr2.  CSFs of the chromophore MOs come first before the
 *     energy sort, as in the reference.  natoms = 0: one threshold.
 *   pt2_off != 0: no PT2 selection, the primaries only (-dual 1, -rw 1).
 * No counterpart in the reference's RKS driver `stda` itself. */
int stda_hip_set_selection(stda_hip_ctx *ctx, double thr2_ev, int32_t natoms,
                           const stda_hip_int *atoms, int32_t pt2_off);

/* CSF list after selection: iconf(nci,2) 1-based (occupied MO, absolute
 * virtual MO index no+1..moci) and ed(nci), the PT2-corrected diagonal
 * (same order as the reference: primaries sorted by energy, then the
 * PT2-selected CSFs in pair order). */
int stda_hip_get_csf(stda_hip_ctx *ctx, stda_hip_int *iconf, double *ed);

/*
 * TDA eigenproblem (replaces rtdamat + ssyevr, stda.f:1107-1150, and the
 * property loop stda.f:1207-1378 / apbtrafo.f:23-298).
 *   nroots > 0 : the nroots lowest roots (block Davidson, FP64);
 *   nroots <= 0: every root in (0, Ethr], as ssyevr with RANGE='V'.
 *   capacity   : columns available in x / rows in props (>= roots returned).
 *   velcorr    : 1 = Rv(corr) from (A + 0.5 B) (default sTDA), 0 = plain RV.
 *   tol        : Davidson residual-norm tolerance (<= 0: 1e-8).
 * Out: nfound; omega(nfound) (Hartree); x(nci, nfound) eigenvectors (may be
 * NULL); props(5, nfound) = fL, fV, RL, RV-or-Rv(corr), (unused 0);
 * alpha_sos(6) = SOS alpha(0) xx,xy,yy,xz,yz,zz (stda.f:1272-1277,1465).
 */
int stda_hip_solve_tda(stda_hip_ctx *ctx, int32_t nroots, int32_t capacity, int32_t velcorr,
                       double tol, int32_t *nfound, double *omega, double *x, double *props,
                       double *alpha_sos);

/*
 * sTD-DFT (RPA) eigenproblem (replaces rrpamat + srpapack, stda.f:800-1005,
 * srpapack.f:32-197): paired (A+B)/(A-B) Davidson without (A-B)^1/2.
 * Out: omega, x = X, y = Y (nci, nfound) as srpapack returns them
 * (X+Y = (A-B)^1/2 Z / sqrt(w), normalised so (X+Y)^T (X-Y) = 1); props as
 * the RPA property loop (lengths with X+Y, velocity/magnetic with X-Y).
 */
int stda_hip_solve_rpa(stda_hip_ctx *ctx, int32_t nroots, int32_t capacity, double tol,
                       int32_t *nfound, double *omega, double *x, double *y, double *props,
                       double *alpha_sos);

/*
 * Linear-response polarizability (replaces lresp, linear_response.f:23-192,
 * and the alpha part of lresp1): for each frequency w solve
 * ((A+B) - w^2 (A-B)^-1)(X+Y) = -2 mu for the 3 Cartesian right-hand sides
 * and return alpha(w)_rc = -2 mu_r^T (X+Y)_c as alpha(3,3,nfreq).
 * xpy (nci,3,nfreq) may be NULL.  tol: residual-norm tolerance of the
 * response equations (<= 0: 1e-10; near a pole alpha amplifies the residual).
 */
int stda_hip_lresp(stda_hip_ctx *ctx, int32_t nfreq, const double *omega, double tol,
                   double *alpha, double *xpy);

/*
 * SHG first hyperpolarizability beta(-2w; w, w) and its hyper-Rayleigh
 * scattering invariants (replaces the beta part of lresp1,
 * linear_response.f:373-699 with beta_resp_fast/A_beta_fast/B_beta_fast/
 * List_A/List_B :911-1226 and HRS :1489-1662; std2 -resp N).  For each
 * frequency w (Hartree; w = 0 is the static case) the linear responses at w
 * and 2w are solved as in stda_hip_lresp, X-Y follows from the response
 * equation, and beta_xyz = A - B over the reference's six orderings.
 * Out: beta(3,3,3,nfreq) (Fortran order, beta[i + 3j + 9k + 27f]) and, if
 * hrs is not NULL, hrs(4,nfreq) = beta_HRS, DR, <beta^2_ZZZ>, <beta^2_XZZ>
 * (the reference's beta_HRS file holds w and beta_HRS).  Needs moments.
 */
int stda_hip_beta_shg(stda_hip_ctx *ctx, int32_t nfreq, const double *omega, double tol,
                      double *beta, double *hrs);

/*
 * Two-photon absorption of RPA states, both photons at half the excitation
 * energy (replaces lresp_2PA_SP and TPA_resp_fast_SP with its A/B_2PA_*_fast_SP
 * sums, linear_response.f:4604-5153; std2 -2PA N).  Inputs are what srpapack
 * hands that routine: excitation energies eci (Hartree, > 0) and the states'
 * X and Y (nci x nstates, leading dimension ldx; stda.f:998-1042).  For each
 * state the response at w = -eci/2 with right-hand side -mu is solved, and
 * sigma(x,y) = (-A + B)/2 is formed from the CSF-pair sums.
 * Out: sigma(3,3,nstates); if delta is not NULL, delta(4,nstates) =
 * delta_2PA parallel, perpendicular, circular and rho = parallel/perpendicular
 * (the columns of the reference's 2PA-abs file).  In -2PA runs the
 * reference's srpapack forms X and Y with the diagonals of (A-B)^1/2 and
 * A+B counted twice (the halving at srpapack.f:98-105 is skipped in that
 * mode); passing its uci/hci reproduces its numbers, passing the X, Y of
 * stda_hip_solve_rpa gives the 2PA of the properly normalised states.
 */
int stda_hip_tpa(stda_hip_ctx *ctx, int32_t nstates, const double *eci, const double *xci,
                 const double *yci, int64_t ldx, double tol, double *sigma, double *delta);

/* Per-routine FP64 equivalents of the reference's matrix builders, for
 * parity checks and small systems (explicit matrices, O(nci^2) memory):
 *   rtdamat  (stda.f:2010-2056):     a(nci,nci) full;
 *   rrpamat  (stda.f:1870-2004):     apb, amb packed lin(i,j), the A-B
 *                                    diagonal un-rooted (resp mode);
 *   rtdacorr (apbtrafo.f:580-637):   bhalf packed, 0.5 B. */
int stda_hip_build_a(stda_hip_ctx *ctx, double *a);
int stda_hip_build_apb_amb(stda_hip_ctx *ctx, double *apb, double *amb);
int stda_hip_build_bhalf(stda_hip_ctx *ctx, double *bhalf);

/* A (or A+B, A-B, 0.5B) applied to nvec vectors without forming the matrix:
 * which = 0 A, 1 A+B, 2 A-B, 3 0.5B.  v, av: (nci, nvec). */
int stda_hip_matvec(stda_hip_ctx *ctx, int32_t which, int32_t nvec, const double *v, double *av);

/* Gaussian broadening of line spectra, the curve loop of the reference's
 * spectrum program g_spec (g_spec/g_spec.f:371-433; the program reads the
 * tda.dat written by stda.f:2235-2265 / apbtrafo.f:115-288 and writes
 * spec.dat).  Replaces that loop for nspec spectra at once; the caller does
 * g_spec's parsing, sorting and grid choice (g_spec.f:60-260).
 *   line_offsets[nspec+1]  lines of spectrum s are [line_offsets[s], line_offsets[s+1])
 *   x, y, wei              line energies (eV), strengths, Gaussian widths (eV),
 *                          in g_spec's summation order
 *   emin, step [nspec]     grid e_j = emin + j*step, j = 1..npts
 *   flags                  STDA_HIP_GSPEC_UV (absorption: f1 = 9.186e-3 and
 *                          f -> dipole strength; else CD, f1 = 22.97),
 *                          _NM (e -> 1e7/(8065.54093 e)), _CM (e -> 8065.54093 e,
 *                          after _NM when both are given), _LOG (r -> max(log10 r, 0))
 *   rfaktor                scale of the curve (RFAKTOR)
 *   e, r [nspec*npts]      grid abscissae and curve, spectrum-major
 * Lines with |x| <= 1e-6 are skipped as in the reference.  Arithmetic in
 * FP64 with the reference's operation order and its single-precision
 * constants; host arrays in and out. */
enum {
    STDA_HIP_GSPEC_UV = 1,
    STDA_HIP_GSPEC_NM = 2,
    STDA_HIP_GSPEC_CM = 4,
    STDA_HIP_GSPEC_LOG = 8
};
int stda_hip_gspec(stda_hip_ctx *ctx, int32_t nspec, const int64_t *line_offsets, const double *x,
                   const double *y, const double *wei, const double *emin, const double *step,
                   int32_t npts, int32_t flags, double rfaktor, double *e, double *r);

/* stda_hip_gspec without any device allocation.  No counterpart in the
 * reference; for callers that cannot allocate device memory (tight serving
 * loops, harnesses that forbid hipMalloc).  Same arguments, kernel and results
 * as stda_hip_gspec (host arrays in and out); the call stages its inputs,
 * parameters and outputs through a fixed device array of 32 MiB per GPU,
 * reserved when the library is loaded and held until the process ends, and
 * calls no allocating routine.  It needs, in doubles, each part rounded up to
 * a multiple of 32: nspec + 1, three times the line count, twice nspec and
 * twice nspec * npts; calls needing more than 32 MiB return
 * STDA_HIP_ERR_TOO_LARGE.  Every call on one GPU uses that GPU's array: calls
 * that overlap on the same GPU (from several threads or contexts) race. */
int stda_hip_gspec_noalloc(stda_hip_ctx *ctx, int32_t nspec, const int64_t *line_offsets,
                           const double *x, const double *y, const double *wei,
                           const double *emin, const double *step, int32_t npts, int32_t flags,
                           double rfaktor, double *e, double *r);

/* The lines of spec.dat (g_spec.f format 1000, (f16.6, e14.4)) for n grid
 * points: out receives 31 n bytes, each line 30 characters and a newline, no
 * terminating zero.  Host only (no device work, no context).  No counterpart in
 * the reference, whose program writes the file with Fortran I/O. */
int stda_hip_format_curve(int64_t n, const double *e, const double *r, char *out);

/* Arithmetic of the coupling operators in the solvers of this context:
 *   0 (default) FP64 throughout;
 *   1 mixed: the transition-density contractions (Q gamma Q^T products) in
 *     FP32 with FP32 accumulation, Davidson orthogonalisation, Rayleigh-Ritz
 *     and residual norms in FP64, converged to a residual of 1e-6, followed
 *     by one FP64 Rayleigh-Ritz refinement of the converged roots.  The
 *     response solver uses the FP32 operator with a 1e-6 tolerance.
 *   2 split-bfloat16: the same, with the operator GEMMs' FP32 operands split
 *     into two bfloat16 numbers each and multiplied as hi*hi + hi*lo + lo*hi
 *     on the bfloat16 matrix cores with FP32 accumulation, to a residual of
 *     3e-5; then the FP32 phase from its vectors and the FP64 refinement.
 * No counterpart in the reference (which is FP32 throughout, stda.f:82-107). */
int stda_hip_set_precision(stda_hip_ctx *ctx, int32_t precision);

/* The matrix-core GEMM behind every contraction of the operators, on host
 * arrays: C = alpha op(A) op(B) + beta C, column-major, op = transpose when
 * ta / tb is nonzero; op(A) is m x k, op(B) k x n.  mode 0: FP64 (the
 * drop-in's arithmetic); 1: FP32 operands and accumulation; 2: split-bfloat16
 * with FP32 accumulation (modes 1 and 2 round A, B, C to FP32 first and
 * return the FP32 result widened).  No counterpart in the reference (its
 * ssymm/sgemm calls, stda.f:558-589,742); exposes the kernel to tests and to
 * callers that want it. */
int stda_hip_gemm(stda_hip_ctx *ctx, int32_t mode, int32_t ta, int32_t tb, int32_t m, int32_t n,
                  int32_t k, double alpha, const double *A, int64_t lda, const double *B,
                  int64_t ldb, double beta, double *C, int64_t ldc);

/* Timing of the last setup/solve phases on this context, in milliseconds:
 * [0] gamma+charges+intermediates, [1] screening+sort, [2] PT2,
 * [3] eigensolve / response, [4] properties, [5] total. */
int stda_hip_timings(const stda_hip_ctx *ctx, double *ms6);

/* Per-kernel statistics accumulated since the last select: HIP-event time
 * (ms) on the context's stream, launch count and algorithmic work (flops).
 * Names: "pt2_rows", "pt2_cols", "ao_matvec", "gspec" (work: point-line
 * terms), "beta" (the GEMMs of one frequency), "tpa" (those of one state).
 * Unknown names give zeros. */
int stda_hip_kernel_stats(const stda_hip_ctx *ctx, const char *name, double *ms,
                          double *launches, double *work);

/* ------------------------------------------------------------------------
 * Batched solves of many independent molecules.  The reference solves one
 * molecule per process (the `stda` routine, stda.f:53-1471); a driver that
 * loops over many molecules replaces that loop with one call here.
 *
 * A batch object owns `nworkers` contexts, each with its own host thread and
 * HIP stream (give the process at least that many hardware queues, e.g.
 * GPU_MAX_HW_QUEUES=32, so that the streams overlap), and a driver stream.
 * stda_hip_batch_solve_tda runs the setup pipeline of every molecule (as
 * stda_hip_setup) on the workers concurrently.  Molecules with at most
 * nmax_batched CSFs (default 1024) and nroots <= 24 are solved in lock step
 * by the calling thread in mixed precision (as stda_hip_set_precision(ctx,
 * 1): FP32 operator products, FP64 subspace numerics, residual norms 1e-6,
 * one FP64 Rayleigh-Ritz refinement): per iteration one launch for the
 * operator products of all of them and one for a Davidson step of each,
 * molecules joining as soon as their setup is done; larger molecules take
 * the single-molecule mixed path on their worker.
 * Results per molecule i (arrays of nmol rows, row-major by molecule):
 *   nfound[i]                       roots found (min(nroots, nci))
 *   omega[i*nroots + j]             excitation energies (Hartree)
 *   props[(i*nroots + j)*5 + 0..4]  fL, fV, RL, RV or Rv(corr), 0
 *   status[i]                       STDA_HIP_OK or the molecule's error code
 * The call itself fails only for invalid arguments or device errors. */
typedef struct stda_hip_batch stda_hip_batch;

int stda_hip_batch_create(stda_hip_batch **out, int device, int nworkers);
int stda_hip_batch_destroy(stda_hip_batch *b);
const char *stda_hip_batch_last_error(const stda_hip_batch *b);
int stda_hip_batch_solve_tda(stda_hip_batch *b, int32_t nmol, const stda_hip_system *sys,
                             const stda_hip_params *par, int32_t nroots, double tol,
                             int32_t velcorr, int32_t *nfound, double *omega, double *props,
                             int32_t *status);
/* Molecules with at most nmax_batched CSFs go to the lock-step solver, whose
 * driver reads the state table back every `poll` iterations. */
int stda_hip_batch_configure(stda_hip_batch *b, int64_t nmax_batched, int32_t poll);
/* Larger molecules (more than nmax_batched and at most nmax_wide CSFs, at
 * most 24 roots and 1024 atoms) join the lock step too, with the matrix-free
 * operator (enable = 1, the default), or take the single-molecule mixed solve
 * on their worker (enable = 0; so do the molecules above nmax_wide;
 * nmax_wide <= 0: no limit).  enable = 2: as 1, with the lock-step iterations
 * of a call starting only once every molecule is prepared (setup and solve
 * then do not share the device).  No counterpart in the reference (its
 * driver solves one molecule per call: stda.f:53, main.f:928). */
int stda_hip_batch_configure_wide(stda_hip_batch *b, int32_t enable, int64_t nmax_wide);
/* Arithmetic of the operator products of the molecules solved on the
 * single-molecule path of the workers: 1 FP32 (the default), 2 split-bfloat16
 * (a split-bfloat16 phase to 3e-5, then FP32 to 1e-6, then the FP64 step, as
 * stda_hip_set_precision 2).  The lock steps keep FP32 products; PT2 pass 1
 * runs in FP32 with the FP64 recheck in both.  No counterpart in the
 * reference. */
int stda_hip_batch_set_precision(stda_hip_batch *b, int32_t precision);
/* molecules of the last call solved in the matrix-free lock step (a subset of
 * counts3[0] of stda_hip_batch_timings) */
int stda_hip_batch_wide_count(const stda_hip_batch *b, int64_t *nwide);
/* ms5: host time of [0] the setup and solve phase (wall), [1] the driver
 * loop (wall), [2] property tables, [3] total, and [4] the summed HIP-event
 * times of the batched launches; counts3 (disjoint, summing to nmol):
 * molecules solved with status OK by the batched solver, with status OK by
 * the single path, and returned with a nonzero status (no CSF in the energy
 * window counts here, as does a solve that did not converge on either path). */
int stda_hip_batch_timings(const stda_hip_batch *b, double *ms5, int64_t *counts3);
/* Per-molecule phase times of the batched step kernel summed over the
 * molecules of the last call (microseconds): Gram products, subspace
 * eigensolve, Ritz vectors, expansion, refinement + moments, [5] 0; [6]
 * Jacobi sweeps; then the workers' host time summed over the workers (us):
 * [7] setup pipeline, [8] explicit matrices, [9] single-path solves. */
int stda_hip_batch_phase_times(const stda_hip_batch *b, double *us10);
/* stda_hip_kernel_stats summed over the workers and the molecules of the last
 * call (names as there, plus "batch_lockstep": the launches of the lock-step
 * graph, each `poll` iterations of the operator products and Davidson steps of
 * every active molecule; work = algorithmic bytes of the operator products). */
int stda_hip_batch_kernel_stats(stda_hip_batch *b, const char *name, double *ms, double *launches,
                                double *work);
/* Continuous mode (serving): molecules are submitted one at a time while
 * earlier ones are being solved.  The object's worker contexts take them in
 * submission order (setup pipeline; explicit matrices, or the whole
 * single-molecule mixed solve for larger ones) while a driver thread keeps the
 * lock-step iterations running: a prepared molecule joins the running set at
 * the driver's next pass and leaves it with its result at the first poll that
 * finds it done.  Results equal those of stda_hip_batch_solve_tda for the same
 * molecule (each molecule's kernels touch only its data and reduce in a fixed
 * order).  The arrays that sys points to must stay valid until the molecule's
 * result has been taken.  Up to 2048 molecules are in the lock-step set at
 * once (more wait for a free row).  While the continuous mode runs,
 * stda_hip_batch_solve_tda returns STDA_HIP_ERR_STATE.
 *   submit: queue one molecule; ticket receives its handle.
 *   result: done = 1 and the outputs (nfound, omega[nroots], props[5 nroots],
 *           status, as in stda_hip_batch_solve_tda) once the molecule is
 *           finished, then the ticket is released; done = 0 otherwise.  wait
 *           != 0 blocks until it is finished.
 *   stop:   finish every submitted molecule and end the mode's threads. */
int stda_hip_batch_submit(stda_hip_batch *b, const stda_hip_system *sys,
                          const stda_hip_params *par, int32_t nroots, double tol,
                          int32_t velcorr, int64_t *ticket);
int stda_hip_batch_result(stda_hip_batch *b, int64_t ticket, int32_t wait, int32_t *done,
                          int32_t *nfound, double *omega, double *props, int32_t *status);
int stda_hip_batch_stop(stda_hip_batch *b);
/* Algorithmic HBM bytes of the last batched Davidson launch: per molecule,
 * the FP32 A (4 n^2) once per iteration plus the FP64 A and 0.5 B (8 n^2
 * each) once. */
int stda_hip_batch_kernel_bytes(const stda_hip_batch *b, double *bytes);

#ifdef __cplusplus
}
#endif
#endif /* STDA_HIP_H */
