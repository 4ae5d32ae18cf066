// Shared declarations of the MI355X excited-state core (host and device).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

namespace stda {

struct Error : std::runtime_error {
    int status;
    Error(int s, const std::string &m) : std::runtime_error(m), status(s) {}
};

#define STDA_HIP_CHECK(expr)                                                              \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess)                                                             \
            throw ::stda::Error(e_ == hipErrorOutOfMemory ? 4 : 3,                        \
                                std::string(#expr) + ": " + hipGetErrorString(e_));     \
    } while (0)

// Sizes of the 4-index coupling tiles (see kernels.hip: coupling_tile_kernel).
// A row tile is TK occupied x TC virtual indices (pairs (k,c)); a column tile
// is TJ occupied x TB virtual indices (pairs (j,b)).
constexpr int TK = 4, TC = 8, TJ = 4, TB = 8;
constexpr int NA_CHUNK = 32;

// --- matrix-core GEMM (gemm.hip) -----------------------------------------------------
// C_b = alpha op(A_b) op(B_b) + beta C_b (column-major, batch strides sA, sB, sC).
enum GemmMode { GEMM_F64 = 0, GEMM_F32 = 1, GEMM_BF16X3 = 2 };
void launch_gemm_f64(hipStream_t s, bool ta, bool tb, int M, int N, int K, double alpha,
                     const double *A, int64_t lda, int64_t sA, const double *B, int64_t ldb,
                     int64_t sB, double beta, double *C, int64_t ldc, int64_t sC, int batch);
// mode GEMM_F32 (FP32 MFMA) or GEMM_BF16X3 (split-bfloat16 MFMA, FP32 accumulation)
void launch_gemm_f32(hipStream_t s, int mode, bool ta, bool tb, int M, int N, int K, float alpha,
                     const float *A, int64_t lda, int64_t sA, const float *B, int64_t ldb,
                     int64_t sB, float beta, float *C, int64_t ldc, int64_t sC, int batch);

// --- kernel launchers (kernels.hip) -------------------------------------------------
void launch_gamma(hipStream_t s, int ncent, const double *xyz4, const double *eta, double ax,
                  double alphak, double betaj, double *gamj, double *gamk);

// charge modes for launch_charges
enum ChargeMode { CH_IA = 0, CH_OCC_PACKED = 1, CH_VIR_PACKED = 2, CH_DIAG = 3 };
// ct: (moci, nao) row-major Loewdin MOs with AOs sorted by atom; atom_ptr(ncent+1).
void launch_charges(hipStream_t s, ChargeMode mode, int ncent, int nao, int no, int nv,
                    const double *ct, const int *atom_ptr, double *q);

void launch_screen(hipStream_t s, int ncent, int no, int nv, const double *eps,
                   const double *jdiag, const double *pia, const double *qia, double ak,
                   double thr, double fthr, int kshift, double smax, double width,
                   double steep, double *de, double *kdiag, unsigned char *flags);

// flags bit 0: primary, bit 1: secondary.  Order-preserving compaction.
void launch_count_flags(hipStream_t s, const unsigned char *flags, int64_t n, int bit,
                        int *block_counts, int nblocks);
void launch_compact_flags(hipStream_t s, const unsigned char *flags, int64_t n, int bit,
                          const int64_t *block_offsets, int nblocks, int32_t *out);
int compact_block_size();

void launch_gather_f64(hipStream_t s, const double *src, const int32_t *idx, int64_t n,
                       double *dst);
void launch_scatter_pairs(hipStream_t s, const double *v, int64_t ldv, const int32_t *pair,
                          int64_t n, int nvec, double *full, int64_t ldfull);
void launch_gather_pairs(hipStream_t s, const double *full, int64_t ldfull, const int32_t *pair,
                         int64_t n, int nvec, const double *diag, const double *v, int64_t ldv,
                         double *out, int64_t ldout);
void launch_csf_moments(hipStream_t s, const double *moments, int64_t npk, const int32_t *pair,
                        int64_t n, int no, int nv, double *mom);

// AO-basis coupling: t_A = sum_{mu in A} D(mu,mu), u = gamk t; per vector.
void launch_atom_diag(hipStream_t s, int nvec, int nao, int ncent, const double *Dt,
                      int64_t strideD, const int *atom_of, const double *gamk, double *u);
// Gt = aJ (gamJ o Dt) + aX (gamK o Dt^T) + diag(aK u_A(mu)), per vector.
void launch_gbuild(hipStream_t s, int nvec, int nao, int ncent, const double *Dt,
                   int64_t strideD, const int *atom_of, const double *gamj, const double *gamk,
                   const double *u, double aJ, double aX, double aK, double *Gt, int64_t strideG);

// Davidson helpers
void launch_residual(hipStream_t s, int64_t n, int k, const double *AX, const double *X,
                     const double *theta, double *R);
void launch_col_norms(hipStream_t s, int64_t n, int k, const double *R, double *norms);
void launch_precondition(hipStream_t s, int64_t n, int k, const double *R, const double *diag,
                         const double *theta, double *T);
void launch_precondition_gather(hipStream_t s, int64_t n, int k, const double *R, const double *col,
                                const double *diag, const double *theta, double *T);
void launch_scale_cols(hipStream_t s, int64_t n, int k, double *V, const double *scale);
// V(:, c) *= 1 / norm[c] (0 where norm[c] == 0)
void launch_inv_scale_cols(hipStream_t s, int64_t n, int k, double *V, const double *norm);
void launch_set_unit(hipStream_t s, int64_t n, int k, double *V, const int32_t *rows);
// Y(:, c) = diag o X(:, c)
void launch_axpby_diag(hipStream_t s, int64_t n, int k, const double *diag, const double *X,
                       int64_t ldx, double *Y, int64_t ldy);

// response preconditioner (2x2 diagonal blocks) and slice reduction
void launch_precond_pair(hipStream_t s, int64_t n, int ncol, double *R1, double *R2,
                         const double *d, const double *w);
void launch_sum_slices(hipStream_t s, int64_t len, int nchunk, const double *W, double *C);

// 4-index coupling tiles: amat(kc, jb) = ak (kc|jb)_K - (kj|cb)_J.
struct TileSet {
    const int32_t *occ;   // ntiles x T_occ occupied indices (-1 = none)
    const int32_t *vir;   // ntiles x T_vir virtual indices (-1 = none)
    int ntiles;
};
// PT2 on the matrix cores (FP64 MFMA) with 8 x 8 x 8 x 8 tiles: rows/cols
// tile lists hold PT2_T occupied and PT2_T virtual indices each.
constexpr int PT2_T = 8;
constexpr int PT2_SLOTS = PT2_T * PT2_T;
// Pass 1: for every secondary of the row tiles its sum of pt over all
// primaries.  drow/dcol (PT2_SLOTS per tile) hold the de of each tile slot,
// NaN where the slot is not a secondary (rows) or primary (cols).  The column
// tiles are split over nsplit work-groups per row tile; the operand columns
// are lda >= ncent apart, lda even, zero-padded (launch_pad_rows); rowpart receives
// nsplit x rows.ntiles x PT2_SLOTS partial sums.  With ptm non-null every pt
// term is also stored at ptm[(PT2_SLOTS rt + row slot) * ldp + PT2_SLOTS ct +
// col slot] (0 for slots that are not a secondary-primary pair), ldp =
// PT2_SLOTS * cols.ntiles.
void launch_pt2_pipe(hipStream_t s, TileSet rows, TileSet cols, int ncent, int nv,
                     const double *pK, const double *q, const double *pJ, const double *qab,
                     int64_t lda, const double *drow, const double *dcol, double ak, int nsplit,
                     double *rowpart, double *ptm, int64_t ldp);
// y (ld x cols) = x (rows x cols, ld rows) with the rows rows..ld zero
void launch_pad_rows(hipStream_t s, const double *x, int rows, int64_t cols, int ld, double *y);
// The same with FP32 operands and MFMA (pt and its sums in FP64); the
// operand columns are lda >= ncent apart, lda a multiple of 4, zero-padded
// (launch_to_f32_padded).
void launch_pt2_pipe_f32(hipStream_t s, TileSet rows, TileSet cols, int ncent, int nv,
                         const float *pK, const float *q, const float *pJ, const float *qab,
                         int64_t lda, const double *drow, const double *dcol, double ak,
                         int nsplit, double *rowpart, double *ptm, int64_t ldp);
// y (ld x cols, FP32) = x (rows x cols, ld rows, FP64), rows rows..ld zero
void launch_to_f32_padded(hipStream_t s, const double *x, int rows, int64_t cols, int ld,
                          float *y);
// rowsum[pair] of every secondary slot: its partial sums added in split order.
void launch_pt2_rows_finish(hipStream_t s, TileSet rows, int nv, int nsplit, const double *drow,
                            const double *rowpart, double *rowsum);
// Pass 2 from the stored terms: colsum[c] = sum over the rows rrow[0..nrej)
// (ascending) of ptm[r * ldp + c] for every column c < ldp, the rows split
// over nsplit work-groups per column block and the splits summed in order.
void launch_pt2_colsum(hipStream_t s, int nrej, const int32_t *rrow, int64_t ldp,
                       const double *ptm, int nsplit, double *part, double *colsum);
// PT2 pass 2 over a sparse row set: row tile = one occupied k (rk) with 16
// virtual c (rc, -1 padded); column tile = 16 occupied j (cj) x 16 virtual b (cb).
void launch_pt2_cols_sparse(hipStream_t s, int nrt, const int32_t *rk, const int32_t *rc, int nct,
                            const int32_t *cj, const int32_t *cb, int ncent, int nv,
                            const double *pK, const double *q, const double *pJ, const double *qab,
                            const double *de, const unsigned char *rowmask, unsigned char rowbit,
                            const unsigned char *colmask, unsigned char colbit, double ak,
                            int nsplit, double *part, double *colsum);
// Cholesky QR factor of a block Gram matrix S (nt <= 64), dependent columns
// dropped: LinvT = L^-T on the kept columns, alive[j] = 1.0 if column j is kept
// (pivot > thr * ref[j * ref_stride], or > thr when ref is null).  S =
// G[m:m+nt, :] - G[0:m, :]^T G[0:m, :] with G ((m + nt) x nt, ld ldg).
// skip (device, optional; needs ref = squared norms before the projection):
// 1.0 when T LinvT needs no further projection against V, else 0.0.
void launch_chol_inv(hipStream_t s, int nt, const double *G, int m, int ldg, const double *ref,
                     int ref_stride, double thr, double *LinvT, double *alive,
                     double *skip = nullptr);
// column 2-norms of an (n x k) block through nsl row slices (part: nsl * k)
void launch_col_norms_split(hipStream_t s, int64_t n, int k, const double *R, double *part,
                            int nsl, double *norms);
// R = AX - X diag(theta) and the column norms of R (same slicing)
void launch_residual_norms(hipStream_t s, int64_t n, int k, const double *AX, const double *X,
                           const double *theta, double *R, double *part, int nsl, double *norms);
// T(:, q) = R(:, c) / (theta_c - diag) for the q-th column c with norms[c] >= tol, q < room
// quadratic response (hyper.hip): output column c of the dense X/Y stacks reads column
// src[c] of the (X+Y) and (A+B)(X+Y) arrays; inv_w[c] = 1/w of its frequency (0 for w = 0)
struct BetaCols {
    int64_t src[6];
    double inv_w[6];
    double scale;   // applied to X and Y (right-hand sides -mu instead of -2 mu: 0.5)
};
void launch_beta_xy(hipStream_t s, int64_t n, const int32_t *pair, int no, int nv,
                    const double *XP, const double *AX, const double *rhs, const BetaCols &cols,
                    int ncols, double *Xd, double *Yd);
// columns of V (n x k, leading dimension ldv) into dense occupied x virtual matrices
void launch_scatter_pairs(hipStream_t s, int64_t n, const int32_t *pair, int no, int nv, int k,
                          const double *V, int64_t ldv, double *D);
void launch_mu_blocks(hipStream_t s, const double *moments, int64_t npk, int no, int nv,
                      double *muoo, double *muvv);
// upstream transforms (transform.hip).  AO_LMAX: the largest Cartesian exponent per
// direction the integral kernel's recurrence tables are sized for (checked on the host).
constexpr int AO_LMAX = 4;
void launch_pack_upper(hipStream_t s, int64_t n, const double *P, double *out);
void launch_scale_rows(hipStream_t s, int64_t m, int64_t k, const double *f, double *T);
void launch_ao_ints(hipStream_t s, int nao, const int64_t *ptr, const double *alpha,
                    const double *coef, const double *cen, const int32_t *lmn, double *S,
                    double *ao9);
// spectrum broadening (spectrum.hip); flags as STDA_HIP_GSPEC_* of the public header
enum { GSPEC_UV = 1, GSPEC_NM = 2, GSPEC_CM = 4, GSPEC_LOG = 8 };
void launch_gspec(hipStream_t s, int nspec, const int64_t *off, const double *x, const double *y,
                  const double *w, const double *emin, const double *step, int npts, int flags,
                  double rfaktor, double *eout, double *rout);
void launch_precondition_compact(hipStream_t s, int64_t n, int k, const double *R,
                                 const double *diag, const double *theta, const double *norms,
                                 double tol, int room, double *T);
// FP32 operator kernels (mixed precision)
void launch_scatter_pairs_f32(hipStream_t s, const double *v, int64_t ldv, const int32_t *pair,
                              int64_t n, int nvec, float *full, int64_t ldfull);
void launch_gather_pairs_f32(hipStream_t s, const float *full, int64_t ldfull, const int32_t *pair,
                             int64_t n, int nvec, const double *diag, const double *v, int64_t ldv,
                             double *out, int64_t ldout);
void launch_atom_diag_f32(hipStream_t s, int nvec, int nao, int ncent, const float *Dt,
                          int64_t strideD, const int *atom_ptr, const float *gamk, float *u);
void launch_gbuild_f32(hipStream_t s, int nvec, int nao, int ncent, const float *Dt, int64_t strideD,
                       const int *atom_of, const float *gamj, const float *gamk, const float *u,
                       float aJ, float aX, float aK, float *Gt, int64_t strideG);
void launch_to_f32(hipStream_t s, const double *x, int64_t n, float *y);
// dst(:, I) = src(:, idx[I]) for I < n, columns of `rows` doubles
void launch_gather_cols(hipStream_t s, int rows, int64_t n, const int32_t *idx, const double *src,
                        double *dst);
// A(I, I) = d[I] for the n x n column-major A
void launch_set_diag(hipStream_t s, int64_t n, const double *d, double *A);

// Tall-skinny FP64 products (kernels.hip).  C (m x p, ld m) = A^T B with
// A (n x m), B (n x p), m <= TS_MMAX, p <= TS_PMAX, via per-slice partial
// sums in part (tsgemm_tn_slices(n, m, p) * m * p doubles) reduced in slice order.
constexpr int TS_MMAX = 256, TS_PMAX = 64;
int tsgemm_tn_slices(int64_t n, int m, int p);
void launch_tsgemm_tn(hipStream_t s, int64_t n, int m, int p, const double *A, int64_t lda,
                      const double *B, int64_t ldb, double *part, double *C);
// the two launches of launch_tsgemm_tn: slice partials, then their ordered sum
void launch_tsgemm_tn_partial(hipStream_t s, int64_t n, int m, int p, const double *A,
                              int64_t lda, const double *B, int64_t ldb, double *part);
// slice partials of [A1 A2]^T B (m1 + m2 columns, same lda); with skip1
// (device) non-null and *skip1 != 0 the A1 rows come out zero, A1 unread
void launch_tsgemm_tn2_partial(hipStream_t s, int64_t n, int m1, const double *A1, int m2,
                               const double *A2, int p, int64_t lda, const double *B, int64_t ldb,
                               double *part, const double *skip1 = nullptr);
void launch_tsgemm_tn_reduce(hipStream_t s, int64_t n, int m, int p, const double *part,
                             double *C);
// out (n x p) = alpha A C + beta out with A (n x m), C (m x p); any m, p
// (32 columns per launch); with skip (device) non-null and *skip != 0 the
// launches return at once and out is left as it is
void launch_tsgemm_nn(hipStream_t s, int64_t n, int m, int p, double alpha, const double *A,
                      int64_t lda, const double *C, int64_t ldc, double beta, double *out,
                      int64_t ldo, const double *skip = nullptr);
// two such products with the same C in one launch: out0 = alpha A0 C + beta
// out0, out1 = alpha A1 C + beta out1 (A1 null: the first only)
void launch_tsgemm_nn2(hipStream_t s, int64_t n, int m, int p, double alpha, const double *A0,
                       const double *A1, int64_t lda, const double *C, int64_t ldc, double beta,
                       double *out0, double *out1, int64_t ldo, const double *skip = nullptr);

// --- batched Davidson (batch.hip): molecules in lock step ------------------------------
constexpr int BATCH_MD = 64;                        // subspace limit
constexpr int BATCH_KMAX = 24;                      // roots per molecule
constexpr int BATCH_OUT = 20 * BATCH_KMAX + 12;     // doubles of output per molecule
constexpr int BATCH_STATE = 64;                     // ints of state per molecule
// Output of one molecule: omega[KMAX] | T[9 KMAX] = mom^T X | T2[9 KMAX] = mom^T Xnew
// | FP64 residual norms[KMAX] | info[12]: iterations, converged, roots found, last
// FP32 residual, wall-clock ticks of the step phases (Gram, subspace eigensolve,
// Ritz vectors, expansion, refinement + moments), unused, Jacobi sweeps, unused.
// State: phase, m, mnew, iterations, roots kept, unused x3.
enum BatchPhase { B_INIT = 0, B_ITER = 1, B_REFINE = 2, B_VEL = 3, B_DONE = 5 };
struct BatchMol {
    int64_t n;                 // CSFs
    int k, nguess, maxdim, maxiter;
    int velcorr, pad;
    double tol;                // residual tolerance of the FP32-operator iterations
    const float *A32;          // n x n FP32 copy of A (iterations)
    const double *A64;         // n x n FP64 A (final Rayleigh-Ritz)
    const double *B64;         // n x n 0.5 B (velocity correction) or null
    const double *diag;        // n, diagonal of A (preconditioner)
    const double *mom;         // n x 9 CSF moments or null
    const int32_t *guess;      // nguess start rows
    double *V, *AV;            // n x maxdim
    double *X, *AX, *W;        // n x BATCH_KMAX
    double *H;                 // BATCH_MD x BATCH_MD
    double *out;               // BATCH_OUT
    int *state;                // BATCH_STATE
};
// one step of every molecule (one work-group each)
void launch_batch_step(hipStream_t s, int nmol, const BatchMol *mols);

// --- grouped GEMM (gemm.hip) -----------------------------------------------------------
// One problem of a grouped launch: C_b = alpha op(A_b) op(B_b) + beta C_b with
// operand b at A + b sA (and B, C likewise).  The count follows `kind`:
// GC_FIXED uses `count`; GC_OP32 / GC_OP64 read the vector count of the FP32 /
// FP64 operator product of a lock-step molecule from its state (wide_op_count).
// cols_per_count > 0: one entry with N = cols_per_count * count columns;
// otherwise count batch entries of N columns, bpw of them in turn per
// work-group (gb groups of them in the grid).
enum GemmCount { GC_FIXED = 0, GC_OP32 = 1, GC_OP64 = 2 };
struct GemmProb {
    int M, N, K;
    int cols_per_count;
    int kind, count;
    int gm, gn, gb;            // 64 x 64 tile grid of the launch: rows, columns, batch groups
    int bpw;                   // batch entries per work-group (in turn), gb = ceil(max count / bpw)
    int epi;                   // 1: C(r, c) = aJ gam(A(r), A(c)) (alpha AB)(r, c) (+ aK u(A(c)), r = c)
    int pad2;
    const int *state;
    const void *A, *B;
    void *C;
    int64_t lda, ldb, ldc, sA, sB, sC;
    double alpha, beta;
    // epilogue: gam (na x na), atom of each row/column, u (na per batch entry)
    const void *ep_gam;
    const int *ep_atom;
    const void *ep_u;
    int ep_na, ep_pad;
    double ep_aJ, ep_aK;
};
// Work-group b of a grouped launch belongs to item p with start[p] <= b <
// start[p + 1] (start: nitems + 1 ascending offsets, start[0] = 0).
__device__ inline int grouped_item(const int *start, int nitems, int b) {
    int lo = 0, hi = nitems - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (start[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}
// start[p] .. start[p + 1]: the gm * gn * gb tiles of problem p (tile row fastest)
void launch_gemm_grouped(hipStream_t s, int mode, bool ta, bool tb, int nprob,
                         const GemmProb *probs, const int *start, int ntiles);

// --- wide lock-step molecules (wide.hip): matrix-free operator, grouped kernels ---------
constexpr int WIDE_RC = 1024;                        // rows of a chunk tile
constexpr int WIDE_MD = 96;                          // subspace limit (LDS Jacobi)
constexpr int WIDE_PART = WIDE_MD * 40;              // doubles of partial sums per chunk
constexpr int WIDE_AUX = 16384;                      // doubles of coefficients per molecule
constexpr int WIDE_GB = 32;                          // AO columns per gbuild work-group
constexpr int WIDE_MAXATOMS = 1024;                  // atoms of a wide molecule (LDS bound)
// State beyond the shared fields (phase, m, mnew, iterations, roots kept):
enum WideState {
    WS_NEXT = 5, WS_RESTART = 6, WS_ROOM = 7, WS_ACC1 = 8, WS_ACC2 = 9, WS_FINISH = 10,
    WS_NUN = 11, WS_MP = 12, WS_UN = 16, WS_KEEP = 40
};
// Coefficients in WideMol::aux
enum WideAux {
    WA_Y = 0,                                      // WIDE_MD x KMAX Ritz coefficients
    WA_TH = WA_Y + WIDE_MD * BATCH_KMAX,           // KMAX Ritz values (ascending)
    WA_C = WA_TH + BATCH_KMAX,                     // WIDE_MD x KMAX projection coefficients
    WA_L = WA_C + WIDE_MD * BATCH_KMAX,            // KMAX x KMAX Cholesky factor
    WA_NR0 = WA_L + BATCH_KMAX * BATCH_KMAX,       // KMAX squared norms of the new directions
    WA_YP = 6144,                                  // WIDE_MD x WIDE_MD eigenvectors of the last subspace
    WA_END = WA_YP + WIDE_MD * WIDE_MD
};
static_assert(WA_NR0 + BATCH_KMAX <= WA_YP, "aux layout");
static_assert(WA_END <= WIDE_AUX, "aux area");
struct WideMol {
    int64_t n, npair;          // CSFs; occupied x virtual pairs of the window
    int nao, no, nv, na;
    int k, nguess, maxdim, maxiter;
    int nvb;                   // vectors of the FP32 operator buffers (>= nguess, k)
    int velcorr, nchunk, pad;
    double tol;                // residual tolerance of the FP32-operator iterations
    double ak, ax;
    const float *C32;          // nao x (no + nv) Loewdin MOs, occupied then virtual
    const double *C64;
    const float *gj32, *gk32;  // na x na gamma^J, gamma^K
    const double *gj64, *gk64;
    const int32_t *atom_of;    // nao
    const int32_t *atom_ptr;   // na + 1
    const int32_t *pair;       // n: CSF -> pair index (occupied * nv + virtual)
    const int32_t *guess;      // nguess start rows
    const double *corr;        // n: diagonal term of A added to the coupling part
    const double *diag;        // n: diagonal of A (preconditioner)
    const double *mom;         // n x 9 CSF moments or null
    // FP32 operator work (nvb vectors): transition densities in the pair grid,
    // nao x no half transforms, nao x nao densities (G in place), pair-grid results
    float *xf32, *t32, *d32, *r32, *u32;
    // FP64 operator work (k vectors): the same, G separate (exchange term)
    double *xf64, *t64, *d64, *g64, *r64, *u64;
    double *V, *AV;            // n x maxdim (maxdim <= WIDE_MD)
    double *X, *AX, *W;        // n x BATCH_KMAX
    double *H;                 // WIDE_MD x WIDE_MD
    double *aux;               // WIDE_AUX
    double *part;              // nchunk x WIDE_PART
    double *out;               // BATCH_OUT
    int *state;                // BATCH_STATE
};
// vector count of a wide molecule's FP32 (B_ITER: the new block) or FP64
// (B_REFINE, B_VEL: the kept roots) operator product this iteration
__host__ __device__ inline int wide_op_count(const int *st, bool fp64) {
    const int ph = st[0];
    if (!fp64) return ph == B_ITER ? st[2] : 0;
    return (ph == B_REFINE || ph == B_VEL) ? st[4] : 0;
}
__host__ __device__ inline int gemm_prob_count(const GemmProb &P) {
    if (P.kind == GC_OP32) return wide_op_count(P.state, false);
    if (P.kind == GC_OP64) return wide_op_count(P.state, true);
    return P.count;
}
// Work-group offsets of the grouped wide kernels, nmol + 1 entries each
// (grouped_item): row chunks (nchunk per molecule), vectors (nvb), and
// G-build tiles (nvb * ceil(nao / WIDE_GB): vector, then AO column block).
struct WideStarts {
    const int *chunk, *vec, *gb;
    int nchunks, nvecs, ngb;
};
// operator product launches of one lock-step iteration (scatter, atom sums, G build, gather)
void launch_wide_scatter(hipStream_t s, int nmol, const WideMol *mols, WideStarts st);
void launch_wide_atomdiag(hipStream_t s, int nmol, const WideMol *mols, WideStarts st, bool fp64);
void launch_wide_gbuild(hipStream_t s, int nmol, const WideMol *mols, WideStarts st);
void launch_wide_gather(hipStream_t s, int nmol, const WideMol *mols, WideStarts st);
// the Davidson step of every wide molecule (a fixed sequence of grouped launches)
void launch_wide_step(hipStream_t s, int nmol, const WideMol *mols, WideStarts st);
// operator products of every active molecule; tiles = (molecule, BATCH_MV_ROWS-row block)
constexpr int BATCH_MV_ROWS = 64;
void launch_batch_matvec(hipStream_t s, int ntiles, const BatchMol *mols, const int2 *tiles);
// Explicit matrix: A(row_csf, col_csf) = ak (ia|jb) - (ij|ab) for every CSF pair
// (diagonal written later); csf_of_pair maps pair -> CSF index or -1.
void launch_build_a(hipStream_t s, TileSet tiles, int ncent, int no, int nv, const double *pK,
                    const double *q, const double *pJ, const double *qab,
                    const int32_t *csf_of_pair, double ak, double *A, int64_t lda);
// Exchange-type term (ib|ja)_K for every CSF pair (explicit A+-B, 0.5B; small systems).
void launch_exchange(hipStream_t s, int64_t nci, int ncent, int nv, const int32_t *pair,
                     const double *pK, const double *q, double *E, int64_t lde);

}  // namespace stda
