// Batched sTDA eigensolver for many molecules in lock step: per iteration one
// launch computes the operator products of every active molecule (grid over
// the molecules' row blocks) and one launch advances every molecule's
// Davidson by a step (one work-group per molecule: Gram products, subspace
// eigensolve, Ritz vectors, residuals, preconditioning, orthonormalisation,
// and at the end the FP64 refinement and transition moments).  Mixed precision as in the single-molecule mixed path: the
// operator products use the FP32 copy of A with FP32 accumulation; the
// subspace, Gram products, Rayleigh-Ritz, residual norms and
// orthonormalisation are FP64, and the converged block gets one FP64
// Rayleigh-Ritz step with the FP64 A (stda.f:1134-1150 is the eigenproblem
// solved; the properties follow stda.f:1207-1378 and apbtrafo.f:58-69).
//
// The molecule's matrices and workspace are global memory; only the subspace
// eigenproblem and the staged vector rows live in LDS.  Every reduction has a
// fixed order, so a molecule's result does not depend on the rest of the batch.
#include "common.h"
#include "lockstep.h"

#include <type_traits>

#include <hip/hip_runtime.h>

namespace stda {

namespace {

// D[i + ldd c] = sum_r P[r + ldp i] Q[r + ldq c] for i < ni, c < nc (FP64).
template <typename OutT>
__device__ void dots(BatchSmem &sm, int64_t n, const double *P, int64_t ldp, int ni,
                     const double *Q, int64_t ldq, int nc, OutT *D, int64_t ldd) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int i0 = 0; i0 < ni; i0 += DI)
        for (int c0 = 0; c0 < nc; c0 += DC) {
            const int ti = min(DI, ni - i0), tc = min(DC, nc - c0);
            double acc[DI * DC];
#pragma unroll
            for (int e = 0; e < DI * DC; ++e) acc[e] = 0.0;
            for (int64_t r = threadIdx.x; r < n; r += BT) {
                double p[DI], q[DC];
#pragma unroll
                for (int ii = 0; ii < DI; ++ii) p[ii] = ii < ti ? P[r + ldp * (i0 + ii)] : 0.0;
#pragma unroll
                for (int cc = 0; cc < DC; ++cc) q[cc] = cc < tc ? Q[r + ldq * (c0 + cc)] : 0.0;
#pragma unroll
                for (int ii = 0; ii < DI; ++ii)
#pragma unroll
                    for (int cc = 0; cc < DC; ++cc) acc[ii * DC + cc] = fma(p[ii], q[cc], acc[ii * DC + cc]);
            }
            const double t = wave_reduce<DI * DC>(acc);
            if (lane < DI * DC) sm.red[w * 64 + bitrev(lane, ilog2(DI * DC))] = t;
            __syncthreads();
            if ((int)threadIdx.x < DI * DC) {
                const int ii = threadIdx.x / DC, cc = threadIdx.x % DC;
                if (ii < ti && cc < tc) {
                    double s = 0.0;
                    for (int ww = 0; ww < NWV; ++ww) s += sm.red[ww * 64 + threadIdx.x];
                    D[(i0 + ii) + ldd * (c0 + cc)] = (OutT)s;
                }
            }
            __syncthreads();
        }
}


// Ritz vectors of the kk lowest subspace roots: X = V Y, AX = AV Y,
// W = AX - theta X; column norms of W -> sm.nrm.  V, AV are n x m.
__device__ void ritz(BatchSmem &sm, int64_t n, int m, int kk, const double *V, const double *AV,
                     double *X, double *AX, double *W) {
    for (int c0 = 0; c0 < kk; c0 += 8) {
        const int tc = min(8, kk - c0);
        int pc[8];
        double th[8];
#pragma unroll
        for (int cc = 0; cc < 8; ++cc) {
            const int c = min(c0 + cc, kk - 1);
            pc[cc] = sm.perm[c];
            th[cc] = sm.ths[c];
        }
        double nacc[8];
#pragma unroll
        for (int cc = 0; cc < 8; ++cc) nacc[cc] = 0.0;
        for (int64_t r = threadIdx.x; r < n; r += BT) {
            double x[8], ax[8];
#pragma unroll
            for (int cc = 0; cc < 8; ++cc) x[cc] = ax[cc] = 0.0;
            for (int j = 0; j < m; ++j) {
                const double v = V[r + n * j], av = AV[r + n * j];
#pragma unroll
                for (int cc = 0; cc < 8; ++cc) {
                    const double yv = sm.y[j + LDU * pc[cc]];
                    x[cc] = fma(v, yv, x[cc]);
                    ax[cc] = fma(av, yv, ax[cc]);
                }
            }
#pragma unroll
            for (int cc = 0; cc < 8; ++cc)
                if (cc < tc) {
                    const double wv = ax[cc] - th[cc] * x[cc];
                    X[r + n * (c0 + cc)] = x[cc];
                    AX[r + n * (c0 + cc)] = ax[cc];
                    W[r + n * (c0 + cc)] = wv;
                    nacc[cc] = fma(wv, wv, nacc[cc]);
                }
        }
        block_sum<8>(sm, nacc, tc, sm.nrm + c0);
    }
    if ((int)threadIdx.x < kk) sm.nrm[threadIdx.x] = sqrt(sm.nrm[threadIdx.x]);
    __syncthreads();
}

// 2-norms of the nc columns of P (n x nc, ld) -> sm.nrm[0:nc), nc <= MD
__device__ void col_norms(BatchSmem &sm, int64_t n, const double *P, int64_t ld, int nc) {
    for (int c0 = 0; c0 < nc; c0 += 8) {
        const int tc = min(8, nc - c0);
        double acc[8];
#pragma unroll
        for (int cc = 0; cc < 8; ++cc) acc[cc] = 0.0;
        for (int64_t r = threadIdx.x; r < n; r += BT)
#pragma unroll
            for (int cc = 0; cc < 8; ++cc)
                if (cc < tc) {
                    const double v = P[r + ld * (c0 + cc)];
                    acc[cc] = fma(v, v, acc[cc]);
                }
        block_sum<8>(sm, acc, tc, sm.nrm + c0);
    }
    if ((int)threadIdx.x < nc) sm.nrm[threadIdx.x] = sqrt(sm.nrm[threadIdx.x]);
    __syncthreads();
}

// T[:, q] -= sum_j Q[:, j] C[j + LDU q] for q < nt (C in LDS, sm.y)
__device__ void subtract_proj(BatchSmem &sm, int64_t n, const double *Q, int mq, double *T, int nt) {
    for (int q0 = 0; q0 < nt; q0 += 8) {
        const int tq = min(8, nt - q0);
        for (int64_t r = threadIdx.x; r < n; r += BT) {
            double acc[8];
#pragma unroll
            for (int qq = 0; qq < 8; ++qq) acc[qq] = 0.0;
            for (int j = 0; j < mq; ++j) {
                const double v = Q[r + n * j];
#pragma unroll
                for (int qq = 0; qq < 8; ++qq)
                    acc[qq] = fma(v, sm.y[j + LDU * min(q0 + qq, nt - 1)], acc[qq]);
            }
#pragma unroll
            for (int qq = 0; qq < 8; ++qq)
                if (qq < tq) T[r + n * (q0 + qq)] -= acc[qq];
        }
    }
    __syncthreads();
}

// Orthonormalise the nt <= KM columns of T (n x nt) among themselves:
// S = T^T T, Cholesky S = L L^T with pivots <= 1e-20 dropped, T <- T L^-T,
// kept columns packed to the front.  Returns the number kept.  The
// factorisation runs on wave 0 (lane i owns row i of L), the transform on all
// threads (one row of T each).
__device__ int chol_qr(BatchSmem &sm, int64_t n, double *T, int nt) {
    const int tid = threadIdx.x;
    dots(sm, n, T, n, nt, T, n, nt, sm.u, LDU);   // S in sm.u
    double *L = sm.y;
    if (tid < 64) {
        const int i = tid;
        for (int j = 0; j < nt; ++j) {
            // pivot (every lane computes it from the same data)
            double d = sm.u[j + LDU * j];
            for (int p = 0; p < j; ++p) d -= L[j + LDU * p] * L[j + LDU * p];
            const bool keep = d > 1.0e-20;
            const double ljj = keep ? sqrt(d) : 0.0;
            if (i > j && i < nt) {
                double x = 0.0;
                if (keep) {
                    x = sm.u[i + LDU * j];
                    for (int p = 0; p < j; ++p) x -= L[i + LDU * p] * L[j + LDU * p];
                    x /= ljj;
                }
                L[i + LDU * j] = x;
            }
            if (i == j) {
                L[j + LDU * j] = ljj;
                sm.perm[j] = keep ? 1 : 0;
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
    }
    __syncthreads();
    int kept = 0;
    for (int j = 0; j < nt; ++j) kept += sm.perm[j];
    // T <- T L^-T: row t, forward substitution t'_j = (t_j - sum_p L_jp t'_p) / L_jj
    for (int64_t r = tid; r < n; r += BT) {
        double t[KM];
#pragma unroll
        for (int j = 0; j < KM; ++j) t[j] = j < nt ? T[r + n * j] : 0.0;
#pragma unroll
        for (int j = 0; j < KM; ++j) {
            if (j < nt) {
                if (sm.perm[j]) {
                    double x = t[j];
#pragma unroll
                    for (int p = 0; p < j; ++p) x -= L[j + LDU * p] * t[p];
                    t[j] = x / L[j + LDU * j];
                } else {
                    t[j] = 0.0;
                }
            }
        }
        int q = 0;
#pragma unroll
        for (int j = 0; j < KM; ++j)
            if (j < nt && sm.perm[j]) T[r + n * q++] = t[j];
    }
    __syncthreads();
    return kept;
}

}  // namespace

// ---------------------------------------------------------------------------------------
// Step kernel: one work-group per molecule advances its state machine by one
// step.  The operator products come from batch_matvec_kernel, launched
// before each step over all active molecules.
//   B_INIT   start vectors (unit vectors on the lowest diagonal rows)
//   B_ITER   AV[:, m:m+mnew] is ready: Gram block of H, Jacobi on H, Ritz
//            vectors and residual norms, convergence test, restart,
//            preconditioned residuals, block Gram-Schmidt + Cholesky QR
//   B_REFINE AX = A64 X is ready: FP64 Rayleigh-Ritz in span(X), moments
//   B_VEL    W = (0.5 B) V is ready: velocity-corrected moments
__global__ __launch_bounds__(BT) void batch_step_kernel(const BatchMol *__restrict__ mols) {
    __shared__ BatchSmem sm;
    const BatchMol M = mols[blockIdx.x];
    int *st = M.state;
    const int phase = st[0];
    if (phase == B_DONE) return;
    const int tid = threadIdx.x;
    const int64_t n = M.n;
    const int k = M.k, maxdim = M.maxdim, nguess = M.nguess;
    double *V = M.V, *AV = M.AV, *X = M.X, *AX = M.AX, *W = M.W, *H = M.H;
    double *out = M.out;
    double *info = out + 20 * KM;
    // wall-clock ticks of the step's phases (thread 0): Gram, Jacobi, Ritz, expansion, final
    long long tk = wall_clock64();
    auto lap = [&](int ph) {
        const long long t = wall_clock64();
        if (tid == 0) info[4 + ph] += (double)(t - tk);
        tk = t;
    };

    if (phase == B_INIT) {
        for (int64_t e = tid; e < n * nguess; e += BT) V[e] = 0.0;
        for (int e = tid; e < MD * MD; e += BT) H[e] = 0.0;
        for (int e = tid; e < BATCH_OUT; e += BT) out[e] = 0.0;
        __syncthreads();
        if (tid < nguess) V[M.guess[tid] + n * tid] = 1.0;
        if (tid == 0) {
            st[1] = 0;        // m
            st[2] = nguess;   // mnew
            st[3] = 0;        // iterations
            st[4] = 0;        // kk
            st[0] = B_ITER;
        }
        return;
    }

    if (phase == B_ITER) {
        int m = st[1], mnew = st[2];
        const int it = st[3] + 1;
        const int mt = m + mnew;
        dots(sm, n, V, n, mt, AV + n * m, n, mnew, H + MD * m, MD);   // H[0:mt, m:mt]
        for (int e = tid; e < m * mnew; e += BT) {
            const int i = e % m, c = e / m;
            H[(m + c) + MD * i] = H[i + MD * (m + c)];
        }
        __syncthreads();
        m = mt;
        for (int e = tid; e < m * m; e += BT) {
            const int i = e % m, j = e / m;
            sm.u[i + LDU * j] = 0.5 * (H[i + MD * j] + H[j + MD * i]);
        }
        __syncthreads();
        lap(0);
        const int sw = jacobi(sm, m);
        lap(1);
        const int kk = min(k, m);
        if (tid < kk) sm.ths[tid] = sm.theta[sm.perm[tid]];
        __syncthreads();
        ritz(sm, n, m, kk, V, AV, X, AX, W);
        lap(2);
        int nun = 0;
        double rmax = 0.0;
        for (int c = 0; c < kk; ++c) {
            const double rn = sm.nrm[c];
            rmax = fmax(rmax, rn);
            if (!(rn < M.tol)) {
                if (tid == 0) sm.un[nun] = c;
                ++nun;
            }
        }
        bool finish = (nun == 0 && kk == k) || m == n || it >= M.maxiter;
        int acc = 0;
        if (!finish) {
            if (m + nun > maxdim) {   // restart with the Ritz vectors
                for (int64_t e = tid; e < n * kk; e += BT) {
                    V[e] = X[e];
                    AV[e] = AX[e];
                }
                for (int e = tid; e < MD * MD; e += BT) {
                    const int i = e % MD, j = e / MD;
                    H[e] = (i == j && i < kk) ? sm.ths[i] : 0.0;
                }
                m = kk;
            }
            const int room = min(nun, maxdim - m);
            __syncthreads();
            // preconditioned residuals (theta - diag)^-1 r as new directions
            for (int q = 0; q < room; ++q) {
                const int col = sm.un[q];
                const double th = sm.ths[col];
                for (int64_t r = tid; r < n; r += BT) {
                    double d = th - M.diag[r];
                    if (fabs(d) < 1.0e-4) d = d < 0.0 ? -1.0e-4 : 1.0e-4;
                    V[r + n * (m + q)] = W[r + n * col] / d;
                }
            }
            __syncthreads();
            col_norms(sm, n, V + n * m, n, room);
            for (int q = 0; q < room; ++q) {
                const double nq = sm.nrm[q];
                const double f = nq > 0.0 ? 1.0 / nq : 0.0;
                for (int64_t r = tid; r < n; r += BT) V[r + n * (m + q)] *= f;
            }
            __syncthreads();
            for (int pass = 0; pass < 2 && m > 0; ++pass) {
                dots(sm, n, V, n, m, V + n * m, n, room, sm.y, LDU);
                subtract_proj(sm, n, V, m, V + n * m, room);
            }
            acc = room;
            for (int pass = 0; pass < 2 && acc > 0; ++pass) acc = chol_qr(sm, n, V + n * m, acc);
            lap(3);
            finish = acc == 0;   // no new direction left: converged to working precision
        }
        if (tid == 0) {
            info[0] = (double)it;
            info[3] = rmax;
            info[10] += (double)sw;
            st[3] = it;
            if (finish) {
                info[1] = (nun == 0 || m == n || acc == 0) ? 1.0 : 0.0;
                info[2] = (double)kk;
                st[4] = kk;
                st[0] = B_REFINE;   // X holds the Ritz vectors
            } else {
                st[1] = m;
                st[2] = acc;
            }
        }
        return;
    }

    if (phase == B_REFINE) {   // AX = A64 X: FP64 Rayleigh-Ritz in span(X)
        const int kk = st[4];
        dots(sm, n, X, n, kk, AX, n, kk, sm.u, LDU);
        for (int e = tid; e < kk * kk; e += BT) {
            const int i = e % kk, j = e / kk;
            if (i < j) {
                const double a = 0.5 * (sm.u[i + LDU * j] + sm.u[j + LDU * i]);
                sm.u[i + LDU * j] = a;
                sm.u[j + LDU * i] = a;
            }
        }
        __syncthreads();
        jacobi(sm, kk);
        if (tid < kk) sm.ths[tid] = sm.theta[sm.perm[tid]];
        __syncthreads();
        ritz(sm, n, kk, kk, X, AX, V, AV, W);   // refined vectors -> V, A V -> AV
        if (tid < kk) {
            out[tid] = sm.ths[tid];
            out[19 * KM + tid] = sm.nrm[tid];
        }
        if (M.mom) dots(sm, n, M.mom, n, 9, V, n, kk, out + KM, 9);
        lap(4);
        if (tid == 0) st[0] = (M.velcorr && M.B64 && M.mom) ? B_VEL : B_DONE;
        return;
    }

    if (phase == B_VEL) {   // W = (0.5 B) V: X_new = W / (omega + 1e-8) (apbtrafo.f:58-69)
        const int kk = st[4];
        for (int c = 0; c < kk; ++c) {
            const double f = 1.0 / (out[c] + 1.0e-8);
            for (int64_t r = tid; r < n; r += BT) W[r + n * c] *= f;
        }
        __syncthreads();
        dots(sm, n, M.mom, n, 9, W, n, kk, out + 10 * KM, 9);
        lap(4);
        if (tid == 0) st[0] = B_DONE;
    }
}

// ---------------------------------------------------------------------------------------
// Operator products of all active molecules on the matrix cores: work-group
// (molecule, 64-row block).  B_ITER: AV[:, m:m+mnew] = A32 V[:, m:m+mnew]
// (FP32 operands and accumulation, v_mfma_f32_16x16x4f32); B_REFINE: AX = A64 X;
// B_VEL: W = B64 V[:, 0:kk] (FP64, v_mfma_f64_16x16x4f64).  The product is
// out(64 rows, nc) = A(rows, 0:n) V(0:n, 0:nc): wave w owns rows 16 w .. +16
// and 32 columns (two 16 x 16 blocks) per pass over the columns; K = n
// advances in chunks of 16, A (column-major, coalesced over the rows) and the
// vector rows (widened or rounded to the product's type) staged through LDS.
// MFMA 16x16x4 layouts (lane l, g = l / 16, c = l % 16): A[c][g], B[g][c];
// the FP64 result register i holds D[4i + g][c], the FP32 one D[4g + i][c].
constexpr int MV_BT = 256;
constexpr int MV_ROWS = BATCH_MV_ROWS, MV_COLS = 32, MV_KC = 16;
typedef double mv_v4d __attribute__((ext_vector_type(4)));
typedef float mv_v4f __attribute__((ext_vector_type(4)));

template <typename T>
struct MvSmem {
    T a[MV_KC][MV_ROWS + 1];
    T v[MV_KC][MV_COLS + 1];
};

template <typename T>
__device__ void matvec_rows_mfma(MvSmem<T> &sm, int64_t n, int64_t r0, const T *A,
                                 const double *Vin, int nc, double *out) {
    using Acc = typename std::conditional<sizeof(T) == 8, mv_v4d, mv_v4f>::type;
    const int t = threadIdx.x, l = t & 63, g = l >> 4, c = l & 15, w = t >> 6;
    for (int c0 = 0; c0 < nc; c0 += MV_COLS) {
        Acc acc[2] = {Acc{0, 0, 0, 0}, Acc{0, 0, 0, 0}};
        for (int64_t k0 = 0; k0 < n; k0 += MV_KC) {
            __syncthreads();
            for (int e = t; e < MV_ROWS * MV_KC; e += MV_BT) {
                const int rr = e % MV_ROWS, kk = e / MV_ROWS;
                const int64_t gr = r0 + rr, gk = k0 + kk;
                sm.a[kk][rr] = (gr < n && gk < n) ? A[gr + n * gk] : T(0);
            }
            for (int e = t; e < MV_KC * MV_COLS; e += MV_BT) {
                const int kk = e % MV_KC, cc = e / MV_KC;
                const int64_t gk = k0 + kk;
                sm.v[kk][cc] = (gk < n && c0 + cc < nc) ? (T)Vin[gk + n * (c0 + cc)] : T(0);
            }
            __syncthreads();
#pragma unroll
            for (int kk = 0; kk < MV_KC; kk += 4) {
                const T a = sm.a[kk + g][16 * w + c];
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    const T b = sm.v[kk + g][16 * j + c];
                    if constexpr (sizeof(T) == 8)
                        acc[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[j], 0, 0, 0);
                    else
                        acc[j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc[j], 0, 0, 0);
                }
            }
        }
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int row = sizeof(T) == 8 ? 4 * e + g : 4 * g + e;
                const int64_t gr = r0 + 16 * w + row;
                const int col = c0 + 16 * j + c;
                if (gr < n && col < nc) out[gr + n * col] = (double)acc[j][e];
            }
    }
}

__global__ __launch_bounds__(MV_BT) void batch_matvec_kernel(const BatchMol *__restrict__ mols,
                                                             const int2 *__restrict__ tiles) {
    __shared__ union {
        MvSmem<float> f;
        MvSmem<double> d;
    } sm;
    const int2 t = tiles[blockIdx.x];
    const BatchMol &M = mols[t.x];
    const int *st = M.state;
    const int phase = st[0];
    const int64_t n = M.n, r0 = (int64_t)t.y * MV_ROWS;
    if (phase == B_ITER) {
        const int m = st[1], mnew = st[2];
        matvec_rows_mfma<float>(sm.f, n, r0, M.A32, M.V + n * m, mnew, M.AV + n * m);
    } else if (phase == B_REFINE) {
        matvec_rows_mfma<double>(sm.d, n, r0, M.A64, M.X, st[4], M.AX);
    } else if (phase == B_VEL) {
        matvec_rows_mfma<double>(sm.d, n, r0, M.B64, M.V, st[4], M.W);
    }
}

void launch_batch_step(hipStream_t s, int nmol, const BatchMol *mols) {
    if (nmol <= 0) return;
    hipLaunchKernelGGL(batch_step_kernel, dim3(nmol), dim3(BT), 0, s, mols);
}

void launch_batch_matvec(hipStream_t s, int ntiles, const BatchMol *mols, const int2 *tiles) {
    if (ntiles <= 0) return;
    hipLaunchKernelGGL(batch_matvec_kernel, dim3(ntiles), dim3(MV_BT), 0, s, mols, tiles);
}

}  // namespace stda
