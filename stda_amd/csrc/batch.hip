// Batched sTDA eigensolver for many small molecules: one workgroup runs the
// whole block Davidson of one molecule (matvecs, Gram products, subspace
// eigensolve, Ritz vectors, residuals, preconditioning, orthonormalisation,
// FP64 refinement and transition moments) inside a single launch for the
// whole batch.  Mixed precision as in the single-molecule mixed path: the
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

#include <hip/hip_runtime.h>

namespace stda {

namespace {

constexpr int BT = 512;            // threads per molecule (8 waves)
constexpr int NWV = BT / 64;
constexpr int MD = BATCH_MD;
constexpr int KM = BATCH_KMAX;
constexpr int JB = 128;            // vector rows staged per LDS pass of a matvec
constexpr int CW32 = 32;           // columns per FP32 matvec pass
constexpr int CW64 = KM;           // columns per FP64 matvec pass
constexpr int DI = 4, DC = 8;      // tile of the dot-product routine
constexpr int JU = (MD / 2) * MD / BT;   // Jacobi items per thread and phase
static_assert(JU * BT == (MD / 2) * MD, "Jacobi items tile the phase");

struct BatchSmem {
    double u[MD * MD];     // subspace matrix under Jacobi rotation; staged vector rows
    double y[MD * MD];     // subspace eigenvectors; projection coefficients
    double theta[MD];      // eigenvalues in Jacobi order
    double ths[MD];        // eigenvalues of the kept roots, ascending
    double rc[MD / 2], rs[MD / 2];
    double red[NWV * 64];
    double nrm[MD];
    int pa[MD / 2], pb[MD / 2];
    int perm[MD];
    int un[KM];
    int ctl[4];
};

__device__ __forceinline__ double wsum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// out[q] = sum over the workgroup of v[q], q < nv (nv <= NV <= 64); out may be
// LDS or global.  Ends with a barrier.
template <int NV>
__device__ void block_sum(BatchSmem &sm, const double (&v)[NV], int nv, double *out) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NV; ++q) {
        double s = wsum(v[q]);
        if (lane == 0) sm.red[w * 64 + q] = s;
    }
    __syncthreads();
    if ((int)threadIdx.x < nv) {
        double s = 0.0;
        for (int ww = 0; ww < NWV; ++ww) s += sm.red[ww * 64 + threadIdx.x];
        out[threadIdx.x] = s;
    }
    __syncthreads();
}

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
#pragma unroll
            for (int e = 0; e < DI * DC; ++e) {
                double s = wsum(acc[e]);
                if (lane == 0) sm.red[w * 64 + e] = s;
            }
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

// out[:, c] = A V[:, c] for c < nc with A (n x n, column-major, symmetric) in
// precision T: the vector rows are staged through LDS in precision T, the
// products are accumulated in T, the result is stored as FP64.
template <typename T, int CW>
__device__ void matvec(BatchSmem &sm, int64_t n, const T *A, const double *V, int64_t ldv, int nc,
                       double *out, int64_t ldo) {
    static_assert(JB * CW * sizeof(T) <= sizeof(double) * MD * MD, "stage fits the LDS area");
    T *stage = reinterpret_cast<T *>(sm.u);
    for (int c0 = 0; c0 < nc; c0 += CW) {
        const int cw = min(CW, nc - c0);
        for (int64_t r0 = 0; r0 < n; r0 += BT) {
            const int64_t r = r0 + threadIdx.x;
            T acc[CW];
#pragma unroll
            for (int cc = 0; cc < CW; ++cc) acc[cc] = (T)0;
            for (int64_t j0 = 0; j0 < n; j0 += JB) {
                const int jb = (int)min((int64_t)JB, n - j0);
                __syncthreads();
                for (int e = threadIdx.x; e < JB * CW; e += BT) {
                    const int jj = e % JB, cc = e / JB;
                    stage[jj * CW + cc] =
                        (jj < jb && cc < cw) ? (T)V[(j0 + jj) + ldv * (c0 + cc)] : (T)0;
                }
                __syncthreads();
                if (r < n) {
                    const T *Ar = A + r + n * j0;
                    for (int jj = 0; jj < jb; ++jj) {
                        const T a = Ar[n * jj];
#pragma unroll
                        for (int cc = 0; cc < CW; ++cc) acc[cc] = fma(a, stage[jj * CW + cc], acc[cc]);
                    }
                }
            }
            if (r < n) {
#pragma unroll
                for (int cc = 0; cc < CW; ++cc)
                    if (cc < cw) out[r + ldo * (c0 + cc)] = (double)acc[cc];
            }
        }
    }
    __syncthreads();
}

// Eigen-decomposition of the symmetric m x m matrix in sm.u (leading dimension
// MD), m <= MD: cyclic Jacobi (two-sided rotations, Golub & Van Loan,
// Matrix Computations, Alg. 8.5.2/8.5.3) with the round-robin schedule, the
// m/2 disjoint rotations of a step applied in parallel.  A rotation is
// skipped when |a_pq| < 1e-17 (|a_pp| + |a_qq|); the sweeps stop when a whole
// sweep rotates nothing.  Out: sm.theta (Jacobi order), eigenvectors in the
// columns of sm.y, sm.perm = indices in ascending eigenvalue order.
__device__ int jacobi(BatchSmem &sm, int m) {
    const int tid = threadIdx.x;
    for (int e = tid; e < MD * MD; e += BT) sm.y[e] = (e % MD == e / MD) ? 1.0 : 0.0;
    const int me = m + (m & 1), h = me / 2;
    __syncthreads();
    int sweep = 0;
    for (; sweep < 40 && m > 1; ++sweep) {
        if (tid == 0) sm.ctl[0] = 0;
        __syncthreads();
        for (int r = 0; r < me - 1; ++r) {
            if (tid < h) {
                int a, b;
                if (tid == 0) {
                    a = me - 1;
                    b = r;
                } else {
                    a = (r + tid) % (me - 1);
                    b = (r - tid + me - 1) % (me - 1);
                }
                if (a > b) { const int t = a; a = b; b = t; }
                double c = 1.0, s = 0.0;
                if (b < m) {
                    const double app = sm.u[a + MD * a], aqq = sm.u[b + MD * b];
                    const double apq = sm.u[a + MD * b];
                    if (fabs(apq) > 1e-300 && fabs(apq) >= 1e-17 * (fabs(app) + fabs(aqq))) {
                        const double tau = (aqq - app) / (2.0 * apq);
                        const double t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
                        c = 1.0 / sqrt(1.0 + t * t);
                        s = t * c;
                        sm.ctl[0] = 1;
                    }
                }
                sm.pa[tid] = a;
                sm.pb[tid] = b;
                sm.rc[tid] = c;
                sm.rs[tid] = s;
            }
            __syncthreads();
            // rows a, b: J^T A.  Each thread owns JU items (h m <= JU BT); all
            // loads are issued before the stores.
            {
                int ia[JU], ib[JU];
                double cs[JU], sn[JU], xa[JU], xb[JU];
#pragma unroll
                for (int u = 0; u < JU; ++u) {
                    const int e = tid + u * BT;
                    sn[u] = 0.0;
                    if (e < h * m) {
                        const int p = e / m, j = e - p * m;
                        sn[u] = sm.rs[p];
                        cs[u] = sm.rc[p];
                        ia[u] = sm.pa[p] + MD * j;
                        ib[u] = sm.pb[p] + MD * j;
                    }
                }
#pragma unroll
                for (int u = 0; u < JU; ++u)
                    if (sn[u] != 0.0) { xa[u] = sm.u[ia[u]]; xb[u] = sm.u[ib[u]]; }
#pragma unroll
                for (int u = 0; u < JU; ++u)
                    if (sn[u] != 0.0) {
                        sm.u[ia[u]] = cs[u] * xa[u] - sn[u] * xb[u];
                        sm.u[ib[u]] = sn[u] * xa[u] + cs[u] * xb[u];
                    }
            }
            __syncthreads();
            // columns a, b: (J^T A) J and Y J
            {
                int ia[JU], ib[JU];
                double cs[JU], sn[JU], xa[JU], xb[JU], ya[JU], yb[JU];
#pragma unroll
                for (int u = 0; u < JU; ++u) {
                    const int e = tid + u * BT;
                    sn[u] = 0.0;
                    if (e < h * m) {
                        const int p = e / m, i = e - p * m;
                        sn[u] = sm.rs[p];
                        cs[u] = sm.rc[p];
                        ia[u] = i + MD * sm.pa[p];
                        ib[u] = i + MD * sm.pb[p];
                    }
                }
#pragma unroll
                for (int u = 0; u < JU; ++u)
                    if (sn[u] != 0.0) {
                        xa[u] = sm.u[ia[u]]; xb[u] = sm.u[ib[u]];
                        ya[u] = sm.y[ia[u]]; yb[u] = sm.y[ib[u]];
                    }
#pragma unroll
                for (int u = 0; u < JU; ++u)
                    if (sn[u] != 0.0) {
                        sm.u[ia[u]] = cs[u] * xa[u] - sn[u] * xb[u];
                        sm.u[ib[u]] = sn[u] * xa[u] + cs[u] * xb[u];
                        sm.y[ia[u]] = cs[u] * ya[u] - sn[u] * yb[u];
                        sm.y[ib[u]] = sn[u] * ya[u] + cs[u] * yb[u];
                    }
            }
            __syncthreads();
        }
        const bool rotated = sm.ctl[0] != 0;
        __syncthreads();
        if (!rotated) break;
    }
    if (tid < m) sm.theta[tid] = sm.u[tid + MD * tid];
    __syncthreads();
    if (tid < m) {
        const double v = sm.theta[tid];
        int rank = 0;
        for (int j = 0; j < m; ++j) {
            const double x = sm.theta[j];
            rank += (x < v) || (x == v && j < tid);
        }
        sm.perm[rank] = tid;
    }
    __syncthreads();
    return sweep;
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
                    const double yv = sm.y[j + MD * pc[cc]];
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

// T[:, q] -= sum_j Q[:, j] C[j + MD q] for q < nt (C in LDS, sm.y)
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
                    acc[qq] = fma(v, sm.y[j + MD * min(q0 + qq, nt - 1)], acc[qq]);
            }
#pragma unroll
            for (int qq = 0; qq < 8; ++qq)
                if (qq < tq) T[r + n * (q0 + qq)] -= acc[qq];
        }
    }
    __syncthreads();
}

}  // namespace

__global__ __launch_bounds__(BT) void davidson_batch_kernel(const BatchMol *__restrict__ mols) {
    __shared__ BatchSmem sm;
    const BatchMol M = mols[blockIdx.x];
    const int tid = threadIdx.x;
    const int64_t n = M.n;
    const int k = M.k, maxdim = M.maxdim, nguess = M.nguess;
    double *V = M.V, *AV = M.AV, *X = M.X, *AX = M.AX, *W = M.W, *H = M.H;

    // start vectors: unit vectors on the rows with the lowest diagonal
    for (int64_t e = tid; e < n * nguess; e += BT) V[e] = 0.0;
    for (int e = tid; e < MD * MD; e += BT) H[e] = 0.0;
    __syncthreads();
    if (tid < nguess) V[M.guess[tid] + n * tid] = 1.0;
    __syncthreads();

    int m = 0, mnew = nguess, it = 0, kk = 0, conv = 0, sweeps = 0;
    double rmax = 0.0;
    // wall-clock ticks per phase (thread 0): matvec, Gram, Jacobi, Ritz, expansion, final
    double tph[6] = {0, 0, 0, 0, 0, 0};
    long long tk = wall_clock64();
    auto lap = [&](int ph) {
        const long long t = wall_clock64();
        tph[ph] += (double)(t - tk);
        tk = t;
    };
    while (true) {
        ++it;
        matvec<float, CW32>(sm, n, M.A32, V + n * m, n, mnew, AV + n * m, n);
        lap(0);
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
            sm.u[i + MD * j] = 0.5 * (H[i + MD * j] + H[j + MD * i]);
        }
        __syncthreads();
        lap(1);
        sweeps += jacobi(sm, m);
        lap(2);
        kk = min(k, m);
        if (tid < kk) sm.ths[tid] = sm.theta[sm.perm[tid]];
        __syncthreads();
        ritz(sm, n, m, kk, V, AV, X, AX, W);
        lap(3);
        int nun = 0;
        rmax = 0.0;
        for (int c = 0; c < kk; ++c) {
            const double rn = sm.nrm[c];
            rmax = fmax(rmax, rn);
            if (!(rn < M.tol)) {
                if (tid == 0) sm.un[nun] = c;
                ++nun;
            }
        }
        if ((nun == 0 && kk == k) || m == n) { conv = 1; break; }
        if (it >= M.maxiter) break;
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
        // classical Gram-Schmidt against the subspace, twice
        for (int pass = 0; pass < 2 && m > 0; ++pass) {
            dots(sm, n, V, n, m, V + n * m, n, room, sm.y, MD);
            subtract_proj(sm, n, V, m, V + n * m, room);
        }
        // then among the new directions, dropping (near-)dependent ones
        int acc = 0;
        for (int q = 0; q < room; ++q) {
            double *t = V + n * (m + q);
            for (int pass = 0; pass < 2 && acc > 0; ++pass) {
                dots(sm, n, V + n * m, n, acc, t, n, 1, sm.y, MD);
                subtract_proj(sm, n, V + n * m, acc, t, 1);
            }
            col_norms(sm, n, t, n, 1);
            const double nq = sm.nrm[0];
            if (nq > 1.0e-10) {
                double *dst = V + n * (m + acc);
                const double f = 1.0 / nq;
                for (int64_t r = tid; r < n; r += BT) dst[r] = t[r] * f;
                ++acc;
            }
            __syncthreads();
        }
        mnew = acc;
        lap(4);
        if (mnew == 0) { conv = 1; break; }
    }

    // FP64 Rayleigh-Ritz refinement in the span of the kk Ritz vectors
    matvec<double, CW64>(sm, n, M.A64, X, n, kk, AX, n);
    dots(sm, n, X, n, kk, AX, n, kk, sm.u, MD);
    for (int e = tid; e < kk * kk; e += BT) {
        const int i = e % kk, j = e / kk;
        if (i < j) {
            const double a = 0.5 * (sm.u[i + MD * j] + sm.u[j + MD * i]);
            sm.u[i + MD * j] = a;
            sm.u[j + MD * i] = a;
        }
    }
    __syncthreads();
    jacobi(sm, kk);
    if (tid < kk) sm.ths[tid] = sm.theta[sm.perm[tid]];
    __syncthreads();
    ritz(sm, n, kk, kk, X, AX, V, AV, W);   // refined vectors -> V, A V -> AV
    double *out = M.out;
    if (tid < kk) {
        out[tid] = sm.ths[tid];
        out[19 * KM + tid] = sm.nrm[tid];
    }
    if (tid == 0) {
        out[20 * KM + 0] = (double)it;
        out[20 * KM + 1] = (double)conv;
        out[20 * KM + 2] = (double)kk;
        out[20 * KM + 3] = rmax;
        out[20 * KM + 10] = (double)sweeps;
    }
    __syncthreads();
    // transition moments of the roots and of the velocity-corrected vectors
    // X_new = (0.5 B) X / (omega + 1e-8)  (apbtrafo.f:58-69)
    for (int e = tid; e < 18 * KM; e += BT) out[KM + e] = 0.0;
    __syncthreads();
    if (M.mom) {
        dots(sm, n, M.mom, n, 9, V, n, kk, out + KM, 9);
        if (M.velcorr && M.B64) {
            matvec<double, CW64>(sm, n, M.B64, V, n, kk, W, n);
            for (int c = 0; c < kk; ++c) {
                const double f = 1.0 / (sm.ths[c] + 1.0e-8);
                for (int64_t r = tid; r < n; r += BT) W[r + n * c] *= f;
            }
            __syncthreads();
            dots(sm, n, M.mom, n, 9, W, n, kk, out + 10 * KM, 9);
        }
    }
    lap(5);
    if (tid == 0)
        for (int q = 0; q < 6; ++q) out[20 * KM + 4 + q] = tph[q];
}

void launch_davidson_batch(hipStream_t s, int nmol, const BatchMol *mols) {
    if (nmol <= 0) return;
    hipLaunchKernelGGL(davidson_batch_kernel, dim3(nmol), dim3(BT), 0, s, mols);
}

}  // namespace stda
