// Device helpers shared by the lock-step solvers (batch.hip: explicit
// operator, one work-group per molecule; wide.hip: matrix-free operator,
// grouped kernels): work-group reductions and the LDS Jacobi eigensolver
// of the subspace matrix.
#pragma once

#include "common.h"

#include <hip/hip_runtime.h>

namespace stda {

namespace {

constexpr int BT = 512;            // threads per molecule (8 waves)
constexpr int NWV = BT / 64;
constexpr int MD = BATCH_MD;
constexpr int KM = BATCH_KMAX;
constexpr int DI = 4, DC = 8;      // tile of the dot-product routine
constexpr int LDU = MD + 1;        // leading dimension of the LDS matrices (odd: no bank conflicts)

// LDS of a subspace of at most MDX vectors (leading dimension MDX + 1, odd:
// the rows of a column fall in distinct banks)
template <int MDX>
struct LockSmem {
    static constexpr int LDX = MDX + 1;
    double u[MDX * LDX];   // subspace matrix under Jacobi rotation; staged vector rows
    double y[MDX * LDX];   // subspace eigenvectors; projection coefficients
    double theta[MDX];     // eigenvalues in Jacobi order
    double ths[MDX];       // eigenvalues of the kept roots, ascending
    double rc[MDX / 2], rs[MDX / 2], rt[MDX / 2];
    double red[NWV * 64];
    double nrm[MDX];
    int pa[MDX / 2], pb[MDX / 2];
    int perm[MDX];
    int un[KM];
    int ctl[4];
};
using BatchSmem = LockSmem<MD>;

constexpr int ilog2(int x) { return x <= 1 ? 0 : 1 + ilog2(x / 2); }

__device__ __forceinline__ int bitrev(int x, int bits) {
    return bits == 0 ? 0 : (int)(__builtin_bitreverse32((unsigned)x) >> (32 - bits));
}

// Sums NV (a power of two <= 64) per-lane values over the wave by recursive
// halving: step s exchanges half of the remaining values with lane ^ 2^s, so
// log2(NV) steps move NV - 1 values in all; the last value is then summed
// over the remaining lane groups.  Lane l returns the total of value
// bitrev(l mod NV, log2 NV).  v is destroyed.
template <int NV>
__device__ __forceinline__ double wave_reduce(double (&v)[NV]) {
    constexpr int L = ilog2(NV);
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int s = 0; s < L; ++s) {
        const int half = NV >> (s + 1);
        // all-ones mask in the upper lanes; the halves are chosen with bit
        // operations so that v stays in registers (no indexed access)
        const long long um = -(long long)((lane >> s) & 1);
#pragma unroll
        for (int i = 0; i < half; ++i) {
            const long long lo = __double_as_longlong(v[i]);
            const long long hi = __double_as_longlong(v[i + half]);
            const double send = __longlong_as_double((lo & um) | (hi & ~um));
            const double keep = __longlong_as_double((hi & um) | (lo & ~um));
            v[i] = keep + __shfl_xor(send, 1 << s, 64);
        }
    }
    double t = v[0];
#pragma unroll
    for (int s = L; s < 6; ++s) t += __shfl_xor(t, 1 << s, 64);
    return t;
}

// out[q] = sum over the workgroup of v[q], q < nv (nv <= NV <= 64); out may be
// LDS or global.  Ends with a barrier.  v is destroyed.
template <int NV>
__device__ void block_sum(BatchSmem &sm, double (&v)[NV], int nv, double *out) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const double t = wave_reduce<NV>(v);
    if (lane < NV) sm.red[w * 64 + bitrev(lane, ilog2(NV))] = t;
    __syncthreads();
    if ((int)threadIdx.x < nv) {
        double s = 0.0;
        for (int ww = 0; ww < NWV; ++ww) s += sm.red[ww * 64 + threadIdx.x];
        out[threadIdx.x] = s;
    }
    __syncthreads();
}


// Eigen-decomposition of the symmetric m x m matrix in sm.u (leading dimension
// LDX), m <= MDX: cyclic Jacobi (two-sided rotations, Golub & Van Loan,
// Matrix Computations, Alg. 8.5.2/8.5.3) with the round-robin schedule, the
// m/2 disjoint rotations of a step applied in parallel.  A rotation is
// skipped when |a_pq| < rtol (|a_pp| + |a_qq|) (rtol = 1e-17 unless given); the
// sweeps stop when a whole sweep rotates nothing.  Out: sm.theta (Jacobi order), eigenvectors in the
// columns of sm.y, sm.perm = indices in ascending eigenvalue order.
template <int MDX>
__device__ int jacobi_t(LockSmem<MDX> &sm, int m, double rtol = 1.0e-17) {
    constexpr int LDX = LockSmem<MDX>::LDX;
    const int tid = threadIdx.x;
    for (int e = tid; e < LDX * MDX; e += BT) sm.y[e] = (e % LDX == e / LDX) ? 1.0 : 0.0;
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
                double c = 1.0, s = 0.0, t = 0.0;
                if (b < m) {
                    const double app = sm.u[a + LDX * a], aqq = sm.u[b + LDX * b];
                    const double apq = sm.u[a + LDX * b];
                    if (fabs(apq) > 1e-300 && fabs(apq) >= rtol * (fabs(app) + fabs(aqq))) {
                        const double tau = (aqq - app) / (2.0 * apq);
                        t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
                        c = 1.0 / sqrt(1.0 + t * t);
                        s = t * c;
                        sm.ctl[0] = 1;
                    }
                }
                sm.pa[tid] = a;
                sm.pb[tid] = b;
                sm.rc[tid] = c;
                sm.rs[tid] = s;
                sm.rt[tid] = t;
            }
            __syncthreads();
            // J^T A J for all pairs at once, by 2x2 blocks (pair P rows, pair Q
            // columns): A_PQ <- J_P^T A_PQ J_Q; diagonal blocks take the
            // closed form a_pp - t a_pq, a_qq + t a_pq, a_pq = 0.  Then Y <- Y J.
#pragma unroll
            for (int u = 0; u < ((MDX / 2) * (MDX / 2) + BT - 1) / BT; ++u) {
                const int e = tid + u * BT;
                const int P = e / (MDX / 2), Q = e % (MDX / 2);
                if (P >= h || Q >= h) continue;
                const double sP = sm.rs[P], sQ = sm.rs[Q];
                if (sP == 0.0 && sQ == 0.0) continue;
                const int aP = sm.pa[P], bP = sm.pb[P], aQ = sm.pa[Q], bQ = sm.pb[Q];
                if (P == Q) {
                    const double tP = sm.rt[P], apq = sm.u[aP + LDX * bP];
                    sm.u[aP + LDX * aP] -= tP * apq;
                    sm.u[bP + LDX * bP] += tP * apq;
                    sm.u[aP + LDX * bP] = 0.0;
                    sm.u[bP + LDX * aP] = 0.0;
                    continue;
                }
                const double cP = sm.rc[P], cQ = sm.rc[Q];
                const bool vb = bP < m, vq = bQ < m;   // the odd-size dummy index is never stored
                const double x00 = sm.u[aP + LDX * aQ];
                const double x01 = vq ? sm.u[aP + LDX * bQ] : 0.0;
                const double x10 = vb ? sm.u[bP + LDX * aQ] : 0.0;
                const double x11 = (vb && vq) ? sm.u[bP + LDX * bQ] : 0.0;
                const double r00 = cP * x00 - sP * x10, r01 = cP * x01 - sP * x11;
                const double r10 = sP * x00 + cP * x10, r11 = sP * x01 + cP * x11;
                sm.u[aP + LDX * aQ] = cQ * r00 - sQ * r01;
                if (vq) sm.u[aP + LDX * bQ] = sQ * r00 + cQ * r01;
                if (vb) sm.u[bP + LDX * aQ] = cQ * r10 - sQ * r11;
                if (vb && vq) sm.u[bP + LDX * bQ] = sQ * r10 + cQ * r11;
            }
#pragma unroll
            for (int u = 0; u < (MDX * (MDX / 2) + BT - 1) / BT; ++u) {
                const int e = tid + u * BT;
                const int i = e % MDX, Q = e / MDX;
                if (Q >= h || i >= m) continue;
                const double sQ = sm.rs[Q];
                if (sQ == 0.0) continue;
                const int aQ = sm.pa[Q], bQ = sm.pb[Q];
                const double cQ = sm.rc[Q];
                const double ya = sm.y[i + LDX * aQ], yb = sm.y[i + LDX * bQ];
                sm.y[i + LDX * aQ] = cQ * ya - sQ * yb;
                sm.y[i + LDX * bQ] = sQ * ya + cQ * yb;
            }
            __syncthreads();
        }
        const bool rotated = sm.ctl[0] != 0;
        __syncthreads();
        if (!rotated) break;
    }
    if (tid < m) sm.theta[tid] = sm.u[tid + LDX * tid];
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

__device__ __forceinline__ int jacobi(BatchSmem &sm, int m) { return jacobi_t<MD>(sm, m); }

}  // namespace

}  // namespace stda
