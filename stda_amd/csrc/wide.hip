// Lock-step sTDA Davidson for molecules too large for an explicit A (the
// "wide" molecules of a batch).  Every kernel here is grouped: one launch
// covers every active wide molecule, each work-group taking one tile of one
// molecule, so an iteration of the whole set is a fixed sequence of launches
// (captured once in a hipGraph by the driver) however many molecules and
// however uneven their sizes.
//
// Operator product (stda.f:2029-2053 is the matrix it applies, here without
// forming it): for the vectors of the current block, in the Loewdin AO basis,
//   X (pair grid) -> T = Cv X^T -> D = T Co^T -> G = aJ GamJ o D + aX GamK o D^T
//   + aK diag(u), u = gamK t, t_A = sum_{mu on A} D(mu,mu) -> W = G Co -> R = Cv^T W
//   -> AV = R(CSF pairs) + corr o V
// The four transforms are grouped MFMA GEMMs (gemm.hip); scatter, atom sums,
// G build and gather are the kernels below.  FP32 for the Davidson iterations
// (the atom sums come from T, and the D transform stores G itself), FP64 for
// the final Rayleigh-Ritz (A) and the velocity correction (0.5 B, with the
// exchange-type term that needs D transposed: D, then a separate G build).
//
// Davidson step (stda.f:1134-1150 is the eigenproblem; the properties follow
// stda.f:1207-1378 and apbtrafo.f:58-69), per iteration:
//   gram     H(:, new) = V^T AV(:, new)            chunk partials
//   sub      sum partials, Jacobi on H (LDS), Ritz coefficients Y, theta
//   ritz     X = V Y, AX = AV Y, W = AX - theta X  chunk partials of |W|^2
//   ctl      residual norms, convergence, restart decision
//   expand   (restart: V, AV <- X, AX) preconditioned residuals -> V(:, m:),
//            chunk partials of [V T]^T T
//   red      projection coefficients C = V^T T, squared norms of T
//   proj x2  T -= V C, partials of V^T T (then of T^T T)
//   chol     Cholesky of T^T T, columns with a relative pivot below 1e-20 dropped
//   tri      T <- T L^-T (kept columns packed), partials of T^T T
//   chol     second Cholesky (CholQR2), tri again
//   commit   next phase, subspace size
// Every sum over rows is a fixed sequence (chunk partials summed in chunk
// order), so a molecule's result does not depend on the rest of the batch.

#include "common.h"
#include "lockstep.h"

#include <hip/hip_runtime.h>

namespace stda {

namespace {

constexpr int WT = 256;                 // threads of the chunk and vector kernels
constexpr int WMD = WIDE_MD;            // subspace limit
constexpr int WLD = LockSmem<WMD>::LDX;
constexpr int GR = 32;                  // rows staged per step of gram_staged
constexpr int GLD = WMD + 1;           // LDS row length of gram_staged (odd: rows in distinct banks)
constexpr int GB_OWN = 3;               // output blocks per lane in gram_rows

// (molecule, index within the molecule) of this work-group in a grouped launch
__device__ __forceinline__ int2 locate(const int *start, int nmol) {
    const int q = grouped_item(start, nmol, (int)blockIdx.x);
    return make_int2(q, (int)blockIdx.x - start[q]);
}

// ----------------------------------------------------------------------------------
// Operator product kernels

// xf(pair[r], v) = source(r, v) for the rows of one chunk and every vector of the product
__global__ __launch_bounds__(WT) void wide_scatter_kernel(const WideMol *__restrict__ mols,
                                                          const int *__restrict__ start, int nmol) {
    const int2 t = locate(start, nmol);
    const WideMol &M = mols[t.x];
    const int *st = M.state;
    const int64_t n = M.n, r0 = (int64_t)t.y * WIDE_RC, r1 = min(n, r0 + WIDE_RC);
    const int c32 = wide_op_count(st, false), c64 = wide_op_count(st, true);
    if (c32 > 0) {
        const double *src = M.V + n * st[1];
        for (int v = 0; v < c32; ++v)
            for (int64_t r = r0 + threadIdx.x; r < r1; r += WT)
                M.xf32[M.pair[r] + M.npair * v] = (float)src[r + n * v];
    } else if (c64 > 0) {
        const double *src = st[0] == B_REFINE ? M.X : M.V;
        for (int v = 0; v < c64; ++v)
            for (int64_t r = r0 + threadIdx.x; r < r1; r += WT)
                M.xf64[M.pair[r] + M.npair * v] = src[r + n * v];
    }
}

// u(A, v) = sum_B gamK(A, B) t_B, t_B = sum_{mu on B} D(mu, mu): one work-group per (molecule, vector)
template <typename T>
__device__ void atomdiag_one(const WideMol &M, const T *D, const T *gk, T *u) {
    __shared__ T tb[WIDE_MAXATOMS];
    const int nao = M.nao, na = M.na;
    for (int B = threadIdx.x; B < na; B += WT) {
        T s = T(0);
        for (int mu = M.atom_ptr[B]; mu < M.atom_ptr[B + 1]; ++mu) s += D[mu + (int64_t)nao * mu];
        tb[B] = s;
    }
    __syncthreads();
    for (int A = threadIdx.x; A < na; A += WT) {
        T s = T(0);
        const T *g = gk + (int64_t)na * A;   // symmetric
        for (int B = 0; B < na; ++B) s += g[B] * tb[B];
        u[A] = s;
    }
}

// FP32 products: t_B from the half transform, D(mu, mu) = sum_j T(mu, j) Co(mu, j),
// before D itself is formed (its transform stores G directly, gemm.hip GammaEpi)
__device__ void atomdiag_from_t(const WideMol &M, const float *T, float *u) {
    __shared__ float dd[2048];
    __shared__ float tb[WIDE_MAXATOMS];
    const int nao = M.nao, na = M.na, no = M.no;
    const float *Co = M.C32;
    for (int mu0 = 0; mu0 < nao; mu0 += 2048) {
        const int cnt = min(2048, nao - mu0);
        __syncthreads();
        for (int x = threadIdx.x; x < cnt; x += WT) {
            const int mu = mu0 + x;
            float s = 0.0f;
#pragma unroll 8
            for (int j = 0; j < no; ++j) s += T[mu + (int64_t)nao * j] * Co[mu + (int64_t)nao * j];
            dd[x] = s;
        }
        __syncthreads();
        for (int B = threadIdx.x; B < na; B += WT) {
            float s = mu0 == 0 ? 0.0f : tb[B];
            const int m0 = max(M.atom_ptr[B], mu0), m1 = min(M.atom_ptr[B + 1], mu0 + cnt);
            for (int mu = m0; mu < m1; ++mu) s += dd[mu - mu0];
            tb[B] = s;
        }
    }
    __syncthreads();
    for (int A = threadIdx.x; A < na; A += WT) {
        float s = 0.0f;
        const float *g = M.gk32 + (int64_t)na * A;   // symmetric
        for (int B = 0; B < na; ++B) s += g[B] * tb[B];
        u[A] = s;
    }
}

template <bool F64>
__global__ __launch_bounds__(WT) void wide_atomdiag_kernel(const WideMol *__restrict__ mols,
                                                           const int *__restrict__ start, int nmol) {
    const int2 t = locate(start, nmol);
    const WideMol &M = mols[t.x];
    const int v = t.y;
    if (!F64) {
        if (v < wide_op_count(M.state, false))
            atomdiag_from_t(M, M.t32 + (int64_t)M.nao * M.no * v, M.u32 + (int64_t)M.na * v);
    } else if (v < wide_op_count(M.state, true)) {
        const int64_t nd = (int64_t)M.nao * M.nao;
        atomdiag_one<double>(M, M.d64 + nd * v, M.gk64, M.u64 + (int64_t)M.na * v);
    }
}

// G(nu, mu) for the AO columns mu0 .. mu0 + 32 of one vector.  aX != 0 needs
// D(mu, nu): 32 x 32 blocks of D are staged transposed through LDS.
template <typename T>
__device__ void gbuild_block(const WideMol &M, const T *D, T *G, const T *gj, const T *gk,
                             const T *u, T aJ, T aX, T aK, int mu0) {
    __shared__ T tr[WIDE_GB][WIDE_GB + 1];
    const int nao = M.nao, na = M.na;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // 32 x 8
    for (int nu0 = 0; nu0 < nao; nu0 += WIDE_GB) {
        if (aX != T(0)) {
            __syncthreads();
            for (int r = ty; r < WIDE_GB; r += 8) {
                const int mu = mu0 + tx, nu = nu0 + r;
                tr[r][tx] = (mu < nao && nu < nao) ? D[mu + (int64_t)nao * nu] : T(0);
            }
            __syncthreads();
        }
        for (int r = ty; r < WIDE_GB; r += 8) {
            const int nu = nu0 + tx, mu = mu0 + r;
            if (nu >= nao || mu >= nao) continue;
            const int An = M.atom_of[nu], Am = M.atom_of[mu];
            const int64_t g = An + (int64_t)na * Am;
            const int64_t e = nu + (int64_t)nao * mu;
            T val = T(0);
            if (aJ != T(0)) val += aJ * gj[g] * D[e];
            if (aX != T(0)) val += aX * gk[g] * tr[tx][r];
            if (nu == mu && aK != T(0)) val += aK * u[Am];
            G[e] = val;
        }
    }
}

__global__ __launch_bounds__(WT) void wide_gbuild_kernel(const WideMol *__restrict__ mols,
                                                         const int *__restrict__ start, int nmol) {
    const int2 t = locate(start, nmol);
    const WideMol &M = mols[t.x];
    const int nblk = (M.nao + WIDE_GB - 1) / WIDE_GB;
    const int v = t.y / nblk, mu0 = (t.y % nblk) * WIDE_GB;
    const int64_t nd = (int64_t)M.nao * M.nao;
    if (v < wide_op_count(M.state, true)) {   // FP32 products form G in their D transform
        const bool vel = M.state[0] == B_VEL;   // 0.5 B: aJ = 0, aX = -ax / 2, aK = ak / 2
        const double aJ = vel ? 0.0 : -1.0, aX = vel ? -0.5 * M.ax : 0.0;
        const double aK = vel ? 0.5 * M.ak : M.ak;
        gbuild_block<double>(M, M.d64 + nd * v, M.g64 + nd * v, M.gj64, M.gk64,
                             M.u64 + (int64_t)M.na * v, aJ, aX, aK, mu0);
    }
}

// destination(r, v) = R(pair[r], v) + corr(r) source(r, v); 0.5 B: R(pair[r], v) / (omega_v + 1e-8)
__global__ __launch_bounds__(WT) void wide_gather_kernel(const WideMol *__restrict__ mols,
                                                         const int *__restrict__ start, int nmol) {
    const int2 t = locate(start, nmol);
    const WideMol &M = mols[t.x];
    const int *st = M.state;
    const int64_t n = M.n, r0 = (int64_t)t.y * WIDE_RC, r1 = min(n, r0 + WIDE_RC);
    const int c32 = wide_op_count(st, false), c64 = wide_op_count(st, true);
    if (c32 > 0) {
        const double *src = M.V + n * st[1];
        double *dst = M.AV + n * st[1];
        for (int v = 0; v < c32; ++v)
            for (int64_t r = r0 + threadIdx.x; r < r1; r += WT)
                dst[r + n * v] = (double)M.r32[M.pair[r] + M.npair * v] + M.corr[r] * src[r + n * v];
    } else if (c64 > 0) {
        if (st[0] == B_REFINE) {
            for (int v = 0; v < c64; ++v)
                for (int64_t r = r0 + threadIdx.x; r < r1; r += WT)
                    M.AX[r + n * v] = M.r64[M.pair[r] + M.npair * v] + M.corr[r] * M.X[r + n * v];
        } else {   // B_VEL: X_new = (0.5 B X) / (omega + 1e-8)  (apbtrafo.f:58-69)
            for (int v = 0; v < c64; ++v) {
                const double f = 1.0 / (M.out[v] + 1.0e-8);
                for (int64_t r = r0 + threadIdx.x; r < r1; r += WT)
                    M.W[r + n * v] = M.r64[M.pair[r] + M.npair * v] * f;
            }
        }
    }
}

// ----------------------------------------------------------------------------------
// Row-chunk helpers (256 threads)

struct GramSmem {
    double p[GR][GLD];
    double q[GR][GLD];
};

// Chunk Gram products out[i + WMD c] = sum_{r0 <= r < r1} A(r, i) B(r, c),
// i < ni <= WMD, c < nc <= 40 (ceil(ni/4) ceil(nc/4) <= 64 GB_OWN), over row
// blocks of GR staged in LDS: P (np columns) into sm.p, Q (nq columns) into
// sm.q; A = sm.p (or sm.q when QQ), B = sm.q.  `xf` may rewrite the staged Q
// block (and store it) before the products (XF: it does, a barrier follows).
// Only the ceil(ni/4) x ceil(nc/4) output blocks of 4 x 4 are formed: lane l
// of every wave owns blocks l, l + 64, l + 128 (block b: rows 4 (b % nbi),
// columns 4 (b / nbi)), and wave w takes rows w, w + 4, ... of each block, so
// all four SIMDs share the products; the waves' partial sums are added in LDS
// at the end, in wave order.  The loads of the next row block are in flight
// while the current one is multiplied.
template <bool QQ, bool XF, typename Xf>
__device__ void gram_staged(GramSmem &sm, int64_t n, int64_t r0, int64_t r1, const double *P,
                            int np, const double *Q, int nq, int ni, int nc, double *out, Xf xf) {
    constexpr int PER = GR * WMD / WT;   // staged elements per thread and matrix
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int nbi = (ni + 3) >> 2, nbc = (nc + 3) >> 2, nblk = nbi * nbc;
    int bi[GB_OWN], bc[GB_OWN];
#pragma unroll
    for (int j = 0; j < GB_OWN; ++j) {
        const int b = lane + 64 * j;
        bi[j] = b < nblk ? 4 * (b % nbi) : -1;
        bc[j] = b < nblk ? 4 * (b / nbi) : 0;
    }
    double acc[GB_OWN][16];
#pragma unroll
    for (int j = 0; j < GB_OWN; ++j)
#pragma unroll
        for (int e = 0; e < 16; ++e) acc[j][e] = 0.0;
    double rp[PER], rq[PER];
    auto load = [&](int64_t rb) __attribute__((always_inline)) {
#pragma unroll
        for (int p = 0; p < PER; ++p) {
            const int e = t + WT * p, rr = e % GR, col = e / GR;
            const int64_t r = rb + rr;
            const bool in = r < r1;
            rp[p] = (in && col < np) ? P[r + n * col] : 0.0;
            rq[p] = (in && col < nq) ? Q[r + n * col] : 0.0;
        }
    };
    load(r0);
    for (int64_t rb = r0; rb < r1; rb += GR) {
        __syncthreads();
#pragma unroll
        for (int p = 0; p < PER; ++p) {
            const int e = t + WT * p, rr = e % GR, col = e / GR;
            sm.p[rr][col] = rp[p];
            sm.q[rr][col] = rq[p];
        }
        __syncthreads();
        if (rb + GR < r1) load(rb + GR);
        if (XF) {
            xf(rb);
            __syncthreads();
        }
#pragma unroll 1
        for (int rr = w; rr < GR; rr += WT / 64) {
#pragma unroll
            for (int j = 0; j < GB_OWN; ++j) {
                if (bi[j] < 0) continue;
                double a[4], b[4];
#pragma unroll
                for (int x = 0; x < 4; ++x) {
                    a[x] = QQ ? sm.q[rr][bi[j] + x] : sm.p[rr][bi[j] + x];
                    b[x] = sm.q[rr][bc[j] + x];
                }
#pragma unroll
                for (int x = 0; x < 4; ++x)
#pragma unroll
                    for (int y = 0; y < 4; ++y) acc[j][4 * x + y] = fma(a[x], b[y], acc[j][4 * x + y]);
            }
        }
    }
    // wave partial sums: waves 1, 2, 3 in turn through LDS into wave 0
    double *buf = &sm.p[0][0];   // 64 lanes x 16 GB_OWN values
    for (int w2 = 1; w2 < WT / 64; ++w2) {
        __syncthreads();
        if (w == w2)
#pragma unroll
            for (int j = 0; j < GB_OWN; ++j)
#pragma unroll
                for (int e = 0; e < 16; ++e) buf[(j * 16 + e) * 64 + lane] = acc[j][e];
        __syncthreads();
        if (w == 0)
#pragma unroll
            for (int j = 0; j < GB_OWN; ++j)
#pragma unroll
                for (int e = 0; e < 16; ++e) acc[j][e] += buf[(j * 16 + e) * 64 + lane];
    }
    if (w == 0)
#pragma unroll
        for (int j = 0; j < GB_OWN; ++j) {
            if (bi[j] < 0) continue;
#pragma unroll
            for (int x = 0; x < 4; ++x)
#pragma unroll
                for (int y = 0; y < 4; ++y) {
                    const int i = bi[j] + x, c = bc[j] + y;
                    if (i < ni && c < nc) out[i + WMD * c] = acc[j][4 * x + y];
                }
        }
    __syncthreads();
}

__device__ __forceinline__ void gram_rows(GramSmem &sm, int64_t n, int64_t r0, int64_t r1, const double *P,
                          int ni, const double *Q, int nc, double *out) {
    gram_staged<false, true>(sm, n, r0, r1, P, ni, Q, nc, ni, nc, out, [](int64_t) {});
}

// sums of NV per-thread values over the 256 threads: out[q] for q < nv (in thread order)
template <int NV>
__device__ void chunk_sum(double (&v)[NV], int nv, double *red, double *out) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const double s = wave_reduce<NV>(v);
    if (lane < NV) red[w * 64 + bitrev(lane, ilog2(NV))] = s;
    __syncthreads();
    if ((int)threadIdx.x < nv) {
        double a = 0.0;
        for (int ww = 0; ww < WT / 64; ++ww) a += red[ww * 64 + threadIdx.x];
        out[threadIdx.x] = a;
    }
    __syncthreads();
}

// sum of one element over the chunk partials, in chunk order (loads issued 8 at a time)
__device__ __forceinline__ double chunk_total(const double *part, int nch, int64_t idx) {
    double s = 0.0;
    int ch = 0;
    for (; ch + 8 <= nch; ch += 8) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = part[(int64_t)(ch + u) * WIDE_PART + idx];
#pragma unroll
        for (int u = 0; u < 8; ++u) s += v[u];
    }
    for (; ch < nch; ++ch) s += part[(int64_t)ch * WIDE_PART + idx];
    return s;
}

// ----------------------------------------------------------------------------------
// Davidson step kernels

__device__ __forceinline__ bool iter_active(const int *st) {
    return st[0] == B_ITER && st[WS_FINISH] == 0;
}

// gram: chunk partial sums of the subspace products (B_INIT: start vectors)
__global__ __launch_bounds__(WT) void wide_gram_kernel(const WideMol *__restrict__ mols,
                                                       const int *__restrict__ start, int nmol) {
    __shared__ GramSmem sm;
    const int2 t = locate(start, nmol);
    const WideMol &M = mols[t.x];
    const int *st = M.state;
    const int ph = st[0];
    const int64_t n = M.n, r0 = (int64_t)t.y * WIDE_RC, r1 = min(n, r0 + WIDE_RC);
    double *part = M.part + (int64_t)t.y * WIDE_PART;
    if (ph == B_INIT) {   // unit vectors on the lowest diagonal rows
        for (int q = 0; q < M.nguess; ++q)
            for (int64_t r = r0 + threadIdx.x; r < r1; r += WT) M.V[r + n * q] = 0.0;
        __syncthreads();
        if ((int)threadIdx.x < M.nguess) {
            const int64_t g = M.guess[threadIdx.x];
            if (g >= r0 && g < r1) M.V[g + n * threadIdx.x] = 1.0;
        }
        return;
    }
    // one call site: the operands by phase
    const double *P = nullptr, *Q = nullptr;
    int ni = 0, nc = 0;
    if (ph == B_ITER) {
        const int m = st[1], mnew = st[2];
        P = M.V; ni = m + mnew; Q = M.AV + n * m; nc = mnew;
    } else if (ph == B_REFINE) {
        P = M.X; ni = st[4]; Q = M.AX; nc = st[4];
    } else if (ph == B_VEL) {
        P = M.mom; ni = 9; Q = M.W; nc = st[4];
    } else {
        return;
    }
    gram_rows(sm, n, r0, r1, P, ni, Q, nc, part);
}

// sub: subspace eigenproblem (one work-group of BT threads per molecule)
__global__ __launch_bounds__(BT) void wide_sub_kernel(const WideMol *__restrict__ mols) {
    constexpr int PT = WMD * WMD / BT;   // matrix elements per thread
    __shared__ LockSmem<WMD> sm;
    const WideMol &M = mols[blockIdx.x];
    int *st = M.state;
    const int ph = st[0];
    const int tid = threadIdx.x, nch = M.nchunk;
    double *H = M.H, *aux = M.aux, *out = M.out;
    double *info = out + 20 * KM;
    if (ph == B_INIT) {
        for (int e = tid; e < WMD * WMD; e += BT) H[e] = 0.0;
        for (int e = tid; e < BATCH_OUT; e += BT) out[e] = 0.0;
        if (tid == 0) {
            st[1] = 0;
            st[2] = M.nguess;
            st[3] = 0;
            st[4] = 0;
            st[WS_NEXT] = B_ITER;
            st[WS_FINISH] = 0;
            st[WS_MP] = 0;
        }
        return;
    }
    if (ph == B_VEL) {   // velocity-corrected moments mom^T X_new
        const int kk = st[4];
        for (int e = tid; e < 9 * kk; e += BT) {
            const int q = e % 9, c = e / 9;
            out[10 * KM + q + 9 * c] = chunk_total(M.part, nch, q + WMD * c);
        }
        if (tid == 0) st[WS_NEXT] = B_DONE;
        return;
    }
    if (ph != B_ITER && ph != B_REFINE) return;
    int m, mp = 0;
    if (ph == B_ITER) {
        const int c0 = st[1], ncol = st[2];
        m = c0 + ncol;
        mp = st[WS_MP];
        // subspace matrix in sm.u: the earlier block from H, the new columns
        // (and their mirror) from the chunk partials; the new part goes to H
        for (int e = tid; e < m * m; e += BT) {
            const int i = e % m, j = e / m;
            double h;
            if (j >= c0) {
                h = chunk_total(M.part, nch, i + WMD * (j - c0));
                H[i + WMD * j] = h;
            } else if (i >= c0) {
                h = chunk_total(M.part, nch, j + WMD * (i - c0));
                H[i + WMD * j] = h;
            } else {
                h = H[i + WMD * j];
            }
            sm.u[i + WLD * j] = h;
        }
        // the last subspace's eigenvectors (mp x mp) in sm.y
        for (int e = tid; e < mp * mp; e += BT) {
            const int i = e % mp, j = e / mp;
            sm.y[i + WLD * j] = aux[WA_YP + i + WMD * j];
        }
        __syncthreads();
        // Warm start: the first mp basis vectors are those of the last
        // subspace, whose eigenvectors Yp diagonalise that block, so with
        // Q = diag(Yp, I) the matrix Q^T H Q is an arrow (diag(theta_old) and
        // the new rows and columns); Jacobi runs on it and Y = Q Y'.
        double z[PT];
#pragma unroll
        for (int u = 0; u < PT; ++u) {   // Hs Q, Hs = (H + H^T) / 2
            const int e = tid + BT * u, i = e % WMD, c = e / WMD;
            double v = 0.0;
            if (i < m && c < m) {
                if (c < mp)
                    for (int j = 0; j < mp; ++j)
                        v += 0.5 * (sm.u[i + WLD * j] + sm.u[j + WLD * i]) * sm.y[j + WLD * c];
                else
                    v = 0.5 * (sm.u[i + WLD * c] + sm.u[c + WLD * i]);
            }
            z[u] = v;
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < PT; ++u) {
            const int e = tid + BT * u, i = e % WMD, c = e / WMD;
            if (i < m && c < m) sm.u[i + WLD * c] = z[u];
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < PT; ++u) {   // Q^T (Hs Q)
            const int e = tid + BT * u, r = e % WMD, c = e / WMD;
            double v = 0.0;
            if (r < m && c < m) {
                if (r < mp)
                    for (int j = 0; j < mp; ++j) v += sm.y[j + WLD * r] * sm.u[j + WLD * c];
                else
                    v = sm.u[r + WLD * c];
            }
            z[u] = v;
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < PT; ++u) {
            const int e = tid + BT * u, i = e % WMD, c = e / WMD;
            if (i < m && c < m) sm.u[i + WLD * c] = z[u];
        }
        __syncthreads();
        for (int e = tid; e < m * m; e += BT) {   // exact symmetry
            const int i = e % m, j = e / m;
            if (i < j) {
                const double v = 0.5 * (sm.u[i + WLD * j] + sm.u[j + WLD * i]);
                sm.u[i + WLD * j] = v;
                sm.u[j + WLD * i] = v;
            }
        }
    } else {   // B_REFINE: X^T A64 X, symmetrised
        m = st[4];
        for (int e = tid; e < m * m; e += BT) {
            const int i = e % m, j = e / m;
            sm.u[i + WLD * j] = 0.5 * (chunk_total(M.part, nch, i + WMD * j) +
                                       chunk_total(M.part, nch, j + WMD * i));
        }
    }
    __syncthreads();
    const int sw = jacobi_t<WMD>(sm, m, 1.0e-14);
    const int kk = min(M.k, m);
    if (tid < kk) aux[WA_TH + tid] = sm.theta[sm.perm[tid]];
    if (ph == B_ITER) {
        // Y = Q Y' in ascending order, kept whole for the next warm start and
        // its kk lowest columns as the Ritz coefficients (Yp staged in sm.u,
        // free after the sweeps)
        for (int e = tid; e < mp * mp; e += BT) {
            const int i = e % mp, j = e / mp;
            sm.u[i + WLD * j] = aux[WA_YP + i + WMD * j];
        }
        __syncthreads();
        double z[PT];
#pragma unroll
        for (int u = 0; u < PT; ++u) {
            const int e = tid + BT * u, j = e % WMD, c = e / WMD;
            double v = 0.0;
            if (j < m && c < m) {
                const int pc = sm.perm[c];
                if (j < mp)
                    for (int i = 0; i < mp; ++i) v += sm.u[j + WLD * i] * sm.y[i + WLD * pc];
                else
                    v = sm.y[j + WLD * pc];
            }
            z[u] = v;
        }
#pragma unroll
        for (int u = 0; u < PT; ++u) {
            const int e = tid + BT * u, j = e % WMD, c = e / WMD;
            if (j < m && c < m) {
                aux[WA_YP + j + WMD * c] = z[u];
                if (c < kk) aux[WA_Y + j + WMD * c] = z[u];
            }
        }
    } else {
        for (int e = tid; e < m * kk; e += BT) {
            const int j = e % m, c = e / m;
            aux[WA_Y + j + WMD * c] = sm.y[j + WLD * sm.perm[c]];
        }
    }
    if (tid == 0) {
        if (ph == B_ITER) {
            st[1] = m;
            st[4] = kk;
            st[WS_MP] = m;
        }
        info[10] += (double)sw;
    }
}

// ritz: Ritz vectors, their products and residuals over one chunk; partial |W|^2
__global__ __launch_bounds__(WT) void wide_ritz_kernel(const WideMol *__restrict__ mols,
                                                       const int *__restrict__ start, int nmol) {
    __shared__ double y[WMD * KM];
    __shared__ double th[KM];
    __shared__ double red[WT];
    const int2 t = locate(start, nmol);
    const WideMol &M = mols[t.x];
    const int *st = M.state;
    const int ph = st[0];
    if (ph != B_ITER && ph != B_REFINE) return;
    const int64_t n = M.n, r0 = (int64_t)t.y * WIDE_RC, r1 = min(n, r0 + WIDE_RC);
    const int kk = st[4], m = ph == B_ITER ? st[1] : kk;
    const double *P = ph == B_ITER ? M.V : M.X, *AP = ph == B_ITER ? M.AV : M.AX;
    double *Xo = ph == B_ITER ? M.X : M.V, *AXo = ph == B_ITER ? M.AX : M.AV;
    // y[j * KM + c] = Y(j, c): one broadcast row per j
    for (int e = threadIdx.x; e < WMD * KM; e += WT) {
        const int j = e / KM, c = e % KM;
        y[e] = (j < m && c < kk) ? M.aux[WA_Y + j + WMD * c] : 0.0;
    }
    if ((int)threadIdx.x < KM) th[threadIdx.x] = (int)threadIdx.x < kk ? M.aux[WA_TH + threadIdx.x] : 0.0;
    __syncthreads();
    double nacc[32];
#pragma unroll
    for (int c = 0; c < 32; ++c) nacc[c] = 0.0;
    for (int64_t r = r0 + threadIdx.x; r < r1; r += WT) {
        double x[KM], ax[KM];
#pragma unroll
        for (int c = 0; c < KM; ++c) x[c] = ax[c] = 0.0;
#pragma unroll 8
        for (int j = 0; j < m; ++j) {
            const double v = P[r + n * j], av = AP[r + n * j];
#pragma unroll
            for (int c = 0; c < KM; ++c) {
                x[c] = fma(v, y[j * KM + c], x[c]);
                ax[c] = fma(av, y[j * KM + c], ax[c]);
            }
        }
#pragma unroll
        for (int c = 0; c < KM; ++c)
            if (c < kk) {
                const double w = ax[c] - th[c] * x[c];
                Xo[r + n * c] = x[c];
                AXo[r + n * c] = ax[c];
                M.W[r + n * c] = w;
                nacc[c] = fma(w, w, nacc[c]);
            }
    }
    chunk_sum<32>(nacc, kk, red, M.part + (int64_t)t.y * WIDE_PART);
}

// ctl: residual norms, convergence test, restart decision (one work-group per molecule)
__global__ __launch_bounds__(WT) void wide_ctl_kernel(const WideMol *__restrict__ mols) {
    const WideMol &M = mols[blockIdx.x];
    int *st = M.state;
    const int ph = st[0];
    if (ph != B_ITER && ph != B_REFINE) return;
    __shared__ double rn[KM];
    const int kk = st[4], nch = M.nchunk, tid = threadIdx.x;
    double *out = M.out, *info = out + 20 * KM;
    if (tid < kk) {
        const double s = chunk_total(M.part, nch, tid);
        rn[tid] = sqrt(s);
    }
    __syncthreads();
    if (ph == B_REFINE) {
        if (tid < kk) {
            out[tid] = M.aux[WA_TH + tid];
            out[19 * KM + tid] = rn[tid];
        }
        if (tid == 0) st[WS_NEXT] = M.mom ? B_REFINE : B_DONE;   // moments follow
        return;
    }
    __shared__ int rs_flag;
    if (tid == 0) {
        const int m = st[1], it = st[3] + 1;
        int nun = 0;
        double rmax = 0.0;
        for (int c = 0; c < kk; ++c) {
            rmax = fmax(rmax, rn[c]);
            if (!(rn[c] < M.tol)) st[WS_UN + nun++] = c;
        }
        const bool finish = (nun == 0 && kk == M.k) || m == M.n || it >= M.maxiter;
        st[3] = it;
        info[0] = (double)it;
        info[3] = rmax;
        st[WS_NUN] = nun;
        rs_flag = 0;
        if (finish) {
            info[1] = (nun == 0 || m == M.n) ? 1.0 : 0.0;
            info[2] = (double)kk;
            st[WS_FINISH] = 1;
            st[WS_NEXT] = B_REFINE;
        } else {
            const bool restart = m + nun > M.maxdim;
            const int mb = restart ? kk : m;
            st[WS_RESTART] = restart ? 1 : 0;
            if (restart) st[WS_MP] = 0;   // basis = Ritz vectors, H diagonal: no rotation needed
            st[1] = mb;
            st[WS_ROOM] = min(nun, M.maxdim - mb);
            st[WS_NEXT] = B_ITER;
            rs_flag = restart ? 1 : 0;
        }
    }
    __syncthreads();
    if (rs_flag)
        for (int e = tid; e < WMD * WMD; e += WT) {
            const int i = e % WMD, j = e / WMD;
            M.H[e] = (i == j && i < kk) ? M.aux[WA_TH + i] : 0.0;
        }
}

// (restart: V, AV <- X, AX) and the preconditioned residuals (theta - diag)^-1 r
// of the unconverged roots as the new columns V(:, m:m+room), rows r0 .. r1
__device__ __forceinline__ void expand_rows(const WideMol &M, int64_t n, int64_t r0, int64_t r1) {
    const int *st = M.state;
    const int m = st[1], room = st[WS_ROOM], kk = st[4];
    if (st[WS_RESTART])
        for (int c = 0; c < kk; ++c)
            for (int64_t r = r0 + threadIdx.x; r < r1; r += WT) {
                M.V[r + n * c] = M.X[r + n * c];
                M.AV[r + n * c] = M.AX[r + n * c];
            }
    for (int q = 0; q < room; ++q) {
        const int col = st[WS_UN + q];
        const double th = M.aux[WA_TH + col];
        for (int64_t r = r0 + threadIdx.x; r < r1; r += WT) {
            double d = th - M.diag[r];
            if (fabs(d) < 1.0e-4) d = d < 0.0 ? -1.0e-4 : 1.0e-4;
            M.V[r + n * (m + q)] = M.W[r + n * col] / d;
        }
    }
}

// expand: new directions (B_ITER) or moment partials (B_REFINE)
__global__ __launch_bounds__(WT) void wide_expand_kernel(const WideMol *__restrict__ mols,
                                                         const int *__restrict__ start, int nmol) {
    __shared__ GramSmem sm;
    const int2 t = locate(start, nmol);
    const WideMol &M = mols[t.x];
    const int *st = M.state;
    const int ph = st[0];
    const int64_t n = M.n, r0 = (int64_t)t.y * WIDE_RC, r1 = min(n, r0 + WIDE_RC);
    double *part = M.part + (int64_t)t.y * WIDE_PART;
    const double *P = nullptr, *Q = nullptr;
    int ni = 0, nc = 0;
    if (ph == B_REFINE) {
        if (!M.mom) return;
        P = M.mom; ni = 9; Q = M.V; nc = st[4];
    } else {
        if (!iter_active(st)) return;
        const int m = st[1], room = st[WS_ROOM];
        P = M.V; ni = m + room; Q = M.V + n * m; nc = room;
        expand_rows(M, n, r0, r1);
        __syncthreads();
    }
    gram_rows(sm, n, r0, r1, P, ni, Q, nc, part);
}


// red: sums of the partials of expand (B_ITER: C = V^T T, |T|^2; B_REFINE: moments)
__global__ __launch_bounds__(WT) void wide_red_kernel(const WideMol *__restrict__ mols) {
    const WideMol &M = mols[blockIdx.x];
    int *st = M.state;
    const int ph = st[0], nch = M.nchunk, tid = threadIdx.x;
    if (ph == B_REFINE) {
        if (!M.mom) return;
        const int kk = st[4];
        for (int e = tid; e < 9 * kk; e += WT) {
            const int q = e % 9, c = e / 9;
            const double s = chunk_total(M.part, nch, q + WMD * c);
            M.out[KM + q + 9 * c] = s;
        }
        if (tid == 0) st[WS_NEXT] = (M.velcorr && M.mom) ? B_VEL : B_DONE;
        return;
    }
    if (!iter_active(st)) return;
    const int m = st[1], room = st[WS_ROOM];
    for (int e = tid; e < (m + 1) * room; e += WT) {
        const int i = e % (m + 1), q = e / (m + 1);
        const int row = i < m ? i : m + q;   // the last row: the diagonal of T^T T
        const double s = chunk_total(M.part, nch, row + WMD * q);
        if (i < m) M.aux[WA_C + i + WMD * q] = s;
        else M.aux[WA_NR0 + q] = s;
    }
}

// proj: T -= V C with V and T staged once through LDS, then the partial
// products from the same staged blocks: V^T T (pass 0) or T^T T (pass 1)
template <int PASS>
__global__ __launch_bounds__(WT) void wide_proj_kernel(const WideMol *__restrict__ mols,
                                                       const int *__restrict__ start, int nmol) {
    __shared__ GramSmem sm;
    __shared__ double c[WMD * KM];
    const int2 t = locate(start, nmol);
    const WideMol &M = mols[t.x];
    const int *st = M.state;
    if (!iter_active(st)) return;
    const int64_t n = M.n, r0 = (int64_t)t.y * WIDE_RC, r1 = min(n, r0 + WIDE_RC);
    const int m = st[1], room = st[WS_ROOM];
    for (int e = threadIdx.x; e < WMD * KM; e += WT) {
        const int j = e / KM, q = e % KM;
        c[e] = (j < m && q < room) ? M.aux[WA_C + j + WMD * q] : 0.0;
    }
    __syncthreads();
    double *T = M.V + n * m;
    // staged block: sm.p = V(rows, 0:m), sm.q = T(rows, 0:room)
    auto sub = [&](int64_t rb) {
        double x[GR * KM / WT];
#pragma unroll
        for (int u = 0; u < GR * KM / WT; ++u) {
            const int e = threadIdx.x + WT * u, rr = e % GR, q = e / GR;
            double v = 0.0;
            if (q < room) {
                v = sm.q[rr][q];
                for (int j = 0; j < m; ++j) v = fma(-sm.p[rr][j], c[j * KM + q], v);
            }
            x[u] = v;
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < GR * KM / WT; ++u) {
            const int e = threadIdx.x + WT * u, rr = e % GR, q = e / GR;
            if (q < room) {
                sm.q[rr][q] = x[u];
                if (rb + rr < r1) T[rb + rr + n * q] = x[u];
            }
        }
    };
    double *part = M.part + (int64_t)t.y * WIDE_PART;
    if (PASS == 0) gram_staged<false, true>(sm, n, r0, r1, M.V, m, T, room, m, room, part, sub);
    else gram_staged<true, true>(sm, n, r0, r1, M.V, m, T, room, room, room, part, sub);
}

// red2: C = sum of the V^T T partials (second projection)
__global__ __launch_bounds__(WT) void wide_red2_kernel(const WideMol *__restrict__ mols) {
    const WideMol &M = mols[blockIdx.x];
    const int *st = M.state;
    if (!iter_active(st)) return;
    const int m = st[1], room = st[WS_ROOM], nch = M.nchunk;
    for (int e = threadIdx.x; e < m * room; e += WT) {
        const int i = e % m, q = e / m;
        const double s = chunk_total(M.part, nch, i + WMD * q);
        M.aux[WA_C + i + WMD * q] = s;
    }
}

// chol: Cholesky S = L L^T of the summed T^T T partials; a pivot d_j <=
// 1e-20 s_j drops column j (s_j: squared norm before the projections in
// pass 0, 1 after the first orthonormalisation).  Wave 0, lane i owns row i.
template <int PASS>
__global__ __launch_bounds__(64) void wide_chol_kernel(const WideMol *__restrict__ mols) {
    const WideMol &M = mols[blockIdx.x];
    int *st = M.state;
    if (!iter_active(st)) return;
    __shared__ double S[KM * KM], L[KM * KM];
    const int nt = PASS == 0 ? st[WS_ROOM] : st[WS_ACC1], nch = M.nchunk;
    const int i = threadIdx.x;
    for (int e = i; e < nt * nt; e += 64) {
        const int a = e % nt, b = e / nt;
        const double s = chunk_total(M.part, nch, a + WMD * b);
        S[a + KM * b] = s;
    }
    __syncthreads();
    int kept = 0;
    for (int j = 0; j < nt; ++j) {
        double d = S[j + KM * j];
        for (int p = 0; p < j; ++p) d -= L[j + KM * p] * L[j + KM * p];
        const double scale = PASS == 0 ? M.aux[WA_NR0 + j] : 1.0;
        const bool keep = d > 1.0e-20 * scale;
        const double ljj = keep ? sqrt(d) : 0.0;
        if (i > j && i < nt) {
            double x = 0.0;
            if (keep) {
                x = S[i + KM * j];
                for (int p = 0; p < j; ++p) x -= L[i + KM * p] * L[j + KM * p];
                x /= ljj;
            }
            L[i + KM * j] = x;
        }
        if (i == j) L[j + KM * j] = ljj;
        kept += keep ? 1 : 0;
        if (i == 0) st[WS_KEEP + j] = keep ? 1 : 0;
        __syncthreads();
    }
    for (int e = i; e < KM * KM; e += 64) M.aux[WA_L + e] = L[e];
    if (i == 0) st[PASS == 0 ? WS_ACC1 : WS_ACC2] = kept;
}

// tri: T <- T L^-T with the kept columns packed to the front; pass 0 also
// leaves the partials of the new T^T T
template <int PASS>
__global__ __launch_bounds__(WT) __attribute__((amdgpu_waves_per_eu(1, 1))) void wide_tri_kernel(const WideMol *__restrict__ mols,
                                                      const int *__restrict__ start, int nmol) {
    __shared__ GramSmem sm;
    __shared__ double L[KM * KM];
    __shared__ int keep[KM];
    const int2 t = locate(start, nmol);
    const WideMol &M = mols[t.x];
    const int *st = M.state;
    if (!iter_active(st)) return;
    const int64_t n = M.n, r0 = (int64_t)t.y * WIDE_RC, r1 = min(n, r0 + WIDE_RC);
    const int m = st[1], nt = PASS == 0 ? st[WS_ROOM] : st[WS_ACC1];
    for (int e = threadIdx.x; e < KM * KM; e += WT) L[e] = M.aux[WA_L + e];
    if ((int)threadIdx.x < KM) keep[threadIdx.x] = (int)threadIdx.x < nt ? st[WS_KEEP + threadIdx.x] : 0;
    __syncthreads();
    double *T = M.V + n * m;
    for (int64_t r = r0 + threadIdx.x; r < r1; r += WT) {
        double v[KM];
#pragma unroll
        for (int j = 0; j < KM; ++j) v[j] = j < nt ? T[r + n * j] : 0.0;
#pragma unroll
        for (int j = 0; j < KM; ++j) {
            if (j < nt) {
                if (keep[j]) {
                    double x = v[j];
#pragma unroll
                    for (int p = 0; p < j; ++p) x -= L[j + KM * p] * v[p];
                    v[j] = x / L[j + KM * j];
                } else {
                    v[j] = 0.0;
                }
            }
        }
        int q = 0;
#pragma unroll
        for (int j = 0; j < KM; ++j)
            if (j < nt && keep[j]) T[r + n * q++] = v[j];
    }
    if (PASS == 0) {
        __syncthreads();
        const int acc = st[WS_ACC1];
        gram_rows(sm, n, r0, r1, T, acc, T, acc, M.part + (int64_t)t.y * WIDE_PART);
    }
}

// commit: the phase and subspace size for the next iteration
__global__ void wide_commit_kernel(const WideMol *__restrict__ mols) {
    const WideMol &M = mols[blockIdx.x];
    int *st = M.state;
    if (threadIdx.x != 0) return;
    const int ph = st[0];
    if (ph == B_DONE) return;
    if (ph == B_ITER && st[WS_FINISH] == 0) {
        const int acc = st[WS_ACC2];
        st[2] = acc;
        if (acc == 0) {   // no new direction left: converged to working precision
            double *info = M.out + 20 * KM;
            info[1] = 1.0;
            info[2] = (double)st[4];
            st[WS_NEXT] = B_REFINE;
        }
    }
    const int next = st[WS_NEXT];
    if (next == B_REFINE && ph == B_ITER) st[WS_FINISH] = 0;
    st[0] = next;
}

}  // namespace

void launch_wide_scatter(hipStream_t s, int nmol, const WideMol *mols, WideStarts st) {
    if (nmol <= 0 || st.nchunks <= 0) return;
    hipLaunchKernelGGL(wide_scatter_kernel, dim3(st.nchunks), dim3(WT), 0, s, mols, st.chunk, nmol);
}

void launch_wide_atomdiag(hipStream_t s, int nmol, const WideMol *mols, WideStarts st, bool fp64) {
    if (nmol <= 0 || st.nvecs <= 0) return;
    if (fp64)
        hipLaunchKernelGGL(wide_atomdiag_kernel<true>, dim3(st.nvecs), dim3(WT), 0, s, mols, st.vec,
                           nmol);
    else
        hipLaunchKernelGGL(wide_atomdiag_kernel<false>, dim3(st.nvecs), dim3(WT), 0, s, mols,
                           st.vec, nmol);
}

void launch_wide_gbuild(hipStream_t s, int nmol, const WideMol *mols, WideStarts st) {
    if (nmol <= 0 || st.ngb <= 0) return;
    hipLaunchKernelGGL(wide_gbuild_kernel, dim3(st.ngb), dim3(WT), 0, s, mols, st.gb, nmol);
}

void launch_wide_gather(hipStream_t s, int nmol, const WideMol *mols, WideStarts st) {
    if (nmol <= 0 || st.nchunks <= 0) return;
    hipLaunchKernelGGL(wide_gather_kernel, dim3(st.nchunks), dim3(WT), 0, s, mols, st.chunk, nmol);
}

void launch_wide_step(hipStream_t s, int nmol, const WideMol *mols, WideStarts st) {
    if (nmol <= 0 || st.nchunks <= 0) return;
    const dim3 gc(st.nchunks), gm(nmol);
    const int *ck = st.chunk;
    hipLaunchKernelGGL(wide_gram_kernel, gc, dim3(WT), 0, s, mols, ck, nmol);
    hipLaunchKernelGGL(wide_sub_kernel, gm, dim3(BT), 0, s, mols);
    hipLaunchKernelGGL(wide_ritz_kernel, gc, dim3(WT), 0, s, mols, ck, nmol);
    hipLaunchKernelGGL(wide_ctl_kernel, gm, dim3(WT), 0, s, mols);
    hipLaunchKernelGGL(wide_expand_kernel, gc, dim3(WT), 0, s, mols, ck, nmol);
    hipLaunchKernelGGL(wide_red_kernel, gm, dim3(WT), 0, s, mols);
    hipLaunchKernelGGL(wide_proj_kernel<0>, gc, dim3(WT), 0, s, mols, ck, nmol);
    hipLaunchKernelGGL(wide_red2_kernel, gm, dim3(WT), 0, s, mols);
    hipLaunchKernelGGL(wide_proj_kernel<1>, gc, dim3(WT), 0, s, mols, ck, nmol);
    hipLaunchKernelGGL(wide_chol_kernel<0>, gm, dim3(64), 0, s, mols);
    hipLaunchKernelGGL(wide_tri_kernel<0>, gc, dim3(WT), 0, s, mols, ck, nmol);
    hipLaunchKernelGGL(wide_chol_kernel<1>, gm, dim3(64), 0, s, mols);
    hipLaunchKernelGGL(wide_tri_kernel<1>, gc, dim3(WT), 0, s, mols, ck, nmol);
    hipLaunchKernelGGL(wide_commit_kernel, gm, dim3(64), 0, s, mols);
    STDA_HIP_CHECK(hipGetLastError());
}

}  // namespace stda
