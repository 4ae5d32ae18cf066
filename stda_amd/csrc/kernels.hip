// HIP kernels of the MI355X excited-state core (gfx950, wave64).
//
// Index conventions (0-based): occupied i in [0,no), virtual a in [0,nv),
// pair p = i*nv + a; charge arrays are (ncent, npairs) with the atom index
// fastest; packed occupied/virtual pairs use lin0(x,y) = max*(max+1)/2 + min,
// the 0-based form of the reference's lin() (stda.f:2174-2181).
#include "../../include/stda_hip.h"
#include "common.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <type_traits>

namespace stda {

__device__ __forceinline__ int64_t lin0(int64_t x, int64_t y) {
    int64_t a = x > y ? x : y, b = x > y ? y : x;
    return a * (a + 1) / 2 + b;
}

__device__ __forceinline__ double wave_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// ----------------------------------------------------------------------------
// gamma^J, gamma^K (stda.f:509-532): Mataga-Nishimoto-Ohno-Klopman forms
//   gJ = (R^bJ + (ax eta)^-bJ)^(-1/bJ),  gK = (R^aK + eta^-aK)^(-1/aK)
// eta = (eta_ZA + eta_ZB)/2.  ax = 0 gives gJ = 0 exactly (1/0^b = +Inf).
__global__ void gamma_kernel(int n, const double *__restrict__ xyz4, const double *__restrict__ eta,
                             double ax, double ak_exp, double bj_exp, double *gamj, double *gamk) {
    int j = blockIdx.x * 16 + threadIdx.x;
    int i = blockIdx.y * 16 + threadIdx.y;
    if (i >= n || j >= n) return;
    int zi = (int)xyz4[4 * i + 3], zj = (int)xyz4[4 * j + 3];
    double e = 0.5 * (eta[zi - 1] + eta[zj - 1]);
    double dx = xyz4[4 * i] - xyz4[4 * j];
    double dy = xyz4[4 * i + 1] - xyz4[4 * j + 1];
    double dz = xyz4[4 * i + 2] - xyz4[4 * j + 2];
    double r = sqrt(dx * dx + dy * dy + dz * dz);
    double xj = e * ax;
    double gj = (xj == 0.0) ? 0.0 : 1.0 / pow(pow(r, bj_exp) + 1.0 / pow(xj, bj_exp), 1.0 / bj_exp);
    double gk = 1.0 / pow(pow(r, ak_exp) + 1.0 / pow(e, ak_exp), 1.0 / ak_exp);
    gamj[i + (int64_t)n * j] = gj;
    gamk[i + (int64_t)n * j] = gk;
}

void launch_gamma(hipStream_t s, int ncent, const double *xyz4, const double *eta, double ax,
                  double alphak, double betaj, double *gamj, double *gamk) {
    dim3 b(16, 16), g((ncent + 15) / 16, (ncent + 15) / 16);
    hipLaunchKernelGGL(gamma_kernel, g, b, 0, s, ncent, xyz4, eta, ax, alphak, betaj, gamj, gamk);
}

// ----------------------------------------------------------------------------
// Loewdin transition charges q^A_pq = sum_{mu on A} C'(mu,p) C'(mu,q)
// (lo12pop, stda.f:1727-1744).  ct is (moci, nao) row-major with the AOs
// sorted by atom, so the AOs of atom A are ct[p][atom_ptr[A]:atom_ptr[A+1]).
// One block per "outer" MO p, staged in LDS; the block covers CH_TI inner MOs.
constexpr int CH_TI = 32;
constexpr int CH_LDS = 8192;   // doubles of the outer MO kept in LDS (64 KiB)

__global__ void __launch_bounds__(256) charges_kernel(int mode, int ncent, int nao, int no, int nv,
                                                      const double *__restrict__ ct,
                                                      const int *__restrict__ atom_ptr,
                                                      double *__restrict__ q) {
    __shared__ double sp[CH_LDS];
    int outer = blockIdx.y;
    int p, qbeg, qend;
    int64_t colbase;
    if (mode == CH_IA) {
        p = outer; qbeg = no; qend = no + nv; colbase = (int64_t)outer * nv - no;
    } else if (mode == CH_OCC_PACKED) {
        p = outer; qbeg = 0; qend = outer + 1; colbase = (int64_t)outer * (outer + 1) / 2;
    } else if (mode == CH_VIR_PACKED) {
        p = no + outer; qbeg = no; qend = no + outer + 1;
        colbase = (int64_t)outer * (outer + 1) / 2 - no;
    } else {
        p = outer; qbeg = outer; qend = outer + 1; colbase = 0;
    }
    int q0 = qbeg + blockIdx.x * CH_TI;
    if (q0 >= qend) return;
    int nq = min(CH_TI, qend - q0);
    const double *cp = ct + (int64_t)p * nao;
    bool in_lds = nao <= CH_LDS;
    if (in_lds) {
        for (int m = threadIdx.x; m < nao; m += blockDim.x) sp[m] = cp[m];
        __syncthreads();
    }
    int total = nq * ncent;
    for (int u = threadIdx.x; u < total; u += blockDim.x) {
        int A = u % ncent;
        int qi = q0 + u / ncent;
        const double *cq = ct + (int64_t)qi * nao;
        int m0 = atom_ptr[A], m1 = atom_ptr[A + 1];
        double s = 0.0;
        if (in_lds)
            for (int m = m0; m < m1; ++m) s += sp[m] * cq[m];
        else
            for (int m = m0; m < m1; ++m) s += cp[m] * cq[m];
        q[A + (int64_t)ncent * (colbase + qi)] = s;
    }
}

void launch_charges(hipStream_t s, ChargeMode mode, int ncent, int nao, int no, int nv,
                    const double *ct, const int *atom_ptr, double *q) {
    int nouter, maxinner;
    switch (mode) {
        case CH_IA: nouter = no; maxinner = nv; break;
        case CH_OCC_PACKED: nouter = no; maxinner = no; break;
        case CH_VIR_PACKED: nouter = nv; maxinner = nv; break;
        default: nouter = no + nv; maxinner = 1; break;
    }
    if (nouter == 0) return;
    dim3 g((maxinner + CH_TI - 1) / CH_TI, nouter);
    hipLaunchKernelGGL(charges_kernel, g, dim3(256), 0, s, (int)mode, ncent, nao, no, nv, ct,
                       atom_ptr, q);
}

// ----------------------------------------------------------------------------
// CSF screening (stda.f:593-632): de = (e_a - e_i) - (ii|aa)_J + ak (ia|ia)_K
// (+ the optional K-dependent shift, stda.f:1750-1768).  One wavefront per pair.
// flags: bit0 primary (de <= thr), bit1 secondary (thr < de < fthr).
__global__ void __launch_bounds__(256) screen_kernel(int ncent, int no, int nv,
                                                     const double *__restrict__ eps,
                                                     const double *__restrict__ jdiag,
                                                     const double *__restrict__ pia,
                                                     const double *__restrict__ qia, double ak,
                                                     double thr, double fthr, int kshift,
                                                     double sau, double wau, double steep,
                                                     double *de_out, double *kdiag,
                                                     unsigned char *flags) {
    int lane = threadIdx.x & 63;
    int64_t npair = (int64_t)no * nv;
    int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t p = wave; p < npair; p += nwaves) {
        const double *a = pia + (int64_t)ncent * p;
        const double *b = qia + (int64_t)ncent * p;
        double s = 0.0;
        for (int k = lane; k < ncent; k += 64) s += a[k] * b[k];
        s = wave_sum(s);
        if (lane == 0) {
            int i = (int)(p / nv), l = (int)(p % nv);
            double de = eps[no + l] - eps[i];
            de = de - jdiag[l + (int64_t)nv * i];
            de = de + ak * s;
            if (kshift) de = de + sau / (1.0 + pow(s / wau, steep));
            de_out[p] = de;
            kdiag[p] = s;
            unsigned char f = 0;
            if (de <= thr) f |= 1;
            if (de > thr && de < fthr) f |= 2;
            flags[p] = f;
        }
    }
}

void launch_screen(hipStream_t s, int ncent, int no, int nv, const double *eps,
                   const double *jdiag, const double *pia, const double *qia, double ak,
                   double thr, double fthr, int kshift, double smax, double width,
                   double steep, double *de, double *kdiag, unsigned char *flags) {
    int64_t npair = (int64_t)no * nv;
    int64_t blocks = (npair + 3) / 4;
    if (blocks > 65536) blocks = 65536;
    double sau = smax * 0.036749320, wau = width * 0.036749320;   // stda.f:1764-1765
    hipLaunchKernelGGL(screen_kernel, dim3((unsigned)blocks), dim3(256), 0, s, ncent, no, nv, eps,
                       jdiag, pia, qia, ak, thr, fthr, kshift, sau, wau, steep, de, kdiag, flags);
}

// ----------------------------------------------------------------------------
// Order-preserving stream compaction of the pair indices whose flag has `bit`.
constexpr int CMP_THREADS = 256, CMP_ITEMS = 4, CMP_BLOCK = CMP_THREADS * CMP_ITEMS;
int compact_block_size() { return CMP_BLOCK; }

__global__ void __launch_bounds__(CMP_THREADS) count_kernel(const unsigned char *flags, int64_t n,
                                                            int bit, int *counts) {
    __shared__ int ws[CMP_THREADS / 64];
    int64_t base = (int64_t)blockIdx.x * CMP_BLOCK;
    int c = 0;
    for (int r = 0; r < CMP_ITEMS; ++r) {
        int64_t idx = base + (int64_t)r * CMP_THREADS + threadIdx.x;
        if (idx < n && (flags[idx] >> bit & 1)) ++c;
    }
    for (int off = 32; off > 0; off >>= 1) c += __shfl_down(c, off, 64);
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int w = 0; w < CMP_THREADS / 64; ++w) t += ws[w];
        counts[blockIdx.x] = t;
    }
}

__global__ void __launch_bounds__(CMP_THREADS) compact_kernel(const unsigned char *flags, int64_t n,
                                                              int bit, const int64_t *offsets,
                                                              int32_t *out) {
    // items of thread t: base + t*ITEMS + r (contiguous per thread keeps order)
    __shared__ int scan[CMP_THREADS];
    int64_t base = (int64_t)blockIdx.x * CMP_BLOCK + (int64_t)threadIdx.x * CMP_ITEMS;
    int c = 0;
    for (int r = 0; r < CMP_ITEMS; ++r) {
        int64_t idx = base + r;
        if (idx < n && (flags[idx] >> bit & 1)) ++c;
    }
    scan[threadIdx.x] = c;
    __syncthreads();
    for (int off = 1; off < CMP_THREADS; off <<= 1) {
        int v = threadIdx.x >= off ? scan[threadIdx.x - off] : 0;
        __syncthreads();
        scan[threadIdx.x] += v;
        __syncthreads();
    }
    int64_t pos = offsets[blockIdx.x] + scan[threadIdx.x] - c;
    for (int r = 0; r < CMP_ITEMS; ++r) {
        int64_t idx = base + r;
        if (idx < n && (flags[idx] >> bit & 1)) out[pos++] = (int32_t)idx;
    }
}

void launch_count_flags(hipStream_t s, const unsigned char *flags, int64_t n, int bit,
                        int *block_counts, int nblocks) {
    hipLaunchKernelGGL(count_kernel, dim3(nblocks), dim3(CMP_THREADS), 0, s, flags, n, bit,
                       block_counts);
}

void launch_compact_flags(hipStream_t s, const unsigned char *flags, int64_t n, int bit,
                          const int64_t *block_offsets, int nblocks, int32_t *out) {
    hipLaunchKernelGGL(compact_kernel, dim3(nblocks), dim3(CMP_THREADS), 0, s, flags, n, bit,
                       block_offsets, out);
}

// ----------------------------------------------------------------------------
// Gathers / scatters between the CSF list and the full occupied x virtual grid.
__global__ void gather_f64_kernel(const double *src, const int32_t *idx, int64_t n, double *dst) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[idx[i]];
}

void launch_gather_f64(hipStream_t s, const double *src, const int32_t *idx, int64_t n,
                       double *dst) {
    if (n == 0) return;
    hipLaunchKernelGGL(gather_f64_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, src,
                       idx, n, dst);
}

__global__ void scatter_pairs_kernel(const double *v, int64_t ldv, const int32_t *pair, int64_t n,
                                     double *full, int64_t ldfull) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int vec = blockIdx.y;
    if (i < n) full[pair[i] + ldfull * vec] = v[i + ldv * vec];
}

void launch_scatter_pairs(hipStream_t s, const double *v, int64_t ldv, const int32_t *pair,
                          int64_t n, int nvec, double *full, int64_t ldfull) {
    if (n == 0 || nvec == 0) return;
    hipLaunchKernelGGL(scatter_pairs_kernel, dim3((unsigned)((n + 255) / 256), nvec), dim3(256), 0,
                       s, v, ldv, pair, n, full, ldfull);
}

// out = full[pair] + diag o v
__global__ void gather_pairs_kernel(const double *full, int64_t ldfull, const int32_t *pair,
                                    int64_t n, const double *diag, const double *v, int64_t ldv,
                                    double *out, int64_t ldout) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int vec = blockIdx.y;
    if (i >= n) return;
    double r = full[pair[i] + ldfull * vec];
    if (diag) r += diag[i] * v[i + ldv * vec];
    out[i + ldout * vec] = r;
}

void launch_gather_pairs(hipStream_t s, const double *full, int64_t ldfull, const int32_t *pair,
                         int64_t n, int nvec, const double *diag, const double *v, int64_t ldv,
                         double *out, int64_t ldout) {
    if (n == 0 || nvec == 0) return;
    hipLaunchKernelGGL(gather_pairs_kernel, dim3((unsigned)((n + 255) / 256), nvec), dim3(256), 0,
                       s, full, ldfull, pair, n, diag, v, ldv, out, ldout);
}

// mom(n, 9): moments of each CSF, element M(io, iv) of the packed arrays
// (stda.f:1229-1231: ij = lin(io, iv)).
__global__ void csf_moments_kernel(const double *moments, int64_t npk, const int32_t *pair,
                                   int64_t n, int no, int nv, double *mom) {
    int64_t I = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= n) return;
    int64_t p = pair[I];
    int64_t io = p / nv;            // 0-based occupied MO
    int64_t iv = no + p % nv;       // 0-based absolute virtual MO
    int64_t ij = lin0(io, iv);
    for (int c = 0; c < 9; ++c) mom[I + n * c] = moments[c * npk + ij];
}

void launch_csf_moments(hipStream_t s, const double *moments, int64_t npk, const int32_t *pair,
                        int64_t n, int no, int nv, double *mom) {
    if (n == 0) return;
    hipLaunchKernelGGL(csf_moments_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                       moments, npk, pair, n, no, nv, mom);
}

// ----------------------------------------------------------------------------
// AO-basis (Loewdin-basis) form of the coupling operators.  For a trial
// vector X (no x nv) the transition density is D = Co X Cv^T (nao x nao),
// stored transposed as Dt = Cv X^T Co^T.  Then
//   sum_jb (ia|jb)_K X_jb = [Co^T diag(u_A(mu)) Cv]_ia,  u = gamK t,
//                          t_A = sum_{mu on A} D(mu,mu)
//   sum_jb (ij|ab)_J X_jb = [Co^T (GamJ o D) Cv]_ia
//   sum_jb (ib|ja)_K X_jb = [Co^T (GamK o D^T) Cv]_ia
// with Gam_{mu nu} = gamma_{A(mu) A(nu)}, because the Loewdin charges
// factor per AO: q^A_pq = sum_{mu on A} C'(mu,p) C'(mu,q).
__global__ void atom_diag_kernel(int nao, int ncent, const double *Dt, int64_t strideD,
                                 const int *atom_ptr, const double *gamk, double *u) {
    extern __shared__ double st[];
    int v = blockIdx.x;
    const double *D = Dt + strideD * v;
    for (int A = threadIdx.x; A < ncent; A += blockDim.x) {
        double s = 0.0;
        for (int m = atom_ptr[A]; m < atom_ptr[A + 1]; ++m) s += D[m + (int64_t)nao * m];
        st[A] = s;
    }
    __syncthreads();
    for (int A = threadIdx.x; A < ncent; A += blockDim.x) {
        double s = 0.0;
        const double *g = gamk + (int64_t)ncent * A;   // symmetric
        for (int B = 0; B < ncent; ++B) s += g[B] * st[B];
        u[A + (int64_t)ncent * v] = s;
    }
}

void launch_atom_diag(hipStream_t s, int nvec, int nao, int ncent, const double *Dt,
                      int64_t strideD, const int *atom_ptr, const double *gamk, double *u) {
    if (nvec == 0) return;
    hipLaunchKernelGGL(atom_diag_kernel, dim3(nvec), dim3(256), ncent * sizeof(double), s, nao,
                       ncent, Dt, strideD, atom_ptr, gamk, u);
}

constexpr int GB_T = 32;
__global__ void __launch_bounds__(256) gbuild_kernel(int nao, int ncent, const double *Dt,
                                                     int64_t strideD, const int *atom_of,
                                                     const double *gamj, const double *gamk,
                                                     const double *u, double aJ, double aX,
                                                     double aK, double *Gt, int64_t strideG) {
    __shared__ double tr[GB_T][GB_T + 1];
    int v = blockIdx.z;
    const double *D = Dt + strideD * v;
    double *G = Gt + strideG * v;
    int nu0 = blockIdx.x * GB_T, mu0 = blockIdx.y * GB_T;   // Gt(nu, mu), column-major
    int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;       // 32 x 8
    if (aX != 0.0) {
        // stage Dt(mu0+.., nu0+..) so that tr[nu][mu] = Dt(mu, nu)
        for (int r = ty; r < GB_T; r += 8) {
            int mu = mu0 + tx, nu = nu0 + r;
            tr[r][tx] = (mu < nao && nu < nao) ? D[mu + (int64_t)nao * nu] : 0.0;
        }
        __syncthreads();
    }
    for (int r = ty; r < GB_T; r += 8) {
        int nu = nu0 + tx, mu = mu0 + r;
        if (nu >= nao || mu >= nao) continue;
        int An = atom_of[nu], Am = atom_of[mu];
        int64_t g = An + (int64_t)ncent * Am;
        double val = 0.0;
        if (aJ != 0.0) val += aJ * gamj[g] * D[nu + (int64_t)nao * mu];
        if (aX != 0.0) val += aX * gamk[g] * tr[tx][r];
        if (nu == mu && aK != 0.0) val += aK * u[Am + (int64_t)ncent * v];
        G[nu + (int64_t)nao * mu] = val;
    }
}

void launch_gbuild(hipStream_t s, int nvec, int nao, int ncent, const double *Dt,
                   int64_t strideD, const int *atom_of, const double *gamj, const double *gamk,
                   const double *u, double aJ, double aX, double aK, double *Gt, int64_t strideG) {
    if (nvec == 0) return;
    dim3 g((nao + GB_T - 1) / GB_T, (nao + GB_T - 1) / GB_T, nvec);
    hipLaunchKernelGGL(gbuild_kernel, g, dim3(256), 0, s, nao, ncent, Dt, strideD, atom_of, gamj,
                       gamk, u, aJ, aX, aK, Gt, strideG);
}

// ----------------------------------------------------------------------------
// Davidson vector operations on (n x k) column-major blocks.
__global__ void residual_kernel(int64_t n, const double *AX, const double *X, const double *theta,
                                double *R) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int c = blockIdx.y;
    if (i < n) R[i + n * c] = AX[i + n * c] - theta[c] * X[i + n * c];
}

void launch_residual(hipStream_t s, int64_t n, int k, const double *AX, const double *X,
                     const double *theta, double *R) {
    if (k == 0) return;
    hipLaunchKernelGGL(residual_kernel, dim3((unsigned)((n + 255) / 256), k), dim3(256), 0, s, n,
                       AX, X, theta, R);
}

__global__ void col_norms_kernel(int64_t n, const double *R, double *norms) {
    __shared__ double ws[4];
    int c = blockIdx.x;
    double s = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
        double v = R[i + n * c];
        s += v * v;
    }
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) norms[c] = sqrt(ws[0] + ws[1] + ws[2] + ws[3]);
}

void launch_col_norms(hipStream_t s, int64_t n, int k, const double *R, double *norms) {
    if (k == 0) return;
    hipLaunchKernelGGL(col_norms_kernel, dim3(k), dim3(256), 0, s, n, R, norms);
}

// T = R / (theta - diag): diagonal (Davidson) preconditioner, guarded.
__global__ void precondition_kernel(int64_t n, const double *R, const double *diag,
                                    const double *theta, double *T) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int c = blockIdx.y;
    if (i >= n) return;
    double d = theta[c] - diag[i];
    if (fabs(d) < 1.0e-4) d = d < 0.0 ? -1.0e-4 : 1.0e-4;
    T[i + n * c] = R[i + n * c] / d;
}

// T(:, c) = R(:, col[c]) / (theta[c] - diag), |denominator| >= 1e-4; col[c]
// holds the column index as a double (packed with theta in one upload)
__global__ void precondition_gather_kernel(int64_t n, const double *R, const double *col,
                                           const double *diag, const double *theta, double *T) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int c = blockIdx.y;
    if (i >= n) return;
    const int64_t src = (int64_t)col[c];
    double d = theta[c] - diag[i];
    if (fabs(d) < 1.0e-4) d = d < 0.0 ? -1.0e-4 : 1.0e-4;
    T[i + n * c] = R[i + n * src] / d;
}

// New directions from the unconverged residuals (norm not below tol), packed
// in column order: T(:, q) = R(:, c) / (theta_c - diag) for the q-th
// unconverged column c, q < room.
__global__ void precondition_compact_kernel(int64_t n, const double *R, const double *diag,
                                            const double *theta, const double *norms, double tol,
                                            int room, double *T) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (norms[c] < tol) return;
    int q = 0;
    for (int c2 = 0; c2 < c; ++c2) q += !(norms[c2] < tol);
    if (q >= room || i >= n) return;
    double d = theta[c] - diag[i];
    if (fabs(d) < 1.0e-4) d = d < 0.0 ? -1.0e-4 : 1.0e-4;
    T[i + n * q] = R[i + n * c] / d;
}

void launch_precondition_compact(hipStream_t s, int64_t n, int k, const double *R,
                                 const double *diag, const double *theta, const double *norms,
                                 double tol, int room, double *T) {
    if (k == 0 || room == 0) return;
    hipLaunchKernelGGL(precondition_compact_kernel, dim3((unsigned)((n + 255) / 256), k), dim3(256),
                       0, s, n, R, diag, theta, norms, tol, room, T);
}

void launch_precondition_gather(hipStream_t s, int64_t n, int k, const double *R, const double *col,
                                const double *diag, const double *theta, double *T) {
    if (k == 0) return;
    hipLaunchKernelGGL(precondition_gather_kernel, dim3((unsigned)((n + 255) / 256), k), dim3(256),
                       0, s, n, R, col, diag, theta, T);
}

void launch_precondition(hipStream_t s, int64_t n, int k, const double *R, const double *diag,
                         const double *theta, double *T) {
    if (k == 0) return;
    hipLaunchKernelGGL(precondition_kernel, dim3((unsigned)((n + 255) / 256), k), dim3(256), 0, s,
                       n, R, diag, theta, T);
}

__global__ void scale_cols_kernel(int64_t n, double *V, const double *scale) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int c = blockIdx.y;
    if (i < n) V[i + n * c] *= scale[c];
}

__global__ void inv_scale_cols_kernel(int64_t n, double *V, const double *norm) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int c = blockIdx.y;
    const double d = norm[c];
    if (i < n) V[i + n * c] *= d > 0.0 ? 1.0 / d : 0.0;
}

void launch_inv_scale_cols(hipStream_t s, int64_t n, int k, double *V, const double *norm) {
    if (k == 0) return;
    hipLaunchKernelGGL(inv_scale_cols_kernel, dim3((unsigned)((n + 255) / 256), k), dim3(256), 0, s,
                       n, V, norm);
}

void launch_scale_cols(hipStream_t s, int64_t n, int k, double *V, const double *scale) {
    if (k == 0) return;
    hipLaunchKernelGGL(scale_cols_kernel, dim3((unsigned)((n + 255) / 256), k), dim3(256), 0, s, n,
                       V, scale);
}

__global__ void set_unit_kernel(int64_t n, double *V, const int32_t *rows) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int c = blockIdx.y;
    if (i < n) V[i + n * c] = (i == rows[c]) ? 1.0 : 0.0;
}

void launch_set_unit(hipStream_t s, int64_t n, int k, double *V, const int32_t *rows) {
    if (k == 0) return;
    hipLaunchKernelGGL(set_unit_kernel, dim3((unsigned)((n + 255) / 256), k), dim3(256), 0, s, n, V,
                       rows);
}

// Y(:, c) = diag o X(:, c)
__global__ void axpby_diag_kernel(int64_t n, const double *diag, const double *X, int64_t ldx,
                                  double *Y, int64_t ldy) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int c = blockIdx.y;
    if (i < n) Y[i + ldy * c] = diag[i] * X[i + ldx * c];
}

void launch_axpby_diag(hipStream_t s, int64_t n, int k, const double *diag, const double *X,
                       int64_t ldx, double *Y, int64_t ldy) {
    if (k == 0) return;
    hipLaunchKernelGGL(axpby_diag_kernel, dim3((unsigned)((n + 255) / 256), k), dim3(256), 0, s, n,
                       diag, X, ldx, Y, ldy);
}

// ----------------------------------------------------------------------------
// 4-index coupling tiles.
//
// For a row tile R = Kset x Cset (secondary-type pairs (k,c)) and a column
// tile C = Jset x Bset (primary-type pairs (j,b)) the block forms, for every
// 4-tuple (k,c,j,b), both contractions over atoms
//   K(kc,jb) = sum_A pK(A,kc) q(A,jb)           ((kc|jb), stda.f:1811)
//   J(kc,jb) = sum_A pJ(A,lin(k,j)) qab(A,lin(c,b))   ((kj|cb), stda.f:1814-1815)
// i.e. two small GEMMs over the atom index whose output layouts differ by a
// permutation of the tuple; here each thread owns the same tuples in both,
// so the permutation is implicit.  Operands are staged through LDS in
// atom chunks of NA_CHUNK, stored [atom][vector] so that a wavefront reads
// consecutive vectors.
constexpr int NT = 256;
constexpr int NRV = TK * TC;     // row-tile vectors (pK)
constexpr int NCV = TJ * TB;     // col-tile vectors (q)
constexpr int NJV = TK * TJ;     // pJ vectors
constexpr int NQV = TC * TB;     // qab vectors
static_assert(TK * TC * TJ * TB == 4 * NT, "each thread owns TK tuples");
static_assert(TK == 4, "thread mapping assumes TK == 4");

struct TileSmem {
    double pk[NA_CHUNK][NRV + 1];
    double qc[NA_CHUNK][NCV + 1];
    double pj[NA_CHUNK][NJV + 1];
    double qv[NA_CHUNK][NQV + 1];
    int occr[TK], virr[TC], occc[TJ], virc[TB];
};

__device__ __forceinline__ void stage_tile(TileSmem &sm, int a0, int na, int ncent, int nv,
                                           const double *__restrict__ pK,
                                           const double *__restrict__ q,
                                           const double *__restrict__ pJ,
                                           const double *__restrict__ qab) {
    // each (vector, atom) element loaded by one thread; atoms fastest for coalescing
    for (int e = threadIdx.x; e < NRV * NA_CHUNK; e += NT) {
        int v = e / NA_CHUNK, a = e % NA_CHUNK;
        int k = sm.occr[v / TC], c = sm.virr[v % TC];
        double x = 0.0;
        if (a < na && k >= 0 && c >= 0) x = pK[a0 + a + (int64_t)ncent * ((int64_t)k * nv + c)];
        sm.pk[a][v] = x;
    }
    for (int e = threadIdx.x; e < NCV * NA_CHUNK; e += NT) {
        int v = e / NA_CHUNK, a = e % NA_CHUNK;
        int j = sm.occc[v / TB], b = sm.virc[v % TB];
        double x = 0.0;
        if (a < na && j >= 0 && b >= 0) x = q[a0 + a + (int64_t)ncent * ((int64_t)j * nv + b)];
        sm.qc[a][v] = x;
    }
    for (int e = threadIdx.x; e < NJV * NA_CHUNK; e += NT) {
        int v = e / NA_CHUNK, a = e % NA_CHUNK;
        int k = sm.occr[v / TJ], j = sm.occc[v % TJ];
        double x = 0.0;
        if (a < na && k >= 0 && j >= 0) x = pJ[a0 + a + (int64_t)ncent * lin0(k, j)];
        sm.pj[a][v] = x;
    }
    for (int e = threadIdx.x; e < NQV * NA_CHUNK; e += NT) {
        int v = e / NA_CHUNK, a = e % NA_CHUNK;
        int c = sm.virr[v / TB], b = sm.virc[v % TB];
        double x = 0.0;
        if (a < na && c >= 0 && b >= 0) x = qab[a0 + a + (int64_t)ncent * lin0(c, b)];
        sm.qv[a][v] = x;
    }
}

// thread t owns tuples (k=0..3, c, j, b) with b = t%TB, j = (t/TB)%TJ, c = t/(TB*TJ)
__device__ __forceinline__ void tile_contract(TileSmem &sm, int ncent, int nv,
                                              const double *pK, const double *q,
                                              const double *pJ, const double *qab, double accK[TK],
                                              double accJ[TK]) {
    int t = threadIdx.x;
    int b = t % TB, j = (t / TB) % TJ, c = t / (TB * TJ);
    for (int k = 0; k < TK; ++k) { accK[k] = 0.0; accJ[k] = 0.0; }
    for (int a0 = 0; a0 < ncent; a0 += NA_CHUNK) {
        int na = min(NA_CHUNK, ncent - a0);
        __syncthreads();
        stage_tile(sm, a0, na, ncent, nv, pK, q, pJ, qab);
        __syncthreads();
        for (int a = 0; a < na; ++a) {
            double qjb = sm.qc[a][j * TB + b];
            double qcb = sm.qv[a][c * TB + b];
#pragma unroll
            for (int k = 0; k < TK; ++k) {
                accK[k] += sm.pk[a][k * TC + c] * qjb;
                accJ[k] += sm.pj[a][k * TJ + j] * qcb;
            }
        }
    }
}

__device__ __forceinline__ void load_tile_ids(TileSmem &sm, TileSet rows, int rt, TileSet cols,
                                              int ctile) {
    __syncthreads();
    if (threadIdx.x < TK) sm.occr[threadIdx.x] = rows.occ[rt * TK + threadIdx.x];
    if (threadIdx.x < TC) sm.virr[threadIdx.x] = rows.vir[rt * TC + threadIdx.x];
    if (threadIdx.x < TJ) sm.occc[threadIdx.x] = cols.occ[ctile * TJ + threadIdx.x];
    if (threadIdx.x < TB) sm.virc[threadIdx.x] = cols.vir[ctile * TB + threadIdx.x];
    __syncthreads();
}

// Explicit A: one block per (row tile, column tile); stores the coupling of
// every pair of CSFs (the diagonal is overwritten afterwards with ed).
__global__ void __launch_bounds__(NT) build_a_kernel(TileSet tiles, int ncent, int no, int nv,
                                                     const double *pK, const double *q,
                                                     const double *pJ, const double *qab,
                                                     const int32_t *csf_of_pair, double ak,
                                                     double *A, int64_t lda) {
    __shared__ TileSmem sm;
    int rt = blockIdx.x, ctile = blockIdx.y;
    int t = threadIdx.x;
    int b = t % TB, j = (t / TB) % TJ, c = t / (TB * TJ);
    TileSet cols{tiles.occ, tiles.vir, tiles.ntiles};
    // the same tile list serves rows and columns; TK == TJ and TC == TB here
    load_tile_ids(sm, tiles, rt, cols, ctile);
    double accK[TK], accJ[TK];
    tile_contract(sm, ncent, nv, pK, q, pJ, qab, accK, accJ);
    int jj = sm.occc[j], bb = sm.virc[b], cc = sm.virr[c];
    if (jj < 0 || bb < 0 || cc < 0) return;
    int32_t col = csf_of_pair[(int64_t)jj * nv + bb];
    if (col < 0) return;
#pragma unroll
    for (int k = 0; k < TK; ++k) {
        int kk = sm.occr[k];
        if (kk < 0) continue;
        int32_t row = csf_of_pair[(int64_t)kk * nv + cc];
        if (row < 0) continue;
        A[row + lda * col] = ak * accK[k] - accJ[k];
    }
}

void launch_build_a(hipStream_t s, TileSet tiles, int ncent, int no, int nv, const double *pK,
                    const double *q, const double *pJ, const double *qab,
                    const int32_t *csf_of_pair, double ak, double *A, int64_t lda) {
    if (tiles.ntiles == 0) return;
    hipLaunchKernelGGL(build_a_kernel, dim3(tiles.ntiles, tiles.ntiles), dim3(NT), 0, s, tiles,
                       ncent, no, nv, pK, q, pJ, qab, csf_of_pair, ak, A, lda);
}

// E(I,J) = (ib|ja)_K = pK(:, i*nv+b) . q(:, j*nv+a) for CSFs I=(i,a), J=(j,b)
// (stda.f:1952-1956).  One wavefront per (I, J) pair; small systems only.
__global__ void __launch_bounds__(256) exchange_kernel(int64_t nci, int ncent, int nv,
                                                       const int32_t *pair, const double *pK,
                                                       const double *q, double *E, int64_t lde) {
    int lane = threadIdx.x & 63;
    int64_t w = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t e = w; e < nci * nci; e += nw) {
        int64_t I = e % nci, J = e / nci;
        int64_t pI = pair[I], pJ_ = pair[J];
        int64_t i = pI / nv, a = pI % nv, j = pJ_ / nv, b = pJ_ % nv;
        const double *x = pK + (int64_t)ncent * (i * nv + b);
        const double *y = q + (int64_t)ncent * (j * nv + a);
        double s = 0.0;
        for (int k = lane; k < ncent; k += 64) s += x[k] * y[k];
        s = wave_sum(s);
        if (lane == 0) E[I + lde * J] = s;
    }
}

void launch_exchange(hipStream_t s, int64_t nci, int ncent, int nv, const int32_t *pair,
                     const double *pK, const double *q, double *E, int64_t lde) {
    if (nci == 0) return;
    int64_t blocks = (nci * nci + 3) / 4;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(exchange_kernel, dim3((unsigned)blocks), dim3(256), 0, s, nci, ncent, nv,
                       pair, pK, q, E, lde);
}

}  // namespace stda

namespace stda {

// ----------------------------------------------------------------------------
// PT2 couplings on the matrix cores (FP64 MFMA, v_mfma_f64_16x16x4f64).
//
// Work-group tile: 8 occupied (k) x 8 virtual (c) secondary-side pairs times
// 8 occupied (j) x 8 virtual (b) primary-side pairs = 4096 tuples.  Wave w
// owns k in [4(w/2), +4) and j in [4(w%2), +4) and all c, b: four 16x16
// output blocks per coupling, indexed by the virtual halves (ch, bh).
//   J block: rows (k', j'), cols (c', b')   A = pJ(atom, lin(k,j)), B = qab(atom, lin(c,b))
//   K block: rows (k', c'), cols (j', b')   A = pK(atom, kc),       B = q(atom, jb)
// MFMA f64 16x16x4 lane layout: lane l supplies A[l%16][l/16] and
// B[l/16][l%16]; register i of lane l holds D[4i + l/16][l%16].  The J
// values are moved to the K layout through LDS before the epilogue
//   amat = ak K - J,  pt = amat^2 / (de(kc) - de(jb) + 1e-10).
typedef double v4d __attribute__((ext_vector_type(4)));
constexpr int M_T = 8;          // per-dimension tile size
constexpr int M_CH = 16;        // atoms per staged chunk (4 MFMA k-steps)
constexpr int M_LD = 64 + 1;    // padded LDS row (doubles)

}  // namespace stda

namespace stda {

// Preconditioner of the response equations: with d the diagonal (ed) solve
//   d x1 - w x2 = R1,  -w x1 + d x2 = R2
// per element; w is the frequency of each right-hand side column.
__global__ void precond_pair_kernel(int64_t n, int ncol, double *R1, double *R2, const double *d,
                                    const double *w) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int c = blockIdx.y;
    if (i >= n) return;
    double dd = d[i], ww = w[c];
    double det = dd * dd - ww * ww;
    if (fabs(det) < 1.0e-6) det = det < 0.0 ? -1.0e-6 : 1.0e-6;
    double r1 = R1[i + n * c], r2 = R2[i + n * c];
    R1[i + n * c] = (dd * r1 + ww * r2) / det;
    R2[i + n * c] = (ww * r1 + dd * r2) / det;
}

void launch_precond_pair(hipStream_t s, int64_t n, int ncol, double *R1, double *R2,
                         const double *d, const double *w) {
    if (ncol == 0) return;
    hipLaunchKernelGGL(precond_pair_kernel, dim3((unsigned)((n + 255) / 256), ncol), dim3(256), 0,
                       s, n, ncol, R1, R2, d, w);
}

// C(i) = sum_c W(i, c) over nchunk slices of length len, in slice order.
__global__ void sum_slices_kernel(int64_t len, int nchunk, const double *W, double *C) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= len) return;
    double s = 0.0;
    for (int c = 0; c < nchunk; ++c) s += W[i + len * c];
    C[i] = s;
}

void launch_sum_slices(hipStream_t s, int64_t len, int nchunk, const double *W, double *C) {
    if (len == 0) return;
    hipLaunchKernelGGL(sum_slices_kernel, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, s, len,
                       nchunk, W, C);
}

}  // namespace stda

namespace stda {

// ----------------------------------------------------------------------------
// Software-pipelined PT2 pass 1 (stda.f:1800-1838).  A work-group owns one
// secondary row tile and streams a range of primary column tiles.  The
// (tile, atom chunk) steps are flattened; while the matrix cores work on step
// s from one LDS buffer, the operands of step s+1 are fetched into registers
// and then written to the other buffer, so a single barrier per step orders
// everything.
template <typename T, int CH>
struct PipeBuf {
    T pk[CH][M_LD];
    T qc[CH][M_LD];
    T pj[CH][M_LD];
    T qv[CH][M_LD];
};
template <typename T, int CH>
struct PipeSmem {
    PipeBuf<T, CH> buf[2];
    T xch[4][256];
    double red[4][64];
};
typedef float v4f __attribute__((ext_vector_type(4)));
typedef double v2d __attribute__((ext_vector_type(2)));
// MFMA 16x16x4 step and accumulator type per operand type.  Output layouts
// (lane l, g = l / 16, c = l % 16): FP64 register i holds D[4i + g][c], FP32
// register i holds D[4g + i][c].
template <typename T> struct PipeMfma;
template <> struct PipeMfma<double> {
    typedef v4d acc;
    static __device__ __forceinline__ acc step(double a, double b, acc c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
};
template <> struct PipeMfma<float> {
    typedef v4f acc;
    static __device__ __forceinline__ acc step(float a, float b, acc c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
};

// Column index (pair or packed pair) of this thread's VPT staged vectors of
// each operand (vectors vb + VS r), or -1.  o = 0 pK(k,c), 1 q(j,b),
// 2 pJ(k,j), 3 qab(c,b); the row tile gives k, c and the column tile j, b.
template <int VS, int VPT>
__device__ __forceinline__ void stage_cols(TileSet rows, int rt, TileSet cols, int ct, int nv,
                                           int vb, int32_t idx[4][VPT]) {
#pragma unroll
    for (int r = 0; r < VPT; ++r) {
        int v = vb + VS * r, x = v / M_T, y = v % M_T;
        int kx = rows.occ[rt * M_T + x], cy = rows.vir[rt * M_T + y];
        int jx = cols.occ[ct * M_T + x], by = cols.vir[ct * M_T + y];
        int jy = cols.occ[ct * M_T + y], cx = rows.vir[rt * M_T + x];
        idx[0][r] = (kx >= 0 && cy >= 0) ? kx * nv + cy : -1;
        idx[1][r] = (jx >= 0 && by >= 0) ? jx * nv + by : -1;
        idx[2][r] = (kx >= 0 && jy >= 0) ? (int32_t)lin0(kx, jy) : -1;
        idx[3][r] = (cx >= 0 && by >= 0) ? (int32_t)lin0(cx, by) : -1;
    }
}

// de of a row pair (k,c) or column pair (j,b) of a tile; NaN when the pair is
// outside the tile or not selected by the mask.
__device__ __forceinline__ double pair_de(int o, int v, int nv, const double *de,
                                          const unsigned char *mask, unsigned char bit) {
    if (o < 0 || v < 0) return __builtin_nan("");
    int64_t p = (int64_t)o * nv + v;
    return (mask[p] & bit) ? de[p] : __builtin_nan("");
}

// block b of n is dispatched to XCD b % 8; map it to the tile range of that XCD
__device__ __forceinline__ int xcd_remap(int b, int n) {
    const int nx = 8;
    int q = n / nx, r = n % nx;
    int x = b % nx, i = b / nx;
    return x * q + min(x, r) + i;
}

// Work-group item = split sp x row tile own; it streams column tiles
// [ct0, ct1) of that split.  de of the tile slots comes from drow/dcol (NaN
// for slots outside the selection), read once per row tile and prefetched
// with each column tile.  T = double: FP64 operands and MFMA; T = float:
// FP32 operands and MFMA (the reference's own precision for these couplings,
// stda.f:1781-1788), pt and its sums in FP64 either way.
//
// A lane's K values are the rows (k, c) = (4kh + kk, 4chv + cc) and columns
// (j, b) = (4jh + hi, 4bh + lo) of block blk = 2chv + bh, with (kk, cc) =
// (i, g) for the FP64 layout and (g, i) for the FP32 one (register i); its J
// values are (kk, jj = the other of i and g, cc = hi, bb = lo), moved to the
// K lanes through LDS.
template <typename T>
__global__ void __launch_bounds__(256, 2) pt2_pipe_kernel(
    TileSet rows, TileSet cols, int ncent, int nv, const T *__restrict__ pK,
    const T *__restrict__ q, const T *__restrict__ pJ, const T *__restrict__ qab, int64_t lda,
    const double *__restrict__ drow, const double *__restrict__ dcol, double ak, int nsplit,
    double *__restrict__ rowpart, double *__restrict__ ptm, int64_t ldp) {
    typedef typename PipeMfma<T>::acc Acc;
    constexpr bool F64 = sizeof(T) == 8;
    // atoms per step: FP32 steps take 32 (the same 2048 matrix-core cycles per
    // wave as the FP64 steps of 16, and the same LDS bytes); each load brings
    // 16 bytes, i.e. W consecutive atoms of a vector (lda a multiple of W, the
    // padding zero).  TPV threads share a vector, VS vectors per pass, VPT
    // passes cover the 64 vectors of an operand.
    constexpr int W = F64 ? 2 : 4;
    constexpr int CH = F64 ? M_CH : 2 * M_CH, TPV = CH / W, VS = 256 / TPV, VPT = 64 / VS;
    typedef typename std::conditional<F64, v2d, v4f>::type LoadT;
    __shared__ PipeSmem<T, CH> sm;
    // XCD-aware order: work-groups are dealt round-robin to the 8 XCDs, so the
    // items are regrouped to give each XCD a contiguous range (split-major):
    // the work-groups of one XCD then stream the same column tiles, whose
    // operands hit that XCD's L2.
    const int nown = rows.ntiles;
    const int item = xcd_remap((int)blockIdx.x, (int)gridDim.x);
    const int sp = item / nown, own = item % nown;
    const int ct0 = (int)((int64_t)cols.ntiles * sp / nsplit);
    const int ct1 = (int)((int64_t)cols.ntiles * (sp + 1) / nsplit);
    const int nch = (ncent + CH - 1) / CH;
    const int64_t nsteps = (int64_t)(ct1 - ct0) * nch;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int kh = w >> 1, jh = w & 1;
    const int r16 = lane & 15, g = lane >> 4, hi = r16 >> 2, lo = r16 & 3;
    const int sa = threadIdx.x % TPV, svb = threadIdx.x / TPV;
    const int iPJ = M_T * (4 * kh + hi) + (4 * jh + lo);
    const int iPK0 = M_T * (4 * kh + hi) + lo;
    const int iQ0 = M_T * (4 * jh + hi) + lo;
    const int iQV0 = M_T * hi + lo;   // + 32 ch + 4 bh
    const T *src[4] = {pK, q, pJ, qab};
    // (kk, cc) of register i: the row slot (4kh + kk) * 8 + 4chv + cc
    auto kk_of = [&](int i) { return F64 ? i : g; };
    auto cc_of = [&](int i) { return F64 ? g : i; };

    // this lane's rows and columns (4jh + hi, 4bh + lo)
    double dr[2][4], dc[2] = {0.0, 0.0}, dcn[2] = {0.0, 0.0};
#pragma unroll
    for (int chv = 0; chv < 2; ++chv)
#pragma unroll
        for (int i = 0; i < 4; ++i)
            dr[chv][i] = drow[(int64_t)own * PT2_SLOTS + (4 * kh + kk_of(i)) * M_T + 4 * chv + cc_of(i)];

    int32_t idx[4][VPT];
    LoadT pre[4][VPT];
    auto fetch = [&](int ch) {
        const int a = ch * CH + W * sa;
        const bool in = a < ncent;
#pragma unroll
        for (int o = 0; o < 4; ++o)
#pragma unroll
            for (int r = 0; r < VPT; ++r)
                pre[o][r] = (in && idx[o][r] >= 0)
                                ? *reinterpret_cast<const LoadT *>(src[o] + lda * idx[o][r] + a)
                                : LoadT(0);
    };
    auto store = [&](int b) {
#pragma unroll
        for (int r = 0; r < VPT; ++r) {
            const int v = svb + VS * r;
#pragma unroll
            for (int u = 0; u < W; ++u) {
                T x[4];
#pragma unroll
                for (int o = 0; o < 4; ++o) x[o] = pre[o][r][u];
                sm.buf[b].pk[W * sa + u][v] = x[0];
                sm.buf[b].qc[W * sa + u][v] = x[1];
                sm.buf[b].pj[W * sa + u][v] = x[2];
                sm.buf[b].qv[W * sa + u][v] = x[3];
            }
        }
    };
    auto set_tile = [&](int other) {
        stage_cols<VS, VPT>(rows, own, cols, other, nv, svb, idx);
#pragma unroll
        for (int bh = 0; bh < 2; ++bh)
            dcn[bh] = dcol[(int64_t)other * PT2_SLOTS + (4 * jh + hi) * M_T + 4 * bh + lo];
    };

    if (nsteps > 0) {
        set_tile(ct0);
        fetch(0);
        store(0);
        dc[0] = dcn[0];
        dc[1] = dcn[1];
    }
    __syncthreads();

    Acc accJ[4], accK[4];
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        accJ[b] = Acc{0, 0, 0, 0};
        accK[b] = Acc{0, 0, 0, 0};
    }
    // partial row sums of register i's rows of half chv, index 4 chv + i
    double part[8] = {0, 0, 0, 0, 0, 0, 0, 0};

    for (int64_t step = 0; step < nsteps; ++step) {
        const int ch = (int)(step % nch);
        const int other = ct0 + (int)(step / nch);
        const int cur = (int)(step & 1);
        const bool more = step + 1 < nsteps;
        if (more) {
            int ch2 = (int)((step + 1) % nch);
            if (ch2 == 0) set_tile(other + 1);
            fetch(ch2);
        }
        const PipeBuf<T, CH> &B = sm.buf[cur];
#pragma unroll
        for (int s4 = 0; s4 < CH / 4; ++s4) {
            int arow = 4 * s4 + g;
            T aJ = B.pj[arow][iPJ];
            T aK0 = B.pk[arow][iPK0], aK1 = B.pk[arow][iPK0 + 4];
            T bQ0 = B.qc[arow][iQ0], bQ1 = B.qc[arow][iQ0 + 4];
            accJ[0] = PipeMfma<T>::step(aJ, B.qv[arow][iQV0], accJ[0]);
            accJ[1] = PipeMfma<T>::step(aJ, B.qv[arow][iQV0 + 4], accJ[1]);
            accJ[2] = PipeMfma<T>::step(aJ, B.qv[arow][iQV0 + 32], accJ[2]);
            accJ[3] = PipeMfma<T>::step(aJ, B.qv[arow][iQV0 + 36], accJ[3]);
            accK[0] = PipeMfma<T>::step(aK0, bQ0, accK[0]);
            accK[1] = PipeMfma<T>::step(aK0, bQ1, accK[1]);
            accK[2] = PipeMfma<T>::step(aK1, bQ0, accK[2]);
            accK[3] = PipeMfma<T>::step(aK1, bQ1, accK[3]);
        }
        if (ch == nch - 1) {
#pragma unroll
            for (int blk = 0; blk < 4; ++blk) {
                const int chv = blk >> 1, bh = blk & 1;
                // J of register i: (kk, jj, cc, bb) = (kk_of(i), cc_of(i), hi, lo)
                // stored at ((kk * 4 + cc) * 4 + jj) * 4 + bb
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    sm.xch[w][((kk_of(i) * 4 + hi) * 4 + cc_of(i)) * 4 + lo] = accJ[blk][i];
                __builtin_amdgcn_wave_barrier();
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const T J = sm.xch[w][((kk_of(i) * 4 + cc_of(i)) * 4 + hi) * 4 + lo];
                    const T amat = T(ak) * accK[blk][i] - J;
                    const double am = (double)amat;
                    double pt = am * am / (dr[chv][i] - dc[bh] + 1.0e-10);
                    const bool ok = !__builtin_isnan(pt);
                    if (ok) part[chv * 4 + i] += pt;
                    if (ptm)
                        ptm[((int64_t)own * PT2_SLOTS + (4 * kh + kk_of(i)) * M_T + 4 * chv + cc_of(i)) * ldp +
                            (int64_t)other * PT2_SLOTS + (4 * jh + hi) * M_T + 4 * bh + lo] =
                            ok ? pt : 0.0;
                }
                __builtin_amdgcn_wave_barrier();
                accJ[blk] = Acc{0, 0, 0, 0};
                accK[blk] = Acc{0, 0, 0, 0};
            }
            dc[0] = dcn[0];
            dc[1] = dcn[1];
        }
        if (more) store(cur ^ 1);
        __syncthreads();
    }

#pragma unroll
    for (int q8 = 0; q8 < 8; ++q8) {
        double v = part[q8];
        for (int o = 8; o > 0; o >>= 1) v += __shfl_xor(v, o, 16);
        part[q8] = v;
    }
    if (r16 == 0)
        for (int q8 = 0; q8 < 8; ++q8) sm.red[w][g * 8 + q8] = part[q8];   // q8 = 4 chv + i
    __syncthreads();
    if (threadIdx.x < PT2_SLOTS) {
        // slot t = 8 k + c, k = 4 kh2 + kk, c = 4 c2 + cc: the lane group g and
        // register i that hold it (g = cc, i = kk in FP64; g = kk, i = cc in FP32)
        const int t = threadIdx.x;
        const int k = t / M_T, c = t % M_T;
        const int kh2 = k / 4, kk = k % 4, c2 = c / 4, cc = c % 4;
        const int gg = F64 ? cc : kk, ii = F64 ? kk : cc;
        const double sum = sm.red[2 * kh2][gg * 8 + c2 * 4 + ii] + sm.red[2 * kh2 + 1][gg * 8 + c2 * 4 + ii];
        rowpart[((int64_t)sp * nown + own) * PT2_SLOTS + t] = sum;
    }
}

// rowsum[pair] = sum over the splits, in order, of the partial row sums of
// each secondary slot.
__global__ void pt2_rows_finish_kernel(TileSet rows, int nv, int nsplit,
                                       const double *__restrict__ drow,
                                       const double *__restrict__ rowpart,
                                       double *__restrict__ rowsum) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nslot = (int64_t)rows.ntiles * PT2_SLOTS;
    if (e >= nslot || __builtin_isnan(drow[e])) return;
    const int64_t rt = e / PT2_SLOTS;
    const int slot = (int)(e % PT2_SLOTS);
    const int k = rows.occ[rt * M_T + slot / M_T], c = rows.vir[rt * M_T + slot % M_T];
    double s = rowpart[e];
    for (int sp = 1; sp < nsplit; ++sp) s += rowpart[sp * nslot + e];
    rowsum[(int64_t)k * nv + c] = s;
}

template <typename T>
static void pt2_pipe_launch(hipStream_t s, TileSet rows, TileSet cols, int ncent, int nv,
                            const T *pK, const T *q, const T *pJ, const T *qab, int64_t lda,
                            const double *drow, const double *dcol, double ak, int nsplit,
                            double *rowpart, double *ptm, int64_t ldp) {
    if (rows.ntiles == 0 || cols.ntiles == 0) return;
    if (nsplit < 1 || nsplit > cols.ntiles)
        throw Error(STDA_HIP_ERR_ARG, "PT2 pass 1: split count outside [1, column tiles]");
    const int64_t nitem = (int64_t)rows.ntiles * nsplit;
    if (nitem > INT32_MAX) throw Error(STDA_HIP_ERR_TOO_LARGE, "PT2 pass 1: too many row tiles");
    hipLaunchKernelGGL(pt2_pipe_kernel<T>, dim3((unsigned)nitem), dim3(256), 0, s, rows, cols,
                       ncent, nv, pK, q, pJ, qab, lda, drow, dcol, ak, nsplit, rowpart, ptm, ldp);
}

void launch_pt2_pipe(hipStream_t s, TileSet rows, TileSet cols, int ncent, int nv,
                     const double *pK, const double *q, const double *pJ, const double *qab,
                     int64_t lda, const double *drow, const double *dcol, double ak, int nsplit,
                     double *rowpart, double *ptm, int64_t ldp) {
    if (lda % 2 != 0 || lda < ncent)
        throw Error(STDA_HIP_ERR_ARG, "PT2 pass 1: operand rows must be padded to 2");
    pt2_pipe_launch<double>(s, rows, cols, ncent, nv, pK, q, pJ, qab, lda, drow, dcol, ak,
                            nsplit, rowpart, ptm, ldp);
}

// y (ld x cols) = x (rows x cols, ld rows), the rows rows..ld zero
__global__ void pad_rows_kernel(const double *__restrict__ x, int rows, int64_t cols, int ld,
                                double *__restrict__ y) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)ld * cols) return;
    const int64_t c = e / ld;
    const int r = (int)(e % ld);
    y[e] = r < rows ? x[(int64_t)rows * c + r] : 0.0;
}

void launch_pad_rows(hipStream_t s, const double *x, int rows, int64_t cols, int ld, double *y) {
    const int64_t n = (int64_t)ld * cols;
    if (n == 0) return;
    hipLaunchKernelGGL(pad_rows_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, x,
                       rows, cols, ld, y);
}

void launch_pt2_pipe_f32(hipStream_t s, TileSet rows, TileSet cols, int ncent, int nv,
                         const float *pK, const float *q, const float *pJ, const float *qab,
                         int64_t lda, const double *drow, const double *dcol, double ak,
                         int nsplit, double *rowpart, double *ptm, int64_t ldp) {
    if (lda % 4 != 0 || lda < ncent)
        throw Error(STDA_HIP_ERR_ARG, "PT2 pass 1 (FP32): operand rows must be padded to 4");
    pt2_pipe_launch<float>(s, rows, cols, ncent, nv, pK, q, pJ, qab, lda, drow, dcol, ak, nsplit,
                           rowpart, ptm, ldp);
}

// y (ld x cols) = x (rows x cols, ld rows) in FP32, the rows rows..ld zero
__global__ void to_f32_padded_kernel(const double *__restrict__ x, int rows, int64_t cols, int ld,
                                     float *__restrict__ y) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)ld * cols) return;
    const int64_t c = e / ld;
    const int r = (int)(e % ld);
    y[e] = r < rows ? (float)x[(int64_t)rows * c + r] : 0.0f;
}

void launch_to_f32_padded(hipStream_t s, const double *x, int rows, int64_t cols, int ld,
                          float *y) {
    const int64_t n = (int64_t)ld * cols;
    if (n == 0) return;
    hipLaunchKernelGGL(to_f32_padded_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                       x, rows, cols, ld, y);
}

void launch_pt2_rows_finish(hipStream_t s, TileSet rows, int nv, int nsplit, const double *drow,
                            const double *rowpart, double *rowsum) {
    const int64_t nslot = (int64_t)rows.ntiles * PT2_SLOTS;
    if (nslot == 0) return;
    hipLaunchKernelGGL(pt2_rows_finish_kernel, dim3((unsigned)((nslot + 255) / 256)), dim3(256), 0,
                       s, rows, nv, nsplit, drow, rowpart, rowsum);
}

// Pass 2 from the stored terms: block (cb, sp) sums columns [256 cb, +256)
// over the rows of split sp; every row is a contiguous line of ldp terms.
__global__ void __launch_bounds__(256) pt2_colsum_kernel(int nrej, const int32_t *__restrict__ rrow,
                                                         int64_t ldp, const double *__restrict__ ptm,
                                                         int nsplit, double *__restrict__ part) {
    const int64_t col = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int sp = blockIdx.y;
    const int r0 = (int)((int64_t)nrej * sp / nsplit), r1 = (int)((int64_t)nrej * (sp + 1) / nsplit);
    if (col >= ldp) return;
    double s = 0.0;
    int r = r0;
    for (; r + 4 <= r1; r += 4) {
        const double a0 = ptm[(int64_t)rrow[r] * ldp + col];
        const double a1 = ptm[(int64_t)rrow[r + 1] * ldp + col];
        const double a2 = ptm[(int64_t)rrow[r + 2] * ldp + col];
        const double a3 = ptm[(int64_t)rrow[r + 3] * ldp + col];
        s += a0;
        s += a1;
        s += a2;
        s += a3;
    }
    for (; r < r1; ++r) s += ptm[(int64_t)rrow[r] * ldp + col];
    part[(int64_t)sp * ldp + col] = s;
}

// colsum[c] = the split partials of column c added in split order (loads
// issued eight at a time)
__global__ void pt2_colsum_finish_kernel(int64_t ldp, int nsplit, const double *__restrict__ part,
                                         double *__restrict__ colsum) {
    const int64_t col = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= ldp) return;
    double s = part[col];
    int sp = 1;
    for (; sp + 8 <= nsplit; sp += 8) {
        double x[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) x[u] = part[(int64_t)(sp + u) * ldp + col];
#pragma unroll
        for (int u = 0; u < 8; ++u) s += x[u];
    }
    for (; sp < nsplit; ++sp) s += part[(int64_t)sp * ldp + col];
    colsum[col] = s;
}

void launch_pt2_colsum(hipStream_t s, int nrej, const int32_t *rrow, int64_t ldp,
                       const double *ptm, int nsplit, double *part, double *colsum) {
    if (ldp <= 0) return;
    if (nsplit < 1 || nsplit > std::max(nrej, 1))
        throw Error(STDA_HIP_ERR_ARG, "PT2 pass 2: split count outside [1, rejected rows]");
    const unsigned nb = (unsigned)((ldp + 255) / 256);
    hipLaunchKernelGGL(pt2_colsum_kernel, dim3(nb, nsplit), dim3(256), 0, s, nrej, rrow, ldp, ptm,
                       nsplit, part);
    hipLaunchKernelGGL(pt2_colsum_finish_kernel, dim3(nb), dim3(256), 0, s, ldp, nsplit, part,
                       colsum);
}

}  // namespace stda

namespace stda {

// ----------------------------------------------------------------------------
// PT2 pass 2 for sparse row sets.  A row tile is ONE occupied k with up to
// 16 virtual c (the rejected secondaries of k); a column tile is 16 occupied
// j x 16 virtual b.  Wave w takes b in [4w, 4w+4):
//   J blocks nb = 0..3: rows j (16), cols (c = 4nb + r16/4, b = 4w + r16%4)
//   K blocks mb = 0..3: rows c (16), cols (j = 4mb + r16/4, b = 4w + r16%4)
// The J values are exchanged through LDS (J[c][j][b]) into the K layout.
// Work-group (ct, sp) owns column tile ct and streams the row tiles of split
// sp (a contiguous range); its 256 partial column sums go to
// part[(sp * nct + ct) * 256 + 16 j_local + b_local] and are summed over the
// splits in order by pt2_cols_finish_kernel (deterministic).
constexpr int S_CH = 8;          // atoms per staged chunk (2 MFMA k-steps)
constexpr int S_LD = 16 * 16 + 1;
struct SparseSmem {
    double pj[S_CH][17];         // (k, j)        16
    double pk[S_CH][17];         // (k, c)        16
    double qv[S_CH][S_LD];       // (c, b)        256 = 16 c x 16 b
    double qc[S_CH][S_LD];       // (j, b)        256 = 16 j x 16 b
    double xch[4][16 * 16 * 4];  // per wave: [c][j][b_local]
    double red[4][64];
};

__global__ void __launch_bounds__(256, 2) pt2_cols_sparse_kernel(
    int nrt, const int32_t *__restrict__ rk, const int32_t *__restrict__ rc, int nct,
    const int32_t *__restrict__ cj, const int32_t *__restrict__ cb, int ncent, int nv,
    const double *__restrict__ pK, const double *__restrict__ q, const double *__restrict__ pJ,
    const double *__restrict__ qab, const double *__restrict__ de,
    const unsigned char *__restrict__ rowmask, unsigned char rowbit,
    const unsigned char *__restrict__ colmask, unsigned char colbit, double ak,
    double *__restrict__ part) {
    __shared__ SparseSmem sm;
    const int ct = blockIdx.x;
    const int nsp = gridDim.y, sp = blockIdx.y;
    const int rt0 = (int)((int64_t)nrt * sp / nsp), rt1 = (int)((int64_t)nrt * (sp + 1) / nsp);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int r16 = lane & 15, g = lane >> 4, hi = r16 >> 2, lo = r16 & 3;
    const int t = threadIdx.x;
    double cpart[4] = {0.0, 0.0, 0.0, 0.0};   // columns (j = 4mb + hi, b = 4w + lo)
    // this lane's columns: de (NaN when not a selected primary)
    double dcol[4];
    const int bcol = cb[ct * 16 + 4 * w + lo];
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) dcol[mb] = pair_de(cj[ct * 16 + 4 * mb + hi], bcol, nv, de, colmask, colbit);

    for (int rt = rt0; rt < rt1; ++rt) {
        const int k = rk[rt];
        v4d accJ[4], accK[4];
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            accJ[b] = v4d{0.0, 0.0, 0.0, 0.0};
            accK[b] = v4d{0.0, 0.0, 0.0, 0.0};
        }
        for (int a0 = 0; a0 < ncent; a0 += S_CH) {
            __syncthreads();
            // stage: 544 vectors x 8 atoms; thread -> (atom = t % 8, vector slot = t / 8)
            const int sa = t % S_CH, sv = t / S_CH;   // sv in 0..31
            const int a = a0 + sa;
            const bool in = a < ncent;
            if (sv < 16) {
                int j = cj[ct * 16 + sv];
                sm.pj[sa][sv] = (in && j >= 0) ? pJ[(int64_t)ncent * lin0(k, j) + a] : 0.0;
            } else {
                int c = rc[rt * 16 + sv - 16];
                sm.pk[sa][sv - 16] = (in && c >= 0) ? pK[(int64_t)ncent * ((int64_t)k * nv + c) + a] : 0.0;
            }
#pragma unroll
            for (int r = 0; r < 8; ++r) {
                int v = sv + 32 * r;   // 0..255: (x = v / 16, y = v % 16)
                int x = v >> 4, y = v & 15;
                int c = rc[rt * 16 + x], b = cb[ct * 16 + y], j = cj[ct * 16 + x];
                sm.qv[sa][v] = (in && c >= 0 && b >= 0) ? qab[(int64_t)ncent * lin0(c, b) + a] : 0.0;
                sm.qc[sa][v] = (in && j >= 0 && b >= 0) ? q[(int64_t)ncent * ((int64_t)j * nv + b) + a] : 0.0;
            }
            __syncthreads();
#pragma unroll
            for (int s4 = 0; s4 < S_CH / 4; ++s4) {
                const int arow = 4 * s4 + g;
                const double aJ = sm.pj[arow][r16];
                const double aK = sm.pk[arow][r16];
#pragma unroll
                for (int nb = 0; nb < 4; ++nb) {
                    // J block nb: column (c = 4nb + hi, b = 4w + lo) -> qab slot 16 c + b
                    double bV = sm.qv[arow][16 * (4 * nb + hi) + 4 * w + lo];
                    accJ[nb] = __builtin_amdgcn_mfma_f64_16x16x4f64(aJ, bV, accJ[nb], 0, 0, 0);
                    // K block nb: column (j = 4nb + hi, b = 4w + lo) -> q slot 16 j + b
                    double bQ = sm.qc[arow][16 * (4 * nb + hi) + 4 * w + lo];
                    accK[nb] = __builtin_amdgcn_mfma_f64_16x16x4f64(aK, bQ, accK[nb], 0, 0, 0);
                }
            }
        }
        // J to the K layout: J block nb, lane, reg i -> (c = 4nb + hi, j = 4i + g, b = lo)
#pragma unroll
        for (int nb = 0; nb < 4; ++nb)
#pragma unroll
            for (int i = 0; i < 4; ++i)
                sm.xch[w][((4 * nb + hi) * 16 + 4 * i + g) * 4 + lo] = accJ[nb][i];
        __builtin_amdgcn_wave_barrier();
        // rows of this lane in the K layout: c = 4i + g
        double drow[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) drow[i] = pair_de(k, rc[rt * 16 + 4 * i + g], nv, de, rowmask, rowbit);
#pragma unroll
        for (int mb = 0; mb < 4; ++mb)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                // K block mb, lane, reg i -> (c = 4i + g, j = 4mb + hi, b = lo)
                double J = sm.xch[w][((4 * i + g) * 16 + 4 * mb + hi) * 4 + lo];
                double amat = ak * accK[mb][i] - J;
                double pt = amat * amat / (drow[i] - dcol[mb] + 1.0e-10);
                if (!__builtin_isnan(pt)) cpart[mb] += pt;
            }
        __builtin_amdgcn_wave_barrier();
    }
    // reduce over the four row groups g (lanes 16 apart)
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
        double v = cpart[mb];
        v += __shfl_xor(v, 16, 64);
        v += __shfl_xor(v, 32, 64);
        cpart[mb] = v;
    }
    if (g == 0) {
        double *dst = part + ((int64_t)sp * nct + ct) * 256;
#pragma unroll
        for (int mb = 0; mb < 4; ++mb) dst[(4 * mb + hi) * 16 + 4 * w + lo] = cpart[mb];
    }
}

__global__ void pt2_cols_finish_kernel(int nct, int nsp, const int32_t *__restrict__ cj,
                                       const int32_t *__restrict__ cb, int nv,
                                       const unsigned char *__restrict__ colmask,
                                       unsigned char colbit, const double *__restrict__ part,
                                       double *__restrict__ colsum) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)nct * 256) return;
    const int ct = (int)(e / 256), slot = (int)(e % 256);
    const int j = cj[ct * 16 + slot / 16], b = cb[ct * 16 + slot % 16];
    if (j < 0 || b < 0) return;
    const int64_t p = (int64_t)j * nv + b;
    if (!(colmask[p] & colbit)) return;
    double s = 0.0;
    for (int sp = 0; sp < nsp; ++sp) s += part[((int64_t)sp * nct + ct) * 256 + slot];
    colsum[p] = s;
}

void launch_pt2_cols_sparse(hipStream_t s, int nrt, const int32_t *rk, const int32_t *rc, int nct,
                            const int32_t *cj, const int32_t *cb, int ncent, int nv,
                            const double *pK, const double *q, const double *pJ, const double *qab,
                            const double *de, const unsigned char *rowmask, unsigned char rowbit,
                            const unsigned char *colmask, unsigned char colbit, double ak,
                            int nsplit, double *part, double *colsum) {
    if (nrt == 0 || nct == 0) return;
    if (nsplit < 1) nsplit = 1;
    if (nsplit > nrt) nsplit = nrt;
    hipLaunchKernelGGL(pt2_cols_sparse_kernel, dim3(nct, nsplit), dim3(256), 0, s, nrt, rk, rc, nct,
                       cj, cb, ncent, nv, pK, q, pJ, qab, de, rowmask, rowbit, colmask, colbit, ak,
                       part);
    const int64_t tot = (int64_t)nct * 256;
    hipLaunchKernelGGL(pt2_cols_finish_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s,
                       nct, nsplit, cj, cb, nv, colmask, colbit, part, colsum);
}

// Column norms: one block per (column, slice of rows); partial sums of squares
// are reduced in slice order by a second kernel (deterministic).
__global__ void col_sumsq_partial_kernel(int64_t n, int64_t slice, const double *R, double *part) {
    __shared__ double ws[4];
    int c = blockIdx.y, sl = blockIdx.x;
    int64_t i0 = (int64_t)sl * slice, i1 = min(n, i0 + slice);
    double s = 0.0;
    for (int64_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
        double v = R[i + n * c];
        s += v * v;
    }
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) part[sl + (int64_t)gridDim.x * c] = ws[0] + ws[1] + ws[2] + ws[3];
}

// R = AX - X diag(theta) with the slice partial sums of squares of its
// columns (col_norm_final_kernel completes the norms)
__global__ void residual_sumsq_partial_kernel(int64_t n, int64_t slice, const double *AX,
                                              const double *X, const double *theta, double *R,
                                              double *part) {
    __shared__ double ws[4];
    const int c = blockIdx.y, sl = blockIdx.x;
    const int64_t i0 = (int64_t)sl * slice, i1 = min(n, i0 + slice);
    const double th = theta[c];
    double s = 0.0;
    for (int64_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
        const double v = AX[i + n * c] - th * X[i + n * c];
        R[i + n * c] = v;
        s += v * v;
    }
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) part[sl + (int64_t)gridDim.x * c] = ws[0] + ws[1] + ws[2] + ws[3];
}

__global__ void col_norm_final_kernel(int nsl, int k, const double *part, double *norms) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= k) return;
    double s = 0.0;
    for (int sl = 0; sl < nsl; ++sl) s += part[sl + (int64_t)nsl * c];
    norms[c] = sqrt(s);
}

// Cholesky factor of the Gram matrix S = T^T T (nt x nt, nt <= 64) of a
// block of directions, dependent ones dropped, and the transposed inverse of
// the factor on the kept columns: LinvT (nt x nt, ld nt), so that T LinvT is
// orthonormal on the kept columns and zero on the dropped ones.  Column j is
// dropped when its pivot (the squared norm of its part orthogonal to the kept
// columns before it) is at most thr * ref[j * ref_stride] (thr when ref is
// null); alive[j] = 1.0 for the kept columns, 0.0 for the dropped ones.
// Thread i owns row i of the factor, then column i of its inverse (forward
// substitution).  S is
// given through G ((m + nt) x nt, ld ldg) = [V T]^T T of directions T before
// their projection T - V G[0:m, :] against orthonormal V: S = G[m:m+nt, :] -
// G[0:m, :]^T G[0:m, :] is the Gram matrix of the projected block (m = 0:
// G = T^T T).
//
// With skip non-null (and ref the squared norms of the directions before
// their projection), skip[0] = 1.0 when max_k sum_j |Linv(k, j)| |t_j| <= 1e3,
// else 0.0.  After two projections the part of t_j along V is a few ulps of
// |t_j|, and column k of T LinvT carries at most that sum of them: at the
// bound it stays below about 1e-13 and another projection of T LinvT against
// V would change nothing that matters, so the caller may leave it out.
__global__ __launch_bounds__(256) void chol_inv_kernel(int nt, const double *__restrict__ G, int m,
                                                       int ldg, const double *__restrict__ ref,
                                                       int ref_stride, double thr,
                                                       double *__restrict__ LinvT,
                                                       double *__restrict__ alive,
                                                       double *__restrict__ skip) {
    // L and Z (the inverse factor) share one buffer, which first holds the
    // staged rows of G
    constexpr int SBUF = 2 * 64 * 65;
    __shared__ double buf[SBUF];
    double *L = buf, *Z = buf + 64 * 65;
    __shared__ int keep[64];
    __shared__ double dg[64], lim[64];
    const int tid = threadIdx.x;
    if (tid < nt) lim[tid] = thr * (ref ? ref[(int64_t)tid * ref_stride] : 1.0);
    // S = G[m:m+nt, :] - G[0:m, :]^T G[0:m, :]: the correction over chunks of
    // SBUF / (nt + 1) rows of G (all of them for m <= 395 at nt = 20), loaded
    // eight per thread before any is stored, each thread owning up to 16
    // elements of S
    const int ld = nt + 1, rows_per = SBUF / ld;
    double acc[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) acc[u] = 0.0;
    for (int q0 = 0; q0 < m; q0 += rows_per) {
        const int mq = min(rows_per, m - q0), tot = mq * nt;
        __syncthreads();
        for (int e0 = 0; e0 < tot; e0 += 256 * 8) {
            double v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int e = e0 + tid + 256 * u;
                v[u] = e < tot ? G[(q0 + e % mq) + (int64_t)ldg * (e / mq)] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int e = e0 + tid + 256 * u;
                if (e < tot) buf[(e % mq) * ld + e / mq] = v[u];
            }
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int e = tid + 256 * u;
            if (e < nt * nt) {
                const int r = e % nt, c = e / nt;
                double s = 0.0;
                for (int q = 0; q < mq; ++q) s += buf[q * ld + r] * buf[q * ld + c];
                acc[u] -= s;
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < 16; ++u) {
        const int e = tid + 256 * u;
        if (e < nt * nt) {
            const int r = e % nt, c = e / nt;
            L[r * 65 + c] = G[(m + r) + (int64_t)ldg * c] + acc[u];
        }
    }
    __syncthreads();
    // Right-looking factor, one barrier per step: step j takes the pivot of
    // column j and subtracts l_ij l_kj (l = column j over its root, formed
    // where it is used) from the trailing lower triangle; column j itself is
    // scaled in step j + 1, when nothing reads it.  Every element sees the
    // same operations in the same order as in the left-looking form.
    for (int j = 0; j < nt; ++j) {
        const double d = L[j * 65 + j];
        const bool kp = d > lim[j] && d > 0.0;
        const double ljj = kp ? sqrt(d) : 0.0;
        if (j > 0 && tid >= j && tid < nt) {
            const int pj = j - 1;
            L[tid * 65 + pj] = keep[pj] ? L[tid * 65 + pj] / dg[pj] : 0.0;
        }
        if (tid == 0) {
            keep[j] = kp ? 1 : 0;
            dg[j] = ljj;
        }
        if (kp) {
            const int w = nt - j - 1;
            for (int e = tid; e < w * w; e += 256) {
                const int i = j + 1 + e % w, k = j + 1 + e / w;
                if (k <= i) L[i * 65 + k] -= (L[i * 65 + j] / ljj) * (L[k * 65 + j] / ljj);
            }
        }
        __syncthreads();
    }
    if (tid < nt) L[tid * 65 + tid] = dg[tid];
    // inverse on the kept indices, row by row (forward substitution over all
    // columns at once): step r subtracts L_ir (row r over its pivot) from the
    // rows below; row r is divided in step r + 1
    for (int e = tid; e < nt * nt; e += 256) {
        const int r = e / nt, q = e % nt;
        Z[r * 65 + q] = (r == q && keep[r]) ? 1.0 : 0.0;
    }
    __syncthreads();
    for (int r = 0; r < nt; ++r) {
        if (r > 0 && tid < r && keep[r - 1] && keep[tid]) Z[(r - 1) * 65 + tid] /= dg[r - 1];
        if (keep[r]) {
            const double lrr = dg[r];
            const int w = nt - r - 1, h = r + 1;
            for (int e = tid; e < w * h; e += 256) {
                const int i = r + 1 + e / h, q = e % h;
                if (keep[i] && keep[q]) Z[i * 65 + q] -= L[i * 65 + r] * (Z[r * 65 + q] / lrr);
            }
        }
        __syncthreads();
    }
    if (tid < nt && keep[nt - 1] && keep[tid]) Z[(nt - 1) * 65 + tid] /= dg[nt - 1];
    __syncthreads();
    for (int e = tid; e < nt * nt; e += 256) LinvT[e] = Z[(e / nt) * 65 + e % nt];
    if (tid < nt) alive[tid] = keep[tid] ? 1.0 : 0.0;
    if (skip && ref) {
        // row k of Z holds the coefficients of column k of T LinvT; lim[]
        // is no longer read, so it takes the row sums
        if (tid < nt) {
            double a = 0.0;
            for (int j = 0; j <= tid; ++j)
                a += fabs(Z[tid * 65 + j]) * sqrt(fmax(ref[(int64_t)j * ref_stride], 0.0));
            lim[tid] = a;
        }
        __syncthreads();
        if (tid == 0) {
            double a = 0.0;
            for (int k = 0; k < nt; ++k) a = fmax(a, lim[k]);
            skip[0] = a <= 1.0e3 ? 1.0 : 0.0;   // NaN rows fail the test
        }
    }
}

void launch_chol_inv(hipStream_t s, int nt, const double *G, int m, int ldg, const double *ref,
                     int ref_stride, double thr, double *LinvT, double *alive, double *skip) {
    if (nt <= 0) return;
    if (nt > 64) throw Error(STDA_HIP_ERR_ARG, "chol_inv: more than 64 columns");
    if (skip && !ref) throw Error(STDA_HIP_ERR_ARG, "chol_inv: the skip test needs the norms");
    hipLaunchKernelGGL(chol_inv_kernel, dim3(1), dim3(256), 0, s, nt, G, m, ldg, ref, ref_stride,
                       thr, LinvT, alive, skip);
}

void launch_col_norms_split(hipStream_t s, int64_t n, int k, const double *R, double *part,
                            int nsl, double *norms) {
    if (k == 0) return;
    int64_t slice = (n + nsl - 1) / nsl;
    hipLaunchKernelGGL(col_sumsq_partial_kernel, dim3(nsl, k), dim3(256), 0, s, n, slice, R, part);
    hipLaunchKernelGGL(col_norm_final_kernel, dim3((k + 63) / 64), dim3(64), 0, s, nsl, k, part, norms);
}

void launch_residual_norms(hipStream_t s, int64_t n, int k, const double *AX, const double *X,
                           const double *theta, double *R, double *part, int nsl, double *norms) {
    if (k == 0) return;
    int64_t slice = (n + nsl - 1) / nsl;
    hipLaunchKernelGGL(residual_sumsq_partial_kernel, dim3(nsl, k), dim3(256), 0, s, n, slice, AX, X,
                       theta, R, part);
    hipLaunchKernelGGL(col_norm_final_kernel, dim3((k + 63) / 64), dim3(64), 0, s, nsl, k, part, norms);
}

}  // namespace stda

namespace stda {

// ----------------------------------------------------------------------------
// FP32 variants of the AO-basis operator kernels (mixed-precision solves).
__global__ void scatter_pairs_f32_kernel(const double *v, int64_t ldv, const int32_t *pair,
                                         int64_t n, float *full, int64_t ldfull) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int vec = blockIdx.y;
    if (i < n) full[pair[i] + ldfull * vec] = (float)v[i + ldv * vec];
}

void launch_scatter_pairs_f32(hipStream_t s, const double *v, int64_t ldv, const int32_t *pair,
                              int64_t n, int nvec, float *full, int64_t ldfull) {
    if (n == 0 || nvec == 0) return;
    hipLaunchKernelGGL(scatter_pairs_f32_kernel, dim3((unsigned)((n + 255) / 256), nvec), dim3(256),
                       0, s, v, ldv, pair, n, full, ldfull);
}

// out = (double) full[pair] + diag o v   (diagonal term in FP64)
__global__ void gather_pairs_f32_kernel(const float *full, int64_t ldfull, const int32_t *pair,
                                        int64_t n, const double *diag, const double *v,
                                        int64_t ldv, double *out, int64_t ldout) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int vec = blockIdx.y;
    if (i >= n) return;
    double r = (double)full[pair[i] + ldfull * vec];
    if (diag) r += diag[i] * v[i + ldv * vec];
    out[i + ldout * vec] = r;
}

void launch_gather_pairs_f32(hipStream_t s, const float *full, int64_t ldfull, const int32_t *pair,
                             int64_t n, int nvec, const double *diag, const double *v, int64_t ldv,
                             double *out, int64_t ldout) {
    if (n == 0 || nvec == 0) return;
    hipLaunchKernelGGL(gather_pairs_f32_kernel, dim3((unsigned)((n + 255) / 256), nvec), dim3(256), 0,
                       s, full, ldfull, pair, n, diag, v, ldv, out, ldout);
}

__global__ void atom_diag_f32_kernel(int nao, int ncent, const float *Dt, int64_t strideD,
                                     const int *atom_ptr, const float *gamk, float *u) {
    extern __shared__ float stf[];
    int v = blockIdx.x;
    const float *D = Dt + strideD * v;
    for (int A = threadIdx.x; A < ncent; A += blockDim.x) {
        float s = 0.0f;
        for (int m = atom_ptr[A]; m < atom_ptr[A + 1]; ++m) s += D[m + (int64_t)nao * m];
        stf[A] = s;
    }
    __syncthreads();
    for (int A = threadIdx.x; A < ncent; A += blockDim.x) {
        float s = 0.0f;
        const float *g = gamk + (int64_t)ncent * A;
        for (int B = 0; B < ncent; ++B) s += g[B] * stf[B];
        u[A + (int64_t)ncent * v] = s;
    }
}

void launch_atom_diag_f32(hipStream_t s, int nvec, int nao, int ncent, const float *Dt,
                          int64_t strideD, const int *atom_ptr, const float *gamk, float *u) {
    if (nvec == 0) return;
    hipLaunchKernelGGL(atom_diag_f32_kernel, dim3(nvec), dim3(256), ncent * sizeof(float), s, nao,
                       ncent, Dt, strideD, atom_ptr, gamk, u);
}

__global__ void __launch_bounds__(256) gbuild_f32_kernel(int nao, int ncent, const float *Dt,
                                                         int64_t strideD, const int *atom_of,
                                                         const float *gamj, const float *gamk,
                                                         const float *u, float aJ, float aX,
                                                         float aK, float *Gt, int64_t strideG) {
    __shared__ float tr[GB_T][GB_T + 1];
    int v = blockIdx.z;
    const float *D = Dt + strideD * v;
    float *G = Gt + strideG * v;
    int nu0 = blockIdx.x * GB_T, mu0 = blockIdx.y * GB_T;
    int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    if (aX != 0.0f) {
        for (int r = ty; r < GB_T; r += 8) {
            int mu = mu0 + tx, nu = nu0 + r;
            tr[r][tx] = (mu < nao && nu < nao) ? D[mu + (int64_t)nao * nu] : 0.0f;
        }
        __syncthreads();
    }
    for (int r = ty; r < GB_T; r += 8) {
        int nu = nu0 + tx, mu = mu0 + r;
        if (nu >= nao || mu >= nao) continue;
        int64_t g = atom_of[nu] + (int64_t)ncent * atom_of[mu];
        float val = 0.0f;
        if (aJ != 0.0f) val += aJ * gamj[g] * D[nu + (int64_t)nao * mu];
        if (aX != 0.0f) val += aX * gamk[g] * tr[tx][r];
        if (nu == mu && aK != 0.0f) val += aK * u[atom_of[mu] + (int64_t)ncent * v];
        G[nu + (int64_t)nao * mu] = val;
    }
}

void launch_gbuild_f32(hipStream_t s, int nvec, int nao, int ncent, const float *Dt, int64_t strideD,
                       const int *atom_of, const float *gamj, const float *gamk, const float *u,
                       float aJ, float aX, float aK, float *Gt, int64_t strideG) {
    if (nvec == 0) return;
    dim3 g((nao + GB_T - 1) / GB_T, (nao + GB_T - 1) / GB_T, nvec);
    hipLaunchKernelGGL(gbuild_f32_kernel, g, dim3(256), 0, s, nao, ncent, Dt, strideD, atom_of, gamj,
                       gamk, u, aJ, aX, aK, Gt, strideG);
}

__global__ void to_f32_kernel(const double *x, int64_t n, float *y) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = (float)x[i];
}

void launch_to_f32(hipStream_t s, const double *x, int64_t n, float *y) {
    if (n == 0) return;
    hipLaunchKernelGGL(to_f32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, x, n, y);
}

}  // namespace stda

namespace stda {

__global__ void gather_cols_kernel(int rows, int64_t n, const int32_t *idx, const double *src,
                                   double *dst) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)rows * n) return;
    const int64_t I = e / rows, a = e - I * rows;
    dst[e] = src[a + (int64_t)rows * idx[I]];
}

void launch_gather_cols(hipStream_t s, int rows, int64_t n, const int32_t *idx, const double *src,
                        double *dst) {
    const int64_t tot = (int64_t)rows * n;
    if (tot == 0) return;
    hipLaunchKernelGGL(gather_cols_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s,
                       rows, n, idx, src, dst);
}

__global__ void set_diag_kernel(int64_t n, const double *d, double *A) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) A[i * (n + 1)] = d[i];
}

void launch_set_diag(hipStream_t s, int64_t n, const double *d, double *A) {
    if (n == 0) return;
    hipLaunchKernelGGL(set_diag_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, d, A);
}

}  // namespace stda

namespace stda {

// ----------------------------------------------------------------------------
// Tall-skinny FP64 products of the solvers (n rows, a few dozen columns).
// Memory-bound: each input element is read once per launch (the B columns of
// C = A^T B once per 32 columns of A).
//
// C (m x p) = A^T B, A (n x m), B (n x p): work-group (s, ib) reduces the
// rows of slice s for the 32 columns ib*32.. of A into part[s*m*p + i + m*j],
// on the matrix cores (v_mfma_f64_16x16x4f64, the rows as the reduction
// index).  Rows are staged through LDS in TS_CH = 32-row chunks with
// coalesced loads, the next chunk's loads in flight (registers) while the
// current one is multiplied; the chunk's 8 steps of 4 rows are dealt to the
// four waves (step s to wave s % 4).  (32-row chunks leave more work-groups
// per CU than 64-row ones: 3.39 against 2.71 TB/s at n = 1.2 M.)  Lane
// l = 16g + c supplies A^T[16mb + c][row 4s + g]
// and B[row 4s + g][16pb + c]; register i of accumulator (mb, pb) holds
// C(16mb + 4i + g, 16pb + c).  The wave partials are added in wave order at
// the end.  tsgemm_reduce_kernel then sums the slices in slice order.
constexpr int TS_ROWS = 32, TS_CH = 32;

// (five waves per SIMD: the PB = 2 form then fits 86 registers without
// spilling; 3.47 against 3.39 TB/s at n = 1.2 M)
template <int PB>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) void tsgemm_tn_partial_kernel(int64_t n, int m, int p,
                                                                const double *__restrict__ A,
                                                                int m1,
                                                                const double *__restrict__ A2,
                                                                int64_t lda,
                                                                const double *__restrict__ B,
                                                                int64_t ldb, int64_t rows,
                                                                double *__restrict__ part,
                                                                const double *__restrict__ skip1) {
    // the m columns of A: A's first m1 columns, then those of A2 (same lda);
    // columns below lo read as zero (skip1 set: A's columns are not read)
    constexpr int PJ = 16 * PB, QA = TS_CH * 32 / 256, QB = TS_CH * PJ / 256;
    // LDS: the staged chunk (As[64][33], Bs[64][PJ+1]), then the wave partials
    constexpr int LDS_STAGE = TS_CH * 33 + TS_CH * (PJ + 1), LDS_RED = 3 * 64 * 8 * PB;
    __shared__ double lds[LDS_STAGE > LDS_RED ? LDS_STAGE : LDS_RED];
    double *As = lds, *Bs = lds + TS_CH * 33;
    const int tid = threadIdx.x, l = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int i0 = 32 * blockIdx.y, mi = min(32, m - i0);
    const int c = l & 15, g = l >> 4;
    const int64_t rb = (int64_t)blockIdx.x * rows, re = min(n, rb + rows);
    const int lo = (skip1 && *skip1 != 0.0) ? m1 : 0;
    if (i0 + mi <= lo) {   // every column of this block reads as zero
        double *dst = part + (int64_t)blockIdx.x * m * p;
        for (int e = tid; e < mi * p; e += 256) dst[(i0 + e % mi) + (int64_t)m * (e / mi)] = 0.0;
        return;
    }
    double pa[QA], pb[QB];
    auto fetch = [&](int64_t r0) {
        const int nr = (int)min((int64_t)TS_CH, re - r0);
#pragma unroll
        for (int q = 0; q < QA; ++q) {
            const int e = tid + 256 * q, rr = e % TS_CH, i = e / TS_CH, col = i0 + i;
            const double *ac = col < m1 ? A + lda * col : A2 + lda * (col - m1);
            pa[q] = (rr < nr && i < mi && col >= lo) ? ac[r0 + rr] : 0.0;
        }
#pragma unroll
        for (int q = 0; q < QB; ++q) {
            const int e = tid + 256 * q, rr = e % TS_CH, j = e / TS_CH;
            pb[q] = (rr < nr && j < p) ? B[(r0 + rr) + ldb * j] : 0.0;
        }
    };
    v4d acc[2][PB];
#pragma unroll
    for (int mb = 0; mb < 2; ++mb)
#pragma unroll
        for (int u = 0; u < PB; ++u) acc[mb][u] = v4d{0.0, 0.0, 0.0, 0.0};
    if (rb < re) fetch(rb);
    for (int64_t r0 = rb; r0 < re; r0 += TS_CH) {
        __syncthreads();   // the previous chunk is consumed
#pragma unroll
        for (int q = 0; q < QA; ++q) {
            const int e = tid + 256 * q;
            As[(e % TS_CH) * 33 + e / TS_CH] = pa[q];
        }
#pragma unroll
        for (int q = 0; q < QB; ++q) {
            const int e = tid + 256 * q;
            Bs[(e % TS_CH) * (PJ + 1) + e / TS_CH] = pb[q];
        }
        __syncthreads();
        if (r0 + TS_CH < re) fetch(r0 + TS_CH);
        // rows past the slice end are zero in the staged chunk
#pragma unroll
        for (int st = 0; st < TS_CH / 16; ++st) {
            const int rr = 4 * (w + 4 * st) + g;
            const double a0 = As[rr * 33 + c], a1 = As[rr * 33 + 16 + c];
#pragma unroll
            for (int u = 0; u < PB; ++u) {
                const double b = Bs[rr * (PJ + 1) + 16 * u + c];
                acc[0][u] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b, acc[0][u], 0, 0, 0);
                acc[1][u] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b, acc[1][u], 0, 0, 0);
            }
        }
    }
    __syncthreads();   // every wave is done with the staged chunk
    double *red = lds;  // red[(w-1)][v][l], v = 4 (PB mb + u) + i < 8 PB
    if (w > 0) {
#pragma unroll
        for (int mb = 0; mb < 2; ++mb)
#pragma unroll
            for (int u = 0; u < PB; ++u)
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    red[((w - 1) * 8 * PB + 4 * (PB * mb + u) + i) * 64 + l] = acc[mb][u][i];
    }
    __syncthreads();
    if (w == 0) {
        double *dst = part + (int64_t)blockIdx.x * m * p;
#pragma unroll
        for (int mb = 0; mb < 2; ++mb)
#pragma unroll
            for (int u = 0; u < PB; ++u)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int v = 4 * (PB * mb + u) + i;
                    const int ii = 16 * mb + 4 * i + g, j = 16 * u + c;
                    const double s = ((acc[mb][u][i] + red[(0 * 8 * PB + v) * 64 + l]) +
                                      red[(1 * 8 * PB + v) * 64 + l]) + red[(2 * 8 * PB + v) * 64 + l];
                    if (ii < mi && j < p) dst[(i0 + ii) + (int64_t)m * j] = s;
                }
    }
}

// C[e] = sum_s part[s*len + e]: a work-group of 16 groups x 16 elements;
// group g sums slices s = g, g+16, ... in order (loads issued eight at a time),
// then the 16 group sums are added in group order (deterministic).
__global__ __launch_bounds__(256) void tsgemm_reduce_kernel(int64_t len, int ns,
                                                            const double *__restrict__ part,
                                                            double *__restrict__ C) {
    __shared__ double red[16][17];
    const int l = threadIdx.x & 15, g = threadIdx.x >> 4;
    const int64_t e = (int64_t)blockIdx.x * 16 + l;
    double s = 0.0;
    if (e < len) {
        int q = g;
        for (; q + 7 * 16 < ns; q += 8 * 16) {
            double x[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) x[u] = part[(int64_t)(q + 16 * u) * len + e];
#pragma unroll
            for (int u = 0; u < 8; ++u) s += x[u];
        }
        for (; q < ns; q += 16) s += part[(int64_t)q * len + e];
    }
    red[g][l] = s;
    __syncthreads();
    if (g == 0 && e < len) {
        double t = red[0][l];
#pragma unroll
        for (int h = 1; h < 16; ++h) t += red[h][l];
        C[e] = t;
    }
}

// Rows per slice: whole 32-row chunks, aiming at about 1024 work-groups in
// all (slices x 32-column blocks of A), with the partial sums kept below a
// quarter of the input bytes.
static int64_t tsgemm_tn_rows(int64_t n, int m, int p) {
    const int64_t cb = (m + 31) / 32;
    int64_t ns = std::max<int64_t>(1, 1024 / cb);
    const int64_t cap = (n * (m + p)) / (4 * (int64_t)m * p);
    ns = std::max<int64_t>(1, std::min(ns, cap));
    const int64_t r = (n + ns - 1) / ns;
    return std::max<int64_t>(TS_ROWS, (r + TS_ROWS - 1) / TS_ROWS * TS_ROWS);
}

int tsgemm_tn_slices(int64_t n, int m, int p) {
    if (n <= 0 || m <= 0 || p <= 0) return 0;
    const int64_t rows = tsgemm_tn_rows(n, m, p);
    return (int)((n + rows - 1) / rows);
}

void launch_tsgemm_tn2_partial(hipStream_t s, int64_t n, int m1, const double *A1, int m2,
                               const double *A2, int p, int64_t lda, const double *B, int64_t ldb,
                               double *part, const double *skip1) {
    const int m = m1 + m2;
    const int ns = tsgemm_tn_slices(n, m, p);
    if (ns == 0) return;
    const int64_t rows = tsgemm_tn_rows(n, m, p);
    const dim3 grid(ns, (m + 31) / 32);
    if (!A2) A2 = A1;
    if (p <= 16)
        hipLaunchKernelGGL(tsgemm_tn_partial_kernel<1>, grid, dim3(256), 0, s, n, m, p, A1, m1, A2, lda, B, ldb, rows, part, skip1);
    else if (p <= 32)
        hipLaunchKernelGGL(tsgemm_tn_partial_kernel<2>, grid, dim3(256), 0, s, n, m, p, A1, m1, A2, lda, B, ldb, rows, part, skip1);
    else
        hipLaunchKernelGGL(tsgemm_tn_partial_kernel<4>, grid, dim3(256), 0, s, n, m, p, A1, m1, A2, lda, B, ldb, rows, part, skip1);
}

void launch_tsgemm_tn_partial(hipStream_t s, int64_t n, int m, int p, const double *A,
                              int64_t lda, const double *B, int64_t ldb, double *part) {
    launch_tsgemm_tn2_partial(s, n, m, A, 0, nullptr, p, lda, B, ldb, part, nullptr);
}

void launch_tsgemm_tn_reduce(hipStream_t s, int64_t n, int m, int p, const double *part,
                             double *C) {
    const int ns = tsgemm_tn_slices(n, m, p);
    if (ns == 0) return;
    const int64_t len = (int64_t)m * p;
    hipLaunchKernelGGL(tsgemm_reduce_kernel, dim3((unsigned)((len + 15) / 16)), dim3(256), 0, s, len,
                       ns, part, C);
}

void launch_tsgemm_tn(hipStream_t s, int64_t n, int m, int p, const double *A, int64_t lda,
                      const double *B, int64_t ldb, double *part, double *C) {
    launch_tsgemm_tn_partial(s, n, m, p, A, lda, B, ldb, part);
    launch_tsgemm_tn_reduce(s, n, m, p, part, C);
}

// out (n x p) = alpha A C + beta out, A (n x m), C (m x p), p <= 32, on the
// matrix cores (v_mfma_f64_16x16x4f64).  A work-group covers 64 rows, wave w
// the 16-row block 16w..16w+15, and NB column blocks of 16.  C is staged
// through LDS in chunks of NN_CH = 32 rows; each lane loads its 8 A values
// of a chunk (row 16w + c of the block, inner indices g + 4u: lane l = 16g + c
// supplies the A value of row c, inner index g, of MFMA step u) before the
// multiplies.  (Smaller chunks take less LDS and fewer registers, so more
// work-groups share a CU: 4.50 TB/s at n = 1.2 M against 4.15 with 64-row
// and 3.36 with 128-row chunks.)  The products take the C values as the
// first operand (the transposed block), so register i of accumulator j holds
// row 16w + c, column 16j + 4i + g.  beta == 0 never reads out.
constexpr int NN_CH = 32;
template <int NB>
__global__ __launch_bounds__(256) void tsgemm_nn_kernel(int64_t n, int m, int p, double alpha,
                                                        const double *__restrict__ A0,
                                                        const double *__restrict__ A1, int64_t lda,
                                                        const double *__restrict__ C, int64_t ldc,
                                                        double beta, double *out0, double *out1,
                                                        int64_t ldo, const double *__restrict__ skip) {
    if (skip && *skip != 0.0) return;   // the caller's flag: this product is not needed
    // blockIdx.y selects the product: (A0, out0) or (A1, out1), same C
    const double *__restrict__ A = blockIdx.y ? A1 : A0;
    double *out = blockIdx.y ? out1 : out0;
    constexpr int PJ = 16 * NB;
    __shared__ double cs[NN_CH * PJ];
    const int tid = threadIdx.x, lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int c = lane & 15, g = lane >> 4;
    const int64_t r0 = (int64_t)blockIdx.x * 64 + 16 * w;
    const int64_t row = r0 + c;
    const bool live = row < n;
    v4d acc[NB];
#pragma unroll
    for (int j = 0; j < NB; ++j) acc[j] = v4d{0.0, 0.0, 0.0, 0.0};
    for (int i0 = 0; i0 < m; i0 += NN_CH) {
        const int mi = min(NN_CH, m - i0);
        __syncthreads();
        for (int e = tid; e < NN_CH * PJ; e += 256) {
            const int i = e / PJ, j = e % PJ;
            cs[e] = (i < mi && j < p) ? C[(i0 + i) + ldc * j] : 0.0;
        }
        double a[NN_CH / 4];
        const double *ap = A + (live ? row : 0) + lda * (int64_t)(i0 + g);
#pragma unroll
        for (int u = 0; u < NN_CH / 4; ++u) {
            a[u] = (live && g + 4 * u < mi) ? *ap : 0.0;
            ap += 4 * lda;
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < NN_CH / 4; ++u) {
            if (4 * u < mi) {
#pragma unroll
                for (int j = 0; j < NB; ++j)
                    acc[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(cs[(4 * u + g) * PJ + 16 * j + c], a[u],
                                                                  acc[j], 0, 0, 0);
            }
        }
    }
    // the C values are the first operand, so acc[j] holds the transposed block:
    // lane c is the row (coalesced stores), register i with g the column
#pragma unroll
    for (int j = 0; j < NB; ++j)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int64_t r = r0 + c;
            const int col = 16 * j + 4 * i + g;
            if (r < n && col < p) {
                double v = alpha * acc[j][i];
                if (beta != 0.0) v += beta * out[r + ldo * col];
                out[r + ldo * col] = v;
            }
        }
}

void launch_tsgemm_nn2(hipStream_t s, int64_t n, int m, int p, double alpha, const double *A0,
                       const double *A1, int64_t lda, const double *C, int64_t ldc, double beta,
                       double *out0, double *out1, int64_t ldo, const double *skip) {
    if (n == 0 || p == 0) return;
    const dim3 grid((unsigned)((n + 63) / 64), A1 ? 2 : 1);
    for (int j0 = 0; j0 < p; j0 += 32) {
        const int pj = std::min(32, p - j0);
        const double *Cj = C + ldc * j0;
        double *oj0 = out0 + ldo * j0;
        double *oj1 = out1 ? out1 + ldo * j0 : oj0;
        if (pj <= 16)
            hipLaunchKernelGGL(tsgemm_nn_kernel<1>, grid, dim3(256), 0, s, n, m, pj, alpha, A0,
                               A1 ? A1 : A0, lda, Cj, ldc, beta, oj0, oj1, ldo, skip);
        else
            hipLaunchKernelGGL(tsgemm_nn_kernel<2>, grid, dim3(256), 0, s, n, m, pj, alpha, A0,
                               A1 ? A1 : A0, lda, Cj, ldc, beta, oj0, oj1, ldo, skip);
    }
}

void launch_tsgemm_nn(hipStream_t s, int64_t n, int m, int p, double alpha, const double *A,
                      int64_t lda, const double *C, int64_t ldc, double beta, double *out,
                      int64_t ldo, const double *skip) {
    launch_tsgemm_nn2(s, n, m, p, alpha, A, nullptr, lda, C, ldc, beta, out, nullptr, ldo, skip);
}

}  // namespace stda
