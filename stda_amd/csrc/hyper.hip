// Quadratic response (SHG first hyperpolarizability): the device parts of
// lresp1's beta (linear_response.f:373-699, 911-1226).
//
// The reference sums, for every pair of selected CSFs sharing a virtual (i,a),(j,a)
// or an occupied (i,a),(i,b), products X[ia] mu[ij] Y[ja] and X[ia] mu[ab] Y[ib]
// over index lists (List_A / List_B).  Here the response vectors are scattered
// into dense occupied x virtual matrices (zero for CSFs outside the selection,
// which removes exactly the pairs the lists leave out), so that
//   A - B = -<X_x, mu_y^oo Y_z - Y_z mu_y^vv>_F
// becomes two GEMMs per (y, z, frequency) and one inner-product GEMM for all
// (x, frequency) pairs (rocBLAS, host side).  The kernels below build the dense
// operands; both are HBM-bound scatters/gathers.

#include "../../include/stda_hip.h"
#include "common.h"

namespace stda {

namespace {

__device__ inline int64_t lin0h(int64_t a, int64_t b) {   // 0-based packed index, a,b 0-based
    const int64_t hi = a > b ? a : b, lo = a > b ? b : a;
    return lo + hi * (hi + 1) / 2;
}

// X = ((X+Y) + (X-Y))/2, Y = (X+Y) - X with X-Y = ((A+B)(X+Y) - rhs) / w (linear_response.f:375-400:
// X-Y = w (A-B)^-1 (X+Y), and (A-B)^-1 (X+Y) = ((A+B)(X+Y) - rhs) / w^2 by the response equation);
// column c of the output reads column src[c] of XP/AX, component c % 3 of rhs; X and Y
// are multiplied by cols.scale.
__global__ void beta_xy_kernel(int64_t n, const int32_t *__restrict__ pair, int no, int nv,
                               const double *__restrict__ XP, const double *__restrict__ AX,
                               const double *__restrict__ rhs, BetaCols cols, double *Xd,
                               double *Yd) {
    const int64_t L = (int64_t)no * nv;
    const int c = blockIdx.y;
    const int64_t src = cols.src[c];
    const double iw = cols.inv_w[c];
    for (int64_t I = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; I < n;
         I += (int64_t)gridDim.x * blockDim.x) {
        const double xpy = XP[I + n * src];
        const double xmy = iw == 0.0 ? 0.0 : (AX[I + n * src] - rhs[I + n * (c % 3)]) * iw;
        const double x = 0.5 * (xpy + xmy);
        const int64_t p = pair[I];
        const int64_t idx = (p / nv) + (int64_t)no * (p % nv);
        Xd[idx + L * c] = cols.scale * x;
        Yd[idx + L * c] = cols.scale * (xpy - x);
    }
}

__global__ void scatter_pairs_kernel(int64_t n, const int32_t *__restrict__ pair, int no, int nv,
                                     const double *__restrict__ V, int64_t ldv, double *D) {
    const int64_t L = (int64_t)no * nv;
    const int c = blockIdx.y;
    for (int64_t I = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; I < n;
         I += (int64_t)gridDim.x * blockDim.x) {
        const int64_t p = pair[I];
        D[(p / nv) + (int64_t)no * (p % nv) + L * c] = V[I + ldv * c];
    }
}

// mu^oo (no x no) and mu^vv (nv x nv) of the three dipole components from the packed moments
__global__ void mu_blocks_kernel(const double *__restrict__ moments, int64_t npk, int no, int nv,
                                 double *muoo, double *muvv) {
    const int q = blockIdx.y;
    const int64_t noo = (int64_t)no * no, nvv = (int64_t)nv * nv;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < noo + nvv;
         e += (int64_t)gridDim.x * blockDim.x) {
        if (e < noo) {
            const int64_t i = e % no, j = e / no;
            muoo[q * noo + e] = moments[q * npk + lin0h(i, j)];
        } else {
            const int64_t f = e - noo, a = f % nv, b = f / nv;
            muvv[q * nvv + f] = moments[q * npk + lin0h(no + a, no + b)];
        }
    }
}

unsigned blocks_for(int64_t n) {
    const int64_t b = (n + 255) / 256;
    return (unsigned)(b < 4096 ? (b > 0 ? b : 1) : 4096);
}

}  // namespace

void launch_beta_xy(hipStream_t s, int64_t n, const int32_t *pair, int no, int nv,
                    const double *XP, const double *AX, const double *rhs, const BetaCols &cols,
                    int ncols, double *Xd, double *Yd) {
    if (n == 0 || ncols <= 0) return;
    if (ncols > 6) throw Error(STDA_HIP_ERR_ARG, "launch_beta_xy: at most 6 columns");
    hipLaunchKernelGGL(beta_xy_kernel, dim3(blocks_for(n), ncols), dim3(256), 0, s, n, pair, no,
                       nv, XP, AX, rhs, cols, Xd, Yd);
    STDA_HIP_CHECK(hipGetLastError());
}

void launch_scatter_pairs(hipStream_t s, int64_t n, const int32_t *pair, int no, int nv, int k,
                          const double *V, int64_t ldv, double *D) {
    if (n == 0 || k <= 0) return;
    hipLaunchKernelGGL(scatter_pairs_kernel, dim3(blocks_for(n), k), dim3(256), 0, s, n, pair, no,
                       nv, V, ldv, D);
    STDA_HIP_CHECK(hipGetLastError());
}

void launch_mu_blocks(hipStream_t s, const double *moments, int64_t npk, int no, int nv,
                      double *muoo, double *muvv) {
    const int64_t tot = (int64_t)no * no + (int64_t)nv * nv;
    if (tot == 0) return;
    hipLaunchKernelGGL(mu_blocks_kernel, dim3(blocks_for(tot), 3), dim3(256), 0, s, moments, npk,
                       no, nv, muoo, muvv);
    STDA_HIP_CHECK(hipGetLastError());
}

}  // namespace stda
