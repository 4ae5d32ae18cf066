// Device parts of the upstream transforms (the AO -> MO moments of onetri and
// shrink, stda.f:351-412 and 1695-1706, and the Loewdin orthogonalisation of
// makel, stda.f:430-452 and 1615-1691): the products themselves are rocBLAS
// GEMMs and a rocSOLVER eigensolve on the host side; these kernels pack and
// scale (HBM-bound element-wise passes), and compute the one-electron AO
// integrals the transforms start from (libcint.f:252-352, 653-756;
// intpack.f90:1208-1283), one thread per AO pair.

#include "common.h"

// the integrals keep the operation order of the host recurrences (no contraction)
#pragma clang fp contract(off)

namespace stda {

namespace {

// out[lin(i, j)] = P(j, i) for j <= i: the packed upper triangle (0-based
// lin(i, j) = j + i (i + 1) / 2), P (n x n) column-major.
__global__ void pack_upper_kernel(int64_t n, const double *__restrict__ P, double *out) {
    const int64_t npk = n * (n + 1) / 2;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < npk;
         e += (int64_t)gridDim.x * blockDim.x) {
        // invert e = j + i (i + 1) / 2 with 0 <= j <= i
        int64_t i = (int64_t)((sqrt(8.0 * (double)e + 1.0) - 1.0) * 0.5);
        while (i * (i + 1) / 2 > e) --i;
        while ((i + 1) * (i + 2) / 2 <= e) ++i;
        const int64_t j = e - i * (i + 1) / 2;
        out[e] = P[j + n * i];
    }
}

// T(i, :) *= f[i] for the rows of T (m x k, leading dimension m)
__global__ void scale_rows_kernel(int64_t m, int64_t k, const double *__restrict__ f, double *T) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < m * k;
         e += (int64_t)gridDim.x * blockDim.x)
        T[e] *= f[e % m];
}

// One-electron integrals over contracted Cartesian Gaussians for AO pair (mu, nu):
// the Obara-Saika 1-D overlaps S[t][u] (t <= l_mu, u <= l_nu + 1) of each
// primitive pair and direction, then
//   s = S[la][lb], x = S[la][lb+1] + B s (r about the origin), g = lb S[la][lb-1] - 2 b S[la][lb+1]
// and S = sx sy sz, R_k = -(x_k s s), V_k = -(g_k s s), M = s (x g - g x) (cyclic),
// contracted with the primitive coefficients, ket primitives summed first.
// Output column-major: out[q][mu + nao nu] for q = S, R x y z, M x y z, V x y z.
__global__ void __launch_bounds__(256)
ao_ints_kernel(int nao, const int64_t *__restrict__ ptr, const double *__restrict__ alpha,
               const double *__restrict__ coef, const double *__restrict__ cen,
               const int32_t *__restrict__ lmn, double *S, double *ao9) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)nao * nao) return;
    const int mu = (int)(e % nao), nu = (int)(e / nao);
    double A[3], B[3];
    int la[3], lb[3];
    for (int d = 0; d < 3; ++d) {
        A[d] = cen[3 * mu + d];
        B[d] = cen[3 * nu + d];
        la[d] = lmn[3 * mu + d];
        lb[d] = lmn[3 * nu + d];
    }
    double acc[10];
    for (int q = 0; q < 10; ++q) acc[q] = 0.0;
    for (int64_t i = ptr[mu]; i < ptr[mu + 1]; ++i) {
        const double a = alpha[i];
        double inner[10];
        for (int q = 0; q < 10; ++q) inner[q] = 0.0;
        for (int64_t j = ptr[nu]; j < ptr[nu + 1]; ++j) {
            const double b = alpha[j];
            const double p = a + b, inv2p = 0.5 / p, mup = a * b / p;
            double sd[3], xd[3], gd[3];
            for (int d = 0; d < 3; ++d) {
                const double P = (a * A[d] + b * B[d]) / p;
                const double XPA = P - A[d], XPB = P - B[d], XAB = A[d] - B[d];
                double T[AO_LMAX + 1][AO_LMAX + 3];
                const int tmax = la[d], umax = lb[d] + 1;
                T[0][0] = sqrt(M_PI / p) * exp(-mup * XAB * XAB);
                for (int t = 0; t <= tmax; ++t) {
                    if (t > 0)
                        T[t][0] = XPA * T[t - 1][0] + (t >= 2 ? (t - 1) * T[t - 2][0] : 0.0) * inv2p;
                    for (int u = 1; u <= umax; ++u) {
                        double v = XPB * T[t][u - 1];
                        if (t > 0) v = v + t * T[t - 1][u - 1] * inv2p;
                        if (u > 1) v = v + (u - 1) * T[t][u - 2] * inv2p;
                        T[t][u] = v;
                    }
                }
                const double s0 = T[tmax][lb[d]], sp1 = T[tmax][lb[d] + 1];
                const double sm1 = lb[d] > 0 ? T[tmax][lb[d] - 1] : 0.0;
                sd[d] = s0;
                xd[d] = sp1 + B[d] * s0;
                gd[d] = lb[d] * sm1 - 2.0 * b * sp1;
            }
            const double w = coef[i] * coef[j];
            const double v[10] = {sd[0] * sd[1] * sd[2],
                                  -(xd[0] * sd[1] * sd[2]), -(sd[0] * xd[1] * sd[2]),
                                  -(sd[0] * sd[1] * xd[2]),
                                  sd[0] * (xd[1] * gd[2] - gd[1] * xd[2]),
                                  sd[1] * (xd[2] * gd[0] - gd[2] * xd[0]),
                                  sd[2] * (xd[0] * gd[1] - gd[0] * xd[1]),
                                  -(gd[0] * sd[1] * sd[2]), -(sd[0] * gd[1] * sd[2]),
                                  -(sd[0] * sd[1] * gd[2])};
            for (int q = 0; q < 10; ++q) inner[q] = inner[q] + v[q] * w;
        }
        for (int q = 0; q < 10; ++q) acc[q] = acc[q] + inner[q];
    }
    const int64_t nn = (int64_t)nao * nao;
    S[e] = acc[0];
    for (int q = 0; q < 9; ++q) ao9[q * nn + e] = acc[1 + q];
}

unsigned grid_for(int64_t n) {
    const int64_t b = (n + 255) / 256;
    return (unsigned)(b < 8192 ? (b > 0 ? b : 1) : 8192);
}

}  // namespace

void launch_pack_upper(hipStream_t s, int64_t n, const double *P, double *out) {
    if (n <= 0) return;
    hipLaunchKernelGGL(pack_upper_kernel, dim3(grid_for(n * (n + 1) / 2)), dim3(256), 0, s, n, P,
                       out);
    STDA_HIP_CHECK(hipGetLastError());
}

void launch_ao_ints(hipStream_t s, int nao, const int64_t *ptr, const double *alpha,
                    const double *coef, const double *cen, const int32_t *lmn, double *S,
                    double *ao9) {
    const int64_t n = (int64_t)nao * nao;
    if (n == 0) return;
    hipLaunchKernelGGL(ao_ints_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, nao,
                       ptr, alpha, coef, cen, lmn, S, ao9);
    STDA_HIP_CHECK(hipGetLastError());
}

void launch_scale_rows(hipStream_t s, int64_t m, int64_t k, const double *f, double *T) {
    if (m <= 0 || k <= 0) return;
    hipLaunchKernelGGL(scale_rows_kernel, dim3(grid_for(m * k)), dim3(256), 0, s, m, k, f, T);
    STDA_HIP_CHECK(hipGetLastError());
}

}  // namespace stda
