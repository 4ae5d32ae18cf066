// Device parts of the upstream transforms (the AO -> MO moments of onetri and
// shrink, stda.f:351-412 and 1695-1706, and the Loewdin orthogonalisation of
// makel, stda.f:430-452 and 1615-1691): the products themselves are rocBLAS
// GEMMs and a rocSOLVER eigensolve on the host side; these kernels pack and
// scale.  Both are HBM-bound element-wise passes.

#include "common.h"

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

void launch_scale_rows(hipStream_t s, int64_t m, int64_t k, const double *f, double *T) {
    if (m <= 0 || k <= 0) return;
    hipLaunchKernelGGL(scale_rows_kernel, dim3(grid_for(m * k)), dim3(256), 0, s, m, k, f, T);
    STDA_HIP_CHECK(hipGetLastError());
}

}  // namespace stda
