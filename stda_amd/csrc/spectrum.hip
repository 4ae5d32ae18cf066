// Gaussian broadening of line spectra: the curve loop of the reference's
// spectrum program g_spec (g_spec/g_spec.f:371-433).
//
// For spectrum s with lines [off[s], off[s+1]) (energies x in eV, strengths y,
// widths w) and grid points j = 1..npts:
//   e_j = emin[s] + j step[s]
//   r_j = sum over lines with |x| > 1e-6, in line order, of
//         f3 f2 exp(-(e_j - x)^2 / w^2),
//   f3 = 1 / (f1 sqrt(pi) w),  f2 = x y (CD) or x (c y / (3.67493e-2 x)) (UV),
// then r_j *= rfaktor and the optional nm / cm / log10 transforms of
// g_spec.f:414-429.  The reference's single-precision literals keep their
// float values, and the whole file is compiled without floating-point
// contraction so that every product and sum rounds as the reference's do.
//
// Layout: one 256-thread workgroup per 256 grid points of one spectrum
// (grid.x over points, grid.y over spectra); the lines of the spectrum go
// through LDS in chunks of 256, each thread sums all of them in order for its
// own point.  Work per (point, line): one exp and five flops; the kernel is
// bound by FP64 VALU throughput (exp), HBM traffic is the lines once per
// workgroup plus 16 bytes per point.

#include "common.h"

#pragma clang fp contract(off)

namespace stda {

namespace {

constexpr int GS_BLOCK = 256;
constexpr double GS_SQRT_PI = (double)1.7724539f;     // g_spec.f:384 (default real)
constexpr double GS_F1_CD = (double)22.97f;           // g_spec.f:376
constexpr double GS_F1_UV = 9.186e-3;                 // g_spec.f:377
constexpr double GS_EV_CM = (double)8065.54093f;      // g_spec.f:419-420
constexpr double GS_DIP = (double)(6.4607f * 1.5f);   // g_spec.f:395, folded in single precision

__global__ void __launch_bounds__(GS_BLOCK)
gspec_kernel(const int64_t *__restrict__ off, const double *__restrict__ x,
             const double *__restrict__ y, const double *__restrict__ w,
             const double *__restrict__ emin, const double *__restrict__ step, int npts,
             int flags, double rfaktor, double *__restrict__ eout, double *__restrict__ rout) {
    __shared__ double sx[GS_BLOCK], samp[GS_BLOCK], swid[GS_BLOCK];
    __shared__ int sok[GS_BLOCK];
    const int s = blockIdx.y, t = threadIdx.x;
    const int j = blockIdx.x * GS_BLOCK + t + 1;
    const bool uv = flags & GSPEC_UV;
    const double f1 = uv ? GS_F1_UV : GS_F1_CD;
    const int64_t l0 = off[s], l1 = off[s + 1];
    const double ee = emin[s] + (double)j * step[s];
    double r = 0.0;
    for (int64_t c = l0; c < l1; c += GS_BLOCK) {
        const int64_t i = c + t;
        __syncthreads();
        if (i < l1) {
            const double xi = x[i], yi = y[i], wi = w[i];
            const bool ok = fabs(xi) > 1.0e-6;
            double amp = 0.0, wid = 0.0;
            if (ok) {
                const double f3 = 1.0 / (f1 * GS_SQRT_PI * wi);
                wid = 1.0 / (wi * wi);
                double f2;
                if (uv) {
                    const double dip2 = GS_DIP * yi / (3.67493e-2 * xi);
                    f2 = xi * dip2;
                } else {
                    f2 = xi * yi;
                }
                amp = f3 * f2;
            }
            sx[t] = xi;
            samp[t] = amp;
            swid[t] = wid;
            sok[t] = ok ? 1 : 0;
        }
        __syncthreads();
        const int cnt = (int)((l1 - c) < GS_BLOCK ? (l1 - c) : GS_BLOCK);
        for (int q = 0; q < cnt; ++q) {
            if (!sok[q]) continue;
            const double edif = ee - sx[q];
            r = r + samp[q] * exp(-(swid[q] * (edif * edif)));
        }
    }
    if (j > npts) return;
    r = r * rfaktor;
    double e = ee;
    if (flags & GSPEC_NM) e = 1.0e7 / (GS_EV_CM * e);
    if (flags & GSPEC_CM) e = GS_EV_CM * e;
    if (flags & GSPEC_LOG) {
        double v = log10(r);
        if (v < 0.0) v = 0.0;
        r = v;
    }
    eout[(int64_t)s * npts + (j - 1)] = e;
    rout[(int64_t)s * npts + (j - 1)] = r;
}

}  // namespace

void launch_gspec(hipStream_t st, int nspec, const int64_t *off, const double *x, const double *y,
                  const double *w, const double *emin, const double *step, int npts, int flags,
                  double rfaktor, double *eout, double *rout) {
    if (nspec <= 0 || npts <= 0) return;
    const int chunk = 65535;   // grid.y limit
    for (int s0 = 0; s0 < nspec; s0 += chunk) {
        const int ns = nspec - s0 < chunk ? nspec - s0 : chunk;
        dim3 grid((unsigned)((npts + GS_BLOCK - 1) / GS_BLOCK), (unsigned)ns);
        hipLaunchKernelGGL(gspec_kernel, grid, dim3(GS_BLOCK), 0, st, off + s0, x, y, w, emin + s0,
                           step + s0, npts, flags, rfaktor, eout + (int64_t)s0 * npts,
                           rout + (int64_t)s0 * npts);
        STDA_HIP_CHECK(hipGetLastError());
    }
}

}  // namespace stda
