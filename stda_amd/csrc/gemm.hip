// Column-major GEMM on the matrix cores for the contractions of the
// excited-state operators (the transition-density transforms of the
// matrix-free A, A+B, A-B and 1/2 B, and the contracted charges p = gamma q):
//
//   C_b = alpha op(A_b) op(B_b) + beta C_b,   b = 0 .. batch-1
//   op(A) is M x K: A[m + k lda], or A[k + m lda] when transposed
//   op(B) is K x N: B[k + n ldb], or B[n + k ldb] when transposed
//
// Three arithmetic modes share one tiling:
//   F64     FP64 operands and accumulation, v_mfma_f64_16x16x4f64;
//   F32     FP32 operands and accumulation, v_mfma_f32_16x16x4f32;
//   BF16X3  FP32 operands split x = hi + lo into two bfloat16 numbers
//           (round to nearest even; |x - hi - lo| <= 2^-18 |x|), the product
//           formed as hi hi + hi lo + lo hi by v_mfma_f32_16x16x32_bf16 with
//           FP32 accumulation (the lo lo term, ~2^-18 relative, is dropped).
//
// Work-group: 256 threads (four waves) own a BT x BT tile of C (BT = 128, or
// 64 when the 128-tiles would leave compute units idle); wave w owns rows
// BT/2 (w / 2) .. + BT/2 and columns BT/2 (w % 2) .. + BT/2, i.e. (BT/32)^2
// MFMA blocks of 16 x 16.  K advances in chunks (16 for F64/F32, 32 for BF16X3):
// each thread loads its share of the next chunk into registers (coalesced
// along the contiguous index of each operand) while the waves multiply the
// current chunk out of LDS.  Operand layouts (lane l, g = l / 16, c = l % 16):
//   16x16x4:  A[c][g], B[g][c];       16x16x32 bf16: A[c][8g..8g+7], B[8g..8g+7][c]
//   FP64 out: register i holds D[4i + g][c];   FP32 out: register i holds D[4g + i][c].
// The tiles are formed as C^T blocks (B^T A^T: the B values as the first
// operand), so that lane c holds row c of C and the stores are coalesced.
// Edges are zero-filled on load and masked on store; blockIdx.z is the batch.

#include "common.h"

#include <cmath>
#include <map>
#include <mutex>
#include <tuple>
#include <type_traits>

namespace stda {

namespace {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

constexpr int GTHREADS = 256;

// bfloat16 of x rounded to nearest even (finite inputs), as raw bits
__device__ __forceinline__ unsigned short bf16_bits(float x) {
    unsigned int u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf16_value(unsigned short b) {
    return __uint_as_float((unsigned int)b << 16);
}

template <int MODE> struct GemmTraits;
template <> struct GemmTraits<GEMM_F64> {
    using T = double;
    static constexpr int KC = 16;
};
// K chunk of a mode and tile: FP64 128-tiles take 8 (half the prefetch registers)
template <int MODE, int BT>
constexpr int chunk_of() {
    return (MODE == GEMM_F64 && BT == 128) ? 8 : GemmTraits<MODE>::KC;
}
template <> struct GemmTraits<GEMM_F32> {
    using T = float;
    static constexpr int KC = 16;
};
template <> struct GemmTraits<GEMM_BF16X3> {
    using T = float;
    static constexpr int KC = 32;
};

// LDS: F64/F32 keep [k][m] (BT + 1-element rows); BF16X3 keeps [m][k] of hi and
// lo bf16 (40-element rows, 80 bytes: each lane reads 8 consecutive k as 16 bytes).
constexpr int BK_LD = 32 + 8;

template <int MODE, int BT>
struct GemmSmem {
    typename GemmTraits<MODE>::T a[16][BT + 1], b[16][BT + 1];
};
template <int BT>
struct GemmSmem<GEMM_BF16X3, BT> {
    unsigned short ah[BT][BK_LD], al[BT][BK_LD], bh[BT][BK_LD], bl[BT][BK_LD];
};

// Epilogue of the transition-density transform D = T Co^T of the matrix-free
// operator: the stored element is G(nu, mu) = aJ gamJ(A(nu), A(mu)) D(nu, mu)
// (+ aK u(A(mu)) on the diagonal), i.e. the Coulomb-type and atom-diagonal
// parts of G formed as D leaves the matrix cores (row nu = tile row, mu = column).
template <typename T>
struct GammaEpi {
    const T *gam;
    const int *atom_of;
    const T *u;
    int na;
    T aJ, aK;
};

// One BT x BT tile of C at (m0, n0), computed by the whole work-group.
template <int MODE, bool TA, bool TB, int BT, bool EPI = false>
__device__ __forceinline__ void gemm_tile(GemmSmem<MODE, BT> &sm, int M, int N, int K,
                                          typename GemmTraits<MODE>::T alpha,
                                          const typename GemmTraits<MODE>::T *__restrict__ A,
                                          int64_t lda,
                                          const typename GemmTraits<MODE>::T *__restrict__ B,
                                          int64_t ldb, typename GemmTraits<MODE>::T beta,
                                          typename GemmTraits<MODE>::T *__restrict__ C,
                                          int64_t ldc, int m0, int n0,
                                          GammaEpi<typename GemmTraits<MODE>::T> ge = {}) {
    using T = typename GemmTraits<MODE>::T;
    constexpr int KC = chunk_of<MODE, BT>();
    constexpr int PER = BT * KC / GTHREADS;   // elements of each operand per thread and chunk
    constexpr int WT = BT / 2;                // wave tile edge
    constexpr int NB = WT / 16;               // MFMA blocks per wave tile edge
    const int t = threadIdx.x, l = t & 63, g = l >> 4, c = l & 15;
    const int w = t >> 6, wr = w >> 1, wc = w & 1;

    // (tile row, k) of this thread's p-th element of the A chunk, and (k, tile column) of B
    auto a_idx = [&](int p, int &mm, int &kk) {
        const int e = t + GTHREADS * p;
        if (TA) { kk = e % KC; mm = e / KC; } else { mm = e % BT; kk = e / BT; }
    };
    auto b_idx = [&](int p, int &kk, int &nn) {
        const int e = t + GTHREADS * p;
        if (TB) { nn = e % BT; kk = e / BT; } else { kk = e % KC; nn = e / KC; }
    };
    T ra[PER], rb[PER];
    auto load = [&](int k0) {
#pragma unroll
        for (int p = 0; p < PER; ++p) {
            int mm, kk;
            a_idx(p, mm, kk);
            const int gm = m0 + mm, gk = k0 + kk;
            ra[p] = (gm < M && gk < K) ? (TA ? A[gk + lda * gm] : A[gm + lda * gk]) : T(0);
            int kb, nn;
            b_idx(p, kb, nn);
            const int gn = n0 + nn, gkb = k0 + kb;
            rb[p] = (gn < N && gkb < K) ? (TB ? B[gn + ldb * gkb] : B[gkb + ldb * gn]) : T(0);
        }
    };
    auto store = [&]() {
#pragma unroll
        for (int p = 0; p < PER; ++p) {
            int mm, kk, kb, nn;
            a_idx(p, mm, kk);
            b_idx(p, kb, nn);
            if constexpr (MODE == GEMM_BF16X3) {
                const unsigned short ah = bf16_bits(ra[p]);
                sm.ah[mm][kk] = ah;
                sm.al[mm][kk] = bf16_bits(ra[p] - bf16_value(ah));
                const unsigned short bh = bf16_bits(rb[p]);
                sm.bh[nn][kb] = bh;
                sm.bl[nn][kb] = bf16_bits(rb[p] - bf16_value(bh));
            } else {
                sm.a[kk][mm] = ra[p];
                sm.b[kb][nn] = rb[p];
            }
        }
    };

    using Acc = typename std::conditional<MODE == GEMM_F64, v4d, v4f>::type;
    Acc acc[NB][NB];
#pragma unroll
    for (int i = 0; i < NB; ++i)
#pragma unroll
        for (int j = 0; j < NB; ++j) acc[i][j] = Acc{0, 0, 0, 0};

    load(0);
    for (int k0 = 0; k0 < K; k0 += KC) {
        __syncthreads();   // the previous chunk has been consumed
        store();
        __syncthreads();
        if (k0 + KC < K) load(k0 + KC);   // next chunk in flight during the products
        if constexpr (MODE == GEMM_BF16X3) {
            v8bf bh[NB], bl[NB];
#pragma unroll
            for (int j = 0; j < NB; ++j) {
                const int q = wc * WT + j * 16 + c;
                bh[j] = *reinterpret_cast<const v8bf *>(&sm.bh[q][8 * g]);
                bl[j] = *reinterpret_cast<const v8bf *>(&sm.bl[q][8 * g]);
            }
#pragma unroll
            for (int i = 0; i < NB; ++i) {
                const int r = wr * WT + i * 16 + c;
                const v8bf ah = *reinterpret_cast<const v8bf *>(&sm.ah[r][8 * g]);
                const v8bf al = *reinterpret_cast<const v8bf *>(&sm.al[r][8 * g]);
#pragma unroll
                for (int j = 0; j < NB; ++j) {
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh[j], al, acc[i][j], 0, 0, 0);
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bl[j], ah, acc[i][j], 0, 0, 0);
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh[j], ah, acc[i][j], 0, 0, 0);
                }
            }
        } else {
#pragma unroll
            for (int kk = 0; kk < KC; kk += 4) {
                T av[NB], bv[NB];
#pragma unroll
                for (int i = 0; i < NB; ++i) {
                    av[i] = sm.a[kk + g][wr * WT + i * 16 + c];
                    bv[i] = sm.b[kk + g][wc * WT + i * 16 + c];
                }
#pragma unroll
                for (int i = 0; i < NB; ++i)
#pragma unroll
                    for (int j = 0; j < NB; ++j) {
                        if constexpr (MODE == GEMM_F64)
                            acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(bv[j], av[i], acc[i][j], 0, 0, 0);
                        else
                            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(bv[j], av[i], acc[i][j], 0, 0, 0);
                    }
            }
        }
    }
    // epilogue: alpha acc (+ beta C), masked at the edges.  The products above
    // take B as the MFMA's first operand, so acc[i][j] holds the transposed
    // 16 x 16 block: lane c is the row of C and register e a column, and the
    // 16 lanes of a group store 16 consecutive rows of one column (coalesced
    // in column-major C)
#pragma unroll
    for (int i = 0; i < NB; ++i)
#pragma unroll
        for (int j = 0; j < NB; ++j)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int col = (MODE == GEMM_F64) ? 4 * e + g : 4 * g + e;
                const int gm = m0 + wr * WT + i * 16 + c;
                const int gn = n0 + wc * WT + j * 16 + col;
                if (gm < M && gn < N) {
                    T v = alpha * acc[i][j][e];
                    T *p = C + gm + ldc * (int64_t)gn;
                    if constexpr (EPI) {
                        const int am = ge.atom_of[gm], an = ge.atom_of[gn];
                        v = ge.aJ * ge.gam[am + (int64_t)ge.na * an] * v;
                        if (gm == gn) v += ge.aK * ge.u[an];
                    } else if (beta != T(0)) {
                        v += beta * *p;
                    }
                    *p = v;
                }
            }
}

template <int MODE, bool TA, bool TB, int BT>
__global__ void __launch_bounds__(GTHREADS, 2)   // at least two waves per SIMD
gemm_kernel(int M, int N, int K, typename GemmTraits<MODE>::T alpha,
            const typename GemmTraits<MODE>::T *__restrict__ A, int64_t lda, int64_t sA,
            const typename GemmTraits<MODE>::T *__restrict__ B, int64_t ldb, int64_t sB,
            typename GemmTraits<MODE>::T beta, typename GemmTraits<MODE>::T *__restrict__ C,
            int64_t ldc, int64_t sC) {
    __shared__ __align__(16) GemmSmem<MODE, BT> sm;
    gemm_tile<MODE, TA, TB, BT>(sm, M, N, K, alpha, A + sA * blockIdx.z, lda, B + sB * blockIdx.z,
                                ldb, beta, C + sC * blockIdx.z, ldc, (int)blockIdx.x * BT,
                                (int)blockIdx.y * BT);
}

// Grouped form: the work-groups start[p] .. start[p + 1] compute the 64 x 64
// tiles of problem p (tile row fastest, then tile column, then batch group).
// A problem's column or batch count can follow the state of a lock-step
// molecule (GemmProb::kind), so that one captured launch serves every
// iteration; tiles beyond the current count return at once.
template <int MODE, bool TA, bool TB, int BT>
__device__ __forceinline__ void grouped_tile(GemmSmem<MODE, BT> &sm, const GemmProb &P, int local) {
    using T = typename GemmTraits<MODE>::T;
    const int tm = local % P.gm, rest = local / P.gm;
    const int tn = rest % P.gn, bg = rest / P.gn;
    const int cnt = gemm_prob_count(P);
    int N = P.N, b0 = bg, b1 = bg + 1;
    if (P.cols_per_count > 0) {
        N = P.cols_per_count * cnt;
        if (bg != 0) return;
    } else {   // batch entries bg * bpw .. + bpw of this work-group, in turn
        b0 = bg * P.bpw;
        b1 = min(cnt, b0 + P.bpw);
        if (b0 >= b1) return;
    }
    if (tn * BT >= N) return;
    for (int b = b0; b < b1; ++b) {
        const T *A = static_cast<const T *>(P.A) + P.sA * (int64_t)b;
        const T *B = static_cast<const T *>(P.B) + P.sB * (int64_t)b;
        T *C = static_cast<T *>(P.C) + P.sC * (int64_t)b;
        if (b > b0) __syncthreads();   // LDS of the previous entry consumed
        if (P.epi) {
            GammaEpi<T> ge{static_cast<const T *>(P.ep_gam), P.ep_atom,
                           static_cast<const T *>(P.ep_u) + (int64_t)P.ep_na * b, P.ep_na,
                           (T)P.ep_aJ, (T)P.ep_aK};
            gemm_tile<MODE, TA, TB, BT, true>(sm, P.M, N, P.K, (T)P.alpha, A, P.lda, B, P.ldb,
                                              (T)P.beta, C, P.ldc, tm * BT, tn * BT, ge);
        } else {
            gemm_tile<MODE, TA, TB, BT>(sm, P.M, N, P.K, (T)P.alpha, A, P.lda, B, P.ldb,
                                        (T)P.beta, C, P.ldc, tm * BT, tn * BT);
        }
    }
}

// 64 x 64 tiles throughout: on the 300-600 AO transforms of the lock step the
// edge waste of 128-tiles, and their LDS, cost more than they gain (measured).
template <int MODE, bool TA, bool TB>
__global__ void __launch_bounds__(GTHREADS, 2)
gemm_grouped_kernel(const GemmProb *__restrict__ probs, const int *__restrict__ start, int nprob) {
    __shared__ __align__(16) GemmSmem<MODE, 64> sm;
    const int p = grouped_item(start, nprob, (int)blockIdx.x);
    const GemmProb &P = probs[p];
    grouped_tile<MODE, TA, TB, 64>(sm, P, (int)blockIdx.x - start[p]);
}

template <int MODE, typename T>
void launch_tile(hipStream_t s, int bt, bool ta, bool tb, int M, int N, int K, T alpha, const T *A,
                 int64_t lda, int64_t sA, const T *B, int64_t ldb, int64_t sB, T beta, T *C,
                 int64_t ldc, int64_t sC, int batch);

// Tile-size autotuner.  The first call of a shape class (mode, transposes, and
// the binary orders of M, N, K and the batch count) that is large enough to
// matter times both tile sizes on its own operands, writing into a scratch
// output, and keeps the faster one for the rest of the process; smaller
// classes, and a class while another thread tunes it, use the size rule
// (128-tiles when they still give every compute unit two work-groups).
using TileKey = std::tuple<int, int, int, int, int, int, int>;
std::mutex g_tile_mu;
std::map<TileKey, int> g_tile_choice;   // 0: being tuned

int rule_tile(int M, int N, int batch) {
    const int64_t big = (int64_t)((M + 127) / 128) * ((N + 127) / 128) * batch;
    return big >= 512 ? 128 : 64;
}

int ilog2(int64_t v) {
    int r = 0;
    while (v > 1) { v >>= 1; ++r; }
    return r;
}

template <int MODE, typename T>
int choose_tile(hipStream_t s, bool ta, bool tb, int M, int N, int K, const T *A, int64_t lda,
                int64_t sA, const T *B, int64_t ldb, int64_t sB, int64_t ldc, int64_t sC, int batch) {
    const double flops = 2.0 * M * (double)N * K * batch;
    if (flops < 2.0e9) return rule_tile(M, N, batch);
    const TileKey key{MODE, ta, tb, ilog2(M), ilog2(N), ilog2(K), ilog2(batch)};
    {
        std::lock_guard<std::mutex> lk(g_tile_mu);
        auto it = g_tile_choice.find(key);
        if (it != g_tile_choice.end()) return it->second ? it->second : rule_tile(M, N, batch);
        g_tile_choice[key] = 0;
    }
    int best = rule_tile(M, N, batch);
    void *scratch = nullptr;
    const size_t bytes = sizeof(T) * (size_t)(ldc * (int64_t)(N - 1) + M + sC * (int64_t)(batch - 1));
    hipEvent_t e[3] = {nullptr, nullptr, nullptr};
    bool ok = hipMallocAsync(&scratch, bytes, s) == hipSuccess;
    for (auto &x : e) ok = ok && hipEventCreate(&x) == hipSuccess;
    if (ok) {
        float ms[2] = {0.f, 0.f};
        const int cand[2] = {64, 128};
        for (int c = 0; c < 2 && ok; ++c) {
            // one untimed run (code load, caches), then the timed one
            launch_tile<MODE>(s, cand[c], ta, tb, M, N, K, T(1), A, lda, sA, B, ldb, sB, T(0),
                              static_cast<T *>(scratch), ldc, sC, batch);
            ok = hipEventRecord(e[0], s) == hipSuccess;
            launch_tile<MODE>(s, cand[c], ta, tb, M, N, K, T(1), A, lda, sA, B, ldb, sB, T(0),
                              static_cast<T *>(scratch), ldc, sC, batch);
            ok = ok && hipEventRecord(e[1], s) == hipSuccess &&
                 hipEventSynchronize(e[1]) == hipSuccess &&
                 hipEventElapsedTime(&ms[c], e[0], e[1]) == hipSuccess;
        }
        if (ok) best = ms[1] < ms[0] ? 128 : 64;
    }
    for (auto &x : e)
        if (x) (void)hipEventDestroy(x);
    if (scratch) (void)hipFreeAsync(scratch, s);
    (void)hipGetLastError();
    std::lock_guard<std::mutex> lk(g_tile_mu);
    g_tile_choice[key] = best;
    return best;
}

template <int MODE, typename T>
void launch_gemm_mode(hipStream_t s, bool ta, bool tb, int M, int N, int K, T alpha, const T *A,
                      int64_t lda, int64_t sA, const T *B, int64_t ldb, int64_t sB, T beta, T *C,
                      int64_t ldc, int64_t sC, int batch) {
    if (M <= 0 || N <= 0 || batch <= 0) return;
    const int bt = choose_tile<MODE>(s, ta, tb, M, N, K, A, lda, sA, B, ldb, sB, ldc, sC, batch);
    launch_tile<MODE>(s, bt, ta, tb, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC,
                      batch);
}

template <int MODE, typename T>
void launch_tile(hipStream_t s, int bt, bool ta, bool tb, int M, int N, int K, T alpha, const T *A,
                 int64_t lda, int64_t sA, const T *B, int64_t ldb, int64_t sB, T beta, T *C,
                 int64_t ldc, int64_t sC, int batch) {
    const unsigned gx = (unsigned)((M + bt - 1) / bt), gy = (unsigned)((N + bt - 1) / bt);
    if (gy > 65535u || batch > 65535)
        throw Error(9, "gemm: more than 65535 column tiles or batch entries per launch");
    const dim3 grid(gx, gy, (unsigned)batch);
#define STDA_GEMM_LAUNCH(BTV)                                                                   \
    if (!ta && !tb)                                                                             \
        hipLaunchKernelGGL((gemm_kernel<MODE, false, false, BTV>), grid, dim3(GTHREADS), 0, s, M, \
                           N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC);              \
    else if (!ta && tb)                                                                         \
        hipLaunchKernelGGL((gemm_kernel<MODE, false, true, BTV>), grid, dim3(GTHREADS), 0, s, M,  \
                           N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC);              \
    else if (ta && !tb)                                                                         \
        hipLaunchKernelGGL((gemm_kernel<MODE, true, false, BTV>), grid, dim3(GTHREADS), 0, s, M,  \
                           N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC);              \
    else                                                                                        \
        hipLaunchKernelGGL((gemm_kernel<MODE, true, true, BTV>), grid, dim3(GTHREADS), 0, s, M,   \
                           N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC);
    if (bt == 128) {
        STDA_GEMM_LAUNCH(128)
    } else {
        STDA_GEMM_LAUNCH(64)
    }
#undef STDA_GEMM_LAUNCH
    STDA_HIP_CHECK(hipGetLastError());
}

}  // namespace

void launch_gemm_f64(hipStream_t s, bool ta, bool tb, int M, int N, int K, double alpha,
                     const double *A, int64_t lda, int64_t sA, const double *B, int64_t ldb,
                     int64_t sB, double beta, double *C, int64_t ldc, int64_t sC, int batch) {
    launch_gemm_mode<GEMM_F64>(s, ta, tb, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc,
                               sC, batch);
}

void launch_gemm_grouped(hipStream_t s, int mode, bool ta, bool tb, int nprob,
                         const GemmProb *probs, const int *start, int ntiles) {
    if (ntiles <= 0 || nprob <= 0) return;
#define STDA_GROUPED(MD)                                                                        \
    if (!ta && !tb)                                                                             \
        hipLaunchKernelGGL((gemm_grouped_kernel<MD, false, false>), dim3(ntiles), dim3(GTHREADS), \
                           0, s, probs, start, nprob);                                                 \
    else if (!ta && tb)                                                                         \
        hipLaunchKernelGGL((gemm_grouped_kernel<MD, false, true>), dim3(ntiles), dim3(GTHREADS),  \
                           0, s, probs, start, nprob);                                                 \
    else if (ta && !tb)                                                                         \
        hipLaunchKernelGGL((gemm_grouped_kernel<MD, true, false>), dim3(ntiles), dim3(GTHREADS),  \
                           0, s, probs, start, nprob);                                                 \
    else                                                                                        \
        hipLaunchKernelGGL((gemm_grouped_kernel<MD, true, true>), dim3(ntiles), dim3(GTHREADS),   \
                           0, s, probs, start, nprob);
    if (mode == GEMM_F64) {
        STDA_GROUPED(GEMM_F64)
    } else if (mode == GEMM_BF16X3) {
        STDA_GROUPED(GEMM_BF16X3)
    } else {
        STDA_GROUPED(GEMM_F32)
    }
#undef STDA_GROUPED
    STDA_HIP_CHECK(hipGetLastError());
}

void launch_gemm_f32(hipStream_t s, int mode, bool ta, bool tb, int M, int N, int K, float alpha,
                     const float *A, int64_t lda, int64_t sA, const float *B, int64_t ldb,
                     int64_t sB, float beta, float *C, int64_t ldc, int64_t sC, int batch) {
    if (mode == GEMM_BF16X3)
        launch_gemm_mode<GEMM_BF16X3>(s, ta, tb, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C,
                                      ldc, sC, batch);
    else
        launch_gemm_mode<GEMM_F32>(s, ta, tb, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C,
                                   ldc, sC, batch);
}

}  // namespace stda
