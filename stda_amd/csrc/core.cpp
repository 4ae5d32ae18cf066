// Host side of the MI355X excited-state core: the setup pipeline that
// replaces stda.f:454-790, the matrix-free operators, the FP64 block
// Davidson / paired RPA Davidson / response solvers, and the C ABI.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <optional>
#include <string>
#include <thread>
#include <vector>

#include "../../include/stda_hip.h"
#include "common.h"
#include <rocsolver/rocsolver.h>

#include "devbuf.h"
#include "selsort.h"

#include <dlfcn.h>

namespace stda {

// hardness table of setrep (stda.f:2060-2167), Hartree
static const double kEta[94] = {
    0.472592880, 0.922033910, 0.174528880, 0.257007330, 0.339490860, 0.421954120, 0.504381930,
    0.586918630, 0.669313510, 0.751916070, 0.179641050, 0.221572760, 0.263485780, 0.305396450,
    0.347340140, 0.389247250, 0.431156700, 0.473082690, 0.171054690, 0.202762440, 0.210073220,
    0.217396470, 0.224710390, 0.232015010, 0.239339690, 0.246656380, 0.253982550, 0.261288630,
    0.268594760, 0.275925650, 0.307629990, 0.339315800, 0.372359850, 0.402735490, 0.434457760,
    0.466117080, 0.155850790, 0.186493240, 0.193562100, 0.200633110, 0.207705220, 0.214772540,
    0.221846140, 0.228918720, 0.235986210, 0.243056120, 0.250130180, 0.257199370, 0.287847800,
    0.318486730, 0.349124310, 0.379765930, 0.410408080, 0.441057770, 0.050193320, 0.067625700,
    0.085044450, 0.102477360, 0.119911050, 0.137327720, 0.154762970, 0.172182650, 0.189612880,
    0.207047600, 0.224467520, 0.241896450, 0.259325030, 0.276760940, 0.294182310, 0.311595870,
    0.329022740, 0.345922980, 0.363880480, 0.381305860, 0.398774760, 0.416142980, 0.433645100,
    0.451040140, 0.468489860, 0.485845500, 0.125267300, 0.142686770, 0.160116150, 0.177558890,
    0.194975570, 0.212407780, 0.072635250, 0.094221580, 0.099202950, 0.104186210, 0.142356330,
    0.163942940, 0.185519410, 0.223701390};

constexpr double kHartreeEv = 27.211385050;   // stda.f:142
constexpr double kRot = 235.7220;             // stda.f:1300
// PT2 pass 1: work-groups resident at a time (256 compute units x 2 of the
// kernel's 77 KiB LDS footprint), the unit of its split choice.
constexpr int kPt2Resident = 512;
// PT2: largest matrix of pass-1 terms kept in device memory for pass 2.
constexpr size_t kPt2StoreBytes = size_t(2) << 30;
// PT2 pass 1 in FP32: a secondary whose row sum lies within this fraction of
// thrp has its row tile evaluated again in FP64.  The couplings are FP32 dot
// products over the atoms, accumulated in order in 4-atom steps: relative
// error at most about (Na / 4) 2^-24, 3e-6 for 200 atoms; the row sums of
// 60-200-atom clusters measure at most 6e-8 against FP64.  The margin is 30
// times the bound.
constexpr double kPt2Recheck = 1.0e-4;

// Symmetric eigensolver for the small subspace problems.  The image's LAPACK
// (MKL's dsyevd) is loaded at run time by absolute path with RTLD_LOCAL, so
// that its directory never enters the library search path of the process
// (it holds an older libstdc++ than the ROCm libraries need).  Without it, a
// Householder tridiagonalisation + implicit QL (tred2/tql2 of EISPACK,
// Wilkinson & Reinsch, Handbook for Automatic Computation II) is used.
using dsyevd_t = void (*)(const char *, const char *, const int *, double *, const int *, double *,
                          double *, const int *, int *, const int *, int *);
using dgemm_t = void (*)(const char *, const char *, const int *, const int *, const int *,
                         const double *, const double *, const int *, const double *, const int *,
                         const double *, double *, const int *);
using dpotrf_t = void (*)(const char *, const int *, double *, const int *, int *);
using dtrsm_t = void (*)(const char *, const char *, const char *, const char *, const int *,
                         const int *, const double *, const double *, const int *, double *,
                         const int *);
struct HostLapack {
    dsyevd_t syevd = nullptr;
    dgemm_t gemm = nullptr;
    dpotrf_t potrf = nullptr;
    dtrsm_t trsm = nullptr;
};
static HostLapack find_lapack() {
    HostLapack L;
    const char *env = getenv("STDA_HIP_LAPACK");
    const char *cands[] = {env ? env : "", "/opt/conda/lib/libmkl_rt.so.1",
                           "/opt/conda/lib/libmkl_rt.so"};
    for (const char *path : cands) {
        if (!path || !*path) continue;
        void *h = dlopen(path, RTLD_NOW | RTLD_LOCAL);
        if (!h) continue;
        using set_t = int (*)(int);
        if (auto st = (set_t)dlsym(h, "MKL_Set_Threading_Layer")) st(1);   // sequential
        L.syevd = (dsyevd_t)dlsym(h, "dsyevd_");
        if (!L.syevd) continue;
        L.gemm = (dgemm_t)dlsym(h, "dgemm_");
        L.potrf = (dpotrf_t)dlsym(h, "dpotrf_");
        L.trsm = (dtrsm_t)dlsym(h, "dtrsm_");
        break;
    }
    return L;
}

// Looked up once per process; contexts driven from several host threads share it.
static const HostLapack &host_lapack() {
    static const HostLapack L = find_lapack();
    return L;
}
static dsyevd_t load_dsyevd() { return host_lapack().syevd; }

static void tred2_tql2(int n, std::vector<double> &a, std::vector<double> &d) {
    // a: column-major n x n symmetric on input, eigenvectors on output
    std::vector<double> e(n, 0.0);
    d.assign(n, 0.0);
    auto A = [&](int i, int j) -> double & { return a[i + (size_t)n * j]; };
    for (int i = n - 1; i > 0; --i) {
        int l = i - 1;
        double h = 0.0, scale = 0.0;
        if (l > 0) {
            for (int k = 0; k <= l; ++k) scale += std::fabs(A(i, k));
            if (scale == 0.0) {
                e[i] = A(i, l);
            } else {
                for (int k = 0; k <= l; ++k) { A(i, k) /= scale; h += A(i, k) * A(i, k); }
                double f = A(i, l);
                double g = f >= 0.0 ? -std::sqrt(h) : std::sqrt(h);
                e[i] = scale * g;
                h -= f * g;
                A(i, l) = f - g;
                f = 0.0;
                for (int j = 0; j <= l; ++j) {
                    A(j, i) = A(i, j) / h;
                    g = 0.0;
                    for (int k = 0; k <= j; ++k) g += A(j, k) * A(i, k);
                    for (int k = j + 1; k <= l; ++k) g += A(k, j) * A(i, k);
                    e[j] = g / h;
                    f += e[j] * A(i, j);
                }
                double hh = f / (h + h);
                for (int j = 0; j <= l; ++j) {
                    f = A(i, j);
                    e[j] = g = e[j] - hh * f;
                    for (int k = 0; k <= j; ++k) A(j, k) -= (f * e[k] + g * A(i, k));
                }
            }
        } else {
            e[i] = A(i, l);
        }
        d[i] = h;
    }
    d[0] = 0.0;
    e[0] = 0.0;
    for (int i = 0; i < n; ++i) {
        int l = i - 1;
        if (d[i] != 0.0) {
            for (int j = 0; j <= l; ++j) {
                double g = 0.0;
                for (int k = 0; k <= l; ++k) g += A(i, k) * A(k, j);
                for (int k = 0; k <= l; ++k) A(k, j) -= g * A(k, i);
            }
        }
        d[i] = A(i, i);
        A(i, i) = 1.0;
        for (int j = 0; j <= l; ++j) A(j, i) = A(i, j) = 0.0;
    }
    // implicit QL on the tridiagonal (d, e)
    for (int i = 1; i < n; ++i) e[i - 1] = e[i];
    e[n - 1] = 0.0;
    for (int l = 0; l < n; ++l) {
        int iter = 0, m;
        do {
            for (m = l; m < n - 1; ++m) {
                double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
                if (std::fabs(e[m]) <= 1e-15 * dd) break;
            }
            if (m != l) {
                if (iter++ == 60) throw Error(STDA_HIP_ERR_DEVICE, "tql2: no convergence");
                double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
                double r = std::hypot(g, 1.0);
                g = d[m] - d[l] + e[l] / (g + (g >= 0.0 ? std::fabs(r) : -std::fabs(r)));
                double s = 1.0, c = 1.0, p = 0.0;
                int i;
                for (i = m - 1; i >= l; --i) {
                    double f = s * e[i], b = c * e[i];
                    e[i + 1] = (r = std::hypot(f, g));
                    if (r == 0.0) { d[i + 1] -= p; e[m] = 0.0; break; }
                    s = f / r; c = g / r;
                    g = d[i + 1] - p;
                    r = (d[i] - g) * s + 2.0 * c * b;
                    d[i + 1] = g + (p = s * r);
                    g = c * r - b;
                    for (int k = 0; k < n; ++k) {
                        f = A(k, i + 1);
                        A(k, i + 1) = s * A(k, i) + c * f;
                        A(k, i) = c * A(k, i) - s * f;
                    }
                }
                if (r == 0.0 && i >= l) continue;
                d[l] -= p; e[l] = g; e[m] = 0.0;
            }
        } while (m != l);
    }
    // sort ascending
    std::vector<int> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    std::sort(idx.begin(), idx.end(), [&](int x, int y) { return d[x] < d[y]; });
    std::vector<double> a2(a.size()), d2(n);
    for (int j = 0; j < n; ++j) {
        d2[j] = d[idx[j]];
        for (int i = 0; i < n; ++i) a2[i + (size_t)n * j] = a[i + (size_t)n * idx[j]];
    }
    a.swap(a2);
    d.swap(d2);
}

void host_syev(int n, std::vector<double> &a, std::vector<double> &w) {
    w.assign(n, 0.0);
    if (n == 0) return;
    dsyevd_t fn = load_dsyevd();
    if (!fn) {
        tred2_tql2(n, a, w);
        return;
    }
    int lwork = -1, liwork = -1, info = 0, iw = 0;
    double wq = 0.0;
    fn("V", "U", &n, a.data(), &n, w.data(), &wq, &lwork, &iw, &liwork, &info);
    lwork = (int)wq + 1;
    liwork = iw + 1;
    std::vector<double> work(lwork);
    std::vector<int> iwork(liwork);
    fn("V", "U", &n, a.data(), &n, w.data(), work.data(), &lwork, iwork.data(), &liwork, &info);
    if (info != 0) throw Error(STDA_HIP_ERR_DEVICE, "dsyevd failed, info=" + std::to_string(info));
}

// C(m x n) = alpha op(A) op(B) + beta C, column-major (op = transpose when t*).
void host_gemm(bool ta, bool tb, int m, int n, int k, double alpha, const double *A, int lda,
               const double *B, int ldb, double beta, double *C, int ldc) {
    if (m <= 0 || n <= 0) return;
    if (auto fn = host_lapack().gemm) {
        fn(ta ? "T" : "N", tb ? "T" : "N", &m, &n, &k, &alpha, A, &lda, B, &ldb, &beta, C, &ldc);
        return;
    }
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < m; ++i) {
            double s = 0.0;
            for (int l = 0; l < k; ++l)
                s += (ta ? A[l + (size_t)lda * i] : A[i + (size_t)lda * l]) *
                     (tb ? B[j + (size_t)ldb * l] : B[l + (size_t)ldb * j]);
            C[i + (size_t)ldc * j] = alpha * s + (beta == 0.0 ? 0.0 : beta * C[i + (size_t)ldc * j]);
        }
}

// Lower Cholesky factor in place (the strict upper triangle is zeroed); false
// when the matrix is not positive definite.
bool host_potrf_lower(int n, std::vector<double> &a) {
    int info = 0;
    if (auto fn = host_lapack().potrf) {
        fn("L", &n, a.data(), &n, &info);
    } else {
        for (int j = 0; j < n && info == 0; ++j) {
            double d = a[j + (size_t)n * j];
            for (int l = 0; l < j; ++l) d -= a[j + (size_t)n * l] * a[j + (size_t)n * l];
            if (!(d > 0.0)) { info = j + 1; break; }
            d = std::sqrt(d);
            a[j + (size_t)n * j] = d;
            for (int i = j + 1; i < n; ++i) {
                double s = a[i + (size_t)n * j];
                for (int l = 0; l < j; ++l) s -= a[i + (size_t)n * l] * a[j + (size_t)n * l];
                a[i + (size_t)n * j] = s / d;
            }
        }
    }
    if (info != 0) return false;
    for (int j = 1; j < n; ++j)
        for (int i = 0; i < j; ++i) a[i + (size_t)n * j] = 0.0;
    return true;
}

// B(n x k) <- L^-T B for the lower triangular L(n x n).
void host_solve_lt(int n, int k, const std::vector<double> &L, double *B, int ldb) {
    if (n <= 0 || k <= 0) return;
    if (auto fn = host_lapack().trsm) {
        const double one = 1.0;
        fn("L", "L", "T", "N", &n, &k, &one, L.data(), &n, B, &ldb);
        return;
    }
    for (int c = 0; c < k; ++c) {
        double *b = B + (size_t)ldb * c;
        for (int i = n - 1; i >= 0; --i) {
            double s = b[i];
            for (int l = i + 1; l < n; ++l) s -= L[l + (size_t)n * i] * b[l];
            b[i] = s / L[i + (size_t)n * i];
        }
    }
}

struct Timer {
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    double ms() const {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
            .count();
    }
};

}  // namespace stda

using namespace stda;

struct stda_hip_ctx {
    // Owns the stream and the BLAS handle; declared first so that it is
    // destroyed last, after every buffer below has been freed in the order
    // of this stream.
    struct Owner {
        hipStream_t stream = nullptr;
        rocblas_handle blas = nullptr;
        ~Owner() {
            if (blas) rocblas_destroy_handle(blas);
            if (stream) {
                block_cache_drain(stream);
                (void)hipStreamSynchronize(stream);
                (void)hipStreamDestroy(stream);
            }
        }
    } owner;
    int device = 0;
    hipStream_t stream = nullptr;
    rocblas_handle blas = nullptr;
    std::string last_error;
    double timings[6] = {0, 0, 0, 0, 0, 0};

    // system
    int ncent = 0, nao = 0, no = 0, nv = 0;
    int64_t npair = 0;
    double ak = 2.0, ax = 0.0, thr = 0.0, thrp = 1e-4, fthr = 0.0;
    bool small_ax = false;
    bool have_moments = false;
    std::vector<double> moments_h;   // 9 * npk (host copy, for window-mode helpers)
    int64_t npk = 0;

    // device-resident intermediates
    DBuf<double> ct;          // (nao, moci) Loewdin MOs, AOs sorted by atom (column-major)
    DBuf<int> atom_ptr, atom_of;
    DBuf<double> gamj, gamk, eta, xyz;
    DBuf<double> qdiag, jdiag, eps;
    DBuf<double> qia, pia, qij, pij, qab;
    DBuf<double> de, kdiag;
    DBuf<unsigned char> flags;
    DBuf<double> moments;

    // CSF space
    int64_t nci = 0, nprim = 0, nsec = 0, npt = 0;
    std::vector<int32_t> pair_h;   // CSF -> pair index
    std::vector<double> ed_h;
    DBuf<int32_t> pair;
    DBuf<double> ed, corr, corr_small;   // diagonal corrections of the AO-path operators
    DBuf<double> mom;                    // (nci, 9) CSF moments

    // matrix-free work buffers
    DBuf<double> xfull, tbuf, dbuf, gbuf, wbuf, rbuf, ubuf;
    // mixed precision: FP32 copies of the operator data and FP32 work buffers
    int precision = 0;    // 0 FP64, 1 FP32 operator + FP64 Davidson numerics and refinement,
                          // 2 split-bfloat16 operator, then FP32, then the FP64 refinement
    int gemm32 = GEMM_F32;   // arithmetic of the FP32 operator's GEMMs in the current phase
    DBuf<float> ct32, gamj32, gamk32;
    DBuf<float> xfull32, tbuf32, dbuf32, gbuf32, wbuf32, rbuf32, ubuf32;
    int nb_max32 = 0;
    bool have32 = false;
    DBuf<double> tnbuf;   // split-K partial products of dgemm_tn
    DBuf<double> tn2buf;  // the two blocks of dgemm_tn2 on its split-K path
    DBuf<double> normbuf; // slice partials of column norms
    // PT2 selection: the kept pass-1 terms and the FP32 operand copies, kept
    // (and grown) from one molecule to the next instead of allocated per call
    DBuf<double> pt2_terms;
    DBuf<float> pt2_pk32, pt2_q32, pt2_pj32, pt2_qab32;
    DBuf<double> pt2_pk64, pt2_q64, pt2_pj64, pt2_qab64;   // odd atom counts: padded to even
    int nb_max = 0;
    int nb_chunk = 0;     // vectors per operator block (from the free memory at first use)
    bool xfull_clean = false;

    bool ready = false;
    bool loaded = false;
    double alphak = 0.0, betaj = 0.0;
    int kshift = 0;
    double ks_max = 0.5, ks_width = 0.1, ks_steep = 4.0;
    // selection variants (stda_hip_set_selection): dual thresholds, no PT2
    double dual_thr2_ev = 0.0;
    std::vector<int32_t> dual_atoms;   // 1-based, as listed
    bool pt2_off = false;
    std::vector<int> aptr_h;           // AO ranges per atom (AOs sorted by atom)

    // per-kernel statistics: HIP-event time on the context's stream and the
    // algorithmic work (flops or bytes) of the launches
    struct Stat { double ms = 0, launches = 0, work = 0; };
    std::vector<std::pair<std::string, Stat>> stats;
    // the same summed over the molecules set up since the last reset (batches)
    std::vector<std::pair<std::string, Stat>> stats_total;
    Stat &stat(const std::string &name) {
        for (auto &kv : stats)
            if (kv.first == name) return kv.second;
        stats.emplace_back(name, Stat{});
        return stats.back().second;
    }
    void fold_stats() {
        for (auto &kv : stats) {
            Stat *t = nullptr;
            for (auto &tv : stats_total)
                if (tv.first == kv.first) t = &tv.second;
            if (!t) {
                stats_total.emplace_back(kv.first, Stat{});
                t = &stats_total.back().second;
            }
            t->ms += kv.second.ms;
            t->launches += kv.second.launches;
            t->work += kv.second.work;
        }
        stats.clear();
    }

    // event pairs of timed launches, resolved lazily so that timing never
    // synchronises the stream
    struct PendingTimer {
        std::string name;
        double work;
        hipEvent_t a, b;
        double weight;   // launches this sample stands for
    };
    // tall-skinny products issued; one in eight is timed, picked by a hash of
    // the count so that no fixed position of a periodic call sequence is missed
    unsigned long long ts_count = 0;
    bool sample_ts() {
        unsigned long long z = (ts_count++) + 0x9e3779b97f4a7c15ULL;
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
        return ((z ^ (z >> 31)) & 7) == 0;
    }
    std::vector<PendingTimer> pending;
    std::vector<hipEvent_t> event_pool;   // recycled timing events
    hipEvent_t take_event() {
        if (!event_pool.empty()) {
            hipEvent_t e = event_pool.back();
            event_pool.pop_back();
            return e;
        }
        hipEvent_t e = nullptr;
        STDA_HIP_CHECK(hipEventCreate(&e));
        return e;
    }
    void flush_timers() {
        for (auto &p : pending) {
            float ms = 0.f;
            if (hipEventSynchronize(p.b) == hipSuccess &&
                hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
                auto &st = stat(p.name);
                st.ms += p.weight * ms;
                st.launches += p.weight;
                st.work += p.weight * p.work;
            }
            event_pool.push_back(p.a);
            event_pool.push_back(p.b);
        }
        pending.clear();
    }
    // pinned host staging of the solvers' small per-iteration transfers (grown
    // on demand; every region is reused only after a stream synchronisation
    // that follows its last transfer)
    double *pin = nullptr;
    size_t pin_n = 0;
    double *pinned(size_t n) {
        if (n > pin_n) {
            STDA_HIP_CHECK(hipStreamSynchronize(stream));
            if (pin) (void)hipHostFree(pin);
            pin = nullptr;
            pin_n = 0;
            STDA_HIP_CHECK(hipHostMalloc(reinterpret_cast<void **>(&pin), n * sizeof(double), 0));
            pin_n = n;
        }
        return pin;
    }
    ~stda_hip_ctx() {
        flush_timers();
        for (hipEvent_t e : event_pool) (void)hipEventDestroy(e);
        if (pin) (void)hipHostFree(pin);
    }

    void sync() { STDA_HIP_CHECK(hipStreamSynchronize(stream)); }
};

namespace {
// Adds host wall time (ms) since t0 to the named statistic.
struct HostSpan {
    stda_hip_ctx *c;
    const char *name;
    Timer t;
    HostSpan(stda_hip_ctx *ctx, const char *n) : c(ctx), name(n) {}
    ~HostSpan() {
        auto &st = c->stat(name);
        st.ms += t.ms();
        st.launches += 1;
    }
};

// Records the HIP-event duration of the work enqueued on the context's stream
// between construction and finish().
struct EventTimer {
    stda_hip_ctx *c;
    hipEvent_t a = nullptr, b = nullptr;
    explicit EventTimer(stda_hip_ctx *ctx) : c(ctx) {
        a = c->take_event();
        b = c->take_event();
        STDA_HIP_CHECK(hipEventRecord(a, c->stream));
    }
    // weight: the number of launches this one stands for (sampled timing)
    void finish(const char *name, double work, double weight = 1.0) {
        STDA_HIP_CHECK(hipEventRecord(b, c->stream));
        c->pending.push_back({name, work, a, b, weight});
        a = b = nullptr;
        if (c->pending.size() > 512) c->flush_timers();
    }
    ~EventTimer() {
        if (a) c->event_pool.push_back(a);
        if (b) c->event_pool.push_back(b);
    }
};
}  // namespace

namespace {

// ------------------------------------------------------------------------------------------
// matrix-free operators
enum Op { OP_A = 0, OP_APB = 1, OP_AMB = 2, OP_BHALF = 3 };

// C (m x p, column-major, leading dimension m) = A^T B with A (n x m, lda),
// B (n x p, ldb).  For the tall-skinny products of the solvers (n up to
// millions, m and p at most a few hundred) the reduction dimension is split
// into up to 256 slices computed as one batched GEMM, then the slices are
// summed in a fixed order: all CUs work on the product and the result is
// deterministic.
void dgemm_tn(stda_hip_ctx *c, int64_t m, int64_t p, int64_t n, const double *A, int64_t lda,
              const double *B, int64_t ldb, double *C) {
    if (m == 0 || p == 0) return;
    if (m <= TS_MMAX && p <= TS_PMAX) {   // the common shapes: hand-written tall-skinny kernel
        c->tnbuf.ensure((size_t)tsgemm_tn_slices(n, (int)m, (int)p) * m * p);
        std::optional<EventTimer> ev;
        if (c->sample_ts()) ev.emplace(c);
        launch_tsgemm_tn_partial(c->stream, n, (int)m, (int)p, A, lda, B, ldb, c->tnbuf.get());
        if (ev) ev->finish("tsgemm_tn", 8.0 * (double)n * (double)(m + p), 8.0);   // inputs once
        launch_tsgemm_tn_reduce(c->stream, n, (int)m, (int)p, c->tnbuf.get(), C);
        return;
    }
    const int64_t KC_MIN = 2048;
    if (n < 4 * KC_MIN) {
        dgemm(c->blas, true, false, m, p, n, 1.0, A, lda, B, ldb, 0.0, C, m);
        return;
    }
    int64_t kc = std::max<int64_t>(KC_MIN, (n + 255) / 256);
    int64_t nch = n / kc, rem = n - nch * kc;
    int64_t slots = nch + (rem > 0 ? 1 : 0);
    c->tnbuf.ensure((size_t)(slots * m * p));
    dgemm_batched(c->blas, true, false, m, p, kc, 1.0, A, lda, kc, B, ldb, kc, 0.0,
                  c->tnbuf.get(), m, m * p, nch);
    if (rem > 0)
        dgemm(c->blas, true, false, m, p, rem, 1.0, A + nch * kc, lda, B + nch * kc, ldb, 0.0,
              c->tnbuf.get() + nch * m * p, m);
    launch_sum_slices(c->stream, m * p, (int)slots, c->tnbuf.get(), C);
}

// C ((m1 + m2) x p, ld m1 + m2) = [A1 A2]^T B for two column blocks with the
// same leading dimension (one launch pair on the hand-written path)
void dgemm_tn2(stda_hip_ctx *c, int64_t m1, const double *A1, int64_t m2, const double *A2,
               int64_t p, int64_t n, int64_t lda, const double *B, int64_t ldb, double *C) {
    const int64_t m = m1 + m2;
    if (m == 0 || p == 0) return;
    if (m <= TS_MMAX && p <= TS_PMAX) {
        c->tnbuf.ensure((size_t)tsgemm_tn_slices(n, (int)m, (int)p) * m * p);
        std::optional<EventTimer> ev;
        if (c->sample_ts()) ev.emplace(c);
        launch_tsgemm_tn2_partial(c->stream, n, (int)m1, A1, (int)m2, A2, (int)p, lda, B, ldb,
                                  c->tnbuf.get());
        if (ev) ev->finish("tsgemm_tn", 8.0 * (double)n * (double)(m + p), 8.0);
        launch_tsgemm_tn_reduce(c->stream, n, (int)m, (int)p, c->tnbuf.get(), C);
        return;
    }
    // larger shapes: the split-K path of dgemm_tn (a plain GEMM would run the
    // whole reduction over n in a few work-groups)
    if (A2 == A1 + lda * m1) {   // one contiguous block of columns
        dgemm_tn(c, m, p, n, A1, lda, B, ldb, C);
        return;
    }
    c->tn2buf.ensure((size_t)m * p);
    double *t = c->tn2buf.get();
    dgemm_tn(c, m1, p, n, A1, lda, B, ldb, t);
    dgemm_tn(c, m2, p, n, A2, lda, B, ldb, t + m1 * p);
    if (m1 > 0)
        STDA_HIP_CHECK(hipMemcpy2DAsync(C, m * 8, t, m1 * 8, m1 * 8, p, hipMemcpyDeviceToDevice,
                                        c->stream));
    if (m2 > 0)
        STDA_HIP_CHECK(hipMemcpy2DAsync(C + m1, m * 8, t + m1 * p, m2 * 8, m2 * 8, p,
                                        hipMemcpyDeviceToDevice, c->stream));
}

// out (n x p) = alpha A C + beta out (A n x m, C m x p) for the tall-skinny
// products of the solvers: hand-written kernel for p <= 32, rocBLAS beyond
void gemm_nn(stda_hip_ctx *c, int64_t n, int64_t p, int64_t m, double alpha, const double *A,
             int64_t lda, const double *C, int64_t ldc, double beta, double *out, int64_t ldo) {
    if (n == 0 || p == 0) return;
    if (p <= 32) {
        std::optional<EventTimer> ev;
        if (c->sample_ts()) ev.emplace(c);
        launch_tsgemm_nn(c->stream, n, (int)m, (int)p, alpha, A, lda, C, ldc, beta, out, ldo);
        // A and out once each (out twice when accumulated), C once
        if (ev)
            ev->finish("tsgemm_nn", 8.0 * ((double)n * (double)(m + (beta != 0.0 ? 2 * p : p)) +
                                           (double)m * (double)p), 8.0);
    } else {
        dgemm(c->blas, false, false, n, p, m, alpha, A, lda, C, ldc, beta, out, ldo);
    }
}

// out0 = A0 C and out1 = A1 C (beta 0) for two tall-skinny products sharing C
void gemm_nn2(stda_hip_ctx *c, int64_t n, int64_t p, int64_t m, const double *A0, const double *A1,
              int64_t lda, const double *C, int64_t ldc, double *out0, double *out1, int64_t ldo) {
    if (n == 0 || p == 0) return;
    if (p <= 32) {
        std::optional<EventTimer> ev;
        if (c->sample_ts()) ev.emplace(c);
        launch_tsgemm_nn2(c->stream, n, (int)m, (int)p, 1.0, A0, A1, lda, C, ldc, 0.0, out0, out1, ldo);
        if (ev)
            ev->finish("tsgemm_nn", 2.0 * 8.0 * ((double)n * (double)(m + p) + (double)m * (double)p),
                       8.0);
    } else {
        dgemm(c->blas, false, false, n, p, m, 1.0, A0, lda, C, ldc, 0.0, out0, ldo);
        dgemm(c->blas, false, false, n, p, m, 1.0, A1, lda, C, ldc, 0.0, out1, ldo);
    }
}

// column 2-norms of R (n x k) into norms (device), deterministic
void col_norms(stda_hip_ctx *c, int64_t n, int k, const double *R, double *norms) {
    int nsl = (int)std::min<int64_t>(256, std::max<int64_t>(1, n / 8192));
    c->normbuf.ensure((size_t)nsl * std::max(k, 1));
    launch_col_norms_split(c->stream, n, k, R, c->normbuf.get(), nsl, norms);
}

// R = AX - X diag(theta) and its column norms (device), deterministic
void residual_norms(stda_hip_ctx *c, int64_t n, int k, const double *AX, const double *X,
                    const double *theta, double *R, double *norms) {
    int nsl = (int)std::min<int64_t>(256, std::max<int64_t>(1, n / 8192));
    c->normbuf.ensure((size_t)nsl * std::max(k, 1));
    launch_residual_norms(c->stream, n, k, AX, X, theta, R, c->normbuf.get(), nsl, norms);
}

void ensure_work(stda_hip_ctx *c, int nb) {
    if (nb <= c->nb_max) return;
    size_t nao = c->nao, no = c->no;
    c->xfull.resize((size_t)c->npair * nb);
    c->xfull.zero(c->stream);
    c->tbuf.resize(nao * no * nb);
    c->dbuf.resize(nao * nao * nb);
    c->gbuf.resize(nao * nao * nb);
    c->wbuf.resize(nao * no * nb);
    c->rbuf.resize((size_t)c->npair * nb);
    c->ubuf.resize((size_t)c->ncent * nb);
    c->nb_max = nb;
}

int chunk_vectors(stda_hip_ctx *c) {
    if (c->nb_chunk > 0) return c->nb_chunk;
    size_t per = sizeof(double) * ((size_t)2 * c->npair + (size_t)2 * c->nao * c->no +
                                   (size_t)2 * c->nao * c->nao + c->ncent);
    size_t budget = std::min<size_t>(device_available_bytes() / 2, (size_t)24 << 30);
    int nb = (int)std::max<size_t>(1, budget / per);
    c->nb_chunk = std::min(nb, 64);
    return c->nb_chunk;
}

void ensure_fp32(stda_hip_ctx *c) {
    if (c->have32) return;
    const size_t moci = (size_t)c->no + c->nv, na = c->ncent;
    c->ct32.resize((size_t)c->nao * moci);
    c->gamj32.resize(na * na);
    c->gamk32.resize(na * na);
    launch_to_f32(c->stream, c->ct.get(), (int64_t)c->nao * moci, c->ct32.get());
    launch_to_f32(c->stream, c->gamj.get(), (int64_t)(na * na), c->gamj32.get());
    launch_to_f32(c->stream, c->gamk.get(), (int64_t)(na * na), c->gamk32.get());
    c->have32 = true;
}

void ensure_work32(stda_hip_ctx *c, int nb) {
    if (nb <= c->nb_max32) return;
    size_t nao = c->nao, no = c->no;
    c->xfull32.resize((size_t)c->npair * nb);
    c->xfull32.zero(c->stream);
    c->tbuf32.resize(nao * no * nb);
    c->dbuf32.resize(nao * nao * nb);
    c->gbuf32.resize(nao * nao * nb);
    c->wbuf32.resize(nao * no * nb);
    c->rbuf32.resize((size_t)c->npair * nb);
    c->ubuf32.resize((size_t)c->ncent * nb);
    c->nb_max32 = nb;
}

// The same operators with the contractions in FP32 (FP32 Loewdin MOs, gamma
// and GEMMs, FP32 accumulation); the diagonal correction term stays FP64.
void apply_ops_f32(stda_hip_ctx *c, int op1, int op2, int nvec, const double *V, int64_t ldv,
                   double *out1, int64_t ld1, double *out2, int64_t ld2) {
    const int64_t n = c->nci;
    hipStream_t s = c->stream;
    const int nao = c->nao, no = c->no, nv = c->nv, na = c->ncent;
    ensure_fp32(c);
    const float *Co = c->ct32.get();
    const float *Cv = c->ct32.get() + (size_t)nao * no;
    const int nbmax = std::min(128, 2 * chunk_vectors(c));
    for (int v0 = 0; v0 < nvec; v0 += nbmax) {
        int nb = std::min(nbmax, nvec - v0);
        ensure_work32(c, nb);
        const double *Vb = V + ldv * v0;
        EventTimer evm(c);
        launch_scatter_pairs_f32(s, Vb, ldv, c->pair.get(), n, nb, c->xfull32.get(), c->npair);
        sgemm(c->blas, false, false, nao, (int64_t)no * nb, nv, 1.0f, Cv, nao, c->xfull32.get(), nv,
              0.0f, c->tbuf32.get(), nao, c->gemm32);
        sgemm_batched(c->blas, false, true, nao, nao, no, 1.0f, c->tbuf32.get(), nao,
                      (int64_t)nao * no, Co, nao, 0, 0.0f, c->dbuf32.get(), nao, (int64_t)nao * nao,
                      nb, c->gemm32);
        bool need_u = false;
        int nops = 0;
        for (int pass = 0; pass < 2; ++pass) {
            int op = pass == 0 ? op1 : op2;
            if (op < 0) continue;
            double *out = pass == 0 ? out1 : out2;
            int64_t ld = pass == 0 ? ld1 : ld2;
            float aJ = -1.0f, aX = 0.0f, aK = 0.0f;
            const double *corr = c->corr.get();
            switch (op) {
                case 0: aK = (float)c->ak; break;
                case 1:
                    if (c->small_ax) { aJ = 0.0f; aK = (float)(2.0 * c->ak); corr = c->corr_small.get(); }
                    else { aX = (float)-c->ax; aK = (float)(2.0 * c->ak); }
                    break;
                case 2:
                    if (c->small_ax) {
                        launch_axpby_diag(s, n, nb, c->corr_small.get(), Vb, ldv, out + ld * v0, ld);
                        continue;
                    }
                    aX = (float)c->ax;
                    break;
                case 3:
                    aJ = 0.0f; aX = (float)(-0.5 * c->ax); aK = (float)(0.5 * c->ak); corr = nullptr;
                    break;
            }
            ++nops;
            if (aK != 0.0f && !need_u) {
                launch_atom_diag_f32(s, nb, nao, na, c->dbuf32.get(), (int64_t)nao * nao,
                                     c->atom_ptr.get(), c->gamk32.get(), c->ubuf32.get());
                need_u = true;
            }
            launch_gbuild_f32(s, nb, nao, na, c->dbuf32.get(), (int64_t)nao * nao, c->atom_of.get(),
                              c->gamj32.get(), c->gamk32.get(), c->ubuf32.get(), aJ, aX, aK,
                              c->gbuf32.get(), (int64_t)nao * nao);
            sgemm_batched(c->blas, false, false, nao, no, nao, 1.0f, c->gbuf32.get(), nao,
                          (int64_t)nao * nao, Co, nao, 0, 0.0f, c->wbuf32.get(), nao,
                          (int64_t)nao * no, nb, c->gemm32);
            sgemm(c->blas, true, false, nv, (int64_t)no * nb, nao, 1.0f, Cv, nao, c->wbuf32.get(),
                  nao, 0.0f, c->rbuf32.get(), nv, c->gemm32);
            launch_gather_pairs_f32(s, c->rbuf32.get(), c->npair, c->pair.get(), n, nb, corr, Vb,
                                    ldv, out + ld * v0, ld);
        }
        double per = 2.0 * nao * (double)no * nv + 2.0 * (double)nao * nao * no;
        evm.finish("ao_matvec_f32", nb * (per + nops * per));
    }
}

// out(I, v) = [op V](I, v) for nvec column vectors (n = nci rows).
// Two operators sharing one transition density can be requested at once
// (op2 >= 0 writes out2), which is how the RPA solver gets (A+B)V and (A-B)V.
void apply_ops(stda_hip_ctx *c, int op1, int op2, int nvec, const double *V, int64_t ldv,
               double *out1, int64_t ld1, double *out2, int64_t ld2, bool fp32 = false) {
    if (fp32) {
        apply_ops_f32(c, op1, op2, nvec, V, ldv, out1, ld1, out2, ld2);
        return;
    }
    const int64_t n = c->nci;
    hipStream_t s = c->stream;
    const int nao = c->nao, no = c->no, nv = c->nv, na = c->ncent;
    const double *Co = c->ct.get();
    const double *Cv = c->ct.get() + (size_t)nao * no;
    const int nbmax = chunk_vectors(c);
    for (int v0 = 0; v0 < nvec; v0 += nbmax) {
        int nb = std::min(nbmax, nvec - v0);
        ensure_work(c, nb);
        const double *Vb = V + ldv * v0;
        EventTimer evm(c);
        // embed the CSF vectors in the full occupied x virtual grid
        launch_scatter_pairs(s, Vb, ldv, c->pair.get(), n, nb, c->xfull.get(), c->npair);
        // T = Cv X^T (nao x no) for all vectors at once: X^T is (nv, no) column-major
        dgemm(c->blas, false, false, nao, (int64_t)no * nb, nv, 1.0, Cv, nao, c->xfull.get(), nv,
              0.0, c->tbuf.get(), nao);
        // Dt = T Co^T (nao x nao) per vector
        dgemm_batched(c->blas, false, true, nao, nao, no, 1.0, c->tbuf.get(), nao,
                      (int64_t)nao * no, Co, nao, 0, 0.0, c->dbuf.get(), nao, (int64_t)nao * nao,
                      nb);
        bool need_u = false;
        for (int pass = 0; pass < 2; ++pass) {
            int op = pass == 0 ? op1 : op2;
            if (op < 0) continue;
            double *out = pass == 0 ? out1 : out2;
            int64_t ld = pass == 0 ? ld1 : ld2;
            double aJ = -1.0, aX = 0.0, aK = 0.0;
            const double *corr = c->corr.get();
            switch (op) {
                case OP_A: aK = c->ak; break;
                case OP_APB:
                    if (c->small_ax) { aJ = 0.0; aK = 2.0 * c->ak; corr = c->corr_small.get(); }
                    else { aX = -c->ax; aK = 2.0 * c->ak; }
                    break;
                case OP_AMB:
                    if (c->small_ax) {   // A-B diagonal (stda.f:1893-1924): out = corr_small o V
                        launch_axpby_diag(s, n, nb, c->corr_small.get(), Vb, ldv, out + ld * v0, ld);
                        continue;
                    }
                    aX = c->ax;
                    break;
                case OP_BHALF:
                    aJ = 0.0; aX = -0.5 * c->ax; aK = 0.5 * c->ak; corr = nullptr;
                    break;
            }
            if (aK != 0.0 && !need_u) {
                launch_atom_diag(s, nb, nao, na, c->dbuf.get(), (int64_t)nao * nao,
                                 c->atom_ptr.get(), c->gamk.get(), c->ubuf.get());
                need_u = true;
            }
            launch_gbuild(s, nb, nao, na, c->dbuf.get(), (int64_t)nao * nao, c->atom_of.get(),
                          c->gamj.get(), c->gamk.get(), c->ubuf.get(), aJ, aX, aK, c->gbuf.get(),
                          (int64_t)nao * nao);
            // W = Gt Co (nao x no) per vector, then R = Cv^T W (nv x no)
            dgemm_batched(c->blas, false, false, nao, no, nao, 1.0, c->gbuf.get(), nao,
                          (int64_t)nao * nao, Co, nao, 0, 0.0, c->wbuf.get(), nao,
                          (int64_t)nao * no, nb);
            dgemm(c->blas, true, false, nv, (int64_t)no * nb, nao, 1.0, Cv, nao, c->wbuf.get(),
                  nao, 0.0, c->rbuf.get(), nv);
            launch_gather_pairs(s, c->rbuf.get(), c->npair, c->pair.get(), n, nb, corr, Vb, ldv,
                                out + ld * v0, ld);
        }
        // algorithmic flops of the four transforms per vector (and per extra operator)
        double per = 2.0 * nao * (double)no * nv + 2.0 * (double)nao * nao * no;
        int nops = (op1 >= 0) + (op2 >= 0 && !(op2 == OP_AMB && c->small_ax));
        evm.finish("ao_matvec", nb * (per + nops * per));
    }
}

// ------------------------------------------------------------------------------------------
// small dense helpers (host)
std::vector<double> sym_sqrt(const std::vector<double> &A, int m, bool inverse, double *minev) {
    std::vector<double> U(A), w;
    host_syev(m, U, w);
    if (minev) *minev = m ? w[0] : 0.0;
    std::vector<double> R((size_t)m * m, 0.0);
    for (int k = 0; k < m; ++k) {
        double f = w[k] > 0 ? std::sqrt(w[k]) : 0.0;
        if (inverse) f = f > 0 ? 1.0 / f : 0.0;
        for (int j = 0; j < m; ++j) {
            double ujk = U[j + (size_t)m * k] * f;
            for (int i = 0; i < m; ++i) R[i + (size_t)m * j] += U[i + (size_t)m * k] * ujk;
        }
    }
    return R;
}

std::vector<double> matmul(const std::vector<double> &A, const std::vector<double> &B, int m,
                           int k, int n) {   // (m x k)(k x n), column-major
    std::vector<double> C((size_t)m * n, 0.0);
    for (int j = 0; j < n; ++j)
        for (int l = 0; l < k; ++l) {
            double b = B[l + (size_t)k * j];
            if (b == 0.0) continue;
            for (int i = 0; i < m; ++i) C[i + (size_t)m * j] += A[i + (size_t)m * l] * b;
        }
    return C;
}

// Orthonormalise the columns of T (n x nt, device) against V (n x m, device,
// orthonormal) and among themselves; returns the number of columns kept,
// packed at the front of T.
int orthonormalize(stda_hip_ctx *c, const double *V, int m, double *T, int nt, DBuf<double> &tmp) {
    const int64_t n = c->nci;
    if (nt == 0) return 0;
    tmp.ensure((size_t)std::max(m, 1) * nt + (size_t)nt * nt + (size_t)n * nt);
    double *C = tmp.get();
    // unit columns first, so that the drop threshold below measures how much
    // of each direction is new, not how small the residual it came from is
    {
        DBuf<double> nd(nt);
        col_norms(c, n, nt, T, nd.get());
        launch_inv_scale_cols(c->stream, n, nt, T, nd.get());
    }
    for (int pass = 0; pass < 2 && m > 0; ++pass) {
        dgemm_tn(c, m, nt, n, V, n, T, n, C);
        gemm_nn(c, n, nt, m, -1.0, V, n, C, m, 1.0, T, n);
    }
    // symmetric orthonormalisation within T: S = T^T T = U s U^T, T <- T U s^-1/2
    double *S = tmp.get() + (size_t)std::max(m, 1) * nt;
    double *T2 = S + (size_t)nt * nt;
    dgemm_tn(c, nt, nt, n, T, n, T, n, S);
    std::vector<double> Sh((size_t)nt * nt), w;
    STDA_HIP_CHECK(hipMemcpyAsync(Sh.data(), S, Sh.size() * 8, hipMemcpyDeviceToHost, c->stream));
    c->sync();
    host_syev(nt, Sh, w);
    double wmax = nt ? std::max(w[nt - 1], 0.0) : 0.0;
    std::vector<double> Z;
    int kept = 0;
    for (int k = nt - 1; k >= 0; --k) {
        if (w[k] <= 1e-14 * std::max(1.0, wmax) || w[k] <= 1e-20) continue;
        double f = 1.0 / std::sqrt(w[k]);
        for (int i = 0; i < nt; ++i) Z.push_back(Sh[i + (size_t)nt * k] * f);
        ++kept;
    }
    if (kept == 0) return 0;
    DBuf<double> Zd;
    Zd.upload(Z.data(), Z.size(), c->stream);
    gemm_nn(c, n, kept, nt, 1.0, T, n, Zd.get(), nt, 0.0, T2, n);
    // A direction kept above with little new norm (Gram eigenvalue near the
    // threshold) is mostly the rounding error of the projections, scaled up:
    // it is far from orthogonal to V.  Project twice more and orthonormalise
    // the block again, dropping the directions that lost more than ~30 % of
    // their norm to V (Gram eigenvalue below 0.5).  With one projection and
    // a plain renormalisation instead, V drifted from orthonormality
    // (off-diagonal 1e-4 after 20 iterations) when the new directions were
    // nearly dependent, e.g. response residuals at close frequencies, and
    // the response solver stalled near 1e-6.
    for (int pass = 0; pass < 2 && m > 0; ++pass) {
        dgemm_tn(c, m, kept, n, V, n, T2, n, C);
        gemm_nn(c, n, kept, m, -1.0, V, n, C, m, 1.0, T2, n);
    }
    dgemm_tn(c, kept, kept, n, T2, n, T2, n, S);
    std::vector<double> S2((size_t)kept * kept), w2;
    STDA_HIP_CHECK(hipMemcpyAsync(S2.data(), S, S2.size() * 8, hipMemcpyDeviceToHost, c->stream));
    c->sync();
    host_syev(kept, S2, w2);
    std::vector<double> Z2;
    int kept2 = 0;
    for (int k = kept - 1; k >= 0; --k) {
        if (!(w2[k] > 0.5)) continue;
        const double f = 1.0 / std::sqrt(w2[k]);
        for (int i = 0; i < kept; ++i) Z2.push_back(S2[i + (size_t)kept * k] * f);
        ++kept2;
    }
    if (kept2 == 0) return 0;
    Zd.upload(Z2.data(), Z2.size(), c->stream);
    gemm_nn(c, n, kept2, kept, 1.0, T2, n, Zd.get(), kept, 0.0, T, n);
    return kept2;
}

// The same for a block of nt <= 64 new directions T (n x nt), without a host
// round trip: classical Gram-Schmidt against V twice, Cholesky QR, one more
// projection against V, Cholesky QR again.  Each projection's product [V T]^T T
// also carries the Gram matrix of the projected block (T^T T - C^T C), so the
// Cholesky factors need no product of their own.  Dependent directions
// become zero columns marked dead (alive[j] = 0.0, device): a direction is
// dropped when less than 1e-7 of its norm is new.  The third projection is
// left out on the device when the first factor shows that it cannot matter
// (a flag from chol_inv_kernel: the inverse factor amplifies what the first
// two projections left along V by at most 1e3; the V rows of its product then
// come out zero and its update returns at once).  T2: n x nt scratch; work:
// 3 (m + nt) nt + nt^2 + 1 doubles.
void orthonormalize_dev(stda_hip_ctx *c, const double *V, int m, double *T, int nt, double *T2,
                        double *work, double *alive) {
    const int64_t n = c->nci;
    if (nt == 0) return;
    const int ld = m + nt;
    double *G1 = work, *G2 = G1 + (size_t)ld * nt, *G3 = G2 + (size_t)ld * nt;
    double *Li = G3 + (size_t)ld * nt, *skip = Li + (size_t)nt * nt;
    hipStream_t s = c->stream;
    const bool cond = m > 0 && ld <= TS_MMAX && nt <= 32;   // hand-written kernels take the flag
    dgemm_tn2(c, m, V, nt, T, nt, n, n, T, n, G1);   // [V T]^T T: projections and squared norms
    const double *Gs = G1;
    if (m > 0) {
        gemm_nn(c, n, nt, m, -1.0, V, n, G1, ld, 1.0, T, n);
        dgemm_tn2(c, m, V, nt, T, nt, n, n, T, n, G2);
        gemm_nn(c, n, nt, m, -1.0, V, n, G2, ld, 1.0, T, n);
        Gs = G2;
    }
    launch_chol_inv(s, nt, Gs, m, ld, G1 + m, ld + 1, 1.0e-14, Li, alive, cond ? skip : nullptr);
    gemm_nn(c, n, nt, nt, 1.0, T, n, Li, nt, 0.0, T2, n);
    if (cond) {   // not timed: how much of it runs is known on the device only
        c->tnbuf.ensure((size_t)tsgemm_tn_slices(n, ld, nt) * ld * nt);
        launch_tsgemm_tn2_partial(s, n, m, V, nt, T2, nt, n, T2, n, c->tnbuf.get(), skip);
        launch_tsgemm_tn_reduce(s, n, ld, nt, c->tnbuf.get(), G3);
        launch_tsgemm_nn(s, n, m, nt, -1.0, V, n, G3, ld, 1.0, T2, n, skip);
    } else {
        dgemm_tn2(c, m, V, nt, T2, nt, n, n, T2, n, G3);
        if (m > 0) gemm_nn(c, n, nt, m, -1.0, V, n, G3, ld, 1.0, T2, n);
    }
    launch_chol_inv(s, nt, G3, m, ld, nullptr, 0, 1.0e-14, Li, alive);
    gemm_nn(c, n, nt, nt, 1.0, T2, n, Li, nt, 0.0, T, n);
}

std::vector<int32_t> lowest_diag_rows(const std::vector<double> &d, int k) {
    std::vector<int32_t> idx(d.size());
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return d[a] < d[b]; });
    idx.resize(std::min<size_t>(idx.size(), (size_t)k));
    return idx;
}

// ------------------------------------------------------------------------------------------
// FP64 block Davidson for the k lowest eigenpairs of A (TDA).
struct DavOpts {
    bool fp32 = false;             // FP32 operator
    const double *V0 = nullptr;    // device (n x nv0) start vectors instead of unit vectors
    int nv0 = 0;
    int maxiter = 300;
    bool allow_unconverged = false;
};

void davidson_tda(stda_hip_ctx *c, int k, double tol, std::vector<double> &theta,
                  DBuf<double> &X, const DavOpts &o = DavOpts()) {
    const int64_t n = c->nci;
    hipStream_t s = c->stream;
    k = (int)std::min<int64_t>(k, n);
    int nguess = (int)std::min<int64_t>(n, std::max(k + 8, k + k / 2));
    int maxdim = (int)std::min<int64_t>(n, std::max(6 * k, k + 64));
    if (maxdim < nguess) maxdim = nguess;
    // new blocks of up to 64 directions are orthonormalised on the device
    // (dependent ones stay as zero columns, marked dead; the host learns which
    // with the next Gram block, in the same synchronisation)
    const bool devortho = k <= 64;
    DBuf<double> V((size_t)n * maxdim), AV((size_t)n * maxdim), R((size_t)n * maxdim);
    DBuf<double> AX((size_t)n * k), nrm(k), tmp, T2;
    // Gram block, followed by the alive flags of the new block (one download)
    DBuf<double> Hc((size_t)maxdim * maxdim + maxdim), yt((size_t)maxdim * k + k);
    DBuf<double> ow;
    if (devortho) {
        T2.resize((size_t)n * k);
        ow.resize(3 * (size_t)(maxdim + k) * k + (size_t)k * k + 1);
    }
    X.resize((size_t)n * k);
    // pinned staging: Gram block and alive flags | Y and theta | residual norms
    const size_t o_hc = 0, o_yt = o_hc + (size_t)maxdim * maxdim + maxdim,
                 o_rn = o_yt + (size_t)maxdim * k + k;
    double *pin = c->pinned(o_rn + (size_t)k);
    std::vector<char> alive(maxdim, 0);
    int m = 0, mnew = nguess;
    bool flags_pending = false;   // alive flags of the new block still on the device
    if (o.V0) {
        int nv0 = std::min(o.nv0, maxdim);
        STDA_HIP_CHECK(hipMemcpyAsync(V.get(), o.V0, (size_t)n * nv0 * 8, hipMemcpyDeviceToDevice, s));
        if (devortho && nv0 <= k) {
            orthonormalize_dev(c, V.get(), 0, V.get(), nv0, T2.get(), ow.get(),
                               Hc.get() + (size_t)nv0 * nv0);
            mnew = nv0;
            flags_pending = true;
        } else {
            mnew = orthonormalize(c, V.get(), 0, V.get(), nv0, tmp);
            std::fill(alive.begin(), alive.begin() + mnew, 1);
        }
    } else {
        std::vector<int32_t> rows = lowest_diag_rows(c->ed_h, nguess);
        DBuf<int32_t> rowsd;
        rowsd.upload(rows.data(), rows.size(), s);
        launch_set_unit(s, n, nguess, V.get(), rowsd.get());
        std::fill(alive.begin(), alive.begin() + nguess, 1);
    }
    std::vector<double> H((size_t)maxdim * maxdim, 0.0), Hs, w;
    std::vector<int> idx;
    const int maxiter = o.maxiter;
    for (int it = 0; it < maxiter; ++it) {
        {
            HostSpan hs(c, "dav_matvec");
            apply_ops(c, OP_A, -1, mnew, V.get() + n * m, n, AV.get() + n * m, n, nullptr, 0, o.fp32);
        }
        HostSpan hs_it(c, "dav_iter_other");
        // H(0:m+mnew, m:m+mnew) = V^T AV, with the alive flags of the new block
        const int mt = m + mnew;
        dgemm_tn(c, mt, mnew, n, V.get(), n, AV.get() + n * m, n, Hc.get());
        STDA_HIP_CHECK(hipMemcpyAsync(pin + o_hc, Hc.get(),
                                      ((size_t)mt * mnew + (flags_pending ? mnew : 0)) * 8,
                                      hipMemcpyDeviceToHost, s));
        c->sync();
        if (flags_pending) {
            int nnew = 0;
            const double *fl = pin + o_hc + (size_t)mt * mnew;
            for (int j = 0; j < mnew; ++j) nnew += (alive[m + j] = fl[j] != 0.0 ? 1 : 0);
            flags_pending = false;
            // No new direction survived although residuals were above the
            // tolerance.  For the FP32 operator this is its working precision
            // (an FP64 refinement follows, X and theta are those of the previous
            // iteration); in FP64 the iteration cannot progress: an error.
            if (nnew == 0 && it > 0) {
                if (o.fp32 || o.allow_unconverged) return;
                throw Error(STDA_HIP_ERR_NOT_CONVERGED, "Davidson did not converge (no new directions)");
            }
        }
        for (int j = 0; j < mnew; ++j)
            for (int i = 0; i < mt; ++i) {
                double v = pin[o_hc + i + (size_t)mt * j];
                H[i + (size_t)maxdim * (m + j)] = v;
                H[(m + j) + (size_t)maxdim * i] = v;
            }
        m = mt;
        // Rayleigh-Ritz on the host (FP64) over the live columns
        idx.clear();
        for (int i = 0; i < m; ++i)
            if (alive[i]) idx.push_back(i);
        const int na = (int)idx.size();
        Hs.assign((size_t)na * na, 0.0);
        for (int b = 0; b < na; ++b)
            for (int a2 = 0; a2 < na; ++a2)
                Hs[a2 + (size_t)na * b] = 0.5 * (H[idx[a2] + (size_t)maxdim * idx[b]] +
                                                 H[idx[b] + (size_t)maxdim * idx[a2]]);
        {
            HostSpan hs(c, "dav_rr_eig");
            host_syev(na, Hs, w);
        }
        const int kk = std::min(k, na);
        double *Y = pin + o_yt;
        std::fill(Y, Y + (size_t)m * kk, 0.0);
        for (int j = 0; j < kk; ++j)
            for (int a2 = 0; a2 < na; ++a2) Y[idx[a2] + (size_t)m * j] = Hs[a2 + (size_t)na * j];
        for (int j = 0; j < kk; ++j) Y[(size_t)m * kk + j] = w[j];
        STDA_HIP_CHECK(hipMemcpyAsync(yt.get(), Y, ((size_t)m * kk + kk) * 8, hipMemcpyHostToDevice, s));
        theta.assign(w.begin(), w.begin() + kk);
        const double *th = yt.get() + (size_t)m * kk;
        gemm_nn2(c, n, kk, m, V.get(), AV.get(), n, yt.get(), m, X.get(), AX.get(), n);
        residual_norms(c, n, kk, AX.get(), X.get(), th, R.get(), nrm.get());
        double *rn = pin + o_rn;
        STDA_HIP_CHECK(hipMemcpyAsync(rn, nrm.get(), (size_t)kk * 8, hipMemcpyDeviceToHost, s));
        c->sync();
        std::vector<int> un;
        for (int j = 0; j < kk; ++j)
            if (!(rn[j] < tol)) un.push_back(j);
        if (un.empty() && kk == k) return;
        if (na == n) return;   // full space: exact
        if (it + 1 == maxiter && o.allow_unconverged) return;
        // preconditioned unconverged residuals become the new directions
        int nun = (int)un.size();
        if (m + nun > maxdim) {   // restart with the Ritz vectors
            STDA_HIP_CHECK(hipMemcpyAsync(V.get(), X.get(), (size_t)n * kk * 8,
                                          hipMemcpyDeviceToDevice, s));
            STDA_HIP_CHECK(hipMemcpyAsync(AV.get(), AX.get(), (size_t)n * kk * 8,
                                          hipMemcpyDeviceToDevice, s));
            std::fill(H.begin(), H.end(), 0.0);
            for (int j = 0; j < kk; ++j) H[j + (size_t)maxdim * j] = theta[j];
            std::fill(alive.begin(), alive.end(), 0);
            std::fill(alive.begin(), alive.begin() + kk, 1);
            m = kk;
        }
        int room = std::min(nun, maxdim - m);
        // the first room unconverged residuals, preconditioned, in column
        // order (the device applies the same test to the same norms)
        launch_precondition_compact(s, n, kk, R.get(), c->ed.get(), th, nrm.get(), tol, room,
                                    V.get() + n * m);
        {
            HostSpan hs(c, "dav_ortho");
            if (devortho) {
                orthonormalize_dev(c, V.get(), m, V.get() + n * m, room, T2.get(), ow.get(),
                                   Hc.get() + (size_t)(m + room) * room);
                mnew = room;
                flags_pending = true;
            } else {
                mnew = orthonormalize(c, V.get(), m, V.get() + n * m, room, tmp);
                std::fill(alive.begin() + m, alive.begin() + m + mnew, 1);
                if (mnew == 0) {   // no new directions
                    if (o.fp32 || o.allow_unconverged) return;
                    throw Error(STDA_HIP_ERR_NOT_CONVERGED, "Davidson did not converge (no new directions)");
                }
            }
        }
    }
    throw Error(STDA_HIP_ERR_NOT_CONVERGED, "Davidson did not converge");
}

// ------------------------------------------------------------------------------------------
// Paired Davidson for the RPA problem (A-B)(A+B)|X+Y> = w^2 |X+Y>, using
// the Hermitian reduced form (Am^1/2 Ap Am^1/2) of the subspace problem so
// that (A-B)^1/2 is never formed in the full space.
void davidson_rpa(stda_hip_ctx *c, int k, double tol, std::vector<double> &omega,
                  DBuf<double> &XpY, DBuf<double> &XmY, const DavOpts &o = DavOpts()) {
    const int64_t n = c->nci;
    hipStream_t s = c->stream;
    k = (int)std::min<int64_t>(k, n);
    int nguess = (int)std::min<int64_t>(n, std::max(k + 8, k + k / 2));
    // Thick restart: the Ritz pairs of the lowest kt = k + max(8, k/2) roots are
    // kept, so that roots in a dense cluster (the 400-atom workload has 50 roots
    // within 0.1 eV) keep their neighbours through a restart.
    const int kt_max = (int)std::min<int64_t>(n, k + std::max(8, k / 2));
    int maxdim = (int)std::min<int64_t>(n, std::max(10 * k, 2 * k + 128));
    if (maxdim < nguess) maxdim = nguess;
    DBuf<double> V((size_t)n * maxdim), P((size_t)n * maxdim), M((size_t)n * maxdim);
    DBuf<double> R((size_t)n * 2 * k), PX((size_t)n * k), MX((size_t)n * k), tmp, Yd, nrm(2 * k);
    XpY.resize((size_t)n * k);
    XmY.resize((size_t)n * k);
    int m = 0, mnew = nguess;
    if (o.V0) {
        int nv0 = std::min(o.nv0, maxdim);
        STDA_HIP_CHECK(hipMemcpyAsync(V.get(), o.V0, (size_t)n * nv0 * 8, hipMemcpyDeviceToDevice, s));
        mnew = orthonormalize(c, V.get(), 0, V.get(), nv0, tmp);
    } else {
        std::vector<int32_t> rows = lowest_diag_rows(c->ed_h, nguess);
        DBuf<int32_t> rowsd;
        rowsd.upload(rows.data(), rows.size(), s);
        launch_set_unit(s, n, nguess, V.get(), rowsd.get());
    }
    std::vector<double> Ap((size_t)maxdim * maxdim, 0.0), Am((size_t)maxdim * maxdim, 0.0);
    for (int it = 0; it < o.maxiter; ++it) {
        apply_ops(c, OP_APB, OP_AMB, mnew, V.get() + n * m, n, P.get() + n * m, n, M.get() + n * m,
                  n, o.fp32);
        int mt = m + mnew;
        DBuf<double> Hc((size_t)mt * mnew * 2);
        dgemm_tn(c, mt, mnew, n, V.get(), n, P.get() + n * m, n, Hc.get());
        dgemm_tn(c, mt, mnew, n, V.get(), n, M.get() + n * m, n, Hc.get() + (size_t)mt * mnew);
        std::vector<double> hc((size_t)mt * mnew * 2);
        Hc.download(hc.data(), hc.size(), s);
        c->sync();
        for (int j = 0; j < mnew; ++j)
            for (int i = 0; i < mt; ++i) {
                double vp = hc[i + (size_t)mt * j], vm = hc[(size_t)mt * mnew + i + (size_t)mt * j];
                Ap[i + (size_t)maxdim * (m + j)] = vp; Ap[(m + j) + (size_t)maxdim * i] = vp;
                Am[i + (size_t)maxdim * (m + j)] = vm; Am[(m + j) + (size_t)maxdim * i] = vm;
            }
        m = mt;
        std::vector<double> ap((size_t)m * m), am((size_t)m * m);
        for (int j = 0; j < m; ++j)
            for (int i = 0; i < m; ++i) {
                ap[i + (size_t)m * j] = 0.5 * (Ap[i + (size_t)maxdim * j] + Ap[j + (size_t)maxdim * i]);
                am[i + (size_t)m * j] = 0.5 * (Am[i + (size_t)maxdim * j] + Am[j + (size_t)maxdim * i]);
            }
        double minev = 0.0;
        std::vector<double> sq = sym_sqrt(am, m, false, &minev);
        if (minev <= 0.0)
            throw Error(STDA_HIP_ERR_AMB_NOT_PD, "matrix power impossible (A-B not positive definite)");
        std::vector<double> Hm = matmul(matmul(sq, ap, m, m, m), sq, m, m, m), w2;
        host_syev(m, Hm, w2);
        int kk = std::min(k, m);
        omega.assign(kk, 0.0);
        std::vector<double> xr((size_t)m * kk), yr((size_t)m * kk);   // (X+Y)_red, (X-Y)_red
        std::vector<double> Z(Hm.begin(), Hm.begin() + (size_t)m * kk);
        std::vector<double> xp = matmul(sq, Z, m, m, kk);
        for (int j = 0; j < kk; ++j) {
            if (w2[j] <= 0.0) throw Error(STDA_HIP_ERR_NO_ROOTS, "RPA diag failed (w^2 <= 0)");
            omega[j] = std::sqrt(w2[j]);
            double f = 1.0 / std::sqrt(omega[j]);
            for (int i = 0; i < m; ++i) xr[i + (size_t)m * j] = xp[i + (size_t)m * j] * f;
        }
        std::vector<double> apx = matmul(ap, xr, m, m, kk);
        for (int j = 0; j < kk; ++j)
            for (int i = 0; i < m; ++i) yr[i + (size_t)m * j] = apx[i + (size_t)m * j] / omega[j];
        // full-space Ritz vectors and residuals
        std::vector<double> both((size_t)m * 2 * kk);
        std::copy(xr.begin(), xr.end(), both.begin());
        std::copy(yr.begin(), yr.end(), both.begin() + (size_t)m * kk);
        Yd.upload(both.data(), both.size(), s);
        gemm_nn(c, n, kk, m, 1.0, V.get(), n, Yd.get(), m, 0.0, XpY.get(), n);
        gemm_nn(c, n, kk, m, 1.0, V.get(), n, Yd.get() + (size_t)m * kk, m, 0.0,
              XmY.get(), n);
        gemm_nn(c, n, kk, m, 1.0, P.get(), n, Yd.get(), m, 0.0, PX.get(), n);
        gemm_nn(c, n, kk, m, 1.0, M.get(), n, Yd.get() + (size_t)m * kk, m, 0.0,
              MX.get(), n);
        DBuf<double> om;
        om.upload(omega.data(), kk, s);
        launch_residual(s, n, kk, PX.get(), XmY.get(), om.get(), R.get());            // (A+B)(X+Y) - w(X-Y)
        launch_residual(s, n, kk, MX.get(), XpY.get(), om.get(), R.get() + n * kk);   // (A-B)(X-Y) - w(X+Y)
        col_norms(c, n, 2 * kk, R.get(), nrm.get());
        std::vector<double> rn(2 * kk);
        nrm.download(rn.data(), 2 * kk, s);
        c->sync();
        std::vector<int> un;
        for (int j = 0; j < 2 * kk; ++j)
            if (!(rn[j] < tol)) un.push_back(j);
        if (un.empty() && kk == k) return;
        if (m == n) return;
        if (it + 1 == o.maxiter && o.allow_unconverged) return;
        int nun = (int)un.size();
        for (int q = 0; q < nun; ++q)
            if (un[q] != q)
                STDA_HIP_CHECK(hipMemcpyAsync(R.get() + n * q, R.get() + n * un[q], n * 8,
                                              hipMemcpyDeviceToDevice, s));
        std::vector<double> thu(nun);
        for (int q = 0; q < nun; ++q) thu[q] = omega[un[q] % kk];
        DBuf<double> thud;
        thud.upload(thu.data(), nun, s);
        launch_precondition(s, n, nun, R.get(), c->ed.get(), thud.get(), R.get());
        if (m + nun > maxdim) {
            // restart: V <- orthonormal basis of span(X+Y, X-Y) of the lowest kt
            // Ritz pairs, P/M transformed alike
            const int kt = std::min(m, kt_max);
            std::vector<double> G((size_t)m * 2 * kt);
            {
                std::vector<double> Zt(Hm.begin(), Hm.begin() + (size_t)m * kt);
                std::vector<double> xt = matmul(sq, Zt, m, m, kt);
                for (int j = 0; j < kt; ++j) {
                    const double wj = std::sqrt(std::max(w2[j], 1.0e-300));
                    for (int i = 0; i < m; ++i) xt[i + (size_t)m * j] /= std::sqrt(wj);
                }
                std::vector<double> at = matmul(ap, xt, m, m, kt);
                for (int j = 0; j < kt; ++j) {
                    const double wj = std::sqrt(std::max(w2[j], 1.0e-300));
                    for (int i = 0; i < m; ++i) {
                        G[i + (size_t)m * j] = xt[i + (size_t)m * j];
                        G[i + (size_t)m * (kt + j)] = at[i + (size_t)m * j] / wj;
                    }
                }
            }
            int nc = 2 * kt;
            std::vector<double> S = matmul([&] {
                std::vector<double> t((size_t)nc * m);
                for (int j = 0; j < nc; ++j)
                    for (int i = 0; i < m; ++i) t[j + (size_t)nc * i] = G[i + (size_t)m * j];
                return t;
            }(), G, nc, m, nc), w;
            host_syev(nc, S, w);
            std::vector<double> Zc;
            int kept = 0;
            for (int q = nc - 1; q >= 0; --q) {
                if (w[q] <= 1e-12 * w[nc - 1]) continue;
                std::vector<double> col(nc);
                for (int i = 0; i < nc; ++i) col[i] = S[i + (size_t)nc * q] / std::sqrt(w[q]);
                std::vector<double> gz = matmul(G, col, m, nc, 1);
                Zc.insert(Zc.end(), gz.begin(), gz.end());
                ++kept;
            }
            DBuf<double> Zd, V2((size_t)n * kept), P2((size_t)n * kept), M2((size_t)n * kept);
            Zd.upload(Zc.data(), Zc.size(), s);
            gemm_nn(c, n, kept, m, 1.0, V.get(), n, Zd.get(), m, 0.0, V2.get(), n);
            gemm_nn(c, n, kept, m, 1.0, P.get(), n, Zd.get(), m, 0.0, P2.get(), n);
            gemm_nn(c, n, kept, m, 1.0, M.get(), n, Zd.get(), m, 0.0, M2.get(), n);
            STDA_HIP_CHECK(hipMemcpyAsync(V.get(), V2.get(), (size_t)n * kept * 8, hipMemcpyDeviceToDevice, s));
            STDA_HIP_CHECK(hipMemcpyAsync(P.get(), P2.get(), (size_t)n * kept * 8, hipMemcpyDeviceToDevice, s));
            STDA_HIP_CHECK(hipMemcpyAsync(M.get(), M2.get(), (size_t)n * kept * 8, hipMemcpyDeviceToDevice, s));
            // reduced matrices in the new basis: Zc^T Ap Zc
            std::vector<double> zt((size_t)kept * m);
            for (int j = 0; j < kept; ++j)
                for (int i = 0; i < m; ++i) zt[j + (size_t)kept * i] = Zc[i + (size_t)m * j];
            std::vector<double> ap2 = matmul(zt, matmul(ap, Zc, m, m, kept), kept, m, kept);
            std::vector<double> am2 = matmul(zt, matmul(am, Zc, m, m, kept), kept, m, kept);
            std::fill(Ap.begin(), Ap.end(), 0.0);
            std::fill(Am.begin(), Am.end(), 0.0);
            for (int j = 0; j < kept; ++j)
                for (int i = 0; i < kept; ++i) {
                    Ap[i + (size_t)maxdim * j] = ap2[i + (size_t)kept * j];
                    Am[i + (size_t)maxdim * j] = am2[i + (size_t)kept * j];
                }
            m = kept;
            c->sync();
        }
        int room = std::min(nun, maxdim - m);
        STDA_HIP_CHECK(hipMemcpyAsync(V.get() + n * m, R.get(), (size_t)n * room * 8,
                                      hipMemcpyDeviceToDevice, s));
        mnew = orthonormalize(c, V.get(), m, V.get() + n * m, room, tmp);
        // every new direction lies in the space although residuals are above the
        // tolerance: the iteration cannot progress
        if (mnew == 0) {
            if (o.allow_unconverged) return;
            throw Error(STDA_HIP_ERR_NOT_CONVERGED, "RPA Davidson did not converge (no new directions)");
        }
    }
    throw Error(STDA_HIP_ERR_NOT_CONVERGED, "RPA Davidson did not converge");
}

// ------------------------------------------------------------------------------------------
// properties (stda.f:1207-1305; apbtrafo.f:160-265)
void gather_moments(stda_hip_ctx *c) {
    if (!c->have_moments) return;
    c->mom.resize((size_t)c->nci * 9);
    launch_csf_moments(c->stream, c->moments.get(), c->npk, c->pair.get(), c->nci, c->no, c->nv,
                       c->mom.get());
}

// T(9, k) = mom^T U
std::vector<double> moment_products(stda_hip_ctx *c, const double *U, int k) {
    std::vector<double> T((size_t)9 * k, 0.0);
    if (!c->have_moments || k == 0) return T;
    DBuf<double> Td((size_t)9 * k);
    dgemm_tn(c, 9, k, c->nci, c->mom.get(), c->nci, U, c->nci, Td.get());
    Td.download(T.data(), T.size(), c->stream);
    c->sync();
    return T;
}

}  // namespace

// ============================================================================================
// setup: stda.f:454-790
static void do_load(stda_hip_ctx *c, const stda_hip_system *sys, const stda_hip_params *par) {
    if (!sys || !par || sys->ncent <= 0 || sys->nao <= 0 || sys->no < 0 || sys->nv < 0 ||
        !sys->clow || !sys->epsi || !sys->iaoat || !sys->xyz)
        throw Error(STDA_HIP_ERR_ARG, "invalid system description");
    if (sys->no == 0 || sys->nv == 0)
        throw Error(STDA_HIP_ERR_NO_CSF, "no CSF, increase energy threshold (-e option)");
    hipStream_t s = c->stream;
    c->ready = false;
    c->loaded = false;
    c->nb_max = 0;
    c->nb_chunk = 0;
    c->ncent = sys->ncent; c->nao = sys->nao; c->no = sys->no; c->nv = sys->nv;
    const int na = c->ncent, nao = c->nao, no = c->no, nv = c->nv, moci = no + nv;
    c->npair = (int64_t)no * nv;
    if (c->npair > INT32_MAX) throw Error(STDA_HIP_ERR_TOO_LARGE, "no*nv exceeds int32 pair indexing");
    c->ak = par->triplet ? 0.0 : 2.0;                    // stda.f:482-483
    c->ax = par->ax;
    c->small_ax = std::fabs(par->ax) < 1.0e-6;           // stda.f:1893
    double betaj = par->betaj < -99.0 ? 0.20 + 1.830 * par->ax : par->betaj;   // stda.f:490-500
    double alphak = par->alphak < -99.0 ? 1.420 + 0.480 * par->ax : par->alphak;
    c->alphak = alphak;
    c->betaj = betaj;
    c->kshift = par->kshift;
    if (par->kshift) {
        c->ks_max = par->kshift_max_ev; c->ks_width = par->kshift_width_ev; c->ks_steep = par->kshift_steep;
    }
    c->thr = par->thr_ev / kHartreeEv;
    c->thrp = par->thrp;
    c->fthr = par->fthr_ev > 0.0 ? par->fthr_ev / kHartreeEv : 1.7976931348623157e308;
    for (int i = 0; i < na; ++i) {
        int z = (int)sys->xyz[4 * i + 3];
        if (z < 1 || z > 94) throw Error(STDA_HIP_ERR_ARG, "nuclear charge outside 1..94");
    }

    // AOs sorted by atom (stable), Loewdin MOs transposed into that order
    std::vector<int> order(nao);
    std::iota(order.begin(), order.end(), 0);
    for (int m = 0; m < nao; ++m)
        if (sys->iaoat[m] < 1 || sys->iaoat[m] > na) throw Error(STDA_HIP_ERR_ARG, "iaoat out of range");
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return sys->iaoat[a] < sys->iaoat[b]; });
    std::vector<int> aptr(na + 1, 0), aof(nao);
    for (int m = 0; m < nao; ++m) aptr[sys->iaoat[m]]++;
    for (int a = 0; a < na; ++a) aptr[a + 1] += aptr[a];
    for (int m = 0; m < nao; ++m) aof[m] = sys->iaoat[order[m]] - 1;
    std::vector<double> cth((size_t)nao * moci);
    for (int p = 0; p < moci; ++p)
        for (int m = 0; m < nao; ++m) cth[m + (size_t)nao * p] = sys->clow[order[m] + (size_t)nao * p];
    c->ct.upload(cth.data(), cth.size(), s);
    c->atom_ptr.upload(aptr.data(), aptr.size(), s);
    c->aptr_h = aptr;
    c->atom_of.upload(aof.data(), aof.size(), s);
    c->eta.upload(kEta, 94, s);
    c->xyz.upload(sys->xyz, (size_t)4 * na, s);
    c->eps.upload(sys->epsi, moci, s);
    c->have_moments = sys->moments != nullptr;
    c->npk = (int64_t)moci * (moci + 1) / 2;
    if (c->have_moments) {
        c->moments.upload(sys->moments, (size_t)9 * c->npk, s);
        c->moments_h.assign(sys->moments, sys->moments + (size_t)9 * c->npk);
    }
    c->sync();
    c->loaded = true;
}

static void do_select(stda_hip_ctx *c) {
    if (!c->loaded) throw Error(STDA_HIP_ERR_STATE, "load must be called before select");
    Timer ttot;
    hipStream_t s = c->stream;
    c->ready = false;
    c->nb_max = 0;
    c->nb_max32 = 0;
    c->have32 = false;
    c->flush_timers();
    c->fold_stats();
    const int na = c->ncent, nao = c->nao, no = c->no, nv = c->nv, moci = no + nv;
    const double alphak = c->alphak, betaj = c->betaj;
    Timer t0;

    // gamma (stda.f:509-532)
    c->gamj.resize((size_t)na * na);
    c->gamk.resize((size_t)na * na);
    launch_gamma(s, na, c->xyz.get(), c->eta.get(), c->ax, alphak, betaj, c->gamj.get(), c->gamk.get());
    // q_ii, q_aa and (ii|aa)^J (stda.f:465-566)
    c->qdiag.resize((size_t)na * moci);
    launch_charges(s, CH_DIAG, na, nao, no, nv, c->ct.get(), c->atom_ptr.get(), c->qdiag.get());
    DBuf<double> pii((size_t)na * no);
    dgemm(c->blas, false, false, na, no, na, 1.0, c->gamj.get(), na, c->qdiag.get(), na, 0.0, pii.get(), na);
    c->jdiag.resize((size_t)nv * no);
    dgemm(c->blas, true, false, nv, no, na, 1.0, c->qdiag.get() + (size_t)na * no, na, pii.get(), na, 0.0,
          c->jdiag.get(), nv);
    // q_ia over all pairs and p_ia = gamK q_ia (stda.f:576-589)
    c->qia.resize((size_t)na * c->npair);
    c->pia.resize((size_t)na * c->npair);
    launch_charges(s, CH_IA, na, nao, no, nv, c->ct.get(), c->atom_ptr.get(), c->qia.get());
    dgemm(c->blas, false, false, na, c->npair, na, 1.0, c->gamk.get(), na, c->qia.get(), na, 0.0,
          c->pia.get(), na);
    // packed q_ij -> p_ij = gamJ q_ij, packed q_ab (stda.f:728-763)
    int64_t nij = (int64_t)no * (no + 1) / 2, nab = (int64_t)nv * (nv + 1) / 2;
    c->qij.resize((size_t)na * nij);
    c->pij.resize((size_t)na * nij);
    c->qab.resize((size_t)na * nab);
    launch_charges(s, CH_OCC_PACKED, na, nao, no, nv, c->ct.get(), c->atom_ptr.get(), c->qij.get());
    launch_charges(s, CH_VIR_PACKED, na, nao, no, nv, c->ct.get(), c->atom_ptr.get(), c->qab.get());
    dgemm(c->blas, false, false, na, nij, na, 1.0, c->gamj.get(), na, c->qij.get(), na, 0.0,
          c->pij.get(), na);
    c->qij.release();
    c->sync();
    c->timings[0] = t0.ms();

    // screening (stda.f:593-632)
    Timer t1;
    c->de.resize(c->npair);
    c->kdiag.resize(c->npair);
    c->flags.resize(c->npair);
    launch_screen(s, na, no, nv, c->eps.get(), c->jdiag.get(), c->pia.get(), c->qia.get(), c->ak,
                  c->thr, c->fthr, c->kshift, c->ks_max, c->ks_width, c->ks_steep,
                  c->de.get(), c->kdiag.get(), c->flags.get());
    std::vector<unsigned char> fl(c->npair);
    std::vector<double> deh(c->npair);
    c->flags.download(fl.data(), fl.size(), s);
    c->de.download(deh.data(), deh.size(), s);
    c->sync();
    // occupied MOs in the order the screening loop visits them (stda.f:600-631), or for
    // dual thresholds the chromophore MOs first (stda-rw_dual.f:712-826)
    std::vector<int> occ_order;
    if (!c->dual_atoms.empty()) {
        // chromophore MOs: squared Loewdin coefficients on the listed atoms sum above 0.1
        // (stda-rw_dual.f:676-698; 0.1 is a default-real literal there)
        std::vector<double> cth((size_t)nao * no), w(no, 0.0);
        STDA_HIP_CHECK(hipMemcpyAsync(cth.data(), c->ct.get(), cth.size() * 8,
                                      hipMemcpyDeviceToHost, s));
        c->sync();
        std::vector<double> cnt(na, 0.0);
        for (int32_t a : c->dual_atoms) {
            if (a < 1 || a > na) throw Error(STDA_HIP_ERR_ARG, "dual atom list: atom out of range");
            cnt[a - 1] += 1.0;
        }
        for (int a = 0; a < na; ++a) {
            if (cnt[a] == 0.0) continue;
            for (int i = 0; i < no; ++i)
                for (int m = c->aptr_h[a]; m < c->aptr_h[a + 1]; ++m) {
                    const double v = cth[m + (size_t)nao * i];
                    w[i] += cnt[a] * (v * v);
                }
        }
        const double t1 = (c->thr * kHartreeEv) / kHartreeEv;   // thresh(1), eV and back
        const double t2 = c->dual_thr2_ev / kHartreeEv;
        const double f2 = ((1.0 + (double)0.8f * c->ax) * t2) * 2.0;   // stda-rw_dual.f:776
        std::vector<double> ti(no), fi(no);
        std::vector<char> chrom(no);
        for (int i = 0; i < no; ++i) {
            chrom[i] = w[i] > (double)0.1f;
            ti[i] = chrom[i] ? t1 : t2;
            fi[i] = chrom[i] ? c->fthr : f2;
        }
        for (int g = 1; g >= 0; --g)
            for (int i = 0; i < no; ++i)
                if (chrom[i] == g) occ_order.push_back(i);
        for (int64_t p = 0; p < c->npair; ++p) {
            const int i = (int)(p / nv);
            unsigned char f = 0;
            if (deh[p] <= ti[i]) f |= 1;
            if (deh[p] > ti[i] && deh[p] < fi[i]) f |= 2;
            fl[p] = f;
        }
    } else {
        for (int i = 0; i < no; ++i) occ_order.push_back(i);
    }
    if (c->pt2_off)
        for (auto &f : fl) f &= (unsigned char)~2u;
    if (!c->dual_atoms.empty() || c->pt2_off) c->flags.upload(fl.data(), fl.size(), s);
    std::vector<int32_t> prim, sec;
    for (int i : occ_order)
        for (int a = 0; a < nv; ++a) {
            const int64_t p = (int64_t)i * nv + a;
            if (fl[p] & 1) prim.push_back((int32_t)p);
            if (fl[p] & 2) sec.push_back((int32_t)p);
        }
    c->nprim = (int64_t)prim.size();
    c->nsec = (int64_t)sec.size();
    if (c->nprim < 1) throw Error(STDA_HIP_ERR_NO_CSF, "no CSF, increase energy threshold (-e option)");
    // energy order of the primaries: the reference's selection sort, ties included
    {
        std::vector<double> keys(prim.size());
        for (size_t q = 0; q < prim.size(); ++q) keys[q] = deh[prim[q]];
        reference_selection_order(std::move(keys), prim);
    }
    c->timings[1] = t1.ms();

    // PT2 selection (stda.f:1773-1863)
    Timer t2;
    std::vector<double> pt2(c->npair, 0.0);
    std::vector<unsigned char> incl(c->npair, 0);
    if (c->nsec > 0) {
        // tiles of the (k,c) grid holding secondaries and of the (j,b) grid
        // holding primaries; tile_of maps a block (k / PT2_T, c / PT2_T) to its tile
        const int nvb = (nv + PT2_T - 1) / PT2_T;
        auto make_tiles = [&](unsigned char bit, std::vector<int32_t> &occ,
                              std::vector<int32_t> &vir, std::vector<int32_t> &tile_of) {
            tile_of.assign((size_t)((no + PT2_T - 1) / PT2_T) * nvb, -1);
            int nt = 0;
            for (int k0 = 0; k0 < no; k0 += PT2_T)
                for (int c0 = 0; c0 < nv; c0 += PT2_T) {
                    bool any = false;
                    for (int k = k0; k < std::min(no, k0 + PT2_T) && !any; ++k)
                        for (int cc = c0; cc < std::min(nv, c0 + PT2_T); ++cc)
                            if (fl[(int64_t)k * nv + cc] & bit) { any = true; break; }
                    if (!any) continue;
                    tile_of[(size_t)(k0 / PT2_T) * nvb + c0 / PT2_T] = nt;
                    for (int k = 0; k < PT2_T; ++k) occ.push_back(k0 + k < no ? k0 + k : -1);
                    for (int cc = 0; cc < PT2_T; ++cc) vir.push_back(c0 + cc < nv ? c0 + cc : -1);
                    ++nt;
                }
            return nt;
        };
        // de of every tile slot; NaN where the slot's pair lacks `bit`
        auto slot_de = [&](int nt, const std::vector<int32_t> &occ, const std::vector<int32_t> &vir,
                           unsigned char bit) {
            std::vector<double> d((size_t)nt * PT2_SLOTS, std::nan(""));
            for (int t = 0; t < nt; ++t)
                for (int x = 0; x < PT2_T; ++x)
                    for (int y = 0; y < PT2_T; ++y) {
                        const int k = occ[(size_t)t * PT2_T + x], cc = vir[(size_t)t * PT2_T + y];
                        if (k < 0 || cc < 0) continue;
                        const int64_t p = (int64_t)k * nv + cc;
                        if (fl[p] & bit) d[(size_t)t * PT2_SLOTS + x * PT2_T + y] = deh[p];
                    }
            return d;
        };
        // slot of pair p in the tile grid of tile_of
        auto slot_of = [&](const std::vector<int32_t> &tile_of, int64_t p) {
            const int k = (int)(p / nv), cc = (int)(p % nv);
            const int t = tile_of[(size_t)(k / PT2_T) * nvb + cc / PT2_T];
            return (int64_t)t * PT2_SLOTS + (k % PT2_T) * PT2_T + cc % PT2_T;
        };
        std::vector<int32_t> ro, rv, co, cv, rtile, ctile;
        const int nrt = make_tiles(2, ro, rv, rtile);
        const int nct = make_tiles(1, co, cv, ctile);
        const std::vector<double> drh = slot_de(nrt, ro, rv, 2), dch = slot_de(nct, co, cv, 1);
        DBuf<int32_t> rod, rvd, cod, cvd;
        DBuf<double> drd, dcd;
        rod.upload(ro.data(), ro.size(), s); rvd.upload(rv.data(), rv.size(), s);
        cod.upload(co.data(), co.size(), s); cvd.upload(cv.data(), cv.size(), s);
        drd.upload(drh.data(), drh.size(), s); dcd.upload(dch.data(), dch.size(), s);
        TileSet rows{rod.get(), rvd.get(), nrt}, cols{cod.get(), cvd.get(), nct};
        // Column tiles split so that row tiles x splits fill the device in the
        // fewest whole rounds (kPt2Resident work-groups resident at a time);
        // a function of the tile counts only, so a molecule's sums do not
        // depend on what else runs.
        int nsplit = 1;
        {
            double best = std::ceil((double)nrt / kPt2Resident);
            for (int ns = 2; ns <= std::min(16, nct / 2); ++ns) {
                const double t = std::ceil((double)nrt * ns / kPt2Resident) / ns;
                if (t < best - 1e-9) { best = t; nsplit = ns; }
            }
        }
        // every pt term kept for pass 2 when the (secondary slot x primary
        // slot) matrix fits kPt2StoreBytes
        // (STDA_HIP_PT2_STORE_MAX, bytes, overrides the limit)
        const int64_t ldp = (int64_t)PT2_SLOTS * nct;
        double store_max = (double)kPt2StoreBytes;
        if (const char *e = std::getenv("STDA_HIP_PT2_STORE_MAX")) store_max = std::strtod(e, nullptr);
        const bool store = (double)PT2_SLOTS * nrt * (double)ldp * 8.0 <= store_max;
        if (store) c->pt2_terms.ensure((size_t)PT2_SLOTS * nrt * (size_t)ldp);
        double *ptm = store ? c->pt2_terms.get() : nullptr;
        DBuf<double> rowsum(c->npair), rowpart((size_t)nsplit * nrt * PT2_SLOTS);
        std::vector<double> rs(c->npair);
        const double work1 = 4.0 * na * (double)c->nsec * (double)c->nprim;
        const int64_t nij = (int64_t)no * (no + 1) / 2, nab = (int64_t)nv * (nv + 1) / 2;
        // FP64 operands with an even column stride (16-byte loads): the arrays
        // themselves for an even atom count, zero-padded copies otherwise
        const double *op64[4] = {c->pia.get(), c->qia.get(), c->pij.get(), c->qab.get()};
        int na2 = na;
        auto fp64_operands = [&]() {
            if (na % 2 == 0 || na2 != na) return;
            na2 = na + 1;
            DBuf<double> *buf[4] = {&c->pt2_pk64, &c->pt2_q64, &c->pt2_pj64, &c->pt2_qab64};
            const int64_t ncol[4] = {c->npair, c->npair, nij, nab};
            for (int o = 0; o < 4; ++o) {
                buf[o]->ensure((size_t)na2 * ncol[o]);
                launch_pad_rows(s, op64[o], na, ncol[o], na2, buf[o]->get());
                op64[o] = buf[o]->get();
            }
        };
        if (c->precision == 0) {
            fp64_operands();
            EventTimer ev1(c);
            launch_pt2_pipe(s, rows, cols, na, nv, op64[0], op64[1], op64[2], op64[3], na2,
                            drd.get(), dcd.get(), c->ak, nsplit, rowpart.get(), ptm, ldp);
            ev1.finish("pt2_rows", work1);
            launch_pt2_rows_finish(s, rows, nv, nsplit, drd.get(), rowpart.get(), rowsum.get());
            rowsum.download(rs.data(), rs.size(), s);
            c->sync();
        } else {
            // Mixed precision: pass 1 on FP32 copies of the operands (the
            // reference's own real*4 couplings, stda.f:1781-1788), then the row
            // tiles holding a secondary whose sum lies within kPt2Recheck of thrp
            // again in FP64, with the same split: their sums are then those of
            // the FP64 pass bit for bit, so the selection is the FP64 one.
            // columns padded to a multiple of 4 atoms (16-byte loads)
            const int na4 = (na + 3) / 4 * 4;
            c->pt2_pk32.ensure((size_t)na4 * c->npair);
            c->pt2_q32.ensure((size_t)na4 * c->npair);
            c->pt2_pj32.ensure((size_t)na4 * nij);
            c->pt2_qab32.ensure((size_t)na4 * nab);
            launch_to_f32_padded(s, c->pia.get(), na, c->npair, na4, c->pt2_pk32.get());
            launch_to_f32_padded(s, c->qia.get(), na, c->npair, na4, c->pt2_q32.get());
            launch_to_f32_padded(s, c->pij.get(), na, nij, na4, c->pt2_pj32.get());
            launch_to_f32_padded(s, c->qab.get(), na, nab, na4, c->pt2_qab32.get());
            EventTimer ev1(c);
            launch_pt2_pipe_f32(s, rows, cols, na, nv, c->pt2_pk32.get(), c->pt2_q32.get(),
                                c->pt2_pj32.get(), c->pt2_qab32.get(), na4, drd.get(), dcd.get(),
                                c->ak, nsplit, rowpart.get(), ptm, ldp);
            ev1.finish("pt2_rows_f32", work1);
            launch_pt2_rows_finish(s, rows, nv, nsplit, drd.get(), rowpart.get(), rowsum.get());
            rowsum.download(rs.data(), rs.size(), s);
            c->sync();
            std::vector<char> redo(nrt, 0);
            int nredo = 0;
            for (int32_t p : sec)
                if (std::fabs(rs[p] - c->thrp) <= kPt2Recheck * c->thrp) {
                    const int t = (int)(slot_of(rtile, p) / PT2_SLOTS);
                    nredo += redo[t] ? 0 : 1;
                    redo[t] = 1;
                }
            if (nredo > 0) {
                std::vector<int32_t> ro2, rv2;
                std::vector<double> dr2;
                for (int t = 0; t < nrt; ++t) {
                    if (!redo[t]) continue;
                    ro2.insert(ro2.end(), ro.begin() + (size_t)t * PT2_T, ro.begin() + (size_t)(t + 1) * PT2_T);
                    rv2.insert(rv2.end(), rv.begin() + (size_t)t * PT2_T, rv.begin() + (size_t)(t + 1) * PT2_T);
                    dr2.insert(dr2.end(), drh.begin() + (size_t)t * PT2_SLOTS,
                               drh.begin() + (size_t)(t + 1) * PT2_SLOTS);
                }
                DBuf<int32_t> ro2d, rv2d;
                DBuf<double> dr2d, part2((size_t)nsplit * nredo * PT2_SLOTS);
                ro2d.upload(ro2.data(), ro2.size(), s);
                rv2d.upload(rv2.data(), rv2.size(), s);
                dr2d.upload(dr2.data(), dr2.size(), s);
                TileSet rows2{ro2d.get(), rv2d.get(), nredo};
                fp64_operands();
                EventTimer ev3(c);
                launch_pt2_pipe(s, rows2, cols, na, nv, op64[0], op64[1], op64[2], op64[3], na2,
                                dr2d.get(), dcd.get(), c->ak, nsplit, part2.get(), nullptr, 0);
                ev3.finish("pt2_rows_recheck", 4.0 * na * (double)nredo * PT2_SLOTS * (double)c->nprim);
                launch_pt2_rows_finish(s, rows2, nv, nsplit, dr2d.get(), part2.get(), rowsum.get());
                rowsum.download(rs.data(), rs.size(), s);
                c->sync();
            }
        }
        int64_t nexcl = 0;
        for (int32_t p : sec) {
            if (rs[p] > c->thrp) incl[p] = 1;
            else { fl[p] |= 4; ++nexcl; }
        }
        if (nexcl > 0 && store) {
            // pass 2 from the stored terms: column sums over the rejected rows
            std::vector<int32_t> rrow;
            rrow.reserve((size_t)nexcl);
            for (int32_t p : sec)
                if (fl[p] & 4) rrow.push_back((int32_t)slot_of(rtile, p));
            std::sort(rrow.begin(), rrow.end());
            DBuf<int32_t> rrowd;
            rrowd.upload(rrow.data(), rrow.size(), s);
            const int64_t ncb = (ldp + 255) / 256;
            const int nsp2 = (int)std::max<int64_t>(
                1, std::min<int64_t>(nexcl, (1024 + ncb - 1) / ncb));
            DBuf<double> part((size_t)nsp2 * ldp), colsum((size_t)ldp);
            EventTimer ev2(c);
            launch_pt2_colsum(s, (int)nexcl, rrowd.get(), ldp, ptm, nsp2, part.get(),
                              colsum.get());
            ev2.finish("pt2_cols", 4.0 * na * (double)nexcl * (double)c->nprim);
            std::vector<double> cs((size_t)ldp);
            colsum.download(cs.data(), cs.size(), s);
            c->sync();
            for (int32_t p : prim) pt2[p] = cs[(size_t)slot_of(ctile, p)];
        } else if (nexcl > 0) {
            c->flags.upload(fl.data(), fl.size(), s);
            // rejected secondaries grouped by occupied k, 16 virtual c per row tile
            std::vector<int32_t> rk, rc;
            for (int64_t idx = 0; idx < (int64_t)sec.size();) {
                int32_t p = sec[idx];
                if (!(fl[p] & 4)) { ++idx; continue; }
                int k = p / nv;
                rk.push_back(k);
                int filled = 0;
                while (idx < (int64_t)sec.size() && sec[idx] / nv == k && filled < 16) {
                    if (fl[sec[idx]] & 4) { rc.push_back(sec[idx] % nv); ++filled; }
                    ++idx;
                }
                for (; filled < 16; ++filled) rc.push_back(-1);
            }
            std::vector<int32_t> cjv, cbv;
            int nct2 = 0;
            for (int j0 = 0; j0 < no; j0 += 16)
                for (int b0 = 0; b0 < nv; b0 += 16) {
                    bool any = false;
                    for (int j = j0; j < std::min(no, j0 + 16) && !any; ++j)
                        for (int b = b0; b < std::min(nv, b0 + 16); ++b)
                            if (fl[(int64_t)j * nv + b] & 1) { any = true; break; }
                    if (!any) continue;
                    for (int x = 0; x < 16; ++x) cjv.push_back(j0 + x < no ? j0 + x : -1);
                    for (int x = 0; x < 16; ++x) cbv.push_back(b0 + x < nv ? b0 + x : -1);
                    ++nct2;
                }
            DBuf<int32_t> rkd, rcd, cjd, cbd;
            rkd.upload(rk.data(), rk.size(), s); rcd.upload(rc.data(), rc.size(), s);
            cjd.upload(cjv.data(), cjv.size(), s); cbd.upload(cbv.data(), cbv.size(), s);
            DBuf<double> colsum(c->npair);
            // split the row stream so that about 1024 work-groups are in flight
            const int nrt2 = (int)rk.size();
            const int nsp2 = std::max(1, std::min(nrt2, (1024 + nct2 - 1) / std::max(nct2, 1)));
            DBuf<double> part((size_t)nsp2 * nct2 * 256);
            EventTimer ev2(c);
            launch_pt2_cols_sparse(s, nrt2, rkd.get(), rcd.get(), nct2, cjd.get(), cbd.get(), na,
                                   nv, c->pia.get(), c->qia.get(), c->pij.get(), c->qab.get(),
                                   c->de.get(), c->flags.get(), 4, c->flags.get(), 1, c->ak, nsp2,
                                   part.get(), colsum.get());
            ev2.finish("pt2_cols", 4.0 * na * (double)nexcl * (double)c->nprim);
            std::vector<double> cs(c->npair);
            colsum.download(cs.data(), cs.size(), s);
            c->sync();
            for (int32_t p : prim) pt2[p] = cs[p];
        }
    }
    c->timings[2] = t2.ms();

    // CSF list: sorted primaries (ed lowered by PT2), then selected secondaries
    c->pair_h.clear();
    c->ed_h.clear();
    for (int32_t p : prim) { c->pair_h.push_back(p); c->ed_h.push_back(deh[p] - pt2[p]); }
    for (int32_t p : sec)
        if (incl[p]) { c->pair_h.push_back(p); c->ed_h.push_back(deh[p]); }
    c->nci = (int64_t)c->pair_h.size();
    c->npt = c->nci - c->nprim;
    c->pair.upload(c->pair_h.data(), c->nci, s);
    c->ed.upload(c->ed_h.data(), c->nci, s);
    // diagonal corrections of the AO-path operators: ed - ak K_ii + J_ii (and
    // ed - ak K_ii for the small-ax A+B / diagonal A-B)
    std::vector<double> kd(c->npair), jd((size_t)nv * no);
    c->kdiag.download(kd.data(), kd.size(), s);
    c->jdiag.download(jd.data(), jd.size(), s);
    c->sync();
    std::vector<double> corr(c->nci), corr_s(c->nci);
    for (int64_t I = 0; I < c->nci; ++I) {
        int64_t p = c->pair_h[I];
        int64_t i = p / nv, a = p % nv;
        corr[I] = c->ed_h[I] - c->ak * kd[p] + jd[a + (int64_t)nv * i];
        corr_s[I] = c->ed_h[I] - c->ak * kd[p];
    }
    c->corr.upload(corr.data(), corr.size(), s);
    c->corr_small.upload(corr_s.data(), corr_s.size(), s);
    gather_moments(c);
    c->sync();
    c->timings[5] = ttot.ms();
    c->ready = true;
}

// explicit A (nci x nci) on the device, column-major
static void build_a_device(stda_hip_ctx *c, DBuf<double> &A) {
    const int64_t n = c->nci;
    hipStream_t s = c->stream;
    std::vector<int32_t> csf(c->npair, -1);
    for (int64_t I = 0; I < n; ++I) csf[c->pair_h[I]] = (int32_t)I;
    std::vector<int32_t> occ, vir;
    int nt = 0;
    for (int k0 = 0; k0 < c->no; k0 += TK)
        for (int c0 = 0; c0 < c->nv; c0 += TC) {
            bool any = false;
            for (int k = k0; k < std::min(c->no, k0 + TK) && !any; ++k)
                for (int cc = c0; cc < std::min(c->nv, c0 + TC); ++cc)
                    if (csf[(int64_t)k * c->nv + cc] >= 0) { any = true; break; }
            if (!any) continue;
            for (int k = 0; k < TK; ++k) occ.push_back(k0 + k < c->no ? k0 + k : -1);
            for (int cc = 0; cc < TC; ++cc) vir.push_back(c0 + cc < c->nv ? c0 + cc : -1);
            ++nt;
        }
    DBuf<int32_t> od, vd, csfd;
    od.upload(occ.data(), occ.size(), s);
    vd.upload(vir.data(), vir.size(), s);
    csfd.upload(csf.data(), csf.size(), s);
    A.resize((size_t)n * n);
    A.zero(s);
    launch_build_a(s, TileSet{od.get(), vd.get(), nt}, c->ncent, c->no, c->nv, c->pia.get(),
                   c->qia.get(), c->pij.get(), c->qab.get(), csfd.get(), c->ak, A.get(), n);
    // diagonal = ed (rtdamat: hci(i,i) = ed(i))
    launch_set_diag(s, n, c->ed.get(), A.get());
    c->sync();
}

// (ia|jb)_K for all CSF pairs (n x n) and (ib|ja)_K (n x n), device
static void build_k_and_exchange(stda_hip_ctx *c, DBuf<double> &K, DBuf<double> &E) {
    const int64_t n = c->nci;
    const int na = c->ncent;
    hipStream_t s = c->stream;
    DBuf<double> pq((size_t)na * n), qq((size_t)na * n);
    launch_gather_cols(s, na, n, c->pair.get(), c->pia.get(), pq.get());
    launch_gather_cols(s, na, n, c->pair.get(), c->qia.get(), qq.get());
    K.resize((size_t)n * n);
    dgemm(c->blas, true, false, n, n, na, 1.0, pq.get(), na, qq.get(), na, 0.0, K.get(), n);
    E.resize((size_t)n * n);
    launch_exchange(s, n, na, c->nv, c->pair.get(), c->pia.get(), c->qia.get(), E.get(), n);
    c->sync();
}

static std::vector<double> download_matrix(stda_hip_ctx *c, DBuf<double> &A, int64_t n) {
    std::vector<double> h((size_t)n * n);
    A.download(h.data(), h.size(), c->stream);
    c->sync();
    return h;
}

// TDA / RPA property tables (5 per root) and alpha_SOS
// Property table (fL, fV, RL, RV or Rv(corr), 0 per root) and alpha_SOS from
// the moment products T = mom^T X (9 per root: xl yl zl xm ym zm xv yv zv)
// and, with the velocity correction, T2 = mom^T Xnew.
static void tda_props_from_moments(double ak, const std::vector<double> &w, const double *T,
                                   const double *T2, double *props, double *alpha) {
    const int k = (int)w.size();
    double sak = std::sqrt(ak);
    double a6[6] = {0, 0, 0, 0, 0, 0};
    for (int j = 0; j < k; ++j) {
        const double *t = &T[(size_t)9 * j];
        double de = w[j], ef = 1.0 / (de + 1.0e-8);
        double l[3] = {t[0] * sak, t[1] * sak, t[2] * sak};
        double m[3] = {t[3] * sak, t[4] * sak, t[5] * sak};
        double v[3] = {t[6] * sak, t[7] * sak, t[8] * sak};
        double fl = de * (2.0 / 3.0) * (l[0] * l[0] + l[1] * l[1] + l[2] * l[2]);
        double fv = ef * (2.0 / 3.0) * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        double rl = -kRot * (l[0] * m[0] + l[1] * m[1] + l[2] * m[2]);
        double rv = -kRot * (v[0] * m[0] + v[1] * m[1] + v[2] * m[2]) * ef;
        if (T2) {   // apbtrafo.f:180-262
            const double *u = &T2[(size_t)9 * j];
            double hilf = 1.0 - std::exp(-150.0 * de * de);
            double efmod = 1.0 / (std::pow(de, hilf) + 1.0e-8);
            double s = 0.0;
            for (int q = 0; q < 3; ++q)
                s += t[6 + q] * t[3 + q] + u[6 + q] * t[3 + q] + t[6 + q] * u[3 + q];
            rv = -kRot * s * ak * efmod;
        }
        if (props) {
            props[5 * j + 0] = fl; props[5 * j + 1] = fv; props[5 * j + 2] = rl;
            props[5 * j + 3] = rv; props[5 * j + 4] = 0.0;
        }
        a6[0] += l[0] * l[0] * ef; a6[1] += l[0] * l[1] * ef; a6[2] += l[1] * l[1] * ef;
        a6[3] += l[0] * l[2] * ef; a6[4] += l[1] * l[2] * ef; a6[5] += l[2] * l[2] * ef;
    }
    if (alpha)
        for (int q = 0; q < 6; ++q) alpha[q] = 2.0 * a6[q];
}

static void tda_properties(stda_hip_ctx *c, const double *X, const std::vector<double> &w,
                           int velcorr, double *props, double *alpha) {
    int k = (int)w.size();
    std::vector<double> T = moment_products(c, X, k);
    std::vector<double> T2;
    if (velcorr && c->have_moments && k > 0) {
        DBuf<double> Xn((size_t)c->nci * k);
        apply_ops(c, OP_BHALF, -1, k, X, c->nci, Xn.get(), c->nci, nullptr, 0);
        std::vector<double> sc(k);
        for (int j = 0; j < k; ++j) sc[j] = 1.0 / (w[j] + 1.0e-8);   // apbtrafo.f:63-69
        DBuf<double> scd;
        scd.upload(sc.data(), k, c->stream);
        launch_scale_cols(c->stream, c->nci, k, Xn.get(), scd.get());
        T2 = moment_products(c, Xn.get(), k);
    }
    tda_props_from_moments(c->ak, w, T.data(), T2.empty() ? nullptr : T2.data(), props, alpha);
}

// Mixed precision TDA: FP32 operator to a residual of max(tol, 1e-6), then one
// FP64 Rayleigh-Ritz step in the span of the converged vectors.  With the
// split-bfloat16 operator (precision 2) a first phase runs to 3e-5 (about its
// noise floor) and the FP32 phase starts from its Ritz vectors.
static void solve_tda_mixed(stda_hip_ctx *c, int nroots, double tol, std::vector<double> &w,
                            DBuf<double> &X) {
    DavOpts lo;
    lo.fp32 = true;
    if (c->precision == 2) {
        HostSpan hs(c, "mixed_bf16_phase");
        c->gemm32 = GEMM_BF16X3;
        davidson_tda(c, nroots, std::max(tol, 3.0e-5), w, X, lo);
        c->gemm32 = GEMM_F32;
        lo.V0 = X.get();
        lo.nv0 = (int)w.size();
    }
    {
        HostSpan hs(c, "mixed_fp32_phase");
        DBuf<double> Xb = std::move(X);
        davidson_tda(c, nroots, std::max(tol, 1.0e-6), w, X, lo);
    }
    DavOpts ref;
    ref.V0 = X.get();
    ref.nv0 = (int)w.size();
    ref.maxiter = 1;
    ref.allow_unconverged = true;
    DBuf<double> X0 = std::move(X);
    HostSpan hs(c, "mixed_fp64_refine");
    davidson_tda(c, nroots, tol, w, X, ref);
}

static void rpa_properties(stda_hip_ctx *c, const double *XpY, const double *XmY,
                           const std::vector<double> &w, double *props, double *alpha) {
    int k = (int)w.size();
    std::vector<double> Tp = moment_products(c, XpY, k), Tm = moment_products(c, XmY, k);
    double sak = std::sqrt(c->ak);
    double a6[6] = {0, 0, 0, 0, 0, 0};
    for (int j = 0; j < k; ++j) {
        const double *tp = &Tp[(size_t)9 * j], *tm = &Tm[(size_t)9 * j];
        double de = w[j], ef = 1.0 / (de + 1.0e-8);
        double l[3] = {tp[0] * sak, tp[1] * sak, tp[2] * sak};
        double m[3] = {tm[3] * sak, tm[4] * sak, tm[5] * sak};
        double v[3] = {tm[6] * sak, tm[7] * sak, tm[8] * sak};
        if (props) {
            props[5 * j + 0] = de * (2.0 / 3.0) * (l[0] * l[0] + l[1] * l[1] + l[2] * l[2]);
            props[5 * j + 1] = ef * (2.0 / 3.0) * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
            props[5 * j + 2] = -kRot * (l[0] * m[0] + l[1] * m[1] + l[2] * m[2]);
            props[5 * j + 3] = -kRot * (v[0] * m[0] + v[1] * m[1] + v[2] * m[2]) * ef;
            props[5 * j + 4] = 0.0;
        }
        a6[0] += l[0] * l[0] * ef; a6[1] += l[0] * l[1] * ef; a6[2] += l[1] * l[1] * ef;
        a6[3] += l[0] * l[2] * ef; a6[4] += l[1] * l[2] * ef; a6[5] += l[2] * l[2] * ef;
    }
    if (alpha)
        for (int q = 0; q < 6; ++q) alpha[q] = 2.0 * a6[q];
}

// ============================================================================================
// C ABI
// ------------------------------------------------------------------------------------------
// Fixed device staging area of stda_hip_gspec_noalloc: one copy per GPU, held
// until the process ends.  In host code `scratch` is only a host-side stand-in
// for the symbol; its device address is taken once per device, with the device
// current, by hipGetSymbolAddress at context creation and kept in
// g_scratch_ptr, which is what copies and kernel arguments use.
constexpr size_t kScratchBytes = (size_t)32 << 20;
static __device__ __align__(256) double scratch[kScratchBytes / sizeof(double)];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};

// Address of this device's copy of `scratch` (the device must be current).
void scratch_init(int device) {
    if (device < 0 || device >= kMaxDevices) throw Error(STDA_HIP_ERR_ARG, "device index too large");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        throw Error(STDA_HIP_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") +
                                             hipGetErrorString(e));
    g_scratch_ptr[device] = static_cast<double *>(p);
}

double *scratch_of(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return device >= 0 && device < kMaxDevices ? g_scratch_ptr[device] : nullptr;
}
}  // namespace

#define STDA_TRY(ctx, ...)                                            \
    try {                                                             \
        AllocScope alloc_scope_((ctx)->stream);                       \
        __VA_ARGS__;                                                  \
        return STDA_HIP_OK;                                           \
    } catch (const Error &e) {                                        \
        if (ctx) (ctx)->last_error = e.what();                        \
        return e.status;                                              \
    } catch (const std::exception &e) {                               \
        if (ctx) (ctx)->last_error = e.what();                        \
        return STDA_HIP_ERR_DEVICE;                                   \
    }

extern "C" {

int stda_hip_index_size(void) { return (int)sizeof(stda_hip_int); }

int stda_hip_create(stda_hip_ctx **out, int device) {
    if (!out) return STDA_HIP_ERR_ARG;
    stda_hip_ctx *c = new stda_hip_ctx();
    try {
        if (device >= 0) STDA_HIP_CHECK(hipSetDevice(device));
        STDA_HIP_CHECK(hipGetDevice(&c->device));
        scratch_init(c->device);
        STDA_HIP_CHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        c->owner.stream = c->stream;
        STDA_BLAS_CHECK(rocblas_create_handle(&c->blas));
        c->owner.blas = c->blas;
        STDA_BLAS_CHECK(rocblas_set_stream(c->blas, c->stream));
        STDA_BLAS_CHECK(rocblas_set_pointer_mode(c->blas, rocblas_pointer_mode_host));
        // bitwise-reproducible GEMMs: a job's result must not depend on scheduling
        STDA_BLAS_CHECK(rocblas_set_atomics_mode(c->blas, rocblas_atomics_not_allowed));
        // keep freed blocks in the device pool instead of returning them at
        // every synchronisation
        hipMemPool_t pool;
        STDA_HIP_CHECK(hipDeviceGetDefaultMemPool(&pool, c->device));
        uint64_t keep = UINT64_MAX;
        STDA_HIP_CHECK(hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep));
    } catch (const Error &e) {
        delete c;
        return e.status;
    }
    *out = c;
    return STDA_HIP_OK;
}

int stda_hip_destroy(stda_hip_ctx *c) {
    if (!c) return STDA_HIP_OK;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    delete c;   // buffers first (freed on the stream), then stream and handle (Owner)
    return STDA_HIP_OK;
}

const char *stda_hip_error_string(int st) {
    switch (st) {
        case STDA_HIP_OK: return "ok";
        case STDA_HIP_ERR_ARG: return "invalid argument";
        case STDA_HIP_ERR_NO_CSF: return "no CSF, increase energy threshold (-e option)";
        case STDA_HIP_ERR_DEVICE: return "HIP device or library error";
        case STDA_HIP_ERR_NOMEM: return "device allocation failed";
        case STDA_HIP_ERR_STATE: return "setup must be called before solving";
        case STDA_HIP_ERR_NOT_CONVERGED: return "iterative solver did not converge";
        case STDA_HIP_ERR_AMB_NOT_PD: return "matrix power impossible";
        case STDA_HIP_ERR_NO_ROOTS: return "internal error in diag";
        case STDA_HIP_ERR_TOO_LARGE: return "problem too large for this entry point";
        default: return "unknown status";
    }
}

const char *stda_hip_last_error(const stda_hip_ctx *c) { return c ? c->last_error.c_str() : ""; }

int stda_hip_load(stda_hip_ctx *c, const stda_hip_system *sys, const stda_hip_params *par) {
    if (!c) return STDA_HIP_ERR_ARG;
    STDA_TRY(c, {
        STDA_HIP_CHECK(hipSetDevice(c->device));
        do_load(c, sys, par);
    })
}

int stda_hip_select(stda_hip_ctx *c, int64_t *nci, int64_t *nprim, int64_t *nsec) {
    if (!c) return STDA_HIP_ERR_ARG;
    if (!c->loaded) return STDA_HIP_ERR_STATE;
    STDA_TRY(c, {
        STDA_HIP_CHECK(hipSetDevice(c->device));
        do_select(c);
        if (nci) *nci = c->nci;
        if (nprim) *nprim = c->nprim;
        if (nsec) *nsec = c->nsec;
    })
}

int stda_hip_setup(stda_hip_ctx *c, const stda_hip_system *sys, const stda_hip_params *par,
                   int64_t *nci, int64_t *nprim, int64_t *nsec) {
    int st = stda_hip_load(c, sys, par);
    if (st != STDA_HIP_OK) return st;
    return stda_hip_select(c, nci, nprim, nsec);
}

int stda_hip_kernel_stats(const stda_hip_ctx *c, const char *name, double *ms, double *launches,
                          double *work) {
    if (!c || !name) return STDA_HIP_ERR_ARG;
    const_cast<stda_hip_ctx *>(c)->flush_timers();
    for (const auto &kv : c->stats)
        if (kv.first == name) {
            if (ms) *ms = kv.second.ms;
            if (launches) *launches = kv.second.launches;
            if (work) *work = kv.second.work;
            return STDA_HIP_OK;
        }
    if (ms) *ms = 0;
    if (launches) *launches = 0;
    if (work) *work = 0;
    return STDA_HIP_OK;
}

int stda_hip_get_csf(stda_hip_ctx *c, stda_hip_int *iconf, double *ed) {
    if (!c) return STDA_HIP_ERR_ARG;
    if (!c->ready) return STDA_HIP_ERR_STATE;
    for (int64_t I = 0; I < c->nci; ++I) {
        int64_t p = c->pair_h[I];
        if (iconf) {
            iconf[I] = (stda_hip_int)(p / c->nv) + 1;
            iconf[I + c->nci] = (stda_hip_int)(c->no + p % c->nv) + 1;
        }
        if (ed) ed[I] = c->ed_h[I];
    }
    return STDA_HIP_OK;
}

int stda_hip_solve_tda(stda_hip_ctx *c, int32_t nroots, int32_t capacity, int32_t velcorr,
                       double tol, int32_t *nfound, double *omega, double *x, double *props,
                       double *alpha_sos) {
    if (!c) return STDA_HIP_ERR_ARG;
    if (!c->ready) return STDA_HIP_ERR_STATE;
    STDA_TRY(c, {
        STDA_HIP_CHECK(hipSetDevice(c->device));
        Timer t3;
        if (tol <= 0.0) tol = 1.0e-8;
        std::vector<double> w;
        DBuf<double> X;
        const int64_t n = c->nci;
        if (nroots > 0 && c->precision >= 1) {
            solve_tda_mixed(c, nroots, tol, w, X);
        } else if (nroots > 0) {
            davidson_tda(c, nroots, tol, w, X);
        } else {
            // every root in (0, Ethr] (ssyevr RANGE='V', stda.f:1143-1150)
            int64_t cnt = 0;
            for (double e : c->ed_h) cnt += e <= c->thr;
            int k = (int)std::min<int64_t>(n, cnt + 8);
            while (true) {
                davidson_tda(c, k, tol, w, X);
                if (w.back() > c->thr || k >= n) break;
                k = (int)std::min<int64_t>(n, 2 * (int64_t)k);
            }
            int keep = 0;
            for (double e : w) keep += (e > 0.0 && e <= c->thr);
            int first = 0;
            while (first < (int)w.size() && w[first] <= 0.0) ++first;
            // column by column in ascending order: source and destination of
            // one column are first >= 1 columns apart, so no copy overlaps
            for (int j = 0; first > 0 && j < keep; ++j)
                STDA_HIP_CHECK(hipMemcpyAsync(X.get() + n * j, X.get() + n * (first + j),
                                              (size_t)n * 8, hipMemcpyDeviceToDevice, c->stream));
            w = std::vector<double>(w.begin() + first, w.begin() + first + keep);
        }
        int k = (int)w.size();
        if (k < 1) throw Error(STDA_HIP_ERR_NO_ROOTS, "internal error in diag");
        if (capacity < k) throw Error(STDA_HIP_ERR_ARG, "capacity smaller than the number of roots");
        c->timings[3] = t3.ms();
        Timer t4;
        *nfound = k;
        std::copy(w.begin(), w.end(), omega);
        if (x) {
            X.download(x, (size_t)n * k, c->stream);
            c->sync();
        }
        tda_properties(c, X.get(), w, velcorr, props, alpha_sos);
        c->timings[4] = t4.ms();
    })
}

int stda_hip_solve_rpa(stda_hip_ctx *c, int32_t nroots, int32_t capacity, double tol,
                       int32_t *nfound, double *omega, double *x, double *y, double *props,
                       double *alpha_sos) {
    if (!c) return STDA_HIP_ERR_ARG;
    if (!c->ready) return STDA_HIP_ERR_STATE;
    STDA_TRY(c, {
        STDA_HIP_CHECK(hipSetDevice(c->device));
        Timer t3;
        if (tol <= 0.0) tol = 1.0e-8;
        std::vector<double> w;
        DBuf<double> XpY, XmY;
        const int64_t n = c->nci;
        if (nroots > 0 && c->precision >= 1) {
            DavOpts lo;
            lo.fp32 = true;
            DBuf<double> Vb;
            if (c->precision == 2) {   // split-bfloat16 phase, then FP32 from its vectors
                HostSpan hs(c, "mixed_bf16_phase");
                c->gemm32 = GEMM_BF16X3;
                davidson_rpa(c, nroots, std::max(tol, 3.0e-5), w, XpY, XmY, lo);
                c->gemm32 = GEMM_F32;
                const int kb = (int)w.size();
                Vb.resize((size_t)n * 2 * kb);
                STDA_HIP_CHECK(hipMemcpyAsync(Vb.get(), XpY.get(), (size_t)n * kb * 8, hipMemcpyDeviceToDevice, c->stream));
                STDA_HIP_CHECK(hipMemcpyAsync(Vb.get() + n * kb, XmY.get(), (size_t)n * kb * 8, hipMemcpyDeviceToDevice, c->stream));
                lo.V0 = Vb.get();
                lo.nv0 = 2 * kb;
            }
            {
                HostSpan hs(c, "mixed_fp32_phase");
                davidson_rpa(c, nroots, std::max(tol, 1.0e-6), w, XpY, XmY, lo);
            }
            // refinement: FP64 paired Rayleigh-Ritz in span(X+Y, X-Y)
            int k2 = (int)w.size();
            DBuf<double> V0((size_t)n * 2 * k2);
            STDA_HIP_CHECK(hipMemcpyAsync(V0.get(), XpY.get(), (size_t)n * k2 * 8, hipMemcpyDeviceToDevice, c->stream));
            STDA_HIP_CHECK(hipMemcpyAsync(V0.get() + n * k2, XmY.get(), (size_t)n * k2 * 8, hipMemcpyDeviceToDevice, c->stream));
            DavOpts ref;
            ref.V0 = V0.get();
            ref.nv0 = 2 * k2;
            ref.maxiter = 1;
            ref.allow_unconverged = true;
            HostSpan hs(c, "mixed_fp64_refine");
            davidson_rpa(c, nroots, tol, w, XpY, XmY, ref);
        } else if (nroots > 0) {
            davidson_rpa(c, nroots, tol, w, XpY, XmY);
        } else {
            int64_t cnt = 0;
            for (double e : c->ed_h) cnt += e <= c->thr;
            int k = (int)std::min<int64_t>(n, cnt + 8);
            while (true) {
                davidson_rpa(c, k, tol, w, XpY, XmY);
                if (w.back() > c->thr || k >= n) break;
                k = (int)std::min<int64_t>(n, 2 * (int64_t)k);
            }
            int keep = 0;
            for (double e : w) keep += (e * e <= c->thr * c->thr);   // srpapack.f:84-89
            w.resize(keep);
        }
        int k = (int)w.size();
        if (k < 1) throw Error(STDA_HIP_ERR_NO_ROOTS, "RPA diag failed");
        if (capacity < k) throw Error(STDA_HIP_ERR_ARG, "capacity smaller than the number of roots");
        c->timings[3] = t3.ms();
        Timer t4;
        *nfound = k;
        std::copy(w.begin(), w.end(), omega);
        if (x || y) {   // X = ((X+Y) + (X-Y))/2, Y = (X+Y) - X (srpapack.f:155-162)
            std::vector<double> hp((size_t)n * k), hm((size_t)n * k);
            XpY.download(hp.data(), hp.size(), c->stream);
            XmY.download(hm.data(), hm.size(), c->stream);
            c->sync();
            for (size_t q = 0; q < hp.size(); ++q) {
                double xv = 0.5 * (hp[q] + hm[q]);
                if (x) x[q] = xv;
                if (y) y[q] = hp[q] - xv;
            }
        }
        rpa_properties(c, XpY.get(), XmY.get(), w, props, alpha_sos);
        c->timings[4] = t4.ms();
    })
}

int stda_hip_lresp(stda_hip_ctx *c, int32_t nfreq, const double *omega, double tol,
                   double *alpha, double *xpy_out) {
    if (!c) return STDA_HIP_ERR_ARG;
    if (!c->ready) return STDA_HIP_ERR_STATE;
    if (nfreq < 1 || !omega || !alpha) return STDA_HIP_ERR_ARG;
    if (!c->have_moments) return STDA_HIP_ERR_ARG;
    STDA_TRY(c, {
        STDA_HIP_CHECK(hipSetDevice(c->device));
        Timer t3;
        if (tol <= 0.0) tol = 1.0e-10;
        // Mixed precision: the FP32 operator's noise floor is ~1e-7 of the
        // operator norm times the solution norm, which near resonance exceeds
        // any useful tolerance.  The iteration runs with the FP32 operator until
        // its residuals reach max(tol, 1e-6) or stop falling, then the space's
        // operator images are recomputed in FP64 and it finishes in FP64.
        bool fp32 = c->precision >= 1;
        if (fp32) tol = std::max(tol, 1.0e-8);
        c->gemm32 = c->precision == 2 ? GEMM_BF16X3 : GEMM_F32;
        const double tol32 = std::max(tol, c->precision == 2 ? 3.0e-5 : 1.0e-6);
        double best32 = 1.0e300;
        int since_best = 0;
        const int64_t n = c->nci;
        hipStream_t s = c->stream;
        const int nrhs = 3 * nfreq;
        double wmax = 0.0;
        for (int f = 0; f < nfreq; ++f) wmax = std::max(wmax, std::fabs(omega[f]));
        // right-hand sides -2 mu (linear_response.f:98-116): mu = xl of each CSF
        DBuf<double> rhs((size_t)n * 3);
        STDA_HIP_CHECK(hipMemcpyAsync(rhs.get(), c->mom.get(), (size_t)n * 3 * 8, hipMemcpyDeviceToDevice, s));
        std::vector<double> m2(3, -2.0);
        DBuf<double> m2d;
        m2d.upload(m2.data(), 3, s);
        launch_scale_cols(s, n, 3, rhs.get(), m2d.get());
        // Subspace size.  Above many RPA poles M(w) = (A+B) - w^2 (A-B)^-1 is
        // indefinite and the Galerkin residuals oscillate before they fall;
        // the space that resolves the poles near w must be kept whole, so the
        // space is sized not to restart in practice (a 60 000-CSF molecule
        // with 200 poles below w converges at ~600 vectors): V, P = (A+B)V and
        // M = (A-B)V take 24 n bytes per vector, allowed up to an eighth of
        // the available HBM beyond the runtime's reserve (at most 32 GiB) and
        // 2400 vectors.  They are allocated as the space grows (below).
        const size_t budget = std::min<size_t>(device_available_bytes() / 8, (size_t)32 << 30);
        const int need = 3 + 4 * nrhs;   // initial block, kept solutions, one expansion
        int64_t md = std::min<int64_t>((int64_t)(budget / (24.0 * (double)n)), 2400);
        int maxdim = (int)std::min<int64_t>(n, std::max<int64_t>(md, need));
        // STDA_HIP_LRESP_MAXDIM caps the subspace (tests use it to force restarts)
        if (const char *e = getenv("STDA_HIP_LRESP_MAXDIM"))
            if (atoi(e) > 0) maxdim = (int)std::min<int64_t>(maxdim, std::max(atoi(e), need));
        // V, P and M hold `cap` columns and grow (doubling, up to maxdim) when
        // the space needs more, so that a solve holds the memory it uses and
        // concurrent solves share the device; when the device cannot hold
        // more (STDA_HIP_ERR_NOMEM) the space stops growing and restarts at
        // `limit` = its current size instead
        int cap = (int)std::min<int64_t>(maxdim, std::max(2 * need, 128)), limit = maxdim;
        DBuf<double> V((size_t)n * cap), P((size_t)n * cap), M((size_t)n * cap), tmp;
        DBuf<double> R((size_t)n * 2 * nrhs), XP((size_t)n * nrhs), nrm(2 * nrhs), Yd;
        // initial space: the right-hand sides preconditioned by the diagonal
        STDA_HIP_CHECK(hipMemcpyAsync(V.get(), rhs.get(), (size_t)n * 3 * 8, hipMemcpyDeviceToDevice, s));
        std::vector<double> zero3(3, 0.0);
        DBuf<double> z3;
        z3.upload(zero3.data(), 3, s);
        launch_precondition(s, n, 3, V.get(), c->ed.get(), z3.get(), V.get());
        int m = 0, mnew = orthonormalize(c, V.get(), 0, V.get(), 3, tmp);
        int m0 = mnew;   // the initial block, kept through restarts
        std::vector<double> rhsnorm(3);
        {
            DBuf<double> rn(3);
            col_norms(c, n, 3, rhs.get(), rn.get());
            rn.download(rhsnorm.data(), 3, s);
            c->sync();
        }
        std::vector<double> coef;   // (m x 2 nrhs): X+Y coefficients a | X-Y coefficients b
        bool done = mnew == 0;      // no dipole moments at all: alpha = 0
        std::vector<double> Ap((size_t)maxdim * maxdim), Am((size_t)maxdim * maxdim),
            Br((size_t)maxdim * 3);
        std::vector<double> wcol(nrhs);
        for (int f = 0; f < nfreq; ++f)
            for (int r = 0; r < 3; ++r) wcol[f * 3 + r] = omega[f];
        DBuf<double> wd;
        wd.upload(wcol.data(), nrhs, s);
        std::vector<double> onev(3, 1.0);
        DBuf<double> oned;
        oned.upload(onev.data(), 3, s);
        auto grow = [&](int want, int keep) {   // room for `want` columns, the first `keep` kept
            int nc = cap;
            while (nc < want && nc < maxdim) nc = (int)std::min<int64_t>(maxdim, 2 * (int64_t)nc);
            if (nc <= cap) return;
            try {
                DBuf<double> nV((size_t)n * nc), nP((size_t)n * nc), nM((size_t)n * nc);
                const size_t bytes = (size_t)n * keep * 8;
                if (bytes) {
                    STDA_HIP_CHECK(hipMemcpyAsync(nV.get(), V.get(), bytes, hipMemcpyDeviceToDevice, s));
                    STDA_HIP_CHECK(hipMemcpyAsync(nP.get(), P.get(), bytes, hipMemcpyDeviceToDevice, s));
                    STDA_HIP_CHECK(hipMemcpyAsync(nM.get(), M.get(), bytes, hipMemcpyDeviceToDevice, s));
                }
                V = std::move(nV);
                P = std::move(nP);
                M = std::move(nM);
                cap = nc;
            } catch (const Error &e) {
                if (e.status != STDA_HIP_ERR_NOMEM) throw;
                limit = cap;
            }
        };
        for (int it = 0; it < 500 && !done; ++it) {
            if (mnew > 0) {   // operator images and reduced matrices of the new block
                apply_ops(c, OP_APB, OP_AMB, mnew, V.get() + n * m, n, P.get() + n * m, n,
                          M.get() + n * m, n, fp32);
                const int mt = m + mnew;
                DBuf<double> Hc((size_t)mt * mnew * 2 + (size_t)mnew * 3);
                dgemm_tn(c, mt, mnew, n, V.get(), n, P.get() + n * m, n, Hc.get());
                dgemm_tn(c, mt, mnew, n, V.get(), n, M.get() + n * m, n, Hc.get() + (size_t)mt * mnew);
                dgemm_tn(c, mnew, 3, n, V.get() + n * m, n, rhs.get(), n, Hc.get() + (size_t)2 * mt * mnew);
                std::vector<double> hc((size_t)mt * mnew * 2 + (size_t)mnew * 3);
                Hc.download(hc.data(), hc.size(), s);
                c->sync();
                for (int j = 0; j < mnew; ++j)
                    for (int i = 0; i < mt; ++i) {
                        double vp = hc[i + (size_t)mt * j], vm = hc[(size_t)mt * mnew + i + (size_t)mt * j];
                        Ap[i + (size_t)maxdim * (m + j)] = vp; Ap[(m + j) + (size_t)maxdim * i] = vp;
                        Am[i + (size_t)maxdim * (m + j)] = vm; Am[(m + j) + (size_t)maxdim * i] = vm;
                    }
                for (int r = 0; r < 3; ++r)
                    for (int j = 0; j < mnew; ++j) Br[(m + j) + (size_t)maxdim * r] = hc[(size_t)2 * mt * mnew + j + (size_t)mnew * r];
                m = mt;
            }
            // Reduced problem, one factorisation for every frequency: with
            // Am = L L^T and L^T Ap L = Z diag(lam) Z^T (lam: the space's RPA
            // w_n^2), (Ap - w^2 Am^-1) a = V^T r gives a = L Z c with
            // c = (lam - w^2)^-1 Z^T L^T V^T r, and b = w Am^-1 a = w L^-T Z c.
            std::vector<double> ap((size_t)m * m), am((size_t)m * m);
            for (int j = 0; j < m; ++j)
                for (int i = 0; i < m; ++i) {
                    ap[i + (size_t)m * j] = 0.5 * (Ap[i + (size_t)maxdim * j] + Ap[j + (size_t)maxdim * i]);
                    am[i + (size_t)m * j] = 0.5 * (Am[i + (size_t)maxdim * j] + Am[j + (size_t)maxdim * i]);
                }
            std::vector<double> L(am);
            if (!host_potrf_lower(m, L)) throw Error(STDA_HIP_ERR_AMB_NOT_PD, "A-B not positive definite");
            std::vector<double> t1((size_t)m * m), Z((size_t)m * m), lam;
            host_gemm(false, false, m, m, m, 1.0, ap.data(), m, L.data(), m, 0.0, t1.data(), m);
            host_gemm(true, false, m, m, m, 1.0, L.data(), m, t1.data(), m, 0.0, Z.data(), m);
            for (int j = 0; j < m; ++j)
                for (int i = 0; i < j; ++i) {
                    double v = 0.5 * (Z[i + (size_t)m * j] + Z[j + (size_t)m * i]);
                    Z[i + (size_t)m * j] = v; Z[j + (size_t)m * i] = v;
                }
            host_syev(m, Z, lam);
            std::vector<double> W((size_t)m * m), Yz(Z), G((size_t)m * 3);
            host_gemm(false, false, m, m, m, 1.0, L.data(), m, Z.data(), m, 0.0, W.data(), m);
            host_solve_lt(m, m, L, Yz.data(), m);
            host_gemm(true, false, m, 3, m, 1.0, W.data(), m, Br.data(), maxdim, 0.0, G.data(), m);
            const double lamscale = std::max(1.0, std::fabs(lam[m - 1]));
            coef.assign((size_t)m * 2 * nrhs, 0.0);
            std::vector<double> cv((size_t)m * 3);
            for (int f = 0; f < nfreq; ++f) {
                const double w = omega[f];
                for (int r = 0; r < 3; ++r)
                    for (int q = 0; q < m; ++q) {
                        double den = lam[q] - w * w;
                        if (std::fabs(den) < 1.0e-300 * lamscale) den = den < 0.0 ? -1.0e-300 : 1.0e-300;
                        cv[q + (size_t)m * r] = G[q + (size_t)m * r] / den;
                    }
                host_gemm(false, false, m, 3, m, 1.0, W.data(), m, cv.data(), m, 0.0,
                          coef.data() + (size_t)m * 3 * f, m);
                host_gemm(false, false, m, 3, m, w, Yz.data(), m, cv.data(), m, 0.0,
                          coef.data() + (size_t)m * (nrhs + 3 * f), m);
            }
            // residuals R1 = P a - w V b - r ; R2 = M b - w V a
            Yd.upload(coef.data(), coef.size(), s);
            DBuf<double> VB((size_t)n * nrhs), VA((size_t)n * nrhs);
            gemm_nn(c, n, nrhs, m, 1.0, P.get(), n, Yd.get(), m, 0.0, R.get(), n);
            gemm_nn(c, n, nrhs, m, 1.0, V.get(), n, Yd.get() + (size_t)m * nrhs, m, 0.0, VB.get(), n);
            gemm_nn(c, n, nrhs, m, 1.0, V.get(), n, Yd.get(), m, 0.0, VA.get(), n);
            gemm_nn(c, n, nrhs, m, 1.0, M.get(), n, Yd.get() + (size_t)m * nrhs, m, 0.0,
                    R.get() + n * nrhs, n);
            launch_residual(s, n, nrhs, R.get(), VB.get(), wd.get(), R.get());
            launch_residual(s, n, nrhs, R.get() + n * nrhs, VA.get(), wd.get(), R.get() + n * nrhs);
            for (int f = 0; f < nfreq; ++f)   // subtract the right-hand sides: R1 -= r
                launch_residual(s, n, 3, R.get() + n * 3 * f, rhs.get(), oned.get(),
                                R.get() + n * 3 * f);
            col_norms(c, n, 2 * nrhs, R.get(), nrm.get());
            std::vector<double> rn(2 * nrhs);
            nrm.download(rn.data(), rn.size(), s);
            c->sync();
            std::vector<int> un;
            const double tol_now = fp32 ? tol32 : tol;
            double worst = 0.0;
            for (int j = 0; j < 2 * nrhs; ++j) {
                const double lim = tol_now * std::max(1.0, rhsnorm[j % 3]);
                if (!(rn[j] < lim)) un.push_back(j);
                worst = std::max(worst, std::isfinite(rn[j]) ? rn[j] / lim : 1.0e300);
            }
            if (fp32) {
                if (worst < 0.5 * best32) { best32 = worst; since_best = 0; }
                else ++since_best;
            }
            const bool switch64 = fp32 && ((un.empty() && tol32 > tol) || since_best >= 8 ||
                                           (m + (int)un.size() > (3 * limit) / 4 && m < n));
            if (switch64) {
                // FP64 images of the whole space; V^T r is exact already
                fp32 = false;
                c->gemm32 = GEMM_F32;
                apply_ops(c, OP_APB, OP_AMB, m, V.get(), n, P.get(), n, M.get(), n, false);
                DBuf<double> Hf((size_t)m * m * 2);
                dgemm_tn(c, m, m, n, V.get(), n, P.get(), n, Hf.get());
                dgemm_tn(c, m, m, n, V.get(), n, M.get(), n, Hf.get() + (size_t)m * m);
                std::vector<double> hf((size_t)m * m * 2);
                Hf.download(hf.data(), hf.size(), s);
                c->sync();
                for (int j = 0; j < m; ++j)
                    for (int i = 0; i < m; ++i) {
                        Ap[i + (size_t)maxdim * j] = hf[i + (size_t)m * j];
                        Am[i + (size_t)maxdim * j] = hf[(size_t)m * m + i + (size_t)m * j];
                    }
                mnew = 0;
                continue;
            }
            if (un.empty() || m == n) { done = true; break; }
            // 2x2 diagonal preconditioner of the coupled equations, then keep
            // the unconverged columns
            launch_precond_pair(s, n, nrhs, R.get(), R.get() + n * nrhs, c->ed.get(), wd.get());
            int nun = (int)un.size();
            for (int q = 0; q < nun; ++q)
                if (un[q] != q)
                    STDA_HIP_CHECK(hipMemcpyAsync(R.get() + n * q, R.get() + n * un[q], n * 8, hipMemcpyDeviceToDevice, s));
            if (m + nun > cap) grow(m + nun, m);
            if (m + nun > limit) {
                // Restart within half the space: the initial block, the current
                // X+Y and X-Y of every right-hand side, then the space's RPA
                // pairs (X+Y = L z, X-Y ~ L^-T z) from the lowest up to 1.25
                // max(w), orthonormalised in the reduced space.  V is orthonormal,
                // so V Q is; the operator images are P Q and M Q and the reduced
                // matrices Q^T Ap Q and Q^T Am Q.
                std::vector<double> Q;
                auto add_col = [&](const double *x0) {
                    std::vector<double> x(x0, x0 + m);
                    double n0x = 0.0;
                    for (double v : x) n0x += v * v;
                    n0x = std::sqrt(n0x);
                    if (!(n0x > 0.0)) return;
                    for (double &v : x) v /= n0x;
                    const int r = (int)(Q.size() / m);
                    for (int pass = 0; pass < 2; ++pass)
                        for (int q = 0; q < r; ++q) {
                            double d = 0.0;
                            for (int i = 0; i < m; ++i) d += Q[i + (size_t)m * q] * x[i];
                            for (int i = 0; i < m; ++i) x[i] -= d * Q[i + (size_t)m * q];
                        }
                    double nx = 0.0;
                    for (double v : x) nx += v * v;
                    nx = std::sqrt(nx);
                    if (!(nx > 1.0e-8)) return;
                    for (double v : x) Q.push_back(v / nx);
                };
                std::vector<double> e(m, 0.0);
                for (int i = 0; i < m0; ++i) {
                    e[i] = 1.0;
                    add_col(e.data());
                    e[i] = 0.0;
                }
                for (int col = 0; col < 2 * nrhs; ++col) add_col(coef.data() + (size_t)m * col);
                const int room = std::max(0, (limit / 2 - m0 - 2 * nrhs) / 2);
                for (int q = 0, added = 0; q < m && added < room && wmax > 0.0; ++q) {
                    if (!(lam[q] > 0.0) || lam[q] > 1.5625 * wmax * wmax) break;
                    add_col(W.data() + (size_t)m * q);
                    add_col(Yz.data() + (size_t)m * q);
                    ++added;
                }
                const int r = (int)(Q.size() / m);
                DBuf<double> Qd, rs((size_t)n * r);
                Qd.upload(Q.data(), Q.size(), s);
                for (DBuf<double> *B : {&V, &P, &M}) {
                    gemm_nn(c, n, r, m, 1.0, B->get(), n, Qd.get(), m, 0.0, rs.get(), n);
                    STDA_HIP_CHECK(hipMemcpyAsync(B->get(), rs.get(), (size_t)n * r * 8,
                                                  hipMemcpyDeviceToDevice, s));
                }
                auto reduce = [&](const std::vector<double> &a, std::vector<double> &out) {
                    std::vector<double> aq((size_t)m * r), qaq((size_t)r * r);
                    host_gemm(false, false, m, r, m, 1.0, a.data(), m, Q.data(), m, 0.0, aq.data(), m);
                    host_gemm(true, false, r, r, m, 1.0, Q.data(), m, aq.data(), m, 0.0, qaq.data(), r);
                    for (int j = 0; j < r; ++j)
                        for (int i = 0; i < r; ++i) out[i + (size_t)maxdim * j] = qaq[i + (size_t)r * j];
                };
                reduce(ap, Ap);
                reduce(am, Am);
                std::vector<double> br((size_t)maxdim * 3, 0.0);
                host_gemm(true, false, r, 3, m, 1.0, Q.data(), m, Br.data(), maxdim, 0.0, br.data(), maxdim);
                Br.swap(br);
                m = r;
                m0 = std::min(m0, r);
                if (m + nun > limit) nun = limit - m;
            }
            if (nun <= 0) throw Error(STDA_HIP_ERR_NOT_CONVERGED, "response subspace exhausted");
            STDA_HIP_CHECK(hipMemcpyAsync(V.get() + n * m, R.get(), (size_t)n * nun * 8, hipMemcpyDeviceToDevice, s));
            mnew = orthonormalize(c, V.get(), m, V.get() + n * m, nun, tmp);
            // Every new direction already lies in the space although the
            // residuals are above the tolerance: the iteration cannot progress.
            if (mnew == 0)
                throw Error(STDA_HIP_ERR_NOT_CONVERGED,
                            "response solver did not converge (no new directions)");
        }
        if (!done) throw Error(STDA_HIP_ERR_NOT_CONVERGED, "response solver did not converge");
        // alpha_rc = -2 mu_r^T (X+Y)_c (linear_response.f:147-172)
        if (m > 0) {
            Yd.upload(coef.data(), (size_t)m * nrhs, s);
            gemm_nn(c, n, nrhs, m, 1.0, V.get(), n, Yd.get(), m, 0.0, XP.get(), n);
        } else {
            STDA_HIP_CHECK(hipMemsetAsync(XP.get(), 0, (size_t)n * nrhs * 8, s));
        }
        DBuf<double> Ad((size_t)3 * nrhs);
        dgemm_tn(c, 3, nrhs, n, c->mom.get(), n, XP.get(), n, Ad.get());
        std::vector<double> ah((size_t)3 * nrhs);
        Ad.download(ah.data(), ah.size(), s);
        c->sync();
        for (auto &v : ah) v *= -2.0;
        if (xpy_out) XP.download(xpy_out, (size_t)n * nrhs, s);
        c->sync();
        for (int f = 0; f < nfreq; ++f)
            for (int col = 0; col < 3; ++col)
                for (int r = 0; r < 3; ++r) alpha[r + 3 * col + 9 * f] = ah[r + 3 * (f * 3 + col)];
        c->timings[3] = t3.ms();
    })
}

int stda_hip_build_a(stda_hip_ctx *c, double *a) {
    if (!c || !a) return STDA_HIP_ERR_ARG;
    if (!c->ready) return STDA_HIP_ERR_STATE;
    STDA_TRY(c, {
        STDA_HIP_CHECK(hipSetDevice(c->device));
        if (c->nci > 60000) throw Error(STDA_HIP_ERR_TOO_LARGE, "explicit A limited to nci <= 60000");
        DBuf<double> A;
        build_a_device(c, A);
        A.download(a, (size_t)c->nci * c->nci, c->stream);
        c->sync();
    })
}

int stda_hip_build_apb_amb(stda_hip_ctx *c, double *apb, double *amb) {
    if (!c || !apb || !amb) return STDA_HIP_ERR_ARG;
    if (!c->ready) return STDA_HIP_ERR_STATE;
    STDA_TRY(c, {
        STDA_HIP_CHECK(hipSetDevice(c->device));
        const int64_t n = c->nci;
        if (n > 20000) throw Error(STDA_HIP_ERR_TOO_LARGE, "explicit A+-B limited to nci <= 20000");
        DBuf<double> A, K, E;
        build_a_device(c, A);
        build_k_and_exchange(c, K, E);
        std::vector<double> a = download_matrix(c, A, n), k = download_matrix(c, K, n),
                            e = download_matrix(c, E, n);
        const double ak = c->ak, ax = c->ax;
        for (int64_t i = 0; i < n; ++i) {   // rrpamat, stda.f:1896-1962
            for (int64_t j = 0; j < i; ++j) {
                int64_t ij = i * (i + 1) / 2 + j;
                double kk = k[i + n * j], ee = e[i + n * j];
                double jj = ak * kk - a[i + n * j];   // (ij|ab)
                if (c->small_ax) { apb[ij] = 2.0 * ak * kk; amb[ij] = 0.0; }
                else { apb[ij] = 2.0 * ak * kk - ax * ee - jj; amb[ij] = ax * ee - jj; }
            }
            int64_t ii = i * (i + 1) / 2 + i;
            double kd = k[i + n * i];
            if (c->small_ax) { amb[ii] = c->ed_h[i] - ak * kd; apb[ii] = c->ed_h[i] + ak * kd; }
            else { amb[ii] = c->ed_h[i] - ak * kd + ax * kd; apb[ii] = c->ed_h[i] - ax * kd + ak * kd; }
        }
    })
}

int stda_hip_build_bhalf(stda_hip_ctx *c, double *bh) {
    if (!c || !bh) return STDA_HIP_ERR_ARG;
    if (!c->ready) return STDA_HIP_ERR_STATE;
    STDA_TRY(c, {
        STDA_HIP_CHECK(hipSetDevice(c->device));
        const int64_t n = c->nci;
        if (n > 20000) throw Error(STDA_HIP_ERR_TOO_LARGE, "explicit 0.5B limited to nci <= 20000");
        DBuf<double> K, E;
        build_k_and_exchange(c, K, E);
        std::vector<double> k = download_matrix(c, K, n), e = download_matrix(c, E, n);
        for (int64_t i = 0; i < n; ++i) {   // rtdacorr, apbtrafo.f:604-625
            for (int64_t j = 0; j < i; ++j)
                bh[i * (i + 1) / 2 + j] = 0.5 * c->ak * k[i + n * j] - 0.5 * c->ax * e[i + n * j];
            double kd = k[i + n * i];
            bh[i * (i + 1) / 2 + i] = 0.5 * (c->ak * kd - c->ax * kd);
        }
    })
}

int stda_hip_matvec(stda_hip_ctx *c, int32_t which, int32_t nvec, const double *v, double *av) {
    if (!c || !v || !av || which < 0 || which > 3 || nvec < 0) return STDA_HIP_ERR_ARG;
    if (!c->ready) return STDA_HIP_ERR_STATE;
    STDA_TRY(c, {
        STDA_HIP_CHECK(hipSetDevice(c->device));
        const int64_t n = c->nci;
        DBuf<double> V, AV((size_t)n * nvec);
        V.upload(v, (size_t)n * nvec, c->stream);
        apply_ops(c, which, -1, nvec, V.get(), n, AV.get(), n, nullptr, 0);
        AV.download(av, (size_t)n * nvec, c->stream);
        c->sync();
    })
}

// <beta^2_ZZZ>, <beta^2_XZZ>, beta_HRS and DR of a 3x3x3 tensor b[i + 3j + 9k]
// (HRS, linear_response.f:1489-1662); the coefficients are quotients of
// default-real literals in the reference and keep their single-precision values.
static void beta_hrs(const double *b, double out[4]) {
    auto B = [&](int i, int j, int k) { return b[i + 3 * j + 9 * k]; };
    auto cf = [](int num, int den) { return (double)((float)num / (float)den); };
    double iii = 0, iij = 0, jii = 0, ijk = 0, iii_ijj = 0, jii_iij = 0, iii_jji = 0,
           iij_jkk = 0, jii_jkk = 0, iij_kkj = 0, ijk_jik = 0, ijj = 0, iij_jii = 0,
           ijj_ikk = 0, iik_jjk = 0;
    for (int i = 0; i < 3; ++i) iii += B(i, i, i) * B(i, i, i);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            if (i == j) continue;
            iij += B(i, i, j) * B(i, i, j);
            jii += B(j, i, i) * B(j, i, i);
            iii_ijj += B(i, i, i) * B(i, j, j);
            jii_iij += B(j, i, i) * B(i, i, j);
            iii_jji += B(i, i, i) * B(j, j, i);
            ijj += B(i, j, j) * B(i, j, j);
            iij_jii += B(i, i, j) * B(j, i, i);
            for (int k = 0; k < 3; ++k) {
                if (k == i || k == j) continue;
                ijk += B(i, j, k) * B(i, j, k);
                iij_jkk += B(i, i, j) * B(j, k, k);
                jii_jkk += B(j, i, i) * B(j, k, k);
                iij_kkj += B(i, i, j) * B(k, k, j);
                ijk_jik += B(i, j, k) * B(j, i, k);
                ijj_ikk += B(i, j, j) * B(i, k, k);
                iik_jjk += B(i, i, k) * B(j, j, k);
            }
        }
    const double zzz = cf(1, 7) * iii + cf(4, 35) * iij + cf(1, 35) * jii + cf(2, 105) * ijk +
                       cf(2, 35) * iii_ijj + cf(4, 35) * jii_iij + cf(4, 35) * iii_jji +
                       cf(4, 105) * iij_jkk + cf(1, 105) * jii_jkk + cf(4, 105) * iij_kkj +
                       cf(4, 105) * ijk_jik;
    const double xzz = cf(1, 35) * iii + cf(4, 105) * iii_ijj - cf(2, 35) * iii_jji +
                       cf(8, 105) * iij - cf(2, 105) * iij_jkk + cf(2, 35) * ijk -
                       cf(2, 105) * ijk_jik + cf(3, 35) * ijj - cf(2, 35) * iij_jii +
                       cf(1, 35) * ijj_ikk - cf(2, 105) * iik_jjk;
    out[0] = std::sqrt(zzz + xzz);
    out[1] = zzz / xzz;
    out[2] = zzz;
    out[3] = xzz;
}

int stda_hip_beta_shg(stda_hip_ctx *c, int32_t nfreq, const double *omega, double tol,
                      double *beta, double *hrs) {
    if (!c) return STDA_HIP_ERR_ARG;
    if (!c->ready) return STDA_HIP_ERR_STATE;
    if (nfreq < 1 || !omega || !beta || !c->have_moments) return STDA_HIP_ERR_ARG;
    STDA_TRY(c, {
        STDA_HIP_CHECK(hipSetDevice(c->device));
        const int64_t n = c->nci, L = (int64_t)c->no * c->nv;
        const int no = c->no, nv = c->nv;
        hipStream_t s = c->stream;
        // linear responses at w and 2w of every frequency (linear_response.f:306-323;
        // -2w there: the sign enters only X-Y)
        std::vector<double> wl;
        std::vector<int> at(2 * (size_t)nfreq);
        for (int f = 0; f < nfreq; ++f)
            for (int jj = 0; jj < 2; ++jj) {
                const double w = std::fabs(omega[f]) * (jj + 1);
                size_t q = 0;
                while (q < wl.size() && wl[q] != w) ++q;
                if (q == wl.size()) wl.push_back(w);
                at[2 * f + jj] = (int)q;
            }
        const int nw = (int)wl.size();
        std::vector<double> al((size_t)9 * nw), xpy((size_t)n * 3 * nw);
        const int st = stda_hip_lresp(c, nw, wl.data(), tol, al.data(), xpy.data());
        if (st != STDA_HIP_OK) throw Error(st, c->last_error);
        DBuf<double> XP, AX((size_t)n * 3 * nw), rhs((size_t)n * 3);
        XP.upload(xpy.data(), (size_t)n * 3 * nw, s);
        apply_ops(c, OP_APB, -1, 3 * nw, XP.get(), n, AX.get(), n, nullptr, 0);
        STDA_HIP_CHECK(hipMemcpyAsync(rhs.get(), c->mom.get(), (size_t)n * 3 * 8,
                                      hipMemcpyDeviceToDevice, s));
        std::vector<double> m2(3, -2.0);
        DBuf<double> m2d;
        m2d.upload(m2.data(), 3, s);
        launch_scale_cols(s, n, 3, rhs.get(), m2d.get());
        DBuf<double> muoo((size_t)3 * no * no), muvv((size_t)3 * nv * nv);
        launch_mu_blocks(s, c->moments.get(), c->npk, no, nv, muoo.get(), muvv.get());
        DBuf<double> Xd((size_t)6 * L), Yd((size_t)6 * L), M((size_t)18 * L), T(6 * 18);
        std::vector<double> Th(6 * 18);
        for (int f = 0; f < nfreq; ++f) {
            EventTimer ev(c);
            BetaCols cols;
            cols.scale = 1.0;
            for (int jj = 0; jj < 2; ++jj)
                for (int r = 0; r < 3; ++r) {
                    // signed frequency of column (jj, r): w, then -2w (linear_response.f:308)
                    const double ws = jj == 0 ? omega[f] : -2.0 * omega[f];
                    cols.src[3 * jj + r] = 3 * (int64_t)at[2 * f + jj] + r;
                    cols.inv_w[3 * jj + r] = ws == 0.0 ? 0.0 : 1.0 / ws;
                }
            Xd.zero(s);
            Yd.zero(s);
            launch_beta_xy(s, n, c->pair.get(), no, nv, XP.get(), AX.get(), rhs.get(), cols, 6,
                           Xd.get(), Yd.get());
            // M(q; jj, r) = mu_q^oo Y(jj, r) - Y(jj, r) mu_q^vv, column q*6 + 3*jj + r
            for (int q = 0; q < 3; ++q) {
                double *Mq = M.get() + (size_t)q * 6 * L;
                dgemm_batched(c->blas, false, false, no, nv, no, 1.0, muoo.get() + (size_t)q * no * no,
                              no, 0, Yd.get(), no, L, 0.0, Mq, no, L, 6);
                dgemm_batched(c->blas, false, false, no, nv, nv, -1.0, Yd.get(), no, L,
                              muvv.get() + (size_t)q * nv * nv, nv, 0, 1.0, Mq, no, L, 6);
            }
            // T = Xd^T M: <X(jj', x), M(q; jj, r)> for all 6 x 18 pairs
            dgemm(c->blas, true, false, 6, 18, L, 1.0, Xd.get(), L, M.get(), L, 0.0, T.get(), 6);
            ev.finish("beta", 18.0 * 2.0 * ((double)no * no * nv + (double)no * nv * nv) +
                                  2.0 * 6 * 18 * (double)L);
            T.download(Th.data(), 6 * 18, s);
            c->sync();
            // beta(ix,iy,iz) = A - B over the six orderings of beta_resp_fast
            // (linear_response.f:911-1087): term (x, y, z; w_x, w_z) = -<X_x(w_x), M(y; w_z, z)>
            auto term = [&](int x, int y, int z, int wx, int wz) {
                return -Th[(3 * wx + x) + 6 * (y * 6 + 3 * wz + z)];
            };
            double *bf = beta + (size_t)27 * f;
            for (int iz = 0; iz < 3; ++iz)
                for (int iy = 0; iy < 3; ++iy)
                    for (int ix = 0; ix < 3; ++ix)
                        bf[ix + 3 * iy + 9 * iz] = term(ix, iy, iz, 1, 0) + term(ix, iz, iy, 1, 0) +
                                                   term(iy, iz, ix, 0, 1) + term(iy, ix, iz, 0, 0) +
                                                   term(iz, ix, iy, 0, 0) + term(iz, iy, ix, 0, 1);
            if (hrs) beta_hrs(bf, hrs + 4 * (size_t)f);
        }
    })
}

int stda_hip_props_xyz(stda_hip_ctx *c, int32_t k, const double *w, const double *x, int64_t ldx,
                       int32_t velcorr, double *out) {
    if (!c) return STDA_HIP_ERR_ARG;
    if (!c->ready) return STDA_HIP_ERR_STATE;
    if (k < 1 || !w || !x || !out || ldx < c->nci || !c->have_moments) return STDA_HIP_ERR_ARG;
    STDA_TRY(c, {
        STDA_HIP_CHECK(hipSetDevice(c->device));
        const int64_t n = c->nci;
        hipStream_t s = c->stream;
        DBuf<double> X((size_t)n * k);
        for (int j = 0; j < k; ++j)
            STDA_HIP_CHECK(hipMemcpyAsync(X.get() + n * j, x + ldx * j, (size_t)n * 8,
                                          hipMemcpyHostToDevice, s));
        std::vector<double> wv(w, w + k);
        std::vector<double> T = moment_products(c, X.get(), k), T2;
        if (velcorr) {   // the B-corrected vectors of apbtrafo (apbtrafo.f:23-69)
            DBuf<double> Xn((size_t)n * k);
            apply_ops(c, OP_BHALF, -1, k, X.get(), n, Xn.get(), n, nullptr, 0);
            std::vector<double> sc(k);
            for (int j = 0; j < k; ++j) sc[j] = 1.0 / (wv[j] + 1.0e-8);
            DBuf<double> scd;
            scd.upload(sc.data(), k, s);
            launch_scale_cols(s, n, k, Xn.get(), scd.get());
            T2 = moment_products(c, Xn.get(), k);
        }
        // per axis q: f_L, f_V, R_L, R_V or Rv(corr), scaled so that f = (f_x + f_y + f_z)/3
        // and R = R_x + R_y + R_z (stda.f:1360-1376; apbtrafo.f:245-262)
        const double ak = c->ak, sak = std::sqrt(ak);
        for (int j = 0; j < k; ++j) {
            const double *t = &T[(size_t)9 * j];
            const double de = wv[j], ef = 1.0 / (de + 1.0e-8);
            const double hilf = 1.0 - std::exp(-150.0 * de * de);
            const double efmod = 1.0 / (std::pow(de, hilf) + 1.0e-8);
            for (int q = 0; q < 3; ++q) {
                const double l = t[q] * sak, m = t[3 + q] * sak, v = t[6 + q] * sak;
                double *o = out + 12 * (size_t)j + 4 * q;
                o[0] = l * l * (de * 2.0);
                o[1] = v * v * (ef * 2.0);
                o[2] = -kRot * l * m;
                if (velcorr) {
                    const double *u = &T2[(size_t)9 * j];
                    o[3] = -kRot * (t[6 + q] * t[3 + q] + u[6 + q] * t[3 + q] + t[6 + q] * u[3 + q]) *
                           ak * efmod;
                } else {
                    o[3] = -kRot * v * m * ef;
                }
            }
        }
    })
}

int stda_hip_ao_ints(stda_hip_ctx *c, int32_t nao, const int64_t *prim_ptr, const double *alpha,
                     const double *coef, const double *center, const int32_t *lmn, double *S,
                     double *ao9) {
    if (!c || nao < 1 || !prim_ptr || !alpha || !coef || !center || !lmn || !S || !ao9)
        return STDA_HIP_ERR_ARG;
    if (prim_ptr[0] != 0) return STDA_HIP_ERR_ARG;
    for (int32_t m = 0; m < nao; ++m) {
        if (prim_ptr[m + 1] <= prim_ptr[m]) return STDA_HIP_ERR_ARG;
        for (int d = 0; d < 3; ++d)
            if (lmn[3 * m + d] < 0 || lmn[3 * m + d] > AO_LMAX) return STDA_HIP_ERR_ARG;
    }
    const int64_t nprim = prim_ptr[nao];
    STDA_TRY(c, {
        STDA_HIP_CHECK(hipSetDevice(c->device));
        hipStream_t s = c->stream;
        const size_t nn = (size_t)nao * nao;
        DBuf<int64_t> P;
        DBuf<double> AL, CO, CE, SO(nn), AO((size_t)9 * nn);
        DBuf<int32_t> LM;
        P.upload(prim_ptr, (size_t)nao + 1, s);
        AL.upload(alpha, (size_t)nprim, s);
        CO.upload(coef, (size_t)nprim, s);
        CE.upload(center, (size_t)3 * nao, s);
        LM.upload(lmn, (size_t)3 * nao, s);
        EventTimer ev(c);
        launch_ao_ints(s, nao, P.get(), AL.get(), CO.get(), CE.get(), LM.get(), SO.get(), AO.get());
        ev.finish("ao_ints", (double)nn);
        SO.download(S, nn, s);
        AO.download(ao9, 9 * nn, s);
        c->sync();
    })
}

// Moments and Loewdin orbitals from S and the nine AO matrices, which are host
// arrays (kind = hipMemcpyHostToDevice) or device buffers (hipMemcpyDeviceToDevice).
static void transform_impl(stda_hip_ctx *c, int32_t nao, int32_t moci, const double *ca,
                           const double *S, const double *ao9, hipMemcpyKind kind, int32_t bfw,
                           double *clow, double *moments) {
    hipStream_t s = c->stream;
    const int64_t na = nao, nm = moci, npk = nm * (nm + 1) / 2;
    DBuf<double> CA, M((size_t)na * na), T((size_t)na * nm), P((size_t)nm * nm);
    DBuf<double> OUT((size_t)9 * npk);
    CA.upload(ca, (size_t)na * nm, s);
    // moments: shrink(ca^T M_q ca) for the 9 AO matrices (stda.f:351-412, onetri.f:20-55)
    for (int q = 0; q < 9; ++q) {
        STDA_HIP_CHECK(hipMemcpyAsync(M.get(), ao9 + (size_t)q * na * na, (size_t)na * na * 8,
                                      kind, s));
        dgemm(c->blas, false, false, na, nm, na, 1.0, M.get(), na, CA.get(), na, 0.0, T.get(), na);
        dgemm(c->blas, true, false, nm, nm, na, 1.0, CA.get(), na, T.get(), na, 0.0, P.get(), nm);
        launch_pack_upper(s, nm, P.get(), OUT.get() + (size_t)q * npk);
    }
    OUT.download(moments, (size_t)9 * npk, s);
    // Loewdin: clow = S^1/2 ca = U (sqrt(e) o (U^T ca)) (makel, stda.f:1615-1691)
    STDA_HIP_CHECK(hipMemcpyAsync(M.get(), S, (size_t)na * na * 8, kind, s));
    DBuf<double> D(na), E(na);
    DBuf<int> info(1);
    STDA_BLAS_CHECK(rocsolver_dsyevd(c->blas, rocblas_evect_original, rocblas_fill_upper,
                                     (rocblas_int)na, M.get(), (rocblas_int)na, D.get(),
                                     E.get(), info.get()));
    std::vector<double> e(na);
    int infoh = 0;
    D.download(e.data(), na, s);
    STDA_HIP_CHECK(hipMemcpyAsync(&infoh, info.get(), sizeof(int), hipMemcpyDeviceToHost, s));
    c->sync();
    if (infoh != 0) throw Error(STDA_HIP_ERR_DEVICE, "eigensolve of S did not converge");
    for (auto &v : e) {
        if (v < 0.0) {
            if (!bfw) throw Error(STDA_HIP_ERR_ARG, "negative eigenvalue of S: must stop in S^1/2");
            v = 0.0;   // -BFW (stda.f:1660-1670)
        }
        v = std::sqrt(v);
    }
    D.upload(e.data(), na, s);
    dgemm(c->blas, true, false, na, nm, na, 1.0, M.get(), na, CA.get(), na, 0.0, T.get(), na);
    launch_scale_rows(s, na, nm, D.get(), T.get());
    dgemm(c->blas, false, false, na, nm, na, 1.0, M.get(), na, T.get(), na, 0.0, CA.get(), na);
    CA.download(clow, (size_t)na * nm, s);
    c->sync();
    // |<c'|c'> - 1| > 1e-2 stops the reference (stda.f:445-450)
    for (int64_t j = 0; j < nm; ++j) {
        double nr = 0.0;
        for (int64_t i = 0; i < na; ++i) nr += clow[i + na * j] * clow[i + na * j];
        if (std::fabs(nr - 1.0) > 1.0e-2)
            throw Error(STDA_HIP_ERR_ARG, "internal MO norm error (MO " + std::to_string(j + 1) +
                                              ", norm " + std::to_string(nr) + ")");
    }
}

int stda_hip_transform(stda_hip_ctx *c, int32_t nao, int32_t moci, const double *ca,
                       const double *S, const double *ao9, int32_t bfw, double *clow,
                       double *moments) {
    if (!c || nao < 1 || moci < 1 || moci > nao || !ca || !S || !ao9 || !clow || !moments)
        return STDA_HIP_ERR_ARG;
    STDA_TRY(c, {
        STDA_HIP_CHECK(hipSetDevice(c->device));
        transform_impl(c, nao, moci, ca, S, ao9, hipMemcpyHostToDevice, bfw, clow, moments);
    })
}

int stda_hip_prepare_ao(stda_hip_ctx *c, int32_t nao, const int64_t *prim_ptr, const double *alpha,
                        const double *coef, const double *center, const int32_t *lmn,
                        int32_t moci, const double *ca, int32_t bfw, double *clow,
                        double *moments) {
    if (!c || nao < 1 || moci < 1 || moci > nao || !prim_ptr || !alpha || !coef || !center ||
        !lmn || !ca || !clow || !moments)
        return STDA_HIP_ERR_ARG;
    if (prim_ptr[0] != 0) return STDA_HIP_ERR_ARG;
    for (int32_t m = 0; m < nao; ++m) {
        if (prim_ptr[m + 1] <= prim_ptr[m]) return STDA_HIP_ERR_ARG;
        for (int d = 0; d < 3; ++d)
            if (lmn[3 * m + d] < 0 || lmn[3 * m + d] > AO_LMAX) return STDA_HIP_ERR_ARG;
    }
    const int64_t nprim = prim_ptr[nao];
    STDA_TRY(c, {
        STDA_HIP_CHECK(hipSetDevice(c->device));
        hipStream_t s = c->stream;
        const size_t nn = (size_t)nao * nao;
        DBuf<int64_t> P;
        DBuf<double> AL, CO, CE, SO(nn), AO((size_t)9 * nn);
        DBuf<int32_t> LM;
        P.upload(prim_ptr, (size_t)nao + 1, s);
        AL.upload(alpha, (size_t)nprim, s);
        CO.upload(coef, (size_t)nprim, s);
        CE.upload(center, (size_t)3 * nao, s);
        LM.upload(lmn, (size_t)3 * nao, s);
        EventTimer ev(c);
        launch_ao_ints(s, nao, P.get(), AL.get(), CO.get(), CE.get(), LM.get(), SO.get(), AO.get());
        ev.finish("ao_ints", (double)nn);
        transform_impl(c, nao, moci, ca, SO.get(), AO.get(), hipMemcpyDeviceToDevice, bfw, clow,
                       moments);
    })
}

int stda_hip_set_selection(stda_hip_ctx *c, double thr2_ev, int32_t natoms,
                           const stda_hip_int *atoms, int32_t pt2_off) {
    if (!c || natoms < 0 || (natoms > 0 && (!atoms || !(thr2_ev > 0.0)))) return STDA_HIP_ERR_ARG;
    c->dual_atoms.clear();
    for (int32_t q = 0; q < natoms; ++q) {
        if (atoms[q] < 1 || atoms[q] > INT32_MAX) return STDA_HIP_ERR_ARG;
        c->dual_atoms.push_back((int32_t)atoms[q]);
    }
    c->dual_thr2_ev = natoms > 0 ? thr2_ev : 0.0;
    c->pt2_off = pt2_off != 0;
    c->ready = false;   // takes effect at the next select
    return STDA_HIP_OK;
}

int stda_hip_tpa(stda_hip_ctx *c, int32_t nstates, const double *eci, const double *xci,
                 const double *yci, int64_t ldx, double tol, double *sigma, double *delta) {
    if (!c) return STDA_HIP_ERR_ARG;
    if (!c->ready) return STDA_HIP_ERR_STATE;
    if (nstates < 1 || !eci || !xci || !yci || !sigma || !c->have_moments ||
        ldx < c->nci)
        return STDA_HIP_ERR_ARG;
    for (int32_t q = 0; q < nstates; ++q)
        if (!(eci[q] > 0.0)) return STDA_HIP_ERR_ARG;
    STDA_TRY(c, {
        STDA_HIP_CHECK(hipSetDevice(c->device));
        const int64_t n = c->nci, L = (int64_t)c->no * c->nv;
        const int no = c->no, nv = c->nv;
        hipStream_t s = c->stream;
        // responses at half of every excitation energy (lresp_2PA_SP,
        // linear_response.f:4704-4725; right-hand side -mu, here -2 mu and X, Y halved)
        std::vector<double> wl(nstates);
        for (int q = 0; q < nstates; ++q) wl[q] = eci[q] / 2.0;
        std::vector<double> al((size_t)9 * nstates), xpy((size_t)n * 3 * nstates);
        const int st = stda_hip_lresp(c, nstates, wl.data(), tol, al.data(), xpy.data());
        if (st != STDA_HIP_OK) throw Error(st, c->last_error);
        DBuf<double> XP, AX((size_t)n * 3 * nstates), rhs((size_t)n * 3), XN, YN;
        XP.upload(xpy.data(), (size_t)n * 3 * nstates, s);
        apply_ops(c, OP_APB, -1, 3 * nstates, XP.get(), n, AX.get(), n, nullptr, 0);
        STDA_HIP_CHECK(hipMemcpyAsync(rhs.get(), c->mom.get(), (size_t)n * 3 * 8,
                                      hipMemcpyDeviceToDevice, s));
        std::vector<double> m2(3, -2.0);
        DBuf<double> m2d;
        m2d.upload(m2.data(), 3, s);
        launch_scale_cols(s, n, 3, rhs.get(), m2d.get());
        XN.resize((size_t)n * nstates);
        YN.resize((size_t)n * nstates);
        for (int q = 0; q < nstates; ++q) {
            STDA_HIP_CHECK(hipMemcpyAsync(XN.get() + n * q, xci + ldx * q, (size_t)n * 8,
                                          hipMemcpyHostToDevice, s));
            STDA_HIP_CHECK(hipMemcpyAsync(YN.get() + n * q, yci + ldx * q, (size_t)n * 8,
                                          hipMemcpyHostToDevice, s));
        }
        DBuf<double> muoo((size_t)3 * no * no), muvv((size_t)3 * nv * nv);
        launch_mu_blocks(s, c->moments.get(), c->npk, no, nv, muoo.get(), muvv.get());
        // columns 0-2: X, Y of the responses (x, y, z); column 3: the state's X, Y
        DBuf<double> Xd((size_t)4 * L), Yd((size_t)4 * L), M((size_t)12 * L), T(4 * 12);
        std::vector<double> Th(4 * 12);
        for (int q = 0; q < nstates; ++q) {
            EventTimer ev(c);
            const double w = -eci[q] / 2.0;   // linear_response.f:4703
            BetaCols cols;
            cols.scale = 0.5;
            for (int r = 0; r < 3; ++r) {
                cols.src[r] = 3 * (int64_t)q + r;
                cols.inv_w[r] = 1.0 / w;
            }
            Xd.zero(s);
            Yd.zero(s);
            launch_beta_xy(s, n, c->pair.get(), no, nv, XP.get(), AX.get(), rhs.get(), cols, 3,
                           Xd.get(), Yd.get());
            launch_scatter_pairs(s, n, c->pair.get(), no, nv, 1, XN.get() + n * q, n,
                                 Xd.get() + 3 * L);
            launch_scatter_pairs(s, n, c->pair.get(), no, nv, 1, YN.get() + n * q, n,
                                 Yd.get() + 3 * L);
            // M(qd; col) = mu_qd^oo Y_col - Y_col mu_qd^vv, column 4 qd + col
            for (int qd = 0; qd < 3; ++qd) {
                double *Mq = M.get() + (size_t)qd * 4 * L;
                dgemm_batched(c->blas, false, false, no, nv, no, 1.0,
                              muoo.get() + (size_t)qd * no * no, no, 0, Yd.get(), no, L, 0.0, Mq,
                              no, L, 4);
                dgemm_batched(c->blas, false, false, no, nv, nv, -1.0, Yd.get(), no, L,
                              muvv.get() + (size_t)qd * nv * nv, nv, 0, 1.0, Mq, no, L, 4);
            }
            dgemm(c->blas, true, false, 4, 12, L, 1.0, Xd.get(), L, M.get(), L, 0.0, T.get(), 4);
            ev.finish("tpa", 12.0 * 2.0 * ((double)no * no * nv + (double)no * nv * nv) +
                                 2.0 * 4 * 12 * (double)L);
            T.download(Th.data(), 4 * 12, s);
            c->sync();
            // sigma(x,y) = (-A + B)/2 (TPA_resp_fast_SP, linear_response.f:4960-5022):
            // -A1+B1 = <X_x, M(y; n)>, -A2+B2 = <X_y, M(x; n)>,
            // -A3+B3 = <X_n, M(x; y)>, -A4+B4 = <X_n, M(y; x)>
            auto t = [&](int xc, int qd, int col) { return Th[xc + 4 * (qd * 4 + col)]; };
            double sg[9];
            for (int ix = 0; ix < 3; ++ix)
                for (int iy = 0; iy <= ix; ++iy) {
                    const double v =
                        0.5 * (t(ix, iy, 3) + t(iy, ix, 3) + t(3, ix, iy) + t(3, iy, ix));
                    sg[ix + 3 * iy] = sg[iy + 3 * ix] = v;
                }
            for (int e = 0; e < 9; ++e) sigma[9 * (size_t)q + e] = sg[e];
            if (delta) {
                double F = 0, G = 0, H = 0;
                for (int ix = 0; ix < 3; ++ix)
                    for (int iy = 0; iy < 3; ++iy) {
                        F += sg[ix + 3 * ix] * sg[iy + 3 * iy];
                        G += sg[ix + 3 * iy] * sg[ix + 3 * iy];
                        H += sg[ix + 3 * iy] * sg[iy + 3 * ix];
                    }
                F /= 30.0;
                G /= 30.0;
                H /= 30.0;
                double *d = delta + 4 * (size_t)q;
                d[0] = 2.0 * F + 2.0 * G + 2.0 * H;
                d[1] = -F + 4.0 * G - H;
                d[2] = -2.0 * F + 3.0 * G + 3.0 * H;
                d[3] = d[0] / d[1];
            }
        }
    })
}

// Fortran Fw.d / Ew.d output of the spectrum files (host only; no device calls).
// Non-finite numbers as the flang runtime writes them ("Inf", "-Inf", "NaN").
static int fort_special(double v, int w, char *out) {
    const char *t = std::isnan(v) ? "NaN" : (std::isinf(v) ? (v > 0 ? "Inf" : "-Inf") : nullptr);
    if (!t) return 0;
    std::snprintf(out, (size_t)w + 1, "%*s", w, t);
    return 1;
}

static void fort_f(double v, int w, int d, char *out) {   // Fw.d, asterisks on overflow
    if (fort_special(v, w, out)) return;
    char b[400];
    const int len = std::snprintf(b, sizeof b, "%.*f", d, v);
    if (len > w) {
        std::memset(out, '*', (size_t)w);
        out[w] = 0;
    } else {
        std::snprintf(out, (size_t)w + 1, "%*s", w, b);
    }
}

static void fort_e(double v, int w, int d, char *out) {   // Ew.d: 0.dddd, E+xx or +xxx
    if (fort_special(v, w, out)) return;
    char mant[32], b[64];
    int ex = 0;
    if (v == 0.0) {
        std::memset(mant, '0', (size_t)d);
        mant[d] = 0;
    } else {
        char t[64];
        std::snprintf(t, sizeof t, "%.*e", d - 1, std::fabs(v));   // d.ddde+XX
        int q = 0;
        for (const char *p = t; *p && *p != 'e'; ++p)
            if (*p != '.') mant[q++] = *p;
        mant[q] = 0;
        ex = std::atoi(std::strchr(t, 'e') + 1) + 1;
    }
    const char *sg = std::signbit(v) ? "-" : "";
    if (ex >= -99 && ex <= 99)
        std::snprintf(b, sizeof b, "%s0.%sE%c%02d", sg, mant, ex < 0 ? '-' : '+', ex < 0 ? -ex : ex);
    else
        std::snprintf(b, sizeof b, "%s0.%s%c%03d", sg, mant, ex < 0 ? '-' : '+', ex < 0 ? -ex : ex);
    std::snprintf(out, (size_t)w + 1, "%*s", w, b);
}

int stda_hip_format_curve(int64_t n, const double *e, const double *r, char *out) {
    if (n < 0 || (n > 0 && (!e || !r || !out))) return STDA_HIP_ERR_ARG;
    // lines are independent: split them over up to 16 host threads for long inputs
    const int64_t per = 4096;
    const int nth = (int)std::min<int64_t>(
        std::max<int64_t>(1, (n + per - 1) / per),
        std::max(1u, std::min(16u, std::thread::hardware_concurrency())));
    auto work = [&](int t) {
        for (int64_t i = t; i < n; i += nth) {
            char *o = out + 31 * i;
            char line[48];
            fort_f(e[i], 16, 6, line);
            fort_e(r[i], 14, 4, line + 16);
            std::memcpy(o, line, 30);
            o[30] = '\n';
        }
    };
    if (nth == 1) {
        work(0);
    } else {
        std::vector<std::thread> th;
        for (int t = 1; t < nth; ++t) th.emplace_back(work, t);
        work(0);
        for (auto &x : th) x.join();
    }
    return STDA_HIP_OK;
}

static_assert(STDA_HIP_GSPEC_UV == GSPEC_UV && STDA_HIP_GSPEC_NM == GSPEC_NM &&
                  STDA_HIP_GSPEC_CM == GSPEC_CM && STDA_HIP_GSPEC_LOG == GSPEC_LOG,
              "spectrum flags of the header and the kernel agree");

int stda_hip_gspec(stda_hip_ctx *c, int32_t nspec, const int64_t *line_offsets, const double *x,
                   const double *y, const double *wei, const double *emin, const double *step,
                   int32_t npts, int32_t flags, double rfaktor, double *e, double *r) {
    if (!c || nspec < 0 || npts <= 0 || (flags & ~15) != 0) return STDA_HIP_ERR_ARG;
    if (nspec == 0) return STDA_HIP_OK;
    if (!line_offsets || !emin || !step || !e || !r) return STDA_HIP_ERR_ARG;
    if (line_offsets[0] != 0) return STDA_HIP_ERR_ARG;
    for (int32_t q = 0; q < nspec; ++q)
        if (line_offsets[q + 1] < line_offsets[q]) return STDA_HIP_ERR_ARG;
    const int64_t nl = line_offsets[nspec];
    if (nl > 0 && (!x || !y || !wei)) return STDA_HIP_ERR_ARG;
    STDA_TRY(c, {
        STDA_HIP_CHECK(hipSetDevice(c->device));
        hipStream_t st = c->stream;
        const size_t np = (size_t)nspec * npts;
        DBuf<int64_t> off;
        DBuf<double> X, Y, W, E0, ST, EO(np), RO(np);
        off.upload(line_offsets, (size_t)nspec + 1, st);
        X.upload(x, (size_t)nl, st);
        Y.upload(y, (size_t)nl, st);
        W.upload(wei, (size_t)nl, st);
        E0.upload(emin, (size_t)nspec, st);
        ST.upload(step, (size_t)nspec, st);
        EventTimer ev(c);
        launch_gspec(st, nspec, off.get(), X.get(), Y.get(), W.get(), E0.get(), ST.get(), npts,
                     flags, rfaktor, EO.get(), RO.get());
        // work: the (point, line) terms
        ev.finish("gspec", (double)nl * npts);
        EO.download(e, np, st);
        RO.download(r, np, st);
        c->sync();
    })
}

int stda_hip_gspec_noalloc(stda_hip_ctx *c, int32_t nspec, const int64_t *line_offsets,
                           const double *x, const double *y, const double *wei,
                           const double *emin, const double *step, int32_t npts, int32_t flags,
                           double rfaktor, double *e, double *r) {
    if (!c || nspec < 0 || npts <= 0 || (flags & ~15) != 0) return STDA_HIP_ERR_ARG;
    if (nspec == 0) return STDA_HIP_OK;
    if (!line_offsets || !emin || !step || !e || !r) return STDA_HIP_ERR_ARG;
    if (line_offsets[0] != 0) return STDA_HIP_ERR_ARG;
    for (int32_t q = 0; q < nspec; ++q)
        if (line_offsets[q + 1] < line_offsets[q]) return STDA_HIP_ERR_ARG;
    const int64_t nl = line_offsets[nspec];
    if (nl > 0 && (!x || !y || !wei)) return STDA_HIP_ERR_ARG;
    // staging layout in the scratch array (doubles; every part starts on 256 bytes)
    const size_t np = (size_t)nspec * (size_t)npts;
    auto up = [](size_t nd) { return (nd + 31) / 32 * 32; };
    const size_t o_off = 0, o_x = o_off + up((size_t)nspec + 1), o_y = o_x + up((size_t)nl),
                 o_w = o_y + up((size_t)nl), o_e0 = o_w + up((size_t)nl),
                 o_st = o_e0 + up((size_t)nspec), o_eo = o_st + up((size_t)nspec),
                 o_ro = o_eo + up(np), total = o_ro + up(np);
    if (total * sizeof(double) > kScratchBytes) {
        c->last_error = "gspec_noalloc: " + std::to_string(total * sizeof(double)) +
                        " bytes of staging needed, the fixed area holds " +
                        std::to_string(kScratchBytes);
        return STDA_HIP_ERR_TOO_LARGE;
    }
    double *base = scratch_of(c->device);
    if (!base) {
        c->last_error = "gspec_noalloc: no scratch address for this device";
        return STDA_HIP_ERR_STATE;
    }
    try {
        STDA_HIP_CHECK(hipSetDevice(c->device));
        hipStream_t st = c->stream;
        int64_t *doff = reinterpret_cast<int64_t *>(base + o_off);
        STDA_HIP_CHECK(hipMemcpyAsync(doff, line_offsets, ((size_t)nspec + 1) * 8, hipMemcpyHostToDevice, st));
        if (nl > 0) {
            STDA_HIP_CHECK(hipMemcpyAsync(base + o_x, x, (size_t)nl * 8, hipMemcpyHostToDevice, st));
            STDA_HIP_CHECK(hipMemcpyAsync(base + o_y, y, (size_t)nl * 8, hipMemcpyHostToDevice, st));
            STDA_HIP_CHECK(hipMemcpyAsync(base + o_w, wei, (size_t)nl * 8, hipMemcpyHostToDevice, st));
        }
        STDA_HIP_CHECK(hipMemcpyAsync(base + o_e0, emin, (size_t)nspec * 8, hipMemcpyHostToDevice, st));
        STDA_HIP_CHECK(hipMemcpyAsync(base + o_st, step, (size_t)nspec * 8, hipMemcpyHostToDevice, st));
        // every grid point of every spectrum is written by the kernel: nothing to clear
        launch_gspec(st, nspec, doff, base + o_x, base + o_y, base + o_w, base + o_e0,
                     base + o_st, npts, flags, rfaktor, base + o_eo, base + o_ro);
        STDA_HIP_CHECK(hipMemcpyAsync(e, base + o_eo, np * 8, hipMemcpyDeviceToHost, st));
        STDA_HIP_CHECK(hipMemcpyAsync(r, base + o_ro, np * 8, hipMemcpyDeviceToHost, st));
        STDA_HIP_CHECK(hipStreamSynchronize(st));
    } catch (const Error &ex) {
        c->last_error = ex.what();
        return ex.status;
    }
    return STDA_HIP_OK;
}

int stda_hip_set_precision(stda_hip_ctx *c, int32_t precision) {
    if (!c || precision < 0 || precision > 2) return STDA_HIP_ERR_ARG;
    c->precision = precision;
    c->gemm32 = GEMM_F32;
    return STDA_HIP_OK;
}

int stda_hip_gemm(stda_hip_ctx *c, int32_t mode, int32_t ta, int32_t tb, int32_t m, int32_t n,
                  int32_t k, double alpha, const double *A, int64_t lda, const double *B,
                  int64_t ldb, double beta, double *C, int64_t ldc) {
    if (!c || mode < 0 || mode > 2 || m < 0 || n < 0 || k < 0) return STDA_HIP_ERR_ARG;
    if (m == 0 || n == 0) return STDA_HIP_OK;
    const int64_t ar = ta ? k : m, ac = ta ? m : k, br = tb ? n : k, bc = tb ? k : n;
    if (lda < std::max<int64_t>(1, ar) || ldb < std::max<int64_t>(1, br) || ldc < m) return STDA_HIP_ERR_ARG;
    if ((ac > 0 && !A) || (bc > 0 && !B) || !C) return STDA_HIP_ERR_ARG;
    STDA_TRY(c, {
        STDA_HIP_CHECK(hipSetDevice(c->device));
        hipStream_t s = c->stream;
        const size_t na = (size_t)lda * ac, nbb = (size_t)ldb * bc, nc = (size_t)ldc * n;
        if (mode == 0) {
            DBuf<double> dA, dB, dC;
            dA.upload(A, na, s);
            dB.upload(B, nbb, s);
            dC.upload(C, nc, s);
            launch_gemm_f64(s, ta != 0, tb != 0, m, n, k, alpha, dA.get(), lda, 0, dB.get(), ldb, 0,
                            beta, dC.get(), ldc, 0, 1);
            dC.download(C, nc, s);
        } else {
            std::vector<float> hA(A, A + na), hB(B, B + nbb), hC(C, C + nc);
            DBuf<float> dA, dB, dC;
            dA.upload(hA.data(), na, s);
            dB.upload(hB.data(), nbb, s);
            dC.upload(hC.data(), nc, s);
            launch_gemm_f32(s, mode == 2 ? GEMM_BF16X3 : GEMM_F32, ta != 0, tb != 0, m, n, k,
                            (float)alpha, dA.get(), lda, 0, dB.get(), ldb, 0, (float)beta,
                            dC.get(), ldc, 0, 1);
            dC.download(hC.data(), nc, s);
            c->sync();
            for (size_t q = 0; q < nc; ++q) C[q] = hC[q];
        }
        c->sync();
    })
}

int stda_hip_timings(const stda_hip_ctx *c, double *ms6) {
    if (!c || !ms6) return STDA_HIP_ERR_ARG;
    for (int i = 0; i < 6; ++i) ms6[i] = c->timings[i];
    return STDA_HIP_OK;
}

}  // extern "C"
// ============================================================================================
// Batched solves of many independent molecules (C ABI: stda_hip_batch_*).
//
// `nworkers` contexts, each driven by its own host thread and HIP stream,
// take the molecules in ascending size and run the setup pipeline (load,
// gamma, charges, screening, PT2 selection).  A molecule with at most
// nmax_batched CSFs and at most BATCH_KMAX roots then gets its explicit FP64
// A, the FP32 copy, 0.5 B (velocity correction), the diagonal and the CSF
// moments built on the device and is handed to the driver (the calling
// thread).  The driver keeps a set of active molecules in lock step on its
// own stream: per iteration one launch of batch_matvec_kernel (operator
// products of every active molecule) and one of batch_step_kernel (one
// Davidson step per molecule).  Newly prepared molecules join the set at the
// next iteration and finished ones leave it when the driver polls the state
// table (every few iterations), so the set is continuously refilled while
// the workers prepare the rest.  Larger molecules take the single-molecule
// mixed path on their worker.  Allocations are stream-ordered on each
// worker's stream, so the workers never synchronise the device.

struct stda_hip_batch {
    int device = 0;
    std::vector<stda_hip_ctx *> workers;
    hipStream_t dstream = nullptr;       // the driver's stream
    std::string last_error;
    int64_t nmax_batched = 1024;         // CSF limit of the explicit-operator lock step
    bool wide = true;                    // larger molecules in the matrix-free lock step
    bool two_phase = false;              // lock step only after the setup of every molecule
    int64_t nmax_wide = INT64_MAX;       // CSF limit of the matrix-free lock step
    int group = 4;                       // iterations between polls of the state table
    // ms: worker phase (wall), driver loop (wall), property tables, total,
    // summed HIP-event time of the batched launches (matvec + step)
    double timings[5] = {0, 0, 0, 0, 0};
    int64_t counts[3] = {0, 0, 0};       // molecules: batched, single path, failed
    int64_t nwide = 0;                   // of the batched: in the matrix-free lock step
    double kernel_bytes = 0.0;           // algorithmic bytes of the batched operator products
    double lockstep_ms = 0.0;            // HIP-event time of the lock-step graph launches
    double lockstep_launches = 0.0;      // and their number (each `group` iterations)
    // summed step phases per molecule (us), 0, Jacobi sweeps, then the
    // workers' summed host time (us) in setup, explicit matrices, single path
    double phase_us[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    double wall_tick_hz = 1.0e8;
    struct BatchStream *engine = nullptr;   // continuous mode (stda_hip_batch_submit), or null
};

namespace {

struct BatchItem {
    int64_t n = 0;
    int k = 0, maxdim = 0, nguess = 0;
    double ak = 2.0;
    bool moments = false, velcorr = false;
    DBuf<float> A32;
    DBuf<double> A64, B64, diag, mom, work;
    DBuf<int32_t> guess;
    // wide molecules (matrix-free lock step, wide.hip): copies of the
    // operator data of the context and the operator work buffers
    bool wide = false;
    int64_t npair = 0;
    int nao = 0, no = 0, nv = 0, na = 0, nvb = 0, nchunk = 0;
    double ax = 0.0;
    DBuf<float> C32, gj32, gk32, xf32, t32, d32, r32, u32;
    DBuf<double> C64, gj64, gk64, corr, xf64, t64, d64, g64, r64, u64, aux, part;
    DBuf<int32_t> atom_of, atom_ptr, pair;
};

template <typename T>
void copy_device(DBuf<T> &dst, const DBuf<T> &src, size_t count, hipStream_t s) {
    dst.resize(count);
    if (count)
        STDA_HIP_CHECK(hipMemcpyAsync(dst.get(), src.get(), count * sizeof(T),
                                      hipMemcpyDeviceToDevice, s));
}

// Hand-over of a prepared molecule to the wide lock-step solver: its
// operator data is copied out of the context (which goes on to the next
// molecule) and the work buffers are sized (FP32 products for nvb vectors,
// FP64 for k).  The pair-grid buffers are cleared once: only the CSF pairs
// are ever written into them.
void wide_prepare(stda_hip_ctx *c, int k, bool velcorr, BatchItem &it) {
    const int64_t n = c->nci;
    hipStream_t s = c->stream;
    it.wide = true;
    it.n = n;
    it.k = k;
    it.ak = c->ak;
    it.ax = c->ax;
    it.moments = c->have_moments;
    it.velcorr = velcorr && c->have_moments;
    it.maxdim = (int)std::min<int64_t>(n, WIDE_MD);
    it.nguess = std::min(it.maxdim, std::max(k + 8, k + k / 2));
    it.nvb = std::max(it.nguess, k);
    it.nao = c->nao;
    it.no = c->no;
    it.nv = c->nv;
    it.na = c->ncent;
    it.npair = c->npair;
    it.nchunk = (int)((n + WIDE_RC - 1) / WIDE_RC);
    ensure_fp32(c);
    const size_t nao = it.nao, moci = (size_t)it.no + it.nv, na = it.na, np = it.npair;
    copy_device(it.C32, c->ct32, nao * moci, s);
    copy_device(it.C64, c->ct, nao * moci, s);
    copy_device(it.gj32, c->gamj32, na * na, s);
    copy_device(it.gk32, c->gamk32, na * na, s);
    copy_device(it.gj64, c->gamj, na * na, s);
    copy_device(it.gk64, c->gamk, na * na, s);
    copy_device(it.atom_of, c->atom_of, nao, s);
    copy_device(it.atom_ptr, c->atom_ptr, na + 1, s);
    copy_device(it.pair, c->pair, (size_t)n, s);
    copy_device(it.corr, c->corr, (size_t)n, s);
    copy_device(it.diag, c->ed, (size_t)n, s);
    if (it.moments) copy_device(it.mom, c->mom, (size_t)9 * n, s);
    std::vector<int32_t> g = lowest_diag_rows(c->ed_h, it.nguess);
    it.guess.upload(g.data(), g.size(), s);
    const size_t vb = it.nvb, kb = k, no = it.no;
    it.xf32.resize(np * vb);
    it.xf32.zero(s);
    it.t32.resize(nao * no * vb);
    it.d32.resize(nao * nao * vb);
    it.r32.resize(np * vb);
    it.u32.resize(na * vb);
    it.xf64.resize(np * kb);
    it.xf64.zero(s);
    it.t64.resize(nao * no * kb);
    it.d64.resize(nao * nao * kb);
    it.g64.resize(nao * nao * kb);
    it.r64.resize(np * kb);
    it.u64.resize(na * kb);
    it.work.resize((size_t)n * (2 * it.maxdim + 3 * BATCH_KMAX) + (size_t)WIDE_MD * WIDE_MD);
    it.aux.resize(WIDE_AUX);
    it.part.resize((size_t)it.nchunk * WIDE_PART);
}

WideMol wide_desc(BatchItem &it, double tol32, double *out, int *state) {
    WideMol d{};
    const int64_t n = it.n;
    d.n = n;
    d.npair = it.npair;
    d.nao = it.nao;
    d.no = it.no;
    d.nv = it.nv;
    d.na = it.na;
    d.k = it.k;
    d.nguess = it.nguess;
    d.maxdim = it.maxdim;
    d.maxiter = 300;
    d.nvb = it.nvb;
    d.velcorr = it.velcorr ? 1 : 0;
    d.nchunk = it.nchunk;
    d.tol = tol32;
    d.ak = it.ak;
    d.ax = it.ax;
    d.C32 = it.C32.get();
    d.C64 = it.C64.get();
    d.gj32 = it.gj32.get();
    d.gk32 = it.gk32.get();
    d.gj64 = it.gj64.get();
    d.gk64 = it.gk64.get();
    d.atom_of = it.atom_of.get();
    d.atom_ptr = it.atom_ptr.get();
    d.pair = it.pair.get();
    d.guess = it.guess.get();
    d.corr = it.corr.get();
    d.diag = it.diag.get();
    d.mom = it.moments ? it.mom.get() : nullptr;
    d.xf32 = it.xf32.get();
    d.t32 = it.t32.get();
    d.d32 = it.d32.get();
    d.r32 = it.r32.get();
    d.u32 = it.u32.get();
    d.xf64 = it.xf64.get();
    d.t64 = it.t64.get();
    d.d64 = it.d64.get();
    d.g64 = it.g64.get();
    d.r64 = it.r64.get();
    d.u64 = it.u64.get();
    double *w = it.work.get();
    d.V = w;
    d.AV = w + n * it.maxdim;
    d.X = w + 2 * n * it.maxdim;
    d.AX = d.X + n * BATCH_KMAX;
    d.W = d.AX + n * BATCH_KMAX;
    d.H = d.W + n * BATCH_KMAX;
    d.aux = it.aux.get();
    d.part = it.part.get();
    d.out = out;
    d.state = state;
    return d;
}

// The four transforms of a wide molecule's operator product as grouped GEMM
// problems, FP32 (iterations, `cap` = nvb vectors) and FP64 (refinement,
// velocity correction, k vectors): index 2 * stage + precision, stages
// T = Cv X^T (NN), D = T Co^T (NT), W = G Co (NN), R = Cv^T W (TN).
void wide_gemm_probs(const WideMol &d, GemmProb (&P)[8]) {
    const int nao = d.nao, no = d.no, nv = d.nv;
    auto tiles = [](int64_t x) { return (int)((x + 63) / 64); };
    for (int prec = 0; prec < 2; ++prec) {
        const bool f64 = prec == 1;
        const int cap = f64 ? d.k : d.nvb;
        const size_t es = f64 ? 8 : 4;
        const char *C = static_cast<const char *>(f64 ? (const void *)d.C64 : (const void *)d.C32);
        const void *Co = C, *Cv = C + es * (size_t)nao * no;
        void *xf = f64 ? (void *)d.xf64 : (void *)d.xf32;
        void *t = f64 ? (void *)d.t64 : (void *)d.t32;
        void *D = f64 ? (void *)d.d64 : (void *)d.d32;
        void *G = f64 ? (void *)d.g64 : (void *)d.d32;   // FP32: G in place of D
        void *R = f64 ? (void *)d.r64 : (void *)d.r32;
        // FP64 (a few products per molecule): 4 vectors in turn per work-group,
        // so that the launch stays small while no molecule needs it
        const int bpw = f64 ? 4 : 1, gbat = (cap + bpw - 1) / bpw;
        GemmProb base{};
        base.bpw = bpw;
        base.kind = f64 ? GC_OP64 : GC_OP32;
        base.state = d.state;
        base.alpha = 1.0;
        base.beta = 0.0;
        // T = Cv X^T: nao x (no * count), K = nv (FP64: per vector)
        GemmProb g1 = base;
        g1.M = nao; g1.N = no; g1.K = nv; g1.cols_per_count = no;
        g1.A = Cv; g1.lda = nao; g1.B = xf; g1.ldb = nv; g1.C = t; g1.ldc = nao;
        g1.gm = tiles(nao); g1.gn = tiles((int64_t)no * cap); g1.gb = 1;
        if (f64) {
            g1.cols_per_count = 0;
            g1.sB = (int64_t)nv * no;
            g1.sC = (int64_t)nao * no;
            g1.gn = tiles(no);
            g1.gb = gbat;
        }
        // D_v = T_v Co^T: nao x nao, K = no, per vector.  FP32 (operator A):
        // stored as G = -gamJ o D + ak diag(u) (epilogue); FP64: D itself
        GemmProb g2 = base;
        g2.M = nao; g2.N = nao; g2.K = no;
        g2.A = t; g2.lda = nao; g2.sA = (int64_t)nao * no;
        g2.B = Co; g2.ldb = nao; g2.C = D; g2.ldc = nao; g2.sC = (int64_t)nao * nao;
        g2.gm = tiles(nao); g2.gn = tiles(nao); g2.gb = gbat;
        if (!f64) {
            g2.epi = 1;
            g2.ep_gam = d.gj32;
            g2.ep_atom = d.atom_of;
            g2.ep_u = d.u32;
            g2.ep_na = d.na;
            g2.ep_aJ = -1.0;
            g2.ep_aK = d.ak;
        }
        // W_v = G_v Co: nao x no, K = nao, per vector (into the T buffer)
        GemmProb g3 = base;
        g3.M = nao; g3.N = no; g3.K = nao;
        g3.A = G; g3.lda = nao; g3.sA = (int64_t)nao * nao;
        g3.B = Co; g3.ldb = nao; g3.C = t; g3.ldc = nao; g3.sC = (int64_t)nao * no;
        g3.gm = tiles(nao); g3.gn = tiles(no); g3.gb = gbat;
        // R = Cv^T W: nv x (no * count), K = nao
        GemmProb g4 = base;
        g4.M = nv; g4.N = no; g4.K = nao; g4.cols_per_count = no;
        g4.A = Cv; g4.lda = nao; g4.B = t; g4.ldb = nao; g4.C = R; g4.ldc = nv;
        g4.gm = tiles(nv); g4.gn = tiles((int64_t)no * cap); g4.gb = 1;
        if (f64) {
            g4.cols_per_count = 0;
            g4.sB = (int64_t)nao * no;
            g4.sC = (int64_t)nv * no;
            g4.gn = tiles(no);
            g4.gb = gbat;
        }
        P[0 + prec] = g1;
        P[2 + prec] = g2;
        P[4 + prec] = g3;
        P[6 + prec] = g4;
    }
}


// 0.5 B = 0.5 ak (ia|jb)_K - 0.5 ax (ib|ja)_K for all CSF pairs (rtdacorr,
// apbtrafo.f:580-637), n x n column-major on the device
void build_bhalf_device(stda_hip_ctx *c, DBuf<double> &B) {
    const int64_t n = c->nci;
    const int na = c->ncent;
    hipStream_t s = c->stream;
    DBuf<double> pq((size_t)na * n), qq((size_t)na * n);
    launch_gather_cols(s, na, n, c->pair.get(), c->pia.get(), pq.get());
    launch_gather_cols(s, na, n, c->pair.get(), c->qia.get(), qq.get());
    B.resize((size_t)n * n);
    launch_exchange(s, n, na, c->nv, c->pair.get(), c->pia.get(), c->qia.get(), B.get(), n);
    dgemm(c->blas, true, false, n, n, na, 0.5 * c->ak, pq.get(), na, qq.get(), na, -0.5 * c->ax,
          B.get(), n);
}

// one molecule of the setup phase; returns true when it goes to the batched solver
bool batch_prepare(stda_hip_ctx *c, const stda_hip_system *sys, const stda_hip_params *par,
                   int nroots, double tol, int velcorr, int64_t nmax, int64_t nmax_wide, BatchItem &it,
                   int32_t *nfound, double *omega, double *props, double *us3) {
    Timer t0;
    do_load(c, sys, par);
    do_select(c);
    us3[0] += 1e3 * t0.ms();
    Timer t1;
    const int64_t n = c->nci;
    const int k = (int)std::min<int64_t>(nroots, n);
    hipStream_t s = c->stream;
    if (n > nmax && n <= nmax_wide && k <= BATCH_KMAX && c->ncent <= WIDE_MAXATOMS) {
        wide_prepare(c, k, velcorr != 0, it);
        us3[1] += 1e3 * t1.ms();
        return true;
    }
    if (n > nmax || k > BATCH_KMAX) {
        struct Add {
            double *d;
            Timer t;
            ~Add() { *d += 1e3 * t.ms(); }
        } add{us3 + 2, Timer()};
        std::vector<double> w;
        DBuf<double> X;
        solve_tda_mixed(c, nroots, tol, w, X);
        std::vector<double> pr((size_t)5 * w.size());
        tda_properties(c, X.get(), w, velcorr, pr.data(), nullptr);
        *nfound = (int32_t)w.size();
        for (size_t j = 0; j < w.size(); ++j) {
            omega[j] = w[j];
            for (int q = 0; q < 5; ++q) props[5 * j + q] = pr[5 * j + q];
        }
        return false;
    }
    it.n = n;
    it.k = k;
    it.ak = c->ak;
    it.moments = c->have_moments;
    it.velcorr = velcorr && c->have_moments;
    it.maxdim = (int)std::min<int64_t>(n, BATCH_MD);
    it.nguess = std::min(it.maxdim, std::max(k + 8, k + k / 2));
    build_a_device(c, it.A64);
    it.A32.resize((size_t)n * n);
    launch_to_f32(s, it.A64.get(), n * n, it.A32.get());
    if (it.velcorr) build_bhalf_device(c, it.B64);
    it.diag.resize(n);
    STDA_HIP_CHECK(hipMemcpyAsync(it.diag.get(), c->ed.get(), n * 8, hipMemcpyDeviceToDevice, s));
    if (it.moments) {
        it.mom.resize((size_t)9 * n);
        STDA_HIP_CHECK(hipMemcpyAsync(it.mom.get(), c->mom.get(), (size_t)9 * n * 8,
                                      hipMemcpyDeviceToDevice, s));
    }
    std::vector<int32_t> g = lowest_diag_rows(c->ed_h, it.nguess);
    it.guess.upload(g.data(), g.size(), s);
    it.work.resize((size_t)n * (2 * it.maxdim + 3 * BATCH_KMAX) + (size_t)BATCH_MD * BATCH_MD);
    us3[1] += 1e3 * t1.ms();
    return true;
}

BatchMol batch_desc(BatchItem &it, double tol32, double *out, int *state) {
    BatchMol d;
    const int64_t n = it.n;
    d.n = n;
    d.k = it.k;
    d.nguess = it.nguess;
    d.maxdim = it.maxdim;
    d.maxiter = 300;
    d.velcorr = it.velcorr ? 1 : 0;
    d.pad = 0;
    d.tol = tol32;
    d.A32 = it.A32.get();
    d.A64 = it.A64.get();
    d.B64 = it.velcorr ? it.B64.get() : nullptr;
    d.diag = it.diag.get();
    d.mom = it.moments ? it.mom.get() : nullptr;
    d.guess = it.guess.get();
    double *w = it.work.get();
    d.V = w;
    d.AV = w + n * it.maxdim;
    d.X = w + 2 * n * it.maxdim;
    d.AX = d.X + n * BATCH_KMAX;
    d.W = d.AX + n * BATCH_KMAX;
    d.H = d.W + n * BATCH_KMAX;
    d.out = out;
    d.state = state;
    return d;
}

// The device description of an active lock-step set: explicit-operator
// molecules (BatchMol: one work-group each per step, 64-row tiles for the
// products) and wide ones (WideMol: grouped kernels and GEMMs), and the
// captured graph of `group` iterations over both.
struct LockstepSet {
    struct Entry {
        BatchItem *item;
        double tol32;
        double *out;
        int *state;
    };
    DBuf<BatchMol> descd;
    DBuf<int2> tilesd;
    int nsmall = 0, ntiles = 0;
    DBuf<WideMol> wdescd;
    DBuf<int> wstart;       // chunk | vector | G-build offsets, nwide + 1 each
    DBuf<GemmProb> gprob;   // 8 groups of nwide problems
    DBuf<int> gstart;       // 8 groups of nwide + 1 offsets
    int gtotal[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int nwide = 0;
    WideStarts ws{};
    int gemm32 = GEMM_F32;
    hipGraphExec_t gexec = nullptr;

    // the explicit-operator molecules run on a side branch of the graph, beside the wide ones
    hipStream_t side = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;

    LockstepSet() {
        STDA_HIP_CHECK(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));
        STDA_HIP_CHECK(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
        STDA_HIP_CHECK(hipEventCreateWithFlags(&join, hipEventDisableTiming));
    }
    LockstepSet(const LockstepSet &) = delete;
    LockstepSet &operator=(const LockstepSet &) = delete;
    ~LockstepSet() {
        drop_graph();
        if (side) {
            (void)hipStreamSynchronize(side);
            (void)hipStreamDestroy(side);
        }
        if (fork) (void)hipEventDestroy(fork);
        if (join) (void)hipEventDestroy(join);
    }
    void drop_graph() {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        gexec = nullptr;
    }
    bool empty() const { return nsmall == 0 && nwide == 0; }

    void build(std::vector<Entry> entries, hipStream_t ds) {
        drop_graph();
        // largest first: their work-groups start first
        std::stable_sort(entries.begin(), entries.end(),
                         [](const Entry &a, const Entry &c) { return a.item->n > c.item->n; });
        std::vector<BatchMol> desc;
        std::vector<int2> tiles;
        std::vector<WideMol> wdesc;
        for (const Entry &e : entries) {
            if (e.item->wide) {
                wdesc.push_back(wide_desc(*e.item, e.tol32, e.out, e.state));
                continue;
            }
            const int q = (int)desc.size();
            desc.push_back(batch_desc(*e.item, e.tol32, e.out, e.state));
            const int nblk = (int)((e.item->n + BATCH_MV_ROWS - 1) / BATCH_MV_ROWS);
            for (int r = 0; r < nblk; ++r) tiles.push_back(make_int2(q, r));
        }
        nsmall = (int)desc.size();
        ntiles = (int)tiles.size();
        if (nsmall) {
            descd.upload(desc.data(), desc.size(), ds);
            tilesd.upload(tiles.data(), tiles.size(), ds);
        }
        nwide = (int)wdesc.size();
        if (nwide) {
            wdescd.upload(wdesc.data(), wdesc.size(), ds);
            std::vector<int> st((size_t)3 * (nwide + 1), 0);
            int *ck = st.data(), *vc = ck + nwide + 1, *gb = vc + nwide + 1;
            for (int q = 0; q < nwide; ++q) {
                const WideMol &d = wdesc[q];
                ck[q + 1] = ck[q] + d.nchunk;
                vc[q + 1] = vc[q] + d.nvb;
                gb[q + 1] = gb[q] + d.nvb * ((d.nao + WIDE_GB - 1) / WIDE_GB);
            }
            wstart.upload(st.data(), st.size(), ds);
            ws.chunk = wstart.get();
            ws.vec = ws.chunk + nwide + 1;
            ws.gb = ws.vec + nwide + 1;
            ws.nchunks = ck[nwide];
            ws.nvecs = vc[nwide];
            ws.ngb = gb[nwide];
            std::vector<GemmProb> probs((size_t)8 * nwide);
            std::vector<int> gs((size_t)8 * (nwide + 1), 0);
            // the device addresses of the states are those in the descriptors
            for (int q = 0; q < nwide; ++q) {
                GemmProb P[8];
                wide_gemm_probs(wdesc[q], P);
                for (int g = 0; g < 8; ++g) probs[(size_t)g * nwide + q] = P[g];
            }
            for (int g = 0; g < 8; ++g) {
                int *o = gs.data() + (size_t)g * (nwide + 1);
                for (int q = 0; q < nwide; ++q) {
                    const GemmProb &P = probs[(size_t)g * nwide + q];
                    const int64_t cnt = o[q] + (int64_t)P.gm * P.gn * P.gb;
                    if (cnt > INT32_MAX) throw Error(STDA_HIP_ERR_ARG, "wide lock-step: too many GEMM tiles");
                    o[q + 1] = (int)cnt;
                }
                gtotal[g] = o[nwide];
            }
            gprob.upload(probs.data(), probs.size(), ds);
            gstart.upload(gs.data(), gs.size(), ds);
        }
    }

    // one lock-step iteration of both kinds (the explicit-operator molecules
    // on the side stream when there are wide ones: two branches of the graph)
    void iteration(hipStream_t ds) {
        const bool both = nsmall && nwide;
        if (both) {
            STDA_HIP_CHECK(hipEventRecord(fork, ds));
            STDA_HIP_CHECK(hipStreamWaitEvent(side, fork, 0));
        }
        if (nsmall) {
            hipStream_t ss = both ? side : ds;
            launch_batch_matvec(ss, ntiles, descd.get(), tilesd.get());
            launch_batch_step(ss, nsmall, descd.get());
        }
        if (nwide) {
            const WideMol *w = wdescd.get();
            auto gemm = [&](int g, int mode, bool ta, bool tb) {
                launch_gemm_grouped(ds, mode, ta, tb, nwide, gprob.get() + (size_t)g * nwide,
                                    gstart.get() + (size_t)g * (nwide + 1), gtotal[g]);
            };
            launch_wide_scatter(ds, nwide, w, ws);
            gemm(0, gemm32, false, false);
            gemm(1, GEMM_F64, false, false);
            launch_wide_atomdiag(ds, nwide, w, ws, false);
            gemm(2, gemm32, false, true);
            gemm(3, GEMM_F64, false, true);
            launch_wide_atomdiag(ds, nwide, w, ws, true);
            launch_wide_gbuild(ds, nwide, w, ws);
            gemm(4, gemm32, false, false);
            gemm(5, GEMM_F64, false, false);
            gemm(6, gemm32, true, false);
            gemm(7, GEMM_F64, true, false);
            launch_wide_gather(ds, nwide, w, ws);
            launch_wide_step(ds, nwide, w, ws);
        }
        if (both) {
            STDA_HIP_CHECK(hipEventRecord(join, side));
            STDA_HIP_CHECK(hipStreamWaitEvent(ds, join, 0));
        }
    }

    // `group` iterations as one graph launch (captured on first use after a build)
    void launch(hipStream_t ds, int group) {
        if (!gexec) {
            hipGraph_t graph = nullptr;
            STDA_HIP_CHECK(hipStreamBeginCapture(ds, hipStreamCaptureModeThreadLocal));
            for (int j = 0; j < group; ++j) iteration(ds);
            STDA_HIP_CHECK(hipStreamEndCapture(ds, &graph));
            const hipError_t ie = hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            STDA_HIP_CHECK(ie);
        }
        STDA_HIP_CHECK(hipGraphLaunch(gexec, ds));
    }
};

}  // namespace

extern "C" {

int stda_hip_batch_create(stda_hip_batch **out, int device, int nworkers) {
    if (!out || nworkers < 1 || nworkers > 256) return STDA_HIP_ERR_ARG;
    auto *b = new stda_hip_batch();
    for (int w = 0; w < nworkers; ++w) {
        stda_hip_ctx *c = nullptr;
        int st = stda_hip_create(&c, device);
        if (st != STDA_HIP_OK) {
            stda_hip_batch_destroy(b);
            return st;
        }
        c->precision = 1;
        b->workers.push_back(c);
    }
    b->device = b->workers[0]->device;
    if (hipSetDevice(b->device) != hipSuccess ||
        hipStreamCreateWithFlags(&b->dstream, hipStreamNonBlocking) != hipSuccess) {
        stda_hip_batch_destroy(b);
        return STDA_HIP_ERR_DEVICE;
    }
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, b->device) == hipSuccess &&
        khz > 0)
        b->wall_tick_hz = 1.0e3 * khz;
    *out = b;
    return STDA_HIP_OK;
}

int stda_hip_batch_destroy(stda_hip_batch *b) {
    if (!b) return STDA_HIP_OK;
    (void)stda_hip_batch_stop(b);
    if (b->dstream) {
        block_cache_drain(b->dstream);
        (void)hipStreamSynchronize(b->dstream);
        (void)hipStreamDestroy(b->dstream);
    }
    for (auto *c : b->workers) stda_hip_destroy(c);
    delete b;
    return STDA_HIP_OK;
}

const char *stda_hip_batch_last_error(const stda_hip_batch *b) {
    return b ? b->last_error.c_str() : "";
}

int stda_hip_batch_configure(stda_hip_batch *b, int64_t nmax_batched, int32_t poll) {
    if (!b || nmax_batched < 0 || poll < 1) return STDA_HIP_ERR_ARG;
    b->nmax_batched = nmax_batched;
    b->group = poll;
    return STDA_HIP_OK;
}

int stda_hip_batch_configure_wide(stda_hip_batch *b, int32_t enable, int64_t nmax_wide) {
    if (!b || enable < 0 || enable > 2) return STDA_HIP_ERR_ARG;
    b->wide = enable != 0;
    b->two_phase = enable == 2;
    b->nmax_wide = nmax_wide > 0 ? nmax_wide : INT64_MAX;
    return STDA_HIP_OK;
}

int stda_hip_batch_set_precision(stda_hip_batch *b, int32_t precision) {
    if (!b || precision < 1 || precision > 2) return STDA_HIP_ERR_ARG;
    if (b->engine) {
        b->last_error = "continuous mode is running (stda_hip_batch_stop first)";
        return STDA_HIP_ERR_STATE;
    }
    for (auto *c : b->workers) c->precision = precision;
    return STDA_HIP_OK;
}

int stda_hip_batch_wide_count(const stda_hip_batch *b, int64_t *nwide) {
    if (!b || !nwide) return STDA_HIP_ERR_ARG;
    *nwide = b->nwide;
    return STDA_HIP_OK;
}

int stda_hip_batch_kernel_bytes(const stda_hip_batch *b, double *bytes) {
    if (!b || !bytes) return STDA_HIP_ERR_ARG;
    *bytes = b->kernel_bytes;
    return STDA_HIP_OK;
}

int stda_hip_batch_kernel_stats(stda_hip_batch *b, const char *name, double *ms,
                                double *launches, double *work) {
    if (!b || !name) return STDA_HIP_ERR_ARG;
    double t[3] = {0, 0, 0};
    if (std::string(name) == "batch_lockstep") {
        t[0] = b->lockstep_ms;
        t[1] = b->lockstep_launches;
        t[2] = b->kernel_bytes;
    }
    for (auto *c : b->workers) {
        c->flush_timers();
        for (const auto *v : {&c->stats, &c->stats_total})
            for (const auto &kv : *v)
                if (kv.first == name) {
                    t[0] += kv.second.ms;
                    t[1] += kv.second.launches;
                    t[2] += kv.second.work;
                }
    }
    if (ms) *ms = t[0];
    if (launches) *launches = t[1];
    if (work) *work = t[2];
    return STDA_HIP_OK;
}

int stda_hip_batch_phase_times(const stda_hip_batch *b, double *us10) {
    if (!b || !us10) return STDA_HIP_ERR_ARG;
    std::copy(b->phase_us, b->phase_us + 10, us10);
    return STDA_HIP_OK;
}

int stda_hip_batch_timings(const stda_hip_batch *b, double *ms5, int64_t *counts3) {
    if (!b) return STDA_HIP_ERR_ARG;
    if (ms5) std::copy(b->timings, b->timings + 5, ms5);
    if (counts3) std::copy(b->counts, b->counts + 3, counts3);
    return STDA_HIP_OK;
}

int stda_hip_batch_solve_tda(stda_hip_batch *b, int32_t nmol, const stda_hip_system *sys,
                             const stda_hip_params *par, int32_t nroots, double tol,
                             int32_t velcorr, int32_t *nfound, double *omega, double *props,
                             int32_t *status) {
    if (!b || nmol < 0 || (nmol > 0 && (!sys || !par || !nfound || !omega || !props || !status)) ||
        nroots < 1)
        return STDA_HIP_ERR_ARG;
    if (tol <= 0.0) tol = 1.0e-8;
    const double tol32 = std::max(tol, 1.0e-6);
    if (b->engine) {
        b->last_error = "continuous mode is running (stda_hip_batch_stop first)";
        return STDA_HIP_ERR_STATE;
    }
    try {
        Timer t_all;
        for (auto &x : b->counts) x = 0;
        b->nwide = 0;
        for (auto &x : b->timings) x = 0.0;
        for (auto &x : b->phase_us) x = 0.0;
        b->kernel_bytes = 0.0;
        b->lockstep_ms = 0.0;
        b->lockstep_launches = 0.0;
        for (auto *c : b->workers) {
            c->flush_timers();
            c->stats.clear();
            c->stats_total.clear();
        }
        STDA_HIP_CHECK(hipSetDevice(b->device));
        hipStream_t ds = b->dstream;
        stda_hip_ctx *c0 = b->workers[0];
        std::vector<BatchItem> items(nmol);
        std::vector<char> batched(nmol, 0);
        std::vector<std::string> errs(nmol);
        std::fill(nfound, nfound + nmol, 0);
        std::fill(omega, omega + (size_t)nmol * nroots, 0.0);
        std::fill(props, props + (size_t)nmol * nroots * 5, 0.0);
        std::vector<int> order(nmol);
        std::iota(order.begin(), order.end(), 0);
        // largest first (by pair count): the long single-path solves start
        // early and the short ones fill in behind them, instead of the largest
        // molecules forming the tail of the call.  (Preparing the molecules of
        // the matrix-free lock step first, to overlap its iterations with the
        // single-path solves, measured 3-7 % slower: the device is already
        // saturated and the overlap only stretches both.)
        std::stable_sort(order.begin(), order.end(), [&](int a, int c) {
            return (int64_t)sys[a].no * sys[a].nv > (int64_t)sys[c].no * sys[c].nv;
        });
        // state table and outputs of every molecule, indexed by molecule
        DBuf<int> stated;
        DBuf<double> outd;
        std::vector<int> stateh((size_t)std::max(nmol, 1) * BATCH_STATE, B_DONE);
        {
            AllocScope scope(ds);
            stated.resize((size_t)std::max(nmol, 1) * BATCH_STATE);
            outd.resize((size_t)std::max(nmol, 1) * BATCH_OUT);
        }
        // hand-over from the workers: prepared molecule and the event after its setup
        std::mutex mu;
        std::condition_variable cv;
        std::vector<std::pair<int, hipEvent_t>> ready;
        int workers_left = (int)b->workers.size();
        const int nw = (int)b->workers.size();
        std::vector<std::string> werr(nw);
        std::atomic<int> next{0};
        std::vector<std::array<double, 3>> wus(nw, std::array<double, 3>{0, 0, 0});
        auto work = [&](int wi) {
            stda_hip_ctx *c = b->workers[wi];
            AllocScope scope(c->stream);
            try {
                STDA_HIP_CHECK(hipSetDevice(c->device));
                for (int q = next.fetch_add(1); q < nmol; q = next.fetch_add(1)) {
                    const int i = order[q];
                    try {
                        batched[i] = batch_prepare(c, &sys[i], &par[i], nroots, tol, velcorr,
                                                   b->nmax_batched, b->wide ? b->nmax_wide : 0, items[i], &nfound[i],
                                                   omega + (size_t)i * nroots,
                                                   props + (size_t)i * nroots * 5, wus[wi].data());
                        status[i] = STDA_HIP_OK;
                    } catch (const Error &e) {
                        if (e.status == STDA_HIP_ERR_DEVICE) throw;
                        status[i] = e.status;
                        errs[i] = e.what();
                        continue;
                    }
                    if (batched[i]) {
                        hipEvent_t ev;
                        STDA_HIP_CHECK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
                        STDA_HIP_CHECK(hipEventRecord(ev, c->stream));
                        std::lock_guard<std::mutex> lk(mu);
                        ready.emplace_back(i, ev);
                        cv.notify_all();
                    }
                }
            } catch (const std::exception &e) {
                werr[wi] = e.what();
            }
            std::lock_guard<std::mutex> lk(mu);
            --workers_left;
            cv.notify_all();
        };
        std::vector<std::thread> th;
        // joins the workers on every way out of this scope (declared after
        // everything they use, so it runs before those are destroyed)
        struct JoinAll {
            std::vector<std::thread> &t;
            ~JoinAll() {
                for (auto &x : t)
                    if (x.joinable()) x.join();
            }
        } join_all{th};
        for (int w = 0; w < nw; ++w) th.emplace_back(work, w);

        // driver: lock-step iterations over the active set
        Timer t_drv;
        double drv_ms = 0.0;
        {
            AllocScope scope(ds);
            std::vector<int> active;
            bool dirty = false;
            std::vector<std::array<hipEvent_t, 2>> evs;
            int iter = 0;
            bool feeding = true;
            // The `group` lock-step iterations between two reads of the state
            // table are one captured graph, rebuilt when the active set changes.
            LockstepSet set;
            while (true) {
                std::vector<std::pair<int, hipEvent_t>> got;
                {
                    std::unique_lock<std::mutex> lk(mu);
                    if ((active.empty() || b->two_phase) && workers_left > 0)
                        cv.wait_for(lk, std::chrono::milliseconds(2));
                    got.swap(ready);
                    feeding = workers_left > 0;
                }
                for (auto &gi : got) {
                    STDA_HIP_CHECK(hipStreamWaitEvent(ds, gi.second, 0));
                    (void)hipEventDestroy(gi.second);
                    active.push_back(gi.first);
                    stateh[(size_t)gi.first * BATCH_STATE] = B_INIT;
                    STDA_HIP_CHECK(hipMemcpyAsync(stated.get() + (size_t)gi.first * BATCH_STATE,
                                                  &stateh[(size_t)gi.first * BATCH_STATE],
                                                  sizeof(int), hipMemcpyHostToDevice, ds));
                    dirty = true;
                }
                if (active.empty()) {
                    if (!feeding) {
                        std::lock_guard<std::mutex> lk(mu);
                        if (ready.empty() && workers_left == 0) break;
                    }
                    continue;
                }
                if (b->two_phase && feeding) continue;   // lock step once every molecule is prepared
                if (dirty) {
                    std::vector<LockstepSet::Entry> ent;
                    for (int q : active)
                        ent.push_back({&items[q], tol32, outd.get() + (size_t)q * BATCH_OUT,
                                       stated.get() + (size_t)q * BATCH_STATE});
                    set.build(std::move(ent), ds);
                    dirty = false;
                }
                std::array<hipEvent_t, 2> e{};
                for (auto &x : e) STDA_HIP_CHECK(hipEventCreate(&x));
                STDA_HIP_CHECK(hipEventRecord(e[0], ds));
                set.launch(ds, b->group);
                STDA_HIP_CHECK(hipEventRecord(e[1], ds));
                evs.push_back(e);
                iter += b->group;
                {
                    // poll: drop the finished molecules from the active set
                    STDA_HIP_CHECK(hipMemcpyAsync(stateh.data(), stated.get(),
                                                  stateh.size() * sizeof(int),
                                                  hipMemcpyDeviceToHost, ds));
                    STDA_HIP_CHECK(hipStreamSynchronize(ds));
                    const size_t before = active.size();
                    active.erase(std::remove_if(active.begin(), active.end(),
                                                [&](int i) {
                                                    return stateh[(size_t)i * BATCH_STATE] == B_DONE;
                                                }),
                                 active.end());
                    if (active.size() != before) dirty = true;
                }
                if (iter > 100000) throw Error(STDA_HIP_ERR_DEVICE, "batched driver did not finish");
            }
            STDA_HIP_CHECK(hipStreamSynchronize(ds));
            for (auto &e : evs) {
                float m0 = 0.f;
                STDA_HIP_CHECK(hipEventElapsedTime(&m0, e[0], e[1]));
                b->lockstep_ms += m0;
                b->lockstep_launches += 1.0;
                b->timings[4] += m0;
                for (auto &x : e) (void)hipEventDestroy(x);
            }
        }
        drv_ms = t_drv.ms();
        for (auto &t : th)
            if (t.joinable()) t.join();
        for (int w = 0; w < nw; ++w) {
            if (!werr[w].empty()) throw Error(STDA_HIP_ERR_DEVICE, werr[w]);
            for (int q = 0; q < 3; ++q) b->phase_us[7 + q] += wus[w][q];
        }
        b->timings[0] = t_all.ms();
        b->timings[1] = drv_ms;

        // property tables of the batched molecules
        Timer t3;
        std::vector<double> outh((size_t)std::max(nmol, 1) * BATCH_OUT);
        outd.download(outh.data(), outh.size(), ds);
        STDA_HIP_CHECK(hipStreamSynchronize(ds));
        const double us_per_tick = 1.0e6 / b->wall_tick_hz;
        for (int i = 0; i < nmol; ++i) {
            if (!batched[i] || status[i] != STDA_HIP_OK) continue;
            const double *o = outh.data() + (size_t)i * BATCH_OUT;
            const double *info = o + 20 * BATCH_KMAX;
            const int kk = (int)info[2];
            const bool conv = info[1] != 0.0;
            // the FP32 A once per iteration, the FP64 A and 0.5 B once
            const double n2 = (double)items[i].n * (double)items[i].n;
            if (!items[i].wide)
                b->kernel_bytes += info[0] * 4.0 * n2 + 8.0 * n2 + (items[i].velcorr ? 8.0 * n2 : 0.0);
            else
                ++b->nwide;
            for (int ph = 0; ph < 5; ++ph) b->phase_us[ph] += info[4 + ph] * us_per_tick;
            b->phase_us[6] += info[10];
            std::vector<double> w(o, o + kk);
            std::vector<double> pr((size_t)5 * kk);
            tda_props_from_moments(items[i].ak, w, o + BATCH_KMAX,
                                   items[i].velcorr ? o + 10 * BATCH_KMAX : nullptr, pr.data(),
                                   nullptr);
            nfound[i] = kk;
            for (int j = 0; j < kk; ++j) {
                omega[(size_t)i * nroots + j] = w[j];
                for (int r = 0; r < 5; ++r) props[((size_t)i * nroots + j) * 5 + r] = pr[5 * j + r];
            }
            if (!conv) {
                status[i] = STDA_HIP_ERR_NOT_CONVERGED;
                errs[i] = "batched Davidson did not converge";
            } else {
                ++b->counts[0];   // the three counts are disjoint: batched OK, single OK, nonzero status
            }
        }
        b->timings[2] = t3.ms();
        items.clear();
        (void)c0;
        b->timings[3] = t_all.ms();
        b->last_error.clear();
        for (int i = 0; i < nmol; ++i) {
            if (status[i] != STDA_HIP_OK) {
                ++b->counts[2];
                if (b->last_error.empty())
                    b->last_error = "molecule " + std::to_string(i) + ": " + errs[i];
            } else if (!batched[i]) {
                ++b->counts[1];
            }
        }
        return STDA_HIP_OK;
    } catch (const Error &e) {
        b->last_error = e.what();
        return e.status;
    } catch (const std::exception &e) {
        b->last_error = e.what();
        return STDA_HIP_ERR_DEVICE;
    }
}

}  // extern "C"

// ------------------------------------------------------------------------------------------
// Continuous mode of a batch object.  Submitted molecules are queued; the
// worker contexts take them one at a time (setup pipeline, explicit matrices,
// or the whole single-path solve) while a driver thread keeps the lock-step
// iterations of every prepared molecule running: a molecule joins the active
// set at the driver's next pass after its setup and leaves it, with its
// result, at the first poll that finds it done.  Each molecule owns one of
// kSlots rows of the device state and output tables.
namespace {

struct BatchTicket {
    stda_hip_system sys{};
    stda_hip_params par{};
    int nroots = 0, velcorr = 1;
    double tol = 1.0e-8;
    std::vector<double> omega, props;
    int32_t nfound = 0, status = STDA_HIP_OK;
    std::string err;
    bool done = false;
    BatchItem item;
    int slot = -1;
};

}  // namespace

struct BatchStream {
    static constexpr int kSlots = 2048;
    std::mutex mu;
    std::condition_variable cv;
    std::deque<int64_t> queue;                          // submitted, not yet taken by a worker
    std::vector<std::pair<int64_t, hipEvent_t>> ready;  // prepared for the lock-step driver
    std::map<int64_t, std::unique_ptr<BatchTicket>> tickets;
    int64_t next = 1;
    int busy = 0;                                       // workers inside a preparation
    bool stop = false;
    std::vector<std::thread> threads;
};

namespace {

void stream_worker(stda_hip_batch *b, BatchStream *E, int wi) {
    stda_hip_ctx *c = b->workers[wi];
    AllocScope scope(c->stream);
    (void)hipSetDevice(c->device);
    while (true) {
        int64_t id = 0;
        BatchTicket *t = nullptr;
        {
            std::unique_lock<std::mutex> lk(E->mu);
            E->cv.wait(lk, [&] { return E->stop || !E->queue.empty(); });
            if (E->queue.empty()) return;
            id = E->queue.front();
            E->queue.pop_front();
            t = E->tickets[id].get();
            ++E->busy;
        }
        bool batched = false;
        hipEvent_t ev = nullptr;
        try {
            double us3[3] = {0, 0, 0};
            batched = batch_prepare(c, &t->sys, &t->par, t->nroots, t->tol, t->velcorr,
                                    b->nmax_batched, b->wide ? b->nmax_wide : 0, t->item, &t->nfound, t->omega.data(),
                                    t->props.data(), us3);
            t->status = STDA_HIP_OK;
            if (batched) {
                STDA_HIP_CHECK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
                STDA_HIP_CHECK(hipEventRecord(ev, c->stream));
            }
        } catch (const std::exception &e) {
            const Error *er = dynamic_cast<const Error *>(&e);
            t->status = er ? er->status : STDA_HIP_ERR_DEVICE;
            t->err = e.what();
            t->item = BatchItem();
            batched = false;
        }
        std::lock_guard<std::mutex> lk(E->mu);
        --E->busy;
        if (batched) E->ready.emplace_back(id, ev);
        else t->done = true;
        E->cv.notify_all();
    }
}

void stream_driver(stda_hip_batch *b, BatchStream *E) {
    (void)hipSetDevice(b->device);
    hipStream_t ds = b->dstream;
    AllocScope scope(ds);
    constexpr int S = BatchStream::kSlots;
    std::string fatal;
    try {
        DBuf<int> stated((size_t)S * BATCH_STATE);
        DBuf<double> outd((size_t)S * BATCH_OUT);
        std::vector<int> stateh((size_t)S * BATCH_STATE, B_DONE);
        stated.upload(stateh.data(), stateh.size(), ds);   // free rows read as done
        std::vector<double> outh(BATCH_OUT);
        std::vector<int> free_slots;
        for (int q = S - 1; q >= 0; --q) free_slots.push_back(q);
        std::vector<BatchTicket *> active;
        bool dirty = false;
        LockstepSet set;
        while (true) {
            std::vector<std::pair<int64_t, hipEvent_t>> got;
            std::vector<BatchTicket *> gott;
            {
                std::unique_lock<std::mutex> lk(E->mu);
                if (active.empty())
                    E->cv.wait_for(lk, std::chrono::milliseconds(2),
                                   [&] { return E->stop || !E->ready.empty(); });
                if (active.empty() && E->ready.empty() && E->stop && E->queue.empty() &&
                    E->busy == 0)
                    break;
                while (!E->ready.empty() && free_slots.size() > got.size()) {
                    got.push_back(E->ready.front());
                    gott.push_back(E->tickets[E->ready.front().first].get());
                    E->ready.erase(E->ready.begin());
                }
            }
            for (size_t q = 0; q < got.size(); ++q) {
                STDA_HIP_CHECK(hipStreamWaitEvent(ds, got[q].second, 0));
                (void)hipEventDestroy(got[q].second);
                BatchTicket *t = gott[q];
                t->slot = free_slots.back();
                free_slots.pop_back();
                stateh[(size_t)t->slot * BATCH_STATE] = B_INIT;
                STDA_HIP_CHECK(hipMemcpyAsync(stated.get() + (size_t)t->slot * BATCH_STATE,
                                              &stateh[(size_t)t->slot * BATCH_STATE], sizeof(int),
                                              hipMemcpyHostToDevice, ds));
                active.push_back(t);
                dirty = true;
            }
            if (active.empty()) continue;
            if (dirty) {
                std::vector<LockstepSet::Entry> ent;
                for (BatchTicket *t : active)
                    ent.push_back({&t->item, std::max(t->tol, 1.0e-6),
                                   outd.get() + (size_t)t->slot * BATCH_OUT,
                                   stated.get() + (size_t)t->slot * BATCH_STATE});
                set.build(std::move(ent), ds);
                dirty = false;
            }
            set.launch(ds, b->group);
            // poll the state table (one copy)
            STDA_HIP_CHECK(hipMemcpyAsync(stateh.data(), stated.get(), stateh.size() * sizeof(int),
                                          hipMemcpyDeviceToHost, ds));
            STDA_HIP_CHECK(hipStreamSynchronize(ds));
            std::vector<BatchTicket *> still;
            for (BatchTicket *t : active) {
                if (stateh[(size_t)t->slot * BATCH_STATE] != B_DONE) {
                    still.push_back(t);
                    continue;
                }
                STDA_HIP_CHECK(hipMemcpyAsync(outh.data(), outd.get() + (size_t)t->slot * BATCH_OUT,
                                              (size_t)BATCH_OUT * 8, hipMemcpyDeviceToHost, ds));
                STDA_HIP_CHECK(hipStreamSynchronize(ds));
                const double *o = outh.data();
                const double *info = o + 20 * BATCH_KMAX;
                const int kk = (int)info[2];
                std::vector<double> w(o, o + kk), pr((size_t)5 * kk);
                tda_props_from_moments(t->item.ak, w, o + BATCH_KMAX,
                                       t->item.velcorr ? o + 10 * BATCH_KMAX : nullptr, pr.data(),
                                       nullptr);
                t->nfound = kk;
                for (int j = 0; j < kk; ++j) {
                    t->omega[j] = w[j];
                    for (int r = 0; r < 5; ++r) t->props[(size_t)5 * j + r] = pr[(size_t)5 * j + r];
                }
                if (info[1] == 0.0) {
                    t->status = STDA_HIP_ERR_NOT_CONVERGED;
                    t->err = "batched Davidson did not converge";
                }
                t->item = BatchItem();   // the device work on it has completed
                free_slots.push_back(t->slot);
                std::lock_guard<std::mutex> lk(E->mu);
                t->done = true;
                E->cv.notify_all();
            }
            if (still.size() != active.size()) {
                active.swap(still);
                dirty = true;
            }
        }
    } catch (const std::exception &e) {
        fatal = e.what();
    }
    if (!fatal.empty()) {
        // the device failed: every open ticket ends with the error
        std::lock_guard<std::mutex> lk(E->mu);
        for (auto &kv : E->tickets)
            if (!kv.second->done) {
                kv.second->status = STDA_HIP_ERR_DEVICE;
                kv.second->err = fatal;
                kv.second->done = true;
            }
        E->stop = true;
        E->cv.notify_all();
    }
}

}  // namespace

extern "C" {

int stda_hip_batch_submit(stda_hip_batch *b, const stda_hip_system *sys,
                          const stda_hip_params *par, int32_t nroots, double tol,
                          int32_t velcorr, int64_t *ticket) {
    if (!b || !sys || !par || !ticket || nroots < 1) return STDA_HIP_ERR_ARG;
    if (!b->engine) {
        auto *E = new BatchStream();
        b->engine = E;
        for (int w = 0; w < (int)b->workers.size(); ++w)
            E->threads.emplace_back(stream_worker, b, E, w);
        E->threads.emplace_back(stream_driver, b, E);
    }
    BatchStream *E = b->engine;
    auto t = std::make_unique<BatchTicket>();
    t->sys = *sys;
    t->par = *par;
    t->nroots = nroots;
    t->tol = tol > 0.0 ? tol : 1.0e-8;
    t->velcorr = velcorr;
    t->omega.assign(nroots, 0.0);
    t->props.assign((size_t)5 * nroots, 0.0);
    std::lock_guard<std::mutex> lk(E->mu);
    if (E->stop) return STDA_HIP_ERR_STATE;
    const int64_t id = E->next++;
    E->tickets[id] = std::move(t);
    E->queue.push_back(id);
    E->cv.notify_all();
    *ticket = id;
    return STDA_HIP_OK;
}

int stda_hip_batch_result(stda_hip_batch *b, int64_t ticket, int32_t wait, int32_t *done,
                          int32_t *nfound, double *omega, double *props, int32_t *status) {
    if (!b || !done) return STDA_HIP_ERR_ARG;
    BatchStream *E = b->engine;
    if (!E) return STDA_HIP_ERR_STATE;
    std::unique_lock<std::mutex> lk(E->mu);
    auto it = E->tickets.find(ticket);
    if (it == E->tickets.end()) return STDA_HIP_ERR_ARG;
    BatchTicket *t = it->second.get();
    if (wait) E->cv.wait(lk, [&] { return t->done; });
    *done = t->done ? 1 : 0;
    if (!t->done) return STDA_HIP_OK;
    if (nfound) *nfound = t->nfound;
    if (omega) std::copy(t->omega.begin(), t->omega.end(), omega);
    if (props) std::copy(t->props.begin(), t->props.end(), props);
    if (status) *status = t->status;
    if (t->status != STDA_HIP_OK) b->last_error = t->err;
    E->tickets.erase(it);
    return STDA_HIP_OK;
}

int stda_hip_batch_stop(stda_hip_batch *b) {
    if (!b || !b->engine) return STDA_HIP_OK;
    BatchStream *E = b->engine;
    {
        std::lock_guard<std::mutex> lk(E->mu);
        E->stop = true;
        E->cv.notify_all();
    }
    for (auto &th : E->threads)
        if (th.joinable()) th.join();
    delete E;
    b->engine = nullptr;
    return STDA_HIP_OK;
}

}  // extern "C"
