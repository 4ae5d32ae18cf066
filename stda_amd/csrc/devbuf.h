// Device buffers and BLAS/LAPACK glue for the host side of the core.
#pragma once

#include <rocblas/rocblas.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include "../../include/stda_hip.h"
#include "common.h"

namespace stda {

// Released device blocks kept per stream for reuse on the same stream
// (stream order makes the reuse safe), so that solving molecule after
// molecule does not go through the device allocator, whose calls serialise
// the host threads of concurrent contexts.  At most kCacheBytes per stream
// are kept (larger releases go back to the allocator); a stream's blocks
// are returned with block_cache_drain before the stream is destroyed.
class BlockCache {
  public:
    static constexpr size_t kCacheBytes = size_t(2) << 30;
    // a kept block of at least `bytes` and at most 4x that (plus 1 MiB), or null
    void *take(size_t bytes, hipStream_t s, size_t *got) {
        std::lock_guard<std::mutex> lk(mu_);
        auto it = free_.find(s);
        if (it == free_.end()) return nullptr;
        auto &v = it->second.blocks;
        size_t best = v.size();
        for (size_t q = 0; q < v.size(); ++q)
            if (v[q].second >= bytes && v[q].second <= 4 * bytes + (size_t(1) << 20) &&
                (best == v.size() || v[q].second < v[best].second))
                best = q;
        if (best == v.size()) return nullptr;
        void *p = v[best].first;
        *got = v[best].second;
        it->second.bytes -= *got;
        v[best] = v.back();
        v.pop_back();
        return p;
    }
    void give(void *p, size_t bytes, hipStream_t s) {
        {
            std::lock_guard<std::mutex> lk(mu_);
            auto &e = free_[s];
            if (e.bytes + bytes <= kCacheBytes) {
                e.blocks.emplace_back(p, bytes);
                e.bytes += bytes;
                return;
            }
        }
        (void)hipFreeAsync(p, s);
    }
    void drain(hipStream_t s) {
        std::vector<std::pair<void *, size_t>> v;
        {
            std::lock_guard<std::mutex> lk(mu_);
            auto it = free_.find(s);
            if (it == free_.end()) return;
            v.swap(it->second.blocks);
            free_.erase(it);
        }
        for (auto &b : v) (void)hipFreeAsync(b.first, s);
    }
    // every stream's blocks (when the allocator runs out of memory)
    void drain_all() {
        std::unordered_map<hipStream_t, Entry> all;
        {
            std::lock_guard<std::mutex> lk(mu_);
            all.swap(free_);
        }
        for (auto &kv : all)
            for (auto &b : kv.second.blocks) (void)hipFreeAsync(b.first, kv.first);
    }

  private:
    struct Entry {
        std::vector<std::pair<void *, size_t>> blocks;
        size_t bytes = 0;
    };
    std::mutex mu_;
    std::unordered_map<hipStream_t, Entry> free_;
};

// one per process, never destroyed (no device calls at static destruction)
inline BlockCache &block_cache() {
    static BlockCache *c = new BlockCache();
    return *c;
}

inline void block_cache_drain(hipStream_t s) { block_cache().drain(s); }

// Device memory left free for the runtime's own allocations: the scratch of
// a kernel with a private segment is allocated when it is dispatched,
// outside the stream-ordered pool, and a dispatch that cannot get it aborts
// the process.  The pool keeps released blocks (its release threshold is
// unlimited), so a new block that would leave less than this free first
// returns the cached blocks and the pool's unused memory to the device, and
// fails with STDA_HIP_ERR_NOMEM if that is not enough.  4 GiB unless the
// environment's STDA_HIP_DEVICE_RESERVE_MB (read once per process) says
// otherwise, e.g. for a GPU shared with other processes.
inline size_t device_reserve_bytes() {
    static const size_t r = [] {
        const char *e = std::getenv("STDA_HIP_DEVICE_RESERVE_MB");
        if (e && *e) {
            const double mb = std::strtod(e, nullptr);
            if (mb >= 0.0) return (size_t)std::min(mb * 1048576.0, 1.0e18);
        }
        return size_t(4) << 30;
    }();
    return r;
}

// Device memory a new buffer can draw on beyond the reserve: what the device
// has free plus what the pool holds unused (hipMemGetInfo counts all of the
// pool's memory as used, so after busy phases it alone understates what is
// there).  For sizing work buffers to the memory at hand.
inline size_t device_available_bytes() {
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) fr = 0;
    int dev = 0;
    hipMemPool_t pool = nullptr;
    uint64_t resv = 0, used = 0;
    if (hipGetDevice(&dev) == hipSuccess && hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess &&
        hipMemPoolGetAttribute(pool, hipMemPoolAttrReservedMemCurrent, &resv) == hipSuccess &&
        hipMemPoolGetAttribute(pool, hipMemPoolAttrUsedMemCurrent, &used) == hipSuccess &&
        resv > used)
        fr += (size_t)(resv - used);
    (void)hipGetLastError();
    const size_t rs = device_reserve_bytes();
    return fr > rs ? fr - rs : 0;
}

inline void *device_alloc(size_t bytes, hipStream_t s) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &cs) != hipSuccess) cs = hipStreamCaptureStatusNone;
    size_t fr = 0, tot = 0;
    const size_t rs = device_reserve_bytes();
    if (cs == hipStreamCaptureStatusNone && hipMemGetInfo(&fr, &tot) == hipSuccess &&
        fr < bytes + rs) {
        block_cache().drain_all();
        int dev = 0;
        hipMemPool_t pool = nullptr;
        // the released blocks return to the pool in stream order: wait for
        // them, then hand every unused byte of the pool back
        if (hipDeviceSynchronize() == hipSuccess && hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess)
            (void)hipMemPoolTrimTo(pool, 0);
        (void)hipGetLastError();
        if (hipMemGetInfo(&fr, &tot) == hipSuccess && fr < bytes + rs)
            throw Error(STDA_HIP_ERR_NOMEM,
                        "device allocation of " + std::to_string(bytes >> 20) +
                            " MiB would leave less than the runtime's reserve free (" +
                            std::to_string(fr >> 20) + " MiB free)");
    }
    void *q = nullptr;
    if (hipMallocAsync(&q, bytes, s) != hipSuccess) {
        (void)hipGetLastError();
        block_cache().drain_all();
        STDA_HIP_CHECK(hipMallocAsync(&q, bytes, s));
    }
    return q;
}

// Stream that orders the device allocations of the calling thread.  Every
// entry point sets it to its context's stream (AllocScope), so that buffers
// come from the device's stream-ordered pool (hipMallocAsync/hipFreeAsync)
// and allocating or freeing never synchronises the whole device: contexts
// driven from several host threads keep overlapping.
inline hipStream_t &alloc_stream() {
    static thread_local hipStream_t s = nullptr;
    return s;
}

struct AllocScope {
    hipStream_t prev;
    explicit AllocScope(hipStream_t s) : prev(alloc_stream()) { alloc_stream() = s; }
    ~AllocScope() { alloc_stream() = prev; }
    AllocScope(const AllocScope &) = delete;
    AllocScope &operator=(const AllocScope &) = delete;
};

template <typename T>
class DBuf {
  public:
    DBuf() = default;
    explicit DBuf(size_t n) { resize(n); }
    ~DBuf() { release(); }
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    DBuf(DBuf &&o) noexcept : p_(o.p_), n_(o.n_), cap_(o.cap_), s_(o.s_) {
        o.p_ = nullptr;
        o.n_ = o.cap_ = 0;
    }
    DBuf &operator=(DBuf &&o) noexcept {
        if (this != &o) {
            release();
            p_ = o.p_; n_ = o.n_; cap_ = o.cap_; s_ = o.s_;
            o.p_ = nullptr; o.n_ = o.cap_ = 0;
        }
        return *this;
    }
    // n elements; the allocation is kept when it is large enough (buffers of
    // a context are reused from one molecule to the next without allocating)
    void resize(size_t n) {
        if (n <= cap_) {
            n_ = n;
            return;
        }
        release();
        s_ = alloc_stream();
        size_t got = 0;
        void *q = block_cache().take(n * sizeof(T), s_, &got);
        if (!q) {
            got = n * sizeof(T);
            q = device_alloc(got, s_);
        }
        p_ = static_cast<T *>(q);
        n_ = n;
        cap_ = got / sizeof(T);
    }
    void ensure(size_t n) { if (n > n_) resize(n); }
    // kept for reuse on the stream it was allocated on (or freed in its order)
    void release() {
        if (p_) block_cache().give(p_, cap_ * sizeof(T), s_);
        p_ = nullptr;
        n_ = cap_ = 0;
    }
    T *get() const { return p_; }
    size_t size() const { return n_; }
    void upload(const T *h, size_t n, hipStream_t s) {
        resize(n);
        if (n) STDA_HIP_CHECK(hipMemcpyAsync(p_, h, n * sizeof(T), hipMemcpyHostToDevice, s));
    }
    void download(T *h, size_t n, hipStream_t s) const {
        if (n) STDA_HIP_CHECK(hipMemcpyAsync(h, p_, n * sizeof(T), hipMemcpyDeviceToHost, s));
    }
    void zero(hipStream_t s) {
        if (n_) STDA_HIP_CHECK(hipMemsetAsync(p_, 0, n_ * sizeof(T), s));
    }

  private:
    T *p_ = nullptr;
    size_t n_ = 0, cap_ = 0;
    hipStream_t s_ = nullptr;
};

#define STDA_BLAS_CHECK(expr)                                                           \
    do {                                                                                \
        rocblas_status st_ = (expr);                                                    \
        if (st_ != rocblas_status_success)                                              \
            throw ::stda::Error(3, std::string(#expr) + ": " + rocblas_status_to_string(st_)); \
    } while (0)

// C = alpha op(A) op(B) + beta C (column-major) on the matrix cores (gemm.hip),
// on the stream of the context's BLAS handle (kept for rocSOLVER).
inline hipStream_t handle_stream(rocblas_handle h) {
    hipStream_t s = nullptr;
    STDA_BLAS_CHECK(rocblas_get_stream(h, &s));
    return s;
}

inline void gemm_dims(int64_t m, int64_t n, int64_t k, int64_t batch) {
    if (m > INT32_MAX || n > INT32_MAX || k > INT32_MAX || batch > INT32_MAX)
        throw Error(9, "gemm dimension beyond 2^31");
}

inline void dgemm(rocblas_handle h, bool ta, bool tb, int64_t m, int64_t n, int64_t k,
                  double alpha, const double *A, int64_t lda, const double *B, int64_t ldb,
                  double beta, double *C, int64_t ldc) {
    if (m == 0 || n == 0) return;
    gemm_dims(m, n, k, 1);
    launch_gemm_f64(handle_stream(h), ta, tb, (int)m, (int)n, (int)k, alpha, A, lda, 0, B, ldb,
                    0, beta, C, ldc, 0, 1);
}

inline void dgemm_batched(rocblas_handle h, bool ta, bool tb, int64_t m, int64_t n, int64_t k,
                          double alpha, const double *A, int64_t lda, int64_t sa, const double *B,
                          int64_t ldb, int64_t sb, double beta, double *C, int64_t ldc,
                          int64_t sc, int64_t batch) {
    if (m == 0 || n == 0 || batch == 0) return;
    gemm_dims(m, n, k, batch);
    for (int64_t b0 = 0; b0 < batch; b0 += 65535)
        launch_gemm_f64(handle_stream(h), ta, tb, (int)m, (int)n, (int)k, alpha, A + sa * b0, lda,
                        sa, B + sb * b0, ldb, sb, beta, C + sc * b0, ldc, sc,
                        (int)std::min<int64_t>(65535, batch - b0));
}

// FP32 operands: mode GEMM_F32 or GEMM_BF16X3
inline void sgemm(rocblas_handle h, bool ta, bool tb, int64_t m, int64_t n, int64_t k, float alpha,
                  const float *A, int64_t lda, const float *B, int64_t ldb, float beta, float *C,
                  int64_t ldc, int mode = GEMM_F32) {
    if (m == 0 || n == 0) return;
    gemm_dims(m, n, k, 1);
    launch_gemm_f32(handle_stream(h), mode, ta, tb, (int)m, (int)n, (int)k, alpha, A, lda, 0, B,
                    ldb, 0, beta, C, ldc, 0, 1);
}

inline void sgemm_batched(rocblas_handle h, bool ta, bool tb, int64_t m, int64_t n, int64_t k,
                          float alpha, const float *A, int64_t lda, int64_t sa, const float *B,
                          int64_t ldb, int64_t sb, float beta, float *C, int64_t ldc, int64_t sc,
                          int64_t batch, int mode = GEMM_F32) {
    if (m == 0 || n == 0 || batch == 0) return;
    gemm_dims(m, n, k, batch);
    for (int64_t b0 = 0; b0 < batch; b0 += 65535)
        launch_gemm_f32(handle_stream(h), mode, ta, tb, (int)m, (int)n, (int)k, alpha, A + sa * b0,
                        lda, sa, B + sb * b0, ldb, sb, beta, C + sc * b0, ldc, sc,
                        (int)std::min<int64_t>(65535, batch - b0));
}

// Host symmetric eigensolver (LAPACK dsyevd): eigenvalues ascending, vectors
// in place of the column-major n x n matrix a.
void host_syev(int n, std::vector<double> &a, std::vector<double> &w);

}  // namespace stda
