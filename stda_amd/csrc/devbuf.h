// Device buffers and BLAS/LAPACK glue for the host side of the core.
#pragma once

#include <rocblas/rocblas.h>

#include <cstring>
#include <vector>

#include "common.h"

namespace stda {

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
        STDA_HIP_CHECK(hipMallocAsync(reinterpret_cast<void **>(&p_), n * sizeof(T), s_));
        n_ = cap_ = n;
    }
    void ensure(size_t n) { if (n > n_) resize(n); }
    // freed in the order of the stream it was allocated on
    void release() {
        if (p_) (void)hipFreeAsync(p_, s_);
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

// C = alpha op(A) op(B) + beta C (column-major)
inline void dgemm(rocblas_handle h, bool ta, bool tb, int64_t m, int64_t n, int64_t k,
                  double alpha, const double *A, int64_t lda, const double *B, int64_t ldb,
                  double beta, double *C, int64_t ldc) {
    if (m == 0 || n == 0) return;
    STDA_BLAS_CHECK(rocblas_dgemm(h, ta ? rocblas_operation_transpose : rocblas_operation_none,
                                  tb ? rocblas_operation_transpose : rocblas_operation_none,
                                  (rocblas_int)m, (rocblas_int)n, (rocblas_int)k, &alpha, A,
                                  (rocblas_int)lda, B, (rocblas_int)ldb, &beta, C,
                                  (rocblas_int)ldc));
}

inline void dgemm_batched(rocblas_handle h, bool ta, bool tb, int64_t m, int64_t n, int64_t k,
                          double alpha, const double *A, int64_t lda, int64_t sa, const double *B,
                          int64_t ldb, int64_t sb, double beta, double *C, int64_t ldc,
                          int64_t sc, int64_t batch) {
    if (m == 0 || n == 0 || batch == 0) return;
    STDA_BLAS_CHECK(rocblas_dgemm_strided_batched(
        h, ta ? rocblas_operation_transpose : rocblas_operation_none,
        tb ? rocblas_operation_transpose : rocblas_operation_none, (rocblas_int)m, (rocblas_int)n,
        (rocblas_int)k, &alpha, A, (rocblas_int)lda, sa, B, (rocblas_int)ldb, sb, &beta, C,
        (rocblas_int)ldc, sc, (rocblas_int)batch));
}

inline void sgemm(rocblas_handle h, bool ta, bool tb, int64_t m, int64_t n, int64_t k, float alpha,
                  const float *A, int64_t lda, const float *B, int64_t ldb, float beta, float *C,
                  int64_t ldc) {
    if (m == 0 || n == 0) return;
    STDA_BLAS_CHECK(rocblas_sgemm(h, ta ? rocblas_operation_transpose : rocblas_operation_none,
                                  tb ? rocblas_operation_transpose : rocblas_operation_none,
                                  (rocblas_int)m, (rocblas_int)n, (rocblas_int)k, &alpha, A,
                                  (rocblas_int)lda, B, (rocblas_int)ldb, &beta, C,
                                  (rocblas_int)ldc));
}

inline void sgemm_batched(rocblas_handle h, bool ta, bool tb, int64_t m, int64_t n, int64_t k,
                          float alpha, const float *A, int64_t lda, int64_t sa, const float *B,
                          int64_t ldb, int64_t sb, float beta, float *C, int64_t ldc, int64_t sc,
                          int64_t batch) {
    if (m == 0 || n == 0 || batch == 0) return;
    STDA_BLAS_CHECK(rocblas_sgemm_strided_batched(
        h, ta ? rocblas_operation_transpose : rocblas_operation_none,
        tb ? rocblas_operation_transpose : rocblas_operation_none, (rocblas_int)m, (rocblas_int)n,
        (rocblas_int)k, &alpha, A, (rocblas_int)lda, sa, B, (rocblas_int)ldb, sb, &beta, C,
        (rocblas_int)ldc, sc, (rocblas_int)batch));
}

// Host symmetric eigensolver (LAPACK dsyevd): eigenvalues ascending, vectors
// in place of the column-major n x n matrix a.
void host_syev(int n, std::vector<double> &a, std::vector<double> &w);

}  // namespace stda
