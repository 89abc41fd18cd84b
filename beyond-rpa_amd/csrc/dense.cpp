// Dense FP64 building blocks of the explicit-space exchange: the library's own
// MFMA GEMM and triangular-solve kernels (kernels_dense.hip), its blocked
// Cholesky, and rocSOLVER dsyevd for the eigenvector fallback of the basis step.
#include "dense.h"

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>

int rocblas_handle_of(brpa_ctx *ctx, void **h);
int launch_gemm_f64(brpa_ctx *, bool ta, bool tb, int64_t m, int64_t n, int64_t k, double alpha,
                    const double *A, int64_t lda, const double *B, int64_t ldb, double beta,
                    double *C, int64_t ldc, bool upper);
int launch_trsm_f64(brpa_ctx *, bool brow, int64_t n, int64_t nrhs, const double *L,
                    int64_t ldl, int64_t strideL, double *Y, int64_t ldb, int64_t strideB,
                    int64_t nmat);
int launch_symm_fill(brpa_ctx *, double *M, int64_t n);
int launch_trinv_f64(brpa_ctx *, int64_t n, const double *U, int64_t ldu, int64_t su, double *W,
                     int64_t ldw, int64_t sw, int64_t nmat, double *T);

namespace {

int rb(brpa_ctx *ctx, rocblas_status st, const char *what) {
    if (st != rocblas_status_success)
        return brpa_fail(ctx, BRPA_ERR_SOLVER, "%s failed: rocblas status %d", what, (int)st);
    return BRPA_OK;
}

int hnd(brpa_ctx *ctx, rocblas_handle *h) {
    void *hv = nullptr;
    BRPA_TRY(rocblas_handle_of(ctx, &hv));
    *h = static_cast<rocblas_handle>(hv);
    return BRPA_OK;
}

int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

}  // namespace

int dense_gemm(brpa_ctx *ctx, bool ta, bool tb, int64_t m, int64_t n, int64_t k, double alpha,
               const double *A, int64_t lda, const double *B, int64_t ldb, double beta, double *C,
               int64_t ldc) {
    return launch_gemm_f64(ctx, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, false);
}

int dense_syrk(brpa_ctx *ctx, bool trans, int64_t n, int64_t k, double alpha, const double *A,
               int64_t lda, double beta, double *C, int64_t ldc) {
    // op(A) op(A)^T: A A^T (A n x k) or A^T A (A k x n)
    return launch_gemm_f64(ctx, trans, !trans, n, n, k, alpha, A, lda, A, lda, beta, C, ldc, true);
}

int dense_symm_fill(brpa_ctx *ctx, double *M, int64_t n) { return launch_symm_fill(ctx, M, n); }

int dense_trsm_right(brpa_ctx *ctx, int64_t m, int64_t n, const double *U, int64_t ldu, double *B,
                     int64_t ldb) {
    // B U^{-1} = (U^{-T} B^T)^T: forward substitution with L = U^T on the rows of B
    return launch_trsm_f64(ctx, true, n, m, U, ldu, 0, B, ldb, 0, 1);
}

int dense_trsm_left_t(brpa_ctx *ctx, int64_t m, int64_t n, const double *U, int64_t ldu,
                      int64_t strideU, double *B, int64_t ldb, int64_t strideB, int64_t nmat) {
    return launch_trsm_f64(ctx, false, m, n, U, ldu, strideU, B, ldb, strideB, nmat);
}

int dense_trinv(brpa_ctx *ctx, int64_t n, const double *U, int64_t ldu, int64_t su, double *W,
                int64_t ldw, int64_t sw, int64_t nmat, double *T) {
    return launch_trinv_f64(ctx, n, U, ldu, su, W, ldw, sw, nmat, T);
}

int dense_syevd(brpa_ctx *ctx, int64_t n, double *A, double *D, double *E, int *info) {
    rocblas_handle h;
    BRPA_TRY(hnd(ctx, &h));
    int hinfo = 0;
    // Each exchange worker calls this on its own context (own rocBLAS handle and
    // stream), concurrently with the other workers: the host fault that
    // rocsolver_dsyevd takes under counter collection (rocprofv3 --pmc) also
    // occurs with a single worker, so it is not a concurrency effect, and the
    // eigensolves of different fragments are not serialised.
    BRPA_TRY(rb(ctx,
                rocsolver_dsyevd(h, rocblas_evect_original, rocblas_fill_upper, (rocblas_int)n, A,
                                 (rocblas_int)n, D, E, info),
                "rocsolver_dsyevd"));
    BRPA_TRY(copy_sync(ctx, &hinfo, info, sizeof(int), hipMemcpyDeviceToHost));
    if (hinfo != 0)
        return brpa_fail(ctx, BRPA_ERR_SOLVER, "dsyevd did not converge (info=%d, n=%lld)", hinfo,
                         (long long)n);
    return BRPA_OK;
}

void DenseChol::layout(Arena &ar, int64_t nmax) {
    dfm = ar.take<FragDev>(1);
    fm.linv = ar.take<double>(CHOL_NB * CHOL_NB);
    fm.ldpart = ar.take<double>(cdiv(nmax, CHOL_NB) * 2);
    fm.status = ar.take<int>(1);
    dpre = ar.take<int64_t>(chol_pre_len(nmax));
}

int DenseChol::run(brpa_ctx *ctx, double *M, int64_t n, bool *ok) {
    fm.Pi = M;
    fm.N = (int)n;
    fm.F = 1;
    fm.nsteps = (int)cdiv(n, CHOL_NB);
    fm.keep_l11 = 1;
    BRPA_TRY(launch_fill_int(ctx, fm.status, 1, 0));
    BRPA_TRY(copy_sync(ctx, dfm, &fm, sizeof(FragDev), hipMemcpyHostToDevice));
    BRPA_TRY(chol_single(ctx, dfm, dpre, n, 1));
    int st = 0;
    BRPA_TRY(copy_sync(ctx, &st, fm.status, sizeof(int), hipMemcpyDeviceToHost));
    *ok = (st == 0);
    return BRPA_OK;
}

extern "C" {

int brpa_dgemm(brpa_ctx *ctx, int transa, int transb, int64_t m, int64_t n, int64_t k,
               double alpha, const double *A, int64_t lda, const double *B, int64_t ldb,
               double beta, double *C, int64_t ldc) {
    if (!ctx) return BRPA_ERR_ARG;
    if (m < 0 || n < 0 || k < 0 || !C || (m > 0 && n > 0 && k > 0 && (!A || !B)))
        return brpa_fail(ctx, BRPA_ERR_ARG, "brpa_dgemm: bad arguments");
    if (lda < std::max<int64_t>(1, transa ? k : m) || ldb < std::max<int64_t>(1, transb ? n : k) ||
        ldc < std::max<int64_t>(1, m))
        return brpa_fail(ctx, BRPA_ERR_ARG, "brpa_dgemm: leading dimension too small");
    if ((A && !is_device_ptr(A)) || (B && !is_device_ptr(B)) || !is_device_ptr(C))
        return brpa_fail(ctx, BRPA_ERR_ARG, "brpa_dgemm: operands must be device memory");
    BRPA_HIP(ctx, hipSetDevice(ctx->device));
    stage_reset(ctx);
    kstat_drop(ctx);
    BRPA_TRY(dense_gemm(ctx, transa != 0, transb != 0, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc));
    BRPA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return kstat_collect(ctx);
}

int brpa_dsyrk(brpa_ctx *ctx, int trans, int fill_lower, int64_t n, int64_t k, double alpha,
               const double *A, int64_t lda, double beta, double *C, int64_t ldc) {
    if (!ctx) return BRPA_ERR_ARG;
    if (n < 0 || k < 0 || !C || (n > 0 && k > 0 && !A))
        return brpa_fail(ctx, BRPA_ERR_ARG, "brpa_dsyrk: bad arguments");
    if (lda < std::max<int64_t>(1, trans ? k : n) || ldc < std::max<int64_t>(1, n) ||
        (fill_lower && ldc != n))
        return brpa_fail(ctx, BRPA_ERR_ARG, "brpa_dsyrk: leading dimension not supported");
    if ((A && !is_device_ptr(A)) || !is_device_ptr(C))
        return brpa_fail(ctx, BRPA_ERR_ARG, "brpa_dsyrk: operands must be device memory");
    BRPA_HIP(ctx, hipSetDevice(ctx->device));
    stage_reset(ctx);
    kstat_drop(ctx);
    BRPA_TRY(dense_syrk(ctx, trans != 0, n, k, alpha, A, lda, beta, C, ldc));
    if (fill_lower) BRPA_TRY(dense_symm_fill(ctx, C, n));
    BRPA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return kstat_collect(ctx);
}

int brpa_dtrsm(brpa_ctx *ctx, int side_right, int64_t m, int64_t n, const double *U, int64_t ldu,
               double *B, int64_t ldb) {
    if (!ctx) return BRPA_ERR_ARG;
    if (m < 0 || n < 0 || !U || !B) return brpa_fail(ctx, BRPA_ERR_ARG, "brpa_dtrsm: bad arguments");
    if (ldu < std::max<int64_t>(1, side_right ? n : m) || ldb < std::max<int64_t>(1, m))
        return brpa_fail(ctx, BRPA_ERR_ARG, "brpa_dtrsm: leading dimension too small");
    if (!is_device_ptr(U) || !is_device_ptr(B))
        return brpa_fail(ctx, BRPA_ERR_ARG, "brpa_dtrsm: operands must be device memory");
    BRPA_HIP(ctx, hipSetDevice(ctx->device));
    stage_reset(ctx);
    kstat_drop(ctx);
    if (side_right)
        BRPA_TRY(dense_trsm_right(ctx, m, n, U, ldu, B, ldb));
    else
        BRPA_TRY(dense_trsm_left_t(ctx, m, n, U, ldu, 0, B, ldb, 0, 1));
    BRPA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return kstat_collect(ctx);
}

int brpa_dtrtri(brpa_ctx *ctx, int64_t n, int64_t nmat, const double *U, int64_t ldu, int64_t su,
                double *W, int64_t ldw, int64_t sw) {
    if (!ctx) return BRPA_ERR_ARG;
    if (n < 0 || nmat < 0 || (n > 0 && nmat > 0 && (!U || !W)))
        return brpa_fail(ctx, BRPA_ERR_ARG, "brpa_dtrtri: bad arguments");
    if (n == 0 || nmat == 0) return BRPA_OK;
    if (ldu < n || ldw < n || (nmat > 1 && (su < ldu * n || sw < ldw * n)) || nmat > 65535)
        return brpa_fail(ctx, BRPA_ERR_ARG, "brpa_dtrtri: leading dimension or stride too small");
    if (!is_device_ptr(U) || !is_device_ptr(W))
        return brpa_fail(ctx, BRPA_ERR_ARG, "brpa_dtrtri: operands must be device memory");
    BRPA_HIP(ctx, hipSetDevice(ctx->device));
    Arena ar;
    BRPA_TRY(ws_reserve(ctx, 8 * (size_t)(nmat * trinv_scratch(n)) + 256, &ar));
    double *T = ar.take<double>(nmat * trinv_scratch(n));
    stage_reset(ctx);
    kstat_drop(ctx);
    BRPA_TRY(dense_trinv(ctx, n, U, ldu, su, W, ldw, sw, nmat, T));
    BRPA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return kstat_collect(ctx);
}

}  // extern "C"
