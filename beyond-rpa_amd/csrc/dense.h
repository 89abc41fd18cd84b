// Dense FP64 linear algebra on one context's stream, column-major (the
// reference's Fortran layout): element (r, c) of an m-row matrix with leading
// dimension ld is at r + ld c.  Symmetric and triangular matrices live in the
// upper triangle (the row-major lower triangle the library's Cholesky kernels
// factor in place).
#pragma once
#include "internal.h"

// C (m x n) = alpha op(A) op(B) + beta C (k_gemm_f64)
int dense_gemm(brpa_ctx *ctx, bool ta, bool tb, int64_t m, int64_t n, int64_t k, double alpha,
               const double *A, int64_t lda, const double *B, int64_t ldb, double beta, double *C,
               int64_t ldc);
// upper triangle of C (n x n) = alpha op(A) op(A)^T + beta C; op(A) is n x k
// (trans: A is k x n).  Only the tiles on or above the diagonal are computed;
// below the diagonal inside those tiles C is written too, elsewhere untouched.
int dense_syrk(brpa_ctx *ctx, bool trans, int64_t n, int64_t k, double alpha, const double *A,
               int64_t lda, double beta, double *C, int64_t ldc);
// lower triangle of the n x n matrix M (leading dimension n) from its upper one
int dense_symm_fill(brpa_ctx *ctx, double *M, int64_t n);
// B (m x n) <- B U^{-1}, U upper triangular n x n
int dense_trsm_right(brpa_ctx *ctx, int64_t m, int64_t n, const double *U, int64_t ldu, double *B,
                     int64_t ldb);
// B_w (m x n) <- U_w^{-T} B_w for nmat matrices U_w = U + w strideU (upper, m x m)
// and B_w = B + w strideB
int dense_trsm_left_t(brpa_ctx *ctx, int64_t m, int64_t n, const double *U, int64_t ldu,
                      int64_t strideU, double *B, int64_t ldb, int64_t strideB, int64_t nmat);
// Batched C_z = alpha op(A_z) op(B_z) diag(cs_z) + beta C_z, z = (z1, z2) with
// z1 < nb1, z2 < nb2 and X_z = X + z1 sX1 + z2 sX2 for every operand; upper:
// only the tiles on or above the diagonal of C (m = n); tri: TRI_* (the zero
// triangle of a triangular operand is skipped and must be stored as zeros);
// cs may be null.
int launch_gemm_f64_batched(brpa_ctx *ctx, bool ta, bool tb, int64_t m, int64_t n, int64_t k,
                            double alpha, const double *A, int64_t lda, int64_t sA1, int64_t sA2,
                            const double *B, int64_t ldb, int64_t sB1, int64_t sB2, double beta,
                            double *C, int64_t ldc, int64_t sC1, int64_t sC2, int64_t nb1,
                            int64_t nb2, bool upper, int tri, const double *cs, int64_t scs1,
                            int64_t scs2);
// W_z = U_z^{-1} for nmat upper triangular n x n matrices (U_z = U + z su, read
// in its upper triangle only; W_z = W + z sw, written whole: upper triangle and
// zeros below).  T: scratch of nmat * trinv_scratch(n) doubles.
int64_t trinv_scratch(int64_t n);
int dense_trinv(brpa_ctx *ctx, int64_t n, const double *U, int64_t ldu, int64_t su, double *W,
                int64_t ldw, int64_t sw, int64_t nmat, double *T);
// eigenpairs of the symmetric n x n matrix A (upper triangle), ascending D;
// A <- eigenvectors (rocSOLVER dsyevd: the fallback of the basis step).  E (n doubles) and info (1 int) are device scratch.
int dense_syevd(brpa_ctx *ctx, int64_t n, double *A, double *D, double *E, int *info);

// Cholesky factor of one dense n x n matrix in place with the library's
// blocked kernels: M = U^T U, U in the column-major upper triangle.
struct DenseChol {
    FragDev fm{};
    FragDev *dfm = nullptr;
    int64_t *dpre = nullptr;
    void layout(Arena &ar, int64_t nmax);
    // *ok = every pivot positive
    int run(brpa_ctx *ctx, double *M, int64_t n, bool *ok);
};
