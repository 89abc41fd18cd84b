// Basis of the effective dielectric matrix, rpa_G (rpa/rpa.f90:31-173): the
// columns G whose span carries Pi(u) for all frequencies, from RWR = Pi(u) at
// the lowest frequency, and the number NVecs of columns that keeps
// Tr(Pi(exact) - Pi(eff)) below the target.
//   RANDOM (1): Gaussian G, two subspace iterations G <- Q(RWR G), then
//               RWR <- G^T RWR G (Saibaba, Alexanderian, Ipsen 2017, Alg. 1)
//   EIGEN  (2): eigenvectors of RWR, largest eigenvalue first; RWR <- diag
//   FULL   (3): G = I, NVecs = NCholesky
// NVecs: scanning p = N..1, the diagonal of the transformed RWR is added while
// the sum stays below the target; NVecs is the last p added (rpa.f90:141-152).
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "internal.h"

int rocblas_handle_of(brpa_ctx *ctx, void **h);
int launch_gauss(brpa_ctx *, double *G, int64_t n, uint64_t seed);
int launch_set_diag_only(brpa_ctx *, double *M, const double *d, int64_t n);
int launch_scale(brpa_ctx *, double *x, double a, int64_t n);
int launch_fill(brpa_ctx *, double *x, int64_t n, double v);
int launch_add_diag(brpa_ctx *, double *M, int64_t n, double a);

namespace {

int rbc(brpa_ctx *ctx, rocblas_status st, const char *what) {
    if (st != rocblas_status_success)
        return brpa_fail(ctx, BRPA_ERR_SOLVER, "%s failed: rocblas status %d", what, (int)st);
    return BRPA_OK;
}

int dgemm(brpa_ctx *ctx, rocblas_handle h, bool ta, int64_t n, const double *A, const double *B,
          double *C) {
    const double one = 1.0, zero = 0.0;
    return rbc(ctx,
               rocblas_dgemm(h, ta ? rocblas_operation_transpose : rocblas_operation_none,
                             rocblas_operation_none, (rocblas_int)n, (rocblas_int)n,
                             (rocblas_int)n, &one, A, (rocblas_int)n, B, (rocblas_int)n, &zero, C,
                             (rocblas_int)n),
               "rocblas_dgemm");
}

}  // namespace

extern "C" {

int brpa_rpa_g(brpa_ctx *ctx, int basis, double *G, double *RWR, int64_t ncholesky,
               double target_error, uint64_t seed, int64_t *nvecs, double *abs_error) {
    if (!ctx) return BRPA_ERR_ARG;
    if (!G || !RWR || !nvecs || ncholesky <= 0 || ncholesky > 32768 || !std::isfinite(target_error))
        return brpa_fail(ctx, BRPA_ERR_ARG, "brpa_rpa_g: bad arguments");
    if (basis != BRPA_BASIS_RANDOM && basis != BRPA_BASIS_EIGEN && basis != BRPA_BASIS_FULL_CHOLESKY)
        return brpa_fail(ctx, BRPA_ERR_ARG,
                         "brpa_rpa_g: invalid basis type for the matrix Pi(u) (%d)", basis);
    const int64_t n = ncholesky;
    BRPA_HIP(ctx, hipSetDevice(ctx->device));
    void *hv = nullptr;
    BRPA_TRY(rocblas_handle_of(ctx, &hv));
    rocblas_handle h = static_cast<rocblas_handle>(hv);
    const bool gdev = is_device_ptr(G), rdev = is_device_ptr(RWR);
    double *dG = G, *dR = RWR, *T, *ev, *work;
    int *info;
    auto layout = [&](Arena &ar) {
        if (!gdev) dG = ar.take<double>(n * n);
        if (!rdev) dR = ar.take<double>(n * n);
        T = ar.take<double>(n * n);
        ev = ar.take<double>(n);
        work = ar.take<double>(n);
        info = ar.take<int>(1);
    };
    Arena count, ar;
    layout(count);
    BRPA_TRY(ws_reserve(ctx, count.off, &ar));
    dG = G;
    dR = RWR;
    layout(ar);
    if (!rdev) BRPA_TRY(copy_sync(ctx, dR, RWR, 8 * n * n, hipMemcpyHostToDevice));
    std::vector<double> diag(n);
    if (basis == BRPA_BASIS_FULL_CHOLESKY) {
        BRPA_TRY(launch_fill(ctx, dG, n * n, 0.0));
        BRPA_TRY(launch_add_diag(ctx, dG, n, 1.0));
    } else if (basis == BRPA_BASIS_EIGEN) {
        // eigenpairs of -RWR (ascending), so that the largest eigenvalue of RWR comes first
        std::lock_guard<std::mutex> lk(solver_mutex());
        BRPA_HIP(ctx, hipMemcpyAsync(dG, dR, 8 * n * n, hipMemcpyDeviceToDevice, ctx->stream));
        BRPA_TRY(launch_scale(ctx, dG, -1.0, n * n));
        BRPA_TRY(rbc(ctx, rocsolver_dsyevd(h, rocblas_evect_original, rocblas_fill_upper,
                                           (rocblas_int)n, dG, (rocblas_int)n, ev, work, info),
                     "rocsolver_dsyevd"));
        int hi = 0;
        BRPA_TRY(copy_sync(ctx, &hi, info, sizeof(int), hipMemcpyDeviceToHost));
        if (hi != 0) return brpa_fail(ctx, BRPA_ERR_SOLVER, "brpa_rpa_g: dsyevd did not converge");
        BRPA_TRY(launch_scale(ctx, ev, -1.0, n));
        BRPA_TRY(launch_set_diag_only(ctx, dR, ev, n));
    } else {
        // Gaussian G, two subspace iterations with QR, then RWR <- G^T RWR G
        BRPA_TRY(launch_gauss(ctx, dG, n * n, seed));
        for (int q = 0; q < 2; ++q) {
            std::lock_guard<std::mutex> lk(solver_mutex());
            BRPA_TRY(dgemm(ctx, h, false, n, dR, dG, T));
            BRPA_TRY(rbc(ctx, rocsolver_dgeqrf(h, (rocblas_int)n, (rocblas_int)n, T,
                                               (rocblas_int)n, ev),
                         "rocsolver_dgeqrf"));
            BRPA_TRY(rbc(ctx, rocsolver_dorgqr(h, (rocblas_int)n, (rocblas_int)n, (rocblas_int)n, T,
                                               (rocblas_int)n, ev),
                         "rocsolver_dorgqr"));
            BRPA_HIP(ctx, hipMemcpyAsync(dG, T, 8 * n * n, hipMemcpyDeviceToDevice, ctx->stream));
        }
        BRPA_TRY(dgemm(ctx, h, false, n, dR, dG, T));
        BRPA_TRY(dgemm(ctx, h, true, n, dG, T, dR));
    }
    // diagonal of the transformed RWR, then the trace rule (extended-precision sums,
    // as the reference's REAL(F128))
    BRPA_HIP(ctx, hipMemcpy2DAsync(diag.data(), sizeof(double), dR, 8 * (n + 1), sizeof(double), n,
                                   hipMemcpyDeviceToHost, ctx->stream));
    BRPA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    int64_t nv = n;
    long double abserr = 0.0L;
    if (basis != BRPA_BASIS_FULL_CHOLESKY) {
        for (int64_t p = n - 1; p >= 0; --p) {
            const long double next = abserr + (long double)diag[p];
            if (next < (long double)target_error) {
                abserr = next;
                nv = p + 1;
            } else {
                break;
            }
        }
    }
    if (!gdev) BRPA_TRY(copy_sync(ctx, G, dG, 8 * n * n, hipMemcpyDeviceToHost));
    if (!rdev) BRPA_TRY(copy_sync(ctx, RWR, dR, 8 * n * n, hipMemcpyDeviceToHost));
    BRPA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    *nvecs = nv;
    if (abs_error) *abs_error = (double)abserr;
    return BRPA_OK;
}

}  // extern "C"
