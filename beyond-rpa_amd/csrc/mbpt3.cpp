// Third-order exchange corrections from factorised amplitudes
// (rpa/rpa_CC_Exchange.f90:178-393):
//
//   MBPT3-1 (rpa_CC_Exchange_MBPT3_1 232-270): U = (R^T V)^T (R^T V),
//     D = V^T diag(d) V, Am = (A_m + A_n) D, Bm = D + 2U + 4U diag(A);
//     Bm^T X + X Bm = Am;  E = -sum_{aibj} (aj|bi) (V X V^T)(ai, bj).
//   MBPT3-2 (rpa_CC_Exchange_MBPT3_2 381-393, rpa_CC_TrW2 339-378):
//     E = 2 sum_{c,k,l} (sum_{a,i} R(l, ai) T(ak, ci))^2,  T = V diag(A) V^T.
//
// The reference solves the Lyapunov equation by Bartels-Stewart (real Schur
// form + dtrsyl, real_linalg.f90:125-171); rocSOLVER has no nonsymmetric
// Schur form, so here the equation is solved by the scaled matrix-sign
// (Newton) iteration, which keeps Bk^T X + X Bk = Ak at every step:
//   B <- (cB + (cB)^-1)/2,  A <- (cA + c^-1 B^-T A B^-1)/2,  c = sqrt(|B^-1|_F/|B|_F),
// until B = I, when X = A/2 (getrf/getri and dgemm on the device; once
// |B - 1|_F <= 1/4 the inverse-free Newton-Schulz step, all dgemm).  X is the
// derivative of the amplitudes with the coupling strength in the V basis and
// is negative definite in practice: -X = U'^T U' by the library's Cholesky
// gives V X V^T = -(V U'^T)(V U'^T)^T, so the trace is the SOSEX contraction
// of V'' = V U'^T with B'' = -1 (brpa_sosex, FP64 or split-int8); a Cholesky
// breakdown falls back to the eigenpairs of X (rocSOLVER dsyevd).
//
// MBPT3-2 runs as T~_k(c, i v + a) = T(ci, ak) = sum_mu (V A)(ci, mu) V(ak, mu)
// (strided-batched dgemm over i) and M_k = R^T T~_k^T (N x v), E = 2 sum_k |M_k|_F^2:
// 2 o^2 v^2 (NT2 + N) flop instead of the reference loop order's 4 o^2 v N NT2.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "internal.h"

int rocblas_handle_of(brpa_ctx *ctx, void **h);
int launch_col_scale(brpa_ctx *, double *out, const double *in, const double *s, int64_t m,
                     int64_t n);
int launch_mbpt3_ab(brpa_ctx *, const double *D, const double *U, const double *A, int64_t n,
                    double *Am, double *Bm);
int launch_scale(brpa_ctx *, double *x, double a, int64_t n);
int launch_add_diag(brpa_ctx *, double *M, int64_t n, double a);
int launch_zero_lower(brpa_ctx *, double *M, int64_t n);
int launch_fill(brpa_ctx *, double *x, int64_t n, double v);
int launch_axpy(brpa_ctx *, double *y, const double *x, int64_t n);
int launch_row_scale(brpa_ctx *, double *out, const double *in, const double *s, int64_t m, int64_t n);

namespace {

int rbc(brpa_ctx *ctx, rocblas_status st, const char *what) {
    if (st != rocblas_status_success)
        return brpa_fail(ctx, BRPA_ERR_SOLVER, "%s failed: rocblas status %d", what, (int)st);
    return BRPA_OK;
}

rocblas_operation op(bool t) { return t ? rocblas_operation_transpose : rocblas_operation_none; }

// column-major C(m x n) = alpha op(A) op(B) + beta C
int gemm(brpa_ctx *ctx, rocblas_handle h, bool ta, bool tb, int64_t m, int64_t n, int64_t k,
         double alpha, const double *A, int64_t lda, const double *B, int64_t ldb, double beta,
         double *C, int64_t ldc) {
    if (m <= 0 || n <= 0) return BRPA_OK;
    return rbc(ctx,
               rocblas_dgemm(h, op(ta), op(tb), (rocblas_int)m, (rocblas_int)n, (rocblas_int)k,
                             &alpha, A, (rocblas_int)lda, B, (rocblas_int)ldb, &beta, C,
                             (rocblas_int)ldc),
               "rocblas_dgemm");
}

int gemm_strided(brpa_ctx *ctx, rocblas_handle h, bool ta, bool tb, int64_t m, int64_t n,
                 int64_t k, double alpha, const double *A, int64_t lda, int64_t sa,
                 const double *B, int64_t ldb, int64_t sb, double beta, double *C, int64_t ldc,
                 int64_t sc, int64_t batch) {
    if (m <= 0 || n <= 0 || batch <= 0) return BRPA_OK;
    return rbc(ctx,
               rocblas_dgemm_strided_batched(h, op(ta), op(tb), (rocblas_int)m, (rocblas_int)n,
                                             (rocblas_int)k, &alpha, A, (rocblas_int)lda, sa, B,
                                             (rocblas_int)ldb, sb, &beta, C, (rocblas_int)ldc, sc,
                                             (rocblas_int)batch),
               "rocblas_dgemm_strided_batched");
}

// |x|_2 of n doubles (device), returned on the host
int nrm2(brpa_ctx *ctx, rocblas_handle h, const double *x, int64_t n, double *out) {
    BRPA_TRY(rbc(ctx, rocblas_set_pointer_mode(h, rocblas_pointer_mode_host), "pointer mode"));
    BRPA_TRY(rbc(ctx, rocblas_dnrm2(h, (rocblas_int)n, x, 1, out), "rocblas_dnrm2"));
    return BRPA_OK;
}

int grow(brpa_ctx *ctx, double **p, size_t *cap, size_t n) {
    if (n <= *cap) return BRPA_OK;
    if (*p) BRPA_HIP(ctx, hipFree(*p));
    *p = nullptr;
    *cap = 0;
    hipError_t e = hipMalloc(p, std::max<size_t>(n, 1) * sizeof(double));
    if (e != hipSuccess) {
        (void)hipGetLastError();
        *p = nullptr;
        return brpa_fail(ctx, BRPA_ERR_OOM, "buffer of %zu doubles: %s", n, hipGetErrorString(e));
    }
    *cap = n;
    return BRPA_OK;
}

int check_common(brpa_ctx *ctx, const char *who, const double *ec, const double *V,
                 const double *A, const double *R, int64_t nocc, int64_t nvir, int64_t naux,
                 int64_t nt) {
    if (!ec || !V || !A || !R || nocc <= 0 || nvir <= 0 || naux <= 0 || nt < 0)
        return brpa_fail(ctx, BRPA_ERR_ARG, "%s: bad arguments", who);
    if (nocc * nvir > INT32_MAX || naux > 65536 || nt > 65536)
        return brpa_fail(ctx, BRPA_ERR_ARG, "%s: sizes beyond the supported range", who);
    return BRPA_OK;
}

}  // namespace

void mbpt3_release(brpa_ctx *ctx) {
    if (ctx->x_V) (void)hipFree(ctx->x_V);
    if (ctx->x_B) (void)hipFree(ctx->x_B);
    ctx->x_V = ctx->x_B = nullptr;
    ctx->x_V_cap = ctx->x_B_cap = 0;
}

extern "C" {

int brpa_mbpt3_1(brpa_ctx *ctx, int precision, double *ec_exchange, const double *V,
                 const double *A, int64_t nvecs_t2, const double *R, int64_t naux,
                 const double *eocc, const double *evir, int64_t nocc, int64_t nvir) {
    if (!ctx) return BRPA_ERR_ARG;
    BRPA_TRY(check_common(ctx, "brpa_mbpt3_1", ec_exchange, V, A, R, nocc, nvir, naux, nvecs_t2));
    if (!eocc || !evir) return brpa_fail(ctx, BRPA_ERR_ARG, "brpa_mbpt3_1: bad arguments");
    if (precision != BRPA_FP64 && precision != BRPA_SPLIT_I8)
        return brpa_fail(ctx, BRPA_ERR_ARG, "brpa_mbpt3_1: precision must be BRPA_FP64 or "
                                            "BRPA_SPLIT_I8");
    const int64_t K = nocc * nvir, N = naux, n = nvecs_t2;
    if (n > 32768)
        return brpa_fail(ctx, BRPA_ERR_ARG, "brpa_mbpt3_1: NVecsT2 = %lld exceeds 32768",
                         (long long)n);
    *ec_exchange = 0.0;
    if (n == 0) return BRPA_OK;
    BRPA_HIP(ctx, hipSetDevice(ctx->device));
    void *hv = nullptr;
    BRPA_TRY(rocblas_handle_of(ctx, &hv));
    rocblas_handle h = static_cast<rocblas_handle>(hv);
    std::vector<double> eo(nocc), ev(nvir), d(K);
    auto hostv = [&](std::vector<double> &out, const double *p, int64_t m) -> int {
        if (is_device_ptr(p)) return copy_sync(ctx, out.data(), p, 8 * m, hipMemcpyDeviceToHost);
        std::memcpy(out.data(), p, 8 * m);
        return BRPA_OK;
    };
    BRPA_TRY(hostv(eo, eocc, nocc));
    BRPA_TRY(hostv(ev, evir, nvir));
    for (int64_t i = 0; i < nocc; ++i)
        for (int64_t a = 0; a < nvir; ++a) d[i * nvir + a] = ev[a] - eo[i];
    for (double x : d)
        if (!std::isfinite(x)) return brpa_fail(ctx, BRPA_ERR_NONFINITE, "brpa_mbpt3_1: non-finite input");
    const bool vdev = is_device_ptr(V), adev = is_device_ptr(A), rdev = is_device_ptr(R);
    const int64_t nsteps = (n + CHOL_NB - 1) / CHOL_NB;
    const double *dV = V, *dA = A, *dR = R;
    double *dd, *W, *U, *Dm, *Vd, *Am, *Bm, *Bi, *T1, *ev_x, *work;
    int *ipiv, *info;
    int64_t *dchol;
    FragDev fm{};
    FragDev *dfm = nullptr;
    auto layout = [&](Arena &ar) {
        if (!vdev) dV = ar.take<double>(K * n);
        if (!adev) dA = ar.take<double>(n);
        if (!rdev) dR = ar.take<double>(K * N);
        dd = ar.take<double>(K);
        W = ar.take<double>(N * n);
        Vd = ar.take<double>(K * n);
        U = ar.take<double>(n * n);
        Dm = ar.take<double>(n * n);
        Am = ar.take<double>(n * n);
        Bm = ar.take<double>(n * n);
        Bi = ar.take<double>(n * n);
        T1 = ar.take<double>(n * n);
        ev_x = ar.take<double>(n);
        work = ar.take<double>(n);
        ipiv = ar.take<int>(n);
        info = ar.take<int>(2);
        dfm = ar.take<FragDev>(1);
        fm.linv = ar.take<double>(CHOL_NB * CHOL_NB);
        fm.ldpart = ar.take<double>(nsteps * 2);
        fm.status = ar.take<int>(1);
        dchol = ar.take<int64_t>(chol_pre_len(n));
    };
    Arena count, ar;
    layout(count);
    BRPA_TRY(ws_reserve(ctx, count.off, &ar));
    dV = V;
    dA = A;
    dR = R;
    layout(ar);
    if (!vdev) BRPA_TRY(copy_sync(ctx, const_cast<double *>(dV), V, 8 * K * n, hipMemcpyHostToDevice));
    if (!adev) BRPA_TRY(copy_sync(ctx, const_cast<double *>(dA), A, 8 * n, hipMemcpyHostToDevice));
    if (!rdev) BRPA_TRY(copy_sync(ctx, const_cast<double *>(dR), R, 8 * K * N, hipMemcpyHostToDevice));
    BRPA_TRY(copy_sync(ctx, dd, d.data(), 8 * K, hipMemcpyHostToDevice));
    stage_reset(ctx);
    BRPA_TRY(stage_begin(ctx, BRPA_STAGE_MBPT3));
    // V is C V[ai][mu] = column-major n x K (V^T); R is column-major N x K (R^T)
    // W = R^T V (N x n), U = W^T W (rpa_CC_Ukl 178-207)
    BRPA_TRY(gemm(ctx, h, false, true, N, n, K, 1.0, dR, N, dV, n, 0.0, W, N));
    BRPA_TRY(gemm(ctx, h, true, false, n, n, N, 1.0, W, N, W, N, 0.0, U, n));
    // D = V^T diag(d) V (rpa_CC_Dkl 210-229)
    BRPA_TRY(launch_col_scale(ctx, Vd, dV, dd, n, K));
    BRPA_TRY(gemm(ctx, h, false, true, n, n, K, 1.0, dV, n, Vd, n, 0.0, Dm, n));
    BRPA_TRY(launch_mbpt3_ab(ctx, Dm, U, dA, n, Am, Bm));
    // Bm^T X + X Bm = Am by the scaled sign iteration; Am -> 2X, Bm -> I
    const double tol = 1e-12 * std::sqrt((double)n);
    bool conv = false;
    int iters = 0;
    double dev = HUGE_VAL;   // |B - 1|_F after the last step
    for (int it = 0; it < 100 && !conv; ++it) {
        if (dev <= 0.25) {
            // Newton-Schulz step on the same block matrix Z = [[B, 0], [A, -B^T]]:
            // Z <- Z (3 - Z^2)/2, i.e. B <- B (3 - B^2)/2 and
            // A <- (3A + B^T A B - A B^2 - (B^2)^T A)/2.  The map is odd, so
            // Bk^T X + X Bk = Ak still holds; it needs no inverse and converges
            // quadratically once |1 - B^2|_2 < 1, which |B - 1|_F <= 1/4 ensures.
            BRPA_TRY(gemm(ctx, h, false, false, n, n, n, 1.0, Bm, n, Bm, n, 0.0, Bi, n));
            BRPA_TRY(gemm(ctx, h, false, false, n, n, n, 1.0, Am, n, Bm, n, 0.0, T1, n));
            BRPA_HIP(ctx, hipMemcpyAsync(Dm, Am, 8 * n * n, hipMemcpyDeviceToDevice, ctx->stream));
            BRPA_TRY(gemm(ctx, h, true, false, n, n, n, 0.5, Bm, n, T1, n, 1.5, Dm, n));
            BRPA_TRY(gemm(ctx, h, false, false, n, n, n, -0.5, Am, n, Bi, n, 1.0, Dm, n));
            BRPA_TRY(gemm(ctx, h, true, false, n, n, n, -0.5, Bi, n, Am, n, 1.0, Dm, n));
            std::swap(Am, Dm);
            BRPA_TRY(launch_scale(ctx, Bi, -1.0, n * n));
            BRPA_TRY(launch_add_diag(ctx, Bi, n, 3.0));
            BRPA_TRY(gemm(ctx, h, false, false, n, n, n, 0.5, Bm, n, Bi, n, 0.0, T1, n));
            std::swap(Bm, T1);
            BRPA_HIP(ctx, hipMemcpyAsync(T1, Bm, 8 * n * n, hipMemcpyDeviceToDevice, ctx->stream));
            BRPA_TRY(launch_add_diag(ctx, T1, n, -1.0));
            BRPA_TRY(nrm2(ctx, h, T1, n * n, &dev));
            if (!std::isfinite(dev))
                return brpa_fail(ctx, BRPA_ERR_SOLVER, "brpa_mbpt3_1: Lyapunov iteration diverged");
            conv = dev <= tol;
            iters = it + 1;
            continue;
        }
        BRPA_HIP(ctx, hipMemcpyAsync(Bi, Bm, 8 * n * n, hipMemcpyDeviceToDevice, ctx->stream));
        int hinfo[2] = {0, 0};
        {
            std::lock_guard<std::mutex> lk(solver_mutex());
            BRPA_TRY(rbc(ctx, rocsolver_dgetrf(h, (rocblas_int)n, (rocblas_int)n, Bi,
                                               (rocblas_int)n, ipiv, info),
                         "rocsolver_dgetrf"));
            BRPA_TRY(rbc(ctx, rocsolver_dgetri(h, (rocblas_int)n, Bi, (rocblas_int)n, ipiv,
                                               info + 1),
                         "rocsolver_dgetri"));
            BRPA_TRY(copy_sync(ctx, hinfo, info, sizeof(hinfo), hipMemcpyDeviceToHost));
        }
        if (hinfo[0] != 0 || hinfo[1] != 0)
            return brpa_fail(ctx, BRPA_ERR_SOLVER,
                             "brpa_mbpt3_1: singular matrix in the Lyapunov iteration (step %d)", it);
        double nb = 0.0, ni = 0.0;
        BRPA_TRY(nrm2(ctx, h, Bm, n * n, &nb));
        BRPA_TRY(nrm2(ctx, h, Bi, n * n, &ni));
        if (!(nb > 0.0) || !std::isfinite(ni))
            return brpa_fail(ctx, BRPA_ERR_SOLVER, "brpa_mbpt3_1: Lyapunov iteration diverged");
        const double c = std::sqrt(ni / nb);
        // A <- (c A + c^-1 B^-T A B^-1)/2
        BRPA_TRY(gemm(ctx, h, false, false, n, n, n, 1.0, Am, n, Bi, n, 0.0, T1, n));
        BRPA_TRY(gemm(ctx, h, true, false, n, n, n, 0.5 / c, Bi, n, T1, n, 0.5 * c, Am, n));
        // B <- (c B + c^-1 B^-1)/2
        BRPA_TRY(launch_scale(ctx, Bm, 0.5 * c, n * n));
        BRPA_TRY(launch_scale(ctx, Bi, 0.5 / c, n * n));
        BRPA_TRY(launch_axpy(ctx, Bm, Bi, n * n));
        BRPA_HIP(ctx, hipMemcpyAsync(T1, Bm, 8 * n * n, hipMemcpyDeviceToDevice, ctx->stream));
        BRPA_TRY(launch_add_diag(ctx, T1, n, -1.0));
        BRPA_TRY(nrm2(ctx, h, T1, n * n, &dev));
        conv = dev <= tol;
        iters = it + 1;
    }
    if (!conv)
        return brpa_fail(ctx, BRPA_ERR_SOLVER, "brpa_mbpt3_1: Lyapunov iteration did not converge");
    // X = Am / 2.  Factor -X = U'^T U' (Cholesky on T1 = -X)
    BRPA_HIP(ctx, hipMemcpyAsync(T1, Am, 8 * n * n, hipMemcpyDeviceToDevice, ctx->stream));
    BRPA_TRY(launch_scale(ctx, T1, -0.5, n * n));
    fm.Pi = T1;
    fm.N = (int)n;
    fm.F = 1;
    fm.nsteps = (int)nsteps;
    fm.keep_l11 = 1;
    BRPA_TRY(launch_fill_int(ctx, fm.status, 1, 0));
    BRPA_TRY(copy_sync(ctx, dfm, &fm, sizeof(FragDev), hipMemcpyHostToDevice));
    BRPA_TRY(chol_single(ctx, dfm, dchol, n, 1));
    int st = 0;
    BRPA_TRY(copy_sync(ctx, &st, fm.status, sizeof(int), hipMemcpyDeviceToHost));
    BRPA_TRY(grow(ctx, &ctx->x_V, &ctx->x_V_cap, (size_t)K * n));
    BRPA_TRY(grow(ctx, &ctx->x_B, &ctx->x_B_cap, (size_t)n));
    if (st == 0) {
        // V X V^T = -(V U'^T)(V U'^T)^T; store V''^T = U' V^T as C V''[ai][mu]
        BRPA_TRY(launch_zero_lower(ctx, T1, n));
        BRPA_TRY(gemm(ctx, h, false, false, n, K, n, 1.0, T1, n, dV, n, 0.0, ctx->x_V, n));
        BRPA_TRY(launch_fill(ctx, ctx->x_B, n, -1.0));
    } else {
        // eigenpairs X = E diag(lambda) E^T; V''^T = E^T V^T
        std::lock_guard<std::mutex> lk(solver_mutex());
        BRPA_HIP(ctx, hipMemcpyAsync(T1, Am, 8 * n * n, hipMemcpyDeviceToDevice, ctx->stream));
        BRPA_TRY(launch_scale(ctx, T1, 0.5, n * n));
        BRPA_TRY(rbc(ctx, rocsolver_dsyevd(h, rocblas_evect_original, rocblas_fill_upper,
                                           (rocblas_int)n, T1, (rocblas_int)n, ev_x, work, info),
                     "rocsolver_dsyevd"));
        int hi = 0;
        BRPA_TRY(copy_sync(ctx, &hi, info, sizeof(int), hipMemcpyDeviceToHost));
        if (hi != 0)
            return brpa_fail(ctx, BRPA_ERR_SOLVER, "brpa_mbpt3_1: dsyevd did not converge");
        BRPA_TRY(gemm(ctx, h, true, false, n, K, n, 1.0, T1, n, dV, n, 0.0, ctx->x_V, n));
        BRPA_HIP(ctx, hipMemcpyAsync(ctx->x_B, ev_x, 8 * n, hipMemcpyDeviceToDevice, ctx->stream));
    }
    BRPA_TRY(stage_end(ctx, BRPA_STAGE_MBPT3));
    BRPA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    BRPA_TRY(stage_collect(ctx, BRPA_STAGE_MBPT3));
    ctx->stage_launches[BRPA_STAGE_MBPT3] = iters;   // reported as the sign-iteration steps
    ctx->mbpt3_chol = (st == 0);
    // E = -sum (aj|bi) (V'' diag(B'') V''^T)(ai, bj): the SOSEX contraction with alpha = 1
    const bool keep = ctx->stage_keep;
    ctx->stage_keep = true;
    // R as the caller gave it: the workspace copy may not survive brpa_sosex's own workspace
    int rc = brpa_sosex(ctx, precision, ec_exchange, 1.0, ctx->x_V, ctx->x_B, R, nocc, nvir, naux,
                        n, 0);
    ctx->stage_keep = keep;
    return rc;
}

int brpa_mbpt3_2(brpa_ctx *ctx, double *ec_exchange, const double *R, const double *V,
                 const double *A, int64_t nocc, int64_t nvir, int64_t naux, int64_t nvecs_t2) {
    if (!ctx) return BRPA_ERR_ARG;
    BRPA_TRY(check_common(ctx, "brpa_mbpt3_2", ec_exchange, V, A, R, nocc, nvir, naux, nvecs_t2));
    const int64_t o = nocc, v = nvir, K = o * v, N = naux, n = nvecs_t2;
    *ec_exchange = 0.0;
    if (n == 0) return BRPA_OK;
    BRPA_HIP(ctx, hipSetDevice(ctx->device));
    void *hv = nullptr;
    BRPA_TRY(rocblas_handle_of(ctx, &hv));
    rocblas_handle h = static_cast<rocblas_handle>(hv);
    const bool vdev = is_device_ptr(V), adev = is_device_ptr(A), rdev = is_device_ptr(R);
    // occupied indices k per chunk: T~_k (v x K) and M_k (N x v) for the chunk
    // stay below ~2 GB
    const int64_t per_k = v * K + N * v;
    const int64_t kc = std::max<int64_t>(1, std::min<int64_t>(o, (int64_t)(256ll << 20) / per_k));
    const double *dV = V, *dA = A, *dR = R;
    double *Y, *Tt, *M;
    auto layout = [&](Arena &ar) {
        if (!vdev) dV = ar.take<double>(K * n);
        if (!adev) dA = ar.take<double>(n);
        if (!rdev) dR = ar.take<double>(K * N);
        Y = ar.take<double>(K * n);
        Tt = ar.take<double>(kc * v * K);
        M = ar.take<double>(kc * N * v);
    };
    Arena count, ar;
    layout(count);
    BRPA_TRY(ws_reserve(ctx, count.off, &ar));
    dV = V;
    dA = A;
    dR = R;
    layout(ar);
    if (!vdev) BRPA_TRY(copy_sync(ctx, const_cast<double *>(dV), V, 8 * K * n, hipMemcpyHostToDevice));
    if (!adev) BRPA_TRY(copy_sync(ctx, const_cast<double *>(dA), A, 8 * n, hipMemcpyHostToDevice));
    if (!rdev) BRPA_TRY(copy_sync(ctx, const_cast<double *>(dR), R, 8 * K * N, hipMemcpyHostToDevice));
    {
        std::vector<double> ha(n);
        if (adev)
            BRPA_TRY(copy_sync(ctx, ha.data(), dA, 8 * n, hipMemcpyDeviceToHost));
        else
            std::memcpy(ha.data(), A, 8 * n);
        for (double x : ha)
            if (!std::isfinite(x))
                return brpa_fail(ctx, BRPA_ERR_NONFINITE, "brpa_mbpt3_2: non-finite A");
    }
    stage_reset(ctx);
    BRPA_TRY(stage_begin(ctx, BRPA_STAGE_MBPT3));
    // Y = V diag(A), C Y[ai][mu]: in the column-major n x K view row mu is scaled by A_mu
    BRPA_TRY(launch_row_scale(ctx, Y, dV, dA, n, K));
    BRPA_TRY(rbc(ctx, rocblas_set_pointer_mode(h, rocblas_pointer_mode_host), "pointer mode"));
    double tr = 0.0;
    for (int64_t k0 = 0; k0 < o; k0 += kc) {
        const int64_t nk = std::min(kc, o - k0);
        for (int64_t kk = 0; kk < nk; ++kk) {
            // T~_k(c, i v + a) = sum_mu Y(ci, mu) V(ak, mu), column-major v x K: for
            // each i the v x v block Y_i^T V_k (Y_i, V_k: column-major n x v views)
            BRPA_TRY(gemm_strided(ctx, h, true, false, v, v, n, 1.0, Y, n, v * n,
                                  dV + (k0 + kk) * v * n, n, 0, 0.0, Tt + kk * v * K, v, v * v,
                                  o));
        }
        // M_k = R^T T~_k^T (N x v), R: column-major N x K
        BRPA_TRY(gemm_strided(ctx, h, false, true, N, v, K, 1.0, dR, N, 0, Tt, v, v * K, 0.0, M,
                              N, N * v, nk));
        double part = 0.0;
        BRPA_TRY(rbc(ctx, rocblas_ddot(h, (rocblas_int)(nk * N * v), M, 1, M, 1, &part),
                     "rocblas_ddot"));
        tr += part;
    }
    BRPA_TRY(stage_end(ctx, BRPA_STAGE_MBPT3));
    BRPA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    BRPA_TRY(stage_collect(ctx, BRPA_STAGE_MBPT3));
    *ec_exchange = 2.0 * tr;
    return BRPA_OK;
}

}  // extern "C"
