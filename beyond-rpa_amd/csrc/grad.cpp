// Adjoint of the direct-RPA energy (no reference counterpart: the reference
// computes no RPA gradients).
//
//   E      = sum_k w_k/(2 pi) [log det(1 + Pi_k) - tr Pi_k],  Pi_k = R^T diag(c_k) R
//   G_k    = w_k/(2 pi) [(1 + Pi_k)^{-1} - 1]
//   dE/dR  = sum_k 2 diag(c_k) R G_k
//   dE/dd  = sum_k (R G_k R^T)_{ai,ai} dc_k/dd,   dc/dd = 4 (u^2 - d^2)/(d^2 + u^2)^2
//   dE/de_a = sum_i dE/dd_ai,  dE/de_i = -sum_a dE/dd_ai
// The frequency grid is held fixed.  Pi_k is recomputed here (FP64 MFMA SYRK),
// not kept from the forward pass.  (1 + Pi_k)^{-1} = U_k^{-1} U_k^{-T} for all k
// at once: the library's batched blocked Cholesky (1 + Pi_k = U_k^T U_k, U_k the
// column-major upper triangle), rocBLAS's batched triangular inverse and one
// strided-batched dgemm; then R (1 + Pi_k)^{-1} by one dgemm per frequency.
#include <rocblas/rocblas.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "internal.h"

int launch_grad_acc(brpa_ctx *, const double *X, const double *Rc, int64_t N, int64_t K,
                    const double *c, const double *dc, double coef, double *dR, double *dd);
int launch_fill(brpa_ctx *, double *p, int64_t n, double v);
int rocblas_handle_of(brpa_ctx *ctx, void **h);
int launch_zero_lower(brpa_ctx *, double *M, int64_t n);

namespace {
constexpr double kPi = 3.14159265358979323846;
int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }
int rb(brpa_ctx *ctx, rocblas_status st, const char *what) {
    if (st != rocblas_status_success)
        return brpa_fail(ctx, BRPA_ERR_SOLVER, "%s failed: rocblas status %d", what, (int)st);
    return BRPA_OK;
}
}  // namespace

extern "C" int brpa_ecrpa_grad(brpa_ctx *ctx, const double *R, int64_t naux, const double *eocc,
                               int64_t nocc, const double *evir, int64_t nvir, const double *freqs,
                               const double *weights, int64_t nfreqs, double *dR, double *deocc,
                               double *devir) {
    if (!ctx) return BRPA_ERR_ARG;
    if (!R || !eocc || !evir || !freqs || !weights || !dR || naux <= 0 || nocc <= 0 || nvir <= 0 ||
        nfreqs <= 0)
        return brpa_fail(ctx, BRPA_ERR_ARG, "brpa_ecrpa_grad: bad arguments");
    const int64_t K = nocc * nvir, N = naux, F = nfreqs;
    if (K > INT32_MAX || N > 65536)
        return brpa_fail(ctx, BRPA_ERR_ARG, "brpa_ecrpa_grad: sizes beyond the supported range");
    std::vector<double> eo(nocc), ev(nvir), fx(F), fw(F);
    auto hv = [&](std::vector<double> &v, const double *p, int64_t n) -> int {
        if (is_device_ptr(p))
            BRPA_TRY(copy_sync(ctx, v.data(), p, 8 * n, hipMemcpyDeviceToHost));
        else
            std::memcpy(v.data(), p, 8 * n);
        for (double x : v)
            if (!std::isfinite(x)) return brpa_fail(ctx, BRPA_ERR_NONFINITE, "non-finite input");
        return BRPA_OK;
    };
    BRPA_TRY(hv(eo, eocc, nocc));
    BRPA_TRY(hv(ev, evir, nvir));
    BRPA_TRY(hv(fx, freqs, F));
    BRPA_TRY(hv(fw, weights, F));
    if (!(*std::min_element(ev.begin(), ev.end()) - *std::max_element(eo.begin(), eo.end()) > 0.0))
        return brpa_fail(ctx, BRPA_ERR_GAP, "Found invalid orbital energy gap.");
    BRPA_HIP(ctx, hipSetDevice(ctx->device));
    void *hv_ = nullptr;
    BRPA_TRY(rocblas_handle_of(ctx, &hv_));
    rocblas_handle h = static_cast<rocblas_handle>(hv_);

    // c_k and dc_k/dd per (k, ai)
    std::vector<double> cc(F * K), dcc(F * K);
    for (int64_t k = 0; k < F; ++k)
        for (int64_t i = 0; i < nocc; ++i)
            for (int64_t a = 0; a < nvir; ++a) {
                const double d = ev[a] - eo[i], u = fx[k], den = d * d + u * u;
                cc[k * K + i * nvir + a] = 4.0 * d / den;
                dcc[k * K + i * nvir + a] = 4.0 * (u * u - d * d) / (den * den);
            }
    const bool rdev = is_device_ptr(R), gdev = is_device_ptr(dR);
    FragDev fd{};
    std::vector<int64_t> pre = {0, 0, 0, 0, 0, 0};
    const int64_t nt = cdiv(N, SYRK_T), nsteps = cdiv(N, CHOL_NB);
    pre[1] = cdiv(F * K, 256);
    pre[3] = nt * (nt + 1) / 2 * cdiv(F, 4);
    pre[5] = cdiv(F * N, 256);
    FragDev *dfd;
    int64_t *dpre;
    double *dRd, *ddR, *dc, *ddc, *X, *dd;
    double *Uinv;
    int64_t *dchol;
    auto layout = [&](Arena &ar) {
        dfd = ar.take<FragDev>(1);
        dpre = ar.take<int64_t>(pre.size());
        dRd = rdev ? const_cast<double *>(R) : ar.take<double>(K * N);
        ddR = gdev ? dR : ar.take<double>(K * N);
        fd.dtab = ar.take<double>(F * K);
        fd.Pi = ar.take<double>(F * N * N);
        fd.diag = ar.take<double>(F * N);
        fd.freqs = ar.take<double>(F);
        fd.weights = ar.take<double>(F);
        fd.eocc = ar.take<double>(nocc);
        fd.evir = ar.take<double>(nvir);
        dc = ar.take<double>(F * K);
        ddc = ar.take<double>(F * K);
        X = ar.take<double>(K * N);
        dd = ar.take<double>(K);
        fd.linv = ar.take<double>(F * CHOL_NB * CHOL_NB);
        fd.ldpart = ar.take<double>(F * nsteps * 2);
        fd.status = ar.take<int>(F);
        dchol = ar.take<int64_t>(chol_pre_len(N));
        Uinv = ar.take<double>(F * N * N);
    };
    Arena count;
    layout(count);
    Arena ar;
    BRPA_TRY(ws_reserve(ctx, count.off, &ar));
    layout(ar);
    fd.R = dRd;
    fd.K = (int)K;
    fd.N = (int)N;
    fd.F = (int)F;
    fd.nocc = (int)nocc;
    fd.nvir = (int)nvir;
    fd.nsteps = (int)nsteps;
    fd.keep_l11 = 1;
    if (!rdev) BRPA_TRY(to_device(ctx, dRd, R, 8 * K * N));
    hipStream_t s = ctx->stream;
    // host inputs go up with completed copies: no pageable host buffer is read
    // after this point, also when a later step fails and the call returns early
    const auto h2d = hipMemcpyHostToDevice;
    BRPA_TRY(copy_sync(ctx, const_cast<double *>(fd.freqs), fx.data(), 8 * F, h2d));
    BRPA_TRY(copy_sync(ctx, const_cast<double *>(fd.weights), fw.data(), 8 * F, h2d));
    BRPA_TRY(copy_sync(ctx, const_cast<double *>(fd.eocc), eo.data(), 8 * nocc, h2d));
    BRPA_TRY(copy_sync(ctx, const_cast<double *>(fd.evir), ev.data(), 8 * nvir, h2d));
    BRPA_TRY(copy_sync(ctx, dc, cc.data(), 8 * F * K, h2d));
    BRPA_TRY(copy_sync(ctx, ddc, dcc.data(), 8 * F * K, h2d));
    BRPA_TRY(copy_sync(ctx, dfd, &fd, sizeof(FragDev), h2d));
    BRPA_TRY(copy_sync(ctx, dpre, pre.data(), 8 * pre.size(), h2d));
    BRPA_HIP(ctx, hipEventRecord(ctx->ev0, s));
    // Pi_k (recomputed), 1 + Pi_k, its Cholesky factor (all k batched)
    BRPA_TRY(launch_dtab(ctx, dfd, dpre, 1, pre[1]));
    BRPA_TRY(launch_syrk_f64(ctx, dfd, dpre + 2, 1, pre[3]));
    BRPA_TRY(launch_set_diag(ctx, dfd, dpre + 4, 1, pre[5], 0));
    BRPA_TRY(launch_fill(ctx, ddR, K * N, 0.0));
    BRPA_TRY(launch_fill(ctx, dd, K, 0.0));
    BRPA_TRY(launch_fill_int(ctx, fd.status, F, 0));
    BRPA_TRY(chol_single(ctx, dfd, dchol, N, F));
    BRPA_TRY(rb(ctx,
                rocblas_dtrtri_strided_batched(h, rocblas_fill_upper, rocblas_diagonal_non_unit,
                                               (rocblas_int)N, fd.Pi, (rocblas_int)N, N * N, Uinv,
                                               (rocblas_int)N, N * N, (rocblas_int)F),
                "rocblas_dtrtri_strided_batched"));
    for (int64_t k = 0; k < F; ++k) BRPA_TRY(launch_zero_lower(ctx, Uinv + k * N * N, N));
    // M_k = U_k^{-1} U_k^{-T} (full, symmetric) in place of Pi_k
    const double one = 1.0, zero = 0.0;
    BRPA_TRY(rb(ctx,
                rocblas_dgemm_strided_batched(h, rocblas_operation_none, rocblas_operation_transpose,
                                              (rocblas_int)N, (rocblas_int)N, (rocblas_int)N, &one,
                                              Uinv, (rocblas_int)N, N * N, Uinv, (rocblas_int)N,
                                              N * N, &zero, fd.Pi, (rocblas_int)N, N * N,
                                              (rocblas_int)F),
                "rocblas_dgemm_strided_batched"));
    for (int64_t k = 0; k < F; ++k) {
        // X(N x K) = M_k Rc
        BRPA_TRY(rb(ctx,
                    rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, (rocblas_int)N,
                                  (rocblas_int)K, (rocblas_int)N, &one, fd.Pi + k * N * N,
                                  (rocblas_int)N, dRd, (rocblas_int)N, &zero, X, (rocblas_int)N),
                    "rocblas_dgemm"));
        BRPA_TRY(launch_grad_acc(ctx, X, dRd, N, K, dc + k * K, ddc + k * K,
                                 fw[k] / (2.0 * kPi), ddR, dd));
    }
    BRPA_HIP(ctx, hipEventRecord(ctx->ev1, s));
    std::vector<double> hdd(K);
    std::vector<int> info(F);
    BRPA_TRY(copy_sync(ctx, hdd.data(), dd, 8 * K, hipMemcpyDeviceToHost));
    BRPA_TRY(copy_sync(ctx, info.data(), fd.status, sizeof(int) * F, hipMemcpyDeviceToHost));
    if (!gdev) BRPA_TRY(from_device(ctx, dR, ddR, 8 * K * N));
    BRPA_HIP(ctx, hipStreamSynchronize(s));
    float ms = 0.f;
    BRPA_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    ctx->last_kernel_ms = ms;
    for (int64_t k = 0; k < F; ++k)
        if (info[k] != 0)
            return brpa_fail(ctx, BRPA_ERR_NOT_SPD, "frequency %lld: 1 + Pi(u) is not positive definite",
                             (long long)k);
    // dE/de from dE/dd (small, host)
    std::vector<long double> go(nocc, 0.0L), gv(nvir, 0.0L);
    for (int64_t i = 0; i < nocc; ++i)
        for (int64_t a = 0; a < nvir; ++a) {
            go[i] -= hdd[i * nvir + a];
            gv[a] += hdd[i * nvir + a];
        }
    std::vector<double> go_d(go.begin(), go.end()), gv_d(gv.begin(), gv.end());
    if (deocc) {
        if (is_device_ptr(deocc))
            BRPA_TRY(copy_sync(ctx, deocc, go_d.data(), 8 * nocc, hipMemcpyHostToDevice));
        else
            std::memcpy(deocc, go_d.data(), 8 * nocc);
    }
    if (devir) {
        if (is_device_ptr(devir))
            BRPA_TRY(copy_sync(ctx, devir, gv_d.data(), 8 * nvir, hipMemcpyHostToDevice));
        else
            std::memcpy(devir, gv_d.data(), 8 * nvir);
    }
    return BRPA_OK;
}
