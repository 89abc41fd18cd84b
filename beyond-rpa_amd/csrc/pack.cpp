// Packed storage of Pi(u): the lower triangle of each symmetric matrix in
// LAPACK packed order (rpa_Pack / rpa_Unpack, rpa/rpa.f90:379-404, on
// real_Pack / real_Unpack, linear-algebra/real_linalg.f90:737-773).
#include "internal.h"

int launch_pack_lower(brpa_ctx *, const double *A, int64_t n, int64_t F, double *AP,
                      int accumulate);
int launch_unpack_lower(brpa_ctx *, const double *AP, int64_t n, int64_t F, double *A);

extern "C" {

int brpa_pack_pi(brpa_ctx *ctx, double *packed, const double *piu, int64_t nfreqs, int64_t naux,
                 int accumulate) {
    if (!ctx) return BRPA_ERR_ARG;
    if (!packed || !piu || nfreqs <= 0 || naux <= 0 || naux > 65536 || nfreqs > 65535)
        return brpa_fail(ctx, BRPA_ERR_ARG, "brpa_pack_pi: bad arguments");
    BRPA_HIP(ctx, hipSetDevice(ctx->device));
    const int64_t N = naux, F = nfreqs, np = N * (N + 1) / 2;
    const bool adev = is_device_ptr(piu), pdev = is_device_ptr(packed);
    const double *dA = piu;
    double *dP = packed;
    auto layout = [&](Arena &ar) {
        if (!adev) dA = ar.take<double>(F * N * N);
        if (!pdev) dP = ar.take<double>(F * np);
    };
    Arena count, ar;
    layout(count);
    BRPA_TRY(ws_reserve(ctx, count.off, &ar));
    dA = piu;
    dP = packed;
    layout(ar);
    if (!adev) BRPA_TRY(copy_sync(ctx, const_cast<double *>(dA), piu, 8 * F * N * N, hipMemcpyHostToDevice));
    if (!pdev && accumulate) BRPA_TRY(copy_sync(ctx, dP, packed, 8 * F * np, hipMemcpyHostToDevice));
    BRPA_TRY(launch_pack_lower(ctx, dA, N, F, dP, accumulate));
    if (!pdev) BRPA_TRY(copy_sync(ctx, packed, dP, 8 * F * np, hipMemcpyDeviceToHost));
    BRPA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return BRPA_OK;
}

int brpa_unpack_pi(brpa_ctx *ctx, double *piu, const double *packed, int64_t nfreqs,
                   int64_t naux) {
    if (!ctx) return BRPA_ERR_ARG;
    if (!packed || !piu || nfreqs <= 0 || naux <= 0 || naux > 65536 || nfreqs > 65535)
        return brpa_fail(ctx, BRPA_ERR_ARG, "brpa_unpack_pi: bad arguments");
    BRPA_HIP(ctx, hipSetDevice(ctx->device));
    const int64_t N = naux, F = nfreqs, np = N * (N + 1) / 2;
    const bool adev = is_device_ptr(piu), pdev = is_device_ptr(packed);
    double *dA = piu;
    const double *dP = packed;
    auto layout = [&](Arena &ar) {
        if (!adev) dA = ar.take<double>(F * N * N);
        if (!pdev) dP = ar.take<double>(F * np);
    };
    Arena count, ar;
    layout(count);
    BRPA_TRY(ws_reserve(ctx, count.off, &ar));
    dA = piu;
    dP = packed;
    layout(ar);
    if (!pdev) BRPA_TRY(copy_sync(ctx, const_cast<double *>(dP), packed, 8 * F * np, hipMemcpyHostToDevice));
    BRPA_TRY(launch_unpack_lower(ctx, dP, N, F, dA));
    if (!adev) BRPA_TRY(copy_sync(ctx, piu, dA, 8 * F * N * N, hipMemcpyDeviceToHost));
    BRPA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return BRPA_OK;
}

}  // extern "C"
