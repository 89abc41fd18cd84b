// AO -> MO transform of the Cholesky vectors for one spin block and one subset
// of packed AO pairs (rpa_CRG, rpa/rpa_core_MO.f90:233-301, with rpa_C 137-190
// and rpa_CRGxy 193-202): one k_ao_mo launch.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "internal.h"

int launch_ao_mo(brpa_ctx *, const double *Co, const double *Cv, const int *pairs,
                 const double *RG, int64_t npq, int64_t nao, int64_t nocc, int64_t nvir,
                 int64_t nvecs, int accumulate, double *R);

extern "C" {

int brpa_ao_mo(brpa_ctx *ctx, double *CRG, const double *RG, int64_t nvecs, int64_t npq,
               const int32_t *pairs, const double *occ_coeffs, const double *virt_coeffs,
               int64_t nao, int64_t nocc, int64_t nvir, int accumulate) {
    if (!ctx) return BRPA_ERR_ARG;
    if (!CRG || !RG || !pairs || !occ_coeffs || !virt_coeffs || nvecs <= 0 || npq < 0 ||
        nao <= 0 || nocc < 0 || nvir < 0)
        return brpa_fail(ctx, BRPA_ERR_ARG, "brpa_ao_mo: bad arguments");
    const int64_t K = nocc * nvir;
    if (K > INT32_MAX || nvecs > 65536 || nao * std::max(nocc, nvir) + nao > INT32_MAX)
        return brpa_fail(ctx, BRPA_ERR_ARG, "brpa_ao_mo: sizes beyond the supported range");
    if (is_device_ptr(pairs))
        return brpa_fail(ctx, BRPA_ERR_ARG, "brpa_ao_mo: the pair table is a host array");
    for (int64_t k = 0; k < 2 * npq; ++k)
        if (pairs[k] < 0 || pairs[k] >= nao)
            return brpa_fail(ctx, BRPA_ERR_ARG, "brpa_ao_mo: pair %lld has an AO index outside "
                             "[0, %lld)", (long long)(k / 2), (long long)nao);
    if (K == 0) return BRPA_OK;
    BRPA_HIP(ctx, hipSetDevice(ctx->device));
    const bool odev = is_device_ptr(occ_coeffs), vdev = is_device_ptr(virt_coeffs);
    const bool gdev = is_device_ptr(RG), cdev = is_device_ptr(CRG);
    const double *dCo = occ_coeffs, *dCv = virt_coeffs, *dRG = RG;
    double *dR = CRG;
    int *dpairs = nullptr;
    auto layout = [&](Arena &ar) {
        dpairs = ar.take<int>(std::max<int64_t>(2 * npq, 2));
        if (!odev) dCo = ar.take<double>(nao * nocc);
        if (!vdev) dCv = ar.take<double>(nao * nvir);
        if (!gdev) dRG = ar.take<double>(std::max<int64_t>(nvecs * npq, 1));
        if (!cdev) dR = ar.take<double>(K * nvecs);
    };
    Arena count, ar;
    layout(count);
    BRPA_TRY(ws_reserve(ctx, count.off, &ar));
    dCo = occ_coeffs;
    dCv = virt_coeffs;
    dRG = RG;
    dR = CRG;
    layout(ar);
    if (npq > 0) BRPA_TRY(copy_sync(ctx, dpairs, pairs, 8 * npq, hipMemcpyHostToDevice));
    if (!odev)
        BRPA_TRY(copy_sync(ctx, const_cast<double *>(dCo), occ_coeffs, 8 * nao * nocc,
                           hipMemcpyHostToDevice));
    if (!vdev)
        BRPA_TRY(copy_sync(ctx, const_cast<double *>(dCv), virt_coeffs, 8 * nao * nvir,
                           hipMemcpyHostToDevice));
    if (!gdev && npq > 0)
        BRPA_TRY(copy_sync(ctx, const_cast<double *>(dRG), RG, 8 * nvecs * npq,
                           hipMemcpyHostToDevice));
    if (!cdev && accumulate)
        BRPA_TRY(copy_sync(ctx, dR, CRG, 8 * K * nvecs, hipMemcpyHostToDevice));
    BRPA_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    BRPA_TRY(launch_ao_mo(ctx, dCo, dCv, dpairs, dRG, npq, nao, nocc, nvir, nvecs, accumulate, dR));
    BRPA_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    if (!cdev) BRPA_TRY(copy_sync(ctx, CRG, dR, 8 * K * nvecs, hipMemcpyDeviceToHost));
    BRPA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    float ms = 0.f;
    BRPA_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    ctx->last_kernel_ms = ms;
    return BRPA_OK;
}

}  // extern "C"
