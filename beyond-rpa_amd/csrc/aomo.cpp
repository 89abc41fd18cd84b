// AO -> MO transform of the Cholesky vectors for one spin block and one subset
// of packed AO pairs (rpa_CRG, rpa/rpa_core_MO.f90:233-301, with rpa_C 137-190
// and rpa_CRGxy 193-202).
//
// The reference forms C(ai, pq) = Co(p,i) Cv(q,a) + Co(q,i) Cv(p,a) over the
// packed pairs and multiplies once: ov Npq NVecs multiply-adds.  Here the packed
// columns are unpacked in HBM into the full symmetric AO-pair matrices of each
// Cholesky vector, RGfull(P, p, q), and the transform runs as two quarter
// transforms (rocBLAS strided-batched dgemm):
//   H(P, q, i) = sum_p RGfull(P, p, q) Co(p, i),   CRG(P, ai) (+)= sum_q H(P, q, i) Cv(q, a)
// i.e. 2 NVecs NAO o (NAO + v) flop, ~NAO v / (2 (NAO + v)) times fewer; the
// vectors go in chunks that keep RGfull below ~2 GB.
#include <rocblas/rocblas.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "internal.h"

int launch_unpack_pairs(brpa_ctx *, const double *RG, const int *pairs, int64_t npq, int64_t NV,
                        int64_t P0, int64_t Pc, int64_t nao, double *full);
int launch_fill(brpa_ctx *, double *x, int64_t n, double v);
int rocblas_handle_of(brpa_ctx *ctx, void **h);

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
    // a subset that holds every unordered AO pair overwrites all of RGfull, so the
    // zero fill is only needed otherwise
    bool covers = npq == nao * (nao + 1) / 2;
    if (covers) {
        std::vector<uint8_t> seen((size_t)nao * nao, 0);
        for (int64_t k = 0; k < npq && covers; ++k) {
            const int64_t p = std::max(pairs[2 * k], pairs[2 * k + 1]);
            const int64_t q = std::min(pairs[2 * k], pairs[2 * k + 1]);
            uint8_t &f = seen[(size_t)p * nao + q];
            covers = !f;
            f = 1;
        }
    }
    BRPA_HIP(ctx, hipSetDevice(ctx->device));
    const bool odev = is_device_ptr(occ_coeffs), vdev = is_device_ptr(virt_coeffs);
    const bool gdev = is_device_ptr(RG), cdev = is_device_ptr(CRG);
    const double *dCo = occ_coeffs, *dCv = virt_coeffs, *dRG = RG;
    double *dR = CRG;
    int *dpairs = nullptr;
    // Cholesky vectors per chunk: RGfull (nao^2 per vector) below ~2 GB
    const int64_t pc = std::max<int64_t>(1, std::min<int64_t>(nvecs, (256ll << 20) / (nao * nao)));
    double *full = nullptr, *H = nullptr;
    auto layout = [&](Arena &ar) {
        dpairs = ar.take<int>(std::max<int64_t>(2 * npq, 2));
        full = ar.take<double>(nao * nao * pc);
        H = ar.take<double>(nocc * nao * pc);
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
    void *hv = nullptr;
    BRPA_TRY(rocblas_handle_of(ctx, &hv));
    rocblas_handle h = static_cast<rocblas_handle>(hv);
    BRPA_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    const double one = 1.0, zero = 0.0, beta = accumulate ? 1.0 : 0.0;
    for (int64_t P0 = 0; P0 < nvecs; P0 += pc) {
        const int64_t m = std::min(pc, nvecs - P0);
        // RGfull[q][p][P] (P fastest), zero outside this subset's pairs
        if (!covers) BRPA_TRY(launch_fill(ctx, full, nao * nao * m, 0.0));
        BRPA_TRY(launch_unpack_pairs(ctx, dRG, dpairs, npq, nvecs, P0, m, nao, full));
        // H[i][q][P] = sum_p RGfull[q][p][P] Co(p, i): per q, (m x nao)(nao x o)
        if (rocblas_dgemm_strided_batched(
                h, rocblas_operation_none, rocblas_operation_none, (rocblas_int)m,
                (rocblas_int)nocc, (rocblas_int)nao, &one, full, (rocblas_int)m, nao * m, dCo,
                (rocblas_int)nao, 0, &zero, H, (rocblas_int)(m * nao), m, (rocblas_int)nao) !=
            rocblas_status_success)
            return brpa_fail(ctx, BRPA_ERR_SOLVER, "brpa_ao_mo: rocblas_dgemm_strided_batched failed");
        // CRG(P0 + P, i v + a) (+)= sum_q H[i][q][P] Cv(q, a): per i, (m x nao)(nao x v)
        if (rocblas_dgemm_strided_batched(
                h, rocblas_operation_none, rocblas_operation_none, (rocblas_int)m,
                (rocblas_int)nvir, (rocblas_int)nao, &one, H, (rocblas_int)m, m * nao, dCv,
                (rocblas_int)nao, 0, &beta, dR + P0, (rocblas_int)nvecs, nvir * nvecs,
                (rocblas_int)nocc) != rocblas_status_success)
            return brpa_fail(ctx, BRPA_ERR_SOLVER, "brpa_ao_mo: rocblas_dgemm_strided_batched failed");
    }
    BRPA_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    if (!cdev) BRPA_TRY(copy_sync(ctx, CRG, dR, 8 * K * nvecs, hipMemcpyDeviceToHost));
    BRPA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    float ms = 0.f;
    BRPA_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    ctx->last_kernel_ms = ms;
    return BRPA_OK;
}

}  // extern "C"
