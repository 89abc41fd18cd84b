// THC route: Cholesky vectors R from the tensor-hypercontraction factors and
// Pi(u) of every frequency from them (rpa_THC_PiU / rpa_THC_PiU_BlockAlgo,
// rpa/rpa_THC.f90:215-339).
//
// The reference walks ai in blocks of MaxBlockDim to bound its memory: per
// block it forms Xgab = Xga o Xgi (ngrid x block), R_block = Zgk^T Xgab, and
// one dgemm per frequency.  Here R is produced whole, once, into HBM by
// k_thc_r (the Hadamard product formed in the operand loader), and all
// frequencies of Pi(u) come from one FP64 MFMA SYRK launch that stages each R
// tile once for four frequencies (k_syrk_f64).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "internal.h"

int launch_thc_r(brpa_ctx *, const double *Xa, const double *Xi, const double *Z, int64_t G,
                 int64_t nocc, int64_t nvir, int64_t N, double *R);
int launch_axpy(brpa_ctx *, double *y, const double *x, int64_t n);
int launch_fill(brpa_ctx *, double *x, int64_t n, double v);

namespace {

int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

int host_copy(brpa_ctx *ctx, std::vector<double> &v, const double *p, int64_t n) {
    v.resize(n);
    if (n == 0) return BRPA_OK;
    if (is_device_ptr(p)) return copy_sync(ctx, v.data(), p, 8 * n, hipMemcpyDeviceToHost);
    std::memcpy(v.data(), p, 8 * n);
    return BRPA_OK;
}

}  // namespace

extern "C" {

int brpa_thc_piu(brpa_ctx *ctx, double *PiU, double *Rkai, const double *Xga, const double *Xgi,
                 const double *Zgk, int64_t ngrid, int64_t nocc, int64_t nvir, int64_t naux,
                 const double *eocc, const double *evir, const double *freqs, int64_t nfreqs,
                 double prefactor, int accumulate) {
    if (!ctx) return BRPA_ERR_ARG;
    if (!PiU || !Xga || !Xgi || !Zgk || !eocc || !evir || !freqs || ngrid <= 0 || nocc < 0 ||
        nvir < 0 || naux <= 0 || nfreqs <= 0 || !std::isfinite(prefactor))
        return brpa_fail(ctx, BRPA_ERR_ARG, "brpa_thc_piu: bad arguments");
    const int64_t K = nocc * nvir, N = naux, F = nfreqs, G = ngrid;
    if (K > INT32_MAX || N > 65536 || F > 4096)
        return brpa_fail(ctx, BRPA_ERR_ARG, "brpa_thc_piu: sizes beyond the supported range");
    BRPA_HIP(ctx, hipSetDevice(ctx->device));
    std::vector<double> eo, ev, fx;
    BRPA_TRY(host_copy(ctx, eo, eocc, nocc));
    BRPA_TRY(host_copy(ctx, ev, evir, nvir));
    BRPA_TRY(host_copy(ctx, fx, freqs, F));
    for (double x : eo)
        if (!std::isfinite(x)) return brpa_fail(ctx, BRPA_ERR_NONFINITE, "brpa_thc_piu: non-finite input");
    for (double x : ev)
        if (!std::isfinite(x)) return brpa_fail(ctx, BRPA_ERR_NONFINITE, "brpa_thc_piu: non-finite input");
    for (double x : fx)
        if (!std::isfinite(x)) return brpa_fail(ctx, BRPA_ERR_NONFINITE, "brpa_thc_piu: non-finite input");
    const bool pdev = is_device_ptr(PiU);
    stage_reset(ctx);
    if (K == 0) {
        // no ai pair in this spin block: the reference skips it (rpa_THC.f90:269)
        if (accumulate) return BRPA_OK;
        if (pdev) {
            BRPA_TRY(launch_fill(ctx, PiU, F * N * N, 0.0));
            BRPA_HIP(ctx, hipStreamSynchronize(ctx->stream));
        } else {
            std::memset(PiU, 0, 8 * F * N * N);
        }
        return BRPA_OK;
    }
    // c_w(ai) = prefactor d/(d^2 + u_w^2) (rpa_THC_ChiRkai 342-370 with Alpha folded in)
    std::vector<double> c(F * K);
    for (int64_t w = 0; w < F; ++w)
        for (int64_t i = 0; i < nocc; ++i)
            for (int64_t a = 0; a < nvir; ++a) {
                const double d = ev[a] - eo[i];
                c[w * K + i * nvir + a] = prefactor * (d / (d * d + fx[w] * fx[w]));
            }
    const bool xadev = is_device_ptr(Xga), xidev = is_device_ptr(Xgi), zdev = is_device_ptr(Zgk);
    const bool rdev = Rkai && is_device_ptr(Rkai);
    // the SYRK writes Pi directly into a device PiU that it overwrites
    const bool direct = pdev && !accumulate;
    const int64_t ntile = cdiv(N, SYRK_T);
    const int64_t total = ntile * (ntile + 1) / 2 * cdiv(F, 4);
    FragDev fd{};
    FragDev *dfd = nullptr;
    int64_t *dpre = nullptr;
    const double *dXa = Xga, *dXi = Xgi, *dZ = Zgk;
    double *dR = nullptr, *dP = nullptr, *dOut = nullptr;
    auto layout = [&](Arena &ar) {
        dfd = ar.take<FragDev>(1);
        dpre = ar.take<int64_t>(2);
        fd.dtab = ar.take<double>(F * K);
        if (!xadev) dXa = ar.take<double>(G * nvir);
        if (!xidev) dXi = ar.take<double>(G * nocc);
        if (!zdev) dZ = ar.take<double>(G * N);
        dR = rdev ? Rkai : ar.take<double>(K * N);
        dP = direct ? PiU : ar.take<double>(F * N * N);
        dOut = (accumulate && !pdev) ? ar.take<double>(F * N * N) : nullptr;
    };
    Arena count;
    layout(count);
    Arena ar;
    BRPA_TRY(ws_reserve(ctx, count.off, &ar));
    dXa = Xga;
    dXi = Xgi;
    dZ = Zgk;
    layout(ar);
    if (!xadev) BRPA_TRY(copy_sync(ctx, const_cast<double *>(dXa), Xga, 8 * G * nvir, hipMemcpyHostToDevice));
    if (!xidev) BRPA_TRY(copy_sync(ctx, const_cast<double *>(dXi), Xgi, 8 * G * nocc, hipMemcpyHostToDevice));
    if (!zdev) BRPA_TRY(copy_sync(ctx, const_cast<double *>(dZ), Zgk, 8 * G * N, hipMemcpyHostToDevice));
    fd.R = dR;
    fd.K = (int)K;
    fd.N = (int)N;
    fd.F = (int)F;
    fd.Pi = dP;
    const int64_t pre[2] = {0, total};
    BRPA_TRY(copy_sync(ctx, const_cast<double *>(fd.dtab), c.data(), 8 * F * K, hipMemcpyHostToDevice));
    BRPA_TRY(copy_sync(ctx, dfd, &fd, sizeof(FragDev), hipMemcpyHostToDevice));
    BRPA_TRY(copy_sync(ctx, dpre, pre, sizeof(pre), hipMemcpyHostToDevice));

    BRPA_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    BRPA_TRY(stage_begin(ctx, BRPA_STAGE_THC_R));
    BRPA_TRY(launch_thc_r(ctx, dXa, dXi, dZ, G, nocc, nvir, N, dR));
    BRPA_TRY(stage_end(ctx, BRPA_STAGE_THC_R));
    BRPA_TRY(stage_begin(ctx, BRPA_STAGE_THC_PI));
    BRPA_TRY(launch_syrk_f64(ctx, dfd, dpre, 1, total));
    BRPA_TRY(stage_end(ctx, BRPA_STAGE_THC_PI));
    BRPA_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    if (accumulate) {
        double *dst = pdev ? PiU : dOut;
        if (!pdev) BRPA_TRY(copy_sync(ctx, dst, PiU, 8 * F * N * N, hipMemcpyHostToDevice));
        BRPA_TRY(launch_axpy(ctx, dst, dP, F * N * N));
        if (!pdev) BRPA_TRY(copy_sync(ctx, PiU, dst, 8 * F * N * N, hipMemcpyDeviceToHost));
    } else if (!pdev) {
        BRPA_TRY(copy_sync(ctx, PiU, dP, 8 * F * N * N, hipMemcpyDeviceToHost));
    }
    if (Rkai && !rdev) BRPA_TRY(copy_sync(ctx, Rkai, dR, 8 * K * N, hipMemcpyDeviceToHost));
    BRPA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    BRPA_TRY(stage_collect(ctx, BRPA_STAGE_THC_R));
    BRPA_TRY(stage_collect(ctx, BRPA_STAGE_THC_PI));
    float ms = 0.f;
    BRPA_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    ctx->last_kernel_ms = ms;
    return BRPA_OK;
}

}  // extern "C"
