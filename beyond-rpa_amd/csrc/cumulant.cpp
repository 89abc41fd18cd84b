// THC SOSEX + 2g cumulant terms (rpa_CCD_corrections_1b2g,
// rpa/rpa_CCD_Corrections.f90:79-290): YU = Yga U_mu by rocBLAS dgemm for a
// chunk of amplitude vectors, then the fused k_thc_1b2g contraction.
#include <rocblas/rocblas.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "internal.h"

int rocblas_handle_of(brpa_ctx *ctx, void **h);
int thc_1b2g_npairs(int64_t G);
int launch_thc_1b2g(brpa_ctx *, const double *YU, const double *X, const double *Z, int64_t G,
                    int64_t o, int64_t nmu, double *part, double *diag, int64_t mu0, double *S);

extern "C" {

int brpa_thc_1b2g(brpa_ctx *ctx, const double *Zgh, const double *Xgi, const double *Yga,
                  const double *Uaim, const double *Am, int64_t ngrid, int64_t nocc, int64_t nvir,
                  int64_t nvecs_t2, double *ec1b, double *ec2g, double *eig_sosex, double *eig_2g,
                  double *yxu_diag) {
    if (!ctx) return BRPA_ERR_ARG;
    if (!Zgh || !Xgi || !Yga || !Uaim || !Am || !ec1b || !ec2g || ngrid <= 0 || nocc <= 0 ||
        nvir <= 0 || nvecs_t2 < 0)
        return brpa_fail(ctx, BRPA_ERR_ARG, "brpa_thc_1b2g: bad arguments");
    const int64_t G = ngrid, o = nocc, v = nvir, nt = nvecs_t2;
    if (G > 65536 || o * v > INT32_MAX || G * std::max(o, v) > INT32_MAX)
        return brpa_fail(ctx, BRPA_ERR_ARG, "brpa_thc_1b2g: sizes beyond the supported range");
    *ec1b = *ec2g = 0.0;
    if (nt == 0) return BRPA_OK;
    BRPA_HIP(ctx, hipSetDevice(ctx->device));
    void *hv = nullptr;
    BRPA_TRY(rocblas_handle_of(ctx, &hv));
    rocblas_handle h = static_cast<rocblas_handle>(hv);
    std::vector<double> a(nt);
    if (is_device_ptr(Am))
        BRPA_TRY(copy_sync(ctx, a.data(), Am, 8 * nt, hipMemcpyDeviceToHost));
    else
        std::memcpy(a.data(), Am, 8 * nt);
    const int npairs = thc_1b2g_npairs(G);
    // amplitude vectors per chunk: YU (G x o per vector) below ~1 GB, at most 65535
    const int64_t nmu = std::max<int64_t>(
        1, std::min<int64_t>({nt, (int64_t)65535, (int64_t)(128ll << 20) / (G * o)}));
    const bool zdev = is_device_ptr(Zgh), xdev = is_device_ptr(Xgi), ydev = is_device_ptr(Yga),
               udev = is_device_ptr(Uaim), ddev = yxu_diag && is_device_ptr(yxu_diag);
    const double *dZ = Zgh, *dX = Xgi, *dY = Yga, *dU = Uaim;
    double *YU, *part, *S, *dD = yxu_diag;
    auto layout = [&](Arena &ar) {
        if (!zdev) dZ = ar.take<double>(G * G);
        if (!xdev) dX = ar.take<double>(G * o);
        if (!ydev) dY = ar.take<double>(G * v);
        if (!udev) dU = ar.take<double>(v * o * nt);
        if (yxu_diag && !ddev) dD = ar.take<double>(G * nt);
        YU = ar.take<double>(G * o * nmu);
        part = ar.take<double>((int64_t)npairs * nmu * 2);
        S = ar.take<double>(2 * nt);
    };
    Arena count, ar;
    layout(count);
    BRPA_TRY(ws_reserve(ctx, count.off, &ar));
    dZ = Zgh;
    dX = Xgi;
    dY = Yga;
    dU = Uaim;
    dD = yxu_diag;
    layout(ar);
    if (!zdev) BRPA_TRY(copy_sync(ctx, const_cast<double *>(dZ), Zgh, 8 * G * G, hipMemcpyHostToDevice));
    if (!xdev) BRPA_TRY(copy_sync(ctx, const_cast<double *>(dX), Xgi, 8 * G * o, hipMemcpyHostToDevice));
    if (!ydev) BRPA_TRY(copy_sync(ctx, const_cast<double *>(dY), Yga, 8 * G * v, hipMemcpyHostToDevice));
    if (!udev)
        BRPA_TRY(copy_sync(ctx, const_cast<double *>(dU), Uaim, 8 * v * o * nt, hipMemcpyHostToDevice));
    BRPA_TRY(rocblas_set_stream(h, ctx->stream) == rocblas_status_success
                 ? BRPA_OK
                 : brpa_fail(ctx, BRPA_ERR_SOLVER, "rocblas_set_stream failed"));
    BRPA_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    const double one = 1.0, zero = 0.0;
    for (int64_t mu0 = 0; mu0 < nt; mu0 += nmu) {
        const int64_t m = std::min(nmu, nt - mu0);
        // YU(:, i + o mu) = Yga U(:, i, mu): column-major G x (o m), K = v
        if (rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, (rocblas_int)G,
                          (rocblas_int)(o * m), (rocblas_int)v, &one, dY, (rocblas_int)G,
                          dU + v * o * mu0, (rocblas_int)v, &zero, YU, (rocblas_int)G) !=
            rocblas_status_success)
            return brpa_fail(ctx, BRPA_ERR_SOLVER, "brpa_thc_1b2g: rocblas_dgemm failed");
        BRPA_TRY(launch_thc_1b2g(ctx, YU, dX, dZ, G, o, m, part, yxu_diag ? dD : nullptr, mu0, S));
    }
    BRPA_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    std::vector<double> hs(2 * nt);
    BRPA_TRY(copy_sync(ctx, hs.data(), S, 8 * 2 * nt, hipMemcpyDeviceToHost));
    if (yxu_diag && !ddev) BRPA_TRY(copy_sync(ctx, yxu_diag, dD, 8 * G * nt, hipMemcpyDeviceToHost));
    float ms = 0.f;
    BRPA_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    ctx->last_kernel_ms = ms;
    // Ec2g = 2 sum_mu (-2 S2g A^2), Ec1b = 2 sum_mu (-S1b A) (rpa_CCD_Corrections.f90:137-160)
    long double e2 = 0.0L, e1 = 0.0L;
    for (int64_t mu = 0; mu < nt; ++mu) {
        const double s2 = hs[2 * mu], s1 = hs[2 * mu + 1];
        e2 += -2.0L * s2 * a[mu] * a[mu];
        e1 += -(long double)s1 * a[mu];
        if (eig_2g) eig_2g[mu] = -4.0 * s2 * a[mu] * a[mu];
        if (eig_sosex) eig_sosex[mu] = -s1 * a[mu];
    }
    *ec2g = (double)(2.0L * e2);
    *ec1b = (double)(2.0L * e1);
    return BRPA_OK;
}

}  // extern "C"
