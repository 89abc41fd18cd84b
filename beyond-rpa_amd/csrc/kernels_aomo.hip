// AO -> MO transform of the Cholesky vectors (rpa_CRG / rpa_C / rpa_Cab / rpa_Caa,
// rpa/rpa_core_MO.f90:40-301): the scatter of the packed AO-pair columns into
// full symmetric AO-pair matrices, before the two quarter transforms (rocBLAS
// dgemm in aomo.cpp).
//
// Layouts: RG(P, pq) at P + nvecs pq (Fortran RG(NVecs, Npq)); pairs: the AO
// indices (p, q) of each packed column; RGfull[q][p][P] with P fastest.
#include "device_common.h"
#include "internal.h"

#define LAUNCH_CHECK(ctx)                                                                      \
    do {                                                                                       \
        hipError_t _e = hipGetLastError();                                                     \
        if (_e != hipSuccess)                                                                  \
            return brpa_fail((ctx), BRPA_ERR_HIP, "kernel launch failed: %s (%s:%d)",          \
                             hipGetErrorString(_e), __FILE__, __LINE__);                       \
    } while (0)

// Unpack the packed AO-pair columns of a chunk of Cholesky vectors into the
// full symmetric AO-pair matrices: RGfull[q][p][P] = RGfull[p][q][P] =
// RG(P0 + P, pq) for the pair (p, q) of column pq (P fastest, Pc per pair).
__global__ void k_unpack_pairs(const double *__restrict__ RG, const int2 *__restrict__ pairs,
                               int64_t npq, int64_t NV, int64_t P0, int64_t Pc, int64_t nao,
                               double *__restrict__ full) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= npq * Pc) return;
    const int64_t pq = e / Pc, P = e - pq * Pc;
    const int2 pr = pairs[pq];
    const double v = RG[(P0 + P) + NV * pq];
    full[((int64_t)pr.y * nao + pr.x) * Pc + P] = v;
    full[((int64_t)pr.x * nao + pr.y) * Pc + P] = v;
}

int launch_unpack_pairs(brpa_ctx *ctx, const double *RG, const int *pairs, int64_t npq, int64_t NV,
                        int64_t P0, int64_t Pc, int64_t nao, double *full) {
    const int64_t n = npq * Pc;
    if (n <= 0) return BRPA_OK;
    hipLaunchKernelGGL(k_unpack_pairs, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream,
                       RG, reinterpret_cast<const int2 *>(pairs), npq, NV, P0, Pc, nao, full);
    LAUNCH_CHECK(ctx);
    return BRPA_OK;
}
