// Tensor-hypercontraction (THC) producer of the Cholesky vectors:
//
//   R[ai][P] = sum_g X(g, a) X(g, i) Z(g, P),   ai = i * nvir + a
//
// (rpa_THC_Rkab, rpa/rpa_THC.f90:373-403, as called by rpa_THC_PiU_BlockAlgo
// 285-339).  The reference forms Xgab(g, ai) = X(g, a) X(g, i) for a block of
// ai and multiplies Z^T by it; here the Hadamard product is formed while the
// GEMM operand is loaded, so no NGrid x ov intermediate exists.  The reduction
// over the grid runs on v_mfma_f64_16x16x4_f64.
//
// Inputs are the reference's column-major arrays: Xga(g, a) at g + G a,
// Xgi(g, i) at g + G i, Zgk(g, P) at g + G P; all three are contiguous in g,
// which is the reduction index, so both MFMA operands load along k.
#include "device_common.h"
#include "internal.h"

#define LAUNCH_CHECK(ctx)                                                                      \
    do {                                                                                       \
        hipError_t _e = hipGetLastError();                                                     \
        if (_e != hipSuccess)                                                                  \
            return brpa_fail((ctx), BRPA_ERR_HIP, "kernel launch failed: %s (%s:%d)",          \
                             hipGetErrorString(_e), __FILE__, __LINE__);                       \
    } while (0)

namespace {

constexpr int TT = 64;        // output tile: 64 ai x 64 P
constexpr int GB = 16;        // grid points per LDS stage
constexpr int PAD = 16;       // LDS row padding (doubles)

// One work-group = one 64 x 64 tile of R; the 4 waves split it 2 x 2 (32 x 32
// each, 2 x 2 MFMA blocks).  MFMA f64 16x16x4 operand layout (lane l):
// A[m = l%16][k = l/16], B[k = l/16][n = l%16], D[r] = C[m = l/16 + 4r][n = l%16].
__global__ __launch_bounds__(256) void k_thc_r(const double *__restrict__ Xa,
                                               const double *__restrict__ Xi,
                                               const double *__restrict__ Z, int64_t G,
                                               int nvir, int64_t K, int64_t N,
                                               double *__restrict__ R) {
    __shared__ double sH[GB][TT + PAD];   // H[g][ai] = X(g, a) X(g, i)
    __shared__ double sZ[GB][TT + PAD];   // Z[g][P]
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int64_t ai0 = (int64_t)blockIdx.x * TT, p0 = (int64_t)blockIdx.y * TT;
    const int wm = wave >> 1, wn = wave & 1;

    // this thread loads grid offset c of rows r = (tid >> 4) + 16 it, it = 0..3
    const int c = tid & 15;
    const double *xa[4], *xi[4], *zp[4];
    bool hok[4], zok[4];
#pragma unroll
    for (int it = 0; it < 4; ++it) {
        const int r = (tid >> 4) + 16 * it;
        const int64_t ai = ai0 + r, p = p0 + r;
        hok[it] = ai < K;
        zok[it] = p < N;
        const int64_t i = hok[it] ? ai / nvir : 0, a = hok[it] ? ai - i * nvir : 0;
        xa[it] = Xa + a * G + c;
        xi[it] = Xi + i * G + c;
        zp[it] = Z + (zok[it] ? p : 0) * G + c;
    }

    d4_t acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = (d4_t){0.0, 0.0, 0.0, 0.0};

    double rh[4], rz[4];
    auto gload = [&](int64_t g0) {
        const bool gin = g0 + c < G;
#pragma unroll
        for (int it = 0; it < 4; ++it) {
            rh[it] = (gin && hok[it]) ? xa[it][g0] * xi[it][g0] : 0.0;
            rz[it] = (gin && zok[it]) ? zp[it][g0] : 0.0;
        }
    };
    auto lstore = [&]() {
#pragma unroll
        for (int it = 0; it < 4; ++it) {
            const int r = (tid >> 4) + 16 * it;
            sH[c][r] = rh[it];
            sZ[c][r] = rz[it];
        }
    };

    gload(0);
    for (int64_t g0 = 0; g0 < G; g0 += GB) {
        __syncthreads();
        lstore();
        __syncthreads();
        if (g0 + GB < G) gload(g0 + GB);
#pragma unroll
        for (int kk = 0; kk < GB; kk += 4) {
            const int kr = kk + (lane >> 4);
            double a[2], b[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) a[i] = sH[kr][32 * wm + 16 * i + (lane & 15)];
#pragma unroll
            for (int j = 0; j < 2; ++j) b[j] = sZ[kr][32 * wn + 16 * j + (lane & 15)];
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
        }
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t ai = ai0 + 32 * wm + 16 * i + (lane >> 4) + 4 * r;
                const int64_t p = p0 + 32 * wn + 16 * j + (lane & 15);
                if (ai < K && p < N) R[ai * N + p] = acc[i][j][r];
            }
}

__global__ void k_axpy(double *__restrict__ y, const double *__restrict__ x, int64_t n) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e < n) y[e] += x[e];
}

}  // namespace

int launch_thc_r(brpa_ctx *ctx, const double *Xa, const double *Xi, const double *Z, int64_t G,
                 int64_t nocc, int64_t nvir, int64_t N, double *R) {
    const int64_t K = nocc * nvir;
    if (K <= 0 || N <= 0) return BRPA_OK;
    const int64_t gx = (K + TT - 1) / TT, gy = (N + TT - 1) / TT;
    if (gx > INT32_MAX || gy > 65535 || nvir > INT32_MAX)
        return brpa_fail(ctx, BRPA_ERR_ARG, "THC R: sizes beyond the supported grid");
    hipLaunchKernelGGL(k_thc_r, dim3((unsigned)gx, (unsigned)gy), dim3(256), 0, ctx->stream, Xa, Xi,
                       Z, G, (int)nvir, K, N, R);
    LAUNCH_CHECK(ctx);
    return BRPA_OK;
}

int launch_axpy(brpa_ctx *ctx, double *y, const double *x, int64_t n) {
    if (n <= 0) return BRPA_OK;
    hipLaunchKernelGGL(k_axpy, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, y, x,
                       n);
    LAUNCH_CHECK(ctx);
    return BRPA_OK;
}
