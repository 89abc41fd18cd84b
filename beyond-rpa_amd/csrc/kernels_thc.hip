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
#include <algorithm>

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

constexpr int TT = 128;       // output tile: 128 ai x 128 P
constexpr int GB = 8;         // grid points per LDS stage
constexpr int CW = GB;        // lanes per operand row in the global loads (g fastest)
constexpr int RPP = 256 / CW; // operand rows per load pass
constexpr int NL = TT / RPP;  // load passes per stage and operand

// LDS column swizzle of row k: col ^ swz(k).  The global loads run along g, so
// the 16 lanes of a store sit on 16 different LDS rows; with an unpadded
// 128-double row (256 dwords, a multiple of the 64 banks) the swizzle puts
// those 16 lanes on 16 different bank pairs, and the two rows one half-wave
// of an MFMA operand read covers (k even, k + 1) into opposite 16-column
// halves, i.e. disjoint banks.  XOR with a constant below 32 keeps every
// column inside its aligned 32-column block, so the map is a bijection of
// each row onto itself.
__device__ __forceinline__ int swz(int k) { return ((k & 1) << 4) | ((k >> 1) << 1); }

// One work-group = one 128 x 128 tile of R; the 4 waves split it 2 x 2 (64 x 64
// each, 4 x 4 MFMA blocks, 32 MFMAs per wave per 8-point stage).  The next
// stage's X(g, a), X(g, i) and Z(g, P) are loaded into registers during the
// MFMAs, and X(g, a) X(g, i) is formed only when it goes to LDS, so that no
// wave waits on those loads before its MFMAs.  Two waves per SIMD (launch
// bound).  MFMA f64 16x16x4 operand layout (lane l): A[m = l%16][k = l/16],
// B[k = l/16][n = l%16], D[r] = C[m = l/16 + 4r][n = l%16].
__global__ __launch_bounds__(256, 2) void k_thc_r(const double *__restrict__ Xa,
                                                  const double *__restrict__ Xi,
                                                  const double *__restrict__ Z, int64_t G,
                                                  int nvir, int64_t K, int64_t N,
                                                  double *__restrict__ R) {
    __shared__ double sH[GB][TT];   // H[g][ai] = X(g, a) X(g, i), swizzled columns
    __shared__ double sZ[GB][TT];   // Z[g][P], swizzled columns
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int64_t ai0 = (int64_t)blockIdx.x * TT, p0 = (int64_t)blockIdx.y * TT;
    const int wm = wave >> 1, wn = wave & 1;

    // this thread loads grid offset c of rows r = tid / CW + RPP it, it < NL;
    // 32-bit element offsets (the launcher checks that every array has fewer
    // than 2^31 elements)
    const int c = tid % CW;
    int oa[NL], oi[NL], oz[NL];
    unsigned hok = 0, zok = 0;
#pragma unroll
    for (int it = 0; it < NL; ++it) {
        const int r = tid / CW + RPP * it;
        const int64_t ai = ai0 + r, p = p0 + r;
        const bool h = ai < K, z = p < N;
        hok |= (unsigned)h << it;
        zok |= (unsigned)z << it;
        const int i = h ? (int)(ai / nvir) : 0, a = h ? (int)(ai - (int64_t)i * nvir) : 0;
        oa[it] = a * (int)G + c;
        oi[it] = i * (int)G + c;
        oz[it] = (z ? (int)p : 0) * (int)G + c;
    }

    d4_t acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = (d4_t){0.0, 0.0, 0.0, 0.0};

    double ra[NL], ri[NL], rz[NL];
    auto gload = [&](int g0) {
        const bool gin = g0 + c < G;
#pragma unroll
        for (int it = 0; it < NL; ++it) {
            const bool h = gin && ((hok >> it) & 1u);
            ra[it] = h ? Xa[oa[it] + g0] : 0.0;
            ri[it] = h ? Xi[oi[it] + g0] : 0.0;
            rz[it] = (gin && ((zok >> it) & 1u)) ? Z[oz[it] + g0] : 0.0;
        }
    };

    gload(0);
    for (int g0 = 0; g0 < (int)G; g0 += GB) {
        __syncthreads();
#pragma unroll
        for (int it = 0; it < NL; ++it) {
            const int r = (tid / CW + RPP * it) ^ swz(c);
            sH[c][r] = ra[it] * ri[it];
            sZ[c][r] = rz[it];
        }
        __syncthreads();
        if (g0 + GB < (int)G) gload(g0 + GB);
#pragma unroll
        for (int kk = 0; kk < GB; kk += 4) {
            const int kr = kk + (lane >> 4);
            const int sw = swz(kr);
            double a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) a[i] = sH[kr][(64 * wm + 16 * i + (lane & 15)) ^ sw];
#pragma unroll
            for (int j = 0; j < 4; ++j) b[j] = sZ[kr][(64 * wn + 16 * j + (lane & 15)) ^ sw];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
        }
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t ai = ai0 + 64 * wm + 16 * i + (lane >> 4) + 4 * r;
                const int64_t p = p0 + 64 * wn + 16 * j + (lane & 15);
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
    const int64_t cols = std::max<int64_t>(std::max<int64_t>(nocc, nvir), N);
    if (gx > INT32_MAX || gy > 65535 || G * cols + 64 > INT32_MAX)
        return brpa_fail(ctx, BRPA_ERR_ARG,
                         "THC R: ngrid x max(nocc, nvir, naux) must stay below 2^31 elements");
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
