// Split-BF16 Pi(u) build on gfx950 bf16 MFMA (v_mfma_f32_16x16x32_bf16).
//
// X_w = diag(sqrt(c_w)) R is cut into 128-row blocks.  In block b, column P
// gets a power-of-two scale s = 2^(e-8) with max_k |X[k][P]| < 2^e, and every
// element is written as s (n0 + n1/256 + n2/65536) with integers |n0| <= 256,
// |n1|, |n2| <= 128: exact in bf16, representation error <= s/2^17.
// Pi_w[p][q] = sum_b s_p s_q sum_k (x_p x_q in slices), and the six slice
// products with level a+b <= 2 are accumulated per level in FP32:
//   L0 = n0 n0,  L1 = n0 n1 + n1 n0,  L2 = n0 n2 + n2 n0 + n1 n1.
// Over 128 rows every partial sum is an integer below 2^24, so the FP32 MFMA
// accumulation is exact; each block is flushed to FP64 as
// s_p s_q (L0 + L1/2^8 + L2/2^16).  The diagonal (and so tr Pi) is replaced
// by the exact FP64 value afterwards (k_diag / k_set_diag), which removes the
// systematic diagonal bias of the dropped level-3/4 products.
#include "device_common.h"
#include "internal.h"

#define LAUNCH_CHECK(ctx)                                                                      \
    do {                                                                                       \
        hipError_t _e = hipGetLastError();                                                     \
        if (_e != hipSuccess)                                                                  \
            return brpa_fail((ctx), BRPA_ERR_HIP, "kernel launch failed: %s (%s:%d)",          \
                             hipGetErrorString(_e), __FILE__, __LINE__);                       \
    } while (0)

__device__ __forceinline__ uint16_t int_to_bf16_bits(double n) {
    float f = (float)n;   // |n| <= 256 integer: exact in float and in bf16
    return (uint16_t)(__float_as_uint(f) >> 16);
}

// Work-groups per fragment: F * nblk * ceil(N/64).  One work-group splits
// a 128-row x 64-column block of X_w.
__global__ __launch_bounds__(256) void k_split(const FragDev *__restrict__ fr,
                                              const int64_t *__restrict__ pre, int nfrag) {
    __shared__ double sX[SPLIT_KB][65];
    __shared__ double sMax[4][64];
    __shared__ int sExp[64];
    int64_t b = blockIdx.x;
    int f = find_frag(pre, nfrag, b);
    const FragDev &F = fr[f];
    const int N = F.N, K = F.K;
    const int ncc = (N + 63) / 64;
    int64_t lb = b - pre[f];
    const int w = (int)(lb / ((int64_t)F.nblk * ncc));
    int64_t rem = lb - (int64_t)w * F.nblk * ncc;
    const int blk = (int)(rem / ncc), cc = (int)(rem - (int64_t)blk * ncc);
    const int k0 = blk * SPLIT_KB, P0 = cc * 64;
    const int tid = threadIdx.x;

    // thread tid always handles column tid%64 (256 is a multiple of 64), rows
    // tid/64 + 4 i; it also accumulates that column's exact diagonal partial
    // sum_k c_w(k) R[k][P]^2 in double-double.
    dd_t dsum = {0.0, 0.0};
    for (int e = tid; e < SPLIT_KB * 64; e += 256) {
        int r = e >> 6, c = e & 63;
        int k = k0 + r, P = P0 + c;
        double x = 0.0;
        if (k < K && P < N) {
            const double cw = F.dtab[(int64_t)w * K + k];
            const double rv = F.R[(int64_t)k * N + P];
            x = sqrt(cw) * rv;
            const double r2 = rv * rv, r2e = fma(rv, rv, -r2);
            const double p = cw * r2, pe = fma(cw, r2, -p) + cw * r2e;
            dsum = dd_add(dsum, p);
            dsum.lo += pe;
        }
        sX[r][c] = x;
    }
    __shared__ double sDh[4][64], sDl[4][64];
    sDh[tid >> 6][tid & 63] = dsum.hi;
    sDl[tid >> 6][tid & 63] = dsum.lo;
    __syncthreads();
    if (tid < 64 && P0 + tid < N) {
        dd_t t = {sDh[0][tid], sDl[0][tid]};
        for (int g = 1; g < 4; ++g) t = dd_add(t, (dd_t){sDh[g][tid], sDl[g][tid]});
        double *q = F.dpart + (((int64_t)w * F.nblk + blk) * N + P0 + tid) * 2;
        q[0] = t.hi;
        q[1] = t.lo;
    }
    {
        int c = tid & 63, g = tid >> 6;
        double m = 0.0;
        for (int r = g * 32; r < g * 32 + 32; ++r) m = fmax(m, fabs(sX[r][c]));
        sMax[g][c] = m;
    }
    __syncthreads();
    if (tid < 64) {
        double m = fmax(fmax(sMax[0][tid], sMax[1][tid]), fmax(sMax[2][tid], sMax[3][tid]));
        int e = 0;
        if (m > 0.0) frexp(m, &e);   // m = f 2^e, 0.5 <= f < 1  ->  m < 2^e
        sExp[tid] = e;
        if (P0 + tid < N)
            F.scale[((int64_t)w * F.nblk + blk) * N + P0 + tid] = (m > 0.0) ? ldexpf(1.0f, e - 8) : 0.0f;
    }
    __syncthreads();
    // thread -> column c = tid/4, rows (tid%4)*32 .. +32 (contiguous in the output)
    const int c = tid >> 2, r0 = (tid & 3) * 32;
    const int P = P0 + c;
    if (P >= N) return;
    const int e = sExp[c];
    uint16_t *s0 = F.sl + (((int64_t)w * 3 + 0) * N + P) * F.Kpad + k0 + r0;
    uint16_t *s1 = F.sl + (((int64_t)w * 3 + 1) * N + P) * F.Kpad + k0 + r0;
    uint16_t *s2 = F.sl + (((int64_t)w * 3 + 2) * N + P) * F.Kpad + k0 + r0;
    for (int r = 0; r < 32; r += 8) {
        uint16_t v0[8], v1[8], v2[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            double x = ldexp(sX[r0 + r + i][c], 8 - e);   // |x| < 256, exact
            double n0 = rint(x);
            double t1 = (x - n0) * 256.0;                  // exact
            double n1 = rint(t1);
            double t2 = (t1 - n1) * 256.0;                 // exact
            double n2 = rint(t2);
            v0[i] = int_to_bf16_bits(n0);
            v1[i] = int_to_bf16_bits(n1);
            v2[i] = int_to_bf16_bits(n2);
        }
        uint4 p0, p1, p2;
        p0.x = v0[0] | ((uint32_t)v0[1] << 16); p0.y = v0[2] | ((uint32_t)v0[3] << 16);
        p0.z = v0[4] | ((uint32_t)v0[5] << 16); p0.w = v0[6] | ((uint32_t)v0[7] << 16);
        p1.x = v1[0] | ((uint32_t)v1[1] << 16); p1.y = v1[2] | ((uint32_t)v1[3] << 16);
        p1.z = v1[4] | ((uint32_t)v1[5] << 16); p1.w = v1[6] | ((uint32_t)v1[7] << 16);
        p2.x = v2[0] | ((uint32_t)v2[1] << 16); p2.y = v2[2] | ((uint32_t)v2[3] << 16);
        p2.z = v2[4] | ((uint32_t)v2[5] << 16); p2.w = v2[6] | ((uint32_t)v2[7] << 16);
        *reinterpret_cast<uint4 *>(s0 + r) = p0;
        *reinterpret_cast<uint4 *>(s1 + r) = p1;
        *reinterpret_cast<uint4 *>(s2 + r) = p2;
    }
}

int launch_split(brpa_ctx *ctx, const FragDev *fr, const int64_t *pre, int nfrag, int64_t total) {
    if (total <= 0) return BRPA_OK;
    hipLaunchKernelGGL(k_split, dim3((unsigned)total), dim3(256), 0, ctx->stream, fr, pre, nfrag);
    LAUNCH_CHECK(ctx);
    return BRPA_OK;
}

// ---------------------------------------------------------------------------
// 128x128 tile pair per work-group, 8 waves as 2 (rows) x 4 (cols) sub-tiles
// of 64x32: 2 waves per SIMD, FP32 level accumulators in AGPRs, the FP64 block
// accumulator, stage prefetch and operand fragments in VGPRs.
// MFMA v_mfma_f32_16x16x32_bf16 operand layout (lane l):
//   A[m = l%16][k = 8*(l/16) .. +7], B[k = 8*(l/16) .. +7][n = l%16],
//   D: 4 floats, D[r] = C[m = 4*(l/16) + r][n = l%16].
// Both operands are columns of X (k contiguous per column), so A and B are
// read from LDS the same way.
constexpr int T16 = 128;
constexpr int KS = 32;          // rows per LDS stage (one MFMA k-depth)
constexpr int LROW = KS + 8;    // bf16 per LDS row (+8 pad: 80 B row stride)

__global__ __launch_bounds__(512) void k_syrk_bf16(const FragDev *__restrict__ fr,
                                                  const int64_t *__restrict__ pre, int nfrag) {
    __shared__ __attribute__((aligned(16))) uint16_t sA[3][T16][LROW];
    __shared__ __attribute__((aligned(16))) uint16_t sB[3][T16][LROW];
    __shared__ float sSP[T16], sSQ[T16];
    int64_t b = blockIdx.x;
    int f = find_frag(pre, nfrag, b);
    const FragDev &F = fr[f];
    const int N = F.N;
    const int nt = (N + T16 - 1) / T16;
    const int npairs = nt * (nt + 1) / 2;
    int64_t lb = b - pre[f];
    const int w = (int)(lb / npairs), pair = (int)(lb - (int64_t)w * npairs);
    int tp, tq;
    tri_decode(pair, tp, tq);
    const int p0 = tp * T16, q0 = tq * T16;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int wi = wave >> 2, wj = wave & 3;     // 64-row half, 32-col quarter

    const uint16_t *slw = F.sl + (int64_t)w * 3 * N * F.Kpad;
    const float *scw = F.scale + (int64_t)w * F.nblk * N;

    f4_t L0[4][2], L1[4][2], L2[4][2];
    double acc[4][2][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            L0[i][j] = L1[i][j] = L2[i][j] = (f4_t){0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int r = 0; r < 4; ++r) acc[i][j][r] = 0.0;
        }

    // global -> register prefetch of one stage: 2 sides x 3 slices x 128 columns
    // x 4 pieces of 8 bf16 (16 bytes) = 3072 pieces, 6 per thread.  Thread t
    // always takes column t/4, piece t%4 (consecutive threads read one
    // column's 64 contiguous bytes) of every slice on both sides.
    uint4 pf[6];
    const int lcol = tid >> 2, lkof = 8 * (tid & 3);
    const bool okP = p0 + lcol < N, okQ = q0 + lcol < N;
    const int64_t sstride = (int64_t)N * F.Kpad;
    const uint16_t *gA = slw + (int64_t)(okP ? p0 + lcol : 0) * F.Kpad + lkof;
    const uint16_t *gB = slw + (int64_t)(okQ ? q0 + lcol : 0) * F.Kpad + lkof;
    auto gload = [&](int k0) {
#pragma unroll
        for (int s = 0; s < 3; ++s) {
            pf[s] = okP ? *reinterpret_cast<const uint4 *>(gA + s * sstride + k0) : make_uint4(0, 0, 0, 0);
            pf[3 + s] = okQ ? *reinterpret_cast<const uint4 *>(gB + s * sstride + k0) : make_uint4(0, 0, 0, 0);
        }
    };
    auto lstore = [&]() {
#pragma unroll
        for (int s = 0; s < 3; ++s) {
            *reinterpret_cast<uint4 *>(&sA[s][lcol][lkof]) = pf[s];
            *reinterpret_cast<uint4 *>(&sB[s][lcol][lkof]) = pf[3 + s];
        }
    };
    auto afrag = [&](int s, int i) {
        return *reinterpret_cast<const bf8_t *>(&sA[s][64 * wi + 16 * i + (lane & 15)][8 * (lane >> 4)]);
    };
    auto bfrag = [&](int s, int j) {
        return *reinterpret_cast<const bf8_t *>(&sB[s][32 * wj + 16 * j + (lane & 15)][8 * (lane >> 4)]);
    };

    const int nblk = F.nblk;
    const int ksteps = SPLIT_KB / KS;
    gload(0);
    for (int blk = 0; blk < nblk; ++blk) {
        for (int ks = 0; ks < ksteps; ++ks) {
            const int k0 = blk * SPLIT_KB + ks * KS;
            __syncthreads();
            lstore();
            if (ks == 0 && tid < 2 * T16) {
                int side = tid / T16, c = tid - side * T16;
                int P = (side ? q0 : p0) + c;
                float sc = (P < N) ? scw[(int64_t)blk * N + P] : 0.0f;
                if (side)
                    sSQ[c] = sc;
                else
                    sSP[c] = sc;
            }
            __syncthreads();
            if (k0 + KS < F.Kpad) gload(k0 + KS);
            // products grouped by A slice and by pairs of 16-row blocks so that
            // few fragments are live at once
            {
                bf8_t b0[2], b1[2], b2[2];
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    b0[j] = bfrag(0, j);
                    b1[j] = bfrag(1, j);
                    b2[j] = bfrag(2, j);
                }
#pragma unroll
                for (int ih = 0; ih < 4; ih += 2) {
                    bf8_t a[2];
#pragma unroll
                    for (int ii = 0; ii < 2; ++ii) a[ii] = afrag(0, ih + ii);
#pragma unroll
                    for (int ii = 0; ii < 2; ++ii)
#pragma unroll
                        for (int j = 0; j < 2; ++j) {
                            const int i = ih + ii;
                            L0[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[ii], b0[j], L0[i][j], 0, 0, 0);
                            L1[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[ii], b1[j], L1[i][j], 0, 0, 0);
                            L2[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[ii], b2[j], L2[i][j], 0, 0, 0);
                        }
#pragma unroll
                    for (int ii = 0; ii < 2; ++ii) a[ii] = afrag(1, ih + ii);
#pragma unroll
                    for (int ii = 0; ii < 2; ++ii)
#pragma unroll
                        for (int j = 0; j < 2; ++j) {
                            const int i = ih + ii;
                            L1[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[ii], b0[j], L1[i][j], 0, 0, 0);
                            L2[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[ii], b1[j], L2[i][j], 0, 0, 0);
                        }
#pragma unroll
                    for (int ii = 0; ii < 2; ++ii) a[ii] = afrag(2, ih + ii);
#pragma unroll
                    for (int ii = 0; ii < 2; ++ii)
#pragma unroll
                        for (int j = 0; j < 2; ++j) {
                            const int i = ih + ii;
                            L2[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[ii], b0[j], L2[i][j], 0, 0, 0);
                        }
                }
            }
        }
        // flush the exact level sums of this block to FP64
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const double sq = (double)sSQ[32 * wj + 16 * j + (lane & 15)];
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const double sp = (double)sSP[64 * wi + 16 * i + 4 * (lane >> 4) + r];
                    double v = (double)L0[i][j][r] + (double)L1[i][j][r] * (1.0 / 256.0) +
                               (double)L2[i][j][r] * (1.0 / 65536.0);
                    acc[i][j][r] = fma(v, sp * sq, acc[i][j][r]);
                }
                L0[i][j] = L1[i][j] = L2[i][j] = (f4_t){0.f, 0.f, 0.f, 0.f};
            }
    }

    double *C = F.Pi + (int64_t)w * N * N;
    const bool diag = (tp == tq);
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                int p = p0 + 64 * wi + 16 * i + 4 * (lane >> 4) + r;
                int q = q0 + 32 * wj + 16 * j + (lane & 15);
                if (p < N && q < N) {
                    C[(int64_t)p * N + q] = acc[i][j][r];
                    if (!diag) C[(int64_t)q * N + p] = acc[i][j][r];
                }
            }
}

int launch_syrk_bf16(brpa_ctx *ctx, const FragDev *fr, const int64_t *pre, int nfrag,
                     int64_t total) {
    if (total <= 0) return BRPA_OK;
    hipLaunchKernelGGL(k_syrk_bf16, dim3((unsigned)total), dim3(512), 0, ctx->stream, fr, pre,
                       nfrag);
    LAUNCH_CHECK(ctx);
    return BRPA_OK;
}
