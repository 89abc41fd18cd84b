// Split-int8 GEMM on gfx950 int8 MFMA (v_mfma_i32_16x16x64_i8) for the
// amplitude contractions of the split form:
//
//   out[m*sm + n*sn] = alpha rs[m] sa[m] sb[n] sum_k a_m[k] b_n[k] + beta out[...]
//
// Operands.  Every "row" m of A and n of B is a K-vector stored as int8 slices
// with one scale over the whole vector, x = s (n0 + n1 2^-7 + ... + n4 2^-28),
// |n0| <= 127, |n1..n4| <= 64 (the slicing of the SOSEX contraction).  The 15
// slice products of level <= 4 go to five exact int32 level sums (|L4| <=
// K * 28544 < 2^31 for K <= 65536) combined once in FP64; the dropped level
// >= 5 products are ~2^-35 relative to the vector maxima.
//
//   k_colslice_i8  scale and slices of vectors stored contiguously (a column of
//                  a column-major matrix), optionally row-scaled on the fly
//   k_gemm_i8      128 x 128 (or 128 x 64) output tiles, waves of 32 x 32,
//                  stages of 64 k straight from HBM to LDS
//                  (global_load_lds_dwordx4, double-buffered, XOR-swizzled)
// Vectors stored across rows of a row-major matrix use the SOSEX block slicer
// (k_blkscale_i8 / k_blksplit_i8 with one block).
#include "device_common.h"
#include "internal.h"

#define LAUNCH_CHECK(ctx)                                                                      \
    do {                                                                                       \
        hipError_t _e = hipGetLastError();                                                     \
        if (_e != hipSuccess)                                                                  \
            return brpa_fail((ctx), BRPA_ERR_HIP, "kernel launch failed: %s (%s:%d)",          \
                             hipGetErrorString(_e), __FILE__, __LINE__);                       \
    } while (0)

typedef __attribute__((address_space(1))) const double gdbl_t;
typedef int v4i_t __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) void lds_void_t;

constexpr int GM_NSL = 5;   // slices per value

// ---------------------------------------------------------------------------
// One work-group per vector c: x_k = X[c*ld + k] (times rs[k] when rs is not
// null), k < len.  scale[c] = max_k |x_k| / 127; slices out[(s*ncols + c)*kpad
// + k], zero for len <= k < kpad.  The same product x_k * rs[k] * (1/scale)
// is formed in both passes, so |n0| <= 127.
__global__ __launch_bounds__(256) void k_colslice_i8(const double *__restrict__ X, int64_t ld,
                                                    int len, int64_t ncols,
                                                    const double *__restrict__ rs, int kpad,
                                                    double *__restrict__ scale,
                                                    int8_t *__restrict__ out) {
    __shared__ double sM[256];
    const int64_t c = blockIdx.x;
    const int tid = threadIdx.x;
    const gdbl_t *x = (const gdbl_t *)(X + c * ld);
    const gdbl_t *r = (const gdbl_t *)rs;
    double m = 0.0;
    for (int k = tid; k < len; k += 256) m = fmax(m, fabs(rs ? x[k] * r[k] : x[k]));
    sM[tid] = m;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (tid < h) sM[tid] = fmax(sM[tid], sM[tid + h]);
        __syncthreads();
    }
    const double sc = sM[0] / 127.0;
    if (tid == 0) scale[c] = sc;
    const double inv = sc > 0.0 ? 1.0 / sc : 0.0;
    for (int k0 = 16 * tid; k0 < kpad; k0 += 16 * 256) {
        uint32_t word[GM_NSL][4];
#pragma unroll
        for (int s = 0; s < GM_NSL; ++s)
#pragma unroll
            for (int q = 0; q < 4; ++q) word[s][q] = 0u;
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int k = k0 + u;
            double t = 0.0;
            if (k < len) t = (rs ? x[k] * r[k] : x[k]) * inv;   // |t| <= 127 (up to rounding)
#pragma unroll
            for (int s = 0; s < GM_NSL; ++s) {
                const double n = rint(t);
                t = (t - n) * 128.0;   // exact
                word[s][u >> 2] |= ((uint32_t)(uint8_t)(int8_t)(int)n) << (8 * (u & 3));
            }
        }
#pragma unroll
        for (int s = 0; s < GM_NSL; ++s)
            *reinterpret_cast<uint4 *>(out + ((int64_t)s * ncols + c) * kpad + k0) =
                make_uint4(word[s][0], word[s][1], word[s][2], word[s][3]);
    }
}

int launch_colslice_i8(brpa_ctx *ctx, const double *X, int64_t ld, int64_t len, int64_t ncols,
                       const double *rs, int64_t kpad, double *scale, int8_t *out) {
    if (ncols <= 0) return BRPA_OK;
    if (len <= 0 || kpad < len || kpad % 16 != 0 || ld < len || ncols > INT32_MAX)
        return brpa_fail(ctx, BRPA_ERR_ARG, "launch_colslice_i8: bad shape");
    hipLaunchKernelGGL(k_colslice_i8, dim3((unsigned)ncols), dim3(256), 0, ctx->stream, X, ld,
                       (int)len, ncols, rs, (int)kpad, scale, out);
    LAUNCH_CHECK(ctx);
    return BRPA_OK;
}

// ---------------------------------------------------------------------------
// Output tile TM x TN = 16 RI WM x 16 RJ WN: WM * WN waves, each owning RI x RJ
// MFMA tiles of 16 x 16; work-groups:
// ceil(M/TM) * ceil(N/TN), tile index = tm * ntn + tn.  The 5 slices of the
// TM A-vectors and of the TN B-vectors of a stage (64 bytes each) are
// 5 (TM + TN) / 16 wave loads of 1 KB; lane l lands in vector l/4, slot l%4,
// and fetches k-piece (l%4) ^ ((l/16)%4), so that the MFMA operand reads (16
// vectors at one k-piece per 16 lanes) hit 16 distinct 4-bank groups.  Rows
// past M (N) read row M-1 (N-1) and are never written.
__device__ __forceinline__ int swzg(int col_local, int piece) {
    return piece ^ ((col_local >> 2) & 3);
}

template <int WM, int WN, int RI, int RJ>
__global__ __launch_bounds__(64 * WM * WN) void k_gemm_i8(
    const int8_t *__restrict__ SA, const double *__restrict__ sa, int64_t M,
    const int8_t *__restrict__ SB, const double *__restrict__ sb, int64_t N, int kpad, int ntn,
    double alpha, const double *__restrict__ rs, double beta, double *__restrict__ out, int64_t sm,
    int64_t sn) {
    constexpr int TM = 16 * RI * WM, TN = 16 * RJ * WN, NWV = WM * WN;
    constexpr int LOADS = GM_NSL * (TM + TN) / 16;
    constexpr int LPW = (LOADS + NWV - 1) / NWV;   // wave loads per wave and stage
    // [buffer][A vectors (slice-major) then B vectors][slot]
    __shared__ __attribute__((aligned(16))) uint4 sbuf[2][GM_NSL * (TM + TN)][4];
    // tiles row by row (an XCD-grouped order measured the same)
    const int64_t bid = blockIdx.x;
    const int tm = (int)(bid / ntn), tn = (int)(bid - (int64_t)(bid / ntn) * ntn);
    const int64_t M0 = (int64_t)tm * TM, N0 = (int64_t)tn * TN;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int wi = wave / WN, wj = wave % WN;   // (16 RI)-row block, (16 RJ)-col block
    const int lcol = lane >> 2;
    const int lpiece = (lane & 3) ^ ((lane >> 4) & 3);
    const int nst = kpad / 64;

    const int8_t *gsrc[LPW];
    int64_t goff[LPW];
#pragma unroll
    for (int t = 0; t < LPW; ++t) {
        const int id = LPW * wave + t;
        gsrc[t] = SA;
        goff[t] = 0;
        if (id < LOADS) {
            const int row = 16 * id + lcol;   // LDS vector row of this lane
            if (row < GM_NSL * TM) {
                const int s = row / TM, cl = row - s * TM;
                int64_t m = M0 + cl;
                m = m < M ? m : M - 1;
                goff[t] = ((int64_t)s * M + m) * kpad + 16 * lpiece;
            } else {
                const int r2 = row - GM_NSL * TM;
                const int s = r2 / TN, cl = r2 - s * TN;
                int64_t n = N0 + cl;
                n = n < N ? n : N - 1;
                gsrc[t] = SB;
                goff[t] = ((int64_t)s * N + n) * kpad + 16 * lpiece;
            }
        }
    }
    auto issue = [&](int buf, int k0) {
#pragma unroll
        for (int t = 0; t < LPW; ++t) {
            const int id = LPW * wave + t;
            if (id < LOADS)
                __builtin_amdgcn_global_load_lds((const void *)(gsrc[t] + goff[t] + k0),
                                                 (lds_void_t *)&sbuf[buf][16 * id][0], 16, 0, 0);
        }
    };
    auto fragA = [&](int buf, int s, int row) {
        const uint4 v = sbuf[buf][s * TM + row][swzg(row & 15, lane >> 4)];
        return __builtin_bit_cast(v4i_t, v);
    };
    auto fragB = [&](int buf, int s, int row) {
        const uint4 v = sbuf[buf][GM_NSL * TM + s * TN + row][swzg(row & 15, lane >> 4)];
        return __builtin_bit_cast(v4i_t, v);
    };
    v4i_t L[GM_NSL][RI][RJ];
#pragma unroll
    for (int l = 0; l < GM_NSL; ++l)
#pragma unroll
        for (int a = 0; a < RI; ++a)
#pragma unroll
            for (int c = 0; c < RJ; ++c) L[l][a][c] = (v4i_t){0, 0, 0, 0};
    issue(0, 0);
    for (int st = 0; st < nst; ++st) {
        const int buf = st & 1;
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();   // stage st in sbuf[buf]; everyone is done with sbuf[buf^1]
        if (st + 1 < nst) issue(buf ^ 1, (st + 1) * 64);
        v4i_t bq[GM_NSL][RJ];
#pragma unroll
        for (int s2 = 0; s2 < GM_NSL; ++s2)
#pragma unroll
            for (int c = 0; c < RJ; ++c)
                bq[s2][c] = fragB(buf, s2, 16 * RJ * wj + 16 * c + (lane & 15));
#pragma unroll
        for (int s1 = 0; s1 < GM_NSL; ++s1) {
            v4i_t a[RI];
#pragma unroll
            for (int r = 0; r < RI; ++r) a[r] = fragA(buf, s1, 16 * RI * wi + 16 * r + (lane & 15));
#pragma unroll
            for (int s2 = 0; s2 < GM_NSL; ++s2) {
                if (s1 + s2 < GM_NSL) {
#pragma unroll
                    for (int r = 0; r < RI; ++r)
#pragma unroll
                        for (int c = 0; c < RJ; ++c)
                            L[s1 + s2][r][c] = __builtin_amdgcn_mfma_i32_16x16x64_i8(
                                a[r], bq[s2][c], L[s1 + s2][r][c], 0, 0, 0);
                }
            }
        }
    }
    const double c1 = 1.0 / 128.0;
#pragma unroll
    for (int c = 0; c < RJ; ++c) {
        const int64_t n = N0 + 16 * RJ * wj + 16 * c + (lane & 15);
        if (n >= N) continue;
        const double sn_ = sb[n];
#pragma unroll
        for (int r = 0; r < RI; ++r)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int64_t m = M0 + 16 * RI * wi + 16 * r + 4 * (lane >> 4) + q;
                if (m >= M) continue;
                // Horner in 2^-7 from the smallest level up
                double v = (double)L[GM_NSL - 1][r][c][q];
#pragma unroll
                for (int l = GM_NSL - 2; l >= 0; --l) v = fma(v, c1, (double)L[l][r][c][q]);
                const double f = rs ? alpha * rs[m] : alpha;
                double *o = out + m * sm + n * sn;
                const double val = f * (v * (sa[m] * sn_));
                *o = beta != 0.0 ? val + beta * *o : val;
            }
    }
}

int launch_gemm_i8(brpa_ctx *ctx, const int8_t *SA, const double *sa, int64_t M, const int8_t *SB,
                   const double *sb, int64_t N, int64_t kpad, double alpha, const double *rs,
                   double beta, double *out, int64_t sm, int64_t sn, int64_t kreal) {
    if (M <= 0 || N <= 0) return BRPA_OK;
    if (kpad <= 0 || kpad % 64 != 0 || kpad > I8_KMAX + 64)
        return brpa_fail(ctx, BRPA_ERR_ARG, "launch_gemm_i8: bad depth %lld", (long long)kpad);
    // 128 x 128 tiles (8 waves of 64 x 32, the whole 160 KB of LDS double-
    // buffered) when both sides fill them, else 128 x 64 (8 waves of 32 x 32)
    // an event pair around the launch (collected by kstat_collect)
    size_t slot = 0;
    BRPA_TRY(kstat_begin(ctx, &slot));
    const bool big = N >= 128 && M >= 128;
    const int64_t TM = 128, TN = big ? 128 : 64;
    const int64_t ntm = (M + TM - 1) / TM, ntn = (N + TN - 1) / TN;
    if (ntm * ntn > INT32_MAX || ntn > INT32_MAX)
        return brpa_fail(ctx, BRPA_ERR_ARG, "launch_gemm_i8: too many tiles");
    if (big)
        hipLaunchKernelGGL((k_gemm_i8<2, 4, 4, 2>), dim3((unsigned)(ntm * ntn)), dim3(512), 0,
                           ctx->stream, SA, sa, M, SB, sb, N, (int)kpad, (int)ntn, alpha, rs, beta,
                           out, sm, sn);
    else
        hipLaunchKernelGGL((k_gemm_i8<4, 2, 2, 2>), dim3((unsigned)(ntm * ntn)), dim3(512), 0,
                           ctx->stream, SA, sa, M, SB, sb, N, (int)kpad, (int)ntn, alpha, rs, beta,
                           out, sm, sn);
    LAUNCH_CHECK(ctx);
    return kstat_end(ctx, slot, BRPA_KERNEL_GEMM_I8,
                     2.0 * (double)M * (double)N * (double)(kreal > 0 ? kreal : kpad));
}
