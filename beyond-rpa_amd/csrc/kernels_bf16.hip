// Split-BF16 Pi(u) build on gfx950 bf16 MFMA (v_mfma_f32_16x16x32_bf16).
//
// X_w = diag(sqrt(c_w)) R is cut into 128-row blocks.  In block b, column P
// gets a power-of-two scale s = 2^(e-8) with max_k |X[k][P]| < 2^e, and every
// element is written as s (n0 + n1 2^-8 + n2 2^-16 + n3 2^-24) with integers
// |n0| <= 256, |n1|,|n2|,|n3| <= 128.  Slice k is stored as the bf16 value
// n_k 2^(-8k), which is exact.  Pi_w[p][q] = sum_b s_p s_q sum_k x_p x_q, and
// the ten slice products of level a+b <= 3 go to three FP32 accumulators:
//   L0  = n0 n0                          (integer sums < 2^24: exact)
//   L1  = (n0 n1 + n1 n0) 2^-8           (integer sums < 2^24 times 2^-8: exact)
//   L23 = levels 2 and 3                 (rounding ~2^-40 relative to Pi)
// Each 128-row block is flushed to FP64 as s_p s_q (L0 + L1 + L23).  The
// dropped level >= 4 products are ~2^-32 relative.  The diagonal (and so
// tr Pi) is then replaced by its exact FP64 value (k_split partials, summed in
// k_set_diag), which removes the systematic diagonal bias of the truncation.
#include "device_common.h"
#include "internal.h"

#define LAUNCH_CHECK(ctx)                                                                      \
    do {                                                                                       \
        hipError_t _e = hipGetLastError();                                                     \
        if (_e != hipSuccess)                                                                  \
            return brpa_fail((ctx), BRPA_ERR_HIP, "kernel launch failed: %s (%s:%d)",          \
                             hipGetErrorString(_e), __FILE__, __LINE__);                       \
    } while (0)

// n 2^-8k with an integer |n| <= 256: exact in float and in bf16 (truncating
// the float's low 16 bits, which are zero, is the exact conversion).
__device__ __forceinline__ uint16_t int_to_bf16_bits(double n) {
    float f = (float)n;
    return (uint16_t)(__float_as_uint(f) >> 16);
}

// Work-groups per fragment: nblk * ceil(N/64).  One work-group reads a
// 128-row x 64-column block of R once and splits X_w = diag(sqrt(c_w)) R on
// that block for every frequency w in turn (R is read once, not once per
// frequency; the slices of all frequencies are written).
__global__ __launch_bounds__(256) void k_split(const FragDev *__restrict__ fr,
                                              const int64_t *__restrict__ pre, int nfrag) {
    __shared__ double sR[SPLIT_KB][65];
    __shared__ double sScw[SPLIT_KB];
    __shared__ double sMax[4][64];
    __shared__ int sExp[64];
    __shared__ double sDh[4][64], sDl[4][64];
    int64_t b = blockIdx.x;
    int f = find_frag(pre, nfrag, b);
    const FragDev &F = fr[f];
    const int N = F.N, K = F.K;
    const int ncc = (N + 63) / 64;
    const int64_t lb = b - pre[f];
    const int blk = (int)(lb / ncc), cc = (int)(lb - (int64_t)blk * ncc);
    const int k0 = blk * SPLIT_KB, P0 = cc * 64;
    const int tid = threadIdx.x;

    for (int e = tid; e < SPLIT_KB * 64; e += 256) {
        const int r = e >> 6, c = e & 63;
        const int k = k0 + r, P = P0 + c;
        sR[r][c] = (k < K && P < N) ? F.R[(int64_t)k * N + P] : 0.0;
    }
    // slice phase: thread -> column c = tid/4, rows (tid%4)*32 .. +32 (contiguous
    // in the output)
    const int cs = tid >> 2, rs0 = (tid & 3) * 32;
    const int Ps = P0 + cs;
    for (int w = 0; w < F.F; ++w) {
        if (tid < SPLIT_KB) {
            const int k = k0 + tid;
            sScw[tid] = k < K ? sqrt(F.dtab[(int64_t)w * K + k]) : 0.0;
        }
        __syncthreads();
        // thread tid handles column tid%64, rows tid/64 + 4 i: that column's
        // exact diagonal partial sum_k c_w(k) R[k][P]^2 in double-double
        dd_t dsum = {0.0, 0.0};
        for (int e = tid; e < SPLIT_KB * 64; e += 256) {
            const int r = e >> 6, c = e & 63;
            const int k = k0 + r, P = P0 + c;
            if (k < K && P < N) {
                const double cw = F.dtab[(int64_t)w * K + k];
                const double rv = sR[r][c];
                const double r2 = rv * rv, r2e = fma(rv, rv, -r2);
                const double p = cw * r2, pe = fma(cw, r2, -p) + cw * r2e;
                dsum = dd_add(dsum, p);
                dsum.lo += pe;
            }
        }
        sDh[tid >> 6][tid & 63] = dsum.hi;
        sDl[tid >> 6][tid & 63] = dsum.lo;
        {
            const int c = tid & 63, g = tid >> 6;
            double m = 0.0;
            for (int r = g * 32; r < g * 32 + 32; ++r) m = fmax(m, fabs(sScw[r] * sR[r][c]));
            sMax[g][c] = m;
        }
        __syncthreads();
        if (tid < 64) {
            if (P0 + tid < N) {
                dd_t t = {sDh[0][tid], sDl[0][tid]};
                for (int g = 1; g < 4; ++g) t = dd_add(t, (dd_t){sDh[g][tid], sDl[g][tid]});
                double *q = F.dpart + (((int64_t)w * F.nblk + blk) * N + P0 + tid) * 2;
                q[0] = t.hi;
                q[1] = t.lo;
            }
            const double m =
                fmax(fmax(sMax[0][tid], sMax[1][tid]), fmax(sMax[2][tid], sMax[3][tid]));
            int e = 0;
            if (m > 0.0) frexp(m, &e);   // m = f 2^e, 0.5 <= f < 1  ->  m < 2^e
            sExp[tid] = e;
            if (P0 + tid < N)
                F.scale[((int64_t)w * F.nblk + blk) * N + P0 + tid] =
                    (m > 0.0) ? ldexpf(1.0f, e - 8) : 0.0f;
        }
        __syncthreads();
        if (Ps < N) {
            const int e = sExp[cs];
            for (int r = 0; r < 32; r += 8) {
                uint16_t v[NSLICE][8];
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    const int rr = rs0 + r + i;
                    double t = ldexp(sScw[rr] * sR[rr][cs], 8 - e);   // |t| < 256, exact
#pragma unroll
                    for (int sl = 0; sl < NSLICE; ++sl) {
                        double n = rint(t);
                        t = (t - n) * 256.0;                      // exact
                        v[sl][i] = int_to_bf16_bits(ldexp(n, -8 * sl));
                    }
                }
#pragma unroll
                for (int sl = 0; sl < NSLICE; ++sl) {
                    uint4 pk;
                    pk.x = v[sl][0] | ((uint32_t)v[sl][1] << 16);
                    pk.y = v[sl][2] | ((uint32_t)v[sl][3] << 16);
                    pk.z = v[sl][4] | ((uint32_t)v[sl][5] << 16);
                    pk.w = v[sl][6] | ((uint32_t)v[sl][7] << 16);
                    uint16_t *dst =
                        F.sl + (((int64_t)w * NSLICE + sl) * N + Ps) * F.Kpad + k0 + rs0 + r;
                    *reinterpret_cast<uint4 *>(dst) = pk;
                }
            }
        }
        // sScw, sMax, sExp and sDh/sDl are rewritten for the next frequency
        __syncthreads();
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
// of 64x32 (2 waves per SIMD).
// MFMA v_mfma_f32_16x16x32_bf16 operand layout (lane l):
//   A[m = l%16][k = 8*(l/16) .. +7], B[k = 8*(l/16) .. +7][n = l%16],
//   D: 4 floats, D[r] = C[m = 4*(l/16) + r][n = l%16].
// Both operands are columns of X (k contiguous per column), so A and B are
// read from LDS the same way.
//
// Stages of 32 rows go straight from HBM to LDS (global_load_lds_dwordx4,
// 16 bytes per lane, no register staging), double-buffered.  One wave load
// fills 1 KB = 16 columns x 64 bytes; lane l lands in column l/4, slot l%4,
// and fetches k-piece (l%4) ^ ((l/16)%4) of that column, so that the MFMA
// operand reads (16 columns at one k-piece per 16 lanes) hit 16 distinct
// 4-bank groups.
constexpr int T16 = 128;
constexpr int KS = 32;          // rows per stage (one MFMA k-depth)

typedef __attribute__((address_space(3))) void lds_void_t;

__device__ __forceinline__ int swz(int col_local, int piece) {
    return piece ^ ((col_local >> 2) & 3);
}

__global__ __launch_bounds__(512) void k_syrk_bf16(const FragDev *__restrict__ fr,
                                                  const int64_t *__restrict__ pre, int nfrag) {
    // [buffer][side][slice][column][slot] of 16-byte pieces
    __shared__ __attribute__((aligned(16))) uint4 sbuf[2][2][NSLICE][T16][4];
    __shared__ float sSc[2][2 * T16];
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

    const uint16_t *slw = F.sl + (int64_t)w * NSLICE * N * F.Kpad;
    const float *scw = F.scale + (int64_t)w * F.nblk * N;
    const int64_t Kpad = F.Kpad;

    f4_t L0[4][2], L1[4][2], L23[4][2];
    double acc[4][2][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            L0[i][j] = L1[i][j] = L23[i][j] = (f4_t){0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int r = 0; r < 4; ++r) acc[i][j][r] = 0.0;
        }

    // This wave's 8 of the 64 wave loads of a stage: load id = 8*wave + t
    // -> side = id/32, slice = (id/8)%4, column group = id%8.  Per-lane global
    // offsets (columns past N are clamped to N-1: they only feed outputs that
    // are never written).
    const int lcol = lane >> 2;
    const int lpiece = (lane & 3) ^ ((lane >> 4) & 3);
    int64_t goff[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) {
        const int id = 8 * wave + t;
        const int side = id >> 5, sl = (id >> 3) & 3, g = id & 7;
        int P = (side ? q0 : p0) + 16 * g + lcol;
        P = P < N ? P : N - 1;
        goff[t] = ((int64_t)sl * N + P) * Kpad + 8 * lpiece;
    }
    auto issue = [&](int buf, int k0) {
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            const int id = 8 * wave + t;
            const int side = id >> 5, sl = (id >> 3) & 3, g = id & 7;
            __builtin_amdgcn_global_load_lds((const void *)(slw + goff[t] + k0),
                                             (lds_void_t *)&sbuf[buf][side][sl][16 * g][0], 16, 0, 0);
        }
    };
    auto frag = [&](int buf, int side, int s, int row) {
        const int cl = row & 15;
        const uint4 v = sbuf[buf][side][s][row][swz(cl, lane >> 4)];
        return __builtin_bit_cast(bf8_t, v);
    };
#define MFMA(a, b, c) c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0)

    const int nblk = F.nblk;
    const int ksteps = SPLIT_KB / KS;
    const int nst = nblk * ksteps;
    issue(0, 0);
    for (int st = 0; st < nst; ++st) {
        const int buf = st & 1, blk = st / ksteps, ks = st - blk * ksteps;
        if (ks == 0 && tid < 2 * T16) {
            int side = tid / T16, c = tid - side * T16;
            int P = (side ? q0 : p0) + c;
            sSc[blk & 1][tid] = (P < N) ? scw[(int64_t)blk * N + P] : 0.0f;
        }
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();   // stage st in sbuf[buf]; everyone is done with sbuf[buf^1]
        if (st + 1 < nst) issue(buf ^ 1, (st + 1) * KS);
        // the ten products of level <= 3, grouped by the A slice and by pairs
        // of 16-row blocks so that few fragments are live at once
        bf8_t bq[NSLICE][2];
#pragma unroll
        for (int s2 = 0; s2 < NSLICE; ++s2)
#pragma unroll
            for (int j = 0; j < 2; ++j) bq[s2][j] = frag(buf, 1, s2, 32 * wj + 16 * j + (lane & 15));
#pragma unroll
        for (int ih = 0; ih < 4; ih += 2) {
            bf8_t a[2];
#pragma unroll
            for (int ii = 0; ii < 2; ++ii) a[ii] = frag(buf, 0, 0, 64 * wi + 16 * (ih + ii) + (lane & 15));
#pragma unroll
            for (int ii = 0; ii < 2; ++ii)
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    MFMA(a[ii], bq[0][j], L0[ih + ii][j]);
                    MFMA(a[ii], bq[1][j], L1[ih + ii][j]);
                    MFMA(a[ii], bq[2][j], L23[ih + ii][j]);
                    MFMA(a[ii], bq[3][j], L23[ih + ii][j]);
                }
#pragma unroll
            for (int ii = 0; ii < 2; ++ii) a[ii] = frag(buf, 0, 1, 64 * wi + 16 * (ih + ii) + (lane & 15));
#pragma unroll
            for (int ii = 0; ii < 2; ++ii)
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    MFMA(a[ii], bq[0][j], L1[ih + ii][j]);
                    MFMA(a[ii], bq[1][j], L23[ih + ii][j]);
                    MFMA(a[ii], bq[2][j], L23[ih + ii][j]);
                }
#pragma unroll
            for (int ii = 0; ii < 2; ++ii) a[ii] = frag(buf, 0, 2, 64 * wi + 16 * (ih + ii) + (lane & 15));
#pragma unroll
            for (int ii = 0; ii < 2; ++ii)
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    MFMA(a[ii], bq[0][j], L23[ih + ii][j]);
                    MFMA(a[ii], bq[1][j], L23[ih + ii][j]);
                }
#pragma unroll
            for (int ii = 0; ii < 2; ++ii) a[ii] = frag(buf, 0, 3, 64 * wi + 16 * (ih + ii) + (lane & 15));
#pragma unroll
            for (int ii = 0; ii < 2; ++ii)
#pragma unroll
                for (int j = 0; j < 2; ++j) MFMA(a[ii], bq[0][j], L23[ih + ii][j]);
        }
        if (ks == ksteps - 1) {
            // flush this block's level sums to FP64
            const float *sc = sSc[blk & 1];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    const double sq = (double)sc[T16 + 32 * wj + 16 * j + (lane & 15)];
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const double sp = (double)sc[64 * wi + 16 * i + 4 * (lane >> 4) + r];
                        double v = (double)L0[i][j][r] + (double)L1[i][j][r] + (double)L23[i][j][r];
                        acc[i][j][r] = fma(v, sp * sq, acc[i][j][r]);
                    }
                    L0[i][j] = L1[i][j] = L23[i][j] = (f4_t){0.f, 0.f, 0.f, 0.f};
                }
        }
    }
#undef MFMA

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
