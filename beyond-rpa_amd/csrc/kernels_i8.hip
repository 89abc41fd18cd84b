// Split-int8 Pi(u) build on gfx950 int8 MFMA (v_mfma_i32_16x16x64_i8).
//
// X_w = diag(sqrt(c_w)) R.  Column P of X_w gets one scale over all K rows,
// s_P = max_k |X_w[k][P]| / 127, and every element is written as
//   X = s_P (n0 + n1 2^-7 + n2 2^-14 + n3 2^-21),
// |n0| <= 127, |n1|, |n2|, |n3| <= 64 (each remainder is at most 1/2 before
// the 2^7 step), all four stored as int8.  The truncation is below 2^-28
// relative to the column maximum.  Pi_w[p][q] = s_p s_q sum_k x_p x_q and the
// ten slice products of level a + b <= 3 go to four int32 accumulators
//   L0 = n0 n0,  L1 = n0 n1 + n1 n0,  L2 = n0 n2 + n1 n1 + n2 n0,
//   L3 = n0 n3 + n1 n2 + n2 n1 + n3 n0,
// which are exact over the whole K: |L3| <= K (2*127*64 + 2*64*64) < 2^31 for
// K <= 87000 (the library caps K at 65536).  One FP64 combination per element,
//   Pi = s_p s_q (L0 + L1 2^-7 + L2 2^-14 + L3 2^-21),
// ends the product; the dropped level >= 4 products are ~2^-28 relative.
// The diagonal (and so tr Pi) is replaced by its exact double-double value
// from the column pass.
//
//   k_colscale_i8  s_w[P] and the exact diagonal, one pass over R for all w
//   k_split_i8     int8 slices of X_w, one read of R for all w
//   k_syrk_i8      128x128 output tiles, 8 waves of 64x32, int8 MFMA
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

// ---------------------------------------------------------------------------
// Column pass.  Work-groups per fragment: ceil(N/64).  Thread tid handles
// column P0 + tid%64 and rows tid/64 + 4 i; frequencies in groups of 16.
// For each (w, P): scale8 = max_k |sqrt(c_w(k)) R[k][P]| / 127 and
// diag = sum_k c_w(k) R[k][P]^2 in double-double (rpa_CC.f90:444-447 sums the
// trace in REAL(F128)).
constexpr int CS_WG = 16;   // frequencies per register group

__global__ __launch_bounds__(256) void k_colscale_i8(const FragDev *__restrict__ fr,
                                                    const int64_t *__restrict__ pre, int nfrag) {
    constexpr int CH = 64;   // rows per table chunk staged in LDS
    __shared__ int64_t s_pre[FIND_LDS_MAX + 1];
    __shared__ double sSD[CS_WG][CH], sCD[CS_WG][CH];
    __shared__ double sM[4][64], sH[4][64], sL[4][64];
    const int64_t b = blockIdx.x;
    const int f = find_frag_lds(pre, nfrag, b, s_pre);
    const FragDev &F = fr[f];
    const int N = F.N, K = F.K;
    const int P0 = (int)(b - (nfrag <= FIND_LDS_MAX ? s_pre[f] : pre[f])) * 64;
    const int tid = threadIdx.x, c = tid & 63, g = tid >> 6;
    const int P = P0 + c;
    const bool colok = P < N;
    const gdbl_t *R = (const gdbl_t *)F.R;
    const gdbl_t *SD = (const gdbl_t *)F.sdtab;
    const gdbl_t *CD = (const gdbl_t *)F.dtab;
    for (int w0 = 0; w0 < F.F; w0 += CS_WG) {
        const int nw = min(CS_WG, F.F - w0);
        double m[CS_WG];
        dd_t ds[CS_WG];
#pragma unroll
        for (int i = 0; i < CS_WG; ++i) {
            m[i] = 0.0;
            ds[i] = (dd_t){0.0, 0.0};
        }
        for (int k0 = 0; k0 < K; k0 += CH) {
            // this chunk's rows of the two tables (zeros past K or past nw)
            __syncthreads();
            for (int e = tid; e < CS_WG * CH; e += 256) {
                const int i = e / CH, r = e - i * CH;
                const bool ok = i < nw && k0 + r < K;
                const int64_t idx = ok ? (int64_t)(w0 + i) * K + k0 + r : 0;
                const double sd = SD[idx], cd = CD[idx];
                sSD[i][r] = ok ? sd : 0.0;
                sCD[i][r] = ok ? cd : 0.0;
            }
            // this thread's 16 rows of the chunk (rows g + 4 j), loads first
            double rv[CH / 4];
#pragma unroll
            for (int j = 0; j < CH / 4; ++j) {
                const int k = k0 + g + 4 * j;
                const bool ok = colok && k < K;
                rv[j] = R[ok ? (int64_t)k * N + P : 0];
                rv[j] = ok ? rv[j] : 0.0;
            }
            __syncthreads();
#pragma unroll 2
            for (int j = 0; j < CH / 4; ++j) {
                const int r = g + 4 * j;
                const double x = rv[j];
                const double r2 = x * x, r2e = fma(x, x, -r2);
#pragma unroll
                for (int i = 0; i < CS_WG; ++i) {
                    const double y = x * sSD[i][r];
                    m[i] = fmax(m[i], fabs(y));
                    const double cw = sCD[i][r];
                    const double p = cw * r2, pe = fma(cw, r2, -p) + cw * r2e;
                    ds[i] = dd_add(ds[i], p);
                    ds[i].lo += pe;
                }
            }
        }
        // combine the 4 row groups, one frequency at a time
        for (int i = 0; i < nw; ++i) {
            double mi = m[0], hi = ds[0].hi, lo = ds[0].lo;
#pragma unroll
            for (int q = 1; q < CS_WG; ++q)
                if (q == i) {
                    mi = m[q];
                    hi = ds[q].hi;
                    lo = ds[q].lo;
                }
            sM[g][c] = mi;
            sH[g][c] = hi;
            sL[g][c] = lo;
            __syncthreads();
            if (tid < 64 && colok) {
                const double mm = fmax(fmax(sM[0][c], sM[1][c]), fmax(sM[2][c], sM[3][c]));
                dd_t t = {sH[0][c], sL[0][c]};
                for (int q = 1; q < 4; ++q) t = dd_add(t, (dd_t){sH[q][c], sL[q][c]});
                const int64_t o = (int64_t)(w0 + i) * N + P;
                F.scale8[o] = mm / 127.0;
                F.diag[o] = t.hi + t.lo;
            }
            __syncthreads();
        }
    }
}

int launch_colscale_i8(brpa_ctx *ctx, const FragDev *fr, const int64_t *pre, int nfrag,
                       int64_t total) {
    if (total <= 0) return BRPA_OK;
    hipLaunchKernelGGL(k_colscale_i8, dim3((unsigned)total), dim3(256), 0, ctx->stream, fr, pre,
                       nfrag);
    LAUNCH_CHECK(ctx);
    return BRPA_OK;
}

// ---------------------------------------------------------------------------
// Split pass.  Work-groups per fragment: (Kpad8 / 128) * ceil(N/64).  A
// 128-row x 64-column tile of R is staged in LDS once and sliced for every
// frequency.  Thread -> column tid/4, rows 32 (tid%4) .. +32, so each
// (w, slice, column) row segment of 128 bytes is written by 4 neighbouring
// threads.  Rows k >= K and columns past N write zeros / nothing.
__global__ __launch_bounds__(256) void k_split_i8(const FragDev *__restrict__ fr,
                                                 const int64_t *__restrict__ pre, int nfrag) {
    __shared__ int64_t s_pre[FIND_LDS_MAX + 1];
    __shared__ double sSD[CS_WG][I8_KB];   // sqrt(c_w) of this work-group's rows
    const int64_t b = blockIdx.x;
    const int f = find_frag_lds(pre, nfrag, b, s_pre);
    const FragDev &F = fr[f];
    const int N = F.N, K = F.K;
    const int ncc = (N + 63) / 64;
    const int64_t lb = b - (nfrag <= FIND_LDS_MAX ? s_pre[f] : pre[f]);
    const int kc = (int)(lb / ncc), cc = (int)(lb - (int64_t)kc * ncc);
    const int k0 = kc * I8_KB, P0 = cc * 64;
    const int tid = threadIdx.x;
    const gdbl_t *R = (const gdbl_t *)F.R;
    const gdbl_t *SD = (const gdbl_t *)F.sdtab;
    // thread -> column P0 + tid/4 and its 32 rows 32 (tid%4) .. +32, held in registers
    const int c = tid >> 2, r0 = (tid & 3) * 32;
    const int P = P0 + c;
    const bool colok = P < N;
    double x[32];
#pragma unroll
    for (int j = 0; j < 32; ++j) {
        const int k = k0 + r0 + j;
        const bool ok = colok && k < K;
        x[j] = R[ok ? (int64_t)k * N + P : 0];
        x[j] = ok ? x[j] : 0.0;
    }
    const int64_t Kp = F.Kpad8;
    for (int w0 = 0; w0 < F.F; w0 += CS_WG) {
        const int nw = min(CS_WG, F.F - w0);
        __syncthreads();
        for (int e = tid; e < CS_WG * I8_KB; e += 256) {
            const int i = e / I8_KB, r = e - i * I8_KB;
            const bool ok = i < nw && k0 + r < K;
            const double v = SD[ok ? (int64_t)(w0 + i) * K + k0 + r : 0];
            sSD[i][r] = ok ? v : 0.0;
        }
        __syncthreads();
        if (!colok) continue;
        for (int i = 0; i < nw; ++i) {
            const int w = w0 + i;
            const double sc = F.scale8[(int64_t)w * N + P];
            const double inv = sc > 0.0 ? 1.0 / sc : 0.0;
            uint32_t pk[NSLICE][8];
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                uint32_t word[NSLICE] = {0u, 0u, 0u, 0u};
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int j = 4 * q + u;
                    double t = x[j] * sSD[i][r0 + j] * inv;   // |t| <= 127 (up to rounding)
#pragma unroll
                    for (int s = 0; s < NSLICE; ++s) {
                        const double n = rint(t);
                        t = (t - n) * 128.0;   // exact
                        word[s] |= ((uint32_t)(uint8_t)(int8_t)(int)n) << (8 * u);
                    }
                }
#pragma unroll
                for (int s = 0; s < NSLICE; ++s) pk[s][q] = word[s];
            }
#pragma unroll
            for (int s = 0; s < NSLICE; ++s) {
                int8_t *dst = F.sl8 + (((int64_t)w * NSLICE + s) * N + P) * Kp + k0 + r0;
                uint4 *d4 = reinterpret_cast<uint4 *>(dst);
                d4[0] = make_uint4(pk[s][0], pk[s][1], pk[s][2], pk[s][3]);
                d4[1] = make_uint4(pk[s][4], pk[s][5], pk[s][6], pk[s][7]);
            }
        }
    }
}

int launch_split_i8(brpa_ctx *ctx, const FragDev *fr, const int64_t *pre, int nfrag,
                    int64_t total) {
    if (total <= 0) return BRPA_OK;
    hipLaunchKernelGGL(k_split_i8, dim3((unsigned)total), dim3(256), 0, ctx->stream, fr, pre,
                       nfrag);
    LAUNCH_CHECK(ctx);
    return BRPA_OK;
}

// ---------------------------------------------------------------------------
// Product.  One TxT output tile pair (p-tile >= q-tile) of one frequency per
// work-group; work-groups per fragment: F * npairs(ceil(N/T)).  The waves
// own 64x32 sub-tiles: T = 128 -> 8 waves (2 x 4), T = 64 -> 2 waves (1 x 2).
//
// v_mfma_i32_16x16x64_i8 lane layout (lane l, 16 bytes per operand):
//   A[m = l%16][k = 16*(l/16) .. +15], B[k = 16*(l/16) .. +15][n = l%16],
//   D: 4 ints, D[r] = C[m = 4*(l/16) + r][n = l%16].
// Both operands are columns of X (k contiguous per column, 64 bytes = one
// MFMA k-depth per stage and slice), read from LDS the same way.
//
// Stages go straight from HBM to LDS (global_load_lds_dwordx4, 16 bytes per
// lane), double-buffered.  One wave load fills 1 KB = 16 columns x 64 bytes;
// lane l lands in column l/4, slot l%4, and fetches k-piece
// (l%4) ^ ((l/16)%4) of that column, so that the MFMA operand reads (16
// columns at one k-piece per 16 lanes) hit 16 distinct 4-bank groups.
constexpr int KS8 = 64;   // k-values (bytes) per stage and slice

typedef __attribute__((address_space(3))) void lds_void_t;

__device__ __forceinline__ int swz8(int col_local, int piece) {
    return piece ^ ((col_local >> 2) & 3);
}

template <int T>
__global__ __launch_bounds__((T / 64) * (T / 32) * 64) void k_syrk_i8(
    const FragDev *__restrict__ fr, const int64_t *__restrict__ pre, int nfrag, int xcd) {
    constexpr int WJN = T / 32;                 // waves along columns
    constexpr int NW = (T / 64) * WJN;          // waves
    constexpr int GPS = T / 16;                 // 16-column groups per side and slice
    constexpr int LPS = NSLICE * GPS;           // wave loads per side and stage
    constexpr int LPW = 2 * LPS / NW;           // wave loads per wave and stage
    // [buffer][side][slice][column][slot] of 16-byte pieces
    __shared__ __attribute__((aligned(16))) uint4 sbuf[2][2][NSLICE][T][4];
    __shared__ int64_t s_pre[FIND_LDS_MAX + 1];
    // xcd = 1: one contiguous range of tiles per XCD (see xcd_remap)
    const int64_t b = xcd ? xcd_remap(blockIdx.x, gridDim.x) : (int64_t)blockIdx.x;
    const int f = find_frag_lds(pre, nfrag, b, s_pre);
    const FragDev &F = fr[f];
    const int N = F.N;
    const int nt = (N + T - 1) / T;
    const int npairs = nt * (nt + 1) / 2;
    const int64_t lb = b - (nfrag <= FIND_LDS_MAX ? s_pre[f] : pre[f]);
    const int w = (int)(lb / npairs), pair = (int)(lb - (int64_t)w * npairs);
    int tp, tq;
    tri_decode(pair, tp, tq);
    const int p0 = tp * T, q0 = tq * T;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int wi = wave / WJN, wj = wave % WJN;   // 64-row block, 32-col block

    const int64_t Kp = F.Kpad8;
    const int8_t *slw = F.sl8 + (int64_t)w * NSLICE * N * Kp;

    v4i_t L0[4][2], L1[4][2], L2[4][2], L3[4][2];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) L0[i][j] = L1[i][j] = L2[i][j] = L3[i][j] = (v4i_t){0, 0, 0, 0};

    // This wave's LPW wave loads of a stage: load id = LPW*wave + t -> side =
    // id/LPS, slice = (id%LPS)/GPS, column group = id%GPS.  Columns past N are
    // clamped to N-1: they only feed outputs that are never written.
    const int lcol = lane >> 2;
    const int lpiece = (lane & 3) ^ ((lane >> 4) & 3);
    int64_t goff[LPW];
#pragma unroll
    for (int t = 0; t < LPW; ++t) {
        const int id = LPW * wave + t;
        const int side = id / LPS, sl = (id % LPS) / GPS, g = id % GPS;
        int P = (side ? q0 : p0) + 16 * g + lcol;
        P = P < N ? P : N - 1;
        goff[t] = ((int64_t)sl * N + P) * Kp + 16 * lpiece;
    }
    auto issue = [&](int buf, int k0) {
#pragma unroll
        for (int t = 0; t < LPW; ++t) {
            const int id = LPW * wave + t;
            const int side = id / LPS, sl = (id % LPS) / GPS, g = id % GPS;
            __builtin_amdgcn_global_load_lds((const void *)(slw + goff[t] + k0),
                                             (lds_void_t *)&sbuf[buf][side][sl][16 * g][0], 16, 0, 0);
        }
    };
    auto frag = [&](int buf, int side, int s, int row) {
        const int cl = row & 15;
        const uint4 v = sbuf[buf][side][s][row][swz8(cl, lane >> 4)];
        return __builtin_bit_cast(v4i_t, v);
    };
#define MFMA8(a, b, c) c = __builtin_amdgcn_mfma_i32_16x16x64_i8(a, b, c, 0, 0, 0)

    const int nst = (int)(Kp / KS8);
    issue(0, 0);
    for (int st = 0; st < nst; ++st) {
        const int buf = st & 1;
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();   // stage st in sbuf[buf]; everyone is done with sbuf[buf^1]
        if (st + 1 < nst) issue(buf ^ 1, (st + 1) * KS8);
        v4i_t bq[NSLICE][2];
#pragma unroll
        for (int s2 = 0; s2 < NSLICE; ++s2)
#pragma unroll
            for (int j = 0; j < 2; ++j) bq[s2][j] = frag(buf, 1, s2, 32 * wj + 16 * j + (lane & 15));
#pragma unroll
        for (int ih = 0; ih < 4; ih += 2) {
            v4i_t a[2];
#pragma unroll
            for (int ii = 0; ii < 2; ++ii) a[ii] = frag(buf, 0, 0, 64 * wi + 16 * (ih + ii) + (lane & 15));
#pragma unroll
            for (int ii = 0; ii < 2; ++ii)
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    MFMA8(a[ii], bq[0][j], L0[ih + ii][j]);
                    MFMA8(a[ii], bq[1][j], L1[ih + ii][j]);
                    MFMA8(a[ii], bq[2][j], L2[ih + ii][j]);
                    MFMA8(a[ii], bq[3][j], L3[ih + ii][j]);
                }
#pragma unroll
            for (int ii = 0; ii < 2; ++ii) a[ii] = frag(buf, 0, 1, 64 * wi + 16 * (ih + ii) + (lane & 15));
#pragma unroll
            for (int ii = 0; ii < 2; ++ii)
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    MFMA8(a[ii], bq[0][j], L1[ih + ii][j]);
                    MFMA8(a[ii], bq[1][j], L2[ih + ii][j]);
                    MFMA8(a[ii], bq[2][j], L3[ih + ii][j]);
                }
#pragma unroll
            for (int ii = 0; ii < 2; ++ii) a[ii] = frag(buf, 0, 2, 64 * wi + 16 * (ih + ii) + (lane & 15));
#pragma unroll
            for (int ii = 0; ii < 2; ++ii)
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    MFMA8(a[ii], bq[0][j], L2[ih + ii][j]);
                    MFMA8(a[ii], bq[1][j], L3[ih + ii][j]);
                }
#pragma unroll
            for (int ii = 0; ii < 2; ++ii) a[ii] = frag(buf, 0, 3, 64 * wi + 16 * (ih + ii) + (lane & 15));
#pragma unroll
            for (int ii = 0; ii < 2; ++ii)
#pragma unroll
                for (int j = 0; j < 2; ++j) MFMA8(a[ii], bq[0][j], L3[ih + ii][j]);
        }
    }
#undef MFMA8

    const double *scw = F.scale8 + (int64_t)w * N;
    double *C = F.Pi + (int64_t)w * N * N;
    const bool diag = (tp == tq);
    const double c1 = 1.0 / 128.0, c2 = c1 * c1, c3 = c2 * c1;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int q = q0 + 32 * wj + 16 * j + (lane & 15);
            const double sq = q < N ? scw[q] : 0.0;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int p = p0 + 64 * wi + 16 * i + 4 * (lane >> 4) + r;
                if (p < N && q < N) {
                    const double v = (double)L0[i][j][r] + c1 * (double)L1[i][j][r] +
                                     c2 * (double)L2[i][j][r] + c3 * (double)L3[i][j][r];
                    const double out = v * (scw[p] * sq);
                    C[(int64_t)p * N + q] = out;
                    if (!diag) C[(int64_t)q * N + p] = out;
                }
            }
        }
}

int syrk_i8_tile(int variant) { return variant == I8V_T64 ? 64 : 128; }

int launch_syrk_i8(brpa_ctx *ctx, const FragDev *fr, const int64_t *pre, int nfrag, int64_t total,
                   int variant) {
    if (total <= 0) return BRPA_OK;
    switch (variant) {
    case I8V_T128:
        hipLaunchKernelGGL(k_syrk_i8<128>, dim3((unsigned)total), dim3(512), 0, ctx->stream, fr,
                           pre, nfrag, 0);
        break;
    case I8V_T128_XCD:
        hipLaunchKernelGGL(k_syrk_i8<128>, dim3((unsigned)total), dim3(512), 0, ctx->stream, fr,
                           pre, nfrag, 1);
        break;
    case I8V_T64:
        hipLaunchKernelGGL(k_syrk_i8<64>, dim3((unsigned)total), dim3(128), 0, ctx->stream, fr,
                           pre, nfrag, 0);
        break;
    default:
        return brpa_fail(ctx, BRPA_ERR_ARG, "unknown split-int8 kernel variant %d", variant);
    }
    LAUNCH_CHECK(ctx);
    return BRPA_OK;
}
