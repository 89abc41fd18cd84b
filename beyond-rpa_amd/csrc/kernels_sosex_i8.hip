// Split-int8 SOSEX contraction on gfx950 int8 MFMA (v_mfma_i32_16x16x64_i8).
//
//   E_x = -alpha sum_{j<=i} (2 - delta_ij) sum_{P,mu} W_ji(P,mu) W_ij(P,mu) B_mu,
//   W_ij(P,mu) = sum_a R(P,a,i) V(mu,a,j)      (rpa/rpa_CC_Exchange.f90:36-116)
//
// Operands.  For every occupied index b the v x N block R_b (rows a, columns P)
// and the v x NT2 block V_b are cut per column over the v rows: column scale
// s = max_a |x| / 127 and five int8 slices, x = s (n0 + n1 2^-7 + ... + n4 2^-28),
// |n0| <= 127, |n1..n4| <= 64.  W_ij is then the sum of the 15 slice products
// of level <= 4, accumulated exactly in five int32 level sums (|L4| <= v *
// 28544 < 2^31 for v <= 65536) and combined once in FP64.  The dropped level
// >= 5 products are ~2^-35 relative.
//
// Fusion.  One work-group computes the 128 x 64 tile of W_ij and of W_ji for one
// pair (i >= j), multiplies them elementwise with B_mu and reduces the tile to
// one double-double partial, pre-multiplied by (2 - delta_ij).  W never goes to
// HBM (the FP64 pair loop writes and rereads 2 N NT2 doubles per pair).
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

constexpr int SX_NSL = 5;    // slices per value
constexpr int SX_KB = 128;   // rows (a) per split work-group; vpad is a multiple

// ---------------------------------------------------------------------------
// Column scales of the blocks: scale[b][c] = max_{a<v} |X[b*v + a][c]| / 127,
// the rows optionally scaled first (X[r][c] * rs[r] when rs is not null).
// Work-groups: nblk * ceil(ncols/64); thread -> column tid%64, rows tid/64 + 4 i.
__global__ __launch_bounds__(256) void k_blkscale_i8(const double *__restrict__ X, int64_t ncols,
                                                    int v, int ncc, const double *__restrict__ rs,
                                                    double *__restrict__ scale) {
    __shared__ double sM[4][64];
    const int b = blockIdx.x / ncc, cc = blockIdx.x - b * ncc;
    const int tid = threadIdx.x, c = tid & 63, g = tid >> 6;
    const int64_t col = (int64_t)cc * 64 + c;
    const bool ok = col < ncols;
    const gdbl_t *G = (const gdbl_t *)X;
    double m = 0.0;
    for (int a0 = g; a0 < v; a0 += 16) {
        double x[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int a = a0 + 4 * u;
            const bool in = ok && a < v;
            x[u] = G[in ? ((int64_t)b * v + a) * ncols + col : 0];
            if (rs) x[u] *= rs[in ? (int64_t)b * v + a : 0];
            x[u] = in ? fabs(x[u]) : 0.0;
        }
        m = fmax(m, fmax(fmax(x[0], x[1]), fmax(x[2], x[3])));
    }
    sM[g][c] = m;
    __syncthreads();
    if (tid < 64 && ok)
        scale[(int64_t)b * ncols + col] =
            fmax(fmax(sM[0][c], sM[1][c]), fmax(sM[2][c], sM[3][c])) / 127.0;
}

// Slices: out[((b*5 + s)*ncols + c)*vpad + a], zero for a >= v.
// Work-groups: nblk * (vpad/128) * ceil(ncols/64); a 128 x 64 tile of the block
// is staged in LDS; thread -> column tid/4, rows 32 (tid%4) .. +32.
__global__ __launch_bounds__(256) void k_blksplit_i8(const double *__restrict__ X, int64_t ncols,
                                                    int v, int vpad, int ncc,
                                                    const double *__restrict__ rs,
                                                    const double *__restrict__ scale,
                                                    int8_t *__restrict__ out) {
    __shared__ double sX[SX_KB][65];
    const int nkc = vpad / SX_KB;
    const int b = blockIdx.x / (nkc * ncc);
    const int rem = blockIdx.x - b * nkc * ncc;
    const int kc = rem / ncc, cc = rem - kc * ncc;
    const int a0 = kc * SX_KB;
    const int64_t c0 = (int64_t)cc * 64;
    const int tid = threadIdx.x;
    const gdbl_t *G = (const gdbl_t *)X;
    {
        constexpr int PER = SX_KB * 64 / 256;
        const int c = tid & 63, r0 = tid >> 6;
        double x[PER];
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int r = r0 + 4 * u;
            const bool in = (a0 + r < v) && (c0 + c < ncols);
            x[u] = G[in ? ((int64_t)b * v + a0 + r) * ncols + c0 + c : 0];
            if (rs) x[u] *= rs[in ? (int64_t)b * v + a0 + r : 0];   // the product of the scale pass
            x[u] = in ? x[u] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < PER; ++u) sX[r0 + 4 * u][c] = x[u];
    }
    __syncthreads();
    const int c = tid >> 2, r0 = (tid & 3) * 32;
    const int64_t col = c0 + c;
    if (col >= ncols) return;
    const double sc = scale[(int64_t)b * ncols + col];
    const double inv = sc > 0.0 ? 1.0 / sc : 0.0;
    uint32_t pk[SX_NSL][8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        uint32_t word[SX_NSL] = {0u, 0u, 0u, 0u, 0u};
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            double t = sX[r0 + 4 * q + i][c] * inv;   // |t| <= 127 (up to one rounding)
#pragma unroll
            for (int s = 0; s < SX_NSL; ++s) {
                const double n = rint(t);
                t = (t - n) * 128.0;   // exact
                word[s] |= ((uint32_t)(uint8_t)(int8_t)(int)n) << (8 * i);
            }
        }
#pragma unroll
        for (int s = 0; s < SX_NSL; ++s) pk[s][q] = word[s];
    }
#pragma unroll
    for (int s = 0; s < SX_NSL; ++s) {
        uint4 *d4 = reinterpret_cast<uint4 *>(out + (((int64_t)b * SX_NSL + s) * ncols + col) * vpad +
                                              a0 + r0);
        d4[0] = make_uint4(pk[s][0], pk[s][1], pk[s][2], pk[s][3]);
        d4[1] = make_uint4(pk[s][4], pk[s][5], pk[s][6], pk[s][7]);
    }
}

// ---------------------------------------------------------------------------
// Fused pair kernel.  Work-groups: npairs * ceil(N/128) * ceil(NT2/64), pair
// index p -> (i >= j) by tri_decode, then the P-tile, then the mu-tile.
// 8 waves as 4 (P) x 2 (mu) of 32 x 32.  Stages of 64 a-values go HBM -> LDS
// with global_load_lds_dwordx4 (double-buffered, XOR-swizzled as in the Pi
// kernels); the 64-byte column pieces of the 5 slices of the R tile (128
// columns) and of the V tile (64 columns) make 60 wave loads per stage.
constexpr int SXT_P = 128, SXT_M = 64;
constexpr int SX_LOADS = SX_NSL * (SXT_P + SXT_M) / 16;   // 60 wave loads of 1 KB per stage

__device__ __forceinline__ int swzx(int col_local, int piece) {
    return piece ^ ((col_local >> 2) & 3);
}

__global__ __launch_bounds__(512) void k_sosex_i8(
    const int8_t *__restrict__ SR, const int8_t *__restrict__ SV, const double *__restrict__ scR,
    const double *__restrict__ scV, const double *__restrict__ B, int64_t N, int64_t NT, int vpad,
    int ntp, int ntm, double *__restrict__ part) {
    // [buffer][P columns then mu columns][slot]: rows 0..5*128-1 are the R
    // slices (slice-major), then 5*64 rows of V slices
    __shared__ __attribute__((aligned(16))) uint4 sbuf[2][SX_NSL * (SXT_P + SXT_M)][4];
    __shared__ double sh[512], sl[512];
    const int64_t bid = blockIdx.x;
    const int64_t per_pair = (int64_t)ntp * ntm;
    const int pr = (int)(bid / per_pair);
    const int64_t rem = bid - (int64_t)pr * per_pair;
    const int tp = (int)(rem / ntm), tm = (int)(rem - (int64_t)tp * ntm);
    int i, j;
    tri_decode(pr, i, j);   // i >= j
    const int64_t P0 = (int64_t)tp * SXT_P, M0 = (int64_t)tm * SXT_M;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int wi = wave >> 1, wj = wave & 1;   // 32-row block (P), 32-col block (mu)

    const int lcol = lane >> 2;
    const int lpiece = (lane & 3) ^ ((lane >> 4) & 3);
    const int nst = vpad / 64;

    // one pass: W = R_x^T-slices x V_y-slices for this tile, combined to FP64
    auto pass = [&](int x, int y, double (&W)[2][2][4]) {
        const int8_t *Rx = SR + (int64_t)x * SX_NSL * N * vpad;
        const int8_t *Vy = SV + (int64_t)y * SX_NSL * NT * vpad;
        int64_t goff[8];
        const int8_t *gsrc[8];
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            const int id = 8 * wave + t;
            goff[t] = 0;
            gsrc[t] = Rx;
            if (id < SX_LOADS) {
                const int row = 16 * id + lcol;   // LDS column row of this lane
                if (row < SX_NSL * SXT_P) {
                    const int s = row / SXT_P, cl = row - s * SXT_P;
                    int64_t P = P0 + cl;
                    P = P < N ? P : N - 1;
                    gsrc[t] = Rx;
                    goff[t] = ((int64_t)s * N + P) * vpad + 16 * lpiece;
                } else {
                    const int r2 = row - SX_NSL * SXT_P;
                    const int s = r2 / SXT_M, cl = r2 - s * SXT_M;
                    int64_t M = M0 + cl;
                    M = M < NT ? M : NT - 1;
                    gsrc[t] = Vy;
                    goff[t] = ((int64_t)s * NT + M) * vpad + 16 * lpiece;
                }
            }
        }
        auto issue = [&](int buf, int k0) {
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                const int id = 8 * wave + t;
                if (id < SX_LOADS)
                    __builtin_amdgcn_global_load_lds((const void *)(gsrc[t] + goff[t] + k0),
                                                     (lds_void_t *)&sbuf[buf][16 * id][0], 16, 0, 0);
            }
        };
        auto fragR = [&](int buf, int s, int row) {
            const uint4 v = sbuf[buf][s * SXT_P + row][swzx(row & 15, lane >> 4)];
            return __builtin_bit_cast(v4i_t, v);
        };
        auto fragV = [&](int buf, int s, int row) {
            const uint4 v = sbuf[buf][SX_NSL * SXT_P + s * SXT_M + row][swzx(row & 15, lane >> 4)];
            return __builtin_bit_cast(v4i_t, v);
        };
        v4i_t L[SX_NSL][2][2];
#pragma unroll
        for (int l = 0; l < SX_NSL; ++l)
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int c = 0; c < 2; ++c) L[l][a][c] = (v4i_t){0, 0, 0, 0};
        __syncthreads();   // the previous pass is done with sbuf
        issue(0, 0);
        for (int st = 0; st < nst; ++st) {
            const int buf = st & 1;
            __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
            if (st + 1 < nst) issue(buf ^ 1, (st + 1) * 64);
            v4i_t bq[SX_NSL][2];
#pragma unroll
            for (int s2 = 0; s2 < SX_NSL; ++s2)
#pragma unroll
                for (int c = 0; c < 2; ++c) bq[s2][c] = fragV(buf, s2, 32 * wj + 16 * c + (lane & 15));
#pragma unroll
            for (int s1 = 0; s1 < SX_NSL; ++s1) {
                v4i_t a[2];
#pragma unroll
                for (int r = 0; r < 2; ++r) a[r] = fragR(buf, s1, 32 * wi + 16 * r + (lane & 15));
#pragma unroll
                for (int s2 = 0; s2 < SX_NSL; ++s2) {
                    if (s1 + s2 < SX_NSL) {
#pragma unroll
                        for (int r = 0; r < 2; ++r)
#pragma unroll
                            for (int c = 0; c < 2; ++c)
                                L[s1 + s2][r][c] = __builtin_amdgcn_mfma_i32_16x16x64_i8(
                                    a[r], bq[s2][c], L[s1 + s2][r][c], 0, 0, 0);
                    }
                }
            }
        }
        const double *sRx = scR + (int64_t)x * N;
        const double *sVy = scV + (int64_t)y * NT;
        const double c1 = 1.0 / 128.0;
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const int64_t M = M0 + 32 * wj + 16 * c + (lane & 15);
            const double sm = M < NT ? sVy[M] : 0.0;
#pragma unroll
            for (int r = 0; r < 2; ++r)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int64_t P = P0 + 32 * wi + 16 * r + 4 * (lane >> 4) + q;
                    const double sp = P < N ? sRx[P] : 0.0;
                    // Horner in 2^-7 from the smallest level up
                    double v = (double)L[SX_NSL - 1][r][c][q];
#pragma unroll
                    for (int l = SX_NSL - 2; l >= 0; --l) v = fma(v, c1, (double)L[l][r][c][q]);
                    W[r][c][q] = v * (sp * sm);
                }
        }
    };

    double W1[2][2][4], W2[2][2][4];
    pass(i, j, W1);   // W_ij
    if (i != j)
        pass(j, i, W2);   // W_ji
    else {
#pragma unroll
        for (int r = 0; r < 2; ++r)
#pragma unroll
            for (int c = 0; c < 2; ++c)
#pragma unroll
                for (int q = 0; q < 4; ++q) W2[r][c][q] = W1[r][c][q];
    }
    dd_t s = {0.0, 0.0};
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        const int64_t M = M0 + 32 * wj + 16 * c + (lane & 15);
        const double bm = M < NT ? B[M] : 0.0;
#pragma unroll
        for (int r = 0; r < 2; ++r)
#pragma unroll
            for (int q = 0; q < 4; ++q) s = dd_add(s, W1[r][c][q] * W2[r][c][q] * bm);
    }
    sh[tid] = s.hi;
    sl[tid] = s.lo;
    __syncthreads();
    for (int k = 256; k > 0; k >>= 1) {
        if (tid < k) {
            dd_t x = {sh[tid], sl[tid]};
            x = dd_add(x, (dd_t){sh[tid + k], sl[tid + k]});
            sh[tid] = x.hi;
            sl[tid] = x.lo;
        }
        __syncthreads();
    }
    if (tid == 0) {
        const double f = (i == j) ? 1.0 : 2.0;
        part[2 * bid] = f * sh[0];
        part[2 * bid + 1] = f * sl[0];
    }
}

// Deterministic double-double sum of n partials (one work-group, fixed order).
__global__ __launch_bounds__(1024) void k_sum_dd(const double *__restrict__ part, int64_t n,
                                                double *__restrict__ out) {
    __shared__ double sh[1024], sl[1024];
    const int tid = threadIdx.x;
    dd_t s = {0.0, 0.0};
    for (int64_t e = tid; e < n; e += 1024) s = dd_add(s, (dd_t){part[2 * e], part[2 * e + 1]});
    sh[tid] = s.hi;
    sl[tid] = s.lo;
    __syncthreads();
    for (int k = 512; k > 0; k >>= 1) {
        if (tid < k) {
            dd_t x = {sh[tid], sl[tid]};
            x = dd_add(x, (dd_t){sh[tid + k], sl[tid + k]});
            sh[tid] = x.hi;
            sl[tid] = x.lo;
        }
        __syncthreads();
    }
    if (tid == 0) {
        out[0] = sh[0];
        out[1] = sl[0];
    }
}

int sosex_i8_vpad(int64_t v) { return (int)(((v + SX_KB - 1) / SX_KB) * SX_KB); }

size_t sosex_i8_slice_bytes(int64_t nblk, int64_t ncols, int64_t v) {
    return (size_t)nblk * SX_NSL * ncols * sosex_i8_vpad(v);
}

int launch_sosex_i8_split(brpa_ctx *ctx, const double *X, int64_t nblk, int64_t v, int64_t ncols,
                          double *scale, int8_t *out, const double *rs) {
    const int ncc = (int)((ncols + 63) / 64);
    const int vpad = sosex_i8_vpad(v);
    hipLaunchKernelGGL(k_blkscale_i8, dim3((unsigned)(nblk * ncc)), dim3(256), 0, ctx->stream, X,
                       ncols, (int)v, ncc, rs, scale);
    LAUNCH_CHECK(ctx);
    hipLaunchKernelGGL(k_blksplit_i8, dim3((unsigned)(nblk * (vpad / SX_KB) * ncc)), dim3(256), 0,
                       ctx->stream, X, ncols, (int)v, vpad, ncc, rs, scale, out);
    LAUNCH_CHECK(ctx);
    return BRPA_OK;
}

// Column maxima of one row-major block in parallel over row chunks of 256:
// pm[rc][c] = max over rows 256 rc .. of |X[r][c] rs[r]|; then scale[c] =
// max_rc pm[rc][c] / 127 (the same products as k_blksplit_i8 forms).
__global__ __launch_bounds__(256) void k_colmax_part(const double *__restrict__ X, int64_t ncols,
                                                    int v, int ncc, const double *__restrict__ rs,
                                                    double *__restrict__ pm) {
    __shared__ double sM[4][64];
    const int rc = blockIdx.x / ncc, cc = blockIdx.x - rc * ncc;
    const int tid = threadIdx.x, c = tid & 63, g = tid >> 6;
    const int64_t col = (int64_t)cc * 64 + c;
    const bool ok = col < ncols;
    const gdbl_t *G = (const gdbl_t *)X;
    double m = 0.0;
#pragma unroll 4
    for (int u = 0; u < 64; ++u) {
        const int a = rc * 256 + g + 4 * u;
        const bool in = ok && a < v;
        double x = G[in ? (int64_t)a * ncols + col : 0];
        if (rs) x *= rs[in ? a : 0];
        m = fmax(m, in ? fabs(x) : 0.0);
    }
    sM[g][c] = m;
    __syncthreads();
    if (tid < 64 && ok)
        pm[(int64_t)rc * ncols + col] = fmax(fmax(sM[0][c], sM[1][c]), fmax(sM[2][c], sM[3][c]));
}

__global__ __launch_bounds__(256) void k_colmax_final(const double *__restrict__ pm, int64_t ncols,
                                                     int nrc, double *__restrict__ scale) {
    const int64_t col = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (col >= ncols) return;
    double m = 0.0;
    for (int rc = 0; rc < nrc; ++rc) m = fmax(m, pm[(int64_t)rc * ncols + col]);
    scale[col] = m / 127.0;
}

int64_t rowslice_i8_pm_len(int64_t v, int64_t ncols) { return (v + 255) / 256 * ncols; }

int launch_rowslice_i8(brpa_ctx *ctx, const double *X, int64_t v, int64_t ncols, const double *rs,
                       double *pm, double *scale, int8_t *out) {
    if (v <= 0 || ncols <= 0) return BRPA_OK;
    const int ncc = (int)((ncols + 63) / 64);
    const int nrc = (int)((v + 255) / 256);
    const int vpad = sosex_i8_vpad(v);
    hipLaunchKernelGGL(k_colmax_part, dim3((unsigned)(nrc * ncc)), dim3(256), 0, ctx->stream, X,
                       ncols, (int)v, ncc, rs, pm);
    LAUNCH_CHECK(ctx);
    hipLaunchKernelGGL(k_colmax_final, dim3((unsigned)((ncols + 255) / 256)), dim3(256), 0,
                       ctx->stream, pm, ncols, nrc, scale);
    LAUNCH_CHECK(ctx);
    hipLaunchKernelGGL(k_blksplit_i8, dim3((unsigned)((vpad / SX_KB) * ncc)), dim3(256), 0,
                       ctx->stream, X, ncols, (int)v, vpad, ncc, rs, scale, out);
    LAUNCH_CHECK(ctx);
    return BRPA_OK;
}

int64_t sosex_i8_nparts(int64_t o, int64_t N, int64_t NT) {
    return o * (o + 1) / 2 * ((N + SXT_P - 1) / SXT_P) * ((NT + SXT_M - 1) / SXT_M);
}

int launch_sosex_i8_pairs(brpa_ctx *ctx, const int8_t *SR, const int8_t *SV, const double *scR,
                          const double *scV, const double *B, int64_t o, int64_t v, int64_t N,
                          int64_t NT, double *part, double *out2) {
    const int ntp = (int)((N + SXT_P - 1) / SXT_P), ntm = (int)((NT + SXT_M - 1) / SXT_M);
    const int64_t nwg = sosex_i8_nparts(o, N, NT);
    if (nwg > (int64_t)UINT32_MAX)
        return brpa_fail(ctx, BRPA_ERR_ARG, "split-int8 SOSEX: %lld work-groups exceed the grid",
                         (long long)nwg);
    hipLaunchKernelGGL(k_sosex_i8, dim3((unsigned)nwg), dim3(512), 0, ctx->stream, SR, SV, scR,
                       scV, B, N, NT, sosex_i8_vpad(v), ntp, ntm, part);
    LAUNCH_CHECK(ctx);
    hipLaunchKernelGGL(k_sum_dd, dim3(1), dim3(1024), 0, ctx->stream, part, nwg, out2);
    LAUNCH_CHECK(ctx);
    return BRPA_OK;
}
