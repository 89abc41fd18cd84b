// Dense FP64 GEMM and triangular solve on v_mfma_f64_16x16x4_f64 (gfx950).
//
// Column-major operands (element (r, c) of a matrix with leading dimension ld
// at r + ld c), the layout of the reference's Fortran arrays and of the
// explicit-space exchange.
//
//   k_gemm_f64<TA, TB>  C = alpha op(A) op(B) + beta C on 128 x 128 tiles; in the
//                       upper mode only tiles on or above the diagonal (SYRK)
//   k_trsm_f64<BROW, IDENT>
//                       L Y = B by blocked forward substitution, L lower
//                       triangular (the row-major lower factor the Cholesky
//                       kernels leave, i.e. U^T of the column-major upper U);
//                       B column-major or row-major, or the identity (then
//                       Y = L^{-1}); batched over matrices with strides
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
constexpr int GT = 128;    // GEMM tile edge
constexpr int GKB = 16;    // k per LDS stage
constexpr int GPAD = 16;   // LDS row padding: rows k and k+1 start 128 B apart mod 256 B
}

// ---------------------------------------------------------------------------
// One 128 x 128 tile of C per work-group, four waves of 64 x 64.  The A and B
// panels of a stage (128 x 16 each) go to LDS as sA[k][row], sB[k][col]; the
// next stage is prefetched into registers.  The MFMA takes the B fragment as
// its first operand, so the accumulator rows run along the lanes
// (D[r] = C(row = lane%16, col = lane/16 + 4r) of a 16 x 16 block) and the
// stores of C are contiguous across lanes.
//
// Loads by orientation: op(A) = A (TA = false) reads A(row, k) along rows, 128
// consecutive doubles per k; TA reads A(k, row) along k, 16 consecutive
// doubles per row.  op(B) = B (TB = false) reads B(k, col) along k; TB reads
// B(col, k) along columns.
template <bool TA, bool TB>
__global__ __launch_bounds__(256) void k_gemm_f64(int m, int n, int k, double alpha,
                                                 const double *__restrict__ A, int64_t lda,
                                                 const double *__restrict__ B, int64_t ldb,
                                                 double beta, double *__restrict__ C, int64_t ldc,
                                                 int upper, int mt) {
    __shared__ double sA[GKB][GT + GPAD];
    __shared__ double sB[GKB][GT + GPAD];
    int ti, tj;
    if (upper) {
        // upper tiles ti <= tj in column order: tj (tj+1)/2 + ti
        tri_decode((int)blockIdx.x, tj, ti);
    } else {
        ti = (int)(blockIdx.x % (unsigned)mt);
        tj = (int)(blockIdx.x / (unsigned)mt);
    }
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int wm = wave >> 1, wn = wave & 1;
    const int m0 = ti * GT, n0 = tj * GT;

    d4_t acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = (d4_t){0.0, 0.0, 0.0, 0.0};

    double ra[8], rb[8];
    auto gload = [&](int k0) {
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int e = tid + 256 * i;
            if (!TA) {
                const int r = e & 127, kk = e >> 7;
                ra[i] = (m0 + r < m && k0 + kk < k) ? A[(int64_t)(m0 + r) + lda * (k0 + kk)] : 0.0;
            } else {
                const int kk = e & 15, r = e >> 4;
                ra[i] = (m0 + r < m && k0 + kk < k) ? A[(int64_t)(k0 + kk) + lda * (m0 + r)] : 0.0;
            }
            if (!TB) {
                const int kk = e & 15, c = e >> 4;
                rb[i] = (n0 + c < n && k0 + kk < k) ? B[(int64_t)(k0 + kk) + ldb * (n0 + c)] : 0.0;
            } else {
                const int c = e & 127, kk = e >> 7;
                rb[i] = (n0 + c < n && k0 + kk < k) ? B[(int64_t)(n0 + c) + ldb * (k0 + kk)] : 0.0;
            }
        }
    };
    auto lstore = [&]() {
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int e = tid + 256 * i;
            if (!TA)
                sA[e >> 7][e & 127] = ra[i];
            else
                sA[e & 15][e >> 4] = ra[i];
            if (!TB)
                sB[e & 15][e >> 4] = rb[i];
            else
                sB[e >> 7][e & 127] = rb[i];
        }
    };

    gload(0);
    for (int k0 = 0; k0 < k; k0 += GKB) {
        __syncthreads();
        lstore();
        __syncthreads();
        if (k0 + GKB < k) gload(k0 + GKB);
#pragma unroll
        for (int kk = 0; kk < GKB; kk += 4) {
            const int kr = kk + (lane >> 4);
            double a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) a[i] = sA[kr][64 * wm + 16 * i + (lane & 15)];
#pragma unroll
            for (int j = 0; j < 4; ++j) b[j] = sB[kr][64 * wn + 16 * j + (lane & 15)];
            // C^T block = op(B)^T op(A)^T: rows of C along the lanes
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(b[j], a[i], acc[i][j], 0, 0, 0);
        }
    }

#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int row = m0 + 64 * wm + 16 * i + (lane & 15);
        if (row >= m) continue;
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int col = n0 + 64 * wn + 16 * j + (lane >> 4) + 4 * r;
                if (col >= n) continue;
                double *p = C + (int64_t)row + ldc * col;
                const double v = alpha * acc[i][j][r];
                *p = beta == 0.0 ? v : fma(beta, *p, v);
            }
    }
}

int launch_gemm_f64(brpa_ctx *ctx, bool ta, bool tb, int64_t m, int64_t n, int64_t k,
                    double alpha, const double *A, int64_t lda, const double *B, int64_t ldb,
                    double beta, double *C, int64_t ldc, bool upper) {
    if (m <= 0 || n <= 0) return BRPA_OK;
    if (m > INT32_MAX / 2 || n > INT32_MAX / 2 || k > INT32_MAX / 2 || (upper && m != n))
        return brpa_fail(ctx, BRPA_ERR_ARG, "gemm: unsupported shape");
    const int64_t mt = (m + GT - 1) / GT, nt = (n + GT - 1) / GT;
    const int64_t nb = upper ? nt * (nt + 1) / 2 : mt * nt;
    if (nb > INT32_MAX) return brpa_fail(ctx, BRPA_ERR_ARG, "gemm: too many tiles");
    const dim3 g((unsigned)nb), b(256);
    const int im = (int)m, in = (int)n, ik = (int)k > 0 ? (int)k : 0, iu = upper ? 1 : 0,
              imt = (int)mt;
    size_t slot = 0;
    BRPA_TRY(kstat_begin(ctx, &slot));
    if (!ta && !tb)
        hipLaunchKernelGGL((k_gemm_f64<false, false>), g, b, 0, ctx->stream, im, in, ik, alpha, A,
                           lda, B, ldb, beta, C, ldc, iu, imt);
    else if (!ta && tb)
        hipLaunchKernelGGL((k_gemm_f64<false, true>), g, b, 0, ctx->stream, im, in, ik, alpha, A,
                           lda, B, ldb, beta, C, ldc, iu, imt);
    else if (ta && !tb)
        hipLaunchKernelGGL((k_gemm_f64<true, false>), g, b, 0, ctx->stream, im, in, ik, alpha, A,
                           lda, B, ldb, beta, C, ldc, iu, imt);
    else
        hipLaunchKernelGGL((k_gemm_f64<true, true>), g, b, 0, ctx->stream, im, in, ik, alpha, A,
                           lda, B, ldb, beta, C, ldc, iu, imt);
    LAUNCH_CHECK(ctx);
    // algorithmic flop: 2 m n k, or n (n + 1) k for the triangle of a SYRK
    const double fl = upper ? (double)n * (double)(n + 1) * (double)ik
                            : 2.0 * (double)m * (double)n * (double)ik;
    return kstat_end(ctx, slot, BRPA_KERNEL_GEMM_F64, fl);
}

// ---------------------------------------------------------------------------
// Forward substitution L Y = B, L (n x n) lower triangular stored as the
// row-major lower triangle (L(r, c) at r * ldl + c, c <= r), Y overwriting B.
// B is n x nrhs: column-major (BROW = false, B(r, c) at r + ldb c) or row-major
// (BROW, B(r, c) at r ldb + c, i.e. the rows of a column-major nrhs x n
// matrix X: X <- X L^{-T}).  IDENT: B = 1 (n = nrhs), the result L^{-1} written
// to Y (column-major, ldb), zero above the diagonal.
//
// One work-group per (matrix, 64 right-hand sides), left-looking over 64-row
// blocks p: acc = B_p - sum_{q<p} L_pq Y_q on MFMA (each wave 16 rows x 64
// columns), then Y_p = L_pp^{-1} acc by substitution in LDS (one thread per
// right-hand side, the 64 x 64 diagonal block staged in LDS).  The Y_q blocks
// were written by the same work-group and come back from L2.
namespace {
constexpr int TB_ = 64;
}

template <bool BROW, bool IDENT>
__global__ __launch_bounds__(256) void k_trsm_f64(int n, int nrhs, const double *__restrict__ L,
                                                 int64_t ldl, int64_t strideL,
                                                 double *__restrict__ Y, int64_t ldb,
                                                 int64_t strideB, int nct) {
    __shared__ double sL[TB_][TB_ + 1];     // L_pq or L_pp, [row][col]
    __shared__ double sY[TB_][TB_ + 1];     // Y_q block [row][rhs]; then the block being solved
    const int mat = (int)(blockIdx.x / (unsigned)nct);
    const int ct = (int)(blockIdx.x % (unsigned)nct);
    const double *Lm = L + (int64_t)mat * strideL;
    double *Ym = Y + (int64_t)mat * strideB;
    const int c0 = ct * TB_;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int nblk = (n + TB_ - 1) / TB_;
    // B(r, c) / Y(r, c) address for row r < n, right-hand side c < nrhs
    auto addr = [&](int r, int c) -> int64_t {
        return BROW ? (int64_t)r * ldb + c : (int64_t)r + ldb * (int64_t)c;
    };
    // L^{-1} is lower triangular: rows above the first column block are zero
    const int pstart = IDENT ? ct : 0;
    if (IDENT) {
        // zero the blocks above the diagonal of this column range
        for (int e = tid; e < c0 * TB_; e += 256) {
            const int r = e / TB_, c = c0 + (e % TB_);
            if (c < nrhs) Ym[addr(r, c)] = 0.0;
        }
    }
    for (int p = pstart; p < nblk; ++p) {
        const int r0 = p * TB_;
        // acc: rows 16 * wave + (lane % 16) of the block, columns lane/16 + 4 r + 16 j
        d4_t acc[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[j] = (d4_t){0.0, 0.0, 0.0, 0.0};
        for (int q = pstart; q < p; ++q) {
            const int q0 = q * TB_;
            __syncthreads();
            for (int e = tid; e < TB_ * TB_; e += 256) {
                const int r = e >> 6, c = e & 63;
                sL[r][c] = (r0 + r < n) ? Lm[(int64_t)(r0 + r) * ldl + q0 + c] : 0.0;
            }
            for (int e = tid; e < TB_ * TB_; e += 256) {
                int r, c;
                if (BROW) {
                    r = e >> 6;
                    c = e & 63;
                } else {
                    c = e >> 6;
                    r = e & 63;
                }
                sY[r][c] = (c0 + c < nrhs) ? Ym[addr(q0 + r, c0 + c)] : 0.0;
            }
            __syncthreads();
#pragma unroll 4
            for (int kk = 0; kk < TB_; kk += 4) {
                const int kr = kk + (lane >> 4);
                const double a = sL[16 * wave + (lane & 15)][kr];
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const double b = sY[kr][16 * j + (lane & 15)];
                    // block (rows of L_pq) x (columns of Y): C^T = Y^T L^T, rows along lanes
                    acc[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(b, a, acc[j], 0, 0, 0);
                }
            }
        }
        __syncthreads();
        // sY <- B_p - acc ; sL <- L_pp
        for (int e = tid; e < TB_ * TB_; e += 256) {
            const int r = e >> 6, c = e & 63;
            sL[r][c] = (r0 + r < n && r0 + c < n) ? Lm[(int64_t)(r0 + r) * ldl + r0 + c] : 0.0;
        }
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int rr = 0; rr < 4; ++rr) {
                const int r = 16 * wave + (lane & 15);
                const int c = 16 * j + (lane >> 4) + 4 * rr;
                double bv;
                if (IDENT)
                    bv = (r0 + r == c0 + c) ? 1.0 : 0.0;
                else
                    bv = (r0 + r < n && c0 + c < nrhs) ? Ym[addr(r0 + r, c0 + c)] : 0.0;
                sY[r][c] = bv - acc[j][rr];
            }
        __syncthreads();
        // substitution: thread c < 64 owns right-hand side c
        if (tid < TB_) {
            const int c = tid;
            const int rmax = min(TB_, n - r0);
            for (int r = 0; r < rmax; ++r) {
                // four interleaved partial sums (a quarter of the dependent FMA chain)
                double s0 = sY[r][c], s1 = 0.0, s2 = 0.0, s3 = 0.0;
                int t = 0;
                for (; t + 3 < r; t += 4) {
                    s0 = fma(-sL[r][t], sY[t][c], s0);
                    s1 = fma(-sL[r][t + 1], sY[t + 1][c], s1);
                    s2 = fma(-sL[r][t + 2], sY[t + 2][c], s2);
                    s3 = fma(-sL[r][t + 3], sY[t + 3][c], s3);
                }
                for (; t < r; ++t) s0 = fma(-sL[r][t], sY[t][c], s0);
                sY[r][c] = ((s0 + s1) + (s2 + s3)) / sL[r][r];
            }
        }
        __syncthreads();
        for (int e = tid; e < TB_ * TB_; e += 256) {
            int r, c;
            if (BROW) {
                r = e >> 6;
                c = e & 63;
            } else {
                c = e >> 6;
                r = e & 63;
            }
            if (r0 + r < n && c0 + c < nrhs) Ym[addr(r0 + r, c0 + c)] = sY[r][c];
        }
        // make this block's writes visible to the work-group's later reads of it
        __threadfence_block();
    }
}

int launch_trsm_f64(brpa_ctx *ctx, bool brow, bool ident, int64_t n, int64_t nrhs,
                    const double *L, int64_t ldl, int64_t strideL, double *Y, int64_t ldb,
                    int64_t strideB, int64_t nmat) {
    if (n <= 0 || nrhs <= 0 || nmat <= 0) return BRPA_OK;
    if (n > INT32_MAX / 2 || nrhs > INT32_MAX / 2 || (ident && (n != nrhs || brow)))
        return brpa_fail(ctx, BRPA_ERR_ARG, "trsm: unsupported shape");
    const int64_t nct = (nrhs + TB_ - 1) / TB_;
    if (nct * nmat > INT32_MAX) return brpa_fail(ctx, BRPA_ERR_ARG, "trsm: too many blocks");
    const dim3 g((unsigned)(nct * nmat)), b(256);
    size_t slot = 0;
    BRPA_TRY(kstat_begin(ctx, &slot));
    if (ident)
        hipLaunchKernelGGL((k_trsm_f64<false, true>), g, b, 0, ctx->stream, (int)n, (int)nrhs, L,
                           ldl, strideL, Y, ldb, strideB, (int)nct);
    else if (brow)
        hipLaunchKernelGGL((k_trsm_f64<true, false>), g, b, 0, ctx->stream, (int)n, (int)nrhs, L,
                           ldl, strideL, Y, ldb, strideB, (int)nct);
    else
        hipLaunchKernelGGL((k_trsm_f64<false, false>), g, b, 0, ctx->stream, (int)n, (int)nrhs, L,
                           ldl, strideL, Y, ldb, strideB, (int)nct);
    LAUNCH_CHECK(ctx);
    // algorithmic flop: n^2 per right-hand side (n^3 / 3 for the inverse)
    const double fl = ident ? (double)n * (double)n * (double)n / 3.0
                            : (double)n * (double)n * (double)nrhs * (double)nmat;
    return kstat_end(ctx, slot, BRPA_KERNEL_TRSM_F64, fl);
}
