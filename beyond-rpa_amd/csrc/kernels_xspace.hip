// Kernels of the exchange correction in the explicit ov space.
//
// The exchange energy of rpa_CC_Exchange_SOSEX (rpa/rpa_CC_Exchange.f90:36-116),
//   E_x = -alpha sum_{j<=i} (2 - delta_ij) sum_{P,mu} W_ji(P,mu) W_ij(P,mu) B_mu,
//   W_ij = R_i V_j^T,
// is the contraction -alpha sum_{aibj} (ai|bj) T2(aj,bi) of the Coulomb matrix
// G = R R^T (ov x ov) with the amplitude matrix T2 = V diag(B) V^T whose two
// occupied indices are swapped.  With both held as ov x ov matrices the
// contraction costs ov^2 N flop (G) instead of 2 o^2 v N NT2 for the pair
// products of the reference's order.
//
//   k_symm_fill     lower triangle of a column-major symmetric matrix from its upper one
//   k_sosex_dot     per tile pair of G: sum_{r,c} G(r,c) (gamma M(pi(r,c)) + delta [r==c])
//   k_ritz_resid    residual norms |A z_j - theta_j z_j| of Ritz pairs
#include "device_common.h"
#include "internal.h"

#define LAUNCH_CHECK(ctx)                                                                      \
    do {                                                                                       \
        hipError_t _e = hipGetLastError();                                                     \
        if (_e != hipSuccess)                                                                  \
            return brpa_fail((ctx), BRPA_ERR_HIP, "kernel launch failed: %s (%s:%d)",          \
                             hipGetErrorString(_e), __FILE__, __LINE__);                       \
    } while (0)

// ---------------------------------------------------------------------------
// M(r, c) = M(c, r) for r > c, n x n column-major (element (r, c) at r + n c):
// the lower triangle is written from the upper one.  One 32 x 32 tile pair per
// work-group: the upper tile (tc, tr) is read along its columns, transposed in
// LDS and written as the lower tile (tr, tc), so both sides are coalesced.
__global__ __launch_bounds__(256) void k_symm_fill(double *__restrict__ M, int64_t n) {
    __shared__ double t[32][33];
    int tr, tc;
    tri_decode((int)blockIdx.x, tr, tc);   // tr >= tc: lower tile (tr, tc)
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    const int64_t r0 = (int64_t)tr * 32, c0 = (int64_t)tc * 32;
    // source: upper tile rows c0.., columns r0.. ; element (c0 + x, r0 + y)
    for (int y = ty; y < 32; y += 8) {
        const int64_t row = c0 + tx, col = r0 + y;
        if (row < n && col < n) t[y][tx] = M[row + n * col];
    }
    __syncthreads();
    // destination (r0 + x, c0 + y) = source (c0 + y, r0 + x) = t[x][y]
    for (int y = ty; y < 32; y += 8) {
        const int64_t row = r0 + tx, col = c0 + y;
        if (row < n && col < n && row > col) M[row + n * col] = t[tx][y];
    }
}

int launch_symm_fill(brpa_ctx *ctx, double *M, int64_t n) {
    if (n <= 1) return BRPA_OK;
    const int64_t nt = (n + 31) / 32;
    hipLaunchKernelGGL(k_symm_fill, dim3((unsigned)(nt * (nt + 1) / 2)), dim3(256), 0, ctx->stream,
                       M, n);
    LAUNCH_CHECK(ctx);
    return BRPA_OK;
}

// ---------------------------------------------------------------------------
// out[j] = |AZ(:, j) - theta[j] Z(:, j)|_2 for the m columns of the n-row
// column-major AZ and Z (leading dimension n).  One work-group per column; the
// partial sums are combined in a fixed order (wave butterfly, then the waves).
__global__ __launch_bounds__(256) void k_ritz_resid(const double *__restrict__ AZ,
                                                    const double *__restrict__ Z,
                                                    const double *__restrict__ theta, int64_t n,
                                                    double *__restrict__ out) {
    __shared__ double sred[4];
    const int64_t j = blockIdx.x;
    const double t = theta[j];
    const double *a = AZ + j * n, *z = Z + j * n;
    double s = 0.0;
    for (int64_t r = threadIdx.x; r < n; r += 256) {
        const double d = a[r] - t * z[r];
        s = fma(d, d, s);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if ((threadIdx.x & 63) == 0) sred[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) out[j] = sqrt((sred[0] + sred[1]) + (sred[2] + sred[3]));
}

int launch_ritz_resid(brpa_ctx *ctx, const double *AZ, const double *Z, const double *theta,
                      int64_t n, int64_t m, double *out) {
    if (m <= 0) return BRPA_OK;
    hipLaunchKernelGGL(k_ritz_resid, dim3((unsigned)m), dim3(256), 0, ctx->stream, AZ, Z, theta,
                       n, out);
    LAUNCH_CHECK(ctx);
    return BRPA_OK;
}

// ---------------------------------------------------------------------------
// Exchange contraction in the explicit ov space.
//
//   part[pair] = w_pair * sum_{(r,c) in tile pair} G(r, c) * Tp(r, c)
//   G(r, c)   = sum_P R[r][P] R[c][P]                      (R row-major K x N)
//   Tp(r, c)  = gamma * M(r', c') + delta * [r == c]
//   r = i v + a, c = j v + b  ->  r' = j v + a, c' = i v + b
// M is symmetric with both triangles stored; it is read at c' + K r', which
// is contiguous across the lanes that hold consecutive columns c.  Tile pairs
// (tp >= tq) of 128 x 128 cover the lower triangle of G; w_pair = 2 off the
// diagonal (G and Tp are both symmetric), 1 on it.  Four waves per work-group,
// each a 64 x 64 quarter of the tile on v_mfma_f64_16x16x4_f64; R panels of 16
// columns are staged in LDS with a register prefetch of the next panel.
//
// MFMA operand layout (lane l): A[m = l%16][k = l/16], B[k = l/16][n = l%16];
// D[r] = C[m = l/16 + 4r][n = l%16].
namespace {
constexpr int XT = 128;       // tile edge
constexpr int XKB = 16;       // columns of R per LDS stage
constexpr int XPAD = 16;      // LDS row padding (rows k and k+1 land 32 banks apart)
}

__global__ __launch_bounds__(256) void k_sosex_dot(const double *__restrict__ R, int K, int N,
                                                  int v, const double *__restrict__ M,
                                                  double gamma, double delta,
                                                  double *__restrict__ part) {
    __shared__ double sA[XKB][XT + XPAD];
    __shared__ double sB[XKB][XT + XPAD];
    __shared__ double sred[4];
    int tp, tq;
    tri_decode((int)blockIdx.x, tp, tq);
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int wm = wave >> 1, wn = wave & 1;
    const int p0 = tp * XT, q0 = tq * XT;
    const bool diag = tp == tq;

    d4_t acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = (d4_t){0.0, 0.0, 0.0, 0.0};

    // thread -> (row, column) of a 128 x 16 panel: 16 consecutive threads read
    // one row's 16 consecutive columns
    double ra[8], rb[8];
    auto gload = [&](int k0) {
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int e = tid + 256 * i;
            const int c = e & 15, r = e >> 4;
            const int k = k0 + c;
            const bool kin = k < N;
            ra[i] = (kin && p0 + r < K) ? R[(int64_t)(p0 + r) * N + k] : 0.0;
            if (!diag) rb[i] = (kin && q0 + r < K) ? R[(int64_t)(q0 + r) * N + k] : 0.0;
        }
    };
    auto lstore = [&]() {
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int e = tid + 256 * i;
            const int c = e & 15, r = e >> 4;
            sA[c][r] = ra[i];
            if (!diag) sB[c][r] = rb[i];
        }
    };
    const double(*pB)[XT + XPAD] = diag ? sA : sB;

    gload(0);
    for (int k0 = 0; k0 < N; k0 += XKB) {
        __syncthreads();
        lstore();
        __syncthreads();
        if (k0 + XKB < N) gload(k0 + XKB);
#pragma unroll
        for (int kk = 0; kk < XKB; kk += 4) {
            const int kr = kk + (lane >> 4);
            double a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) a[i] = sA[kr][64 * wm + 16 * i + (lane & 15)];
#pragma unroll
            for (int j = 0; j < 4; ++j) b[j] = pB[kr][64 * wn + 16 * j + (lane & 15)];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
        }
    }

    // epilogue: (occupied, virtual) of the lane's 4 columns and 16 rows
    int cj[4], cb[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int c = q0 + 64 * wn + 16 * j + (lane & 15);
        cj[j] = c / v;
        cb[j] = c - cj[j] * v;
    }
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int rr = 0; rr < 4; ++rr) {
            const int r = p0 + 64 * wm + 16 * i + (lane >> 4) + 4 * rr;
            if (r >= K) continue;
            const int ri = r / v, ra_ = r - ri * v;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int c = q0 + 64 * wn + 16 * j + (lane & 15);
                if (c >= K) continue;
                const int64_t rp = (int64_t)cj[j] * v + ra_;   // (a, j)
                const int64_t cp = (int64_t)ri * v + cb[j];    // (b, i)
                double t = gamma * M[cp + (int64_t)K * rp];
                if (r == c) t += delta;
                s = fma(acc[i][j][rr], t, s);
            }
        }
    // fixed-order reduction: wave butterfly, then the four waves in order
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane == 0) sred[wave] = s;
    __syncthreads();
    if (tid == 0) {
        const double tot = (sred[0] + sred[1]) + (sred[2] + sred[3]);
        part[blockIdx.x] = diag ? tot : 2.0 * tot;
    }
}

int64_t sosex_dot_nparts(int64_t K) {
    const int64_t nt = (K + XT - 1) / XT;
    return nt * (nt + 1) / 2;
}

int launch_sosex_dot(brpa_ctx *ctx, const double *R, int64_t K, int64_t N, int64_t v,
                     const double *M, double gamma, double delta, double *part) {
    if (K <= 0 || N <= 0) return BRPA_OK;
    if (K > INT32_MAX / 2 || N > INT32_MAX / 2 || sosex_dot_nparts(K) > INT32_MAX)
        return brpa_fail(ctx, BRPA_ERR_ARG, "exchange contraction: sizes beyond the supported range");
    size_t slot = 0;
    BRPA_TRY(kstat_begin(ctx, &slot));
    hipLaunchKernelGGL(k_sosex_dot, dim3((unsigned)sosex_dot_nparts(K)), dim3(256), 0, ctx->stream,
                       R, (int)K, (int)N, (int)v, M, gamma, delta, part);
    LAUNCH_CHECK(ctx);
    // algorithmic flop: the lower triangle of (ai|bj), K (K + 1) N
    return kstat_end(ctx, slot, BRPA_KERNEL_SOSEX_DOT, (double)K * (double)(K + 1) * (double)N);
}
