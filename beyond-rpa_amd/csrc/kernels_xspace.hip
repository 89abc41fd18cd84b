// Kernels of the exchange correction in the explicit ov space.
//
// The exchange energy of rpa_CC_Exchange_SOSEX (rpa/rpa_CC_Exchange.f90:36-116),
//   E_x = -alpha sum_{j<=i} (2 - delta_ij) sum_{P,mu} W_ji(P,mu) W_ij(P,mu) B_mu,
//   W_ij = R_i V_j^T,
// is the contraction -alpha sum_{aibj} (ai|bj) T2(aj,bi) of the Coulomb matrix
// G = R R^T (ov x ov) with the amplitude matrix T2 = V diag(B) V^T whose two
// occupied indices are swapped.  With both held as ov x ov matrices the
// contraction costs ov^2 N flop (G, one SYRK) instead of 2 o^2 v N NT2 for the
// pair products of the reference's order.
//
//   k_symm_fill     lower triangle of a column-major symmetric matrix from its upper one
//   k_sosex_contract  sum_{r,c} G(r,c) (gamma M(pi(r,c)) + delta [r==c]) from an explicit
//                     G = R R^T (one column per work-group)
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
// Exchange contraction from an explicit Coulomb matrix G = R R^T (K x K,
// column-major, upper triangle and diagonal valid):
//   part[c] = sum_{r <= c} w_rc G(r, c) (gamma M(r', c') + delta [r == c]),
//   r = i v + a, c = j v + b, r' = j v + a, c' = i v + b, w = 2 (r < c), 1 (r = c).
// G and the symmetric T2 matrix M are both symmetric and so is the permuted
// M, so the upper triangle with these weights is the full sum.  One work-group
// per column c; its rows r <= c run as (i, a) with i <= j, so the reads of G
// (column c) and of M (column c' = i v + b, rows j v + a) are contiguous in a.
// Fixed-order reduction (per-thread sums, wave butterfly, the waves in order).
__global__ __launch_bounds__(256) void k_sosex_contract(const double *__restrict__ G, int64_t K,
                                                        int v, const double *__restrict__ M,
                                                        double gamma, double delta,
                                                        double *__restrict__ part) {
    __shared__ double sred[4];
    const int64_t c = blockIdx.x;
    const int j = (int)(c / v), b = (int)(c - (int64_t)j * v);
    const double *__restrict__ Gc = G + K * c;
    double s = 0.0;
    for (int i = 0; i <= j; ++i) {
        const int amax = i < j ? v : b + 1;   // rows r = i v + a <= c
        const double *__restrict__ Mc = M + K * ((int64_t)i * v + b) + (int64_t)j * v;
        const double *__restrict__ Gr = Gc + (int64_t)i * v;
        for (int a = threadIdx.x; a < amax; a += 256) {
            const bool dg = i == j && a == b;
            double t = gamma * Mc[a];
            if (dg) t += delta;
            s = fma(dg ? Gr[a] : 2.0 * Gr[a], t, s);
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if ((threadIdx.x & 63) == 0) sred[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) part[c] = (sred[0] + sred[1]) + (sred[2] + sred[3]);
}

int launch_sosex_contract(brpa_ctx *ctx, const double *G, int64_t K, int64_t v, const double *M,
                          double gamma, double delta, double *part) {
    if (K <= 0) return BRPA_OK;
    if (K > INT32_MAX || v <= 0 || K % v != 0)
        return brpa_fail(ctx, BRPA_ERR_ARG, "exchange contraction: bad sizes");
    size_t slot = 0;
    BRPA_TRY(kstat_begin(ctx, &slot));
    hipLaunchKernelGGL(k_sosex_contract, dim3((unsigned)K), dim3(256), 0, ctx->stream, G, K,
                       (int)v, M, gamma, delta, part);
    LAUNCH_CHECK(ctx);
    // algorithmic flop of the contraction itself (one multiply-add per G element
    // of the upper triangle)
    return kstat_end(ctx, slot, BRPA_KERNEL_SOSEX_DOT, (double)K * (double)(K + 1));
}
