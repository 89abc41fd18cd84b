// Elementwise / reduction kernels of the amplitude (T2) and SOSEX paths.
// Matrices here are column-major (the reference's Fortran layout): element
// (r, c) of an m-row matrix is at r + m*c.
#include "device_common.h"
#include "internal.h"

#define LAUNCH_CHECK(ctx)                                                                      \
    do {                                                                                       \
        hipError_t _e = hipGetLastError();                                                     \
        if (_e != hipSuccess)                                                                  \
            return brpa_fail((ctx), BRPA_ERR_HIP, "kernel launch failed: %s (%s:%d)",          \
                             hipGetErrorString(_e), __FILE__, __LINE__);                       \
    } while (0)

static inline unsigned nblocks(int64_t n) { return (unsigned)((n + 255) / 256); }

// Omega(r, c) = u - 1/2, u = SplitMix64(seed + (c*m + r + 1) * golden) / 2^64
// (counter-based: the oracle draws the same bits).  Guess vectors of
// rpa_CC_T2_v5 (rpa/rpa_CC_Doubles.f90:571-588).
__global__ void k_guess(double *__restrict__ O, int64_t m, int64_t n, uint64_t seed) {
    int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= m * n) return;
    uint64_t z = seed + (uint64_t)(e + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    O[e] = (double)(z >> 11) * (1.0 / 9007199254740992.0) - 0.5;
}

int launch_guess(brpa_ctx *ctx, double *O, int64_t m, int64_t n, uint64_t seed) {
    if (m * n <= 0) return BRPA_OK;
    hipLaunchKernelGGL(k_guess, dim3(nblocks(m * n)), dim3(256), 0, ctx->stream, O, m, n, seed);
    LAUNCH_CHECK(ctx);
    return BRPA_OK;
}

// Divide every column of the m x n matrix by its 2-norm
// (rpa_CC_NormalizeColumns, rpa/rpa_CC_Doubles.f90:80-91).  One work-group
// per column, deterministic tree reduction.
__global__ __launch_bounds__(256) void k_normalize_columns(double *__restrict__ X, int64_t m) {
    __shared__ double sh[256];
    double *col = X + (int64_t)blockIdx.x * m;
    double s = 0.0;
    for (int64_t r = threadIdx.x; r < m; r += 256) s = fma(col[r], col[r], s);
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (threadIdx.x < k) sh[threadIdx.x] += sh[threadIdx.x + k];
        __syncthreads();
    }
    const double inv = 1.0 / sqrt(sh[0]);
    for (int64_t r = threadIdx.x; r < m; r += 256) col[r] *= inv;
}

int launch_normalize_columns(brpa_ctx *ctx, double *X, int64_t m, int64_t n) {
    if (m <= 0 || n <= 0) return BRPA_OK;
    hipLaunchKernelGGL(k_normalize_columns, dim3((unsigned)n), dim3(256), 0, ctx->stream, X, m);
    LAUNCH_CHECK(ctx);
    return BRPA_OK;
}

// out(r, c) = s[r] * in(r, c)   (out may alias in)
__global__ void k_row_scale(double *out, const double *in, const double *__restrict__ s, int64_t m,
                            int64_t n) {
    int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= m * n) return;
    out[e] = s[e % m] * in[e];
}

int launch_row_scale(brpa_ctx *ctx, double *out, const double *in, const double *s, int64_t m,
                     int64_t n) {
    if (m * n <= 0) return BRPA_OK;
    hipLaunchKernelGGL(k_row_scale, dim3(nblocks(m * n)), dim3(256), 0, ctx->stream, out, in, s, m, n);
    LAUNCH_CHECK(ctx);
    return BRPA_OK;
}

// io(r, c) += a * s[r] * in(r, c)
__global__ void k_row_scale_acc(double *__restrict__ io, const double *__restrict__ in,
                                const double *__restrict__ s, double a, int64_t m, int64_t n) {
    int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= m * n) return;
    io[e] = fma(a * s[e % m], in[e], io[e]);
}

int launch_row_scale_acc(brpa_ctx *ctx, double *io, const double *in, const double *s, double a,
                         int64_t m, int64_t n) {
    if (m * n <= 0) return BRPA_OK;
    hipLaunchKernelGGL(k_row_scale_acc, dim3(nblocks(m * n)), dim3(256), 0, ctx->stream, io, in, s,
                       a, m, n);
    LAUNCH_CHECK(ctx);
    return BRPA_OK;
}

// W(:, j) = U(:, c0 + j) / sqrt(ev[c0 + j]) for j < nt  (rpa_CC_Doubles.f90:610-614)
__global__ void k_select_scale(double *__restrict__ W, const double *__restrict__ U,
                               const double *__restrict__ ev, int64_t m, int64_t c0, int64_t nt) {
    int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= m * nt) return;
    int64_t j = e / m, r = e - j * m;
    W[e] = U[r + m * (c0 + j)] / sqrt(ev[c0 + j]);
}

int launch_select_scale(brpa_ctx *ctx, double *W, const double *U, const double *ev, int64_t m,
                        int64_t c0, int64_t nt) {
    if (m * nt <= 0) return BRPA_OK;
    hipLaunchKernelGGL(k_select_scale, dim3(nblocks(m * nt)), dim3(256), 0, ctx->stream, W, U, ev,
                       m, c0, nt);
    LAUNCH_CHECK(ctx);
    return BRPA_OK;
}

// x *= a (n elements)
__global__ void k_scale(double *__restrict__ x, double a, int64_t n) {
    int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e < n) x[e] *= a;
}

int launch_scale(brpa_ctx *ctx, double *x, double a, int64_t n) {
    if (n <= 0) return BRPA_OK;
    hipLaunchKernelGGL(k_scale, dim3(nblocks(n)), dim3(256), 0, ctx->stream, x, a, n);
    LAUNCH_CHECK(ctx);
    return BRPA_OK;
}

// M(r, r) += a for the n x n matrix M
__global__ void k_add_diag(double *__restrict__ M, int64_t n, double a) {
    int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (r < n) M[r + n * r] += a;
}

int launch_add_diag(brpa_ctx *ctx, double *M, int64_t n, double a) {
    if (n <= 0) return BRPA_OK;
    hipLaunchKernelGGL(k_add_diag, dim3(nblocks(n)), dim3(256), 0, ctx->stream, M, n, a);
    LAUNCH_CHECK(ctx);
    return BRPA_OK;
}

// M(r, c) = 0 for r > c (n x n, column-major): keeps the upper triangle
__global__ void k_zero_lower(double *__restrict__ M, int64_t n) {
    int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= n * n) return;
    const int64_t c = e / n, r = e - c * n;
    if (r > c) M[e] = 0.0;
}

int launch_zero_lower(brpa_ctx *ctx, double *M, int64_t n) {
    if (n <= 0) return BRPA_OK;
    hipLaunchKernelGGL(k_zero_lower, dim3(nblocks(n * n)), dim3(256), 0, ctx->stream, M, n);
    LAUNCH_CHECK(ctx);
    return BRPA_OK;
}

// SOSEX pair contraction (rpa_CC_EcExchange1, rpa/rpa_CC_Exchange.f90:100-116):
// part[b * nchunk + c] = sum over chunk c of W1_b(P,mu) W2_b(P,mu) B[mu], for
// batch b of npair (N x nt) matrix pairs (stride ld between batches).  Double-
// double per thread and per tree level, deterministic.
constexpr int SOSEX_CHUNK = 65536;

__global__ __launch_bounds__(256) void k_pair_dot(const double *__restrict__ W1,
                                                 const double *__restrict__ W2,
                                                 const double *__restrict__ B, int64_t n_rows,
                                                 int64_t nt, int64_t stride, int nchunk,
                                                 double *__restrict__ part) {
    __shared__ double sh[256], sl[256];
    const int bpair = blockIdx.x / nchunk, c = blockIdx.x - bpair * nchunk;
    const double *w1 = W1 + (int64_t)bpair * stride, *w2 = W2 + (int64_t)bpair * stride;
    const int64_t tot = n_rows * nt;
    const int64_t e0 = (int64_t)c * SOSEX_CHUNK, e1 = min(tot, e0 + (int64_t)SOSEX_CHUNK);
    dd_t s = {0.0, 0.0};
    for (int64_t e = e0 + threadIdx.x; e < e1; e += 256) {
        const double p = w1[e] * w2[e];
        s = dd_add(s, p * B[e / n_rows]);
    }
    sh[threadIdx.x] = s.hi;
    sl[threadIdx.x] = s.lo;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (threadIdx.x < k) {
            dd_t x = {sh[threadIdx.x], sl[threadIdx.x]};
            x = dd_add(x, (dd_t){sh[threadIdx.x + k], sl[threadIdx.x + k]});
            sh[threadIdx.x] = x.hi;
            sl[threadIdx.x] = x.lo;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * (int64_t)blockIdx.x] = sh[0];
        part[2 * (int64_t)blockIdx.x + 1] = sl[0];
    }
}

int sosex_nchunk(int64_t n_rows, int64_t nt) {
    return (int)((n_rows * nt + SOSEX_CHUNK - 1) / SOSEX_CHUNK);
}

int launch_pair_dot(brpa_ctx *ctx, const double *W1, const double *W2, const double *B,
                    int64_t n_rows, int64_t nt, int64_t stride, int npair, double *part) {
    const int nchunk = sosex_nchunk(n_rows, nt);
    if (npair <= 0 || nchunk <= 0) return BRPA_OK;
    hipLaunchKernelGGL(k_pair_dot, dim3((unsigned)(npair * nchunk)), dim3(256), 0, ctx->stream,
                       W1, W2, B, n_rows, nt, stride, nchunk, part);
    LAUNCH_CHECK(ctx);
    return BRPA_OK;
}

// Adjoint accumulation for one frequency (column-major, Rc(P, ai) and
// X = M Rc with M = (1 + Pi_k)^{-1}):
//   dR(P, ai) += 2 c_k(ai) coef (X - Rc)(P, ai)
//   dd(ai)    += coef dc_k(ai) sum_P (X - Rc)(P, ai) Rc(P, ai)
// coef = w_k/(2 pi).  One wave per column ai, fixed-order wave reduction.
__global__ __launch_bounds__(256) void k_grad_acc(const double *__restrict__ X,
                                                 const double *__restrict__ Rc, int64_t N,
                                                 int64_t K, const double *__restrict__ c,
                                                 const double *__restrict__ dc, double coef,
                                                 double *__restrict__ dR, double *__restrict__ dd) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t ai = (int64_t)blockIdx.x * 4 + wave;
    if (ai >= K) return;
    const double *x = X + ai * N, *r = Rc + ai * N;
    double *g = dR + ai * N;
    const double s2 = 2.0 * c[ai] * coef;
    double acc = 0.0;
    for (int64_t P = lane; P < N; P += 64) {
        const double d = x[P] - r[P];
        g[P] = fma(s2, d, g[P]);
        acc = fma(d, r[P], acc);
    }
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if (lane == 0) dd[ai] += coef * dc[ai] * acc;
}

int launch_grad_acc(brpa_ctx *ctx, const double *X, const double *Rc, int64_t N, int64_t K,
                    const double *c, const double *dc, double coef, double *dR, double *dd) {
    if (K <= 0) return BRPA_OK;
    hipLaunchKernelGGL(k_grad_acc, dim3((unsigned)((K + 3) / 4)), dim3(256), 0, ctx->stream, X, Rc,
                       N, K, c, dc, coef, dR, dd);
    LAUNCH_CHECK(ctx);
    return BRPA_OK;
}

// fill n doubles with a value (graph-capture safe)
__global__ void k_fill(double *p, int64_t n, double v) {
    int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = v;
}

int launch_fill(brpa_ctx *ctx, double *p, int64_t n, double v) {
    if (n <= 0) return BRPA_OK;
    hipLaunchKernelGGL(k_fill, dim3(nblocks(n)), dim3(256), 0, ctx->stream, p, n, v);
    LAUNCH_CHECK(ctx);
    return BRPA_OK;
}

// EigRPA_mu = 2 A_mu sum_P RU(P, mu)^2 with RU = R^T V (m x n), one work-group
// per column (rpa_T2_EigRPA, rpa/rpa_CC_Doubles.f90:987-1009).
__global__ __launch_bounds__(256) void k_eig_rpa(const double *__restrict__ RU, int64_t m,
                                                 const double *__restrict__ A,
                                                 double *__restrict__ out) {
    __shared__ double sh[256];
    const double *col = RU + (int64_t)blockIdx.x * m;
    double s = 0.0;
    for (int64_t r = threadIdx.x; r < m; r += 256) s = fma(col[r], col[r], s);
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (threadIdx.x < k) sh[threadIdx.x] += sh[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = 2.0 * sh[0] * A[blockIdx.x];
}

int launch_eig_rpa(brpa_ctx *ctx, const double *RU, int64_t m, int64_t n, const double *A,
                   double *out) {
    if (n <= 0) return BRPA_OK;
    hipLaunchKernelGGL(k_eig_rpa, dim3((unsigned)n), dim3(256), 0, ctx->stream, RU, m, A, out);
    LAUNCH_CHECK(ctx);
    return BRPA_OK;
}

// out(c, r) = in(r, c): in is m x n column-major, out n x m column-major.  32 x 32
// tiles through LDS (one padding column against bank conflicts); 256 threads
// cover a tile in 4 passes of 8 rows.
__global__ __launch_bounds__(256) void k_transpose(const double *__restrict__ in, int64_t m,
                                                   int64_t n, double *__restrict__ out) {
    __shared__ double t[32][33];
    const int64_t r0 = (int64_t)blockIdx.x * 32, c0 = (int64_t)blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int j = ty; j < 32; j += 8) {
        const int64_t r = r0 + tx, c = c0 + j;
        if (r < m && c < n) t[j][tx] = in[r + m * c];
    }
    __syncthreads();
    for (int j = ty; j < 32; j += 8) {
        const int64_t c = c0 + tx, r = r0 + j;
        if (r < m && c < n) out[c + n * r] = t[tx][j];
    }
}

int launch_transpose(brpa_ctx *ctx, const double *in, int64_t m, int64_t n, double *out) {
    if (m <= 0 || n <= 0) return BRPA_OK;
    const int64_t gx = (m + 31) / 32, gy = (n + 31) / 32;
    if (gy > 65535)
        return brpa_fail(ctx, BRPA_ERR_ARG, "transpose: %lld columns exceed the grid limit",
                         (long long)n);
    hipLaunchKernelGGL(k_transpose, dim3((unsigned)gx, (unsigned)gy), dim3(256), 0, ctx->stream, in,
                       m, n, out);
    LAUNCH_CHECK(ctx);
    return BRPA_OK;
}

// out(r, c) = s[c] * in(r, c) for an m x n column-major matrix (out may alias in)
__global__ void k_col_scale(double *out, const double *in, const double *__restrict__ s, int64_t m,
                            int64_t n) {
    int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= m * n) return;
    out[e] = s[e / m] * in[e];
}

int launch_col_scale(brpa_ctx *ctx, double *out, const double *in, const double *s, int64_t m,
                     int64_t n) {
    if (m * n <= 0) return BRPA_OK;
    hipLaunchKernelGGL(k_col_scale, dim3(nblocks(m * n)), dim3(256), 0, ctx->stream, out, in, s, m,
                       n);
    LAUNCH_CHECK(ctx);
    return BRPA_OK;
}

// The two matrices of rpa_CC_Exchange_MBPT3_1 (rpa/rpa_CC_Exchange.f90:248-262),
// n x n column-major, element (m, k) at m + n k:
//   Am(m, k) = (A_m + A_k) D(m, k),   Bm(m, k) = D(m, k) + 2 U(m, k) + 4 U(m, k) A_k
__global__ void k_mbpt3_ab(const double *__restrict__ D, const double *__restrict__ U,
                           const double *__restrict__ A, int64_t n, double *__restrict__ Am,
                           double *__restrict__ Bm) {
    int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= n * n) return;
    const int64_t k = e / n, m = e - k * n;
    const double d = D[e], u = U[e], ak = A[k];
    Am[e] = A[m] * d + d * ak;
    Bm[e] = d + 2.0 * u + 4.0 * u * ak;
}

int launch_mbpt3_ab(brpa_ctx *ctx, const double *D, const double *U, const double *A, int64_t n,
                    double *Am, double *Bm) {
    if (n <= 0) return BRPA_OK;
    hipLaunchKernelGGL(k_mbpt3_ab, dim3(nblocks(n * n)), dim3(256), 0, ctx->stream, D, U, A, n, Am,
                       Bm);
    LAUNCH_CHECK(ctx);
    return BRPA_OK;
}

// LAPACK lower-packed storage of symmetric n x n matrices (dtrttp / dtpttr with
// uplo 'L', the reference's real_Pack / real_Unpack, linear-algebra/
// real_linalg.f90:737-773): column j holds rows i >= j at s_j + (i - j),
// s_j = j n - j (j - 1) / 2.  One work-group per (column j, matrix w).
__device__ __forceinline__ int64_t packed_col_start(int64_t j, int64_t n) {
    return j * n - j * (j - 1) / 2;
}

// AP (+)= lower triangle of A (column-major, element (i, j) at i + n j)
__global__ __launch_bounds__(256) void k_pack_lower(const double *__restrict__ A, int64_t n,
                                                    double *__restrict__ AP, int accumulate) {
    const int64_t j = blockIdx.x, w = blockIdx.y;
    const int64_t np = n * (n + 1) / 2;
    const double *a = A + w * n * n + j * n;
    double *ap = AP + w * np + packed_col_start(j, n) - j;
    for (int64_t i = j + threadIdx.x; i < n; i += 256) ap[i] = accumulate ? ap[i] + a[i] : a[i];
}

// A = symmetric matrix of the lower-packed AP (both triangles written)
__global__ __launch_bounds__(256) void k_unpack_lower(const double *__restrict__ AP, int64_t n,
                                                      double *__restrict__ A) {
    const int64_t j = blockIdx.x, w = blockIdx.y;
    const int64_t np = n * (n + 1) / 2;
    const double *ap = AP + w * np + packed_col_start(j, n) - j;
    double *a = A + w * n * n;
    for (int64_t i = j + threadIdx.x; i < n; i += 256) {
        const double v = ap[i];
        a[i + n * j] = v;
        a[j + n * i] = v;
    }
}

int launch_pack_lower(brpa_ctx *ctx, const double *A, int64_t n, int64_t F, double *AP,
                      int accumulate) {
    if (n <= 0 || F <= 0) return BRPA_OK;
    if (F > 65535) return brpa_fail(ctx, BRPA_ERR_ARG, "pack: more than 65535 matrices");
    hipLaunchKernelGGL(k_pack_lower, dim3((unsigned)n, (unsigned)F), dim3(256), 0, ctx->stream, A,
                       n, AP, accumulate);
    LAUNCH_CHECK(ctx);
    return BRPA_OK;
}

int launch_unpack_lower(brpa_ctx *ctx, const double *AP, int64_t n, int64_t F, double *A) {
    if (n <= 0 || F <= 0) return BRPA_OK;
    if (F > 65535) return brpa_fail(ctx, BRPA_ERR_ARG, "unpack: more than 65535 matrices");
    hipLaunchKernelGGL(k_unpack_lower, dim3((unsigned)n, (unsigned)F), dim3(256), 0, ctx->stream,
                       AP, n, A);
    LAUNCH_CHECK(ctx);
    return BRPA_OK;
}

// Standard normal numbers for the random basis of rpa_G: Box-Muller on two
// SplitMix64 counter draws per element, u1 = (z1 >> 11 + 1) / 2^53 in (0, 1],
// u2 = (z2 >> 11) / 2^53; g = sqrt(-2 ln u1) cos(2 pi u2).
__device__ __forceinline__ uint64_t splitmix(uint64_t seed, uint64_t ctr) {
    uint64_t z = seed + ctr * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__global__ void k_gauss(double *__restrict__ G, int64_t n, uint64_t seed) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= n) return;
    const double u1 = (double)((splitmix(seed, 2 * e + 1) >> 11) + 1) * (1.0 / 9007199254740992.0);
    const double u2 = (double)(splitmix(seed, 2 * e + 2) >> 11) * (1.0 / 9007199254740992.0);
    G[e] = sqrt(-2.0 * log(u1)) * cos(6.283185307179586476925 * u2);
}

int launch_gauss(brpa_ctx *ctx, double *G, int64_t n, uint64_t seed) {
    if (n <= 0) return BRPA_OK;
    hipLaunchKernelGGL(k_gauss, dim3(nblocks(n)), dim3(256), 0, ctx->stream, G, n, seed);
    LAUNCH_CHECK(ctx);
    return BRPA_OK;
}

// M = diag(d) (n x n, column-major)
__global__ void k_set_diag_only(double *__restrict__ M, const double *__restrict__ d, int64_t n) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= n * n) return;
    const int64_t j = e / n, i = e - j * n;
    M[e] = (i == j) ? d[i] : 0.0;
}

int launch_set_diag_only(brpa_ctx *ctx, double *M, const double *d, int64_t n) {
    if (n <= 0) return BRPA_OK;
    hipLaunchKernelGGL(k_set_diag_only, dim3(nblocks(n * n)), dim3(256), 0, ctx->stream, M, d, n);
    LAUNCH_CHECK(ctx);
    return BRPA_OK;
}
