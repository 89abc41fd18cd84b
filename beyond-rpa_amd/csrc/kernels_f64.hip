// FP64 kernels of the E_RPA hot path on gfx950 (CDNA4).
//
//   k_dtab         c_w(ai) = 4 d/(d^2+u_w^2)           (rpa_core_MO.f90:379-401)
//   k_syrk_f64     Pi_w = R^T diag(c_w) R on v_mfma_f64_16x16x4_f64, one R tile in
//                  LDS shared by 4 waves = 4 frequencies  (dgemm of rpa_core_MO.f90:402)
//   k_set_diag     diag_w (tr Pi, rpa_CC.f90:444-447); Pi_w[P][P] += 1 (448-450)
//   k_potrf_diag   Cholesky + inverse of a 64x64 diagonal block (one wave)
//   k_trsm         L21 = A21 L11^{-T} on MFMA
//   k_chol_update  trailing update A22 -= L21 L21^T on MFMA
//   k_energy       E = sum_w w/(2 pi) (2 sum log L_kk - tr Pi_w)  (rpa_CC.f90:451-453)
//
// All kernels walk a flattened work list over fragments: work-group b belongs
// to fragment f with pre[f] <= b < pre[f+1].
#include "device_common.h"
#include "internal.h"

#define LAUNCH_CHECK(ctx)                                                                      \
    do {                                                                                       \
        hipError_t _e = hipGetLastError();                                                     \
        if (_e != hipSuccess)                                                                  \
            return brpa_fail((ctx), BRPA_ERR_HIP, "kernel launch failed: %s (%s:%d)",          \
                             hipGetErrorString(_e), __FILE__, __LINE__);                       \
    } while (0)

static inline unsigned grid1(int64_t total) { return (unsigned)total; }

typedef double d2v_t __attribute__((ext_vector_type(2)));

// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_dtab(const FragDev *__restrict__ fr,
                                             const int64_t *__restrict__ pre, int nfrag) {
    int64_t b = blockIdx.x;
    int f = find_frag(pre, nfrag, b);
    const FragDev &F = fr[f];
    int64_t e = (b - pre[f]) * 256 + threadIdx.x;
    int64_t tot = (int64_t)F.F * F.K;
    if (e >= tot) return;
    int w = (int)(e / F.K);
    int k = (int)(e - (int64_t)w * F.K);
    int i = k / F.nvir, a = k - i * F.nvir;
    double d = F.evir[a] - F.eocc[i];
    double u = F.freqs[w];
    const double c = 4.0 * d / (d * d + u * u);
    const_cast<double *>(F.dtab)[e] = c;
    if (F.sdtab) const_cast<double *>(F.sdtab)[e] = sqrt(c);
}

int launch_dtab(brpa_ctx *ctx, const FragDev *fr, const int64_t *pre, int nfrag, int64_t total) {
    if (total <= 0) return BRPA_OK;
    hipLaunchKernelGGL(k_dtab, dim3(grid1(total)), dim3(256), 0, ctx->stream, fr, pre, nfrag);
    LAUNCH_CHECK(ctx);
    return BRPA_OK;
}

// ---------------------------------------------------------------------------
// FP64 SYRK on MFMA.  One work-group = one 64x64 tile pair (tp >= tq) and a
// group of 4 frequencies; wave w computes the whole tile for frequency
// 4*g + w from the R tiles staged once in LDS.
//
// MFMA v_mfma_f64_16x16x4_f64 operand layout (lane l):
//   A[m = l%16][k = l/16], B[k = l/16][n = l%16],
//   D: 4 doubles, D[r] = C[m = l/16 + 4*r][n = l%16]  (probed on gfx950; the
//   f32 shapes use 4*(l/16) + r instead).
constexpr int KB = 16;        // rows of R per LDS stage
constexpr int LDS_PAD = 16;   // +16 doubles per row: rows r and r+1 land 32 banks apart

__device__ __forceinline__ void syrk_f64_tile(const double *__restrict__ R, int K, int N,
                                              const double *__restrict__ dtab, int F, int w0,
                                              double *__restrict__ Pi, int tp, int tq) {
    __shared__ double sA[KB][SYRK_T + LDS_PAD];
    __shared__ double sB[KB][SYRK_T + LDS_PAD];
    __shared__ double sD[4][KB];
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int p0 = tp * SYRK_T, q0 = tq * SYRK_T;
    const int w = w0 + wave;

    d4_t acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = (d4_t){0.0, 0.0, 0.0, 0.0};

    double ra[4], rb[4], rd = 0.0;
    auto gload = [&](int k0) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            int e = tid + 256 * i;
            int r = e >> 6, c = e & 63;
            int k = k0 + r;
            bool kin = k < K;
            ra[i] = (kin && p0 + c < N) ? R[(int64_t)k * N + p0 + c] : 0.0;
            rb[i] = (kin && q0 + c < N) ? R[(int64_t)k * N + q0 + c] : 0.0;
        }
        if (tid < 4 * KB) {
            int ww = w0 + (tid / KB), k = k0 + (tid % KB);
            rd = (ww < F && k < K) ? dtab[(int64_t)ww * K + k] : 0.0;
        }
    };
    auto lstore = [&]() {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            int e = tid + 256 * i;
            int r = e >> 6, c = e & 63;
            sA[r][c] = ra[i];
            sB[r][c] = rb[i];
        }
        if (tid < 4 * KB) sD[tid / KB][tid % KB] = rd;
    };

    gload(0);
    for (int k0 = 0; k0 < K; k0 += KB) {
        __syncthreads();
        lstore();
        __syncthreads();
        if (k0 + KB < K) gload(k0 + KB);
#pragma unroll
        for (int kk = 0; kk < KB; kk += 4) {
            const int kr = kk + (lane >> 4);
            const double dv = sD[wave][kr];
            double a[4], bv[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) a[i] = sA[kr][16 * i + (lane & 15)] * dv;
#pragma unroll
            for (int j = 0; j < 4; ++j) bv[j] = sB[kr][16 * j + (lane & 15)];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], bv[j], acc[i][j], 0, 0, 0);
        }
    }

    if (w >= F) return;
    double *C = Pi + (int64_t)w * N * N;
    const bool diag = (tp == tq);
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                int p = p0 + 16 * i + (lane >> 4) + 4 * r;
                int q = q0 + 16 * j + (lane & 15);
                if (p < N && q < N) {
                    double v = acc[i][j][r];
                    C[(int64_t)p * N + q] = v;
                    if (!diag) C[(int64_t)q * N + p] = v;
                }
            }
}

// Work-groups per fragment: npairs(ceil(N/64)) * ceil(F/4), pair-major within
// a frequency group.
__global__ __launch_bounds__(256) void k_syrk_f64(const FragDev *__restrict__ fr,
                                                 const int64_t *__restrict__ pre, int nfrag) {
    int64_t b = blockIdx.x;
    int f = find_frag(pre, nfrag, b);
    const FragDev &F = fr[f];
    int64_t lb = b - pre[f];
    int nt = (F.N + SYRK_T - 1) / SYRK_T;
    int npairs = nt * (nt + 1) / 2;
    int g = (int)(lb / npairs), pair = (int)(lb - (int64_t)g * npairs);
    int tp, tq;
    tri_decode(pair, tp, tq);
    syrk_f64_tile(F.R, F.K, F.N, F.dtab, F.F, 4 * g, F.Pi, tp, tq);
}

int launch_syrk_f64(brpa_ctx *ctx, const FragDev *fr, const int64_t *pre, int nfrag, int64_t total) {
    if (total <= 0) return BRPA_OK;
    hipLaunchKernelGGL(k_syrk_f64, dim3(grid1(total)), dim3(256), 0, ctx->stream, fr, pre, nfrag);
    LAUNCH_CHECK(ctx);
    return BRPA_OK;
}

// ---------------------------------------------------------------------------
// diag_w[P] and Pi_w[P][P] = diag_w[P] + 1.  The diagonal is either the one the
// FP64 SYRK (or the caller) computed, or, for the split-BF16 form, the exact
// double-double sum of the per-block partials written by k_split.
// Work-groups per fragment: ceil(F*N/256).
__global__ __launch_bounds__(256) void k_set_diag(const FragDev *__restrict__ fr,
                                                 const int64_t *__restrict__ pre, int nfrag,
                                                 int mode) {
    int64_t b = blockIdx.x;
    int f = find_frag(pre, nfrag, b);
    const FragDev &F = fr[f];
    int64_t e = (b - pre[f]) * 256 + threadIdx.x;
    if (e >= (int64_t)F.F * F.N) return;
    int w = (int)(e / F.N), P = (int)(e - (int64_t)w * F.N);
    double *pp = F.Pi + (int64_t)w * F.N * F.N + (int64_t)P * F.N + P;
    // mode 0: Pi's own diagonal; 1: sum of the split-BF16 double-double block
    // partials; 2: diag already holds the exact value (split-int8 column pass)
    double d;
    if (mode == 1) {
        dd_t s = {0.0, 0.0};
        for (int blk = 0; blk < F.nblk; ++blk) {
            const double *q = F.dpart + (((int64_t)w * F.nblk + blk) * F.N + P) * 2;
            s = dd_add(s, (dd_t){q[0], q[1]});
        }
        d = s.hi + s.lo;
    } else if (mode == 2) {
        d = F.diag[e];
    } else {
        d = *pp;
    }
    F.diag[e] = d;
    *pp = d + 1.0;
}

int launch_set_diag(brpa_ctx *ctx, const FragDev *fr, const int64_t *pre, int nfrag, int64_t total,
                    int mode) {
    if (total <= 0) return BRPA_OK;
    hipLaunchKernelGGL(k_set_diag, dim3(grid1(total)), dim3(256), 0, ctx->stream, fr, pre, nfrag,
                       mode);
    LAUNCH_CHECK(ctx);
    return BRPA_OK;
}

// ---------------------------------------------------------------------------
// Blocked right-looking Cholesky of A_w = 1 + Pi_w (lower triangle, row-major),
// three launches per 64-column step j (j0 = 64 j):
//   k_potrf_diag   one wave per matrix: factors the diagonal block in LDS,
//                  writes its log-det partial and L11^{-1} (own buffer; A11
//                  itself is never written back, so no work-group of the same
//                  step can read a half-factored block)
//   k_trsm         L21 = A21 L11^{-T} as a 64x64x64 MFMA product per row chunk
//   k_chol_update  A22 -= L21 L21^T per lower 64x64 tile pair, MFMA
// Work-groups per fragment: F (diag), F*(nchunks-1) (trsm), F*npairs (update).

// C(64x64) = X Y^T with X, Y 64x64 in LDS; wave (qi,qj) owns a 32x32 quadrant,
// acc[i][j][r] = C[32 qi + 16 i + l/16 + 4 r][32 qj + 16 j + l%16].
// X Y^T of two 64 x 64 LDS panels on MFMA, 4 waves of 32 x 32 (wave = tid / 64 < 4).
__device__ __forceinline__ void mma64_xyt(const double (*sX)[CHOL_NB + 1],
                                          const double (*sY)[CHOL_NB + 1], d4_t acc[2][2],
                                          int tid) {
    const int wave = tid >> 6, lane = tid & 63;
    const int qi = wave >> 1, qj = wave & 1;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = (d4_t){0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
    for (int kk = 0; kk < CHOL_NB; kk += 4) {
        const int kc = kk + (lane >> 4);
        double a[2], bv[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) a[i] = sX[32 * qi + 16 * i + (lane & 15)][kc];
#pragma unroll
        for (int j = 0; j < 2; ++j) bv[j] = sY[32 * qj + 16 * j + (lane & 15)][kc];
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], bv[j], acc[i][j], 0, 0, 0);
    }
}

typedef __attribute__((address_space(1))) const double gdouble_t;
typedef __attribute__((address_space(1))) double gdouble_mut_t;

template <int ROWS, int THREADS, int NP>
__device__ __forceinline__ void load_panels64(double (*const *s)[CHOL_NB + 1],
                                              const double *__restrict__ A, int64_t ld,
                                              const int *row0, int rowmax, int col0, int colmax,
                                              int tid) {
    constexpr int PER = ROWS * CHOL_NB / THREADS;
    constexpr int RSTEP = THREADS / CHOL_NB;
    const int c = tid & (CHOL_NB - 1), r0 = tid / CHOL_NB;
    const gdouble_t *G = (const gdouble_t *)A;
    double v[NP][PER];
#pragma unroll
    for (int p = 0; p < NP; ++p)
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int r = r0 + RSTEP * u;
            const bool ok = row0[p] + r < rowmax && col0 + c < colmax;
            const int64_t idx = ok ? (int64_t)(row0[p] + r) * ld + col0 + c : 0;
            v[p][u] = G[idx];
            v[p][u] = ok ? v[p][u] : 0.0;
        }
#pragma unroll
    for (int p = 0; p < NP; ++p)
#pragma unroll
        for (int u = 0; u < PER; ++u) s[p][r0 + RSTEP * u][c] = v[p][u];
}

template <int ROWS, int THREADS>
__device__ __forceinline__ void load_panel64(double (*s)[CHOL_NB + 1], const double *__restrict__ A,
                                             int64_t ld, int row0, int rowmax, int col0, int colmax,
                                             int tid) {
    double(*sp[1])[CHOL_NB + 1] = {s};
    const int r[1] = {row0};
    load_panels64<ROWS, THREADS, 1>(sp, A, ld, r, rowmax, col0, colmax, tid);
}

// ---------------------------------------------------------------------------
// Panel body of the blocked Cholesky.  NT is the work-group size; every thread
// calls the body (the barriers are uniform) and the work is guarded by the
// thread index.

// L21 chunk r (rows j0 + 64 r ..) of step `step`: L21 = A21 L11^{-T} on MFMA
// (waves 0..3).  sB, sL: [64][65] LDS.
template <int NT>
__device__ __forceinline__ void trsm_body(const FragDev &F, int w, int step, int r, double (*sB)[CHOL_NB + 1],
                          double (*sL)[CHOL_NB + 1], int tid) {
    static_assert(NT >= 256 && NT % 256 == 0, "work-group too small");
    const int N = F.N;
    const int j0 = step * CHOL_NB;
    double *A = F.Pi + (int64_t)w * N * N;
    const double *linv = F.linv + (int64_t)w * CHOL_NB * CHOL_NB;
    const int R0 = j0 + CHOL_NB * r;
    load_panel64<CHOL_NB, NT>(sB, A, N, R0, N, j0, N, tid);
    load_panel64<CHOL_NB, NT>(sL, linv, CHOL_NB, 0, CHOL_NB, 0, CHOL_NB, tid);
    __syncthreads();
    if (tid < 256) {
        d4_t acc[2][2];
        mma64_xyt(sB, sL, acc, tid);   // L21[i][c] = sum_k A21[i][k] Linv[c][k]
        const int wave = tid >> 6, lane = tid & 63, qi = wave >> 1, qj = wave & 1;
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    int row = R0 + 32 * qi + 16 * i + (lane >> 4) + 4 * q;
                    int col = 32 * qj + 16 * j + (lane & 15);
                    if (row < N) A[(int64_t)row * N + j0 + col] = acc[i][j][q];
                }
    }
}

// Number of L21 chunks of step j.
__device__ __forceinline__ int trsm_chunks(int N, int j) {
    const int j0 = j * CHOL_NB;
    return N > j0 + CHOL_NB ? (N - j0 + CHOL_NB - 1) / CHOL_NB - 1 : 0;
}

// ---------------------------------------------------------------------------
// Per-step kernels (flat work lists over fragments and frequencies)

// lane `l`'s value of v, broadcast to the whole wave (two 32-bit readlanes)
__device__ __forceinline__ double wave_bcast(double v, int l) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
    return __hiloint2double(hi, lo);
}

// Diagonal block of step `step` of matrix w with one wave, left-looking: lane
// r keeps row r of the 64 x 64 block in registers.  At column c every lane
// subtracts the dot product of its row with row c (rows of L are complete up
// to column c - 1 and are read from LDS as broadcast ds_read_b128, four
// partial sums), the pivot comes from lane c by readlane, and the finished
// column goes to LDS.  No barrier inside the factorisation (one wave).  Then
// column `lane` of L11^{-1} by forward substitution over the rows of L in LDS
// (reciprocal diagonal precomputed; uniform loops, x_k = 0 above the lane's
// column).  Outputs: linv, log-det partial, status, optional L11 write-back.
__global__ __launch_bounds__(64) void k_potrf_diag(const FragDev *__restrict__ fr,
                                                  const int64_t *__restrict__ pre, int nfrag,
                                                  int step) {
    // row pitch 66 doubles: 16-byte aligned rows; broadcast reads are conflict-free
    constexpr int LP = CHOL_NB + 2;
    __shared__ __attribute__((aligned(16))) double L[CHOL_NB][LP];
    __shared__ double rdiag[CHOL_NB];
    __shared__ int64_t s_pre[FIND_LDS_MAX + 1];
    const int64_t b = blockIdx.x;
    const int f = find_frag_lds(pre, nfrag, b, s_pre);
    const int w = (int)(b - (nfrag <= FIND_LDS_MAX ? s_pre[f] : pre[f]));
    const FragDev &F = fr[f];
    const int lane = threadIdx.x;
    const int N = F.N;
    const int j0 = step * CHOL_NB;
    const int nbj = min(CHOL_NB, N - j0);
    double *A = F.Pi + (int64_t)w * N * N;
    // row `lane` of the block straight into registers, lower triangle only (the
    // other triangle is never read); rows/cols past nbj padded with the identity
    double a[CHOL_NB];
    const gdouble_t *Arow = (const gdouble_t *)(A + (int64_t)(j0 + min(lane, nbj - 1)) * N + j0);
#pragma unroll
    for (int c = 0; c < CHOL_NB; ++c) {
        const bool in = lane < nbj && c < nbj;
        const double v = (in && c <= lane) ? Arow[min(c, nbj - 1)] : 0.0;
        a[c] = in ? v : (lane == c ? 1.0 : 0.0);
    }
    int bad = 0;
#pragma unroll
    for (int c = 0; c < CHOL_NB; ++c) {
        if (c > 0) {
            // a[c] -= sum_{k<c} L[lane][k] L[c][k]; row c of L from LDS
            double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
#pragma unroll
            for (int k = 0; k < c; k += 4) {
                const d2v_t u = *reinterpret_cast<const d2v_t *>(&L[c][k]);
                s0 = fma(a[k], u.x, s0);
                if (k + 1 < c) s1 = fma(a[k + 1], u.y, s1);
                if (k + 2 < c) {
                    const d2v_t u2 = *reinterpret_cast<const d2v_t *>(&L[c][k + 2]);
                    s2 = fma(a[k + 2], u2.x, s2);
                    if (k + 3 < c) s3 = fma(a[k + 3], u2.y, s3);
                }
            }
            a[c] -= (s0 + s1) + (s2 + s3);
        }
        double piv = wave_bcast(a[c], c);
        if (!(piv > 0.0) || !isfinite(piv)) {
            bad = 1;
            piv = 1.0;
        }
        const double d = sqrt(piv);
        const double rd = 1.0 / d;
        // lanes above row c hold values of the upper triangle, never read
        a[c] = (lane == c) ? d : a[c] * rd;
        L[lane][c] = a[c];
        if (lane == c) rdiag[c] = rd;
        __builtin_amdgcn_sched_barrier(0);
    }
    __syncthreads();
    // column `lane` of X = L11^{-1}: x_i = (delta_i,lane - sum_{k<i} L_ik x_k) / L_ii,
    // x_i = 0 for i < lane; a[] now holds x
#pragma unroll
    for (int i = 0; i < CHOL_NB; ++i) {
        double s0 = (i == lane) ? 1.0 : 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
#pragma unroll
        for (int k = 0; k < i; k += 4) {
            const d2v_t u = *reinterpret_cast<const d2v_t *>(&L[i][k]);
            s0 = fma(-u.x, a[k], s0);
            if (k + 1 < i) s1 = fma(-u.y, a[k + 1], s1);
            if (k + 2 < i) {
                const d2v_t u2 = *reinterpret_cast<const d2v_t *>(&L[i][k + 2]);
                s2 = fma(-u2.x, a[k + 2], s2);
                if (k + 3 < i) s3 = fma(-u2.y, a[k + 3], s3);
            }
        }
        a[i] = ((s0 + s1) + (s2 + s3)) * rdiag[i];
        __builtin_amdgcn_sched_barrier(0);
    }
    double *linv = F.linv + (int64_t)w * CHOL_NB * CHOL_NB;
#pragma unroll
    for (int i = 0; i < CHOL_NB; ++i) linv[i * CHOL_NB + lane] = a[i];
    if (F.keep_l11) {
        // the factored block back into Pi (lower triangle and diagonal), one
        // row per pass with the lanes along the columns; no other work item
        // reads this diagonal block
        for (int r = 0; r < nbj; ++r)
            if (lane <= r) A[(int64_t)(j0 + r) * N + j0 + lane] = L[r][lane];
    }
    // log-det partial: one log per lane, then a fixed butterfly in double-double
    dd_t sl = {lane < nbj ? log(L[lane][lane]) : 0.0, 0.0};
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const dd_t o = {__shfl_xor(sl.hi, off), __shfl_xor(sl.lo, off)};
        sl = dd_add(sl, o);
    }
    if (lane == 0) {
        double *part = F.ldpart + ((int64_t)w * F.nsteps + step) * 2;
        part[0] = sl.hi;
        part[1] = sl.lo;
    }
    if (lane == 0 && bad) F.status[w] = 1;
}

int launch_potrf_diag(brpa_ctx *ctx, const FragDev *fr, const int64_t *pre, int nfrag,
                      int64_t total, int step) {
    if (total <= 0) return BRPA_OK;
    hipLaunchKernelGGL(k_potrf_diag, dim3(grid1(total)), dim3(64), 0, ctx->stream, fr, pre, nfrag,
                       step);
    LAUNCH_CHECK(ctx);
    return BRPA_OK;
}

__global__ __launch_bounds__(256) void k_trsm(const FragDev *__restrict__ fr,
                                             const int64_t *__restrict__ pre, int nfrag, int step) {
    __shared__ double sB[CHOL_NB][CHOL_NB + 1];
    __shared__ double sL[CHOL_NB][CHOL_NB + 1];
    __shared__ int64_t s_pre[FIND_LDS_MAX + 1];
    const int64_t b = blockIdx.x;
    const int f = find_frag_lds(pre, nfrag, b, s_pre);
    const FragDev &F = fr[f];
    const int nsub = trsm_chunks(F.N, step);
    const int64_t lb = b - (nfrag <= FIND_LDS_MAX ? s_pre[f] : pre[f]);
    const int w = (int)(lb / nsub), r = 1 + (int)(lb - (int64_t)w * nsub);
    trsm_body<256>(F, w, step, r, sB, sL, threadIdx.x);
}

int launch_trsm(brpa_ctx *ctx, const FragDev *fr, const int64_t *pre, int nfrag, int64_t total,
                int step) {
    if (total <= 0) return BRPA_OK;
    hipLaunchKernelGGL(k_trsm, dim3(grid1(total)), dim3(256), 0, ctx->stream, fr, pre, nfrag, step);
    LAUNCH_CHECK(ctx);
    return BRPA_OK;
}

constexpr int UPD_T = 128;   // trailing-update super-tile edge
constexpr int UPD_KH = 16;   // panel columns per LDS phase

// Trailing update of one 128 x 128 super-tile (TI >= TJ in 128-row units of the
// trailing matrix that starts at t0 = j0 + 64): C -= P_i P_j^T over the panel
// columns of the current outer panel (see below), C preloaded into the
// accumulators.  Four waves of 64 x 64 (two work-groups per CU, two waves per
// SIMD); waves entirely above the diagonal of a diagonal super-tile skip their
// products; element writes keep col <= row.  The panels pass through LDS in
// phases of 16 columns as [row][k] (pitch 18 doubles), double-buffered with one
// barrier per phase; every thread moves pairs of consecutive columns (one
// ds_write_b128), and in a phase lane group l/16 supplies the k values
// 4 (l/16) .. +3 of its row to the four MFMA steps (a permutation of the
// phase's k, the same for both operands), two ds_read_b128 per 16-row block.
// The MFMA takes the P_i fragment first, so the accumulator columns run along
// the lanes and the row-major C is read and written coalesced.
template <bool STRIP>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_chol_update(
    const FragDev *__restrict__ fr, const int64_t *__restrict__ pre, int nfrag, int step) {
    // STRIP: the 64-column strip of a step inside an outer panel; the 128 x 64
    // tile is split over the four waves as 2 x 2 blocks of 64 x 32.  Otherwise
    // 128 x 128 tiles, four waves of 64 x 64.
    constexpr int WC = STRIP ? 32 : 64;    // columns per wave
    constexpr int NJ = WC / 16;            // 16-column MFMA blocks per wave
    constexpr int KPU = UPD_KH + 2;
    __shared__ __attribute__((aligned(16))) double sI[2][UPD_T * KPU];
    __shared__ __attribute__((aligned(16))) double sJ[2][UPD_T * KPU];
    __shared__ int64_t s_pre[FIND_LDS_MAX + 1];
    int64_t b = blockIdx.x;
    int f = find_frag_lds(pre, nfrag, b, s_pre);
    const FragDev &F = fr[f];
    const int N = F.N;
    const int j0 = step * CHOL_NB;
    const int t0 = j0 + CHOL_NB;
    const int nt = (N - t0 + UPD_T - 1) / UPD_T;
    // Outer panels of CHOL_OUTER steps (256 columns): step s of an outer panel
    // (s = step % CHOL_OUTER) updates, for s < CHOL_OUTER - 1, only the next
    // 64-column strip (rows >= t0, columns t0 .. t0+63) with the panels of
    // steps step-s .. step (rank 64 (s+1): the strip has seen no panel of this
    // outer panel yet); the last step updates the rest of the trailing matrix
    // with all CHOL_OUTER panels (rank 256), so A22 is read and written once
    // per outer panel.  Work items: nt row tiles (strip) or nt(nt+1)/2 tile
    // pairs (rest).
    const int s_in = step % CHOL_OUTER;
    constexpr bool strip = STRIP;
    const int kbase = j0 - s_in * CHOL_NB;    // first panel column
    const int nph = (s_in + 1) * CHOL_NB / UPD_KH;
    const int colmax = strip ? min(N, t0 + CHOL_NB) : N;   // C columns < colmax
    const int nitems = strip ? nt : nt * (nt + 1) / 2;
    int64_t lb = b - (nfrag <= FIND_LDS_MAX ? s_pre[f] : pre[f]);
    const int w = (int)(lb / nitems), item = (int)(lb - (int64_t)w * nitems);
    int ti, tj;
    if (strip) {
        ti = item;
        tj = 0;
    } else {
        tri_decode(item, ti, tj);
    }
    double *A = F.Pi + (int64_t)w * N * N;
    const gdouble_t *G = (const gdouble_t *)A;
    const int Ri = t0 + UPD_T * ti, Rj = t0 + UPD_T * tj;
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int wm = wave >> 1, wn = wave & 1;   // rows 64*wm.., cols WC*wn..
    const int r0 = 64 * wm, c0 = WC * wn;
    const bool active = !(ti == tj && c0 > r0 + 63) && Ri + r0 < N && Rj + c0 < colmax;

    // phase h (h < nph): rows Ri.. / Rj.. (128 each) x columns kbase + 16 h .. + 16
    // (all < t0 <= N); pair q = tid + 256 u (u < 4): row q / 8, columns
    // 2 (q % 8) + {0, 1}
    d2v_t vi[4], vj[4];
    auto gload = [&](int h) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int q = tid + 256 * u;
            const int r = q >> 3, col = kbase + UPD_KH * h + 2 * (q & 7);
            const bool oki = Ri + r < N, okj = Rj + r < N;
            const gdouble_t *pi = G + (oki ? (int64_t)(Ri + r) * N + col : 0);
            const gdouble_t *pj = G + (okj ? (int64_t)(Rj + r) * N + col : 0);
            const double i0 = pi[0], i1 = pi[1], ja = pj[0], jb = pj[1];
            vi[u].x = oki ? i0 : 0.0;
            vi[u].y = oki ? i1 : 0.0;
            vj[u].x = okj ? ja : 0.0;
            vj[u].y = okj ? jb : 0.0;
        }
    };
    auto lstore = [&](int buf) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int q = tid + 256 * u;
            const int o = (q >> 3) * KPU + 2 * (q & 7);
            *reinterpret_cast<d2v_t *>(&sI[buf][o]) = vi[u];
            *reinterpret_cast<d2v_t *>(&sJ[buf][o]) = vj[u];
        }
    };

    gload(0);
    lstore(0);
    __syncthreads();
    if (nph > 1) gload(1);
    // the C tile into the accumulators: D[q] = C(row = 16 i + lane/16 * 4 + q ...)
    // in MFMA order, element (row r0 + 16 i + 4 (lane/16) + q, col c0 + 16 j + lane%16)
    // (out-of-range / upper elements read A[0], then 0)
    d4_t acc[4][NJ];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < NJ; ++j)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int row = Ri + r0 + 16 * i + (lane >> 4) + 4 * q;
                const int col = Rj + c0 + 16 * j + (lane & 15);
                const bool ok = active && row < N && col < colmax && col <= row;
                const double v = G[ok ? (int64_t)row * N + col : 0];
                acc[i][j][q] = ok ? v : 0.0;
            }
    const int kq = 4 * (lane >> 4), rl = lane & 15;
    int buf = 0;
    for (int h = 0; h < nph; ++h, buf ^= 1) {
        if (active) {
            // two halves of two MFMA steps (one ds_read_b128 per operand block
            // and half: the register budget)
#pragma unroll
            for (int hh = 0; hh < 2; ++hh) {
                d2v_t a[4], bv[NJ];
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    a[i] = -*reinterpret_cast<const d2v_t *>(
                        &sI[buf][(r0 + 16 * i + rl) * KPU + kq + 2 * hh]);
#pragma unroll
                for (int j = 0; j < NJ; ++j)
                    bv[j] = *reinterpret_cast<const d2v_t *>(
                        &sJ[buf][(c0 + 16 * j + rl) * KPU + kq + 2 * hh]);
#pragma unroll
                for (int st = 0; st < 2; ++st)
#pragma unroll
                    for (int i = 0; i < 4; ++i)
#pragma unroll
                        for (int j = 0; j < NJ; ++j)
                            acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i][st], bv[j][st],
                                                                             acc[i][j], 0, 0, 0);
            }
        }
        // the other buffer was last read before the previous barrier
        if (h + 1 < nph) lstore(buf ^ 1);
        __syncthreads();
        if (h + 2 < nph) gload(h + 2);
    }
    if (!active) return;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < NJ; ++j)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int row = Ri + r0 + 16 * i + (lane >> 4) + 4 * q;
                const int col = Rj + c0 + 16 * j + (lane & 15);
                if (row < N && col < colmax && col <= row)
                    ((gdouble_mut_t *)A)[(int64_t)row * N + col] = acc[i][j][q];
            }
}

int64_t chol_upd_items(int64_t N, int64_t j) {
    const int64_t t0 = (j + 1) * CHOL_NB;
    if (N <= t0) return 0;
    const int64_t t = (N - t0 + UPD_T - 1) / UPD_T;
    return (j % CHOL_OUTER != CHOL_OUTER - 1) ? t : t * (t + 1) / 2;
}

int launch_chol_update(brpa_ctx *ctx, const FragDev *fr, const int64_t *pre, int nfrag,
                       int64_t total, int step) {
    if (total <= 0) return BRPA_OK;
    if (step % CHOL_OUTER != CHOL_OUTER - 1)
        hipLaunchKernelGGL(k_chol_update<true>, dim3(grid1(total)), dim3(256), 0, ctx->stream, fr,
                           pre, nfrag, step);
    else
        hipLaunchKernelGGL(k_chol_update<false>, dim3(grid1(total)), dim3(256), 0, ctx->stream, fr,
                           pre, nfrag, step);
    LAUNCH_CHECK(ctx);
    return BRPA_OK;
}

__global__ __launch_bounds__(256) void k_energy(const FragDev *__restrict__ fr) {
    const FragDev &F = fr[blockIdx.x];
    __shared__ double sh[256], sl[256];
    const int tid = threadIdx.x;
    dd_t etot = {0.0, 0.0};
    for (int w = 0; w < F.F; ++w) {
        dd_t t = {0.0, 0.0};
        for (int P = tid; P < F.N; P += 256) t = dd_add(t, F.diag[(int64_t)w * F.N + P]);
        sh[tid] = t.hi;
        sl[tid] = t.lo;
        __syncthreads();
        for (int s = 128; s > 0; s >>= 1) {
            if (tid < s) {
                dd_t x = {sh[tid], sl[tid]}, y = {sh[tid + s], sl[tid + s]};
                x = dd_add(x, y);
                sh[tid] = x.hi;
                sl[tid] = x.lo;
            }
            __syncthreads();
        }
        if (tid == 0) {
            dd_t tr = {sh[0], sl[0]};
            dd_t ld = {0.0, 0.0};
            // only this fragment's own Cholesky steps wrote a partial
            const int ns = (F.N + CHOL_NB - 1) / CHOL_NB;
            for (int j = 0; j < ns; ++j) {
                const double *p = F.ldpart + ((int64_t)w * F.nsteps + j) * 2;
                ld = dd_add(ld, (dd_t){p[0], p[1]});
            }
            ld.hi *= 2.0;
            ld.lo *= 2.0;
            dd_t eu = dd_add(ld, (dd_t){-tr.hi, -tr.lo});
            F.eu[w] = eu.hi + eu.lo;
            const double c = F.weights[w] / (2.0 * 3.14159265358979323846);
            double ph = c * eu.hi;
            double pl = fma(c, eu.hi, -ph) + c * eu.lo;
            etot = dd_add(etot, ph);
            etot.lo += pl;
        }
        __syncthreads();
    }
    if (tid == 0) F.energy[0] = etot.hi + etot.lo;
}

int launch_energy(brpa_ctx *ctx, const FragDev *fr, int nfrag) {
    if (nfrag <= 0) return BRPA_OK;
    hipLaunchKernelGGL(k_energy, dim3(nfrag), dim3(256), 0, ctx->stream, fr);
    LAUNCH_CHECK(ctx);
    return BRPA_OK;
}

// Fill n ints with a value (used instead of hipMemsetAsync so that the
// sequence stays valid under stream capture).
__global__ void k_fill_int(int *p, int64_t n, int v) {
    int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = v;
}

int launch_fill_int(brpa_ctx *ctx, int *p, int64_t n, int v) {
    if (n <= 0) return BRPA_OK;
    hipLaunchKernelGGL(k_fill_int, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream,
                       p, n, v);
    LAUNCH_CHECK(ctx);
    return BRPA_OK;
}
