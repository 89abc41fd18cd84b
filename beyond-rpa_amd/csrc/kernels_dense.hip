// Dense FP64 GEMM and triangular solve on v_mfma_f64_16x16x4_f64 (gfx950).
//
// Column-major operands (element (r, c) of a matrix with leading dimension ld
// at r + ld c), the layout of the reference's Fortran arrays and of the
// explicit-space exchange.
//
//   k_gemm_f64<TA, TB>  batched C = alpha op(A) op(B) diag(cs) + beta C on
//                       128 x 128 tiles; upper mode (SYRK) and triangular
//                       operands whose zero triangle is skipped; 16-k stages
//                       loaded straight into LDS where alignment allows
//   k_trinv_diag        inverses of the 64 x 64 diagonal blocks of upper
//                       triangular matrices; with GEMM levels on top, the
//                       batched triangular inverse (launch_trinv_f64)
//   k_trsm_f64<BROW>    L Y = B by blocked forward substitution, L lower
//                       triangular (the row-major lower factor the Cholesky
//                       kernels leave, i.e. U^T of the column-major upper U);
//                       B column-major or row-major; batched over matrices
//                       with strides (brpa_dtrsm)
#include <algorithm>

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
constexpr int GGROUP = 8;  // tile rows per group of the tile order (panel reuse in L2)
}

// ---------------------------------------------------------------------------
// Batched C = alpha op(A) op(B) diag(cs) + beta C on 128 x 128 tiles, four
// waves of 64 x 64 per work-group, v_mfma_f64_16x16x4_f64.
//
// Stages of 16 k go to LDS as [row][k] for both operands (row pitch 18
// doubles), double-buffered: the next stage is loaded into registers while
// the current one is multiplied, and stored into the other buffer halfway
// through it, one barrier per stage.  Every thread moves pairs of consecutive
// k of one row (one ds_write_b128).  In a stage, lane l of a wave supplies the k values
// 4 (l / 16) .. +3 of its row to the four MFMA steps, so its operands for the
// whole stage are two ds_read_b128 per 16-row block; the 16 rows of such a
// read land on distinct 4-bank groups (pitch 36 dwords).  The k order inside
// a stage is a permutation of the sum, the same for both operands.  The MFMA
// takes the B fragment as its first operand, so the accumulator rows run along
// the lanes (D[r] = C(row = lane%16, col = lane/16 + 4r) of a 16 x 16 block)
// and the stores of C are contiguous across lanes.
//
// Global loads by orientation: op(A) = A (TA = false) holds rows contiguous,
// so lanes take consecutive rows at two k; TA holds k contiguous, so eight
// lanes read one row's 16 consecutive k.  op(B) = B (TB = false) is k-
// contiguous, TB rows-contiguous.
//
// Batches: work-group x = (item along dimension 1) * tiles + tile, y = item
// along dimension 2; every operand moves by its own stride per item.  Tile
// order: upper (SYRK: tiles ti <= tj only, by column) or groups of GGROUP tile
// rows swept column by column, so that neighbouring work-groups share panels
// (reversed along the tile index whose k range grows, for TRI_B_UPPER and
// TRI_A_LOWER, so that the longest tiles start first).  Triangular operands
// skip the k range where they are zero (they must hold explicit zeros there):
//   TRI_A_UPPER  op(A)(i, k) = 0 for k < i    TRI_A_LOWER  for k > i
//   TRI_B_UPPER  op(B)(k, j) = 0 for k > j    TRI_B_LOWER  for k < j
// cs (optional): column j of the product is scaled by cs[j] before beta C.
// Split k (nsplit > 1, one item only): work-group z takes the z-th chunk of k
// and writes its partial tile to P + z m n (ld m); k_gemm_reduce then adds the
// partials in order and applies alpha, cs and beta.
// With even leading dimensions and strides, 16-byte aligned operands and, for
// an operand contiguous in rows, an even row count (dma = 1), each stage goes
// from global memory straight into LDS (global_load_lds_dwordx4, 16 bytes per
// lane, 1 KB per wave instruction) with no register staging and no LDS store
// instructions (layouts at dma_sources).  Rows past m (n) repeat the last row
// (pair): they only reach products that are not stored.  A last stage with
// fewer than 16 k goes through registers with zeros past kend.  The k order of
// the sum is the same as in the register-staged form.
struct GemmArgs {
    int m, n, k;
    int mt, nt, ntiles;
    int upper, tri, nb1;
    int nsplit, kchunk;
    int dma;   // operands staged by direct global -> LDS loads
    double alpha, beta;
    const double *A;
    int64_t lda, sA1, sA2;
    const double *B;
    int64_t ldb, sB1, sB2;
    double *C;
    int64_t ldc, sC1, sC2;
    const double *cs;
    int64_t scs1, scs2;
    double *P;
};

namespace {
constexpr int KP = GKB + 2;   // LDS row pitch (doubles)
typedef double d2_t __attribute__((ext_vector_type(2)));
}

// Direct-to-LDS staging of one operand of a 128-row tile and 16-k stage.
// k-contiguous (KC): [row][k] with unpadded 16-double rows; wave w loads the
// 8-row blocks 32 w + 8 c (c < 4), lane L the k pair (L % 8) ^ ((row >> 1) & 7)
// of row 8 c + L / 8 of its block.  Rows-contiguous: [k][row] with rows of
// DMA_KPR doubles (a pitch of 4 rows moves by half of the 64 LDS banks, so
// the four lane groups of a fragment read hit disjoint halves); wave w loads
// the k rows 4 w + c, lane L the rows 2 L, 2 L + 1.  dk: the lane's k offset
// within a stage (KC, pair start) or its k row.
constexpr int DMA_KPR = 132;
template <bool KC>
__host__ __device__ __forceinline__ void dma_sources(const double *X, int64_t ldx, int r0,
                                                     int rows, int wave, int lane,
                                                     const double *(&p)[4], int (&dk)[4]) {
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        if (KC) {
            const int row = 32 * wave + 8 * c + (lane >> 3);
            dk[c] = 2 * ((lane & 7) ^ ((row >> 1) & 7));
            p[c] = X + ldx * (int64_t)min(r0 + row, rows - 1) + dk[c];
        } else {
            // rows is even here: the last pair is rows - 2, rows - 1
            dk[c] = 4 * wave + c;
            p[c] = X + min(r0 + 2 * lane, rows - 2) + ldx * (int64_t)dk[c];
        }
    }
}
template <bool KC>
__host__ __device__ __forceinline__ int dma_block(int wave, int c) {
    return KC ? (32 * wave + 8 * c) * GKB : (4 * wave + c) * DMA_KPR;
}
// the lane's 16 bytes of a last stage with k < kend, zeros past it
template <bool KC>
__host__ __device__ __forceinline__ d2_t dma_tail(const double *p, int64_t ldx, int k0, int dk,
                                                  int kend) {
    d2_t v;
    if (KC) {
        v.x = k0 + dk < kend ? p[k0] : 0.0;
        v.y = k0 + dk + 1 < kend ? p[k0 + 1] : 0.0;
    } else {
        const bool ok = k0 + dk < kend;
        v.x = ok ? p[ldx * k0] : 0.0;
        v.y = ok ? p[ldx * k0 + 1] : 0.0;
    }
    return v;
}
// fragment of row `row` at the k pair 2 g4 + hh of a stage
template <bool KC>
__host__ __device__ __forceinline__ d2_t dma_frag(const double *L, int row, int g4, int hh) {
    if (KC) {
        const int kp = 2 * g4 + hh;
        return *reinterpret_cast<const d2_t *>(&L[row * GKB + 2 * (kp ^ ((row >> 1) & 7))]);
    }
    const int kr = 4 * g4 + 2 * hh;
    d2_t v;
    v.x = L[kr * DMA_KPR + row];
    v.y = L[(kr + 1) * DMA_KPR + row];
    return v;
}

// one operand's stage at k0 into the LDS buffer L: direct loads when the stage
// is full, else through registers with zeros past kend
template <bool KC>
__device__ __forceinline__ void dma_stage(const double *(&p)[4], int (&dk)[4],
                                          int64_t ldx, double *L, int wave, int lane, int k0,
                                          int kend) {
    if (k0 + GKB <= kend) {
        // the builtin exists for the device pass only; the host pass never runs this
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
        for (int c = 0; c < 4; ++c)
            __builtin_amdgcn_global_load_lds(p[c] + (KC ? (int64_t)k0 : ldx * k0),
                                             L + dma_block<KC>(wave, c), 16, 0, 0);
#endif
    } else {
#pragma unroll
        for (int c = 0; c < 4; ++c)
            *reinterpret_cast<d2_t *>(L + dma_block<KC>(wave, c) + 2 * lane) =
                dma_tail<KC>(p[c], ldx, k0, dk[c], kend);
    }
}

template <bool TA, bool TB>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_gemm_f64(const GemmArgs g) {
    __shared__ __attribute__((aligned(16))) double sA[2][GT * KP];
    __shared__ __attribute__((aligned(16))) double sB[2][GT * KP];
    const int item1 = (int)(blockIdx.x / (unsigned)g.ntiles);
    const int t = (int)(blockIdx.x - (unsigned)item1 * (unsigned)g.ntiles);
    const int item2 = (int)blockIdx.y;
    int ti, tj;
    if (g.upper) {
        // upper tiles ti <= tj in column order: tj (tj+1)/2 + ti
        tri_decode(t, tj, ti);
    } else {
        const int gsz = GGROUP * g.nt;            // tiles per full group of tile rows
        const int grp = t / gsz, r = t - grp * gsz;
        const int rows = min(GGROUP, g.mt - grp * GGROUP);
        ti = grp * GGROUP + r % rows;
        tj = r / rows;
        // a triangular operand gives tiles k ranges that grow with tj (TRI_B_UPPER)
        // or ti (TRI_A_LOWER): those tiles go first, so the grid ends on short ones
        if (g.tri == TRI_B_UPPER) tj = g.nt - 1 - tj;
        else if (g.tri == TRI_A_LOWER) ti = g.mt - 1 - ti;
    }
    const double *__restrict__ A = g.A + item1 * g.sA1 + item2 * g.sA2;
    const double *__restrict__ B = g.B + item1 * g.sB1 + item2 * g.sB2;
    const int m = g.m, n = g.n;
    const int64_t lda = g.lda, ldb = g.ldb;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int wm = wave >> 1, wn = wave & 1;
    const int m0 = ti * GT, n0 = tj * GT;
    int kbeg = 0, kend = g.k;
    if (g.tri == TRI_A_UPPER) kbeg = m0;
    else if (g.tri == TRI_A_LOWER) kend = min(kend, m0 + GT);
    else if (g.tri == TRI_B_UPPER) kend = min(kend, n0 + GT);
    else if (g.tri == TRI_B_LOWER) kbeg = n0;
    if (g.nsplit > 1) {
        const int z = (int)blockIdx.z;
        kbeg = max(kbeg, z * g.kchunk);
        kend = min(kend, (z + 1) * g.kchunk);
    }

    d4_t acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = (d4_t){0.0, 0.0, 0.0, 0.0};

    // pair q = tid + 256 i (i < 4) of a 128 x 16 panel: (row, k pair kp).
    // Rows-contiguous operands: row q % 128 (fixed per thread), kp q / 128 =
    // kp0 + 2 i; k-contiguous: row q / 8 = row0 + 32 i, kp q % 8 (fixed).  So
    // per operand one base offset and one stride per i.
    const int arow0 = TA ? (tid >> 3) : (tid & 127), akp0 = TA ? (tid & 7) : (tid >> 7);
    const int brow0 = TB ? (tid & 127) : (tid >> 3), bkp0 = TB ? (tid >> 7) : (tid & 7);
    const int64_t astep = TA ? 1 : lda, bstep = TB ? ldb : 1;   // per unit of k
    const double *__restrict__ pa0 = A + (TA ? (int64_t)(2 * akp0) + lda * (m0 + arow0)
                                             : (int64_t)(m0 + arow0) + lda * (2 * akp0));
    const double *__restrict__ pb0 = B + (TB ? (int64_t)(n0 + brow0) + ldb * (2 * bkp0)
                                             : (int64_t)(2 * bkp0) + ldb * (n0 + brow0));
    const int64_t ai_step = TA ? 32 * lda : 4 * lda;   // per i
    const int64_t bi_step = TB ? 4 * ldb : 32 * ldb;
    const int as_lds = arow0 * KP + 2 * akp0, bs_lds = brow0 * KP + 2 * bkp0;
    const int ai_lds = TA ? 32 * KP : 4, bi_lds = TB ? 4 : 32 * KP;

    d2_t ra[4], rb[4];
    auto gload = [&](int k0) {
        const double *pa = pa0 + astep * k0, *pb = pb0 + bstep * k0;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const bool aok = m0 + arow0 + (TA ? 32 * i : 0) < m;
            const int ka = k0 + 2 * akp0 + (TA ? 0 : 4 * i);
            ra[i].x = (aok && ka < kend) ? pa[i * ai_step] : 0.0;
            ra[i].y = (aok && ka + 1 < kend) ? pa[i * ai_step + astep] : 0.0;
            const bool bok = n0 + brow0 + (TB ? 0 : 32 * i) < n;
            const int kb = k0 + 2 * bkp0 + (TB ? 4 * i : 0);
            rb[i].x = (bok && kb < kend) ? pb[i * bi_step] : 0.0;
            rb[i].y = (bok && kb + 1 < kend) ? pb[i * bi_step + bstep] : 0.0;
        }
    };
    auto lstore = [&](int buf) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            *reinterpret_cast<d2_t *>(&sA[buf][as_lds + i * ai_lds]) = ra[i];
            *reinterpret_cast<d2_t *>(&sB[buf][bs_lds + i * bi_lds]) = rb[i];
        }
    };

    if (g.dma) {
        if (kbeg < kend) {
            const int rl = lane & 15, g4 = lane >> 4;
            constexpr bool KCA = TA, KCB = !TB;   // operand contiguous in k
            const double *pa[4], *pb[4];
            int ka[4], kb[4];
            dma_sources<KCA>(A, lda, m0, m, wave, lane, pa, ka);
            dma_sources<KCB>(B, ldb, n0, n, wave, lane, pb, kb);
            dma_stage<KCA>(pa, ka, lda, sA[0], wave, lane, kbeg, kend);
            dma_stage<KCB>(pb, kb, ldb, sB[0], wave, lane, kbeg, kend);
            __syncthreads();
            int buf = 0;
            for (int k0 = kbeg; k0 < kend; k0 += GKB, buf ^= 1) {
                d2_t a[4][2], b[4][2];
#pragma unroll
                for (int hh = 0; hh < 2; ++hh) {
#pragma unroll
                    for (int i = 0; i < 4; ++i)
                        a[i][hh] = dma_frag<KCA>(sA[buf], 64 * wm + 16 * i + rl, g4, hh);
#pragma unroll
                    for (int j = 0; j < 4; ++j)
                        b[j][hh] = dma_frag<KCB>(sB[buf], 64 * wn + 16 * j + rl, g4, hh);
                }
                // the other buffer was last read before the previous barrier
                if (k0 + GKB < kend) {
                    dma_stage<KCA>(pa, ka, lda, sA[buf ^ 1], wave, lane, k0 + GKB, kend);
                    dma_stage<KCB>(pb, kb, ldb, sB[buf ^ 1], wave, lane, k0 + GKB, kend);
                }
#pragma unroll
                for (int hh = 0; hh < 2; ++hh)
#pragma unroll
                    for (int st = 0; st < 2; ++st)
#pragma unroll
                        for (int i = 0; i < 4; ++i)
#pragma unroll
                            for (int j = 0; j < 4; ++j)
                                acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(
                                    b[j][hh][st], a[i][hh][st], acc[i][j], 0, 0, 0);
                // keep the barrier (and its wait for the loads in flight) after the MFMAs
                __builtin_amdgcn_sched_barrier(0);
                __syncthreads();
            }
        }
    } else if (kbeg < kend) {
        const int kq = 4 * (lane >> 4), rl = lane & 15;
        gload(kbeg);
        lstore(0);
        __syncthreads();
        if (kbeg + GKB < kend) gload(kbeg + GKB);
        int buf = 0;
        for (int k0 = kbeg; k0 < kend; k0 += GKB, buf ^= 1) {
            // the stage in two halves of two k each; the next stage goes to the
            // other buffer between them (that buffer was last read before the
            // previous barrier), so its LDS stores overlap the second half's MFMAs
#pragma unroll
            for (int hh = 0; hh < 2; ++hh) {
                d2_t a[4], b[4];
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    a[i] = *reinterpret_cast<const d2_t *>(
                        &sA[buf][(64 * wm + 16 * i + rl) * KP + kq + 2 * hh]);
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    b[j] = *reinterpret_cast<const d2_t *>(
                        &sB[buf][(64 * wn + 16 * j + rl) * KP + kq + 2 * hh]);
                // C^T block = op(B)^T op(A)^T: rows of C along the lanes
#pragma unroll
                for (int st = 0; st < 2; ++st)
#pragma unroll
                    for (int i = 0; i < 4; ++i)
#pragma unroll
                        for (int j = 0; j < 4; ++j)
                            acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(b[j][st], a[i][st],
                                                                             acc[i][j], 0, 0, 0);
                if (hh == 0 && k0 + GKB < kend) lstore(buf ^ 1);
            }
            __syncthreads();
            if (k0 + 2 * GKB < kend) gload(k0 + 2 * GKB);
        }
    }

    if (g.nsplit > 1) {
        // partial tile, reduced by k_gemm_reduce
        double *P = g.P + (int64_t)blockIdx.z * m * n;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int row = m0 + 64 * wm + 16 * i + (lane & 15);
            if (row >= m) continue;
#pragma unroll
            for (int j = 0; j < 4; ++j)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int col = n0 + 64 * wn + 16 * j + (lane >> 4) + 4 * r;
                    if (col < n) P[(int64_t)row + (int64_t)m * col] = acc[i][j][r];
                }
        }
        return;
    }
    double *__restrict__ C = g.C + item1 * g.sC1 + item2 * g.sC2;
    const int64_t ldc = g.ldc;
    const double *cs = g.cs ? g.cs + item1 * g.scs1 + item2 * g.scs2 : nullptr;
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
                double v = g.alpha * acc[i][j][r];
                if (cs) v *= cs[col];
                *p = g.beta == 0.0 ? v : fma(g.beta, *p, v);
            }
    }
}

// C(r, c) = alpha cs[c] sum_z P_z(r, c) + beta C(r, c), partials in split order;
// in the upper mode only the elements of tiles ti <= tj (those the GEMM
// writes).  One thread per element, rows along the threads.
__global__ __launch_bounds__(256) void k_gemm_reduce(const GemmArgs g) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t mn = (int64_t)g.m * g.n;
    if (e >= mn) return;
    const int col = (int)(e / g.m), row = (int)(e - (int64_t)col * g.m);
    if (g.upper && row / GT > col / GT) return;
    double s = 0.0;
    for (int z = 0; z < g.nsplit; ++z) s += g.P[z * mn + e];
    double v = g.alpha * s;
    if (g.cs) v *= g.cs[col];
    double *p = g.C + (int64_t)row + g.ldc * col;
    *p = g.beta == 0.0 ? v : fma(g.beta, *p, v);
}

int launch_gemm_f64_batched(brpa_ctx *ctx, bool ta, bool tb, int64_t m, int64_t n, int64_t k,
                            double alpha, const double *A, int64_t lda, int64_t sA1, int64_t sA2,
                            const double *B, int64_t ldb, int64_t sB1, int64_t sB2, double beta,
                            double *C, int64_t ldc, int64_t sC1, int64_t sC2, int64_t nb1,
                            int64_t nb2, bool upper, int tri, const double *cs, int64_t scs1,
                            int64_t scs2) {
    if (m <= 0 || n <= 0 || nb1 <= 0 || nb2 <= 0) return BRPA_OK;
    if (m > INT32_MAX / 2 || n > INT32_MAX / 2 || k > INT32_MAX / 2 || (upper && m != n) ||
        nb2 > 65535 || tri < 0 || tri > TRI_B_LOWER)
        return brpa_fail(ctx, BRPA_ERR_ARG, "gemm: unsupported shape");
    const int64_t mt = (m + GT - 1) / GT, nt = (n + GT - 1) / GT;
    const int64_t ntiles = upper ? nt * (nt + 1) / 2 : mt * nt;
    if (ntiles * nb1 > INT32_MAX) return brpa_fail(ctx, BRPA_ERR_ARG, "gemm: too many tiles");
    GemmArgs g;
    g.m = (int)m;
    g.n = (int)n;
    g.k = k > 0 ? (int)k : 0;
    g.mt = (int)mt;
    g.nt = (int)nt;
    g.ntiles = (int)ntiles;
    g.upper = upper ? 1 : 0;
    g.tri = tri;
    g.nb1 = (int)nb1;
    g.alpha = alpha;
    g.beta = beta;
    g.A = A;
    g.lda = lda;
    g.sA1 = sA1;
    g.sA2 = sA2;
    g.B = B;
    g.ldb = ldb;
    g.sB1 = sB1;
    g.sB2 = sB2;
    g.C = C;
    g.ldc = ldc;
    g.sC1 = sC1;
    g.sC2 = sC2;
    g.cs = cs;
    g.scs1 = scs1;
    g.scs2 = scs2;
    g.P = nullptr;
    // direct-to-LDS stages: 16-byte aligned sources (even leading dimensions and
    // strides, aligned bases); a rows-contiguous operand also needs an even row
    // count, so that its last row pair lies inside the matrix
    g.dma = (lda % 2 == 0 && ldb % 2 == 0 && sA1 % 2 == 0 && sA2 % 2 == 0 && sB1 % 2 == 0 &&
             sB2 % 2 == 0 && (reinterpret_cast<uintptr_t>(A) & 15) == 0 &&
             (reinterpret_cast<uintptr_t>(B) & 15) == 0 && (ta || m % 2 == 0) &&
             (!tb || n % 2 == 0))
                ? 1
                : 0;
    // split k when the last wave of tiles would leave much of the chip idle:
    // the split count s (<= 8, chunks of >= 256 k) minimises the modelled time
    // 1/eff(s) + (reduction traffic), eff(s) = tiles s / (slots ceil(tiles s /
    // slots)) with two work-groups per CU; partials in the context's split
    // scratch (set by the caller for the duration of a call)
    g.nsplit = 1;
    g.kchunk = g.k;
    const int64_t wgs = ntiles * nb1 * nb2;
    if (nb1 == 1 && nb2 == 1 && ctx->split_ws && g.k >= 512) {
        // CUs of the context's stream (its share of the chip, or all of them)
        int ncu = ctx->ncu;
        if (ncu <= 0 &&
            (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, ctx->device) !=
                 hipSuccess || ncu <= 0)) {
            (void)hipGetLastError();
            ncu = 256;
        }
        const double slots = 2.0 * ncu;
        double best = 1e300;
        int64_t bs = 1;
        const int64_t smax = std::min<int64_t>({8, g.k / 256,
                                                (int64_t)(ctx->split_ws_cap / (size_t)(m * n))});
        for (int64_t sp = 1; sp <= std::max<int64_t>(smax, 1); ++sp) {
            const double w = (double)(wgs * sp);
            const double eff = w / (slots * std::ceil(w / slots));
            const double cost = 1.0 / eff + (sp > 1 ? 48.0 * (double)(sp + 1) / (double)g.k : 0.0);
            if (cost < best - 1e-9) {
                best = cost;
                bs = sp;
            }
        }
        if (bs > 1) {
            const int64_t chunk = ((g.k + bs - 1) / bs + GKB - 1) / GKB * GKB;
            g.nsplit = (int)((g.k + chunk - 1) / chunk);
            g.kchunk = (int)chunk;
            g.P = ctx->split_ws;
        }
    }
    const dim3 grid((unsigned)(ntiles * nb1), (unsigned)nb2, (unsigned)g.nsplit), block(256);
    size_t slot = 0;
    BRPA_TRY(kstat_begin(ctx, &slot));
    if (!ta && !tb)
        hipLaunchKernelGGL((k_gemm_f64<false, false>), grid, block, 0, ctx->stream, g);
    else if (!ta && tb)
        hipLaunchKernelGGL((k_gemm_f64<false, true>), grid, block, 0, ctx->stream, g);
    else if (ta && !tb)
        hipLaunchKernelGGL((k_gemm_f64<true, false>), grid, block, 0, ctx->stream, g);
    else
        hipLaunchKernelGGL((k_gemm_f64<true, true>), grid, block, 0, ctx->stream, g);
    LAUNCH_CHECK(ctx);
    if (g.nsplit > 1) {
        hipLaunchKernelGGL(k_gemm_reduce, dim3((unsigned)((m * n + 255) / 256)), dim3(256), 0,
                           ctx->stream, g);
        LAUNCH_CHECK(ctx);
    }
    // algorithmic flop: 2 m n k, n (n + 1) k on the triangle of a SYRK, m n k
    // (n m k) with a triangular op(A) (op(B)) of order m = k (n = k)
    double fl = upper ? (double)n * (double)(n + 1) * (double)g.k
                      : 2.0 * (double)m * (double)n * (double)g.k;
    if (tri != TRI_NONE) fl *= 0.5;
    return kstat_end(ctx, slot, BRPA_KERNEL_GEMM_F64, fl * (double)nb1 * (double)nb2);
}

int launch_gemm_f64(brpa_ctx *ctx, bool ta, bool tb, int64_t m, int64_t n, int64_t k,
                    double alpha, const double *A, int64_t lda, const double *B, int64_t ldb,
                    double beta, double *C, int64_t ldc, bool upper) {
    return launch_gemm_f64_batched(ctx, ta, tb, m, n, k, alpha, A, lda, 0, 0, B, ldb, 0, 0, beta,
                                   C, ldc, 0, 0, 1, 1, upper, TRI_NONE, nullptr, 0, 0);
}

// ---------------------------------------------------------------------------
// Triangular inverse W = U^{-1} of nmat upper triangular matrices (column-
// major, U(r, c) at r + ldu c + z su; W(r, c) at r + ldw c + z sw), by
// recursive doubling: the 64 x 64 diagonal blocks are inverted by
// substitution (k_trinv_diag, which also writes the zeros below the diagonal),
// then for block sizes b = 64, 128, ...  every pair of neighbouring inverted
// blocks [W11 .; 0 W22] of a 2b block gets its corner
//   W12 = -W11 U12 W22
// by two batched triangular GEMMs (T = U12 W22, W12 = -W11 T).  Only the upper
// triangle of U is read.
namespace {
constexpr int TIB = 64;
}

__global__ __launch_bounds__(256) void k_trinv_diag(int n, const double *__restrict__ U,
                                                   int64_t ldu, int64_t su,
                                                   double *__restrict__ W, int64_t ldw,
                                                   int64_t sw) {
    __shared__ double sU[TIB][TIB + 1];
    __shared__ double sW[TIB][TIB + 1];
    const int p = (int)blockIdx.x, z = (int)blockIdx.y;
    const double *Um = U + (int64_t)z * su;
    double *Wm = W + (int64_t)z * sw;
    const int r0 = p * TIB;
    const int nb = min(TIB, n - r0);
    const int tid = threadIdx.x;
    // the upper triangle of the diagonal block; identity past the matrix end
    for (int e = tid; e < TIB * TIB; e += 256) {
        const int c = e >> 6, r = e & 63;
        double v = 0.0;
        if (r < nb && c < nb) {
            if (r <= c) v = Um[(int64_t)(r0 + r) + ldu * (r0 + c)];
        } else if (r == c) {
            v = 1.0;
        }
        sU[r][c] = v;
    }
    __syncthreads();
    // column c of the inverse by back substitution: U w = e_c, w_i = 0 for i > c
    if (tid < TIB) {
        const int c = tid;
        for (int i = TIB - 1; i > c; --i) sW[i][c] = 0.0;
        for (int i = c; i >= 0; --i) {
            double s0 = (i == c) ? 1.0 : 0.0, s1 = 0.0;
            int q = i + 1;
            for (; q + 1 <= c; q += 2) {
                s0 = fma(-sU[i][q], sW[q][c], s0);
                s1 = fma(-sU[i][q + 1], sW[q + 1][c], s1);
            }
            if (q <= c) s0 = fma(-sU[i][q], sW[q][c], s0);
            sW[i][c] = (s0 + s1) / sU[i][i];
        }
    }
    __syncthreads();
    // the block, and zeros below it in its columns
    for (int e = tid; e < TIB * TIB; e += 256) {
        const int c = e >> 6, r = e & 63;
        if (r < nb && c < nb) Wm[(int64_t)(r0 + r) + ldw * (r0 + c)] = sW[r][c];
    }
    for (int64_t e = tid; e < (int64_t)(n - r0 - nb) * nb; e += 256) {
        const int c = (int)(e / (n - r0 - nb));
        const int r = (int)(e - (int64_t)c * (n - r0 - nb));
        Wm[(int64_t)(r0 + nb + r) + ldw * (r0 + c)] = 0.0;
    }
}

int64_t trinv_scratch(int64_t n) { return n * n; }

int launch_trinv_f64(brpa_ctx *ctx, int64_t n, const double *U, int64_t ldu, int64_t su, double *W,
                     int64_t ldw, int64_t sw, int64_t nmat, double *T) {
    if (n <= 0 || nmat <= 0) return BRPA_OK;
    if (n > INT32_MAX / 2 || nmat > 65535 || ldu < n || ldw < n)
        return brpa_fail(ctx, BRPA_ERR_ARG, "trinv: unsupported shape");
    const int64_t nblk = (n + TIB - 1) / TIB;
    size_t slot = 0;
    BRPA_TRY(kstat_begin(ctx, &slot));
    hipLaunchKernelGGL(k_trinv_diag, dim3((unsigned)nblk, (unsigned)nmat), dim3(256), 0,
                       ctx->stream, (int)n, U, ldu, su, W, ldw, sw);
    LAUNCH_CHECK(ctx);
    BRPA_TRY(kstat_end(ctx, slot, BRPA_KERNEL_TRINV_F64, (double)nmat * nblk * TIB * TIB * TIB / 3.0));
    for (int64_t b = TIB; b < n; b *= 2) {
        // full pairs [r0, r0 + 2b) and one ragged pair whose lower block is shorter
        const int64_t nfull = n / (2 * b);
        const int64_t rr = nfull * 2 * b;
        const int64_t b2 = (rr + b < n) ? n - rr - b : 0;
        const int64_t sT2 = (nfull + 1) * b * b;   // T per matrix
        struct Part {
            int64_t r0, cnt, bb;
        } parts[2] = {{0, nfull, b}, {rr, b2 > 0 ? 1 : 0, b2}};
        for (const Part &pt : parts) {
            if (pt.cnt == 0) continue;
            const int64_t r0 = pt.r0, bb = pt.bb;
            double *Tp = T + (r0 / (2 * b)) * b * b;
            // T = U12 W22: U12 (b x bb) at rows r0.., columns r0 + b..; W22 upper
            BRPA_TRY(launch_gemm_f64_batched(
                ctx, false, false, b, bb, bb, 1.0, U + r0 + ldu * (r0 + b), ldu, 2 * b * (1 + ldu),
                su, W + (r0 + b) * (1 + ldw), ldw, 2 * b * (1 + ldw), sw, 0.0, Tp, b, b * b, sT2,
                pt.cnt, nmat, false, TRI_B_UPPER, nullptr, 0, 0));
            // W12 = -W11 T: W11 (b x b) upper
            BRPA_TRY(launch_gemm_f64_batched(
                ctx, false, false, b, bb, b, -1.0, W + r0 * (1 + ldw), ldw, 2 * b * (1 + ldw), sw, Tp,
                b, b * b, sT2, 0.0, W + r0 + ldw * (r0 + b), ldw, 2 * b * (1 + ldw), sw, pt.cnt,
                nmat, false, TRI_A_UPPER, nullptr, 0, 0));
        }
    }
    return BRPA_OK;
}

// ---------------------------------------------------------------------------
// Forward substitution L Y = B, L (n x n) lower triangular stored as the
// row-major lower triangle (L(r, c) at r * ldl + c, c <= r), Y overwriting B.
// B is n x nrhs: column-major (BROW = false, B(r, c) at r + ldb c) or row-major
// (BROW, B(r, c) at r ldb + c, i.e. the rows of a column-major nrhs x n
// matrix X: X <- X L^{-T}).
//
// One work-group per (matrix, 64 right-hand sides), left-looking over 64-row
// blocks p: acc = B_p - sum_{q<p} L_pq Y_q on MFMA (each wave 16 rows x 64
// columns), then Y_p = L_pp^{-1} acc by substitution in LDS (one thread per
// right-hand side, the 64 x 64 diagonal block staged in LDS).  The Y_q blocks
// were written by the same work-group and come back from L2.
namespace {
constexpr int TB_ = 64;
}

template <bool BROW>
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
    for (int p = 0; p < nblk; ++p) {
        const int r0 = p * TB_;
        // acc: rows 16 * wave + (lane % 16) of the block, columns lane/16 + 4 r + 16 j
        d4_t acc[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[j] = (d4_t){0.0, 0.0, 0.0, 0.0};
        for (int q = 0; q < p; ++q) {
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
                const double bv = (r0 + r < n && c0 + c < nrhs) ? Ym[addr(r0 + r, c0 + c)] : 0.0;
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

int launch_trsm_f64(brpa_ctx *ctx, bool brow, int64_t n, int64_t nrhs, const double *L,
                    int64_t ldl, int64_t strideL, double *Y, int64_t ldb, int64_t strideB,
                    int64_t nmat) {
    if (n <= 0 || nrhs <= 0 || nmat <= 0) return BRPA_OK;
    if (n > INT32_MAX / 2 || nrhs > INT32_MAX / 2)
        return brpa_fail(ctx, BRPA_ERR_ARG, "trsm: unsupported shape");
    const int64_t nct = (nrhs + TB_ - 1) / TB_;
    if (nct * nmat > INT32_MAX) return brpa_fail(ctx, BRPA_ERR_ARG, "trsm: too many blocks");
    const dim3 g((unsigned)(nct * nmat)), b(256);
    size_t slot = 0;
    BRPA_TRY(kstat_begin(ctx, &slot));
    if (brow)
        hipLaunchKernelGGL((k_trsm_f64<true>), g, b, 0, ctx->stream, (int)n, (int)nrhs, L, ldl,
                           strideL, Y, ldb, strideB, (int)nct);
    else
        hipLaunchKernelGGL((k_trsm_f64<false>), g, b, 0, ctx->stream, (int)n, (int)nrhs, L, ldl,
                           strideL, Y, ldb, strideB, (int)nct);
    LAUNCH_CHECK(ctx);
    // algorithmic flop: n^2 per right-hand side
    return kstat_end(ctx, slot, BRPA_KERNEL_TRSM_F64,
                     (double)n * (double)n * (double)nrhs * (double)nmat);
}
