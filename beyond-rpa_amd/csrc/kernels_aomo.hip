// AO -> MO transform of the Cholesky vectors (rpa_CRG / rpa_C / rpa_Cab / rpa_Caa,
// rpa/rpa_core_MO.f90:40-301):
//
//   CRG(P, ai) (+)= sum_pq RG(P, pq) C(ai, pq),
//   C(ai, pq) = Co(p, i) Cv(q, a) + Co(q, i) Cv(p, a)   (p != q)
//             = Co(p, i) Cv(p, a)                        (p == q)
//
// over the packed AO pairs pq of the Cholesky basis; the pair table gives the
// AO indices (p, q) of each packed column.  The reference builds C for a
// subset of pairs (nai x npq) and multiplies; here C is formed while the GEMM
// operand is loaded (no nai x npq intermediate) and the reduction over pq runs
// on v_mfma_f64_16x16x4_f64, with the tile form of the THC producer: 128 x 128
// output tiles, 8 pairs per stage, register prefetch of the raw coefficients
// with the products formed after the MFMAs, XOR-swizzled unpadded LDS rows.
//
// Layouts: Co(p, i) at p + nao i and Cv(q, a) at q + nao a (the reference's
// OccCoeffs(NAO, NOcc) / VirtCoeffs(NAO, NVirt) columns of the active
// orbitals), RG(P, pq) at P + nvecs pq (Fortran RG(NVecs, Npq)), CRG as C
// R[ai][P] (Fortran CRG(NVecs, NVirt*NOcc)), ai = i nvir + a.
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

constexpr int TT = 128;       // output tile: 128 ai x 128 P
constexpr int GB = 8;         // AO pairs per LDS stage
constexpr int CW = GB;        // lanes per operand row of C (pairs fastest)
constexpr int RPP = 256 / CW; // C rows (ai) per load pass
constexpr int NL = TT / RPP;  // load passes per stage and operand

__device__ __forceinline__ int swz(int k) { return ((k & 1) << 4) | ((k >> 1) << 1); }

// One work-group = one 128 x 128 tile of CRG; 4 waves of 64 x 64 (4 x 4 MFMA
// blocks).  MFMA f64 16x16x4 operand layout (lane l): A[m = l%16][k = l/16],
// B[k = l/16][n = l%16], D[r] = C[m = l/16 + 4r][n = l%16].
__global__ __launch_bounds__(256, 2) void k_ao_mo(const double *__restrict__ Co,
                                                  const double *__restrict__ Cv,
                                                  const int2 *__restrict__ pairs,
                                                  const double *__restrict__ RG, int64_t npq,
                                                  int nao, int nvir, int64_t K, int64_t NV,
                                                  int accumulate, double *__restrict__ R) {
    __shared__ double sH[GB][TT];   // C(ai, pq) as [pq][ai], swizzled columns
    __shared__ double sZ[GB][TT];   // RG(P, pq) as [pq][P], swizzled columns
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int64_t ai0 = (int64_t)blockIdx.x * TT, p0 = (int64_t)blockIdx.y * TT;
    const int wm = wave >> 1, wn = wave & 1;

    // C loads: pair offset c, rows r = tid / CW + RPP it
    const int c = tid % CW;
    int oo[NL], ov[NL];   // nao * i and nao * a of each row
    unsigned hok = 0;
#pragma unroll
    for (int it = 0; it < NL; ++it) {
        const int64_t ai = ai0 + tid / CW + RPP * it;
        const bool h = ai < K;
        hok |= (unsigned)h << it;
        const int i = h ? (int)(ai / nvir) : 0, a = h ? (int)(ai - (int64_t)i * nvir) : 0;
        oo[it] = i * nao;
        ov[it] = a * nao;
    }
    // RG loads: P offset rz (lanes along P), pair offsets cz + 2 it
    const int rz = tid & (TT - 1), cz = tid >> 7;
    const bool zok = p0 + rz < NV;

    d4_t acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = (d4_t){0.0, 0.0, 0.0, 0.0};

    // raw prefetch: Co(p,i), Cv(q,a), Co(q,i), Cv(p,a) per row, RG per pair
    double r_op[NL], r_vq[NL], r_oq[NL], r_vp[NL], rz4[NL];
    bool offd = false;
    auto gload = [&](int64_t g0) {
        const bool gin = g0 + c < npq;
        int p = 0, q = 0;
        if (gin) {
            const int2 pr = pairs[g0 + c];
            p = pr.x;
            q = pr.y;
        }
        offd = p != q;
#pragma unroll
        for (int it = 0; it < NL; ++it) {
            const bool h = gin && ((hok >> it) & 1u);
            r_op[it] = h ? Co[oo[it] + p] : 0.0;
            r_vq[it] = h ? Cv[ov[it] + q] : 0.0;
            r_oq[it] = (h && offd) ? Co[oo[it] + q] : 0.0;
            r_vp[it] = (h && offd) ? Cv[ov[it] + p] : 0.0;
            const int64_t g = g0 + cz + 2 * it;
            rz4[it] = (zok && g < npq) ? RG[(p0 + rz) + NV * g] : 0.0;
        }
    };

    gload(0);
    for (int64_t g0 = 0; g0 < npq; g0 += GB) {
        __syncthreads();
#pragma unroll
        for (int it = 0; it < NL; ++it) {
            const int r = (tid / CW + RPP * it) ^ swz(c);
            sH[c][r] = fma(r_op[it], r_vq[it], r_oq[it] * r_vp[it]);
            const int cc = cz + 2 * it;
            sZ[cc][rz ^ swz(cc)] = rz4[it];
        }
        __syncthreads();
        if (g0 + GB < npq) gload(g0 + GB);
#pragma unroll
        for (int kk = 0; kk < GB; kk += 4) {
            const int kr = kk + (lane >> 4);
            const int sw = swz(kr);
            double a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) a[i] = sH[kr][(64 * wm + 16 * i + (lane & 15)) ^ sw];
#pragma unroll
            for (int j = 0; j < 4; ++j) b[j] = sZ[kr][(64 * wn + 16 * j + (lane & 15)) ^ sw];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
        }
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t ai = ai0 + 64 * wm + 16 * i + (lane >> 4) + 4 * r;
                const int64_t p = p0 + 64 * wn + 16 * j + (lane & 15);
                if (ai < K && p < NV) {
                    double *dst = R + ai * NV + p;
                    *dst = accumulate ? *dst + acc[i][j][r] : acc[i][j][r];
                }
            }
}

}  // namespace

int launch_ao_mo(brpa_ctx *ctx, const double *Co, const double *Cv, const int *pairs,
                 const double *RG, int64_t npq, int64_t nao, int64_t nocc, int64_t nvir,
                 int64_t nvecs, int accumulate, double *R) {
    const int64_t K = nocc * nvir;
    if (K <= 0 || nvecs <= 0) return BRPA_OK;
    const int64_t gx = (K + TT - 1) / TT, gy = (nvecs + TT - 1) / TT;
    if (gx > INT32_MAX || gy > 65535 || nao * std::max(nocc, nvir) + nao > INT32_MAX)
        return brpa_fail(ctx, BRPA_ERR_ARG, "AO->MO: sizes beyond the supported grid");
    hipLaunchKernelGGL(k_ao_mo, dim3((unsigned)gx, (unsigned)gy), dim3(256), 0, ctx->stream, Co, Cv,
                       reinterpret_cast<const int2 *>(pairs), RG, npq, (int)nao, (int)nvir, K,
                       nvecs, accumulate, R);
    LAUNCH_CHECK(ctx);
    return BRPA_OK;
}
