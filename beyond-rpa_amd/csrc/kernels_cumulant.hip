// THC SOSEX + 2g cumulant terms (rpa_CCD_corrections_1b2g,
// rpa/rpa_CCD_Corrections.f90:79-290).  For every amplitude vector mu:
//
//   M_mu(g, h) = sum_i YU_mu(g, i) X(h, i),   YU_mu = Y U_mu  (Y = Yga, X = Xgi)
//   S2g_mu = sum_{g,h} Z(g, h) M_mu(g, h)^2
//   S1b_mu = sum_{g,h} Z(g, h) M_mu(g, h) M_mu(h, g)
//
// with Z = Zgh symmetric.  The reference stores M_mu (NGrid x NGrid) per mu and
// walks it by lower-triangular block pairs; here one work-group takes one block
// pair (p >= q) of 64 x 64 grid points and one mu, builds both M(g-block,
// h-block) and M(h-block, g-block)^T in registers on v_mfma_f64_16x16x4_f64
// (K = NOcc, operands straight from L2), contracts them with the Z tile and
// writes two partial sums; M never reaches memory.  The partials are reduced
// per mu in a fixed order (double-double), so the result does not depend on
// scheduling.
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

constexpr int CB = 64;   // grid points per block

// (p, q) with p >= q from the lower-triangle pair index t = p (p + 1) / 2 + q
__device__ __forceinline__ void pair_decode(int t, int &p, int &q) {
    int pp = (int)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
    while ((pp + 1) * (pp + 2) / 2 <= t) ++pp;
    while (pp * (pp + 1) / 2 > t) --pp;
    p = pp;
    q = t - pp * (pp + 1) / 2;
}

// grid: x = block pair, y = mu within the chunk.  YU: column-major G x (o nmu)
// (YU_mu(g, i) at g + G (i + o mu)); X: G x o; Z: G x G symmetric.
// part[(mu * npairs + pair) * 2 + {0, 1}] = (S2g, S1b) of the block pair;
// diag (may be null): M_mu(g, g) at g + G (mu0 + mu).
__global__ __launch_bounds__(256) void k_thc_1b2g(const double *__restrict__ YU,
                                                  const double *__restrict__ X,
                                                  const double *__restrict__ Z, int G, int o,
                                                  int npairs, double *__restrict__ part,
                                                  double *__restrict__ diag, int64_t mu0) {
    __shared__ double red[2][4];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int pair = blockIdx.x, mu = blockIdx.y;
    int pb, qb;
    pair_decode(pair, pb, qb);
    const int g0 = pb * CB, h0 = qb * CB;
    const int wm = wave >> 1, wn = wave & 1;
    const double *yu = YU + (int64_t)G * o * mu;

    d4_t P[2][2], Q[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            P[i][j] = (d4_t){0.0, 0.0, 0.0, 0.0};
            Q[i][j] = (d4_t){0.0, 0.0, 0.0, 0.0};
        }
    // lane rows: m = l % 16 within each 16-row MFMA block, k = l / 16
    const int lm = lane & 15, lk = lane >> 4;
    for (int k0 = 0; k0 < o; k0 += 4) {
        const int k = k0 + lk;
        const bool kin = k < o;
        double ay[2], ax[2], by[2], bx[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const int g = g0 + 32 * wm + 16 * i + lm;
            const bool in = kin && g < G;
            ay[i] = in ? yu[g + (int64_t)G * k] : 0.0;
            ax[i] = in ? X[g + (int64_t)G * k] : 0.0;
        }
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int h = h0 + 32 * wn + 16 * j + lm;
            const bool in = kin && h < G;
            by[j] = in ? yu[h + (int64_t)G * k] : 0.0;
            bx[j] = in ? X[h + (int64_t)G * k] : 0.0;
        }
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                // P(g, h) = M(g, h) = sum_k YU(g, k) X(h, k); Q(g, h) = M(h, g)
                P[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ay[i], bx[j], P[i][j], 0, 0, 0);
                Q[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ax[i], by[j], Q[i][j], 0, 0, 0);
            }
    }
    const bool offd = pb != qb;
    double s2 = 0.0, s1 = 0.0;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int g = g0 + 32 * wm + 16 * i + lk + 4 * r;
                const int h = h0 + 32 * wn + 16 * j + lm;
                if (g < G && h < G) {
                    const double z = Z[h + (int64_t)G * g];   // Z(g, h) = Z(h, g)
                    const double p = P[i][j][r], q = Q[i][j][r];
                    if (offd) {
                        s2 += z * (p * p + q * q);
                        s1 += 2.0 * z * p * q;
                    } else {
                        s2 += z * p * p;
                        s1 += z * p * q;
                        if (diag && g == h) diag[g + (int64_t)G * (mu0 + mu)] = p;
                    }
                }
            }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        s2 += __shfl_xor(s2, off);
        s1 += __shfl_xor(s1, off);
    }
    if (lane == 0) {
        red[0][wave] = s2;
        red[1][wave] = s1;
    }
    __syncthreads();
    if (tid == 0) {
        double* out = part + ((int64_t)mu * npairs + pair) * 2;
        out[0] = (red[0][0] + red[0][1]) + (red[0][2] + red[0][3]);
        out[1] = (red[1][0] + red[1][1]) + (red[1][2] + red[1][3]);
    }
}

// S[(mu0 + mu) * 2 + c] = sum over pairs of part[(mu * npairs + pair) * 2 + c], in
// pair order per thread stripe, then a fixed tree; double-double accumulation
__global__ __launch_bounds__(256) void k_thc_1b2g_reduce(const double *__restrict__ part,
                                                         int npairs, double *__restrict__ S,
                                                         int64_t mu0) {
    __shared__ dd_t sh[2][256];
    const int mu = blockIdx.x, tid = threadIdx.x;
    dd_t a = {0.0, 0.0}, b = {0.0, 0.0};
    for (int t = tid; t < npairs; t += 256) {
        a = dd_add(a, part[((int64_t)mu * npairs + t) * 2]);
        b = dd_add(b, part[((int64_t)mu * npairs + t) * 2 + 1]);
    }
    sh[0][tid] = a;
    sh[1][tid] = b;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (tid < k) {
            sh[0][tid] = dd_add(sh[0][tid], sh[0][tid + k]);
            sh[1][tid] = dd_add(sh[1][tid], sh[1][tid + k]);
        }
        __syncthreads();
    }
    if (tid == 0) {
        S[(mu0 + mu) * 2] = sh[0][0].hi + sh[0][0].lo;
        S[(mu0 + mu) * 2 + 1] = sh[1][0].hi + sh[1][0].lo;
    }
}

}  // namespace

int thc_1b2g_npairs(int64_t G) {
    const int64_t nb = (G + CB - 1) / CB;
    return (int)(nb * (nb + 1) / 2);
}

int launch_thc_1b2g(brpa_ctx *ctx, const double *YU, const double *X, const double *Z, int64_t G,
                    int64_t o, int64_t nmu, double *part, double *diag, int64_t mu0, double *S) {
    if (nmu <= 0 || G <= 0) return BRPA_OK;
    const int npairs = thc_1b2g_npairs(G);
    if (nmu > 65535 || G > 65536)
        return brpa_fail(ctx, BRPA_ERR_ARG, "THC 1b2g: sizes beyond the supported grid");
    hipLaunchKernelGGL(k_thc_1b2g, dim3((unsigned)npairs, (unsigned)nmu), dim3(256), 0, ctx->stream,
                       YU, X, Z, (int)G, (int)o, npairs, part, diag, mu0);
    LAUNCH_CHECK(ctx);
    hipLaunchKernelGGL(k_thc_1b2g_reduce, dim3((unsigned)nmu), dim3(256), 0, ctx->stream, part,
                       npairs, S, mu0);
    LAUNCH_CHECK(ctx);
    return BRPA_OK;
}
