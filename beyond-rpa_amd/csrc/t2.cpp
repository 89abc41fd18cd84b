// Factorised direct-ring amplitudes (rpa_CC_T2_v5) and the SOSEX exchange
// correction (rpa_CC_Exchange_SOSEX) on one GPU.
//
// The reference diagonalises every Pi(u_k) (rpa_CC_Diagonalize_PiU,
// rpa/rpa_CC.f90:1108-1129) only to apply Q_k diag(l/(1+l W_k)) Q_k^T =
// l (1 + l Pi_k)^{-1}; here that operator is applied through a Cholesky
// factor 1 + l Pi_k = U_k^T U_k as Rt_k Rt_k^T with Rt_k = R U_k^{-1}, so the
// frequency loops are plain GEMMs.  Pi_k comes from the FP64 MFMA SYRK of the E_RPA
// path and every Cholesky factorisation (1 + l Pi_k, the Gram matrices of the
// basis, 1 + T^T I T) runs on the library's own blocked kernels; the plain
// GEMMs and triangular solves run on rocBLAS, and the inverse and the
// dense symmetric eigenproblems (S and T^T I T when the Cholesky routes do not
// apply; the reference solves them with LAPACK dsyevd) on rocSOLVER.
// Column-major throughout (Fortran layout): Rc(P, ai) = R[ai][P] with leading
// dimension N.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "internal.h"

int launch_guess(brpa_ctx *, double *O, int64_t m, int64_t n, uint64_t seed);
int launch_normalize_columns(brpa_ctx *, double *X, int64_t m, int64_t n);
int launch_row_scale(brpa_ctx *, double *out, const double *in, const double *s, int64_t m, int64_t n);
int launch_row_scale_acc(brpa_ctx *, double *io, const double *in, const double *s, double a,
                         int64_t m, int64_t n);
int launch_select_scale(brpa_ctx *, double *W, const double *U, const double *ev, int64_t m,
                        int64_t c0, int64_t nt);
int launch_col_scale(brpa_ctx *, double *out, const double *in, const double *s, int64_t m,
                     int64_t n);
int launch_scale(brpa_ctx *, double *x, double a, int64_t n);
int launch_add_diag(brpa_ctx *, double *M, int64_t n, double a);
int launch_zero_lower(brpa_ctx *, double *M, int64_t n);
int launch_fill(brpa_ctx *, double *x, int64_t n, double v);
int launch_eig_rpa(brpa_ctx *, const double *RU, int64_t m, int64_t n, const double *A, double *out);
int launch_transpose(brpa_ctx *, const double *in, int64_t m, int64_t n, double *out);
int launch_pair_dot(brpa_ctx *, const double *W1, const double *W2, const double *B, int64_t n_rows,
                    int64_t nt, int64_t stride, int npair, double *part);
int sosex_nchunk(int64_t n_rows, int64_t nt);
int sosex_i8_vpad(int64_t v);
size_t sosex_i8_slice_bytes(int64_t nblk, int64_t ncols, int64_t v);
// rs: optional row scale (X[r][c] * rs[r] is sliced)
int launch_sosex_i8_split(brpa_ctx *, const double *X, int64_t nblk, int64_t v, int64_t ncols,
                          double *scale, int8_t *out, const double *rs = nullptr);
int launch_colslice_i8(brpa_ctx *, const double *X, int64_t ld, int64_t len, int64_t ncols,
                       const double *rs, int64_t kpad, double *scale, int8_t *out);
// one row-major v x ncols block, per column over the v rows (scaled by rs),
// column maxima in parallel over row chunks (pm: rowslice_i8_pm_len doubles)
int64_t rowslice_i8_pm_len(int64_t v, int64_t ncols);
int launch_rowslice_i8(brpa_ctx *, const double *X, int64_t v, int64_t ncols, const double *rs,
                       double *pm, double *scale, int8_t *out);
// kreal: the contraction length (flop accounting); every launch is timed with
// an event pair, kstat_collect adds the times once the stream is idle
int launch_gemm_i8(brpa_ctx *, const int8_t *SA, const double *sa, int64_t M, const int8_t *SB,
                   const double *sb, int64_t N, int64_t kpad, double alpha, const double *rs,
                   double beta, double *out, int64_t sm, int64_t sn, int64_t kreal);
int64_t sosex_i8_nparts(int64_t o, int64_t N, int64_t NT);
int launch_sosex_i8_pairs(brpa_ctx *, const int8_t *SR, const int8_t *SV, const double *scR,
                          const double *scV, const double *B, int64_t o, int64_t v, int64_t N,
                          int64_t NT, double *part, double *out2);

namespace {

constexpr double kPi = 3.14159265358979323846;

int rb_check(brpa_ctx *ctx, rocblas_status st, const char *what) {
    if (st != rocblas_status_success)
        return brpa_fail(ctx, BRPA_ERR_SOLVER, "%s failed: rocblas status %d", what, (int)st);
    return BRPA_OK;
}

int handle(brpa_ctx *ctx, rocblas_handle *h) {
    if (!ctx->rocblas) {
        rocblas_handle hh;
        if (rocblas_create_handle(&hh) != rocblas_status_success)
            return brpa_fail(ctx, BRPA_ERR_SOLVER, "rocblas_create_handle failed");
        ctx->rocblas = hh;
    }
    *h = static_cast<rocblas_handle>(ctx->rocblas);
    BRPA_TRY(rb_check(ctx, rocblas_set_stream(*h, ctx->stream), "rocblas_set_stream"));
    return BRPA_OK;
}

// Column-major C(m x n) = alpha op(A) op(B) + beta C
int gemm(brpa_ctx *ctx, rocblas_handle h, bool ta, bool tb, int64_t m, int64_t n, int64_t k,
         double alpha, const double *A, int64_t lda, const double *B, int64_t ldb, double beta,
         double *C, int64_t ldc) {
    if (m <= 0 || n <= 0) return BRPA_OK;
    return rb_check(ctx,
                    rocblas_dgemm(h, ta ? rocblas_operation_transpose : rocblas_operation_none,
                                  tb ? rocblas_operation_transpose : rocblas_operation_none,
                                  (rocblas_int)m, (rocblas_int)n, (rocblas_int)k, &alpha, A,
                                  (rocblas_int)lda, B, (rocblas_int)ldb, &beta, C, (rocblas_int)ldc),
                    "rocblas_dgemm");
}

// eigenpairs of the symmetric matrix whose column-major upper triangle A holds.
// rocSOLVER calls are serialised over the whole library (solver_mutex): worker
// contexts of the exchange batch run them from several host threads.
int syevd(brpa_ctx *ctx, rocblas_handle h, int64_t n, double *A, double *D, double *E, int *info) {
    int hinfo = 0;
    {
        std::lock_guard<std::mutex> lk(solver_mutex());
        BRPA_TRY(rb_check(ctx,
                          rocsolver_dsyevd(h, rocblas_evect_original, rocblas_fill_upper,
                                           (rocblas_int)n, A, (rocblas_int)n, D, E, info),
                          "rocsolver_dsyevd"));
        BRPA_TRY(copy_sync(ctx, &hinfo, info, sizeof(int), hipMemcpyDeviceToHost));
    }
    if (hinfo != 0)
        return brpa_fail(ctx, BRPA_ERR_SOLVER, "dsyevd did not converge (info=%d, n=%lld)", hinfo,
                         (long long)n);
    return BRPA_OK;
}

// inverse of A = U^T U from its factor U (column-major upper triangle), in place
int potri(brpa_ctx *ctx, rocblas_handle h, int64_t n, double *U, int *info) {
    std::lock_guard<std::mutex> lk(solver_mutex());
    BRPA_TRY(rb_check(ctx, rocsolver_dpotri(h, rocblas_fill_upper, (rocblas_int)n, U,
                                            (rocblas_int)n, info),
                      "rocsolver_dpotri"));
    BRPA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return BRPA_OK;
}

int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

bool finite_all(const double *p, int64_t n) {
    for (int64_t i = 0; i < n; ++i)
        if (!std::isfinite(p[i])) return false;
    return true;
}

int host_vec(brpa_ctx *ctx, std::vector<double> &v, const double *p, int64_t n) {
    v.resize(n);
    if (is_device_ptr(p))
        BRPA_TRY(copy_sync(ctx, v.data(), p, 8 * n, hipMemcpyDeviceToHost));
    else
        std::memcpy(v.data(), p, 8 * n);
    return BRPA_OK;
}

int grow(brpa_ctx *ctx, double **p, size_t *cap, size_t n) {
    if (n <= *cap) return BRPA_OK;
    if (*p) BRPA_HIP(ctx, hipFree(*p));
    *p = nullptr;
    *cap = 0;
    hipError_t e = hipMalloc(p, std::max<size_t>(n, 1) * sizeof(double));
    if (e != hipSuccess) {
        (void)hipGetLastError();
        *p = nullptr;
        return brpa_fail(ctx, BRPA_ERR_OOM, "T2 result buffer of %zu doubles: %s", n,
                         hipGetErrorString(e));
    }
    *cap = n;
    return BRPA_OK;
}

}  // namespace

int rocblas_handle_of(brpa_ctx *ctx, void **h) {
    rocblas_handle hh;
    BRPA_TRY(handle(ctx, &hh));
    *h = hh;
    return BRPA_OK;
}

void t2_release(brpa_ctx *ctx) {
    if (ctx->t2_A) (void)hipFree(ctx->t2_A);
    if (ctx->t2_V) (void)hipFree(ctx->t2_V);
    ctx->t2_A = ctx->t2_V = nullptr;
    ctx->t2_A_cap = ctx->t2_V_cap = 0;
    if (ctx->rocblas) (void)rocblas_destroy_handle(static_cast<rocblas_handle>(ctx->rocblas));
    ctx->rocblas = nullptr;
}


int t2_build(brpa_ctx *ctx, int precision, const double *R, int64_t naux, const double *eocc,
             int64_t nocc, const double *evir, int64_t nvir, const double *freqs,
             const double *weights, int64_t nfreqs, double lambda, double cutoff, int64_t nguess,
             uint64_t seed, int64_t *nvecs_t2, bool factored) {
    if (!ctx) return BRPA_ERR_ARG;
    if (precision != BRPA_FP64 && precision != BRPA_SPLIT_I8)
        return brpa_fail(ctx, BRPA_ERR_ARG,
                         "brpa_t2_compute: precision must be BRPA_FP64 or BRPA_SPLIT_I8");
    if (!R || !eocc || !evir || !freqs || !weights || !nvecs_t2 || naux <= 0 || nocc <= 0 ||
        nvir <= 0 || nfreqs <= 0 || !(lambda > 0.0))
        return brpa_fail(ctx, BRPA_ERR_ARG, "brpa_t2_compute: bad arguments");
    const int64_t K = nocc * nvir, N = naux, F = nfreqs;
    if (K > INT32_MAX || N > 65536)
        return brpa_fail(ctx, BRPA_ERR_ARG, "brpa_t2_compute: sizes beyond the supported range");
    // split-int8 amplitude contractions: exact int32 level sums need K <= 65536
    const bool split = precision == BRPA_SPLIT_I8;
    if (split && K > I8_KMAX)
        return brpa_fail(ctx, BRPA_ERR_ARG,
                         "brpa_t2_compute: nocc*nvir = %lld exceeds %d, the largest contraction "
                         "with exact int32 sums in the split-int8 form", (long long)K, I8_KMAX);
    // GuessNVecsT2 as given, ceiling(3 NVecsPiU) by default (rpa.f90:1696), with no
    // cap at o*v: the reference's I*Omega has GuessNVecsT2 columns and its Gram
    // matrix is NG x NG whatever o*v is (rpa_CC_Doubles.f90:543, 595-597)
    const int64_t NG = nguess > 0 ? nguess : (int64_t)std::ceil(3.0 * (double)N);
    if (NG > INT32_MAX / 2)
        return brpa_fail(ctx, BRPA_ERR_ARG, "brpa_t2_compute: %lld guess vectors are beyond the "
                         "supported range", (long long)NG);
    // NG >= o*v: the guess space oversamples the whole ov space, S has rank <= o*v,
    // and the basis is found in the ov space itself (see the oversampled route below)
    const bool over = NG >= K;
    std::vector<double> eo, ev, fx, fw;
    BRPA_TRY(host_vec(ctx, eo, eocc, nocc));
    BRPA_TRY(host_vec(ctx, ev, evir, nvir));
    BRPA_TRY(host_vec(ctx, fx, freqs, F));
    BRPA_TRY(host_vec(ctx, fw, weights, F));
    if (!finite_all(eo.data(), nocc) || !finite_all(ev.data(), nvir) || !finite_all(fx.data(), F) ||
        !finite_all(fw.data(), F))
        return brpa_fail(ctx, BRPA_ERR_NONFINITE, "brpa_t2_compute: non-finite input");
    if (!(*std::min_element(ev.begin(), ev.end()) - *std::max_element(eo.begin(), eo.end()) > 0.0))
        return brpa_fail(ctx, BRPA_ERR_GAP, "Found invalid orbital energy gap.");
    BRPA_HIP(ctx, hipSetDevice(ctx->device));
    rocblas_handle h;
    BRPA_TRY(handle(ctx, &h));

    // chi_k(ai) = -4 d/(d^2+u_k^2) (rpa_CC_Chi, rpa/rpa_CC_Doubles.f90:276-292)
    std::vector<double> chi(F * K);
    for (int64_t k = 0; k < F; ++k)
        for (int64_t i = 0; i < nocc; ++i)
            for (int64_t a = 0; a < nvir; ++a) {
                double d = ev[a] - eo[i];
                chi[k * K + i * nvir + a] = -(4.0 * d) / (d * d + fx[k] * fx[k]);
            }

    const bool rdev = is_device_ptr(R);
    // order of the dense symmetric matrices of the basis step: S (NG x NG), or
    // in the oversampled route the ov x ov Gram matrix of the other side
    const int64_t nd = over ? K : NG;
    // oversampled route: chi_k(ai) (w_k l / (4 pi))^(1/2), so that
    // I = -sum_k Y_k^T Y_k with Y_k = Rt_k^T diag(chis_k)
    double *chis = nullptr;
    std::vector<double> hchis;
    if (over) {
        hchis.resize(F * K);
        for (int64_t k = 0; k < F; ++k) {
            const double sk = std::sqrt(std::fabs(fw[k] * lambda) / (4.0 * kPi));
            for (int64_t i = 0; i < K; ++i) hchis[k * K + i] = chi[k * K + i] * sk;
        }
    }
    // workspace layout
    const int64_t nsteps = cdiv(N, CHOL_NB);
    FragDev fd{};
    double *dR = nullptr, *dchi, *omega, *tmp, *iom, *Y, *Sm, *sev, *sE, *Uinv, *Rt;
    int *dinfo;
    int64_t *dchol = nullptr, *dchol2 = nullptr;
    FragDev fm{};                  // one dense matrix for the library's Cholesky
    FragDev *dfm = nullptr;
    // split-int8 operands, 5 slices each: chi_k Rt_k per column P over ai for
    // every k (SCR), Rt_k per row ai over P (SRR), Omega and then T per column
    // over ai (SOT), Y_k per column over P (SY)
    const int64_t Kp = split ? sosex_i8_vpad(K) : 0, Np = cdiv(N, 64) * 64;
    int8_t *SCR = nullptr, *SRR = nullptr, *SOT = nullptr, *SY = nullptr;
    double *scCR = nullptr, *scRR = nullptr, *scOT = nullptr, *scY = nullptr, *pmx = nullptr;
    std::vector<int64_t> pre = {0, 0, 0, 0, 0, 0};
    const int64_t ntile = cdiv(N, SYRK_T);
    pre[1] = cdiv(F * K, 256);                      // dtab
    pre[3] = ntile * (ntile + 1) / 2 * cdiv(F, 4);  // syrk
    pre[5] = cdiv(F * N, 256);                      // set_diag
    int64_t *dpre = nullptr;
    FragDev *dfd = nullptr;
    auto layout = [&](Arena &ar) {
        dfd = ar.take<FragDev>(1);
        dpre = ar.take<int64_t>(pre.size());
        dR = rdev ? const_cast<double *>(R) : ar.take<double>(K * N);
        fd.dtab = ar.take<double>(F * K);
        fd.Pi = ar.take<double>(F * N * N);
        Uinv = ar.take<double>(F * N * N);
        Rt = ar.take<double>(F * K * N);
        fd.diag = ar.take<double>(F * N);
        fd.weights = ar.take<double>(F);
        fd.freqs = ar.take<double>(F);
        fd.eocc = ar.take<double>(nocc);
        fd.evir = ar.take<double>(nvir);
        dchi = ar.take<double>(F * K);
        omega = ar.take<double>(K * NG);
        // oversampled route: tmp holds the explicit ov x ov I and Sm the ov x ov
        // Gram matrix of the other side; no Y and no split operands
        tmp = ar.take<double>(over ? K * K : K * NG);
        iom = ar.take<double>(K * NG);
        Y = over ? nullptr : ar.take<double>(N * NG);
        Sm = ar.take<double>(nd * nd);
        sev = ar.take<double>(nd);
        sE = ar.take<double>(nd);
        dinfo = ar.take<int>(F + 2);
        fd.linv = ar.take<double>(F * CHOL_NB * CHOL_NB);
        fd.ldpart = ar.take<double>(F * nsteps * 2);
        fd.status = ar.take<int>(F);
        dchol = ar.take<int64_t>(chol_pre_len(N));
        dfm = ar.take<FragDev>(1);
        fm.linv = ar.take<double>(CHOL_NB * CHOL_NB);
        fm.ldpart = ar.take<double>(cdiv(nd, CHOL_NB) * 2);
        fm.status = ar.take<int>(1);
        dchol2 = ar.take<int64_t>(chol_pre_len(nd));
        if (over) chis = ar.take<double>(F * K);
        if (split && !over) {
            SCR = ar.take<int8_t>(F * 5 * N * Kp);
            SRR = ar.take<int8_t>(5 * K * Np);
            SOT = ar.take<int8_t>(5 * NG * Kp);
            SY = ar.take<int8_t>(5 * NG * Np);
            scCR = ar.take<double>(F * N);
            scRR = ar.take<double>(K);
            scOT = ar.take<double>(NG);
            scY = ar.take<double>(NG);
            pmx = ar.take<double>(rowslice_i8_pm_len(K, N));
        }
    };
    Arena count;
    layout(count);
    Arena ar;
    BRPA_TRY(ws_reserve(ctx, count.off, &ar));
    layout(ar);
    fd.R = dR;
    fd.K = (int)K;
    fd.N = (int)N;
    fd.F = (int)F;
    fd.nocc = (int)nocc;
    fd.nvir = (int)nvir;
    fd.nsteps = (int)nsteps;
    fd.keep_l11 = 1;
    if (!rdev) BRPA_TRY(to_device(ctx, dR, R, 8 * K * N));
    BRPA_TRY(launch_fill_int(ctx, fd.status, F, 0));
    // host inputs go up with completed (stream-ordered, synchronised) copies:
    // nothing of a pageable host buffer is read after this call returns
    BRPA_TRY(copy_sync(ctx, const_cast<double *>(fd.weights), fw.data(), 8 * F,
                       hipMemcpyHostToDevice));
    BRPA_TRY(copy_sync(ctx, const_cast<double *>(fd.freqs), fx.data(), 8 * F,
                       hipMemcpyHostToDevice));
    BRPA_TRY(copy_sync(ctx, const_cast<double *>(fd.eocc), eo.data(), 8 * nocc,
                       hipMemcpyHostToDevice));
    BRPA_TRY(copy_sync(ctx, const_cast<double *>(fd.evir), ev.data(), 8 * nvir,
                       hipMemcpyHostToDevice));
    BRPA_TRY(copy_sync(ctx, dchi, chi.data(), 8 * F * K, hipMemcpyHostToDevice));
    if (over) BRPA_TRY(copy_sync(ctx, chis, hchis.data(), 8 * F * K, hipMemcpyHostToDevice));
    BRPA_TRY(copy_sync(ctx, dfd, &fd, sizeof(FragDev), hipMemcpyHostToDevice));
    BRPA_TRY(copy_sync(ctx, dpre, pre.data(), 8 * pre.size(), hipMemcpyHostToDevice));
    BRPA_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    if (!ctx->stage_keep) stage_reset(ctx);
    kstat_drop(ctx);   // event pairs left by a call that failed are dropped
    // stages whose events were recorded in this call (collected at the end)
    bool staged[BRPA_NSTAGES] = {};
    auto begin = [&](int st) -> int {
        staged[st] = true;
        return stage_begin(ctx, st);
    };
    BRPA_TRY(begin(BRPA_STAGE_T2_PI));

    // Pi_k for all k (FP64 MFMA SYRK), then 1 + l Pi_k and its Cholesky factor
    BRPA_TRY(launch_dtab(ctx, dfd, dpre, 1, pre[1]));
    BRPA_TRY(launch_syrk_f64(ctx, dfd, dpre + 2, 1, pre[3]));
    if (lambda != 1.0) BRPA_TRY(launch_scale(ctx, fd.Pi, lambda, F * N * N));
    BRPA_TRY(launch_set_diag(ctx, dfd, dpre + 4, 1, pre[5], 0));
    // all F factors in one batched blocked Cholesky (the library's own kernels).
    // The factor L is the lower triangle of the row-major matrix, i.e. the
    // upper triangle U = L^T of the column-major view the solvers below read
    // (A = U^T U): hence rocblas_fill_upper there.
    BRPA_TRY(chol_single(ctx, dfd, dchol, N, F));
    // Rt_k = R U_k^-1 (C Rt_k[ai][P], the layout of R), so that
    // R (1 + l Pi_k)^-1 R^T = Rt_k Rt_k^T: no triangular solves in the
    // frequency loops.  Column-major Rt_k^T = U_k^-T Rc with U_k^-1 from the
    // batched triangular inverse, its lower triangle cleared for the GEMM.
    BRPA_TRY(rb_check(ctx,
                      rocblas_dtrtri_strided_batched(h, rocblas_fill_upper,
                                                     rocblas_diagonal_non_unit, (rocblas_int)N,
                                                     fd.Pi, (rocblas_int)N, N * N, Uinv,
                                                     (rocblas_int)N, N * N, (rocblas_int)F),
                      "rocblas_dtrtri_strided_batched"));
    for (int64_t k = 0; k < F; ++k) {
        BRPA_TRY(launch_zero_lower(ctx, Uinv + k * N * N, N));
        BRPA_TRY(gemm(ctx, h, true, false, N, K, N, 1.0, Uinv + k * N * N, N, dR, N, 0.0,
                      Rt + k * K * N, N));
    }
    BRPA_TRY(stage_end(ctx, BRPA_STAGE_T2_PI));

    // Cholesky factor of one dense n x n matrix M in place with the same
    // kernels: M = U^T U, U in the column-major upper triangle (the row-major
    // lower one the kernels read and write); *okv = every pivot positive
    auto chol_dense = [&](double *M, int64_t n, bool *okv) -> int {
        fm.Pi = M;
        fm.N = (int)n;
        fm.F = 1;
        fm.nsteps = (int)cdiv(n, CHOL_NB);
        fm.keep_l11 = 1;
        BRPA_TRY(launch_fill_int(ctx, fm.status, 1, 0));
        BRPA_TRY(copy_sync(ctx, dfm, &fm, sizeof(FragDev), hipMemcpyHostToDevice));
        BRPA_TRY(chol_single(ctx, dfm, dchol2, n, 1));
        int st = 0;
        BRPA_TRY(copy_sync(ctx, &st, fm.status, sizeof(int), hipMemcpyDeviceToHost));
        *okv = (st == 0);
        return BRPA_OK;
    };

    // the stage times of this call, once the stream is idle
    auto collect = [&]() -> int {
        BRPA_HIP(ctx, hipStreamSynchronize(ctx->stream));
        for (int st = 0; st < BRPA_NSTAGES; ++st)
            if (staged[st]) BRPA_TRY(stage_collect(ctx, st));
        return kstat_collect(ctx);
    };
    const double one = 1.0;
    int64_t nt = 0;
    double *Tkeep = iom;    // K x nt orthonormal basis of the kept subspace
    bool ident = false;     // the kept subspace is the whole ov space, T = 1
    double *TIT = Sm;       // nt x nt, T^T I T
    if (over) {
        // Oversampled route, NG >= o*v.  With X = I Omega (columns normalised,
        // K x NG), S = X^T X (NG x NG) and M = X X^T (K x K) have the same
        // nonzero eigenvalues, and S has NG - K more that are zero; the kept
        // vectors T = X U_s s^{-1/2} of the reference (rpa_CC_Doubles.f90:
        // 595-624) are the eigenvectors of M with eigenvalues above the cutoff.
        // When all K exceed it -- exactly when M - cutoff*1 has a Cholesky
        // factor -- T spans the whole ov space, NT2 = K, and T = 1 gives the
        // same Rayleigh-Ritz pairs; otherwise the K x K eigensolve of M decides.
        // I is formed explicitly: I = -sum_k Y_k^T Y_k with Y_k = Rt_k^T
        // diag(chi_k) (w_k l/(4 pi))^(1/2) costs 2 F K^2 N flop, less than I*Omega
        // through Rt_k (4 F K N NG) whenever K <= NG.
        double *Im = tmp;   // K x K
        BRPA_TRY(begin(BRPA_STAGE_T2_IOMEGA));
        for (int64_t k = 0; k < F; ++k) {
            double *Yk = Rt + k * K * N;   // column-major N x K, scaled in place
            BRPA_TRY(launch_col_scale(ctx, Yk, Yk, chis + k * K, N, K));
            const double a = fw[k] * lambda >= 0.0 ? -1.0 : 1.0;
            BRPA_TRY(gemm(ctx, h, true, false, K, K, N, a, Yk, N, Yk, N, k == 0 ? 0.0 : 1.0, Im,
                          K));
        }
        // I Omega, columns normalised (rpa_CC_Doubles.f90:572-588)
        BRPA_TRY(launch_guess(ctx, omega, K, NG, seed));
        BRPA_TRY(launch_normalize_columns(ctx, omega, K, NG));
        BRPA_TRY(gemm(ctx, h, false, false, K, NG, K, 1.0, Im, K, omega, K, 0.0, iom, K));
        BRPA_TRY(launch_normalize_columns(ctx, iom, K, NG));
        BRPA_TRY(stage_end(ctx, BRPA_STAGE_T2_IOMEGA));
        BRPA_TRY(begin(BRPA_STAGE_T2_BASIS));
        double *Mm = Sm;    // K x K
        BRPA_TRY(gemm(ctx, h, false, true, K, K, NG, 1.0, iom, K, iom, K, 0.0, Mm, K));
        BRPA_TRY(launch_add_diag(ctx, Mm, K, -cutoff));
        bool full = false;
        BRPA_TRY(chol_dense(Mm, K, &full));
        if (full) {
            nt = K;
            ident = true;
            TIT = Im;       // T^T I T = I
        } else {
            BRPA_TRY(begin(BRPA_STAGE_T2_EIG));
            BRPA_TRY(gemm(ctx, h, false, true, K, K, NG, 1.0, iom, K, iom, K, 0.0, Mm, K));
            BRPA_TRY(syevd(ctx, h, K, Mm, sev, sE, dinfo + F));
            BRPA_TRY(stage_end(ctx, BRPA_STAGE_T2_EIG));
            std::vector<double> hs(K);
            BRPA_TRY(copy_sync(ctx, hs.data(), sev, 8 * K, hipMemcpyDeviceToHost));
            // counted from the largest down, stopping at the first one below (600-607)
            for (int64_t i = K - 1; i >= 0; --i) {
                if (hs[i] > cutoff)
                    ++nt;
                else
                    break;
            }
            Tkeep = Mm + K * (K - nt);   // the eigenvectors of the nt largest eigenvalues
        }
        BRPA_TRY(stage_end(ctx, BRPA_STAGE_T2_BASIS));
        if (nt > 0 && !ident) {
            // T^T I T (rpa_CC_MIM_v5, 425-488) from the explicit I
            BRPA_TRY(begin(BRPA_STAGE_T2_TIT));
            double *Z = omega;   // K x nt
            TIT = iom;           // nt x nt
            BRPA_TRY(gemm(ctx, h, false, false, K, nt, K, 1.0, Im, K, Tkeep, K, 0.0, Z, K));
            BRPA_TRY(gemm(ctx, h, true, false, nt, nt, K, 1.0, Tkeep, K, Z, K, 0.0, TIT, nt));
            BRPA_TRY(stage_end(ctx, BRPA_STAGE_T2_TIT));
        }
    } else {
        // I * Omega (rpa_CC_IM_v5, rpa/rpa_CC_Doubles.f90:369-422)
        BRPA_TRY(begin(BRPA_STAGE_T2_IOMEGA));
        BRPA_TRY(launch_guess(ctx, omega, K, NG, seed));
        BRPA_TRY(launch_normalize_columns(ctx, omega, K, NG));
        BRPA_HIP(ctx, hipMemsetAsync(iom, 0, 8 * K * NG, ctx->stream));
        // frequency-independent operand of the split form, once per fragment
        if (split) BRPA_TRY(launch_colslice_i8(ctx, omega, K, K, NG, nullptr, Kp, scOT, SOT));
        for (int64_t k = 0; k < F; ++k) {
            const double *ck = dchi + k * K;
            const double coef = -fw[k] * lambda / (4.0 * kPi);
            const double *Rtk = Rt + k * K * N;
            if (split) {
                // Y = (chi_k Rt_k)^T Omega, I Omega += coef chi_k Rt_k Y; the chi_k Rt_k
                // slices are kept for T^T I T
                int8_t *SCRk = SCR + k * 5 * N * Kp;
                double *scCRk = scCR + k * N;
                BRPA_TRY(launch_rowslice_i8(ctx, Rtk, K, N, ck, pmx, scCRk, SCRk));
                BRPA_TRY(launch_gemm_i8(ctx, SCRk, scCRk, N, SOT, scOT, NG, Kp, 1.0, nullptr, 0.0,
                                        Y, 1, N, K));
                BRPA_TRY(launch_colslice_i8(ctx, Rtk, N, N, K, nullptr, Np, scRR, SRR));
                BRPA_TRY(launch_colslice_i8(ctx, Y, N, N, NG, nullptr, Np, scY, SY));
                BRPA_TRY(launch_gemm_i8(ctx, SRR, scRR, K, SY, scY, NG, Np, coef, ck, 1.0, iom, 1,
                                        K, N));
                continue;
            }
            BRPA_TRY(launch_row_scale(ctx, tmp, omega, ck, K, NG));
            BRPA_TRY(gemm(ctx, h, false, false, N, NG, K, 1.0, Rtk, N, tmp, K, 0.0, Y, N));
            BRPA_TRY(gemm(ctx, h, true, false, K, NG, N, 1.0, Rtk, N, Y, N, 0.0, tmp, K));
            BRPA_TRY(launch_row_scale_acc(ctx, iom, tmp, ck, coef, K, NG));
        }
        BRPA_TRY(launch_normalize_columns(ctx, iom, K, NG));
        BRPA_TRY(stage_end(ctx, BRPA_STAGE_T2_IOMEGA));

        // S = (I Omega)^T (I Omega) (rpa_CC_Doubles.f90:595-609).  The reference
        // keeps the eigenvectors of S with eigenvalues above the cutoff and forms
        // the orthonormal basis T = (I Omega) U_s s^{-1/2}.  When every eigenvalue
        // exceeds the cutoff -- exactly when S - cutoff*1 has a Cholesky factor --
        // T spans all of span(I Omega), and any orthonormal basis of that span
        // gives the same Rayleigh-Ritz pairs below; it is then built by Cholesky-QR
        // (the shifted factor, then one unshifted pass) instead of the NG x NG
        // eigensolve.  Otherwise the eigensolve decides which vectors are kept.
        // Two passes suffice: with every eigenvalue above the cutoff, X = I Omega
        // has kappa(X) <= (NG / cutoff)^(1/2) < 1e8, the first pass leaves
        // X1^T X1 = 1 + cutoff U1^-T U1^-1 + O(eps kappa^2), and the second pass
        // restores orthonormality to O(eps kappa(X1)^2) -- never worse than the
        // eps s_max / s_min of the reference's own eigenvector basis.
        BRPA_TRY(begin(BRPA_STAGE_T2_BASIS));
        // Gram matrices as full GEMMs: rocBLAS dgemm takes a third of dsyrk's time here
        BRPA_TRY(gemm(ctx, h, true, false, NG, NG, K, 1.0, iom, K, iom, K, 0.0, Sm, NG));
        bool cholqr = false;
        BRPA_TRY(launch_add_diag(ctx, Sm, NG, -cutoff));
        BRPA_TRY(chol_dense(Sm, NG, &cholqr));
        if (cholqr) {
            nt = NG;
            for (int pass = 0; pass < 2; ++pass) {
                if (pass > 0) {
                    BRPA_TRY(gemm(ctx, h, true, false, NG, NG, K, 1.0, iom, K, iom, K, 0.0, Sm,
                                  NG));
                    bool okp = false;
                    BRPA_TRY(chol_dense(Sm, NG, &okp));
                    if (!okp)
                        return brpa_fail(ctx, BRPA_ERR_SOLVER,
                                         "Cholesky-QR pass %d of the T2 basis broke down", pass);
                }
                // X <- X U^{-1}, so that X^T X = 1
                BRPA_TRY(rb_check(ctx,
                                  rocblas_dtrsm(h, rocblas_side_right, rocblas_fill_upper,
                                                rocblas_operation_none, rocblas_diagonal_non_unit,
                                                (rocblas_int)K, (rocblas_int)NG, &one, Sm,
                                                (rocblas_int)NG, iom, (rocblas_int)K),
                                  "rocblas_dtrsm"));
            }
        } else {
            BRPA_TRY(begin(BRPA_STAGE_T2_EIG));
            BRPA_TRY(gemm(ctx, h, true, false, NG, NG, K, 1.0, iom, K, iom, K, 0.0, Sm, NG));
            BRPA_TRY(syevd(ctx, h, NG, Sm, sev, sE, dinfo + F));
            BRPA_TRY(stage_end(ctx, BRPA_STAGE_T2_EIG));
            std::vector<double> hs(NG);
            BRPA_TRY(copy_sync(ctx, hs.data(), sev, 8 * NG, hipMemcpyDeviceToHost));
            for (int64_t i = NG - 1; i >= 0; --i) {
                if (hs[i] > cutoff)
                    ++nt;
                else
                    break;
            }
        }
        if (nt > 0) {
            // T = (I Omega) U_s s^{-1/2}; TIT = T^T I T (rpa_CC_MIM_v5, 425-488)
            double *T = tmp;               // K x nt
            double *G = Y;                 // N x nt
            if (!cholqr) {
                double *Wm = omega;        // NG x nt   (omega no longer needed)
                BRPA_TRY(launch_select_scale(ctx, Wm, Sm, sev, NG, NG - nt, nt));
                BRPA_TRY(gemm(ctx, h, false, false, K, nt, NG, 1.0, iom, K, Wm, NG, 0.0, T, K));
                BRPA_HIP(ctx, hipMemcpyAsync(Tkeep, T, 8 * K * nt, hipMemcpyDeviceToDevice,
                                             ctx->stream));
            }
            BRPA_TRY(stage_end(ctx, BRPA_STAGE_T2_BASIS));
            BRPA_TRY(begin(BRPA_STAGE_T2_TIT));
            if (split)
                BRPA_TRY(launch_colslice_i8(ctx, Tkeep, K, K, nt, nullptr, Kp, scOT, SOT));
            for (int64_t k = 0; k < F; ++k) {
                const double *ck = dchi + k * K;
                if (split) {
                    // G = (chi_k Rt_k)^T T
                    BRPA_TRY(launch_gemm_i8(ctx, SCR + k * 5 * N * Kp, scCR + k * N, N, SOT, scOT,
                                            nt, Kp, 1.0, nullptr, 0.0, G, 1, N, K));
                } else {
                    BRPA_TRY(launch_row_scale(ctx, T, Tkeep, ck, K, nt));
                    BRPA_TRY(gemm(ctx, h, false, false, N, nt, K, 1.0, Rt + k * K * N, N, T, K,
                                  0.0, G, N));
                }
                // G = Rt_k^T chi_k T = U_k^-T R^T chi_k T, so G^T G = T^T chi R (1 + l Pi_k)^-1 R^T chi T
                BRPA_TRY(gemm(ctx, h, true, false, nt, nt, N, -fw[k] * lambda / (4.0 * kPi), G, N,
                              G, N, k == 0 ? 0.0 : 1.0, TIT, nt));
            }
            BRPA_TRY(stage_end(ctx, BRPA_STAGE_T2_TIT));
        } else {
            BRPA_TRY(stage_end(ctx, BRPA_STAGE_T2_BASIS));
        }
    }
    *nvecs_t2 = nt;
    ctx->t2_nt = nt;
    ctx->t2_nai = K;
    ctx->t2_factored = false;
    if (nt == 0) return collect();

    BRPA_TRY(grow(ctx, &ctx->t2_A, &ctx->t2_A_cap, nt));
    BRPA_TRY(grow(ctx, &ctx->t2_V, &ctx->t2_V_cap, (size_t)nt * K));
    BRPA_TRY(begin(BRPA_STAGE_T2_CORE));
    bool done = false;
    if (factored) {
        // Energy-only callers need T2 = T C T^T, C = (1/2) TIT (1 + TIT)^{-1}, not the
        // eigenpairs.  The amplitudes A = a/(2(1+a)) are negative (I is negative
        // semidefinite on the subspace), so -C = ((1 + TIT)^{-1} - 1)/2 has a Cholesky
        // factor L, and T2 = -V' V'^T with V' = T L: the pair (A' = -1, V') gives
        // the same exchange energy.  Any breakdown falls back to the eigensolve.
        double *Msave = omega;   // nt x nt copy of TIT (omega is free here)
        BRPA_HIP(ctx, hipMemcpyAsync(Msave, TIT, 8 * nt * nt, hipMemcpyDeviceToDevice,
                                     ctx->stream));
        auto info_ok = [&](int *d, int *okv) -> int {
            int hi = 0;
            BRPA_TRY(copy_sync(ctx, &hi, d, sizeof(int), hipMemcpyDeviceToHost));
            *okv = (hi == 0);
            return BRPA_OK;
        };
        BRPA_TRY(launch_add_diag(ctx, TIT, nt, 1.0));
        bool okb = false;
        BRPA_TRY(chol_dense(TIT, nt, &okb));
        int ok = okb;
        if (ok) {
            BRPA_TRY(potri(ctx, h, nt, TIT, dinfo + F + 1));
            BRPA_TRY(info_ok(dinfo + F + 1, &ok));
        }
        if (ok) {
            BRPA_TRY(launch_scale(ctx, TIT, 0.5, nt * nt));
            BRPA_TRY(launch_add_diag(ctx, TIT, nt, -0.5));
            BRPA_TRY(chol_dense(TIT, nt, &okb));
            ok = okb;
        }
        if (ok) {
            // -C = U'^T U' (U' upper): T2 = -V' V'^T with V' = T U'^T
            BRPA_TRY(launch_zero_lower(ctx, TIT, nt));
            BRPA_TRY(launch_fill(ctx, ctx->t2_A, nt, -1.0));
            // V'^T = U' T^T, stored as V[ai][mu]; with T = 1 it is U' itself
            if (ident)
                BRPA_HIP(ctx, hipMemcpyAsync(ctx->t2_V, TIT, 8 * nt * nt, hipMemcpyDeviceToDevice,
                                             ctx->stream));
            else
                BRPA_TRY(gemm(ctx, h, false, true, nt, K, nt, 1.0, TIT, nt, Tkeep, K, 0.0,
                              ctx->t2_V, nt));
            done = true;
            ctx->t2_factored = true;
        } else {
            BRPA_HIP(ctx, hipMemcpyAsync(TIT, Msave, 8 * nt * nt, hipMemcpyDeviceToDevice,
                                         ctx->stream));
        }
    }
    if (!done) {
        // eigenpairs of T^T I T; A = a / (2 (1 + a)); V = eigvecs^T T^T (629-641)
        BRPA_TRY(syevd(ctx, h, nt, TIT, sev, sE, dinfo + F + 1));
        std::vector<double> ha(nt);
        BRPA_TRY(copy_sync(ctx, ha.data(), sev, 8 * nt, hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < nt; ++i) ha[i] = 0.5 * (ha[i] / (ha[i] + 1.0));
        BRPA_TRY(copy_sync(ctx, ctx->t2_A, ha.data(), 8 * nt, hipMemcpyHostToDevice));
        if (ident)
            BRPA_TRY(launch_transpose(ctx, TIT, nt, nt, ctx->t2_V));
        else
            BRPA_TRY(gemm(ctx, h, true, true, nt, K, nt, 1.0, TIT, nt, Tkeep, K, 0.0, ctx->t2_V,
                          nt));
    }
    BRPA_TRY(stage_end(ctx, BRPA_STAGE_T2_CORE));
    BRPA_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    std::vector<int> info(F);
    BRPA_HIP(ctx, hipMemcpyAsync(info.data(), fd.status, sizeof(int) * F, hipMemcpyDeviceToHost,
                                 ctx->stream));
    BRPA_TRY(collect());
    float ms = 0.f;
    BRPA_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    ctx->last_kernel_ms = ms;
    for (int64_t k = 0; k < F; ++k)
        if (info[k] != 0)
            return brpa_fail(ctx, BRPA_ERR_NOT_SPD, "frequency %lld: 1 + l Pi(u) is not positive definite",
                             (long long)k);
    return BRPA_OK;
}

extern "C" {

int brpa_t2_compute(brpa_ctx *ctx, int precision, const double *R, int64_t naux,
                    const double *eocc, int64_t nocc, const double *evir, int64_t nvir,
                    const double *freqs, const double *weights, int64_t nfreqs, double lambda,
                    double cutoff, int64_t nguess, uint64_t seed, int64_t *nvecs_t2) {
    int rc = t2_build(ctx, precision, R, naux, eocc, nocc, evir, nvir, freqs, weights, nfreqs,
                      lambda, cutoff, nguess, seed, nvecs_t2, false);
    if (rc != BRPA_OK && ctx) ctx->t2_nt = -1;   // no usable amplitudes after a failure
    return rc;
}

int brpa_t2_copyout(brpa_ctx *ctx, double *A, double *V) {
    if (!ctx) return BRPA_ERR_ARG;
    if (ctx->t2_nt < 0) return brpa_fail(ctx, BRPA_ERR_ARG, "no T2 result to copy");
    if (ctx->t2_nt == 0) return BRPA_OK;
    if (A) BRPA_TRY(from_device(ctx, A, ctx->t2_A, 8 * ctx->t2_nt));
    if (V) BRPA_TRY(from_device(ctx, V, ctx->t2_V, 8 * ctx->t2_nt * ctx->t2_nai));
    BRPA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return BRPA_OK;
}

// Amplitude-eigenvalue cutoff of the THC route, rpa_THC_CC_T2
// (rpa/rpa_CC_Doubles.f90:809-983), applied to the eigenpairs left by
// brpa_t2_compute.  The threshold rules and the scan over the NT2 eigenvalues
// are host work on NT2 numbers; V is compacted on the device.
int brpa_t2_cutoff(brpa_ctx *ctx, int cutoff_type, double thresh, int smooth_step,
                   double steepness, int64_t nocc, double *common_thresh, int64_t *nvecs_t2,
                   int64_t *nvecs_scaled) {
    if (!ctx) return BRPA_ERR_ARG;
    if (!common_thresh || !nvecs_t2)
        return brpa_fail(ctx, BRPA_ERR_ARG, "brpa_t2_cutoff: null output argument");
    if (ctx->t2_nt <= 0)
        return brpa_fail(ctx, BRPA_ERR_ARG, "brpa_t2_cutoff: no amplitudes (call brpa_t2_compute)");
    if (ctx->t2_factored)
        return brpa_fail(ctx, BRPA_ERR_ARG, "brpa_t2_cutoff: the amplitudes are not eigenpairs");
    const double common = *common_thresh;
    if (!std::isfinite(common) || !std::isfinite(thresh) ||
        (smooth_step && !(steepness >= 0.0 && std::isfinite(steepness))))
        return brpa_fail(ctx, BRPA_ERR_ARG, "brpa_t2_cutoff: non-finite or negative parameter");
    if (common < 0.0 && cutoff_type == BRPA_T2_CUTOFF_EIG_DIV_NELECTRON && nocc <= 0)
        return brpa_fail(ctx, BRPA_ERR_ARG, "brpa_t2_cutoff: nocc must be positive");
    BRPA_HIP(ctx, hipSetDevice(ctx->device));
    const int64_t nt = ctx->t2_nt, K = ctx->t2_nai;
    std::vector<double> a(nt);
    BRPA_TRY(copy_sync(ctx, a.data(), ctx->t2_A, 8 * nt, hipMemcpyDeviceToHost));
    // threshold K (880-930): the supersystem evaluates it, a subsystem takes the
    // supersystem's value
    double kth = 0.0;
    if (common < 0.0) {
        switch (cutoff_type) {
        case BRPA_T2_CUTOFF_EIG:
            kth = thresh;
            break;
        case BRPA_T2_CUTOFF_EIG_DIV_MAXEIG: {
            double mx = 0.0;
            for (double x : a) mx = std::max(mx, std::fabs(x));
            kth = thresh * mx;
            break;
        }
        case BRPA_T2_CUTOFF_EIG_DIV_NELECTRON:
            kth = thresh / (double)(nocc * 2);
            break;
        case BRPA_T2_CUTOFF_SUM_REJECTED: {
            // the first i with |A(i+1) + ... + A(nt)| < thresh gives K = |A(i+1)|;
            // tail sums accumulated from the end, in the order the scan visits them
            std::vector<double> tail(nt + 1, 0.0);
            for (int64_t i = nt - 1; i >= 0; --i) tail[i] = tail[i + 1] + a[i];
            kth = 0.0;
            for (int64_t i = 1; i < nt; ++i)
                if (std::fabs(tail[i]) < thresh) {
                    kth = std::fabs(a[i]);
                    break;
                }
            break;
        }
        default:
            return brpa_fail(ctx, BRPA_ERR_ARG, "brpa_t2_cutoff: invalid cutoff type %d",
                             cutoff_type);
        }
    } else {
        kth = common;
    }
    // scan in the eigensolver's order (largest |A| first); the first rejected
    // eigenvalue ends the kept set (936-959)
    const double ka = kth, kb = kth * steepness;
    int64_t nnew = nt, nscaled = 0;
    for (int64_t i = 0; i < nt; ++i) {
        const double x = std::fabs(a[i]);
        if (smooth_step) {
            if (x >= ka) continue;
            if (x > kb) {
                // rpa_CC_T2_SmoothStepFunction (29-40)
                const double mu = (2.0 * x - (ka + kb)) / (kb - ka);
                const double mu2 = mu * mu;
                const double z =
                    mu * (35.0 + mu2 * (-35.0 + mu2 * (21.0 - 5.0 * mu2))) / 16.0;
                a[i] *= 0.5 * (1.0 - z);
                ++nscaled;
                continue;
            }
        } else if (x > kth) {
            continue;
        }
        nnew = i;
        break;
    }
    if (nnew <= 0)
        return brpa_fail(ctx, BRPA_ERR_ARG,
                         "brpa_t2_cutoff: threshold %.3e leaves no amplitude eigenvectors", kth);
    if (nnew < nt) {
        // V[ai][mu] keeps its first nnew columns: K rows of nt -> K rows of nnew,
        // through the workspace (the in-place copy would overlap)
        Arena ar;
        BRPA_TRY(ws_reserve(ctx, 8 * (size_t)K * nnew + 256, &ar));
        double *tmp = ar.take<double>((size_t)K * nnew);
        BRPA_HIP(ctx, hipMemcpy2DAsync(tmp, 8 * nnew, ctx->t2_V, 8 * nt, 8 * nnew, K,
                                       hipMemcpyDeviceToDevice, ctx->stream));
        BRPA_HIP(ctx, hipMemcpyAsync(ctx->t2_V, tmp, 8 * (size_t)K * nnew,
                                     hipMemcpyDeviceToDevice, ctx->stream));
    }
    BRPA_TRY(copy_sync(ctx, ctx->t2_A, a.data(), 8 * nnew, hipMemcpyHostToDevice));
    BRPA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ctx->t2_nt = nnew;
    *nvecs_t2 = nnew;
    if (nvecs_scaled) *nvecs_scaled = nscaled;
    if (common < 0.0) *common_thresh = kth;   // kept for the subsystem calculations (980-982)
    return BRPA_OK;
}

// EigRPA_mu = 2 A_mu |R^T V_mu|^2 of the amplitudes kept on the device
// (rpa_T2_EigRPA, rpa/rpa_CC_Doubles.f90:987-1009); sum_mu EigRPA_mu is the
// direct-ring energy of the amplitudes (rpa_THC.f90:126).
int brpa_t2_eig_rpa(brpa_ctx *ctx, const double *R, int64_t naux, double *eig_rpa) {
    if (!ctx) return BRPA_ERR_ARG;
    if (!R || !eig_rpa || naux <= 0)
        return brpa_fail(ctx, BRPA_ERR_ARG, "brpa_t2_eig_rpa: bad arguments");
    if (ctx->t2_nt < 0)
        return brpa_fail(ctx, BRPA_ERR_ARG, "brpa_t2_eig_rpa: no amplitudes (call brpa_t2_compute)");
    if (ctx->t2_factored)
        return brpa_fail(ctx, BRPA_ERR_ARG, "brpa_t2_eig_rpa: the amplitudes are not eigenpairs");
    const int64_t nt = ctx->t2_nt, K = ctx->t2_nai, N = naux;
    if (nt == 0) return BRPA_OK;
    BRPA_HIP(ctx, hipSetDevice(ctx->device));
    rocblas_handle h;
    BRPA_TRY(handle(ctx, &h));
    const bool rdev = is_device_ptr(R), odev = is_device_ptr(eig_rpa);
    Arena count, ar;
    auto layout = [&](Arena &x) {
        x.take<double>((size_t)N * nt);
        if (!rdev) x.take<double>((size_t)K * N);
        if (!odev) x.take<double>(nt);
    };
    layout(count);
    BRPA_TRY(ws_reserve(ctx, count.off, &ar));
    double *RU = ar.take<double>((size_t)N * nt);
    const double *dR = R;
    if (!rdev) {
        double *t = ar.take<double>((size_t)K * N);
        BRPA_TRY(copy_sync(ctx, t, R, 8 * (size_t)K * N, hipMemcpyHostToDevice));
        dR = t;
    }
    double *out = odev ? eig_rpa : ar.take<double>(nt);
    // RU (N x nt) = R^T V: R[ai][P] is column-major N x K, V[ai][mu] column-major nt x K
    BRPA_TRY(gemm(ctx, h, false, true, N, nt, K, 1.0, dR, N, ctx->t2_V, nt, 0.0, RU, N));
    BRPA_TRY(launch_eig_rpa(ctx, RU, N, nt, ctx->t2_A, out));
    if (!odev)
        BRPA_TRY(copy_sync(ctx, eig_rpa, out, 8 * nt, hipMemcpyDeviceToHost));
    BRPA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return BRPA_OK;
}

// The THC route's layout of the amplitudes: Fortran V(ov, NT2), i.e. C
// V[mu][ai] (rpa/rpa_CC_Doubles.f90:968-975), transposed on the device.
int brpa_t2_copyout_ov(brpa_ctx *ctx, double *A, double *V) {
    if (!ctx) return BRPA_ERR_ARG;
    if (ctx->t2_nt < 0) return brpa_fail(ctx, BRPA_ERR_ARG, "no T2 result to copy");
    const int64_t nt = ctx->t2_nt, K = ctx->t2_nai;
    if (nt == 0) return BRPA_OK;
    BRPA_HIP(ctx, hipSetDevice(ctx->device));
    if (A) BRPA_TRY(from_device(ctx, A, ctx->t2_A, 8 * nt));
    if (V) {
        const bool vdev = is_device_ptr(V);
        Arena ar;
        double *dst = V;
        if (!vdev) {
            BRPA_TRY(ws_reserve(ctx, 8 * (size_t)K * nt + 256, &ar));
            dst = ar.take<double>((size_t)K * nt);
        }
        BRPA_TRY(launch_transpose(ctx, ctx->t2_V, nt, K, dst));
        if (!vdev) BRPA_TRY(copy_sync(ctx, V, dst, 8 * (size_t)K * nt, hipMemcpyDeviceToHost));
    }
    BRPA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return BRPA_OK;
}

int brpa_sosex(brpa_ctx *ctx, int precision, double *ec_exchange, double alpha, const double *V,
               const double *A, const double *R, int64_t nocc, int64_t nvir, int64_t naux,
               int64_t nvecs_t2, int s2_operator) {
    if (!ctx) return BRPA_ERR_ARG;
    if (precision != BRPA_FP64 && precision != BRPA_SPLIT_I8)
        return brpa_fail(ctx, BRPA_ERR_ARG,
                         "brpa_sosex: precision must be BRPA_FP64 or BRPA_SPLIT_I8");
    if (!ec_exchange || !V || !A || !R || nocc <= 0 || nvir <= 0 || naux <= 0 || nvecs_t2 < 0)
        return brpa_fail(ctx, BRPA_ERR_ARG, "brpa_sosex: bad arguments");
    if (precision == BRPA_SPLIT_I8 && nvir > 65536)
        return brpa_fail(ctx, BRPA_ERR_ARG,
                         "brpa_sosex: nvir = %lld exceeds 65536, the largest K with exact int32 "
                         "sums in the split-int8 form", (long long)nvir);
    *ec_exchange = 0.0;
    if (!ctx->stage_keep) stage_reset(ctx);
    const int64_t K = nocc * nvir, N = naux, nt = nvecs_t2, o = nocc, v = nvir;
    if (nt == 0) return BRPA_OK;
    if (K > INT32_MAX || N > 65536 || nt > 1000000)
        return brpa_fail(ctx, BRPA_ERR_ARG, "brpa_sosex: sizes beyond the supported range");
    BRPA_HIP(ctx, hipSetDevice(ctx->device));
    rocblas_handle h;
    BRPA_TRY(handle(ctx, &h));
    std::vector<double> hA;
    BRPA_TRY(host_vec(ctx, hA, A, nt));
    if (!finite_all(hA.data(), nt))
        return brpa_fail(ctx, BRPA_ERR_NONFINITE, "brpa_sosex: non-finite amplitudes");
    // B = A, or A/(1-4A^2) for the cumulant route (rpa_CC_Exchange.f90:62-68)
    std::vector<double> hB(nt);
    for (int64_t k = 0; k < nt; ++k) hB[k] = s2_operator ? hA[k] / (1.0 - 4.0 * hA[k] * hA[k]) : hA[k];
    const bool rdev = is_device_ptr(R), vdev = is_device_ptr(V);
    if (precision == BRPA_SPLIT_I8) {
        // fused split-int8 pair kernel: 5 int8 slices per value, 15 products,
        // exact int32 level sums, W_ij and W_ji never stored
        const int64_t nparts = sosex_i8_nparts(o, N, nt);
        double *dR, *dV, *dB, *scR, *scV, *part, *out2;
        int8_t *SR, *SV;
        auto layout = [&](Arena &ar) {
            dR = rdev ? const_cast<double *>(R) : ar.take<double>(K * N);
            dV = vdev ? const_cast<double *>(V) : ar.take<double>(K * nt);
            dB = ar.take<double>(nt);
            scR = ar.take<double>(o * N);
            scV = ar.take<double>(o * nt);
            SR = ar.take<int8_t>(sosex_i8_slice_bytes(o, N, v));
            SV = ar.take<int8_t>(sosex_i8_slice_bytes(o, nt, v));
            part = ar.take<double>(2 * nparts);
            out2 = ar.take<double>(2);
        };
        Arena count;
        layout(count);
        Arena ar;
        BRPA_TRY(ws_reserve(ctx, count.off, &ar));
        layout(ar);
        if (!rdev) BRPA_TRY(to_device(ctx, dR, R, 8 * K * N));
        if (!vdev) BRPA_TRY(to_device(ctx, dV, V, 8 * K * nt));
        BRPA_TRY(copy_sync(ctx, dB, hB.data(), 8 * nt, hipMemcpyHostToDevice));
        BRPA_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
        BRPA_TRY(stage_begin(ctx, BRPA_STAGE_SOSEX));
        BRPA_TRY(launch_sosex_i8_split(ctx, dR, o, v, N, scR, SR));
        BRPA_TRY(launch_sosex_i8_split(ctx, dV, o, v, nt, scV, SV));
        BRPA_TRY(launch_sosex_i8_pairs(ctx, SR, SV, scR, scV, dB, o, v, N, nt, part, out2));
        BRPA_TRY(stage_end(ctx, BRPA_STAGE_SOSEX));
        BRPA_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
        double h2[2];
        BRPA_TRY(copy_sync(ctx, h2, out2, sizeof(h2), hipMemcpyDeviceToHost));
        BRPA_TRY(stage_collect(ctx, BRPA_STAGE_SOSEX));
        float ms = 0.f;
        BRPA_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        ctx->last_kernel_ms = ms;
        // EcExchange1 = -sum W_ji W_ij B per pair; the partials carry (2 - delta_ij)
        *ec_exchange = (double)(-(long double)alpha * ((long double)h2[0] + (long double)h2[1]));
        return BRPA_OK;
    }
    // pairs (i, j >= i) are processed per i in batches of jb occupied indices
    const size_t wbytes_per_j = 2 * 8 * (size_t)N * nt;
    int64_t jb = std::max<int64_t>(1, std::min<int64_t>(o, (int64_t)((8ull << 30) / wbytes_per_j)));
    const int nchunk = sosex_nchunk(N, nt);
    double *dR, *dV, *dB, *W1, *W2, *part;
    auto layout = [&](Arena &ar) {
        dR = rdev ? const_cast<double *>(R) : ar.take<double>(K * N);
        dV = vdev ? const_cast<double *>(V) : ar.take<double>(K * nt);
        dB = ar.take<double>(nt);
        W1 = ar.take<double>(jb * N * nt);
        W2 = ar.take<double>(jb * N * nt);
        part = ar.take<double>(2 * o * o * nchunk);   // (i, j) pair partials
    };
    Arena count;
    layout(count);
    Arena ar;
    BRPA_TRY(ws_reserve(ctx, count.off, &ar));
    layout(ar);
    if (!rdev) BRPA_TRY(to_device(ctx, dR, R, 8 * K * N));
    if (!vdev) BRPA_TRY(to_device(ctx, dV, V, 8 * K * nt));
    BRPA_TRY(copy_sync(ctx, dB, hB.data(), 8 * nt, hipMemcpyHostToDevice));
    BRPA_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    BRPA_TRY(stage_begin(ctx, BRPA_STAGE_SOSEX));
    const double one = 1.0, zero = 0.0;
    for (int64_t i = 0; i < o; ++i) {
        for (int64_t j0 = i; j0 < o; j0 += jb) {
            const int64_t nb = std::min<int64_t>(jb, o - j0);
            // W_ij = R_i V_j^T (N x nt, K = v), j = j0..j0+nb-1 (rpa_CC_Exchange.f90:74-76)
            BRPA_TRY(rb_check(ctx,
                              rocblas_dgemm_strided_batched(
                                  h, rocblas_operation_none, rocblas_operation_transpose,
                                  (rocblas_int)N, (rocblas_int)nt, (rocblas_int)v, &one,
                                  dR + i * v * N, (rocblas_int)N, 0, dV + j0 * v * nt,
                                  (rocblas_int)nt, v * nt, &zero, W1, (rocblas_int)N, N * nt,
                                  (rocblas_int)nb),
                              "rocblas_dgemm_strided_batched"));
            // W_ji = R_j V_i^T (77-78)
            BRPA_TRY(rb_check(ctx,
                              rocblas_dgemm_strided_batched(
                                  h, rocblas_operation_none, rocblas_operation_transpose,
                                  (rocblas_int)N, (rocblas_int)nt, (rocblas_int)v, &one,
                                  dR + j0 * v * N, (rocblas_int)N, v * N, dV + i * v * nt,
                                  (rocblas_int)nt, 0, &zero, W2, (rocblas_int)N, N * nt,
                                  (rocblas_int)nb),
                              "rocblas_dgemm_strided_batched"));
            BRPA_TRY(launch_pair_dot(ctx, W1, W2, dB, N, nt, N * nt, (int)nb,
                                     part + 2 * (i * o + j0) * nchunk));
        }
    }
    BRPA_TRY(stage_end(ctx, BRPA_STAGE_SOSEX));
    BRPA_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    // a completed copy: the partials are read on the host right after it
    std::vector<double> hpart(2 * o * o * nchunk);
    BRPA_TRY(copy_sync(ctx, hpart.data(), part, 8 * hpart.size(), hipMemcpyDeviceToHost));
    long double e = 0.0L;
    for (int64_t i = 0; i < o; ++i) {
        for (int64_t j = i; j < o; ++j) {
            long double sp = 0.0L;
            for (int c = 0; c < nchunk; ++c)
                sp += (long double)hpart[2 * ((i * o + j) * nchunk + c)] +
                      (long double)hpart[2 * ((i * o + j) * nchunk + c) + 1];
            // EcExchange1 = -sum W_ji W_ij B; off-diagonal pairs count twice
            e += (j == i ? 1.0L : 2.0L) * (-sp);
        }
    }
    BRPA_TRY(stage_collect(ctx, BRPA_STAGE_SOSEX));
    float ms = 0.f;
    BRPA_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    ctx->last_kernel_ms = ms;
    *ec_exchange = (double)((long double)alpha * e);
    return BRPA_OK;
}

}  // extern "C"
