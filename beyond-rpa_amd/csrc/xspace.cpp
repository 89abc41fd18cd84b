// Exchange correction of one fragment in the explicit ov space: the energy of
// rpa_CC_T2_v5 (rpa/rpa_CC_Doubles.f90:491-642) followed by
// rpa_CC_Exchange_SOSEX (rpa/rpa_CC_Exchange.f90:36-116), for callers that
// need E_x and not the amplitude eigenpairs.
//
// The reference applies I = -sum_k w_k/(4 pi) chi_k R l(1 + l Pi_k)^{-1} R^T chi_k
// to blocks of vectors frequency by frequency, and contracts the amplitudes
// pair by pair (2 o^2 v N NT2 flop).  Here the ov x ov matrices themselves are
// formed, which is cheaper at every shape of the batched workload
// (ov <= 8.7 N):
//   1. Pi_k and the Cholesky factors 1 + l Pi_k = U_k^T U_k (as for E_RPA);
//   2. Y = [Y_1; ...; Y_F], Y_k = U_k^{-T} R^T diag(chi_k (w_k l/(4 pi))^{1/2}),
//      and I = -Y^T Y  (one SYRK over F N); U_k^{-1} by a batched triangular
//      inverse, Y_k by one batched triangular GEMM with the column scale in
//      its epilogue;
//   3. X = normalised I Omega, Omega the normalised guess vectors;
//   4. the orthonormal basis T of span(X) with the reference's eigenvalue
//      cutoff (Cholesky-QR when every Gram eigenvalue clears it, the
//      eigenvectors otherwise); when NG >= ov and every eigenvalue of X X^T
//      clears it, T = 1;
//   5. TIT = T^T (I T), 1 + TIT = U^T U, W = T U^{-1},
//      T2 = T (1/2) TIT (1 + TIT)^{-1} T^T = (T T^T - W W^T) / 2;
// Every triangular solve is a triangular inverse (GEMM levels) followed by a
// GEMM that skips the zero triangle.
//   6. E_x = -alpha sum_{aibj} (ai|bj) T2(aj,bi) by k_sosex_dot, with
//      (ai|bj) = (R R^T)(ai,bj) formed tile by tile on MFMA.
// T2 = T f(T^T I T) T^T depends only on the subspace span(T), so every basis of
// the reference's kept subspace gives the same energy.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "dense.h"

int launch_guess(brpa_ctx *, double *O, int64_t m, int64_t n, uint64_t seed);
int launch_normalize_columns(brpa_ctx *, double *X, int64_t m, int64_t n);
int launch_select_scale(brpa_ctx *, double *W, const double *U, const double *ev, int64_t m,
                        int64_t c0, int64_t nt);
int launch_col_scale(brpa_ctx *, double *out, const double *in, const double *s, int64_t m,
                     int64_t n);
int launch_scale(brpa_ctx *, double *x, double a, int64_t n);
int launch_add_diag(brpa_ctx *, double *M, int64_t n, double a);
int64_t sosex_dot_nparts(int64_t K);
int launch_sosex_dot(brpa_ctx *, const double *R, int64_t K, int64_t N, int64_t v, const double *M,
                     double gamma, double delta, double *part);

namespace {
constexpr double kPi = 3.14159265358979323846;
// partial tiles of split-k GEMMs (up to 8 splits of the largest C of a call)
constexpr int64_t kSplitWsMin = (int64_t)9 << 20;
int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

bool finite_all(const std::vector<double> &v) {
    for (double x : v)
        if (!std::isfinite(x)) return false;
    return true;
}

int host_copy(brpa_ctx *ctx, std::vector<double> &v, const double *p, int64_t n) {
    v.resize(n);
    if (is_device_ptr(p)) return copy_sync(ctx, v.data(), p, 8 * n, hipMemcpyDeviceToHost);
    std::memcpy(v.data(), p, 8 * n);
    return BRPA_OK;
}
}  // namespace

bool xspace_applicable(const double *weights, int64_t nfreqs, double lambda) {
    if (!(lambda > 0.0)) return false;
    std::vector<double> w(nfreqs);
    if (is_device_ptr(weights)) {
        if (hipMemcpy(w.data(), weights, 8 * nfreqs, hipMemcpyDeviceToHost) != hipSuccess)
            return false;
    } else {
        std::memcpy(w.data(), weights, 8 * nfreqs);
    }
    for (double x : w)
        if (!(x > 0.0)) return false;
    return true;
}

int exchange_xspace(brpa_ctx *ctx, const double *R, int64_t naux, const double *eocc, int64_t nocc,
                    const double *evir, int64_t nvir, const double *freqs, const double *weights,
                    int64_t nfreqs, double lambda, double cutoff, int64_t nguess, uint64_t seed,
                    double alpha, double *ex, int64_t *nvecs_t2) {
    if (!ctx) return BRPA_ERR_ARG;
    if (!R || !eocc || !evir || !freqs || !weights || !ex || !nvecs_t2 || naux <= 0 || nocc <= 0 ||
        nvir <= 0 || nfreqs <= 0 || !(lambda > 0.0))
        return brpa_fail(ctx, BRPA_ERR_ARG, "exchange: bad arguments");
    const int64_t K = nocc * nvir, N = naux, F = nfreqs;
    const int64_t NG = nguess > 0 ? nguess : (int64_t)std::ceil(3.0 * (double)N);
    if (K > 46340 || N > 65536 || NG > 46340 || F * N * K > INT32_MAX)
        return brpa_fail(ctx, BRPA_ERR_ARG, "exchange: sizes beyond the supported range "
                         "(o*v = %lld, N = %lld, NG = %lld)", (long long)K, (long long)N,
                         (long long)NG);
    std::vector<double> eo, ev, fx, fw;
    BRPA_TRY(host_copy(ctx, eo, eocc, nocc));
    BRPA_TRY(host_copy(ctx, ev, evir, nvir));
    BRPA_TRY(host_copy(ctx, fx, freqs, F));
    BRPA_TRY(host_copy(ctx, fw, weights, F));
    if (!finite_all(eo) || !finite_all(ev) || !finite_all(fx) || !finite_all(fw))
        return brpa_fail(ctx, BRPA_ERR_NONFINITE, "exchange: non-finite input");
    if (!(*std::min_element(ev.begin(), ev.end()) - *std::max_element(eo.begin(), eo.end()) > 0.0))
        return brpa_fail(ctx, BRPA_ERR_GAP, "Found invalid orbital energy gap.");
    for (double w : fw)
        if (!(w > 0.0))
            return brpa_fail(ctx, BRPA_ERR_ARG, "exchange: the explicit-space form needs positive "
                             "quadrature weights");
    BRPA_HIP(ctx, hipSetDevice(ctx->device));

    // chi_k(ai) (w_k l/(4 pi))^{1/2}, chi_k = -4 d/(d^2+u_k^2) (rpa_CC_Chi,
    // rpa/rpa_CC_Doubles.f90:276-292)
    std::vector<double> hchis(F * K);
    for (int64_t k = 0; k < F; ++k) {
        const double sk = std::sqrt(fw[k] * lambda / (4.0 * kPi));
        for (int64_t i = 0; i < nocc; ++i)
            for (int64_t a = 0; a < nvir; ++a) {
                const double d = ev[a] - eo[i];
                hchis[k * K + i * nvir + a] = -(4.0 * d) / (d * d + fx[k] * fx[k]) * sk;
            }
    }

    const bool over = NG >= K;          // the guess space oversamples the ov space
    const int64_t ntmax = std::min(K, NG);
    const int64_t nd = over ? K : NG;   // order of the Gram matrix of the basis step
    const int64_t FN = F * N;
    const bool rdev = is_device_ptr(R);
    const int64_t nsteps = cdiv(N, CHOL_NB);
    FragDev fd{};
    FragDev *dfd = nullptr;
    int64_t *dpre = nullptr, *dchol = nullptr;
    double *dR = nullptr, *chis, *Y, *Im, *Om, *X, *Wb, *Sm, *TIT, *Winv, *Tinv, *sev, *sE, *part,
        *splitws;
    int *dinfo;
    DenseChol dc;
    std::vector<int64_t> pre = {0, 0, 0, 0, 0, 0};
    const int64_t ntile = cdiv(N, SYRK_T);
    pre[1] = cdiv(F * K, 256);                      // dtab
    pre[3] = ntile * (ntile + 1) / 2 * cdiv(F, 4);  // syrk
    pre[5] = cdiv(F * N, 256);                      // set_diag
    const int64_t nparts = sosex_dot_nparts(K);
    const int64_t nsplitws = std::max<int64_t>(
        kSplitWsMin, std::min<int64_t>((int64_t)1 << 28, 8 * std::max(K, NG) * std::max(K, NG)));
    auto layout = [&](Arena &ar) {
        dfd = ar.take<FragDev>(1);
        dpre = ar.take<int64_t>(pre.size());
        dR = rdev ? const_cast<double *>(R) : ar.take<double>(K * N);
        fd.dtab = ar.take<double>(F * K);
        fd.Pi = ar.take<double>(F * N * N);
        fd.diag = ar.take<double>(F * N);
        fd.weights = ar.take<double>(F);
        fd.freqs = ar.take<double>(F);
        fd.eocc = ar.take<double>(nocc);
        fd.evir = ar.take<double>(nvir);
        fd.linv = ar.take<double>(F * CHOL_NB * CHOL_NB);
        fd.ldpart = ar.take<double>(F * nsteps * 2);
        fd.status = ar.take<int>(F);
        dchol = ar.take<int64_t>(chol_pre_len(N));
        chis = ar.take<double>(F * K);
        Y = ar.take<double>(std::max(FN * K, K * K));   // stacked Y, then the T2 matrix
        Im = ar.take<double>(K * K);
        Om = ar.take<double>(K * NG);
        X = ar.take<double>(K * NG);
        Wb = ar.take<double>(K * ntmax);
        Sm = ar.take<double>(nd * nd);
        TIT = ar.take<double>(ntmax * ntmax);
        Winv = ar.take<double>(std::max(F * N * N, nd * nd));
        Tinv = ar.take<double>(std::max(F * trinv_scratch(N), trinv_scratch(nd)));
        sev = ar.take<double>(std::max(K, NG));
        sE = ar.take<double>(std::max(K, NG));
        dinfo = ar.take<int>(2);
        part = ar.take<double>(nparts);
        splitws = ar.take<double>(nsplitws);
        dc.layout(ar, nd);
    };
    Arena count;
    layout(count);
    Arena ar;
    BRPA_TRY(ws_reserve(ctx, count.off, &ar));
    layout(ar);
    // the split-k GEMMs of small shapes keep their partial tiles in splitws
    // for the duration of this call
    struct SplitLend {
        brpa_ctx *c;
        ~SplitLend() {
            c->split_ws = nullptr;
            c->split_ws_cap = 0;
        }
    } lend{ctx};
    ctx->split_ws = splitws;
    ctx->split_ws_cap = (size_t)nsplitws;
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
    // host inputs go up with completed copies
    BRPA_TRY(copy_sync(ctx, const_cast<double *>(fd.weights), fw.data(), 8 * F, hipMemcpyHostToDevice));
    BRPA_TRY(copy_sync(ctx, const_cast<double *>(fd.freqs), fx.data(), 8 * F, hipMemcpyHostToDevice));
    BRPA_TRY(copy_sync(ctx, const_cast<double *>(fd.eocc), eo.data(), 8 * nocc, hipMemcpyHostToDevice));
    BRPA_TRY(copy_sync(ctx, const_cast<double *>(fd.evir), ev.data(), 8 * nvir, hipMemcpyHostToDevice));
    BRPA_TRY(copy_sync(ctx, chis, hchis.data(), 8 * F * K, hipMemcpyHostToDevice));
    BRPA_TRY(copy_sync(ctx, dfd, &fd, sizeof(FragDev), hipMemcpyHostToDevice));
    BRPA_TRY(copy_sync(ctx, dpre, pre.data(), 8 * pre.size(), hipMemcpyHostToDevice));
    BRPA_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    if (!ctx->stage_keep) stage_reset(ctx);
    kstat_drop(ctx);
    bool staged[BRPA_NSTAGES] = {};
    auto begin = [&](int st) -> int {
        staged[st] = true;
        return stage_begin(ctx, st);
    };
    auto collect = [&]() -> int {
        BRPA_HIP(ctx, hipStreamSynchronize(ctx->stream));
        for (int st = 0; st < BRPA_NSTAGES; ++st)
            if (staged[st]) BRPA_TRY(stage_collect(ctx, st));
        return kstat_collect(ctx);
    };

    // 1. Pi_k (FP64 MFMA SYRK), 1 + l Pi_k, Cholesky factors of all F in one batch
    BRPA_TRY(begin(BRPA_STAGE_T2_PI));
    BRPA_TRY(launch_dtab(ctx, dfd, dpre, 1, pre[1]));
    BRPA_TRY(launch_syrk_f64(ctx, dfd, dpre + 2, 1, pre[3]));
    if (lambda != 1.0) BRPA_TRY(launch_scale(ctx, fd.Pi, lambda, F * N * N));
    BRPA_TRY(launch_set_diag(ctx, dfd, dpre + 4, 1, pre[5], 0));
    BRPA_TRY(chol_single(ctx, dfd, dchol, N, F));
    // 2. W_k = U_k^{-1} (the factor is the column-major upper triangle of Pi_k),
    // then Y_k = W_k^T R^T diag(chis_k) stacked: column-major (F N) x K, block k
    // at rows k N.. (R^T is R[ai][P] read as a column-major N x K matrix)
    BRPA_TRY(dense_trinv(ctx, N, fd.Pi, N, N * N, Winv, N, N * N, F, Tinv));
    BRPA_TRY(launch_gemm_f64_batched(ctx, true, false, N, K, N, 1.0, Winv, N, N * N, 0, dR, N, 0,
                                     0, 0.0, Y, FN, N, 0, F, 1, false, TRI_A_LOWER, chis, K, 0));
    BRPA_TRY(stage_end(ctx, BRPA_STAGE_T2_PI));

    // 2-3. I = -Y^T Y, then I Omega with normalised columns (rpa_CC_Doubles.f90:572-588)
    BRPA_TRY(begin(BRPA_STAGE_T2_IOMEGA));
    BRPA_TRY(dense_syrk(ctx, true, K, FN, -1.0, Y, FN, 0.0, Im, K));
    BRPA_TRY(dense_symm_fill(ctx, Im, K));
    BRPA_TRY(launch_guess(ctx, Om, K, NG, seed));
    BRPA_TRY(launch_normalize_columns(ctx, Om, K, NG));
    BRPA_TRY(dense_gemm(ctx, false, false, K, NG, K, 1.0, Im, K, Om, K, 0.0, X, K));
    BRPA_TRY(launch_normalize_columns(ctx, X, K, NG));
    BRPA_TRY(stage_end(ctx, BRPA_STAGE_T2_IOMEGA));

    // 4. the kept subspace (rpa_CC_Doubles.f90:595-614).  With X = I Omega,
    // S = X^T X (NG x NG) and X X^T (K x K) share their nonzero eigenvalues.
    // When every eigenvalue of the smaller one exceeds the cutoff -- exactly
    // when it minus cutoff * 1 has a Cholesky factor -- the kept subspace is
    // all of span(X) and any orthonormal basis of it gives the same energy.
    BRPA_TRY(begin(BRPA_STAGE_T2_BASIS));
    int64_t nt = 0;
    bool ident = false;            // T = 1 (the whole ov space is kept)
    const double *T = X;           // K x nt orthonormal basis, leading dimension K
    auto count_kept = [&](int64_t n) -> int {
        std::vector<double> hs(n);
        BRPA_TRY(copy_sync(ctx, hs.data(), sev, 8 * n, hipMemcpyDeviceToHost));
        // counted from the largest down, stopping at the first one below (600-607)
        nt = 0;
        for (int64_t i = n - 1; i >= 0; --i) {
            if (hs[i] > cutoff)
                ++nt;
            else
                break;
        }
        return BRPA_OK;
    };
    if (over) {
        BRPA_TRY(dense_syrk(ctx, false, K, NG, 1.0, X, K, 0.0, Sm, K));
        BRPA_TRY(launch_add_diag(ctx, Sm, K, -cutoff));
        bool full = false;
        BRPA_TRY(dc.run(ctx, Sm, K, &full));
        if (full) {
            nt = K;
            ident = true;
        } else {
            BRPA_TRY(begin(BRPA_STAGE_T2_EIG));
            BRPA_TRY(dense_syrk(ctx, false, K, NG, 1.0, X, K, 0.0, Sm, K));
            BRPA_TRY(dense_syevd(ctx, K, Sm, sev, sE, dinfo));
            BRPA_TRY(stage_end(ctx, BRPA_STAGE_T2_EIG));
            BRPA_TRY(count_kept(K));
            T = Sm + K * (K - nt);   // eigenvectors of the nt largest eigenvalues
        }
    } else {
        BRPA_TRY(dense_syrk(ctx, true, NG, K, 1.0, X, K, 0.0, Sm, NG));
        BRPA_TRY(launch_add_diag(ctx, Sm, NG, -cutoff));
        bool cholqr = false;
        BRPA_TRY(dc.run(ctx, Sm, NG, &cholqr));
        if (cholqr) {
            // Cholesky-QR in two passes: X <- X U^{-1} with the shifted factor, then
            // once more with the unshifted Gram matrix of the result (each product
            // goes to Wb, and X and Wb trade places)
            nt = NG;
            auto right_inv = [&]() -> int {
                BRPA_TRY(dense_trinv(ctx, NG, Sm, NG, 0, Winv, NG, 0, 1, Tinv));
                BRPA_TRY(launch_gemm_f64_batched(ctx, false, false, K, NG, NG, 1.0, X, K, 0, 0,
                                                 Winv, NG, 0, 0, 0.0, Wb, K, 0, 0, 1, 1, false,
                                                 TRI_B_UPPER, nullptr, 0, 0));
                std::swap(X, Wb);
                return BRPA_OK;
            };
            BRPA_TRY(right_inv());
            BRPA_TRY(dense_syrk(ctx, true, NG, K, 1.0, X, K, 0.0, Sm, NG));
            bool ok2 = false;
            BRPA_TRY(dc.run(ctx, Sm, NG, &ok2));
            if (!ok2)
                return brpa_fail(ctx, BRPA_ERR_SOLVER, "Cholesky-QR pass 2 of the T2 basis broke down");
            BRPA_TRY(right_inv());
            T = X;
        } else {
            BRPA_TRY(begin(BRPA_STAGE_T2_EIG));
            BRPA_TRY(dense_syrk(ctx, true, NG, K, 1.0, X, K, 0.0, Sm, NG));
            BRPA_TRY(dense_syevd(ctx, NG, Sm, sev, sE, dinfo));
            BRPA_TRY(stage_end(ctx, BRPA_STAGE_T2_EIG));
            BRPA_TRY(count_kept(NG));
            if (nt > 0) {
                // T = X U_s s^{-1/2} (rpa_CC_Doubles.f90:610-624)
                BRPA_TRY(launch_select_scale(ctx, Om, Sm, sev, NG, NG - nt, nt));
                BRPA_TRY(dense_gemm(ctx, false, false, K, nt, NG, 1.0, X, K, Om, NG, 0.0, Wb, K));
                std::swap(X, Wb);
                T = X;
            }
        }
    }
    BRPA_TRY(stage_end(ctx, BRPA_STAGE_T2_BASIS));
    *nvecs_t2 = nt;
    if (nt == 0) {
        *ex = 0.0;
        return collect();
    }

    // 5. the amplitude matrix T2 = T f(T^T I T) T^T, f(a) = a / (2 (1 + a))
    // (rpa_CC_Doubles.f90:625-641), in the buffer of Y
    double gamma = 1.0, delta = 0.0;
    if (ident) {
        // T = 1: T2 = I (1 + I)^{-1} / 2 = (1 - (1 + I)^{-1}) / 2; the contraction
        // takes (1 + I)^{-1} with gamma = -1/2 and the identity with delta = 1/2
        BRPA_TRY(begin(BRPA_STAGE_T2_CORE));
        BRPA_TRY(launch_add_diag(ctx, Im, K, 1.0));
        bool ok = false;
        BRPA_TRY(dc.run(ctx, Im, K, &ok));
        if (!ok)
            return brpa_fail(ctx, BRPA_ERR_NOT_SPD, "1 + I is not positive definite");
        // (1 + I)^{-1} = U^{-1} U^{-T} = W W^T with W = U^{-1} upper triangular:
        // C(i, j) = sum_{k >= max(i, j)} W(i, k) W(j, k), the upper tiles
        // starting at k = their column
        BRPA_TRY(dense_trinv(ctx, K, Im, K, 0, Wb, K, 0, 1, Tinv));
        BRPA_TRY(launch_gemm_f64_batched(ctx, false, true, K, K, K, 1.0, Wb, K, 0, 0, Wb, K, 0, 0,
                                         0.0, Y, K, 0, 0, 1, 1, true, TRI_B_LOWER, nullptr, 0, 0));
        BRPA_TRY(dense_symm_fill(ctx, Y, K));
        gamma = -0.5;
        delta = 0.5;
        BRPA_TRY(stage_end(ctx, BRPA_STAGE_T2_CORE));
    } else {
        // T^T I T (rpa_CC_MIM_v5, 425-488) from the explicit I
        BRPA_TRY(begin(BRPA_STAGE_T2_TIT));
        BRPA_TRY(dense_gemm(ctx, false, false, K, nt, K, 1.0, Im, K, T, K, 0.0, Om, K));
        BRPA_TRY(dense_gemm(ctx, true, false, nt, nt, K, 1.0, T, K, Om, K, 0.0, TIT, nt));
        BRPA_TRY(stage_end(ctx, BRPA_STAGE_T2_TIT));
        BRPA_TRY(begin(BRPA_STAGE_T2_CORE));
        BRPA_HIP(ctx, hipMemcpyAsync(Wb, TIT, 8 * nt * nt, hipMemcpyDeviceToDevice, ctx->stream));
        BRPA_TRY(launch_add_diag(ctx, TIT, nt, 1.0));
        bool ok = false;
        BRPA_TRY(dc.run(ctx, TIT, nt, &ok));
        if (ok) {
            // W = T U^{-1}: W W^T = T (1 + TIT)^{-1} T^T, T2 = (T T^T - W W^T) / 2
            BRPA_TRY(dense_trinv(ctx, nt, TIT, nt, 0, Winv, nt, 0, 1, Tinv));
            BRPA_TRY(launch_gemm_f64_batched(ctx, false, false, K, nt, nt, 1.0, T, K, 0, 0, Winv,
                                             nt, 0, 0, 0.0, Wb, K, 0, 0, 1, 1, false, TRI_B_UPPER,
                                             nullptr, 0, 0));
            BRPA_TRY(dense_syrk(ctx, false, K, nt, 0.5, T, K, 0.0, Y, K));
            BRPA_TRY(dense_syrk(ctx, false, K, nt, -0.5, Wb, K, 1.0, Y, K));
            BRPA_TRY(dense_symm_fill(ctx, Y, K));
        } else {
            // eigenpairs of T^T I T: A = a / (2 (1 + a)), V = T E, T2 = V diag(A) V^T
            BRPA_HIP(ctx, hipMemcpyAsync(TIT, Wb, 8 * nt * nt, hipMemcpyDeviceToDevice, ctx->stream));
            BRPA_TRY(dense_syevd(ctx, nt, TIT, sev, sE, dinfo + 1));
            std::vector<double> ha(nt);
            BRPA_TRY(copy_sync(ctx, ha.data(), sev, 8 * nt, hipMemcpyDeviceToHost));
            for (int64_t i = 0; i < nt; ++i) ha[i] = 0.5 * (ha[i] / (ha[i] + 1.0));
            BRPA_TRY(copy_sync(ctx, sev, ha.data(), 8 * nt, hipMemcpyHostToDevice));
            BRPA_TRY(dense_gemm(ctx, false, false, K, nt, nt, 1.0, T, K, TIT, nt, 0.0, Om, K));
            BRPA_TRY(launch_col_scale(ctx, Wb, Om, sev, K, nt));
            BRPA_TRY(dense_gemm(ctx, false, true, K, K, nt, 1.0, Wb, K, Om, K, 0.0, Y, K));
        }
        BRPA_TRY(stage_end(ctx, BRPA_STAGE_T2_CORE));
    }

    // 6. E_x = -alpha sum (ai|bj) T2(aj, bi)
    BRPA_TRY(begin(BRPA_STAGE_SOSEX));
    BRPA_TRY(launch_sosex_dot(ctx, dR, K, N, nvir, Y, gamma, delta, part));
    BRPA_TRY(stage_end(ctx, BRPA_STAGE_SOSEX));
    BRPA_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    std::vector<double> hp(nparts);
    BRPA_TRY(copy_sync(ctx, hp.data(), part, 8 * nparts, hipMemcpyDeviceToHost));
    std::vector<int> info(F);
    BRPA_TRY(copy_sync(ctx, info.data(), fd.status, sizeof(int) * F, hipMemcpyDeviceToHost));
    BRPA_TRY(collect());
    float ms = 0.f;
    BRPA_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    ctx->last_kernel_ms = ms;
    for (int64_t k = 0; k < F; ++k)
        if (info[k] != 0)
            return brpa_fail(ctx, BRPA_ERR_NOT_SPD,
                             "frequency %lld: 1 + l Pi(u) is not positive definite", (long long)k);
    long double s = 0.0L;
    for (int64_t p = 0; p < nparts; ++p) s += (long double)hp[p];
    *ex = (double)(-(long double)alpha * s);
    return BRPA_OK;
}
