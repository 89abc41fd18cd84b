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
//      cutoff (Cholesky-QR when every Gram eigenvalue clears it; when a few do
//      not, the complement of their invariant subspace, found by inverse
//      subspace iteration; the eigenvectors otherwise); when NG >= ov and every
//      eigenvalue of X X^T clears it, T = 1;
//   5. TIT = T^T (I T), 1 + TIT = U^T U, W = T U^{-1},
//      T2 = T (1/2) TIT (1 + TIT)^{-1} T^T = (T T^T - W W^T) / 2;
// Every triangular solve is a triangular inverse (GEMM levels) followed by a
// GEMM that skips the zero triangle.
//   6. E_x = -alpha sum_{aibj} (ai|bj) T2(aj,bi): (ai|bj) = (R R^T)(ai,bj) by
//      one SYRK, then the index-permuted contraction (k_sosex_contract).
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
int launch_ritz_resid(brpa_ctx *, const double *AZ, const double *Z, const double *theta,
                      int64_t n, int64_t m, double *out);
int launch_sosex_contract(brpa_ctx *, const double *G, int64_t K, int64_t v, const double *M,
                          double gamma, double delta, double *part);

namespace {
constexpr double kPi = 3.14159265358979323846;
// partial tiles of split-k GEMMs (up to 8 splits of the largest C of a call)
constexpr int64_t kSplitWsMin = (int64_t)9 << 20;
// dropped-subspace basis: first block size, largest block, Ritz pairs kept
// above the dropped ones, subspace steps
constexpr int64_t kDropBlock0 = 192;
constexpr int64_t kDropBlockMax = 512;
constexpr int64_t kDropMargin = 64;
constexpr int kDropSteps = 10;
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

bool xspace_fits(int64_t K, int64_t N, int64_t NG, int64_t F) {
    // ov x ov and NG x NG matrices with 32-bit element offsets within one
    // matrix, and the F N x ov stack Y
    return K <= 46340 && N <= 65536 && NG <= 46340 && F <= INT32_MAX && F * N * K <= INT32_MAX;
}

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
    if (!xspace_fits(K, N, NG, F))
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
    const int64_t nparts = K;   // one partial sum per column of the contraction
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
    // I is symmetric (both triangles stored): I Omega as I^T Omega, whose
    // operands are both contiguous in k
    BRPA_TRY(dense_gemm(ctx, true, false, K, NG, K, 1.0, Im, K, Om, K, 0.0, X, K));
    BRPA_TRY(launch_normalize_columns(ctx, X, K, NG));
    BRPA_TRY(stage_end(ctx, BRPA_STAGE_T2_IOMEGA));

    // 4. the kept subspace (rpa_CC_Doubles.f90:595-614).  With X = I Omega,
    // S = X^T X (NG x NG) and X X^T (K x K) share their nonzero eigenvalues.
    // When every eigenvalue of the smaller one exceeds the cutoff -- exactly
    // when it minus cutoff * 1 has a Cholesky factor -- the kept subspace is
    // all of span(X) and any orthonormal basis of it gives the same energy.
    BRPA_TRY(begin(BRPA_STAGE_T2_BASIS));
    int64_t nt = 0;                // kept dimension (the reference's NT2)
    int64_t nc = 0;                // columns of T (nt, or NG for the padded basis below)
    bool ident = false;            // T2 = (1 - (1 + I)^{-1}) / 2 with the (projected) I
    const double *T = X;           // K x nc basis, leading dimension K
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
    auto gram = [&]() -> int {
        if (over) return dense_syrk(ctx, false, K, NG, 1.0, X, K, 0.0, Sm, K);
        return dense_syrk(ctx, true, NG, K, 1.0, X, K, 0.0, Sm, NG);
    };
    // C (m x n) = op(A) op(B) with a triangular operand (TRI_*), no scale
    auto tri_gemm = [&](bool ta, bool tb, int64_t m, int64_t n, int64_t k, double alpha,
                        const double *A, int64_t lda, const double *B, int64_t ldb, double *C,
                        int64_t ldc, int tri) -> int {
        return launch_gemm_f64_batched(ctx, ta, tb, m, n, k, alpha, A, lda, 0, 0, B, ldb, 0, 0,
                                       0.0, C, ldc, 0, 0, 1, 1, false, tri, nullptr, 0, 0);
    };

    // Dropped-subspace basis.  G (order nd, in Sm) failed the Cholesky test of
    // G - cutoff: some of its eigenvalues lie at or below the cutoff, and on
    // these fragments they are few (1 to ~100 of thousands).  Their invariant
    // subspace Zd (nd x md) is found by inverse subspace iteration with
    // (G + cutoff/4)^{-1} on a block of p >= md + 64 vectors (Cholesky-QR in two
    // passes after each step) and Rayleigh-Ritz; the kept subspace of the
    // reference (the eigenvectors above the cutoff) is the orthogonal
    // complement of Zd, and T2 depends on the kept subspace only:
    //  - oversampled (G = X X^T, order K): T T^T = P = 1 - Zd Zd^T and
    //    T f(T^T I T) T^T = (1 - (1 + P I P)^{-1}) / 2, the whole-space form with
    //    I replaced by P I P = I - Zd E^T - E Zd^T, E = I Zd - Zd (Zd^T I Zd) / 2;
    //  - otherwise (G = X^T X, order NG): the kept subspace is span(X P).  The
    //    (K + md) x NG matrix [X P; Zd^T] has full column rank; its span is
    //    span(X P) plus the last md coordinates, on which the zero-padded I
    //    vanishes, so Cholesky-QR of it and the usual core leave T2 in the top
    //    K x K block: T is its top K rows (nc = NG columns for nt = NG - md kept
    //    directions), the bottom md rows enter the two Gram matrices only.
    // Exactly md eigenvalues of G lie at or below the cutoff when the md Ritz
    // values of Zd do (Courant-Fischer: at least md) and the compression of
    // G - cutoff to the complement of Zd has a Cholesky factor (interlacing: at
    // most md): that factor is the first Cholesky-QR factor, or P G P + Zd Zd^T
    // - cutoff in the oversampled case.  A failure anywhere (a factor that
    // breaks down, no convergence, md too close to p) leaves *done false and
    // the eigensolve runs instead.
    auto drop_basis = [&](bool *done) -> int {
        *done = false;
        const int64_t n = nd, r = std::max(K, nd);
        const int64_t ysz = std::max(FN * K, K * K);
        // the trace of G is NG (normalised columns); its largest eigenvalues are
        // ~10 NG / n on this workload, the residual floor ~3e-15 NG / n
        const double scale = (double)NG / (double)n;
        const double tol_conv = 3e-14 * scale, tol_stall = 1e-13 * scale;
        // G in full and (G + cutoff / 4)^{-1} = W W^T, W = U^{-1} in Winv
        BRPA_TRY(gram());
        BRPA_TRY(dense_symm_fill(ctx, Sm, n));
        BRPA_HIP(ctx, hipMemcpyAsync(TIT, Sm, 8 * n * n, hipMemcpyDeviceToDevice, ctx->stream));
        BRPA_TRY(launch_add_diag(ctx, TIT, n, 0.25 * cutoff));
        bool ok = false;
        BRPA_TRY(dc.run(ctx, TIT, n, &ok));
        if (!ok) return BRPA_OK;
        BRPA_TRY(dense_trinv(ctx, n, TIT, n, 0, Winv, n, 0, 1, Tinv));
        int64_t p = kDropBlock0, md = 0;
        double *Zd = nullptr, *GZd = nullptr, *A2 = nullptr, *B2 = nullptr;
        bool conv = false;
        for (int attempt = 0; attempt < 2 && !conv; ++attempt) {
            if (p > kDropBlockMax || 2 * p > n || 4 * n * p + 4 * r * p > ysz ||
                2 * p * p > K * NG)
                return BRPA_OK;
            double *Z = Y, *Z2 = Y + n * p, *GZ = Y + 2 * n * p;
            GZd = Y + 3 * n * p;
            A2 = Y + 4 * n * p;
            B2 = A2 + 2 * r * p;
            double *Gp = Om, *Wp = Om + p * p;
            BRPA_TRY(launch_guess(ctx, Z, n, p, seed ^ 0x6a09e667f3bcc909ull));
            double res_prev = INFINITY;
            int64_t grow = 0;
            for (int it = 1; it <= kDropSteps && !conv; ++it) {
                // Z <- (G + cutoff / 4)^{-1} Z = W (W^T Z)
                BRPA_TRY(tri_gemm(true, false, n, p, n, 1.0, Winv, n, Z, n, Z2, n, TRI_A_LOWER));
                BRPA_TRY(tri_gemm(false, false, n, p, n, 1.0, Winv, n, Z2, n, Z, n, TRI_A_UPPER));
                for (int pass = 0; pass < 2; ++pass) {
                    BRPA_TRY(dense_syrk(ctx, true, p, n, 1.0, Z, n, 0.0, Gp, p));
                    bool okq = false;
                    BRPA_TRY(dc.run(ctx, Gp, p, &okq));
                    if (!okq) return BRPA_OK;
                    BRPA_TRY(dense_trinv(ctx, p, Gp, p, 0, Wp, p, 0, 1, Tinv));
                    BRPA_TRY(tri_gemm(false, false, n, p, p, 1.0, Z, n, Wp, p, Z2, n, TRI_B_UPPER));
                    std::swap(Z, Z2);
                }
                // Rayleigh-Ritz at steps 4, 6, 8, 10 (the p x p eigensolve costs more
                // than two subspace steps)
                if (it < 4 || (it & 1)) continue;
                // Rayleigh-Ritz: Z^T G Z = E diag(theta) E^T, theta ascending (G Z
                // as G^T Z: G is symmetric with both triangles stored)
                BRPA_TRY(dense_gemm(ctx, true, false, n, p, n, 1.0, Sm, n, Z, n, 0.0, GZ, n));
                BRPA_TRY(dense_gemm(ctx, true, false, p, p, n, 1.0, Z, n, GZ, n, 0.0, Gp, p));
                BRPA_TRY(dense_syevd(ctx, p, Gp, sev, sE, dinfo));
                std::vector<double> th(p);
                BRPA_TRY(copy_sync(ctx, th.data(), sev, 8 * p, hipMemcpyDeviceToHost));
                if (!finite_all(th)) return BRPA_OK;
                int64_t m = 0;
                while (m < p && th[m] <= cutoff) ++m;
                if (m == 0) continue;
                if (m > p - kDropMargin) {
                    grow = m;
                    break;
                }
                // the dropped Ritz pairs: Zd = Z E_d, G Zd = (G Z) E_d
                BRPA_TRY(dense_gemm(ctx, false, false, n, m, p, 1.0, Z, n, Gp, p, 0.0, Z2, n));
                BRPA_TRY(dense_gemm(ctx, false, false, n, m, p, 1.0, GZ, n, Gp, p, 0.0, GZd, n));
                BRPA_TRY(launch_ritz_resid(ctx, GZd, Z2, sev, n, m, sE));
                std::vector<double> rn(m);
                BRPA_TRY(copy_sync(ctx, rn.data(), sE, 8 * m, hipMemcpyDeviceToHost));
                if (!finite_all(rn)) return BRPA_OK;
                const double res = *std::max_element(rn.begin(), rn.end());
                conv = res <= tol_conv || (res <= tol_stall && res > 0.5 * res_prev);
                res_prev = res;
                md = m;
                Zd = Z2;
            }
            if (!conv) {
                if (grow == 0) return BRPA_OK;
                p = cdiv(grow + 2 * kDropMargin, 64) * 64;
            }
        }
        if (!conv) return BRPA_OK;
        // M -= Zd D^T + D Zd^T as one GEMM of depth 2 md: [Zd D] [D Zd]^T
        auto sym_update = [&](double *M, const double *D) -> int {
            BRPA_HIP(ctx, hipMemcpyAsync(A2, Zd, 8 * n * md, hipMemcpyDeviceToDevice, ctx->stream));
            BRPA_HIP(ctx, hipMemcpyAsync(A2 + n * md, D, 8 * n * md, hipMemcpyDeviceToDevice,
                                         ctx->stream));
            BRPA_HIP(ctx, hipMemcpyAsync(B2, D, 8 * n * md, hipMemcpyDeviceToDevice, ctx->stream));
            BRPA_HIP(ctx, hipMemcpyAsync(B2 + n * md, Zd, 8 * n * md, hipMemcpyDeviceToDevice,
                                         ctx->stream));
            return dense_gemm(ctx, false, true, n, n, 2 * md, -1.0, A2, n, B2, n, 1.0, M, n);
        };
        double *Cd = Om;   // md x md
        if (over) {
            // P G P + Zd Zd^T - cutoff = G - Zd D^T - D Zd^T - cutoff,
            // D = G Zd - Zd (Zd^T G Zd + 1) / 2
            BRPA_TRY(dense_gemm(ctx, true, false, md, md, n, 1.0, Zd, n, GZd, n, 0.0, Cd, md));
            BRPA_TRY(launch_add_diag(ctx, Cd, md, 1.0));
            BRPA_TRY(dense_gemm(ctx, false, false, n, md, md, -0.5, Zd, n, Cd, md, 1.0, GZd, n));
            BRPA_TRY(sym_update(Sm, GZd));
            BRPA_TRY(launch_add_diag(ctx, Sm, n, -cutoff));
            bool okv = false;
            BRPA_TRY(dc.run(ctx, Sm, n, &okv));
            if (!okv) return BRPA_OK;
            // I <- P I P = I - Zd E^T - E Zd^T, E = I Zd - Zd (Zd^T I Zd) / 2
            double *E = GZd;
            BRPA_TRY(dense_gemm(ctx, true, false, K, md, K, 1.0, Im, K, Zd, K, 0.0, E, K));
            BRPA_TRY(dense_gemm(ctx, true, false, md, md, K, 1.0, Zd, K, E, K, 0.0, Cd, md));
            BRPA_TRY(dense_gemm(ctx, false, false, K, md, md, -0.5, Zd, K, Cd, md, 1.0, E, K));
            BRPA_TRY(sym_update(Im, E));
            nt = K - md;
            nc = K;
            ident = true;
        } else {
            // [X P; Zd^T]: top block X P = X - (X Zd) Zd^T in Wb (X stays for the
            // eigensolve should the test below fail), bottom block Zd^T
            double *XZ = A2, *B1 = B2;   // K x md, md x NG
            BRPA_HIP(ctx, hipMemcpyAsync(Wb, X, 8 * K * NG, hipMemcpyDeviceToDevice, ctx->stream));
            BRPA_TRY(dense_gemm(ctx, false, false, K, md, NG, 1.0, X, K, Zd, NG, 0.0, XZ, K));
            BRPA_TRY(dense_gemm(ctx, false, true, K, NG, md, -1.0, XZ, K, Zd, NG, 1.0, Wb, K));
            // Gram matrix P G P + Zd Zd^T of the padded matrix, minus the cutoff
            BRPA_TRY(dense_syrk(ctx, true, NG, K, 1.0, Wb, K, 0.0, Sm, NG));
            BRPA_TRY(dense_syrk(ctx, false, NG, md, 1.0, Zd, NG, 1.0, Sm, NG));
            BRPA_TRY(launch_add_diag(ctx, Sm, NG, -cutoff));
            bool okv = false;
            BRPA_TRY(dc.run(ctx, Sm, NG, &okv));
            if (!okv) return BRPA_OK;
            // Cholesky-QR pass 1 with the shifted factor: top X <- (X P) U^{-1},
            // bottom B1 = Zd^T U^{-1}
            BRPA_TRY(dense_trinv(ctx, NG, Sm, NG, 0, Winv, NG, 0, 1, Tinv));
            BRPA_TRY(tri_gemm(false, false, K, NG, NG, 1.0, Wb, K, Winv, NG, X, K, TRI_B_UPPER));
            BRPA_TRY(tri_gemm(true, false, md, NG, NG, 1.0, Zd, NG, Winv, NG, B1, md, TRI_B_UPPER));
            // pass 2 with the Gram matrix of both blocks; only the top block goes on
            BRPA_TRY(dense_syrk(ctx, true, NG, K, 1.0, X, K, 0.0, Sm, NG));
            BRPA_TRY(dense_syrk(ctx, true, NG, md, 1.0, B1, md, 1.0, Sm, NG));
            bool ok2 = false;
            BRPA_TRY(dc.run(ctx, Sm, NG, &ok2));
            if (!ok2)
                return brpa_fail(ctx, BRPA_ERR_SOLVER, "Cholesky-QR pass 2 of the T2 basis broke down");
            BRPA_TRY(dense_trinv(ctx, NG, Sm, NG, 0, Winv, NG, 0, 1, Tinv));
            BRPA_TRY(tri_gemm(false, false, K, NG, NG, 1.0, X, K, Winv, NG, Wb, K, TRI_B_UPPER));
            std::swap(X, Wb);
            T = X;
            nt = NG - md;
            nc = NG;
        }
        *done = true;
        return BRPA_OK;
    };

    BRPA_TRY(gram());
    BRPA_TRY(launch_add_diag(ctx, Sm, nd, -cutoff));
    bool pass = false;
    BRPA_TRY(dc.run(ctx, Sm, nd, &pass));
    bool dropped = false;
    if (!pass) {
        BRPA_TRY(begin(BRPA_STAGE_T2_DROP));
        BRPA_TRY(drop_basis(&dropped));
        BRPA_TRY(stage_end(ctx, BRPA_STAGE_T2_DROP));
    }
    if (over) {
        if (pass) {
            nt = nc = K;
            ident = true;
        } else if (!dropped) {
            BRPA_TRY(begin(BRPA_STAGE_T2_EIG));
            BRPA_TRY(gram());
            BRPA_TRY(dense_syevd(ctx, K, Sm, sev, sE, dinfo));
            BRPA_TRY(stage_end(ctx, BRPA_STAGE_T2_EIG));
            BRPA_TRY(count_kept(K));
            nc = nt;
            T = Sm + K * (K - nt);   // eigenvectors of the nt largest eigenvalues
        }
    } else {
        if (pass) {
            // Cholesky-QR in two passes: X <- X U^{-1} with the shifted factor, then
            // once more with the unshifted Gram matrix of the result (each product
            // goes to Wb, and X and Wb trade places)
            nt = nc = NG;
            auto right_inv = [&]() -> int {
                BRPA_TRY(dense_trinv(ctx, NG, Sm, NG, 0, Winv, NG, 0, 1, Tinv));
                BRPA_TRY(tri_gemm(false, false, K, NG, NG, 1.0, X, K, Winv, NG, Wb, K,
                                  TRI_B_UPPER));
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
        } else if (!dropped) {
            BRPA_TRY(begin(BRPA_STAGE_T2_EIG));
            BRPA_TRY(gram());
            BRPA_TRY(dense_syevd(ctx, NG, Sm, sev, sE, dinfo));
            BRPA_TRY(stage_end(ctx, BRPA_STAGE_T2_EIG));
            BRPA_TRY(count_kept(NG));
            nc = nt;
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
        // T = 1: T2 = I (1 + I)^{-1} / 2 = (1 - (1 + I)^{-1}) / 2 (with I = P I P
        // when a dropped subspace was projected out); the contraction takes
        // (1 + I)^{-1} with gamma = -1/2 and the identity with delta = 1/2
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
        BRPA_TRY(dense_gemm(ctx, true, false, K, nc, K, 1.0, Im, K, T, K, 0.0, Om, K));
        BRPA_TRY(dense_gemm(ctx, true, false, nc, nc, K, 1.0, T, K, Om, K, 0.0, TIT, nc));
        BRPA_TRY(stage_end(ctx, BRPA_STAGE_T2_TIT));
        BRPA_TRY(begin(BRPA_STAGE_T2_CORE));
        BRPA_HIP(ctx, hipMemcpyAsync(Wb, TIT, 8 * nc * nc, hipMemcpyDeviceToDevice, ctx->stream));
        BRPA_TRY(launch_add_diag(ctx, TIT, nc, 1.0));
        bool ok = false;
        BRPA_TRY(dc.run(ctx, TIT, nc, &ok));
        if (ok) {
            // W = T U^{-1}: W W^T = T (1 + TIT)^{-1} T^T, T2 = (T T^T - W W^T) / 2
            BRPA_TRY(dense_trinv(ctx, nc, TIT, nc, 0, Winv, nc, 0, 1, Tinv));
            BRPA_TRY(launch_gemm_f64_batched(ctx, false, false, K, nc, nc, 1.0, T, K, 0, 0, Winv,
                                             nc, 0, 0, 0.0, Wb, K, 0, 0, 1, 1, false, TRI_B_UPPER,
                                             nullptr, 0, 0));
            BRPA_TRY(dense_syrk(ctx, false, K, nc, 0.5, T, K, 0.0, Y, K));
            BRPA_TRY(dense_syrk(ctx, false, K, nc, -0.5, Wb, K, 1.0, Y, K));
            BRPA_TRY(dense_symm_fill(ctx, Y, K));
        } else {
            // eigenpairs of T^T I T: A = a / (2 (1 + a)), V = T E, T2 = V diag(A) V^T
            BRPA_HIP(ctx, hipMemcpyAsync(TIT, Wb, 8 * nc * nc, hipMemcpyDeviceToDevice, ctx->stream));
            BRPA_TRY(dense_syevd(ctx, nc, TIT, sev, sE, dinfo + 1));
            std::vector<double> ha(nc);
            BRPA_TRY(copy_sync(ctx, ha.data(), sev, 8 * nc, hipMemcpyDeviceToHost));
            for (int64_t i = 0; i < nc; ++i) ha[i] = 0.5 * (ha[i] / (ha[i] + 1.0));
            BRPA_TRY(copy_sync(ctx, sev, ha.data(), 8 * nc, hipMemcpyHostToDevice));
            BRPA_TRY(dense_gemm(ctx, false, false, K, nc, nc, 1.0, T, K, TIT, nc, 0.0, Om, K));
            BRPA_TRY(launch_col_scale(ctx, Wb, Om, sev, K, nc));
            BRPA_TRY(dense_gemm(ctx, false, true, K, K, nc, 1.0, Wb, K, Om, K, 0.0, Y, K));
        }
        BRPA_TRY(stage_end(ctx, BRPA_STAGE_T2_CORE));
    }

    // 6. E_x = -alpha sum (ai|bj) T2(aj, bi)
    BRPA_TRY(begin(BRPA_STAGE_SOSEX));
    // (ai|bj) = (R R^T)(ai, bj) as a SYRK into the buffer of I (no longer needed;
    // R[ai][P] is the column-major N x K matrix R^T, contiguous in P), then the
    // contraction with the T2 matrix
    BRPA_TRY(dense_syrk(ctx, true, K, N, 1.0, dR, N, 0.0, Im, K));
    BRPA_TRY(launch_sosex_contract(ctx, Im, K, nvir, Y, gamma, delta, part));
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
