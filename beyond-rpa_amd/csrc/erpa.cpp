// Host orchestration of the E_RPA hot path: Pi(u) build, exact diagonal,
// batched blocked Cholesky log-det and the quadrature, for one fragment or a
// batch of uneven fragments in the same launches.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "internal.h"

namespace {

struct HostFrag {
    int64_t nocc, nvir, naux, nfreqs;
    const double *R, *eocc, *evir, *freqs, *weights;
};

int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

bool finite_all(const double *p, int64_t n) {
    for (int64_t i = 0; i < n; ++i)
        if (!std::isfinite(p[i])) return false;
    return true;
}

int check_frag(brpa_ctx *ctx, const HostFrag &f, int64_t idx) {
    if (f.nocc <= 0 || f.nvir <= 0 || f.naux <= 0 || f.nfreqs <= 0 || !f.R || !f.eocc ||
        !f.evir || !f.freqs || !f.weights)
        return brpa_fail(ctx, BRPA_ERR_ARG, "fragment %lld: bad sizes or null pointer",
                         (long long)idx);
    if (f.nocc * f.nvir > INT32_MAX || f.naux > 65536 || f.nfreqs > 4096)
        return brpa_fail(ctx, BRPA_ERR_ARG, "fragment %lld: sizes beyond the supported range",
                         (long long)idx);
    // energies and grid are small: validate on the host when they are host arrays
    std::vector<double> eo(f.nocc), ev(f.nvir), fx(f.nfreqs), fw(f.nfreqs);
    const bool dev = is_device_ptr(f.eocc);
    if (dev) {
        BRPA_HIP(ctx, hipMemcpy(eo.data(), f.eocc, 8 * f.nocc, hipMemcpyDeviceToHost));
        BRPA_HIP(ctx, hipMemcpy(ev.data(), f.evir, 8 * f.nvir, hipMemcpyDeviceToHost));
    } else {
        std::memcpy(eo.data(), f.eocc, 8 * f.nocc);
        std::memcpy(ev.data(), f.evir, 8 * f.nvir);
    }
    if (is_device_ptr(f.freqs)) {
        BRPA_HIP(ctx, hipMemcpy(fx.data(), f.freqs, 8 * f.nfreqs, hipMemcpyDeviceToHost));
        BRPA_HIP(ctx, hipMemcpy(fw.data(), f.weights, 8 * f.nfreqs, hipMemcpyDeviceToHost));
    } else {
        std::memcpy(fx.data(), f.freqs, 8 * f.nfreqs);
        std::memcpy(fw.data(), f.weights, 8 * f.nfreqs);
    }
    if (!finite_all(eo.data(), f.nocc) || !finite_all(ev.data(), f.nvir) ||
        !finite_all(fx.data(), f.nfreqs) || !finite_all(fw.data(), f.nfreqs))
        return brpa_fail(ctx, BRPA_ERR_NONFINITE, "fragment %lld: non-finite energy or grid",
                         (long long)idx);
    double emin = *std::min_element(ev.begin(), ev.end());
    double omax = *std::max_element(eo.begin(), eo.end());
    if (!(emin - omax > 0.0))
        return brpa_fail(ctx, BRPA_ERR_GAP,
                         "fragment %lld: Found invalid orbital energy gap. Cannot decompose "
                         "energy denominators.",
                         (long long)idx);
    return BRPA_OK;
}

// Copies n doubles of a host-or-device array into host memory.
int host_copy(brpa_ctx *ctx, double *dst, const double *src, int64_t n) {
    if (is_device_ptr(src))
        BRPA_HIP(ctx, hipMemcpy(dst, src, 8 * n, hipMemcpyDeviceToHost));
    else
        std::memcpy(dst, src, 8 * n);
    return BRPA_OK;
}

// Work-group counts per fragment for every stage.
struct Counts {
    int64_t dtab, syrk, setdiag, split, syrk16;
    std::vector<int64_t> pdiag, trsm, update;   // per Cholesky step
};

Counts counts_of(const HostFrag &f, int nsteps_max, int precision) {
    Counts c;
    const int64_t K = f.nocc * f.nvir, N = f.naux, F = f.nfreqs;
    c.dtab = cdiv(F * K, 256);
    const int64_t nt = cdiv(N, SYRK_T);
    c.syrk = (precision == BRPA_FP64) ? nt * (nt + 1) / 2 * cdiv(F, 4) : 0;
    c.setdiag = cdiv(F * N, 256);
    const int64_t kpad = cdiv(K, SPLIT_KB) * SPLIT_KB;
    c.split = (precision == BRPA_SPLIT_BF16) ? F * (kpad / SPLIT_KB) * cdiv(N, 64) : 0;
    const int64_t nt16 = cdiv(N, 128);
    c.syrk16 = (precision == BRPA_SPLIT_BF16) ? F * (nt16 * (nt16 + 1) / 2) : 0;
    c.pdiag.assign(nsteps_max, 0);
    c.trsm.assign(nsteps_max, 0);
    c.update.assign(nsteps_max, 0);
    for (int j = 0; j < nsteps_max; ++j) {
        int64_t j0 = (int64_t)j * CHOL_NB;
        if (N > j0) c.pdiag[j] = F;
        if (N > j0 + CHOL_NB) c.trsm[j] = F * (cdiv(N - j0, CHOL_NB) - 1);
        if (N > j0 + CHOL_NB) {
            int64_t t = cdiv(N - j0 - CHOL_NB, 128);   // 128x128 update super-tiles
            c.update[j] = F * (t * (t + 1) / 2);
        }
    }
    return c;
}

// Device bytes one fragment needs in the E_RPA pipeline.
size_t frag_bytes(const HostFrag &h, int precision) {
    const size_t K = h.nocc * h.nvir, N = h.naux, F = h.nfreqs;
    size_t b = 8 * (F * K + F * N * N + 4 * F * N + 64 * F + F * CHOL_NB * CHOL_NB) + 8 * K * N;
    if (precision == BRPA_SPLIT_BF16) {
        size_t kpad = cdiv(K, SPLIT_KB) * SPLIT_KB;
        b += 2 * F * NSLICE * N * kpad + (4 + 16) * F * (kpad / SPLIT_KB) * N;
    }
    return b;
}

// Runs the whole E_RPA pipeline for one chunk of fragments.  energies: host
// array (nfrag); eu_out (optional): host pointers per fragment for E(u).
int run_chunk(brpa_ctx *ctx, int precision, const std::vector<HostFrag> &fr, double *energies,
              double *const *eu_out, double *kernel_ms) {
    const int nfrag = (int)fr.size();
    if (nfrag <= 0) return BRPA_OK;
    int nsteps_max = 0;
    for (int f = 0; f < nfrag; ++f)
        nsteps_max = std::max<int>(nsteps_max, (int)cdiv(fr[f].naux, CHOL_NB));

    std::vector<Counts> cnt(nfrag);
    for (int f = 0; f < nfrag; ++f) cnt[f] = counts_of(fr[f], nsteps_max, precision);

    // prefix arrays: [dtab, syrk, setdiag, split, syrk16, pdiag_0..S-1, trsm_0..S-1, update_0..]
    const int nstage = 5 + 3 * nsteps_max;
    std::vector<int64_t> pre((size_t)nstage * (nfrag + 1), 0);
    auto P = [&](int s) { return pre.data() + (size_t)s * (nfrag + 1); };
    for (int f = 0; f < nfrag; ++f) {
        const Counts &c = cnt[f];
        int64_t v[5] = {c.dtab, c.syrk, c.setdiag, c.split, c.syrk16};
        for (int s = 0; s < 5; ++s) P(s)[f + 1] = P(s)[f] + v[s];
        for (int j = 0; j < nsteps_max; ++j) {
            P(5 + j)[f + 1] = P(5 + j)[f] + c.pdiag[j];
            P(5 + nsteps_max + j)[f + 1] = P(5 + nsteps_max + j)[f] + c.trsm[j];
            P(5 + 2 * nsteps_max + j)[f + 1] = P(5 + 2 * nsteps_max + j)[f] + c.update[j];
        }
    }

    // small per-fragment arrays (weights, freqs, eocc, evir) are packed into one
    // host buffer and uploaded with a single copy
    size_t nsmall = 0;
    for (int f = 0; f < nfrag; ++f) nsmall += 2 * fr[f].nfreqs + fr[f].nocc + fr[f].nvir;
    std::vector<double> hsmall(nsmall);

    // workspace layout (two passes: count, then assign)
    std::vector<FragDev> hd(nfrag);
    double *dsmall = nullptr;
    auto layout = [&](Arena &ar) {
        dsmall = ar.take<double>(nsmall);
        size_t so = 0;
        for (int f = 0; f < nfrag; ++f) {
            const HostFrag &h = fr[f];
            FragDev &d = hd[f];
            const int64_t K = h.nocc * h.nvir, N = h.naux, F = h.nfreqs;
            d.K = (int)K;
            d.N = (int)N;
            d.F = (int)F;
            d.nocc = (int)h.nocc;
            d.nvir = (int)h.nvir;
            d.nsteps = nsteps_max;
            d.R = is_device_ptr(h.R) ? h.R : ar.take<double>(K * N);
            d.dtab = ar.take<double>(F * K);
            d.Pi = ar.take<double>(F * N * N);
            d.diag = ar.take<double>(F * N);
            d.ldpart = ar.take<double>(F * nsteps_max * 2);
            d.weights = dsmall ? dsmall + so : nullptr;
            d.freqs = dsmall ? dsmall + so + F : nullptr;
            d.eocc = dsmall ? dsmall + so + 2 * F : nullptr;
            d.evir = dsmall ? dsmall + so + 2 * F + h.nocc : nullptr;
            so += 2 * F + h.nocc + h.nvir;
            d.eu = ar.take<double>(F);
            d.energy = ar.take<double>(1);
            d.status = ar.take<int>(F);
            d.linv = ar.take<double>(F * CHOL_NB * CHOL_NB);
            d.Kpad = (int)(cdiv(K, SPLIT_KB) * SPLIT_KB);
            d.nblk = d.Kpad / SPLIT_KB;
            if (precision == BRPA_SPLIT_BF16) {
                d.sl = ar.take<uint16_t>((size_t)F * NSLICE * N * d.Kpad);
                d.scale = ar.take<float>((size_t)F * d.nblk * N);
                d.dpart = ar.take<double>((size_t)F * d.nblk * N * 2);
            } else {
                d.sl = nullptr;
                d.scale = nullptr;
                d.dpart = nullptr;
            }
        }
    };
    Arena count;
    layout(count);
    const size_t desc_bytes = ((sizeof(FragDev) * nfrag + 255) / 256) * 256;
    const size_t pre_bytes = ((sizeof(int64_t) * pre.size() + 255) / 256) * 256;
    Arena ar;
    BRPA_TRY(ws_reserve(ctx, count.off + desc_bytes + pre_bytes, &ar));
    FragDev *d_fr = ar.take<FragDev>(nfrag);
    int64_t *d_pre = ar.take<int64_t>(pre.size());
    layout(ar);

    size_t so = 0;
    for (int f = 0; f < nfrag; ++f) {
        const HostFrag &h = fr[f];
        const FragDev &d = hd[f];
        const int64_t K = h.nocc * h.nvir, N = h.naux, F = h.nfreqs;
        if (d.R != h.R) BRPA_TRY(to_device(ctx, const_cast<double *>(d.R), h.R, 8 * K * N));
        BRPA_TRY(host_copy(ctx, hsmall.data() + so, h.weights, F));
        BRPA_TRY(host_copy(ctx, hsmall.data() + so + F, h.freqs, F));
        BRPA_TRY(host_copy(ctx, hsmall.data() + so + 2 * F, h.eocc, h.nocc));
        BRPA_TRY(host_copy(ctx, hsmall.data() + so + 2 * F + h.nocc, h.evir, h.nvir));
        so += 2 * F + h.nocc + h.nvir;
        BRPA_TRY(launch_fill_int(ctx, d.status, F, 0));
    }
    BRPA_HIP(ctx, hipMemcpyAsync(dsmall, hsmall.data(), 8 * nsmall, hipMemcpyHostToDevice,
                                 ctx->stream));
    BRPA_HIP(ctx, hipMemcpyAsync(d_fr, hd.data(), sizeof(FragDev) * nfrag, hipMemcpyHostToDevice,
                                 ctx->stream));
    BRPA_HIP(ctx, hipMemcpyAsync(d_pre, pre.data(), sizeof(int64_t) * pre.size(),
                                 hipMemcpyHostToDevice, ctx->stream));
    auto DP = [&](int s) { return d_pre + (size_t)s * (nfrag + 1); };
    auto TOT = [&](int s) { return P(s)[nfrag]; };

    BRPA_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    BRPA_TRY(stage_begin(ctx, BRPA_STAGE_AUX));
    BRPA_TRY(launch_dtab(ctx, d_fr, DP(0), nfrag, TOT(0)));
    BRPA_TRY(stage_end(ctx, BRPA_STAGE_AUX));
    if (precision == BRPA_FP64) {
        BRPA_TRY(stage_begin(ctx, BRPA_STAGE_PI));
        BRPA_TRY(launch_syrk_f64(ctx, d_fr, DP(1), nfrag, TOT(1)));
        BRPA_TRY(stage_end(ctx, BRPA_STAGE_PI));
    } else {
        BRPA_TRY(stage_begin(ctx, BRPA_STAGE_SPLIT));
        BRPA_TRY(launch_split(ctx, d_fr, DP(3), nfrag, TOT(3)));
        BRPA_TRY(stage_end(ctx, BRPA_STAGE_SPLIT));
        BRPA_TRY(stage_begin(ctx, BRPA_STAGE_PI));
        BRPA_TRY(launch_syrk_bf16(ctx, d_fr, DP(4), nfrag, TOT(4)));
        BRPA_TRY(stage_end(ctx, BRPA_STAGE_PI));
    }
    BRPA_TRY(stage_begin(ctx, BRPA_STAGE_CHOL));
    BRPA_TRY(launch_set_diag(ctx, d_fr, DP(2), nfrag, TOT(2), precision == BRPA_SPLIT_BF16));
    for (int j = 0; j < nsteps_max; ++j) {
        BRPA_TRY(launch_potrf_diag(ctx, d_fr, DP(5 + j), nfrag, TOT(5 + j), j));
        BRPA_TRY(launch_trsm(ctx, d_fr, DP(5 + nsteps_max + j), nfrag, TOT(5 + nsteps_max + j), j));
        BRPA_TRY(launch_chol_update(ctx, d_fr, DP(5 + 2 * nsteps_max + j), nfrag,
                                    TOT(5 + 2 * nsteps_max + j), j));
    }
    BRPA_TRY(stage_end(ctx, BRPA_STAGE_CHOL));
    BRPA_TRY(launch_energy(ctx, d_fr, nfrag));
    BRPA_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));

    std::vector<int> status;
    for (int f = 0; f < nfrag; ++f) {
        const FragDev &d = hd[f];
        BRPA_TRY(from_device(ctx, energies + f, d.energy, sizeof(double)));
        if (eu_out && eu_out[f]) BRPA_TRY(from_device(ctx, eu_out[f], d.eu, 8 * (size_t)d.F));
    }
    std::vector<std::vector<int>> st(nfrag);
    for (int f = 0; f < nfrag; ++f) {
        st[f].resize(hd[f].F);
        BRPA_HIP(ctx, hipMemcpyAsync(st[f].data(), hd[f].status, sizeof(int) * hd[f].F,
                                     hipMemcpyDeviceToHost, ctx->stream));
    }
    BRPA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    float ms = 0.f;
    BRPA_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    *kernel_ms += ms;
    BRPA_TRY(stage_collect(ctx, BRPA_STAGE_AUX));
    BRPA_TRY(stage_collect(ctx, BRPA_STAGE_PI));
    if (precision != BRPA_FP64) BRPA_TRY(stage_collect(ctx, BRPA_STAGE_SPLIT));
    BRPA_TRY(stage_collect(ctx, BRPA_STAGE_CHOL));
    for (int f = 0; f < nfrag; ++f)
        for (int w = 0; w < hd[f].F; ++w)
            if (st[f][w])
                return brpa_fail(ctx, BRPA_ERR_NOT_SPD,
                                 "fragment %d, frequency %d: 1 + Pi(u) is not positive definite",
                                 f, w);
    return BRPA_OK;
}

// Validates every fragment, then runs the pipeline in chunks whose device
// workspace stays under the budget (BRPA_WS_BUDGET_GB, default 64 GiB).
int run_ecrpa(brpa_ctx *ctx, int precision, const std::vector<HostFrag> &fr, double *energies,
              double *const *eu_out) {
    if (precision != BRPA_FP64 && precision != BRPA_SPLIT_BF16)
        return brpa_fail(ctx, BRPA_ERR_ARG, "unknown precision %d", precision);
    for (size_t f = 0; f < fr.size(); ++f) BRPA_TRY(check_frag(ctx, fr[f], (int64_t)f));
    BRPA_HIP(ctx, hipSetDevice(ctx->device));
    double budget_gb = 64.0;
    if (const char *e = std::getenv("BRPA_WS_BUDGET_GB")) budget_gb = std::atof(e);
    const size_t budget = (size_t)(budget_gb * (1ull << 30));
    double ms = 0.0;
    stage_reset(ctx);
    size_t i = 0;
    while (i < fr.size()) {
        size_t j = i, bytes = 0;
        while (j < fr.size() && (j == i || bytes + frag_bytes(fr[j], precision) <= budget))
            bytes += frag_bytes(fr[j++], precision);
        std::vector<HostFrag> chunk(fr.begin() + i, fr.begin() + j);
        BRPA_TRY(run_chunk(ctx, precision, chunk, energies + i, eu_out ? eu_out + i : nullptr, &ms));
        i = j;
    }
    ctx->last_kernel_ms = ms;
    return BRPA_OK;
}

}  // namespace

extern "C" {

int brpa_ecrpa(brpa_ctx *ctx, int precision, const double *R, int64_t naux, const double *eocc,
               int64_t nocc, const double *evir, int64_t nvir, const double *freqs,
               const double *weights, int64_t nfreqs, double *ecrpa, double *eu) {
    if (!ctx) return BRPA_ERR_ARG;
    if (!ecrpa) return brpa_fail(ctx, BRPA_ERR_ARG, "ecrpa output is null");
    std::vector<HostFrag> fr(1);
    fr[0] = {nocc, nvir, naux, nfreqs, R, eocc, evir, freqs, weights};
    double *eus[1] = {eu};
    return run_ecrpa(ctx, precision, fr, ecrpa, eu ? eus : nullptr);
}

int brpa_ecrpa_batch(brpa_ctx *ctx, int precision, int64_t nfrag, const brpa_frag_t *frags,
                     double *ecrpa) {
    if (!ctx) return BRPA_ERR_ARG;
    if (nfrag < 0 || (nfrag > 0 && (!frags || !ecrpa)))
        return brpa_fail(ctx, BRPA_ERR_ARG, "bad batch arguments");
    std::vector<HostFrag> fr(nfrag);
    for (int64_t f = 0; f < nfrag; ++f)
        fr[f] = {frags[f].nocc, frags[f].nvir, frags[f].naux, frags[f].nfreqs, frags[f].R,
                 frags[f].eocc, frags[f].evir, frags[f].freqs, frags[f].weights};
    return run_ecrpa(ctx, precision, fr, ecrpa, nullptr);
}

}  // extern "C"

extern "C" {

int brpa_gpiug_mo(brpa_ctx *ctx, double *gpiug, const double *crg, int64_t naux,
                  const double *eocc, int64_t nocc, const double *evir, int64_t nvir,
                  double freq, double prefactor) {
    if (!ctx) return BRPA_ERR_ARG;
    if (!gpiug || !crg || !eocc || !evir || naux <= 0 || nocc < 0 || nvir < 0)
        return brpa_fail(ctx, BRPA_ERR_ARG, "brpa_gpiug_mo: bad arguments");
    const int64_t K = nocc * nvir, N = naux;
    if (K == 0) return BRPA_OK;   // rpa_core_MO.f90:387 (Nai > 0)
    if (K > INT32_MAX || N > 65536)
        return brpa_fail(ctx, BRPA_ERR_ARG, "brpa_gpiug_mo: sizes beyond the supported range");
    BRPA_HIP(ctx, hipSetDevice(ctx->device));
    std::vector<double> eo(nocc), ev(nvir);
    if (is_device_ptr(eocc)) {
        BRPA_HIP(ctx, hipMemcpy(eo.data(), eocc, 8 * nocc, hipMemcpyDeviceToHost));
        BRPA_HIP(ctx, hipMemcpy(ev.data(), evir, 8 * nvir, hipMemcpyDeviceToHost));
    } else {
        std::memcpy(eo.data(), eocc, 8 * nocc);
        std::memcpy(ev.data(), evir, 8 * nvir);
    }
    if (!finite_all(eo.data(), nocc) || !finite_all(ev.data(), nvir) || !std::isfinite(freq))
        return brpa_fail(ctx, BRPA_ERR_NONFINITE, "brpa_gpiug_mo: non-finite input");
    // DCRG scaling W_ai = d/(d^2+u^2) (rpa_core_MO.f90:390-401), prefactor folded in
    std::vector<double> c(K);
    for (int64_t i = 0; i < nocc; ++i)
        for (int64_t a = 0; a < nvir; ++a) {
            double d = ev[a] - eo[i];
            c[i * nvir + a] = prefactor * (d / (d * d + freq * freq));
        }
    const bool rdev = is_device_ptr(crg), pdev = is_device_ptr(gpiug);
    Arena count;
    count.take<double>(K);
    if (!rdev) count.take<double>(K * N);
    if (!pdev) count.take<double>(N * N);
    Arena ar;
    BRPA_TRY(ws_reserve(ctx, count.off, &ar));
    double *dc = ar.take<double>(K);
    const double *dR = rdev ? crg : ar.take<double>(K * N);
    double *dP = pdev ? gpiug : ar.take<double>(N * N);
    BRPA_HIP(ctx, hipMemcpyAsync(dc, c.data(), 8 * K, hipMemcpyHostToDevice, ctx->stream));
    if (!rdev) BRPA_TRY(to_device(ctx, const_cast<double *>(dR), crg, 8 * K * N));
    if (!pdev) BRPA_TRY(to_device(ctx, dP, gpiug, 8 * N * N));
    BRPA_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    BRPA_TRY(launch_syrk_single(ctx, dR, dc, (int)K, (int)N, dP, 1));
    BRPA_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    if (!pdev) BRPA_TRY(from_device(ctx, gpiug, dP, 8 * N * N));
    BRPA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    float ms = 0.f;
    BRPA_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    ctx->last_kernel_ms = ms;
    return BRPA_OK;
}

int brpa_ecrpa_analytic(brpa_ctx *ctx, double *ecrpa, const double *piu, const double *freqs,
                        const double *weights, int64_t nfreqs, int64_t naux, double *eu) {
    if (!ctx) return BRPA_ERR_ARG;
    if (!ecrpa || !piu || !weights || nfreqs <= 0 || naux <= 0 || naux > 65536)
        return brpa_fail(ctx, BRPA_ERR_ARG, "brpa_ecrpa_analytic: bad arguments");
    (void)freqs;   // printed by the reference only (rpa_CC.f90:462-463)
    BRPA_HIP(ctx, hipSetDevice(ctx->device));
    const int64_t N = naux, F = nfreqs;
    const int nsteps = (int)cdiv(N, CHOL_NB);
    // prefix arrays: [setdiag, pdiag_0..S-1, trsm_0..S-1, update_0..S-1]
    std::vector<int64_t> pre((size_t)(1 + 3 * nsteps) * 2, 0);
    pre[1] = cdiv(F * N, 256);
    for (int j = 0; j < nsteps; ++j) {
        int64_t j0 = (int64_t)j * CHOL_NB;
        pre[2 * (1 + j) + 1] = (N > j0) ? F : 0;
        pre[2 * (1 + nsteps + j) + 1] = (N > j0 + CHOL_NB) ? F * (cdiv(N - j0, CHOL_NB) - 1) : 0;
        int64_t t = (N > j0 + CHOL_NB) ? cdiv(N - j0 - CHOL_NB, 128) : 0;   // update super-tiles
        pre[2 * (1 + 2 * nsteps + j) + 1] = F * (t * (t + 1) / 2);
    }
    FragDev hd{};
    auto layout = [&](Arena &ar) {
        hd.K = 0;
        hd.N = (int)N;
        hd.F = (int)F;
        hd.nsteps = nsteps;
        hd.Pi = ar.take<double>(F * N * N);
        hd.diag = ar.take<double>(F * N);
        hd.ldpart = ar.take<double>(F * nsteps * 2);
        hd.weights = ar.take<double>(F);
        hd.eu = ar.take<double>(F);
        hd.energy = ar.take<double>(1);
        hd.status = ar.take<int>(F);
        hd.linv = ar.take<double>(F * CHOL_NB * CHOL_NB);
    };
    Arena count;
    count.take<FragDev>(1);
    count.take<int64_t>(pre.size());
    layout(count);
    Arena ar;
    BRPA_TRY(ws_reserve(ctx, count.off, &ar));
    FragDev *d_fr = ar.take<FragDev>(1);
    int64_t *d_pre = ar.take<int64_t>(pre.size());
    layout(ar);
    BRPA_TRY(to_device(ctx, hd.Pi, piu, 8 * F * N * N));
    BRPA_TRY(to_device(ctx, const_cast<double *>(hd.weights), weights, 8 * F));
    BRPA_TRY(launch_fill_int(ctx, hd.status, F, 0));
    BRPA_HIP(ctx, hipMemcpyAsync(d_fr, &hd, sizeof(FragDev), hipMemcpyHostToDevice, ctx->stream));
    BRPA_HIP(ctx, hipMemcpyAsync(d_pre, pre.data(), 8 * pre.size(), hipMemcpyHostToDevice,
                                 ctx->stream));
    BRPA_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    BRPA_TRY(launch_set_diag(ctx, d_fr, d_pre, 1, pre[1], 0));
    for (int j = 0; j < nsteps; ++j) {
        BRPA_TRY(launch_potrf_diag(ctx, d_fr, d_pre + 2 * (1 + j), 1, pre[2 * (1 + j) + 1], j));
        BRPA_TRY(launch_trsm(ctx, d_fr, d_pre + 2 * (1 + nsteps + j), 1,
                             pre[2 * (1 + nsteps + j) + 1], j));
        BRPA_TRY(launch_chol_update(ctx, d_fr, d_pre + 2 * (1 + 2 * nsteps + j), 1,
                                    pre[2 * (1 + 2 * nsteps + j) + 1], j));
    }
    BRPA_TRY(launch_energy(ctx, d_fr, 1));
    BRPA_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    std::vector<int> st(F);
    BRPA_TRY(from_device(ctx, ecrpa, hd.energy, 8));
    if (eu) BRPA_TRY(from_device(ctx, eu, hd.eu, 8 * F));
    BRPA_HIP(ctx, hipMemcpyAsync(st.data(), hd.status, sizeof(int) * F, hipMemcpyDeviceToHost,
                                 ctx->stream));
    BRPA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    float ms = 0.f;
    BRPA_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    ctx->last_kernel_ms = ms;
    for (int64_t w = 0; w < F; ++w)
        if (st[w])
            return brpa_fail(ctx, BRPA_ERR_NOT_SPD,
                             "frequency %lld: 1 + Pi(u) is not positive definite", (long long)w);
    return BRPA_OK;
}

}  // extern "C"
