// Host orchestration of the E_RPA hot path: Pi(u) build, exact diagonal,
// batched blocked Cholesky log-det and the quadrature, for one fragment or a
// batch of uneven fragments in the same launches.
#include <rocblas/rocblas.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "internal.h"

int rocblas_handle_of(brpa_ctx *ctx, void **h);
int launch_col_scale(brpa_ctx *, double *out, const double *in, const double *s, int64_t m,
                     int64_t n);

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
        BRPA_TRY(copy_sync(ctx, eo.data(), f.eocc, 8 * f.nocc, hipMemcpyDeviceToHost));
        BRPA_TRY(copy_sync(ctx, ev.data(), f.evir, 8 * f.nvir, hipMemcpyDeviceToHost));
    } else {
        std::memcpy(eo.data(), f.eocc, 8 * f.nocc);
        std::memcpy(ev.data(), f.evir, 8 * f.nvir);
    }
    if (is_device_ptr(f.freqs)) {
        BRPA_TRY(copy_sync(ctx, fx.data(), f.freqs, 8 * f.nfreqs, hipMemcpyDeviceToHost));
        BRPA_TRY(copy_sync(ctx, fw.data(), f.weights, 8 * f.nfreqs, hipMemcpyDeviceToHost));
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
        BRPA_TRY(copy_sync(ctx, dst, src, 8 * n, hipMemcpyDeviceToHost));
    else
        std::memcpy(dst, src, 8 * n);
    return BRPA_OK;
}

// Work-group counts per fragment for every stage.
struct Counts {
    int64_t dtab, syrk, setdiag, split, syrk16, colscale8, split8;
    int64_t syrk8[I8V_COUNT];   // split-int8 product work-groups, per kernel variant
    std::vector<int64_t> pdiag, trsm, update;   // per Cholesky step
};

Counts counts_of(const HostFrag &f, int nsteps_max, int precision, int i8_variant) {
    Counts c;
    const int64_t K = f.nocc * f.nvir, N = f.naux, F = f.nfreqs;
    c.dtab = cdiv(F * K, 256);
    const int64_t nt = cdiv(N, SYRK_T);
    c.syrk = (precision == BRPA_FP64) ? nt * (nt + 1) / 2 * cdiv(F, 4) : 0;
    c.setdiag = cdiv(F * N, 256);
    const int64_t kpad = cdiv(K, SPLIT_KB) * SPLIT_KB;
    c.split = (precision == BRPA_SPLIT_BF16) ? (kpad / SPLIT_KB) * cdiv(N, 64) : 0;
    const int64_t nt16 = cdiv(N, 128);
    c.syrk16 = (precision == BRPA_SPLIT_BF16) ? F * (nt16 * (nt16 + 1) / 2) : 0;
    const bool i8 = precision == BRPA_SPLIT_I8;
    c.colscale8 = i8 ? cdiv(N, 64) : 0;
    c.split8 = i8 ? cdiv(K, I8_KB) * cdiv(N, 64) : 0;
    for (int v = 0; v < I8V_COUNT; ++v) {
        const int64_t t = cdiv(N, syrk_i8_tile(v));
        c.syrk8[v] = (i8 && v == i8_variant) ? F * (t * (t + 1) / 2) : 0;
    }
    c.pdiag.assign(nsteps_max, 0);
    c.trsm.assign(nsteps_max, 0);
    c.update.assign(nsteps_max, 0);
    for (int j = 0; j < nsteps_max; ++j) {
        int64_t j0 = (int64_t)j * CHOL_NB;
        if (N > j0) c.pdiag[j] = F;
        if (N > j0 + CHOL_NB) c.trsm[j] = F * (cdiv(N - j0, CHOL_NB) - 1);
        c.update[j] = F * chol_upd_items(N, j);
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
    if (precision == BRPA_SPLIT_I8) {
        size_t kpad = cdiv(K, I8_KB) * I8_KB;
        b += F * NSLICE * N * kpad + 8 * F * N + 8 * F * K;
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

    // split-int8 product variant of every fragment (autotuner choice per size bucket)
    std::vector<int> var(nfrag, I8V_T128);
    for (int f = 0; f < nfrag; ++f) {
        const int c = ctx->i8_choice[tune_bucket(fr[f].naux)];
        var[f] = ctx->tune_force >= 0 ? ctx->tune_force : (c >= 0 ? c : I8V_T128);
    }
    std::vector<Counts> cnt(nfrag);
    for (int f = 0; f < nfrag; ++f) cnt[f] = counts_of(fr[f], nsteps_max, precision, var[f]);

    // prefix arrays: [dtab, syrk, setdiag, split, syrk16, colscale8, split8,
    //                 syrk8 per variant, pdiag_0..S-1, trsm_0..S-1, update_0..S-1]
    constexpr int NFIX = 7 + I8V_COUNT;
    const int nstage = NFIX + 3 * nsteps_max;
    std::vector<int64_t> pre((size_t)nstage * (nfrag + 1), 0);
    auto P = [&](int s) { return pre.data() + (size_t)s * (nfrag + 1); };
    for (int f = 0; f < nfrag; ++f) {
        const Counts &c = cnt[f];
        int64_t v[NFIX] = {c.dtab,     c.syrk,     c.setdiag,  c.split,   c.syrk16,
                           c.colscale8, c.split8,  c.syrk8[0], c.syrk8[1], c.syrk8[2]};
        for (int s = 0; s < NFIX; ++s) P(s)[f + 1] = P(s)[f] + v[s];
        for (int j = 0; j < nsteps_max; ++j) {
            P(NFIX + j)[f + 1] = P(NFIX + j)[f] + c.pdiag[j];
            P(NFIX + nsteps_max + j)[f + 1] = P(NFIX + nsteps_max + j)[f] + c.trsm[j];
            P(NFIX + 2 * nsteps_max + j)[f + 1] = P(NFIX + 2 * nsteps_max + j)[f] + c.update[j];
        }
    }

    // Inputs reach the device through one upload of a pinned staging region
    //   [small arrays (weights, freqs, eocc, evir per fragment) | FragDev[] | pre]
    // and results come back in one download of
    //   [energy (nfrag) | E(u) (sum F) | status (sum F)].
    size_t nsmall = 0, nfreq_all = 0;
    for (int f = 0; f < nfrag; ++f) {
        nsmall += 2 * fr[f].nfreqs + fr[f].nocc + fr[f].nvir;
        nfreq_all += fr[f].nfreqs;
    }
    auto pad = [](size_t b) { return ((b + 255) / 256) * 256; };
    const size_t up_small = pad(8 * nsmall), up_fr = pad(sizeof(FragDev) * nfrag),
                 up_pre = pad(sizeof(int64_t) * pre.size());
    const size_t up_bytes = up_small + up_fr + up_pre;
    const size_t dn_e = pad(8 * (size_t)nfrag), dn_eu = pad(8 * nfreq_all),
                 dn_st = pad(4 * nfreq_all);
    const size_t dn_bytes = dn_e + dn_eu + dn_st;
    BRPA_TRY(pin_reserve(ctx, up_bytes + dn_bytes));
    char *h_up = ctx->pin, *h_dn = ctx->pin + up_bytes;
    double *hsmall = reinterpret_cast<double *>(h_up);
    FragDev *hd = reinterpret_cast<FragDev *>(h_up + up_small);
    int64_t *hpre = reinterpret_cast<int64_t *>(h_up + up_small + up_fr);

    // workspace layout (two passes: count, then assign)
    char *d_up = nullptr, *d_dn = nullptr;
    auto layout = [&](Arena &ar) {
        d_up = ar.take<char>(up_bytes);
        d_dn = ar.take<char>(dn_bytes);
        double *dsmall = reinterpret_cast<double *>(d_up);
        double *d_energy = reinterpret_cast<double *>(d_dn);
        double *d_eu = reinterpret_cast<double *>(d_dn + dn_e);
        int *d_status = reinterpret_cast<int *>(d_dn + dn_e + dn_eu);
        size_t so = 0, fo = 0;
        for (int f = 0; f < nfrag; ++f) {
            const HostFrag &h = fr[f];
            FragDev d{};
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
            d.weights = d_up ? dsmall + so : nullptr;
            d.freqs = d_up ? dsmall + so + F : nullptr;
            d.eocc = d_up ? dsmall + so + 2 * F : nullptr;
            d.evir = d_up ? dsmall + so + 2 * F + h.nocc : nullptr;
            so += 2 * F + h.nocc + h.nvir;
            d.eu = d_dn ? d_eu + fo : nullptr;
            d.energy = d_dn ? d_energy + f : nullptr;
            d.status = d_dn ? d_status + fo : nullptr;
            d.mat0 = (int64_t)fo;
            fo += F;
            d.linv = ar.take<double>(F * CHOL_NB * CHOL_NB);
            d.Kpad = (int)(cdiv(K, SPLIT_KB) * SPLIT_KB);
            d.nblk = d.Kpad / SPLIT_KB;
            if (precision == BRPA_SPLIT_BF16) {
                d.sl = ar.take<uint16_t>((size_t)F * NSLICE * N * d.Kpad);
                d.scale = ar.take<float>((size_t)F * d.nblk * N);
                d.dpart = ar.take<double>((size_t)F * d.nblk * N * 2);
            }
            d.Kpad8 = (int)(cdiv(K, I8_KB) * I8_KB);
            if (precision == BRPA_SPLIT_I8) {
                d.sdtab = ar.take<double>((size_t)F * K);
                d.sl8 = ar.take<int8_t>((size_t)F * NSLICE * N * d.Kpad8);
                d.scale8 = ar.take<double>((size_t)F * N);
            }
            if (ar.base) hd[f] = d;
        }
    };
    Arena count;
    layout(count);
    Arena ar;
    BRPA_TRY(ws_reserve(ctx, count.off, &ar));
    layout(ar);
    const FragDev *d_fr = reinterpret_cast<const FragDev *>(d_up + up_small);
    const int64_t *d_pre = reinterpret_cast<const int64_t *>(d_up + up_small + up_fr);

    // host staging: small arrays, descriptors (filled by layout), prefix sums
    bool all_dev = true;
    {
        size_t so = 0;
        for (int f = 0; f < nfrag; ++f) {
            const HostFrag &h = fr[f];
            const int64_t F = h.nfreqs;
            BRPA_TRY(host_copy(ctx, hsmall + so, h.weights, F));
            BRPA_TRY(host_copy(ctx, hsmall + so + F, h.freqs, F));
            BRPA_TRY(host_copy(ctx, hsmall + so + 2 * F, h.eocc, h.nocc));
            BRPA_TRY(host_copy(ctx, hsmall + so + 2 * F + h.nocc, h.evir, h.nvir));
            so += 2 * F + h.nocc + h.nvir;
            if (hd[f].R != h.R) all_dev = false;
        }
        std::memcpy(hpre, pre.data(), sizeof(int64_t) * pre.size());
    }
    auto DP = [&](int s) { return d_pre + (size_t)s * (nfrag + 1); };
    auto TOT = [&](int s) { return P(s)[nfrag]; };

    // the launch sequence; `timed` records the per-stage events (not in graphs)
    auto enqueue = [&](bool timed) -> int {
        auto sb = [&](int st) { return timed ? stage_begin(ctx, st) : BRPA_OK; };
        auto se = [&](int st) { return timed ? stage_end(ctx, st) : BRPA_OK; };
        BRPA_HIP(ctx, hipMemcpyAsync(d_up, h_up, up_bytes, hipMemcpyHostToDevice, ctx->stream));
        // status starts at 0 (a fill kernel: no memset nodes in graphs)
        BRPA_TRY(launch_fill_int(ctx, reinterpret_cast<int *>(d_dn + dn_e + dn_eu),
                                 (int64_t)dn_st / 4, 0));
        BRPA_TRY(sb(BRPA_STAGE_AUX));
        BRPA_TRY(launch_dtab(ctx, d_fr, DP(0), nfrag, TOT(0)));
        BRPA_TRY(se(BRPA_STAGE_AUX));
        if (precision == BRPA_FP64) {
            BRPA_TRY(sb(BRPA_STAGE_PI));
            BRPA_TRY(launch_syrk_f64(ctx, d_fr, DP(1), nfrag, TOT(1)));
            BRPA_TRY(se(BRPA_STAGE_PI));
        } else if (precision == BRPA_SPLIT_I8) {
            BRPA_TRY(sb(BRPA_STAGE_SPLIT));
            BRPA_TRY(launch_colscale_i8(ctx, d_fr, DP(5), nfrag, TOT(5)));
            BRPA_TRY(launch_split_i8(ctx, d_fr, DP(6), nfrag, TOT(6)));
            BRPA_TRY(se(BRPA_STAGE_SPLIT));
            BRPA_TRY(sb(BRPA_STAGE_PI));
            for (int v = 0; v < I8V_COUNT; ++v)
                BRPA_TRY(launch_syrk_i8(ctx, d_fr, DP(7 + v), nfrag, TOT(7 + v), v));
            BRPA_TRY(se(BRPA_STAGE_PI));
        } else {
            BRPA_TRY(sb(BRPA_STAGE_SPLIT));
            BRPA_TRY(launch_split(ctx, d_fr, DP(3), nfrag, TOT(3)));
            BRPA_TRY(se(BRPA_STAGE_SPLIT));
            BRPA_TRY(sb(BRPA_STAGE_PI));
            BRPA_TRY(launch_syrk_bf16(ctx, d_fr, DP(4), nfrag, TOT(4)));
            BRPA_TRY(se(BRPA_STAGE_PI));
        }
        BRPA_TRY(sb(BRPA_STAGE_CHOL));
        const int diag_mode =
            precision == BRPA_SPLIT_BF16 ? 1 : (precision == BRPA_SPLIT_I8 ? 2 : 0);
        BRPA_TRY(launch_set_diag(ctx, d_fr, DP(2), nfrag, TOT(2), diag_mode));
        for (int j = 0; j < nsteps_max; ++j) {
            BRPA_TRY(launch_potrf_diag(ctx, d_fr, DP(NFIX + j), nfrag, TOT(NFIX + j), j));
            BRPA_TRY(launch_trsm(ctx, d_fr, DP(NFIX + nsteps_max + j), nfrag,
                                 TOT(NFIX + nsteps_max + j), j));
            BRPA_TRY(launch_chol_update(ctx, d_fr, DP(NFIX + 2 * nsteps_max + j), nfrag,
                                        TOT(NFIX + 2 * nsteps_max + j), j));
        }
        BRPA_TRY(se(BRPA_STAGE_CHOL));
        BRPA_TRY(launch_energy(ctx, d_fr, nfrag));
        BRPA_HIP(ctx, hipMemcpyAsync(h_dn, d_dn, dn_bytes, hipMemcpyDeviceToHost, ctx->stream));
        return BRPA_OK;
    };

    const bool graph = ctx->use_graphs && all_dev;
    if (graph) {
        // key: everything the captured addresses and launch shapes depend on
        std::string key = std::to_string(precision) + "/" + std::to_string(nfrag) + "/" +
                          std::to_string((uintptr_t)ctx->ws) + "/" +
                          std::to_string((uintptr_t)ctx->pin) + "/" +
                          std::to_string((uintptr_t)ctx->stream);
        for (int f = 0; f < nfrag; ++f)
            key += ":" + std::to_string(fr[f].nocc) + "," + std::to_string(fr[f].nvir) + "," +
                   std::to_string(fr[f].naux) + "," + std::to_string(fr[f].nfreqs) + "," +
                   std::to_string(var[f]);
        hipGraphExec_t exec = nullptr;
        for (auto &g : ctx->graphs)
            if (g.key == key) exec = g.exec;
        if (!exec) {
            BRPA_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
            int rc = enqueue(false);
            hipGraph_t gr = nullptr;
            hipError_t ce = hipStreamEndCapture(ctx->stream, &gr);
            if (rc != BRPA_OK) {
                if (gr) (void)hipGraphDestroy(gr);
                return rc;
            }
            BRPA_HIP(ctx, ce);
            hipError_t ie = hipGraphInstantiate(&exec, gr, nullptr, nullptr, 0);
            (void)hipGraphDestroy(gr);
            BRPA_HIP(ctx, ie);
            ctx->graphs.push_back({key, exec});
            ctx->last_graph_captures += 1;
        } else {
            ctx->last_graph_replays += 1;
        }
        BRPA_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
        BRPA_HIP(ctx, hipGraphLaunch(exec, ctx->stream));
        BRPA_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    } else {
        // fragments whose R is a host array are copied into the workspace first
        for (int f = 0; f < nfrag; ++f)
            if (hd[f].R != fr[f].R)
                BRPA_TRY(to_device(ctx, const_cast<double *>(hd[f].R), fr[f].R,
                                   8 * (size_t)fr[f].nocc * fr[f].nvir * fr[f].naux));
        BRPA_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
        BRPA_TRY(enqueue(true));
        BRPA_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    }
    BRPA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    float ms = 0.f;
    BRPA_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    *kernel_ms += ms;
    if (!graph) {
        BRPA_TRY(stage_collect(ctx, BRPA_STAGE_AUX));
        BRPA_TRY(stage_collect(ctx, BRPA_STAGE_PI));
        if (precision != BRPA_FP64) BRPA_TRY(stage_collect(ctx, BRPA_STAGE_SPLIT));
        BRPA_TRY(stage_collect(ctx, BRPA_STAGE_CHOL));
    }
    const double *h_energy = reinterpret_cast<const double *>(h_dn);
    const double *h_eu = reinterpret_cast<const double *>(h_dn + dn_e);
    const int *h_status = reinterpret_cast<const int *>(h_dn + dn_e + dn_eu);
    size_t fo = 0;
    for (int f = 0; f < nfrag; ++f) {
        energies[f] = h_energy[f];
        if (eu_out && eu_out[f]) {
            if (is_device_ptr(eu_out[f]))
                BRPA_TRY(copy_sync(ctx, eu_out[f], h_eu + fo, 8 * (size_t)fr[f].nfreqs,
                                        hipMemcpyHostToDevice));
            else
                std::memcpy(eu_out[f], h_eu + fo, 8 * (size_t)fr[f].nfreqs);
        }
        for (int w = 0; w < fr[f].nfreqs; ++w)
            if (h_status[fo + w])
                return brpa_fail(ctx, BRPA_ERR_NOT_SPD,
                                 "fragment %d, frequency %d: 1 + Pi(u) is not positive definite",
                                 f, w);
        fo += fr[f].nfreqs;
    }
    return BRPA_OK;
}

// Validates every fragment, then runs the pipeline in chunks whose device
// workspace stays under the budget (BRPA_WS_BUDGET_GB, default 64 GiB).
int run_ecrpa(brpa_ctx *ctx, int precision, const std::vector<HostFrag> &fr, double *energies,
              double *const *eu_out) {
    if (precision != BRPA_FP64 && precision != BRPA_SPLIT_BF16 && precision != BRPA_SPLIT_I8)
        return brpa_fail(ctx, BRPA_ERR_ARG, "unknown precision %d", precision);
    for (size_t f = 0; f < fr.size(); ++f) {
        BRPA_TRY(check_frag(ctx, fr[f], (int64_t)f));
        if (precision == BRPA_SPLIT_I8 && fr[f].nocc * fr[f].nvir > I8_KMAX)
            return brpa_fail(ctx, BRPA_ERR_ARG,
                             "fragment %lld: nocc*nvir = %lld exceeds %d, the largest K with "
                             "exact int32 sums in the split-int8 form",
                             (long long)f, (long long)(fr[f].nocc * fr[f].nvir), I8_KMAX);
    }
    BRPA_HIP(ctx, hipSetDevice(ctx->device));
    // workspace budget: BRPA_WS_BUDGET_GB, else 3/4 of the device memory that is
    // free now plus what the context already holds
    size_t budget;
    if (const char *e = std::getenv("BRPA_WS_BUDGET_GB")) {
        budget = (size_t)(std::atof(e) * (1ull << 30));
    } else {
        size_t free_b = 0, total_b = 0;
        BRPA_HIP(ctx, hipMemGetInfo(&free_b, &total_b));
        budget = (free_b + ctx->ws_cap) / 4 * 3;
    }
    double ms = 0.0;
    stage_reset(ctx);
    ctx->last_graph_captures = ctx->last_graph_replays = 0;
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

int64_t chol_pre_len(int64_t N) { return 3 * cdiv(N, CHOL_NB) * 2; }

int chol_single(brpa_ctx *ctx, const FragDev *d_fr, int64_t *d_pre, int64_t N, int64_t F) {
    const int nsteps = (int)cdiv(N, CHOL_NB);
    // prefix arrays [0, count] per launch: pdiag_0..S-1, trsm_0..S-1, update_0..S-1
    std::vector<int64_t> pre((size_t)chol_pre_len(N), 0);
    for (int j = 0; j < nsteps; ++j) {
        const int64_t j0 = (int64_t)j * CHOL_NB;
        pre[2 * j + 1] = (N > j0) ? F : 0;
        pre[2 * (nsteps + j) + 1] = (N > j0 + CHOL_NB) ? F * (cdiv(N - j0, CHOL_NB) - 1) : 0;
        pre[2 * (2 * nsteps + j) + 1] = F * chol_upd_items(N, j);
    }
    BRPA_TRY(copy_sync(ctx, d_pre, pre.data(), 8 * pre.size(), hipMemcpyHostToDevice));
    for (int j = 0; j < nsteps; ++j) {
        BRPA_TRY(launch_potrf_diag(ctx, d_fr, d_pre + 2 * j, 1, pre[2 * j + 1], j));
        BRPA_TRY(launch_trsm(ctx, d_fr, d_pre + 2 * (nsteps + j), 1, pre[2 * (nsteps + j) + 1], j));
        BRPA_TRY(launch_chol_update(ctx, d_fr, d_pre + 2 * (2 * nsteps + j), 1,
                                    pre[2 * (2 * nsteps + j) + 1], j));
    }
    return BRPA_OK;
}

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

int brpa_autotune(brpa_ctx *ctx, int precision, int64_t nfrag, const brpa_frag_t *frags) {
    if (!ctx) return BRPA_ERR_ARG;
    if (nfrag < 0 || (nfrag > 0 && !frags)) return brpa_fail(ctx, BRPA_ERR_ARG, "bad batch arguments");
    if (precision != BRPA_SPLIT_I8) return BRPA_OK;   // the other forms have one kernel shape
    std::vector<HostFrag> bucket[3];
    for (int64_t f = 0; f < nfrag; ++f)
        bucket[tune_bucket(frags[f].naux)].push_back(
            {frags[f].nocc, frags[f].nvir, frags[f].naux, frags[f].nfreqs, frags[f].R,
             frags[f].eocc, frags[f].evir, frags[f].freqs, frags[f].weights});
    BRPA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    graphs_release(ctx);
    const int saved_graphs = ctx->use_graphs;
    ctx->use_graphs = 0;   // eager runs: the per-stage events time the product kernel
    int rc = BRPA_OK;
    for (int b = 0; b < 3 && rc == BRPA_OK; ++b) {
        if (bucket[b].empty()) continue;
        std::vector<double> e(bucket[b].size()), e0;
        double best = 0.0;
        int best_v = I8V_T128;
        for (int v = 0; v < I8V_COUNT && rc == BRPA_OK; ++v) {
            ctx->tune_force = v;
            // the first run allocates and warms up; the second is timed
            rc = run_ecrpa(ctx, precision, bucket[b], e.data(), nullptr);
            if (rc == BRPA_OK) rc = run_ecrpa(ctx, precision, bucket[b], e.data(), nullptr);
            if (rc != BRPA_OK) break;
            // the int32 level sums are exact, so every tile shape gives the same Pi
            if (v == 0) {
                e0 = e;
            } else if (std::memcmp(e.data(), e0.data(), 8 * e.size()) != 0) {
                rc = brpa_fail(ctx, BRPA_ERR_INTERNAL,
                               "autotuner: split-int8 variant %d differs from variant 0", v);
                break;
            }
            const double ms = ctx->stage_ms[BRPA_STAGE_PI];
            ctx->i8_tune_ms[b][v] = ms;
            if (v == 0 || ms < best) {
                best = ms;
                best_v = v;
            }
        }
        if (rc == BRPA_OK) ctx->i8_choice[b] = best_v;
    }
    ctx->tune_force = -1;
    ctx->use_graphs = saved_graphs;
    return rc;
}

int brpa_autotune_result(const brpa_ctx *ctx, int bucket, int *variant, double *ms_per_variant) {
    if (!ctx || bucket < 0 || bucket >= 3) return BRPA_ERR_ARG;
    if (variant) *variant = ctx->i8_choice[bucket];
    if (ms_per_variant)
        for (int v = 0; v < I8V_COUNT; ++v) ms_per_variant[v] = ctx->i8_tune_ms[bucket][v];
    return BRPA_OK;
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
        BRPA_TRY(copy_sync(ctx, eo.data(), eocc, 8 * nocc, hipMemcpyDeviceToHost));
        BRPA_TRY(copy_sync(ctx, ev.data(), evir, 8 * nvir, hipMemcpyDeviceToHost));
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
    void *hv = nullptr;
    BRPA_TRY(rocblas_handle_of(ctx, &hv));
    rocblas_handle h = static_cast<rocblas_handle>(hv);
    Arena count;
    count.take<double>(K);
    count.take<double>(K * N);
    if (!rdev) count.take<double>(K * N);
    if (!pdev) count.take<double>(N * N);
    Arena ar;
    BRPA_TRY(ws_reserve(ctx, count.off, &ar));
    double *dc = ar.take<double>(K);
    double *dW = ar.take<double>(K * N);
    const double *dR = rdev ? crg : ar.take<double>(K * N);
    double *dP = pdev ? gpiug : ar.take<double>(N * N);
    BRPA_TRY(copy_sync(ctx, dc, c.data(), 8 * K, hipMemcpyHostToDevice));
    if (!rdev) BRPA_TRY(to_device(ctx, const_cast<double *>(dR), crg, 8 * K * N));
    if (!pdev) BRPA_TRY(to_device(ctx, dP, gpiug, 8 * N * N));
    BRPA_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    // the reference's own shape (rpa_core_MO.f90:398-402): DCRG = CRG diag(W),
    // then GPiUG += CRG DCRG^T as one plain dgemm (column-major N x K operands)
    BRPA_TRY(launch_col_scale(ctx, dW, dR, dc, N, K));
    const double one = 1.0;
    const rocblas_status st =
        rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_transpose, (rocblas_int)N,
                      (rocblas_int)N, (rocblas_int)K, &one, dR, (rocblas_int)N, dW, (rocblas_int)N,
                      &one, dP, (rocblas_int)N);
    if (st != rocblas_status_success)
        return brpa_fail(ctx, BRPA_ERR_SOLVER, "rocblas_dgemm failed: rocblas status %d", (int)st);
    BRPA_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    if (!pdev) BRPA_TRY(from_device(ctx, gpiug, dP, 8 * N * N));
    BRPA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    float ms = 0.f;
    BRPA_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    ctx->last_kernel_ms = ms;
    return BRPA_OK;
}

}  // extern "C"

int launch_unpack_lower(brpa_ctx *, const double *AP, int64_t n, int64_t F, double *A);

// E_RPA from stored Pi(u): full (F x N x N) or lower-packed (F x N(N+1)/2) input
// Device buffers of E_RPA from stored Pi(u) for nf frequencies of order N,
// laid out by `ar` (count pass: ar.base == nullptr).  pk_host: a packed host
// input is staged in the buffer as well.
struct StoredBufs {
    FragDev hd{};
    FragDev *d_fr = nullptr;
    int64_t *d_pre = nullptr;
    double *d_pk = nullptr;
    void layout(Arena &ar, int64_t N, int64_t nf, int nsteps, size_t npre, bool pk_host) {
        d_fr = ar.take<FragDev>(1);
        d_pre = ar.take<int64_t>(npre);
        d_pk = pk_host ? ar.take<double>(nf * (N * (N + 1) / 2)) : nullptr;
        hd.K = 0;
        hd.N = (int)N;
        hd.F = (int)nf;
        hd.nsteps = nsteps;
        hd.Pi = ar.take<double>(nf * N * N);
        hd.diag = ar.take<double>(nf * N);
        hd.ldpart = ar.take<double>(nf * nsteps * 2);
        hd.weights = ar.take<double>(nf);
        hd.eu = ar.take<double>(nf);
        hd.energy = ar.take<double>(1);
        hd.status = ar.take<int>(nf);
        hd.linv = ar.take<double>(nf * CHOL_NB * CHOL_NB);
    }
};

// Work-group prefix arrays [setdiag, pdiag_0..S-1, trsm_0..S-1, update_0..S-1]
// of the blocked Cholesky of nf matrices of order N.
static std::vector<int64_t> stored_prefix(int64_t N, int64_t nf, int nsteps) {
    std::vector<int64_t> pre((size_t)(1 + 3 * nsteps) * 2, 0);
    pre[1] = cdiv(nf * N, 256);
    for (int j = 0; j < nsteps; ++j) {
        int64_t j0 = (int64_t)j * CHOL_NB;
        pre[2 * (1 + j) + 1] = (N > j0) ? nf : 0;
        pre[2 * (1 + nsteps + j) + 1] = (N > j0 + CHOL_NB) ? nf * (cdiv(N - j0, CHOL_NB) - 1) : 0;
        pre[2 * (1 + 2 * nsteps + j) + 1] = nf * chol_upd_items(N, j);
    }
    return pre;
}

// E_RPA of nf stored Pi(u) (piu: nf full or packed N x N matrices, host or
// device) in the buffers `b` (laid out by StoredBufs::layout): the chunk's
// energy goes to *ecrpa (host or device), its E(u) to eu (host or device, may
// be NULL).
static int ecrpa_stored_run(brpa_ctx *ctx, StoredBufs &b, const std::vector<int64_t> &pre,
                            double *ecrpa, const double *piu, bool packed, const double *weights,
                            int64_t nf, int64_t N, double *eu) {
    const int nsteps = b.hd.nsteps;
    const int64_t F = nf;
    FragDev &hd = b.hd;
    if (packed) {
        // rpa_Unpack (rpa.f90:390-404) on the device, into the Cholesky's buffer
        const double *src = piu;
        if (b.d_pk) {
            BRPA_TRY(copy_sync(ctx, b.d_pk, piu, 8 * F * (N * (N + 1) / 2), hipMemcpyHostToDevice));
            src = b.d_pk;
        }
        BRPA_TRY(launch_unpack_lower(ctx, src, N, F, hd.Pi));
    } else {
        BRPA_TRY(to_device(ctx, hd.Pi, piu, 8 * F * N * N));
    }
    BRPA_TRY(to_device(ctx, const_cast<double *>(hd.weights), weights, 8 * F));
    BRPA_TRY(launch_fill_int(ctx, hd.status, F, 0));
    // completed copies: the descriptor and prefix sums live on the caller's stack
    BRPA_TRY(copy_sync(ctx, b.d_fr, &hd, sizeof(FragDev), hipMemcpyHostToDevice));
    BRPA_TRY(copy_sync(ctx, b.d_pre, pre.data(), 8 * pre.size(), hipMemcpyHostToDevice));
    const FragDev *d_fr = b.d_fr;
    const int64_t *d_pre = b.d_pre;
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
    BRPA_TRY(copy_sync(ctx, st.data(), hd.status, sizeof(int) * F, hipMemcpyDeviceToHost));
    float ms = 0.f;
    BRPA_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    ctx->last_kernel_ms = ms;
    for (int64_t w = 0; w < F; ++w)
        if (st[w])
            return brpa_fail(ctx, BRPA_ERR_NOT_SPD,
                             "frequency %lld: 1 + Pi(u) is not positive definite", (long long)w);
    return BRPA_OK;
}

static int ecrpa_stored(brpa_ctx *ctx, double *ecrpa, const double *piu, bool packed,
                        const double *weights, int64_t nfreqs, int64_t naux, double *eu,
                        const char *who) {
    if (!ctx) return BRPA_ERR_ARG;
    if (!ecrpa || !piu || !weights || nfreqs <= 0 || naux <= 0 || naux > 65536)
        return brpa_fail(ctx, BRPA_ERR_ARG, "%s: bad arguments", who);
    BRPA_HIP(ctx, hipSetDevice(ctx->device));
    const int64_t N = naux, F = nfreqs;
    const int nsteps = (int)cdiv(N, CHOL_NB);
    const std::vector<int64_t> pre = stored_prefix(N, F, nsteps);
    const bool pk_host = packed && !is_device_ptr(piu);
    StoredBufs b;
    Arena count;
    b.layout(count, N, F, nsteps, pre.size(), pk_host);
    Arena ar;
    BRPA_TRY(ws_reserve(ctx, count.off, &ar));
    b.layout(ar, N, F, nsteps, pre.size(), pk_host);
    return ecrpa_stored_run(ctx, b, pre, ecrpa, piu, packed, weights, F, N, eu);
}

// ---------------------------------------------------------------------------
// Allocation-free E_RPA from stored Pi(u) (brpa_ecrpa_analytic_noalloc).
//
// Its device buffers live in `scratch`, one file-scope array per GPU that the
// library holds until the process ends.  In host code the name `scratch` is a
// host-side stand-in, not a device address: the device address is looked up
// once per device (noalloc_setup, called when a context is created on it) and
// only that address is handed to kernels and copies.  Every call on one GPU
// shares that GPU's array: calls that overlap on one GPU race.
namespace {
constexpr size_t kNoallocBytes = (size_t)BRPA_NOALLOC_SCRATCH_MB << 20;
constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mu;
char *g_noalloc_ptr[kMaxDevices] = {};
}  // namespace

static __device__ __align__(256) char scratch[kNoallocBytes];

int noalloc_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return BRPA_ERR_ARG;
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (g_noalloc_ptr[device]) return BRPA_OK;
    // the caller has made `device` current
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
        (void)hipGetLastError();
        return BRPA_ERR_HIP;
    }
    g_noalloc_ptr[device] = static_cast<char *>(p);
    return BRPA_OK;
}

static int ecrpa_stored_noalloc(brpa_ctx *ctx, double *ecrpa, const double *piu, bool packed,
                                const double *weights, int64_t nfreqs, int64_t naux, double *eu,
                                const char *who) {
    if (!ctx) return BRPA_ERR_ARG;
    if (!ecrpa || !piu || !weights || nfreqs <= 0 || naux <= 0 || naux > 65536)
        return brpa_fail(ctx, BRPA_ERR_ARG, "%s: bad arguments", who);
    char *base = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_noalloc_mu);
        if (ctx->device >= 0 && ctx->device < kMaxDevices) base = g_noalloc_ptr[ctx->device];
    }
    if (!base)
        return brpa_fail(ctx, BRPA_ERR_HIP, "%s: the scratch array of device %d was not set up",
                         who, ctx->device);
    BRPA_HIP(ctx, hipSetDevice(ctx->device));
    const int64_t N = naux, F = nfreqs;
    const int nsteps = (int)cdiv(N, CHOL_NB);
    const bool pk_host = packed && !is_device_ptr(piu);
    // the largest frequency chunk whose buffers fit the scratch array
    auto bytes_for = [&](int64_t nf) {
        StoredBufs t;
        Arena count;
        t.layout(count, N, nf, nsteps, (size_t)(1 + 3 * nsteps) * 2, pk_host);
        return count.off;
    };
    if (bytes_for(1) > kNoallocBytes)
        return brpa_fail(ctx, BRPA_ERR_ARG,
                         "%s: NAux = %lld needs %zu bytes of scratch per frequency, beyond the "
                         "%zu-byte scratch array of the allocation-free entry point",
                         who, (long long)N, bytes_for(1), kNoallocBytes);
    int64_t fc = F;
    while (fc > 1 && bytes_for(fc) > kNoallocBytes) fc = std::max<int64_t>(1, fc / 2);
    while (fc < F && bytes_for(fc + 1) <= kNoallocBytes) ++fc;
    const int64_t msize = packed ? N * (N + 1) / 2 : N * N;
    long double etot = 0.0L;
    double ms_tot = 0.0;
    for (int64_t f0 = 0; f0 < F; f0 += fc) {
        const int64_t nf = std::min(fc, F - f0);
        const std::vector<int64_t> pre = stored_prefix(N, nf, nsteps);
        StoredBufs b;
        Arena ar;
        ar.base = base;
        b.layout(ar, N, nf, nsteps, pre.size(), pk_host);
        // the chunk's energy comes back to the host (the run ends with a
        // completed copy, so e is final on return)
        double e = 0.0;
        BRPA_TRY(ecrpa_stored_run(ctx, b, pre, &e, piu + f0 * msize, packed, weights + f0, nf, N,
                                  eu ? eu + f0 : nullptr));
        etot += (long double)e;
        ms_tot += ctx->last_kernel_ms;
    }
    const double et = (double)etot;
    if (is_device_ptr(ecrpa))
        BRPA_TRY(copy_sync(ctx, ecrpa, &et, 8, hipMemcpyHostToDevice));
    else
        *ecrpa = et;
    ctx->last_kernel_ms = ms_tot;
    return BRPA_OK;
}

extern "C" {

int brpa_ecrpa_analytic(brpa_ctx *ctx, double *ecrpa, const double *piu, const double *freqs,
                        const double *weights, int64_t nfreqs, int64_t naux, double *eu) {
    (void)freqs;   // printed by the reference only (rpa_CC.f90:462-463)
    return ecrpa_stored(ctx, ecrpa, piu, false, weights, nfreqs, naux, eu, "brpa_ecrpa_analytic");
}

int brpa_ecrpa_packed(brpa_ctx *ctx, double *ecrpa, const double *piu_packed, const double *freqs,
                      const double *weights, int64_t nfreqs, int64_t naux, double *eu) {
    (void)freqs;   // printed by the reference only (rpa.f90:1368-1372)
    return ecrpa_stored(ctx, ecrpa, piu_packed, true, weights, nfreqs, naux, eu,
                        "brpa_ecrpa_packed");
}

int brpa_ecrpa_analytic_noalloc(brpa_ctx *ctx, double *ecrpa, const double *piu,
                                const double *freqs, const double *weights, int64_t nfreqs,
                                int64_t naux, double *eu) {
    (void)freqs;
    return ecrpa_stored_noalloc(ctx, ecrpa, piu, false, weights, nfreqs, naux, eu,
                                "brpa_ecrpa_analytic_noalloc");
}

}  // extern "C"
