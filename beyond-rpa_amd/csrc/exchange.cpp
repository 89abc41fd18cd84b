// Exchange correction of many fragments: factorised direct-ring amplitudes
// (rpa_CC_T2) followed by the SOSEX contraction (rpa_CC_Exchange_SOSEX) per
// fragment.
//
// The amplitudes stay in the worker context's device buffers between the two
// steps (no copy-out).  With nworkers > 1 the fragments are spread over that
// many worker contexts (own stream, rocBLAS handle and workspace each), each
// driven by its own host thread and taking the next fragment when it is done,
// so that the latency-bound stages of one fragment (Cholesky steps,
// eigensolver, small GEMMs) overlap with the work of the others.  Every host
// input of a worker goes to the device with a completed copy before the
// worker's kernels use it.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "internal.h"

namespace {
constexpr int MAX_WORKERS = 8;
}

extern "C" {

int brpa_exchange_batch(brpa_ctx *ctx, int precision, int64_t nfrag, const brpa_frag_t *frags,
                        double lambda, double cutoff, double guess_factor, uint64_t seed,
                        double alpha, int s2_operator, int nworkers, double *ec_exchange,
                        int64_t *nvecs_t2) {
    if (!ctx) return BRPA_ERR_ARG;
    // guess_factor: GuessNVecsT2 (3 in the reference); bounded so that the
    // guess count ceil(guess_factor * NAux) is a representable integer
    if (nfrag < 0 || (nfrag > 0 && (!frags || !ec_exchange)) ||
        !(guess_factor > 0.0 && guess_factor <= 1.0e6))
        return brpa_fail(ctx, BRPA_ERR_ARG, "brpa_exchange_batch: bad arguments");
    if (nworkers <= 0) nworkers = 1;
    if (nworkers > MAX_WORKERS)
        return brpa_fail(ctx, BRPA_ERR_ARG, "brpa_exchange_batch: nworkers = %d exceeds %d",
                         nworkers, MAX_WORKERS);
    if (nfrag == 0) return BRPA_OK;
    nworkers = (int)std::min<int64_t>(nworkers, nfrag);
    BRPA_HIP(ctx, hipSetDevice(ctx->device));
    BRPA_HIP(ctx, hipStreamSynchronize(ctx->stream));   // inputs written on the caller's stream
    // With four workers each worker's stream runs on its own quarter of the CUs,
    // so a worker's short kernels (the Cholesky steps of its fragment) never wait
    // behind the grid of another worker's GEMM.  Measured on the 256-fragment
    // batch (MI355X, 256 CUs): 4 workers on CU quarters 13.6 s, 4 or 2 workers
    // on the whole chip 14.2 s, 2 workers on CU halves 14.7 s; 3, 6 and 8
    // workers on CU shares were slower still.  Other counts share the chip.
    // Workers made for another worker count are replaced.
    const int nshares = nworkers == 4 ? 4 : 1;
    if ((int)ctx->workers.size() != nworkers) {
        for (brpa_ctx *w : ctx->workers) brpa_ctx_destroy(w);
        ctx->workers.clear();
    }
    while ((int)ctx->workers.size() < nworkers) {
        brpa_ctx *w = nullptr;
        const int part = nshares > 1 ? (int)ctx->workers.size() : 0;
        int rc = ctx_create_cu_share(ctx->device, part, nshares, &w);
        if (rc != BRPA_OK) return brpa_fail(ctx, rc, "brpa_exchange_batch: worker context");
        ctx->workers.push_back(w);
    }
    // BRPA_EXCHANGE_FORM=pairs keeps the reference's order (amplitude eigenpairs or
    // factored core, then the pair contraction) for every fragment
    const char *form = std::getenv("BRPA_EXCHANGE_FORM");
    const bool pairs = form && std::string(form) == "pairs";
    // fragments are handed out largest first (modelled cost o^2 v^2 (F N + 2 NG),
    // the I and I Omega products), so that the last fragments in flight are
    // small ones and no worker's share of the chip idles long at the end
    std::vector<int64_t> order(nfrag);
    {
        std::vector<double> cost(nfrag);
        for (int64_t f = 0; f < nfrag; ++f) {
            const double K = (double)frags[f].nocc * (double)frags[f].nvir;
            const double N = (double)frags[f].naux;
            cost[f] = K * K * ((double)frags[f].nfreqs * N + 2.0 * guess_factor * N);
            order[f] = f;
        }
        std::stable_sort(order.begin(), order.end(),
                         [&](int64_t a, int64_t b) { return cost[a] > cost[b]; });
    }
    std::atomic<int64_t> next{0};
    std::atomic<bool> failed{false};
    std::mutex mu;
    int first_rc = BRPA_OK;
    int64_t first_f = -1;
    std::string first_msg;
    auto run = [&](brpa_ctx *w) {
        if (hipSetDevice(w->device) != hipSuccess) {
            std::lock_guard<std::mutex> lk(mu);
            if (first_rc == BRPA_OK) {
                first_rc = BRPA_ERR_HIP;
                first_msg = "hipSetDevice failed in a worker";
            }
            failed = true;
            return;
        }
        // per-stage device times summed over this worker's fragments
        stage_reset(w);
        w->stage_keep = true;
        while (!failed) {
            const int64_t q = next.fetch_add(1);
            if (q >= nfrag) break;
            const int64_t f = order[q];
            const brpa_frag_t &fr = frags[f];
            int64_t nt = 0;
            double ex = 0.0;
            int rc;
            if (!fr.R || !fr.eocc || !fr.evir || !fr.freqs || !fr.weights || fr.nocc <= 0 ||
                fr.nvir <= 0 || fr.naux <= 0 || fr.nfreqs <= 0 || fr.nocc > INT32_MAX ||
                fr.nvir > INT32_MAX || fr.naux > 65536 || fr.nfreqs > INT32_MAX) {
                // checked before the guess count and the weights are read
                rc = brpa_fail(w, BRPA_ERR_ARG, "bad fragment arguments (nocc %lld, nvir %lld, "
                               "naux %lld, nfreqs %lld)", (long long)fr.nocc, (long long)fr.nvir,
                               (long long)fr.naux, (long long)fr.nfreqs);
            } else {
                const int64_t ng = (int64_t)std::ceil(guess_factor * (double)fr.naux);
                // the explicit-space form unless the S2 rescaling needs the eigenpairs,
                // the caller asked for the pair order, the fragment is beyond the
                // explicit form's sizes, or a quadrature weight is not positive
                const bool xspace = !s2_operator && !pairs &&
                                    xspace_fits(fr.nocc * fr.nvir, fr.naux, ng, fr.nfreqs) &&
                                    xspace_applicable(fr.weights, fr.nfreqs, lambda);
                if (xspace) {
                    // energy only, in the explicit ov space (FP64)
                    rc = exchange_xspace(w, fr.R, fr.naux, fr.eocc, fr.nocc, fr.evir, fr.nvir,
                                         fr.freqs, fr.weights, fr.nfreqs, lambda, cutoff, ng,
                                         seed + (uint64_t)f, alpha, &ex, &nt);
                } else {
                    // the reference's order: amplitudes (the factored core unless the S2
                    // rescaling needs the eigenpairs), then the pair contraction
                    rc = t2_build(w, precision, fr.R, fr.naux, fr.eocc, fr.nocc, fr.evir,
                                  fr.nvir, fr.freqs, fr.weights, fr.nfreqs, lambda, cutoff, ng,
                                  seed + (uint64_t)f, &nt, !s2_operator);
                    if (rc == BRPA_OK && nt > 0)
                        rc = brpa_sosex(w, precision, &ex, alpha, w->t2_V, w->t2_A, fr.R,
                                        fr.nocc, fr.nvir, fr.naux, nt, s2_operator);
                }
            }
            if (rc != BRPA_OK) {
                // the lowest failing fragment among those that ran is reported
                std::lock_guard<std::mutex> lk(mu);
                if (first_rc == BRPA_OK || f < first_f) {
                    first_rc = rc;
                    first_f = f;
                    first_msg = w->err;
                }
                failed = true;
                break;
            }
            ec_exchange[f] = ex;
            if (nvecs_t2) nvecs_t2[f] = nt;
        }
        w->stage_keep = false;
    };
    if (nworkers == 1) {
        run(ctx->workers[0]);
    } else {
        std::vector<std::thread> th;
        th.reserve(nworkers);
        for (int i = 0; i < nworkers; ++i) th.emplace_back(run, ctx->workers[i]);
        for (auto &t : th) t.join();
    }
    if (first_rc != BRPA_OK)
        return brpa_fail(ctx, first_rc, "brpa_exchange_batch: fragment %lld: %s",
                         (long long)first_f, first_msg.c_str());
    for (int k = 0; k < BRPA_NKERNELS; ++k) {
        ctx->kst_ms[k] = ctx->kst_flops[k] = 0.0;
        ctx->kst_launches[k] = 0;
        for (int i = 0; i < nworkers; ++i) {
            ctx->kst_ms[k] += ctx->workers[i]->kst_ms[k];
            ctx->kst_flops[k] += ctx->workers[i]->kst_flops[k];
            ctx->kst_launches[k] += ctx->workers[i]->kst_launches[k];
        }
    }
    for (int st = 0; st < BRPA_NSTAGES; ++st) {
        ctx->stage_ms[st] = 0.0;
        ctx->stage_launches[st] = 0;
        for (int i = 0; i < nworkers; ++i) {
            ctx->stage_ms[st] += ctx->workers[i]->stage_ms[st];
            ctx->stage_launches[st] += ctx->workers[i]->stage_launches[st];
        }
    }
    return BRPA_OK;
}

}  // extern "C"
