// Exchange correction of many fragments: factorised direct-ring amplitudes
// (rpa_CC_T2) followed by the SOSEX contraction (rpa_CC_Exchange_SOSEX) per
// fragment, with several fragments in flight at once.
//
// The amplitudes stay in the worker context's device buffers between the two
// steps (no copy-out).  Fragments run one after another on one worker: with
// several workers on host threads (own stream, rocBLAS handle and workspace
// each) the rocBLAS/rocSOLVER calls of the amplitude build returned wrong
// results here, while the library's own kernels stayed bit-exact under the
// same concurrency, so nworkers > 1 is rejected until that is resolved.
#include <cmath>

#include "internal.h"

extern "C" {

int brpa_exchange_batch(brpa_ctx *ctx, int precision, int64_t nfrag, const brpa_frag_t *frags,
                        double lambda, double cutoff, double guess_factor, uint64_t seed,
                        double alpha, int s2_operator, int nworkers, double *ec_exchange,
                        int64_t *nvecs_t2) {
    if (!ctx) return BRPA_ERR_ARG;
    if (nfrag < 0 || (nfrag > 0 && (!frags || !ec_exchange)) || !(guess_factor > 0.0))
        return brpa_fail(ctx, BRPA_ERR_ARG, "brpa_exchange_batch: bad arguments");
    if (nfrag == 0) return BRPA_OK;
    if (nworkers <= 0) nworkers = 1;
    if (nworkers > 1)
        return brpa_fail(ctx, BRPA_ERR_ARG,
                         "brpa_exchange_batch: nworkers > 1 is disabled (concurrent rocBLAS/"
                         "rocSOLVER calls on separate handles gave wrong results)");
    BRPA_HIP(ctx, hipSetDevice(ctx->device));
    BRPA_HIP(ctx, hipStreamSynchronize(ctx->stream));   // inputs written on the caller's stream
    if (ctx->workers.empty()) {
        brpa_ctx *w = nullptr;
        int rc = brpa_ctx_create(ctx->device, &w);
        if (rc != BRPA_OK) return brpa_fail(ctx, rc, "brpa_exchange_batch: worker context");
        ctx->workers.push_back(w);
    }
    brpa_ctx *w = ctx->workers[0];
    // per-stage device times summed over the fragments, reported on ctx
    stage_reset(w);
    w->stage_keep = true;
    struct KeepOff {
        brpa_ctx *w;
        ~KeepOff() { w->stage_keep = false; }
    } keep_off{w};
    for (int64_t f = 0; f < nfrag; ++f) {
        const brpa_frag_t &fr = frags[f];
        const int64_t ng = (int64_t)std::ceil(guess_factor * (double)fr.naux);
        int64_t nt = 0;
        double ex = 0.0;
        // energy only: the factored amplitude core (no eigensolve) unless the S2
        // rescaling needs the amplitudes themselves
        int rc = t2_build(w, precision, fr.R, fr.naux, fr.eocc, fr.nocc, fr.evir, fr.nvir,
                          fr.freqs, fr.weights, fr.nfreqs, lambda, cutoff, ng, seed + (uint64_t)f,
                          &nt, !s2_operator);
        if (rc == BRPA_OK && nt > 0)
            rc = brpa_sosex(w, precision, &ex, alpha, w->t2_V, w->t2_A, fr.R, fr.nocc, fr.nvir,
                            fr.naux, nt, s2_operator);
        if (rc != BRPA_OK)
            return brpa_fail(ctx, rc, "brpa_exchange_batch: fragment %lld: %s", (long long)f,
                             w->err.c_str());
        ec_exchange[f] = ex;
        if (nvecs_t2) nvecs_t2[f] = nt;
    }
    for (int st = 0; st < BRPA_NSTAGES; ++st) {
        ctx->stage_ms[st] = w->stage_ms[st];
        ctx->stage_launches[st] = w->stage_launches[st];
    }
    return BRPA_OK;
}

}  // extern "C"
