/*
 * brpa.h -- C ABI of the MI355X (gfx950) correlation-energy hot path.
 *
 * Drop-in for the MO-basis direct-RPA energy and SOSEX correction of
 * modrzejewski/beyond-rpa.  Every entry point keeps the meaning of the
 * reference routine it replaces (cited per function, paths relative to the
 * reference's src/).  The reference's Fortran arrays are column-major with
 * default 64-bit INTEGER; the C layouts below are the same bytes:
 *
 *   Fortran CRG(P, ai) / Rkai(P, a, i)  ==  C row-major R[ai][P], ai = i*nvir + a
 *   Fortran GPiUG(P, Q), PiU(P, Q, u)   ==  C row-major Pi[u][Q][P] (symmetric)
 *   Fortran V(mu, ai) / Vkai(mu, a, i)  ==  C row-major V[ai][mu]
 *
 * Pointers may be host or device (HBM) memory; the library detects device
 * pointers and otherwise stages through its own device workspace.  Calls are
 * synchronous: results are visible to the host on return.  A context owns one
 * HIP stream on one device and must not be shared between threads.
 *
 * Every function returns BRPA_OK (0) or a BRPA_ERR_* code; the message of the
 * last failure is brpa_last_error(ctx).  The reference has no return codes:
 * it prints a message and executes `error stop`, which is what the Fortran
 * shim does with a non-zero code.
 */
#ifndef BRPA_H
#define BRPA_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct brpa_ctx brpa_ctx;

enum {
    BRPA_OK = 0,
    BRPA_ERR_ARG = 1,        /* bad size / null pointer / unsupported option */
    BRPA_ERR_HIP = 2,        /* HIP runtime error (message has the detail) */
    BRPA_ERR_NOT_SPD = 3,    /* Cholesky breakdown of 1 + Pi (reference: DGEQRF info /= 0) */
    BRPA_ERR_GAP = 4,        /* non-positive HOMO-LUMO gap (OrbDiffHist.f90:77-80) */
    BRPA_ERR_NONFINITE = 5,  /* NaN/Inf in an input */
    BRPA_ERR_OOM = 6,        /* device allocation failed */
    BRPA_ERR_GRID = 7,       /* frequency quadrature did not converge (OptFreqQuad.f90:68-71) */
    BRPA_ERR_SOLVER = 8,     /* eigensolver failure (real_linalg.f90:1527-1531) */
    BRPA_ERR_INTERNAL = 9    /* a self-check of the library failed (e.g. kernel variants
                                that must agree bit for bit did not) */
};

/* Arithmetic of the Pi and amplitude contractions.  BRPA_FP64 is the drop-in. */
enum {
    BRPA_FP64 = 0,
    BRPA_SPLIT_BF16 = 1,     /* 4 fixed-point bf16 slices per value, the 10 slice
                                products of level <= 3 on bf16 MFMA with FP32
                                accumulation (levels 0 and 1 exact per 128-row block),
                                FP64 between blocks; exact FP64 diagonal and trace;
                                FP64 Cholesky log-det and quadrature */
    BRPA_SPLIT_I8 = 2        /* 4 fixed-point int8 slices per value (scale per column
                                over the whole K), the 10 slice products of level <= 3
                                on int8 MFMA with exact int32 accumulation (K <= 65536),
                                one FP64 combination per element; exact FP64 diagonal
                                and trace; FP64 Cholesky log-det and quadrature */
};

/* ---- context ------------------------------------------------------------ */
int brpa_ctx_create(int device, brpa_ctx **ctx);
int brpa_ctx_destroy(brpa_ctx *ctx);
const char *brpa_last_error(const brpa_ctx *ctx);
/* Use an external HIP stream (e.g. torch's current stream); NULL restores the
 * context's own stream.  The caller keeps ownership of the stream. */
int brpa_ctx_set_stream(brpa_ctx *ctx, void *hip_stream);
/* Release the device workspace (it grows on demand and is kept otherwise). */
int brpa_ctx_trim(brpa_ctx *ctx);
/* hipGraph replay of the batched E_RPA launch sequence (no counterpart in the
 * reference, which has no GPU path).  When enabled and every fragment's R is
 * device-resident, brpa_ecrpa_batch captures the whole launch sequence of a
 * workspace chunk (uploads from pinned staging, all kernels, result download)
 * once per (precision, fragment shapes, workspace address) and replays it on
 * later calls; inputs are refreshed through the staging buffer every call.
 * Per-stage timings are not recorded in replayed calls. */
int brpa_ctx_set_graphs(brpa_ctx *ctx, int enable);
/* Graphs captured and replayed by the last brpa_ecrpa_batch call. */
int brpa_last_graph_stats(const brpa_ctx *ctx, int *captures, int *replays);
/* Device time in milliseconds of the last hot-path kernel set of the last call
 * (HIP events recorded on the context's stream around the kernels only). */
double brpa_last_kernel_ms(const brpa_ctx *ctx);

/* Device time (ms) of one stage of the last E_RPA call (stages 0-3), of
 * the last amplitude / SOSEX / exchange-batch call (stages 4-9) or of the last
 * THC Pi(u) call (stages 10-11) or MBPT3 call (stage 12, with 9), summed over
 * its launches (the exchange stages include the device idle time inside the
 * stage), and the number of launches.  Stages: */
enum {
    BRPA_STAGE_PI = 0,     /* Pi(u) SYRK kernel (FP64, split-BF16 or split-int8 MFMA) */
    BRPA_STAGE_SPLIT = 1,  /* slicing passes of the split forms (int8: column scales,
                              exact diagonal and slices) */
    BRPA_STAGE_CHOL = 2,   /* blocked Cholesky log-det (panel + trailing updates) */
    BRPA_STAGE_AUX = 3,    /* Delta table, exact diagonal, quadrature */
    BRPA_STAGE_T2_PI = 4,     /* amplitudes: Pi(u_k), the Cholesky factors of 1 + l Pi and
                                 R U_k^-1 */
    BRPA_STAGE_T2_IOMEGA = 5, /* amplitudes: I * Omega over all frequencies */
    BRPA_STAGE_T2_BASIS = 6,  /* amplitudes: orthonormal basis T (Cholesky-QR or eigensolve) */
    BRPA_STAGE_T2_TIT = 7,    /* amplitudes: T^T I T over all frequencies */
    BRPA_STAGE_T2_CORE = 8,   /* amplitudes: factored core or eigenpairs, and V */
    BRPA_STAGE_SOSEX = 9,     /* SOSEX contraction (slicing and pair products) */
    BRPA_STAGE_THC_R = 10,    /* THC producer of R (brpa_thc_piu) */
    BRPA_STAGE_THC_PI = 11,   /* Pi(u) of all frequencies from that R (brpa_thc_piu) */
    BRPA_STAGE_MBPT3 = 12,    /* MBPT3-1: U, D, Lyapunov solve and factor of X (its final
                                 contraction is BRPA_STAGE_SOSEX); MBPT3-2: all products */
    BRPA_STAGE_T2_EIG = 13,   /* amplitudes: the basis eigensolve, inside BRPA_STAGE_T2_BASIS,
                                 when the Gram matrix fails the Cholesky test (its count is
                                 the number of amplitude builds that needed it) */
    BRPA_STAGE_T2_DROP = 14,  /* explicit-space exchange: the dropped-subspace basis (inverse
                                 subspace iteration for the Gram eigenvalues at or below the
                                 cutoff), inside BRPA_STAGE_T2_BASIS (its count is the number
                                 of fragments that took it) */
    BRPA_NSTAGES = 15
};
/* For BRPA_STAGE_MBPT3 after brpa_mbpt3_1 the count is the number of
 * sign-iteration steps of the Lyapunov solve. */
double brpa_last_stage_ms(const brpa_ctx *ctx, int stage);
int brpa_last_stage_launches(const brpa_ctx *ctx, int stage);
/* Split-int8 GEMM launches (the amplitude contractions of the split form) of the
 * last amplitude or exchange-batch call: summed device time (ms, HIP events
 * around each launch), algorithmic FP64-equivalent flop (2 m n k per launch)
 * and the launch count.  No reference counterpart (measurement). */
int brpa_last_gemm_i8_stats(const brpa_ctx *ctx, double *ms, double *flops, int64_t *launches);
/* The same for every measured kernel: BRPA_KERNEL_GEMM_I8 (as above),
 * BRPA_KERNEL_GEMM_F64 (the FP64 MFMA GEMM of the explicit-space exchange and
 * brpa_dgemm / brpa_dsyrk, 2 m n k flop per launch, m n k per launch on the
 * triangle of a SYRK), BRPA_KERNEL_SOSEX_DOT (the contraction of the explicit
 * (ai|bj) = R R^T, itself a GEMM launch, with the occupied-swapped amplitude
 * matrix: (ov)(ov+1) flop, one multiply-add per element of the triangle),
 * BRPA_KERNEL_TRSM_F64 (the
 * triangular solves, n^2 nrhs flop), BRPA_KERNEL_TRINV_F64 (the diagonal
 * blocks of the batched triangular inverses, 64^3/3 flop per block; their
 * off-diagonal blocks are GEMM launches).  No reference counterpart
 * (measurement). */
enum {
    BRPA_KERNEL_GEMM_I8 = 0,
    BRPA_KERNEL_GEMM_F64 = 1,
    BRPA_KERNEL_SOSEX_DOT = 2,
    BRPA_KERNEL_TRSM_F64 = 3,
    BRPA_KERNEL_TRINV_F64 = 4,
    BRPA_NKERNELS = 5
};
int brpa_last_kernel_stats(const brpa_ctx *ctx, int kernel, double *ms, double *flops,
                           int64_t *launches);

/* ---- frequency grid (host code) ----------------------------------------- */
#define BRPA_HISTOGRAM_NBINS 100      /* RPA_HISTOGRAM_NBINS, rpa/rpa_definitions.f90:113 */
#define BRPA_CORE_ORB_THRESH (-3.0)   /* default CoreOrbThresh (Ha), rpa/rpa_definitions.f90:301 */

/* Gap histogram of one spin of one system.  Replaces rpa_DaiHistogram
 * (rpa/OrbDiffHist.f90:42-123) with rpa_NCore (126-146): eocc holds all nocc
 * occupied energies of the spin, ascending, core included; the leading ones
 * below core_orb_thresh are frozen.  The range is [evir[0] - eocc[nocc-1],
 * min(dai_max_thresh, evir[nvir-1] - eocc[ncore])], bin k's value is its left
 * edge, gaps above dai_max_thresh are not counted (dai_max_thresh <= 0: no cap,
 * the reference's huge(ONE) when GridLimitDai is off).  values/weights hold
 * BRPA_HISTOGRAM_NBINS entries; with every occupied orbital frozen they get the
 * reference's special values (weights -1, left out of every error sum). */
int brpa_dai_histogram(const double *eocc, int64_t nocc, const double *evir, int64_t nvir,
                       double core_orb_thresh, double dai_max_thresh, double *values,
                       double *weights);

/* Largest active gap evir[last] - eocc[ncore] over the spins of the whole
 * complex, the cap of its subsystems' histograms.  Replaces rpa_DaiMaxThresh
 * (rpa/OrbDiffHist.f90:10-39).  eocc/evir: the spins' energies one after the
 * other (nocc[s], nvir[s] each). */
int brpa_dai_max_thresh(int64_t nspins, const double *eocc, const int64_t *nocc,
                        const double *evir, const int64_t *nvir, double core_orb_thresh,
                        double *dai_max_thresh);

/* Grid of nfreqs points on nsets histograms (values/dweights: nsets x
 * BRPA_HISTOGRAM_NBINS, the bytes of the reference's daiValues(NBins, NSets)):
 * the 15000-point scan of the Casimir-Polder scale alpha in [0.1, 15] keeping
 * the first strict minimum of sqrt(max over sets of the weighted squared
 * error), then the transformed Gauss-Legendre rule.  Replaces
 * rpa_OptQuadScaling + rpa_FreqQuadError (rpa/OptFreqQuad.f90:75-169) and
 * quad_CasimirPolder (rpa/quadratures.f90:570-643).  info (may be NULL):
 * {alpha, rmsd, max_rel_error}. */
int brpa_opt_quad_scaling(int64_t nfreqs, int64_t nsets, const double *values,
                          const double *dweights, double *freqs, double *weights, double *info);

/* Adaptive grid on nsets histograms: the smallest n in [min_nfreqs,
 * max_nfreqs] meeting both targets (BRPA_ERR_GRID if none does).  Replaces
 * rpa_OptimizeFreqQuad (rpa/OptFreqQuad.f90:12-72).  freqs/weights hold
 * max_nfreqs entries. */
int brpa_optimize_freq_quad(int64_t nsets, const double *values, const double *dweights,
                            double target_error, double target_rel_error, int64_t min_nfreqs,
                            int64_t max_nfreqs, double *freqs, double *weights, int64_t *nfreqs,
                            double *info);

/* Fixed-size grid of one fragment on its own histogram: brpa_dai_histogram
 * (core_orb_thresh, dai_max_thresh as there) then brpa_opt_quad_scaling.
 * eocc: all occupied energies ascending (core included), evir ascending. */
int brpa_freq_grid_ex(const double *eocc, int64_t nocc, const double *evir, int64_t nvir,
                      int64_t nfreqs, double core_orb_thresh, double dai_max_thresh,
                      double *freqs, double *weights, double *info);

/* brpa_freq_grid_ex with the default CoreOrbThresh (BRPA_CORE_ORB_THRESH) and
 * no cap: a fragment computed on its own (rpa_DaiMaxThresh of a single system
 * is its own largest active gap, so the cap changes nothing). */
int brpa_freq_grid(const double *eocc, int64_t nocc, const double *evir, int64_t nvir,
                   int64_t nfreqs, double *freqs, double *weights, double *info);

/* Adaptive grid of one fragment (default CoreOrbThresh, no cap).  Replaces
 * rpa_OptimizeFreqQuad (rpa/OptFreqQuad.f90:12-72) on a single histogram.
 * freqs/weights must hold max_nfreqs entries. */
int brpa_optimize_freq_grid(const double *eocc, int64_t nocc, const double *evir, int64_t nvir,
                            double target_error, double target_rel_error,
                            int64_t min_nfreqs, int64_t max_nfreqs,
                            double *freqs, double *weights, int64_t *nfreqs, double *info);

/* ---- Pi(u) build --------------------------------------------------------- */
/* GPiUG(N,N) += prefactor * CRG * diag(W) * CRG^T, W_ai = d_ai/(d_ai^2+u^2),
 * d_ai = evir[a] - eocc[i], ai = i*nvir + a.  Every entry of GPiUG is updated
 * (one dgemm CRG * DCRG^T, the shape of the reference's real_abT_x call).
 * Replaces rpa_GPiUG_MO (rpa/rpa_core_MO.f90:347-406) for one spin block;
 * prefactor is 4 for closed shells and 2 per spin otherwise (379-383).
 * crg is the Fortran CRG(N, nocc*nvir) (C: R[ai][P]). */
int brpa_gpiug_mo(brpa_ctx *ctx, double *gpiug, const double *crg, int64_t naux,
                  const double *eocc, int64_t nocc, const double *evir, int64_t nvir,
                  double freq, double prefactor);

/* ---- E_RPA from stored Pi(u) ---------------------------------------------- */
/* EcRPA = sum_u w_u/(2 pi) [log det(1 + Pi_u) - tr Pi_u], Pi_u the N x N slices
 * of piu (N*N*nfreqs doubles).  Replaces rpa_CC_EcRPA_Analytic
 * (rpa/rpa_CC.f90:412-471); the log-det is a blocked FP64 Cholesky instead of
 * DGEQRF (1+Pi is SPD), sums in double-double instead of REAL(F128).
 * eu (may be NULL) receives the nfreqs per-frequency terms E(u). */
int brpa_ecrpa_analytic(brpa_ctx *ctx, double *ecrpa, const double *piu,
                        const double *freqs, const double *weights, int64_t nfreqs,
                        int64_t naux, double *eu);

/* Allocation-free form of brpa_ecrpa_analytic.  It has no counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * stream capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * call makes no hipMalloc, hipFree or other allocating call: its device
 * buffers live in one file-scope scratch array of BRPA_NOALLOC_SCRATCH_MB MiB
 * per GPU, which the library holds until the process ends (its device address
 * is looked up when the first context is created on that GPU).  Frequencies
 * are processed in chunks that fit the array; an NAux whose single-frequency
 * buffers (about 8 N^2 bytes + 33 KB) do not fit is rejected with
 * BRPA_ERR_ARG.  Same arguments, results and errors as brpa_ecrpa_analytic
 * (within FP64 rounding when the frequencies need several chunks: the chunk
 * energies are added on the host).  Every call on one GPU shares that GPU's
 * array and leaves its contents behind: calls that overlap on one GPU (from
 * several contexts or threads) race. */
#define BRPA_NOALLOC_SCRATCH_MB 256
int brpa_ecrpa_analytic_noalloc(brpa_ctx *ctx, double *ecrpa, const double *piu,
                                const double *freqs, const double *weights, int64_t nfreqs,
                                int64_t naux, double *eu);

/* ---- fused E_RPA of one fragment ---------------------------------------- */
/* Pi(u_k) build + exact diagonal/trace + Cholesky log-det + quadrature, Pi
 * kept in device memory.  Equivalent to rpa_CC_PiU's frequency loop
 * (rpa/rpa_CC.f90:523-526) followed by rpa_CC_EcRPA_Analytic (785). */
int brpa_ecrpa(brpa_ctx *ctx, int precision, const double *R, int64_t naux,
               const double *eocc, int64_t nocc, const double *evir, int64_t nvir,
               const double *freqs, const double *weights, int64_t nfreqs,
               double *ecrpa, double *eu);

/* ---- batched fragments ---------------------------------------------------- */
typedef struct {
    int64_t nocc, nvir, naux, nfreqs;
    const double *R;        /* (nocc*nvir) x naux, host or device */
    const double *eocc;     /* nocc */
    const double *evir;     /* nvir */
    const double *freqs;    /* nfreqs */
    const double *weights;  /* nfreqs */
} brpa_frag_t;

/* E_RPA of many independent fragments in one set of batched launches (the
 * reference runs subsystems one after another, rpa/rpa_driver.f90:176-223). */
int brpa_ecrpa_batch(brpa_ctx *ctx, int precision, int64_t nfrag, const brpa_frag_t *frags,
                     double *ecrpa);

/* Autotuner (no counterpart in the reference).  Times every tile variant of the
 * split-int8 product kernel (128x128 tiles, the same with one contiguous tile
 * range per XCD, 64x64 tiles) on the given fragments, separately for three
 * NAux buckets (< 640, < 1280, >= 1280), and keeps the fastest per bucket in
 * the context for later calls.  Other precisions have one kernel shape (no-op).
 * Runs the pipeline twice per variant and bucket (warm-up + timed). */
int brpa_autotune(brpa_ctx *ctx, int precision, int64_t nfrag, const brpa_frag_t *frags);
/* Chosen variant of a bucket (-1: not tuned) and the product-kernel time (ms) of
 * each of the 3 variants measured by the last brpa_autotune (0: not measured). */
int brpa_autotune_result(const brpa_ctx *ctx, int bucket, int *variant, double *ms_per_variant);

/* ---- basis of the effective dielectric matrix ------------------------------ */
/* Basis types of rpa_G (rpa/rpa_definitions.f90:92-104). */
enum { BRPA_BASIS_RANDOM = 1, BRPA_BASIS_EIGEN = 2, BRPA_BASIS_FULL_CHOLESKY = 3 };
/* G and NVecs of rpa_G (rpa/rpa.f90:31-173) from RWR = Pi(u) at the lowest
 * frequency (Fortran RWR(N, N), G(N, N); host or device):
 *   BRPA_BASIS_EIGEN (the reference default): G = eigenvectors of RWR, the
 *     largest eigenvalue first; RWR <- diag(eigenvalues);
 *   BRPA_BASIS_RANDOM: G Gaussian (Box-Muller on a SplitMix64 counter hash of
 *     `seed`; the reference draws from its own generator), two subspace
 *     iterations G <- Q(RWR G) (rocSOLVER geqrf + orgqr), RWR <- G^T RWR G;
 *   BRPA_BASIS_FULL_CHOLESKY: G = I, NVecs = N, RWR unchanged.
 * NVecs: scanning p = N..1, the diagonal of the transformed RWR is summed while
 * the sum stays below target_error (extended-precision sum); NVecs is the last
 * p (1-based) added, N when none is (rpa.f90:141-152).  abs_error (may be
 * NULL): that sum, Tr(Pi(exact) - Pi(eff)). */
int brpa_rpa_g(brpa_ctx *ctx, int basis, double *G, double *RWR, int64_t ncholesky,
               double target_error, uint64_t seed, int64_t *nvecs, double *abs_error);

/* ---- packed storage of Pi(u) ----------------------------------------------- */
/* Lower-packed storage of symmetric N x N matrices, LAPACK order (uplo 'L'):
 * column j holds rows i >= j at j*N - j*(j-1)/2 + (i - j), N(N+1)/2 doubles per
 * matrix, the reference's GPiUG(:, u) (rpa/rpa.f90:379-404 rpa_Pack /
 * rpa_Unpack on real_Pack / real_Unpack = dtrttp / dtpttr,
 * linear-algebra/real_linalg.f90:737-773).  Full matrices: nfreqs x N x N,
 * Fortran PiU(N, N, nfreqs), element (i, j) at i + N j.  Host or device.
 * brpa_pack_pi: packed (+)= lower triangle (accumulate = 1 as rpa_Pack, 0
 * overwrites).  brpa_unpack_pi: the full symmetric matrices (rpa_Unpack).
 * brpa_ecrpa_packed: E_RPA from packed Pi(u) as the energy loop of
 * rpa_FreqIntegral_MO (rpa/rpa.f90:1336-1380): unpacked on the device, then the
 * Cholesky log-det of brpa_ecrpa_analytic; eu (may be NULL) receives E(u). */
int brpa_pack_pi(brpa_ctx *ctx, double *packed, const double *piu, int64_t nfreqs, int64_t naux,
                 int accumulate);
int brpa_unpack_pi(brpa_ctx *ctx, double *piu, const double *packed, int64_t nfreqs,
                   int64_t naux);
int brpa_ecrpa_packed(brpa_ctx *ctx, double *ecrpa, const double *piu_packed, const double *freqs,
                      const double *weights, int64_t nfreqs, int64_t naux, double *eu);

/* ---- THC route: Cholesky vectors and Pi(u) from the THC factors --------- */
/* Pi(u_w) (w = 0..nfreqs-1) of one spin block from the tensor-hypercontraction
 * factors, replacing rpa_THC_PiU_BlockAlgo (rpa/rpa_THC.f90:285-339) with
 * rpa_THC_Rkab (373-403) and rpa_THC_ChiRkai (342-370):
 *   R(P, ai) = sum_g Zgk(g, P) Xga(g, a) Xgi(g, i)      (ai = a + nvir*i)
 *   Pi(u_w) (+)= prefactor * R diag(d/(d^2 + u_w^2)) R^T,  d = e_a - e_i.
 * Xga: Fortran Xga(ngrid, nvir) (column a at Xga + ngrid*a), Xgi(ngrid, nocc),
 * Zgk(ngrid, naux); PiU: nfreqs x naux x naux (Fortran PiU(naux, naux, nfreqs));
 * all host or device.  accumulate = 0 overwrites PiU (the reference zeroes it
 * first, rpa_THC.f90:236), 1 adds (the second spin block).  prefactor: 4 for a
 * closed shell, 2 per spin block of an open shell (rpa_THC.f90:254-266).
 * Rkai (may be NULL): receives R as Fortran Rkai(naux, nvir*nocc) == C
 * R[ai][P], the FullRkai = .true. output that the amplitudes read; NULL keeps R
 * inside the library (FullRkai = .false.).  R is built once in HBM by one FP64
 * MFMA kernel that forms Xga(g, a) Xgi(g, i) while loading its operand (no
 * ngrid x ov intermediate); the block size MaxBlockDim of the reference only
 * bounded its memory and has no counterpart.  Device times: stages
 * BRPA_STAGE_THC_R and BRPA_STAGE_THC_PI. */
int brpa_thc_piu(brpa_ctx *ctx, double *PiU, double *Rkai, const double *Xga, const double *Xgi,
                 const double *Zgk, int64_t ngrid, int64_t nocc, int64_t nvir, int64_t naux,
                 const double *eocc, const double *evir, const double *freqs, int64_t nfreqs,
                 double prefactor, int accumulate);

/* ---- THC SOSEX + 2g cumulant terms ---------------------------------------- */
/* The 1b (SOSEX-type) and 2g cumulant energies of the THC route, replacing
 * rpa_CCD_corrections_1b2g (rpa/rpa_CCD_Corrections.f90:79-290):
 *   M_mu(g, h) = sum_{a,i} Y(g, a) U(a, i, mu) X(h, i),
 *   S2g_mu = sum_{g,h} Z(g, h) M_mu(g, h)^2,  S1b_mu = sum_{g,h} Z(g, h) M_mu(g, h) M_mu(h, g),
 *   Ec2g = 2 sum_mu (-2 S2g_mu A_mu^2),  Ec1b = 2 sum_mu (-S1b_mu A_mu)
 * (before the caller's 1/2 on Ec1b, rpa_CCD_Corrections.f90:71, after which
 * it equals SOSEX).  Zgh: Fortran Zgh(ngrid, ngrid), symmetric (Zgk Zgk^T);
 * Xgi(ngrid, nocc), Yga(ngrid, nvir); Uaim: Fortran Uaim(nvir, nocc, NVecsT2),
 * i.e. the V(ov, NVecsT2) of brpa_t2_copyout_ov; Am(NVecsT2).  Host or
 * device.  eig_sosex / eig_2g (host, may be NULL): RPAOutput%EigSOSEX(mu) =
 * -S1b A and Eig2g(mu) = -4 S2g A^2; yxu_diag (may be NULL, host or device,
 * ngrid x NVecsT2): YXUggm(g, mu) = M_mu(g, g), the 2b/2c/2d intermediate.
 * One fused FP64 MFMA kernel per chunk of vectors: M is built and contracted
 * in registers, never stored. */
int brpa_thc_1b2g(brpa_ctx *ctx, const double *Zgh, const double *Xgi, const double *Yga,
                  const double *Uaim, const double *Am, int64_t ngrid, int64_t nocc, int64_t nvir,
                  int64_t nvecs_t2, double *ec1b, double *ec2g, double *eig_sosex, double *eig_2g,
                  double *yxu_diag);

/* ---- AO -> MO transform of the Cholesky vectors -------------------------- */
/* CRG(P, ai) (+)= sum_pq RG(P, pq) C(ai, pq) for one spin block and one subset
 * of packed AO pairs, replacing rpa_CRGxy(rpa_C(...)) of rpa_CRG
 * (rpa/rpa_core_MO.f90:233-301, rpa_C 137-190, rpa_Cab 40-82, rpa_Caa
 * 85-134): C(ai, pq) = Co(p, i) Cv(q, a) + Co(q, i) Cv(p, a) for p != q and
 * Co(p, i) Cv(p, a) for p == q.  pairs: host array of npq (p, q) 0-based AO
 * index pairs, the AO pair of each packed column in the caller's order (the
 * Fortran module builds it from the shell-pair tables as rpa_C enumerates
 * them).  RG: Fortran RG(nvecs, npq); occ_coeffs / virt_coeffs: the active
 * orbitals' AO coefficients, Fortran (nao, nocc) / (nao, nvir); CRG: C
 * R[ai][P] (Fortran CRG(nvecs, nvir*nocc)), ai = i*nvir + a.  accumulate: 0
 * overwrites CRG, 1 adds (the next subset; the reference zeroes CRG once,
 * 269).  The packed columns are unpacked on the device into the full symmetric
 * AO-pair matrices of each Cholesky vector, then two quarter transforms (H =
 * Co^T RGfull, CRG = H Cv) run as strided-batched dgemm: 2 nvecs nao nocc (nao +
 * nvir) flop instead of the packed form's 2 nocc nvir npq nvecs.  Device
 * scratch: up to ~2 GB (vectors in chunks). */
int brpa_ao_mo(brpa_ctx *ctx, double *CRG, const double *RG, int64_t nvecs, int64_t npq,
               const int32_t *pairs, const double *occ_coeffs, const double *virt_coeffs,
               int64_t nao, int64_t nocc, int64_t nvir, int accumulate);

/* ---- factorised direct-ring amplitudes --------------------------------- */
/* T2 = sum_mu A_mu V_mu V_mu^T (ov x ov), the direct-ring doubles amplitudes
 * at coupling strength lambda from nguess random guess vectors (0: the
 * reference default ceil(3 N), capped at ov) and the eigenvalue cutoff of the
 * Gram matrix (reference default 1e-10).  Replaces rpa_CC_T2 -> rpa_CC_T2_v5
 * (rpa/rpa_CC_Doubles.f90:43-77, 491-642).  The reference's A and V are
 * `allocatable, intent(out)`: the result stays on the device and the call
 * returns NVecsT2; the caller allocates A(NVecsT2), V(NVecsT2, ov) and calls
 * brpa_t2_copyout.  The Pi(u) eigenvectors/eigenvalues the reference takes
 * are not needed: Q diag(l/(1+l W)) Q^T = l (1 + l Pi)^{-1} is applied through
 * a Cholesky factor of 1 + l Pi(u) built from R.  Guess vectors: uniform
 * (-1/2, 1/2) from a SplitMix64 counter hash of (seed, column, row), columns
 * normalised (the reference draws them from gfortran's RNG, 571-588).
 * precision: BRPA_FP64 (the drop-in: rocBLAS FP64 GEMMs) or BRPA_SPLIT_I8:
 * the amplitude contractions R^T chi_k Omega, R Y_k and R^T chi_k T run as
 * split-int8 GEMMs (5 int8 slices per value, 15 products, exact int32 level
 * sums, FP64 combination; nocc*nvir <= 65536); Pi(u), the Cholesky factors,
 * the solves, the Gram matrices and the amplitude core stay FP64. */
int brpa_t2_compute(brpa_ctx *ctx, int precision, const double *R, int64_t naux,
                    const double *eocc, int64_t nocc, const double *evir, int64_t nvir,
                    const double *freqs, const double *weights, int64_t nfreqs, double lambda,
                    double cutoff, int64_t nguess, uint64_t seed, int64_t *nvecs_t2);
/* A: NVecsT2 eigenvalues of T2 (ascending); V: Fortran V(NVecsT2, ov) == C V[ai][mu]. */
int brpa_t2_copyout(brpa_ctx *ctx, double *A, double *V);

/* Rules for the amplitude-eigenvalue threshold K of the THC route
 * (rpa/rpa_definitions.f90:139-142). */
enum {
    BRPA_T2_CUTOFF_EIG = 0,              /* K = thresh */
    BRPA_T2_CUTOFF_EIG_DIV_MAXEIG = 1,   /* K = thresh * max |A| */
    BRPA_T2_CUTOFF_EIG_DIV_NELECTRON = 2,/* K = thresh / (2 nocc) */
    BRPA_T2_CUTOFF_SUM_REJECTED = 3      /* K = |A(i+1)|, first i with |A(i+1)+...+A(NT2)| < thresh */
};
/* Small-eigenvalue cutoff of the amplitudes left on the device by
 * brpa_t2_compute: the cutoff section of rpa_THC_CC_T2
 * (rpa/rpa_CC_Doubles.f90:809-983).  *common_thresh < 0: supersystem, K
 * follows cutoff_type and is written back to *common_thresh; >= 0: a subsystem,
 * K = *common_thresh (unchanged).  Scanning the eigenvalues in stored order
 * (largest |A| first), the first with |A| <= K ends the kept set; with
 * smooth_step, |A| >= K keeps A, K*steepness < |A| < K scales A by
 * rpa_CC_T2_SmoothStepFunction (29-40), |A| <= K*steepness ends the set.
 * The device amplitudes become the first *nvecs_t2 pairs (A scaled); then
 * brpa_t2_copyout / brpa_t2_copyout_ov / brpa_sosex see only those.  No kept
 * eigenvector is an error (BRPA_ERR_ARG, amplitudes unchanged), as the
 * reference's error stop (941-944).  nvecs_scaled (may be NULL): eigenvalues
 * in the step interval. */
int brpa_t2_cutoff(brpa_ctx *ctx, int cutoff_type, double thresh, int smooth_step,
                   double steepness, int64_t nocc, double *common_thresh, int64_t *nvecs_t2,
                   int64_t *nvecs_scaled);
/* EigRPA(mu) = 2 A_mu |R^T V_mu|^2 of the device amplitudes (rpa_T2_EigRPA,
 * rpa/rpa_CC_Doubles.f90:987-1009); its sum is the T2 direct-ring energy
 * (rpa_THC.f90:126).  R: C R[ai][P] (ov x naux), host or device; eig_rpa:
 * NVecsT2 doubles, host or device. */
int brpa_t2_eig_rpa(brpa_ctx *ctx, const double *R, int64_t naux, double *eig_rpa);
/* A as brpa_t2_copyout; V in the THC route's layout, Fortran V(ov, NVecsT2) ==
 * C V[mu][ai] (rpa/rpa_CC_Doubles.f90:968-975). */
int brpa_t2_copyout_ov(brpa_ctx *ctx, double *A, double *V);

/* ---- exchange correction ---------------------------------------------------- */
/* EcExchange = -alpha sum_{aibj} (aj|bi) T(ai,bj), T = V^T diag(B) V with
 * B = A (or A/(1-4A^2) when s2_operator), evaluated pair by pair
 * W_ij = R_i V_j^T.  Replaces rpa_CC_Exchange_SOSEX
 * (rpa/rpa_CC_Exchange.f90:36-97).  V: C V[ai][mu] (Fortran Vkai(NVecsT2, v, o)),
 * R: C R[ai][P] (Fortran Rkai(NVecsPiU, v, o)).  precision: BRPA_FP64 (the
 * drop-in: rocBLAS FP64 GEMMs per pair block, double-double pair sums) or
 * BRPA_SPLIT_I8 (one fused gfx950 kernel: W_ij and W_ji tiles from 5 int8
 * slices per value, 15 products with exact int32 level sums, FP64
 * combination, double-double tile sums; W is never stored; nvir <= 65536). */
int brpa_sosex(brpa_ctx *ctx, int precision, double *ec_exchange, double alpha, const double *V,
               const double *A, const double *R, int64_t nocc, int64_t nvir, int64_t naux,
               int64_t nvecs_t2, int s2_operator);

/* Third-order exchange MBPT3-1, replacing rpa_CC_Exchange_MBPT3_1
 * (rpa/rpa_CC_Exchange.f90:232-270, with rpa_CC_Ukl 178-207, rpa_CC_Dkl
 * 210-229, rpa_CC_TrXK 273-336 and real_Lyapunov, linear-algebra/
 * real_linalg.f90:125-171), argument order as there:
 *   U = (R^T V)^T (R^T V), D = V^T diag(e_a - e_i) V,
 *   Bm^T X + X Bm = Am with Am = (A_m + A_n) D, Bm = D + 2U + 4U diag(A),
 *   EcExchange = -sum_{aibj} (aj|bi) (V X V^T)(ai, bj).
 * V: C V[ai][mu] (Fortran Vkai(NVecsT2, v, o)), R: C R[ai][P]; host or device.
 * The Lyapunov equation is solved by a scaled matrix-sign iteration (no
 * Schur form); X is factored (-X = U'^T U', or its eigenpairs) and the trace
 * is the SOSEX contraction of brpa_sosex in `precision` (BRPA_FP64 or
 * BRPA_SPLIT_I8).  OccEnergies/VirtEnergies: the orbital energies the
 * caller's route uses (canonical or semicanonical, rpa_CC.f90:977-987). */
int brpa_mbpt3_1(brpa_ctx *ctx, int precision, double *ec_exchange, const double *V,
                 const double *A, int64_t nvecs_t2, const double *R, int64_t naux,
                 const double *eocc, const double *evir, int64_t nocc, int64_t nvir);
/* Third-order exchange MBPT3-2, replacing rpa_CC_Exchange_MBPT3_2
 * (rpa/rpa_CC_Exchange.f90:381-393 with rpa_CC_TrW2 339-378), argument order
 * as there: EcExchange = 2 sum_{c,k,l} (sum_{a,i} R(l, ai) T(ak, ci))^2 with
 * T = V diag(A) V^T.  FP64 (rocBLAS strided-batched GEMMs over T~_k(c, ia) =
 * T(ci, ak), then M_k = R^T T~_k^T). */
int brpa_mbpt3_2(brpa_ctx *ctx, double *ec_exchange, const double *R, const double *V,
                 const double *A, int64_t nocc, int64_t nvir, int64_t naux, int64_t nvecs_t2);

/* Exchange correction of many fragments: brpa_t2_compute followed by brpa_sosex
 * on the device-resident amplitudes, per fragment (the reference's
 * rpa_CC.f90:884-892 then 963-976 sequence, subsystem after subsystem).
 * Fragment f uses GuessNVecsT2 = ceil(guess_factor * naux) (the reference's
 * default factor is 3.0, rpa_definitions.f90:428; 0 < guess_factor <= 1e6, else
 * BRPA_ERR_ARG) and guess seed seed + f.
 * nworkers (1..8, 0 = 1) worker contexts, each with its own stream, rocBLAS
 * handle and workspace and driven by its own host thread, take the fragments
 * one at a time, so that the latency-bound stages of different fragments
 * overlap; with exactly four workers each worker's stream runs on its own
 * quarter of the CUs.  Across worker counts the results agree to rounding (the
 * split of long GEMM reductions follows the CUs of each worker's share; within
 * one worker count they are reproducible bit for bit).  Per-stage device times
 * (summed over the workers) are left in the BRPA_STAGE_T2_* / BRPA_STAGE_SOSEX
 * slots.  ec_exchange[f] receives E_x; nvecs_t2 (may be NULL) the
 * retained NT2.
 * Unless s2_operator is set (the cumulant rescaling needs the amplitude
 * eigenpairs), every quadrature weight is positive, BRPA_EXCHANGE_FORM is
 * not "pairs" in the environment and the fragment is within the explicit
 * form's sizes (o*v <= 46340, GuessNVecsT2 <= 46340, nfreqs*naux*o*v < 2^31),
 * each fragment runs in the explicit ov
 * space, all FP64: I = -sum_k w_k/(4 pi) chi_k R (1 + l Pi_k)^{-1} R^T chi_k is
 * formed as an ov x ov matrix, the basis of span(I Omega) is chosen by the
 * reference's Gram eigenvalue cutoff, the amplitude matrix
 * T2 = T f(T^T I T) T^T (f(a) = a/(2(1+a))) is formed as an ov x ov matrix, and
 * E_x = -alpha sum (ai|bj) T2(aj,bi) with (ai|bj) = (R R^T)(ai,bj); this
 * equals the reference's pair contraction of the same amplitudes.  Otherwise
 * the reference's order runs (brpa_t2_compute, then brpa_sosex) and precision
 * selects the arithmetic of its amplitude and SOSEX contractions (BRPA_FP64 or
 * BRPA_SPLIT_I8). */
int brpa_exchange_batch(brpa_ctx *ctx, int precision, int64_t nfrag, const brpa_frag_t *frags,
                        double lambda, double cutoff, double guess_factor, uint64_t seed,
                        double alpha, int s2_operator, int nworkers, double *ec_exchange,
                        int64_t *nvecs_t2);

/* ---- dense FP64 building blocks -------------------------------------------------- */
/* The library's own MFMA kernels behind the explicit-space exchange (the
 * energy-only path of brpa_exchange_batch), exposed for callers that hold
 * their matrices in HBM.  Column-major (Fortran) layout, device pointers,
 * run on the context stream, complete on return.
 * brpa_dgemm: C (m x n) = alpha op(A) op(B) + beta C, transa/transb 0 or 1; it
 * replaces the dgemm wrappers the reference routes its products through,
 * real_ab_x / real_aTb_x / real_abT_x / real_aTbT_x
 * (linear-algebra/real_linalg.f90:1079-1110, 1233-1265, 1296-1328, 1363-1395).
 * brpa_dsyrk: the upper triangle of C (n x n) = alpha op(A) op(A)^T + beta C,
 * op(A) n x k (trans: A is k x n); fill_lower = 1 also writes the lower
 * triangle (C symmetric on return).  No single reference counterpart (the
 * reference forms these products with dgemm).
 * brpa_dtrsm: B <- U^{-T} B (side_right = 0, B m x n, U m x m) or B <- B U^{-1}
 * (side_right = 1, U n x n), U upper triangular (e.g. the factor A = U^T U of a
 * Cholesky decomposition, real_Cholesky, real_linalg.f90:39-57).  No single
 * reference counterpart.
 * brpa_dtrtri: W_z = U_z^{-1} for nmat upper triangular n x n matrices U_z = U +
 * z su (only the upper triangle is read), W_z = W + z sw written whole (zeros
 * below the diagonal); the inverses of the 64 x 64 diagonal blocks by
 * substitution, the rest by recursive doubling on the GEMM kernel (LAPACK
 * dtrtri 'U' 'N' semantics; the library's substitute for its triangular
 * solves).  No single reference counterpart. */
int brpa_dgemm(brpa_ctx *ctx, int transa, int transb, int64_t m, int64_t n, int64_t k,
               double alpha, const double *A, int64_t lda, const double *B, int64_t ldb,
               double beta, double *C, int64_t ldc);
int brpa_dsyrk(brpa_ctx *ctx, int trans, int fill_lower, int64_t n, int64_t k, double alpha,
               const double *A, int64_t lda, double beta, double *C, int64_t ldc);
int brpa_dtrsm(brpa_ctx *ctx, int side_right, int64_t m, int64_t n, const double *U, int64_t ldu,
               double *B, int64_t ldb);
int brpa_dtrtri(brpa_ctx *ctx, int64_t n, int64_t nmat, const double *U, int64_t ldu,
                int64_t su, double *W, int64_t ldw, int64_t sw);

/* ---- adjoint ----------------------------------------------------------------- */
/* dE_RPA/dR (C R[ai][P], same shape as R) and dE_RPA/d(eps_occ), dE_RPA/d(eps_vir)
 * of the FP64 energy of brpa_ecrpa, with the frequency grid held fixed.  Pi(u)
 * of every frequency is recomputed in this call, never kept from the forward
 * pass; the call's device workspace holds 2 F N^2 + N K + O(F K) doubles,
 * plus N K for each of R and dR given as a host buffer (F = nfreqs,
 * N = naux, K = nocc*nvir).  No reference counterpart: the
 * reference computes no RPA gradients.  deocc/devir may be NULL. */
int brpa_ecrpa_grad(brpa_ctx *ctx, const double *R, int64_t naux, const double *eocc,
                    int64_t nocc, const double *evir, int64_t nvir, const double *freqs,
                    const double *weights, int64_t nfreqs, double *dR, double *deocc,
                    double *devir);

#ifdef __cplusplus
}
#endif
#endif /* BRPA_H */
