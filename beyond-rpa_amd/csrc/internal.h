// Internal declarations shared by the host orchestration and the kernel files.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <mutex>
#include <string>
#include <vector>

#include "brpa.h"

struct brpa_ctx {
    int device = 0;
    hipStream_t stream = nullptr;       // stream in use
    hipStream_t own_stream = nullptr;   // stream created by the context
    std::string err;
    char *ws = nullptr;                 // device workspace
    size_t ws_cap = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_kernel_ms = 0.0;
    // per-stage device time of the last call (HIP events on ctx->stream)
    hipEvent_t sev[2 * BRPA_NSTAGES] = {};
    double stage_ms[BRPA_NSTAGES] = {};
    int stage_launches[BRPA_NSTAGES] = {};
    // set on an exchange-batch worker: the amplitude and SOSEX calls add to the
    // stage times instead of starting them afresh
    bool stage_keep = false;
    // device time (HIP events around each launch), algorithmic (FP64-
    // equivalent) flop and launch count per measured kernel (BRPA_KERNEL_*)
    // since the last stage_reset; gev holds event pairs not yet collected,
    // gev_flops / gev_kind their flop and kernel
    std::vector<hipEvent_t> gev;
    std::vector<double> gev_flops;
    std::vector<int> gev_kind;
    size_t gev_used = 0;
    double kst_ms[BRPA_NKERNELS] = {}, kst_flops[BRPA_NKERNELS] = {};
    int64_t kst_launches[BRPA_NKERNELS] = {};
    // amplitude (T2) results kept on the device until brpa_t2_copyout
    void *rocblas = nullptr;   // rocblas_handle, created on first use
    double *t2_A = nullptr, *t2_V = nullptr;
    size_t t2_A_cap = 0, t2_V_cap = 0;
    int64_t t2_nt = -1, t2_nai = 0;
    // the kept pair is the factored core (A' = -1, V' = T L), not eigenpairs
    bool t2_factored = false;
    // MBPT3-1: the transformed amplitudes V'' and weights B'' handed to the SOSEX
    // contraction (kept apart from the workspace, which that call may regrow)
    double *x_V = nullptr, *x_B = nullptr;
    size_t x_V_cap = 0, x_B_cap = 0;
    bool mbpt3_chol = false;   // the last MBPT3-1 factored -X by Cholesky (else eigenpairs)
    // pinned bounce buffer of the small completed copies (copy_sync), kBounceBytes
    void *bounce = nullptr;
    // pinned host staging of the batched E_RPA path (uploads and results)
    char *pin = nullptr;
    size_t pin_cap = 0;
    // hipGraph replay of the batched E_RPA launch sequence (brpa_ctx_set_graphs)
    int use_graphs = 0;
    struct GraphEntry {
        std::string key;
        hipGraphExec_t exec = nullptr;
    };
    std::vector<GraphEntry> graphs;
    int last_graph_replays = 0, last_graph_captures = 0;
    // autotuner: split-int8 product variant per fragment-size bucket (-1: default)
    int tune_force = -1;                // variant forced for every fragment (tuning runs)
    int i8_choice[3] = {-1, -1, -1};
    double i8_tune_ms[3][3] = {};      // [bucket][variant], 0 = not measured
    // scratch for the partial tiles of split-k GEMMs (doubles), lent by the
    // caller of a dense sequence for its duration; null: no split
    double *split_ws = nullptr;
    size_t split_ws_cap = 0;
    // CUs the context's stream runs on (0: every CU of the device)
    int ncu = 0;
    // worker contexts (own stream, rocBLAS handle, workspace) of brpa_exchange_batch;
    // their streams run on disjoint CU sets, one share of the chip per worker
    std::vector<brpa_ctx *> workers;
};

// A context whose stream runs on the CUs [part ncu / nparts, (part + 1) ncu /
// nparts) of the device (nparts <= 1: every CU, as brpa_ctx_create; also when
// the runtime refuses the CU mask).
int ctx_create_cu_share(int device, int part, int nparts, brpa_ctx **out);

// Triangular operand of a batched GEMM (launch_gemm_f64_batched): the k range
// where op(A) or op(B) is zero is skipped (the zeros must be stored).
enum { TRI_NONE = 0, TRI_A_UPPER = 1, TRI_A_LOWER = 2, TRI_B_UPPER = 3, TRI_B_LOWER = 4 };

constexpr size_t kBounceBytes = (size_t)1 << 20;   // pinned bounce buffer of copy_sync

// Fragment-size bucket of the autotuner (by NAux).
inline int tune_bucket(int64_t naux) { return naux < 640 ? 0 : (naux < 1280 ? 1 : 2); }

// Pinned host buffer of at least `bytes` owned by the context.
int pin_reserve(brpa_ctx *ctx, size_t bytes);
void graphs_release(brpa_ctx *ctx);

void t2_release(brpa_ctx *ctx);
// Looks up the device address of the allocation-free entry point's scratch
// array on `device` (current) once per process; later calls return at once.
int noalloc_setup(int device);
void mbpt3_release(brpa_ctx *ctx);

// Records the start/end of a stage on the context stream; stage_end adds the
// elapsed time once the stream has been synchronised (stage_collect).
int stage_begin(brpa_ctx *ctx, int stage);
int stage_end(brpa_ctx *ctx, int stage);
int stage_collect(brpa_ctx *ctx, int stage);
void stage_reset(brpa_ctx *ctx);

int brpa_fail(brpa_ctx *ctx, int code, const char *fmt, ...);

// An event pair around one launch of a measured kernel: kstat_begin records
// the first event before the launch, kstat_end the second after it together
// with the kernel's algorithmic flop; kstat_collect adds the elapsed times to
// kst_* once the stream is idle; kstat_drop forgets pairs of a failed call.
int kstat_begin(brpa_ctx *ctx, size_t *slot);
int kstat_end(brpa_ctx *ctx, size_t slot, int kernel, double flops);
int kstat_collect(brpa_ctx *ctx);
void kstat_drop(brpa_ctx *ctx);

// One lock for every rocSOLVER call of the library: contexts driven from
// several host threads (the exchange batch's workers) take turns in it.
std::mutex &solver_mutex();

#define BRPA_HIP(ctx, expr)                                                              \
    do {                                                                                 \
        hipError_t _e = (expr);                                                          \
        if (_e != hipSuccess)                                                            \
            return brpa_fail((ctx), BRPA_ERR_HIP, "%s failed: %s (%s:%d)", #expr,        \
                             hipGetErrorString(_e), __FILE__, __LINE__);                 \
    } while (0)

#define BRPA_TRY(expr)            \
    do {                          \
        int _rc = (expr);         \
        if (_rc != BRPA_OK)       \
            return _rc;           \
    } while (0)

// Bump allocator over the context workspace.  A first pass with base == nullptr
// only counts bytes; ws_reserve then makes the workspace large enough.
struct Arena {
    char *base = nullptr;
    size_t off = 0;
    template <class T>
    T *take(size_t n) {
        size_t bytes = ((n * sizeof(T) + 255) / 256) * 256;
        T *p = base ? reinterpret_cast<T *>(base + off) : nullptr;
        off += bytes;
        return p;
    }
};
int ws_reserve(brpa_ctx *ctx, size_t bytes, Arena *arena);
bool is_device_ptr(const void *p);
// Copy n elements of a host-or-device array into device memory dst.
int to_device(brpa_ctx *ctx, void *dst, const void *src, size_t bytes);
int from_device(brpa_ctx *ctx, void *dst_host_or_dev, const void *src_dev, size_t bytes);
// Blocking copy ordered on ctx->stream.  (A plain hipMemcpy runs on the null
// stream, which is not ordered with the context's non-blocking stream.)
int copy_sync(brpa_ctx *ctx, void *dst, const void *src, size_t bytes, hipMemcpyKind kind);

// ---------------------------------------------------------------------------
// Fragment descriptor used by the batched kernels (device memory).
//   R     : K x N row-major (R[ai][P]), K = nocc*nvir
//   dtab  : F x K, dtab[w][k] = 4 d_k/(d_k^2 + u_w^2)   (prefactor included)
//   Pi    : F x N x N row-major, lower triangle + diagonal is what the
//           Cholesky reads; the SYRK writes both triangles
//   diag  : F x N exact FP64 diagonal of Pi
//   ldpart: F x nsteps x 2 double-double partial log-det sums of the
//           Cholesky diagonal blocks
struct FragDev {
    const double *R;
    const double *dtab;
    double *Pi;
    double *diag;
    double *ldpart;
    const double *weights;   // F (device)
    double *eu;              // F per-frequency E(u) output (device)
    double *energy;          // 1 (device)
    int *status;             // F Cholesky status (0 ok)
    double *linv;            // F x 64 x 64 inverse of the current diagonal block
    double *dpart;           // split-BF16: F x nblk x N x 2 double-double diagonal partials
    int K, N, F, nsteps;
    int nocc, nvir;
    const double *eocc, *evir, *freqs;   // device copies
    // split-BF16 slices (only when precision == BRPA_SPLIT_BF16)
    uint16_t *sl;            // F x NSLICE x N x Kpad, per column P the Kpad rows contiguous
    float *scale;            // F x nblk x N, block scale 2^(e-8)
    int Kpad, nblk;
    // split-int8 slices (only when precision == BRPA_SPLIT_I8)
    const double *sdtab;     // F x K, sqrt(c_w(k))
    int8_t *sl8;             // F x NSLICE x N x Kpad8, per column P the Kpad8 rows contiguous
    double *scale8;          // F x N, column scale max_k |X_w[k][P]| / 127
    int Kpad8;
    // 1: the Cholesky also writes each factored 64x64 diagonal block back into
    // Pi, so the lower triangle holds the complete factor L (for triangular
    // solves after the factorisation)
    int keep_l11;
    int64_t mat0;            // index of this fragment's first matrix in the batch
};

constexpr int CHOL_NB = 64;   // Cholesky block size
constexpr int CHOL_OUTER = 4; // Cholesky steps per outer panel of the trailing update
constexpr int SYRK_T = 64;    // FP64 SYRK tile
constexpr int SPLIT_KB = 128; // split-BF16 exactness block (rows per scale)
constexpr int NSLICE = 4;     // bf16 (or int8) slices per FP64 value in the split forms
constexpr int I8_KB = 128;    // split-int8: rows per split work-group (Kpad8 is a multiple)
constexpr int I8_KMAX = 65536;// split-int8: largest K with exact int32 level sums

// Launch wrappers (defined in the kernel files).  `pre` arrays are device
// prefix sums over fragments of the number of work-groups per fragment.
int launch_dtab(brpa_ctx *, const FragDev *, const int64_t *pre, int nfrag, int64_t total);
int launch_syrk_f64(brpa_ctx *, const FragDev *, const int64_t *pre, int nfrag, int64_t total);
// mode: 0 = Pi's own diagonal, 1 = split-BF16 partials, 2 = diag precomputed
int launch_set_diag(brpa_ctx *, const FragDev *, const int64_t *pre, int nfrag, int64_t total,
                    int mode);
int launch_potrf_diag(brpa_ctx *, const FragDev *, const int64_t *pre, int nfrag, int64_t total, int step);
int launch_trsm(brpa_ctx *, const FragDev *, const int64_t *pre, int nfrag, int64_t total, int step);
int launch_chol_update(brpa_ctx *, const FragDev *, const int64_t *pre, int nfrag, int64_t total, int step);
// Work items per matrix of the trailing update after Cholesky step j (N x N):
// the next 64-column strip inside an outer panel of CHOL_OUTER steps, the rest
// of the trailing matrix after its last step.
int64_t chol_upd_items(int64_t N, int64_t j);
int launch_energy(brpa_ctx *, const FragDev *, int nfrag);
int launch_fill_int(brpa_ctx *, int *p, int64_t n, int v);
int launch_split(brpa_ctx *, const FragDev *, const int64_t *pre, int nfrag, int64_t total);
int launch_syrk_bf16(brpa_ctx *, const FragDev *, const int64_t *pre, int nfrag, int64_t total);
int launch_colscale_i8(brpa_ctx *, const FragDev *, const int64_t *pre, int nfrag, int64_t total);
int launch_split_i8(brpa_ctx *, const FragDev *, const int64_t *pre, int nfrag, int64_t total);
// split-int8 product kernel variants (chosen per fragment-size bucket by the
// autotuner): 128x128 tiles in launch order, the same with one contiguous tile
// range per XCD, 64x64 tiles
enum { I8V_T128 = 0, I8V_T128_XCD = 1, I8V_T64 = 2, I8V_COUNT = 3 };
int syrk_i8_tile(int variant);
int launch_syrk_i8(brpa_ctx *, const FragDev *, const int64_t *pre, int nfrag, int64_t total,
                   int variant);

// Blocked Cholesky of the F matrices d_fr->Pi (1 + Pi already formed) of one
// fragment, in place on ctx->stream; status/linv/ldpart of the descriptor
// must be allocated.  d_pre: device scratch of chol_pre_len(N) int64.
int64_t chol_pre_len(int64_t N);

// brpa_t2_compute with an option for energy-only callers: factored = true
// leaves (A' = -1, V' = T L) with T2 = -V' V'^T (no eigensolve of T^T I T)
// when the factorisation exists, the reference's (A, V) otherwise.
int t2_build(brpa_ctx *ctx, int precision, const double *R, int64_t naux, const double *eocc,
             int64_t nocc, const double *evir, int64_t nvir, const double *freqs,
             const double *weights, int64_t nfreqs, double lambda, double cutoff, int64_t nguess,
             uint64_t seed, int64_t *nvecs_t2, bool factored);
int chol_single(brpa_ctx *ctx, const FragDev *d_fr, int64_t *d_pre, int64_t N, int64_t F);

// Exchange energy of one fragment in the explicit ov space (energy-only form of
// brpa_t2_compute + brpa_sosex, FP64): E_x and the kept amplitude count.
// xspace_fits: the sizes the explicit form takes (K = o v, NG guess vectors,
// F frequencies); larger fragments take the reference's order instead.
bool xspace_fits(int64_t K, int64_t N, int64_t NG, int64_t F);
bool xspace_applicable(const double *weights, int64_t nfreqs, double lambda);
int exchange_xspace(brpa_ctx *ctx, const double *R, int64_t naux, const double *eocc, int64_t nocc,
                    const double *evir, int64_t nvir, const double *freqs, const double *weights,
                    int64_t nfreqs, double lambda, double cutoff, int64_t nguess, uint64_t seed,
                    double alpha, double *ex, int64_t *nvecs_t2);
