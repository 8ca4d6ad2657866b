/*
 * fastfermion_amd.h -- C ABI of the MI355X propagation library (libffamd.so).
 *
 * This is the drop-in boundary for fastfermion's `propagate` hot path.  The
 * reference exposes that path as the pybind11 function ffcore.propagate with a
 * Pauli overload (src/python_bind.cpp:392-425 -> src/pauli/propagate.h:297-300)
 * and a Majorana overload (src/python_bind.cpp:490-527 ->
 * src/majorana_propagation.h:35-45).  The functions below take exactly the data
 * those overloads consume -- a gate list and an observable -- as plain arrays,
 * and return the propagated observable as plain arrays.  No C++ types, no
 * exceptions and no allocator ownership cross this boundary.
 *
 * Conventions
 *   - Every function returns an int status: FFX_OK (0) or a negative FFX_E*
 *     code; ffx_last_error(ctx) then holds a message.
 *   - Input arrays are owned by the caller and only read during the call.
 *   - Output arrays in ffx_terms are owned by the library; release them with
 *     ffx_terms_free.
 *   - A context is bound to one HIP device and owns one HIP stream plus the
 *     device memory pools; use one context per host thread.
 *
 * Key layout (term-major, n_words = W in 1..4, W*64 >= highest index + 1):
 *   Pauli    x[i*W + w] bit b <=> qubit 64w+b is X or Y  (reference xory)
 *            z[i*W + w] bit b <=> qubit 64w+b is Y or Z  (reference yorz)
 *            (src/pauli/algebra.h:30-33)
 *   Majorana m[i*W + w] bit b <=> Majorana operator m_{64w+b} is in the string
 *            (src/majorana_algebra.h:30-36)
 * Keys of one observable must be distinct.
 */
#ifndef FASTFERMION_AMD_H
#define FASTFERMION_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FFX_MAX_WORDS 4

/* maxdegree value that disables the partner degree filter (the reference's
 * maxdegree=None of the Majorana overload, src/python_bind.cpp:505-511). */
#define FFX_NO_DEGREE_LIMIT 0x7FFFFFFF

enum ffx_status {
    FFX_OK = 0,
    FFX_EINVAL = -1,    /* invalid argument (shape, index, duplicate keys, ...) */
    FFX_ENOMEM = -2,    /* device or host allocation failed */
    FFX_EHIP = -3,      /* HIP runtime error */
    FFX_ENODEV = -4,    /* no usable HIP device */
    FFX_ECAPACITY = -5  /* term count exceeds the configured capacity limit */
};

/* Gate kinds of a Pauli circuit (src/pauli/gates.h:71-269). */
enum ffx_gate_kind { FFX_ROT = 0, FFX_H = 1, FFX_S = 2, FFX_CNOT = 3, FFX_SWAP = 4, FFX_CZ = 5 };

/* Coefficient arithmetic on the device. */
enum ffx_precision {
    FFX_FP64 = 0,   /* IEEE double, same rounding as the reference */
    FFX_FP32 = 1,   /* float storage and arithmetic */
    FFX_FP32C = 2   /* float storage with a stored compensation term (float-float) */
};

/* One gate of a Pauli circuit.  ROT(P, theta) is U = exp(-i theta/2 P)
 * (src/pauli/gates.h:217-255); Clifford kinds use q0 (and q1 for two-qubit
 * gates: CNOT control q0 target q1, src/pauli/gates.h:39-51). */
typedef struct ffx_pauli_gate {
    int32_t kind;
    int32_t q0, q1;
    int32_t reserved;
    double theta;
    uint64_t x[FFX_MAX_WORDS];
    uint64_t z[FFX_MAX_WORDS];
} ffx_pauli_gate;

/* One Majorana rotation U = exp(-i theta/2 * i^r m_S) (src/majorana_gates.h:17-50);
 * r = 1 iff deg(S) mod 4 is 2 or 3 (the reference's MROT::_r). */
typedef struct ffx_majorana_gate {
    uint64_t m[FFX_MAX_WORDS];
    double theta;
    int32_t r;
    int32_t reserved;
} ffx_majorana_gate;

/* Truncation and execution options; the first five fields are the reference's
 * propagate keyword arguments (src/pauli/propagate.h:297-300). */
typedef struct ffx_options {
    int32_t maxdegree;      /* partner degree filter (propagate.h:82): partners of degree > maxdegree
                               are dropped, compared literally (a negative value drops every
                               partner); FFX_NO_DEGREE_LIMIT: no filter */
    int32_t topk;           /* keep the topk largest |c| at each truncation; 0: off */
    int32_t max_xweight;    /* X/Y-weight cap (truncate.h:41-46); -1: off */
    int32_t xtrunc_period;  /* the cap applies when (rot_count+1) % period == 0 */
    int32_t batched;        /* 1: truncate once per commuting run (propagate.h:337-346) */
    int32_t precision;      /* enum ffx_precision */
    double mincoeff;        /* erase |c| <= mincoeff after truncation points; 0: off */
} ffx_options;

/* Propagated observable, term-major, owned by the library. */
typedef struct ffx_terms {
    int64_t n_terms;
    int32_t n_words;
    int32_t reserved;
    uint64_t* x;        /* Pauli xory words; Majorana occupation words */
    uint64_t* z;        /* Pauli yorz words; NULL for Majorana */
    double* c_re;
    double* c_im;
} ffx_terms;

/* Work and timing counters of the last propagate call on a context. */
typedef struct ffx_profile {
    int64_t n_gates;         /* rotations applied */
    int64_t updates;         /* term-gate updates: sum of live terms entering each rotation */
    int64_t peak_terms;      /* largest live-term count after a rotation */
    int64_t launches;        /* kernel launches issued */
    double seconds_total;    /* wall time of the device part (upload to download) */
    double seconds_upload;
    double seconds_propagate;
    double seconds_download;
    int64_t outs;            /* sum of live terms leaving each rotation */
    int64_t compactions;     /* erased-slot compactions run */
    int64_t run_launches;    /* launches of the propagation kernels (one rotation, window or layer each) */
    double seconds_kernel;   /* HIP-event time of the timed ones among them */
    int64_t timed_launches;  /* how many were timed (see ffx_set_kernel_timing) */
    double seconds_enqueue;  /* host time spent issuing the propagation launches */
} ffx_profile;

typedef struct ffx_ctx ffx_ctx;

/* Context lifetime (no reference counterpart: the reference is in-process CPU code). */
int ffx_ctx_create(int device, ffx_ctx** out);
int ffx_ctx_destroy(ffx_ctx* ctx);
const char* ffx_last_error(const ffx_ctx* ctx);

/* Pauli propagation.  Replaces pauli_gates::propagate (src/pauli/propagate.h:297-493)
 * behind ffcore.propagate overload #1 (src/python_bind.cpp:392-425).
 * c_im may be NULL (all imaginary parts zero).  When out_expect is non-NULL the
 * product-state expectation <b|O|b> with b = basis_state (n_words words; NULL means
 * |0...0>) is computed on the device; when out is NULL the terms are not
 * downloaded at all (the result stays on the device until the next call). */
int ffx_pauli_propagate(ffx_ctx* ctx, const ffx_pauli_gate* gates, int64_t n_gates, int32_t n_words,
                        const uint64_t* x, const uint64_t* z, const double* c_re, const double* c_im,
                        int64_t n_terms, const ffx_options* opts, ffx_terms* out,
                        const uint64_t* basis_state, double out_expect[2]);

/* Majorana propagation.  Replaces majorana_gates::propagate
 * (src/majorana_propagation.h:18-55) behind ffcore.propagate overload #2
 * (src/python_bind.cpp:490-527).  maxdegree as in ffx_options; the
 * threshold applies after every gate when mincoeff > 0.  When out_expect is
 * non-NULL, <occ|O|occ> for the Fock state with mode occupations occ (NULL: vacuum)
 * is computed on the device (src/fockstate.h:95-114). */
int ffx_majorana_propagate(ffx_ctx* ctx, const ffx_majorana_gate* gates, int64_t n_gates, int32_t n_words,
                           const uint64_t* m, const double* c_re, const double* c_im, int64_t n_terms,
                           int32_t maxdegree, double mincoeff, int32_t precision, ffx_terms* out,
                           const uint64_t* occ, double out_expect[2]);

/* One circuit instance of a batched Pauli propagation. */
typedef struct ffx_instance {
    const ffx_pauli_gate* gates; /* circuit, gate order as in ffx_pauli_propagate */
    int64_t n_gates;
    const uint64_t* x;           /* observable: n_terms x n_words masks, term-major */
    const uint64_t* z;
    const double* c_re;
    const double* c_im;          /* NULL: real coefficients */
    int64_t n_terms;
    uint64_t basis_state[4];     /* readout state |b> (bit q of word q/64) */
} ffx_instance;

/* Batched Pauli propagation (no reference counterpart: the reference runs one
 * circuit per propagate call, src/pauli/propagate.h:297-493, and a batch is a
 * loop over such calls).  Every instance runs the reference schedule and
 * truncation of ffx_pauli_propagate with the shared options, in one of two
 * layouts (the autotuned one, else merged):
 *  - merged: all instances' terms in one device-wide term store, each tagged
 *    with its instance number in key bits above the instances' qubits; one
 *    launch per window step applies every instance's own window (the longest
 *    schedule sets the launch count).  Needs ceil(log2 n_inst) free bits at
 *    the top of the last key word; otherwise the other layout runs.
 *  - one workgroup per instance: all instances in one launch (largest
 *    estimated work first: rotations^2 / qubits), with `slots` term slots per
 *    instance (0: a default); the upload, the launch and the download of the
 *    instances' states run as one captured HIP graph, replayed when the next
 *    call has the same layout.  Instances whose terms outgrow the slot budget
 *    run afterwards on the device-wide engine.
 * Instances that contain Clifford gates or ask for top-k run on the device-wide
 * engine, so every result is complete.  out_expect[2i], [2i+1] receives
 * <b_i|O_i|b_i> (merged: summed with atomics, so the last bits can vary from
 * run to run); out_terms[i] (optional) the live term count at the end;
 * out_fallback (optional) how many instances took the device-wide engine.
 * FFX_BATCH_MERGED=0 / 1 in the environment forces a layout. */
int ffx_pauli_propagate_batch(ffx_ctx* ctx, const ffx_instance* inst, int32_t n_inst, int32_t n_words,
                              const ffx_options* opts, int64_t slots, double* out_expect, int64_t* out_terms,
                              int32_t* out_fallback);

/* Launch configuration of batched propagation chosen by ffx_batch_autotune. */
typedef struct ffx_batch_tuning {
    int32_t block;          /* threads of the workgroup that runs one instance (128, 256 or 512) */
    int32_t window;         /* most commuting rotations fused into one window (1..4) */
    int64_t slots;          /* term slots per instance */
    double ms;              /* time of the sample batch with this configuration */
    int32_t candidates;     /* configurations timed */
    int32_t sample;         /* instances in the sample batch */
    int32_t merged;         /* layout: 1 all instances in one term store, one launch pair per layer-group
                               step; 2 the same store, one launch per window step; 0 one workgroup per
                               instance (block and slots apply to 0 only) */
    int32_t compact_every;  /* layout 1: group steps between device-decided compaction checks */
    int32_t group_max;      /* layout 1: most commuting rotations with overlapping supports in a group (1..8) */
    int32_t split;          /* layout 1: term stores (1..4) the batch is cut into by schedule length, each
                               on its own stream, so that long schedules run beside short ones */
} ffx_batch_tuning;

/* Autotuner of batched propagation (no reference counterpart): times the
 * layer-group layout with compaction checks every 2, 4 or 8 steps and groups of
 * up to 4, 6 or 8 overlapping commuting rotations, the window-step layout, and
 * the one-workgroup layout with every combination of workgroup size (256, 512),
 * window length (3, 4) and slot budget (8192, 32768), on the whole batch
 * (sample <= 0) or on the `sample` instances with the longest circuits, with the
 * options of `opts`; prefers configurations without device-wide fallbacks, and
 * keeps the fastest on the context for this key width, precision and
 * coefficient kind; later ffx_pauli_propagate_batch calls with slots = 0 use
 * it. */
int ffx_batch_autotune(ffx_ctx* ctx, const ffx_instance* inst, int32_t n_inst, int32_t n_words,
                       const ffx_options* opts, int32_t sample, ffx_batch_tuning* out);

/* Number of batched calls on the context that replayed a captured HIP graph
 * (a batch with the same pool layout as the previous one: same instance count,
 * schedule lengths and term counts, e.g. an angle sweep). */
int ffx_batch_graph_replays(const ffx_ctx* ctx, int64_t* out);

/* Expectation of the device-resident result of the last propagate call on a
 * product state (Pauli: qubit bits b; Majorana: mode occupations).
 * Replaces QubitProductState::operator() (src/qubitproductstate.h:82-88),
 * PauliPolynomial::overlapwithzero (src/pauli/algebra.h:549-557) and
 * FockState::operator() (src/fockstate.h:95-114). */
int ffx_expect_last(ffx_ctx* ctx, const uint64_t* state, double out[2]);

void ffx_terms_free(ffx_terms* t);

/* Counters of the last propagate call (the reference's profile_get dictionary,
 * src/python_bind.cpp:427-446, reports CPU phases; these are device phases). */
int ffx_profile_get(const ffx_ctx* ctx, ffx_profile* out);

/* Kernel timing (no reference counterpart): with every = N > 0, one
 * propagation kernel launch in N (chosen by a hash of the launch number, so the
 * sample does not alias with periodic launch patterns) is bracketed by a pair of
 * HIP events on the context's stream; ffx_profile.seconds_kernel /
 * timed_launches report their summed time and count.  Sampling keeps the event
 * overhead out of the other launches.  every = 0 switches timing off. */
int ffx_set_kernel_timing(ffx_ctx* ctx, int every);

/* Number of visible HIP devices (0 when none); never fails. */
int ffx_device_count(void);

/* Bytes of the fixed scratch array of ffx_pauli_propagate_noalloc, held on each
 * device from the first such call there until the process ends. */
#define FFX_NOALLOC_SCRATCH_BYTES (128ull * 1024 * 1024)

/* ffx_pauli_propagate without device memory allocation.  No counterpart in the
 * reference; it is for callers that cannot allocate device memory: code that runs
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc.
 * Same schedule, truncation, kernels and results as ffx_pauli_propagate (no pruned
 * mode).  Every device block of the call -- term store, hash index, counters,
 * epoch arrays, schedule tables, staging -- is carved from a fixed scratch array of
 * FFX_NOALLOC_SCRATCH_BYTES per device; the call calls no hipMalloc or hipFree (its
 * host-side schedule uses the C++ heap).  A problem whose term store would not fit
 * (planned from the key width, precision, gate and term counts: e.g. 524288 slots
 * for 128-bit keys in FP64 with 5600 gates) fails with FFX_ECAPACITY before any
 * launch, or when the live terms outgrow the planned store.  The propagated terms
 * go to caller-owned host arrays: out_x, out_z (out_cap * n_words words each),
 * out_re and optionally out_im (out_cap doubles); with out_cap = 0 (arrays may be
 * NULL) only out_expect / out_n are produced.  *out_n receives the live term count;
 * a result of more than out_cap terms fails with FFX_ECAPACITY after setting it.
 * All calls on one GPU share its scratch array and leave their contents in it:
 * calls that overlap on one GPU (several threads or contexts) race.  The call
 * leaves no device-resident result for ffx_expect_last. */
int ffx_pauli_propagate_noalloc(ffx_ctx* ctx, const ffx_pauli_gate* gates, int64_t n_gates, int32_t n_words,
                                const uint64_t* x, const uint64_t* z, const double* c_re, const double* c_im,
                                int64_t n_terms, const ffx_options* opts, const uint64_t* basis_state,
                                double out_expect[2], uint64_t* out_x, uint64_t* out_z, double* out_re,
                                double* out_im, int64_t out_cap, int64_t* out_n);

/* ------------------------------------------------------------------------
 * Pruned propagation and the learned term pruner (no reference counterpart:
 * the reference truncates by degree, |c|, top-k and X/Y weight only,
 * src/pauli/truncate.h:16-57).  Both change results; they are opt-in and the
 * entry points above never use them.
 * ------------------------------------------------------------------------ */

#define FFX_PRUNER_FEATURES 64
#define FFX_PRUNER_HIDDEN 128
/* MLP 64 -> 128 -> 128 -> 1 (ReLU).  Flat fp32 parameter vector:
 * W1[64][128] (in x out), b1[128], W2[128][128], b2[128], w3[128], b3[1]. */
#define FFX_PRUNER_NPARAMS 24961

/* Options of the pruned mode and the site geometry used by the term features. */
typedef struct ffx_prune_opts {
    double clifford_tol;    /* rotations with |cos theta| <= clifford_tol keep only their sin branch and
                               run as fused Clifford maps (a run of them is one kernel); 0: off */
    float log10_tau;        /* erase terms with log10|c| + predicted log10|influence| < log10_tau */
    int32_t prune_period;   /* run the learned pruner after every prune_period-th rotation; 0: off */
    int32_t rot_per_layer;  /* rotations per circuit layer (feature: remaining layers) */
    int32_t obs_site;       /* qubit / Majorana index of the observable's site (feature: distance) */
    uint8_t row[256];       /* lattice row and column of every qubit (Pauli) or Majorana index */
    uint8_t col[256];
} ffx_prune_opts;

/* ffx_pauli_propagate with the pruned mode of `po` (NULL: identical to ffx_pauli_propagate).
 * The learned pruner needs real coefficients and a model (ffx_pruner_init / _set_params). */
int ffx_pauli_propagate_pruned(ffx_ctx* ctx, const ffx_pauli_gate* gates, int64_t n_gates, int32_t n_words,
                               const uint64_t* x, const uint64_t* z, const double* c_re, const double* c_im,
                               int64_t n_terms, const ffx_options* opts, const ffx_prune_opts* po,
                               ffx_terms* out, const uint64_t* basis_state, double out_expect[2]);

/* ffx_majorana_propagate with the pruned mode of `po`. */
int ffx_majorana_propagate_pruned(ffx_ctx* ctx, const ffx_majorana_gate* gates, int64_t n_gates, int32_t n_words,
                                  const uint64_t* m, const double* c_re, const double* c_im, int64_t n_terms,
                                  int32_t maxdegree, double mincoeff, int32_t precision, const ffx_prune_opts* po,
                                  ffx_terms* out, const uint64_t* occ, double out_expect[2]);

/* Term-contribution logs: one plain FP64 propagation (one launch per rotation,
 * reference schedule and truncation of `opts`; no Clifford blocks, no top-k),
 * then the transposed rotations in the opposite order give every live term q
 * after rotation e its influence v_q(e) on the final expectation <b|O|b>
 * (b = basis_state / occ; NULL: all zeros).  At e = 0, log_every, 2*log_every,
 * ... <= n_rotations, every live term yields one row: FFX_PRUNER_FEATURES
 * features (same as the pruner uses; geometry from `po`) into feats and
 * clamp(log10(|v| + 1e-12), -12, 0)/4 into targets, at most max_rows rows in
 * total.  sums[2p], sums[2p+1] receive sum_q c_q v_q (re, im) at the p-th
 * logged step (n_points of them), each equal to the final expectation up to
 * rounding.  *n_rows: rows written. */
int ffx_pauli_contrib_log(ffx_ctx* ctx, const ffx_pauli_gate* gates, int64_t n_gates, int32_t n_words,
                          const uint64_t* x, const uint64_t* z, const double* c_re, const double* c_im,
                          int64_t n_terms, const ffx_options* opts, const ffx_prune_opts* po,
                          const uint64_t* basis_state, int32_t log_every, int64_t max_rows, float* feats,
                          float* targets, int64_t* n_rows, double* sums, int32_t max_points, int32_t* n_points);

int ffx_majorana_contrib_log(ffx_ctx* ctx, const ffx_majorana_gate* gates, int64_t n_gates, int32_t n_words,
                             const uint64_t* m, const double* c_re, const double* c_im, int64_t n_terms,
                             int32_t maxdegree, double mincoeff, const ffx_prune_opts* po, const uint64_t* occ,
                             int32_t log_every, int64_t max_rows, float* feats, float* targets, int64_t* n_rows,
                             double* sums, int32_t max_points, int32_t* n_points);

/* The context's pruner model (one per context).  init: He-initialised weights
 * from a SplitMix64 seed; set/get: the flat fp32 parameter vector (n must be
 * FFX_PRUNER_NPARAMS).  The device keeps bf16 copies of the weights for its
 * MFMA kernels and the fp32 vector as the master copy for training. */
int ffx_pruner_init(ffx_ctx* ctx, uint64_t seed);
int ffx_pruner_set_params(ffx_ctx* ctx, const float* params, int64_t n);
int ffx_pruner_get_params(ffx_ctx* ctx, float* params, int64_t n);

/* Forward pass on n feature rows (n x FFX_PRUNER_FEATURES floats, rounded to
 * bf16 on upload): out[i] = predicted target of row i. */
int ffx_pruner_forward(ffx_ctx* ctx, const float* feats, int64_t n, float* out);

/* Loss (1/n) sum_i w_i (y_i - t_i)^2 over n rows, w_i = under_weight when the
 * prediction y_i is below the target t_i and 1 otherwise (under_weight = 1: mean
 * squared error), and its gradient with respect to every parameter (the bf16
 * weights' gradient, applied to the fp32 master copy). */
int ffx_pruner_loss_grad(ffx_ctx* ctx, const float* feats, const float* targets, int64_t n, float under_weight,
                         double* loss, float* grad);

/* Adam training on the device with the loss of ffx_pruner_loss_grad: `epochs`
 * passes over n rows in shuffled minibatches of `batch` rows (seeded), learning
 * rate lr.  epoch_loss[e] (optional) receives the loss of epoch e during training. */
int ffx_pruner_train(ffx_ctx* ctx, const float* feats, const float* targets, int64_t n, int32_t epochs,
                     int32_t batch, float lr, float under_weight, uint64_t seed, double* epoch_loss);

#ifdef __cplusplus
}
#endif

#endif
