// ffx_lib.hip -- C ABI of libffamd.so (declared in include/fastfermion_amd.h):
// argument checks, then dispatch on the key width to the drivers instantiated
// by ffx_inst.hip (one object per key width, arithmetic and algebra).
#include "ffx_engine.h"

using namespace ffx_host;

namespace ffx_host {
FFX_EXTERN_WIDTH(1);
FFX_EXTERN_WIDTH(2);
FFX_EXTERN_WIDTH(3);
FFX_EXTERN_WIDTH(4);
}  // namespace ffx_host

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

int ffx_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return n;
}

int ffx_ctx_create(int device, ffx_ctx** out) {
    if (!out) return FFX_EINVAL;
    *out = nullptr;
    const int nd = ffx_device_count();
    if (nd <= 0 || device < 0 || device >= nd) return FFX_ENODEV;
    ffx_ctx* ctx = new (std::nothrow) ffx_ctx();
    if (!ctx) return FFX_ENOMEM;
    ctx->device = device;
    auto fail = [&](int code) {
        ffx_ctx_destroy(ctx);
        return code;
    };
    if (hipSetDevice(device) != hipSuccess) return fail(FFX_ENODEV);
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) return fail(FFX_EHIP);
    if (dev_alloc(ctx, &ctx->ctrl, sizeof(Ctrl)) != FFX_OK) return fail(FFX_ENOMEM);
    if (dev_alloc(ctx, &ctx->total, 64) != FFX_OK) return fail(FFX_ENOMEM);
    if (dev_alloc(ctx, &ctx->partial, 2 * 1024 * sizeof(double)) != FFX_OK) return fail(FFX_ENOMEM);
    if (dev_alloc(ctx, &ctx->result, 64) != FFX_OK) return fail(FFX_ENOMEM);
    if (hipHostMalloc((void**)&ctx->ctrl_host, sizeof(Ctrl), hipHostMallocDefault) != hipSuccess) return fail(FFX_ENOMEM);
    if (hipHostMalloc((void**)&ctx->result_host, 2 * sizeof(double), hipHostMallocDefault) != hipSuccess)
        return fail(FFX_ENOMEM);
    std::memset(ctx->ctrl_host, 0, sizeof(Ctrl));
    if (hipEventCreate(&ctx->ev0) != hipSuccess || hipEventCreate(&ctx->ev1) != hipSuccess) return fail(FFX_EHIP);
    *out = ctx;
    return FFX_OK;
}

int ffx_ctx_destroy(ffx_ctx* ctx) {
    if (!ctx) return FFX_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->aux) ffx_ctx_destroy(ctx->aux);
    free_store(ctx);
    dev_free(ctx->ctrl);
    dev_free(ctx->counts);
    dev_free(ctx->total);
    dev_free(ctx->partial);
    dev_free(ctx->result);
    dev_free(ctx->staging);
    dev_free(ctx->cliff);
    dev_free(ctx->epochs);
    dev_free(ctx->topk_buf);
    dev_free(ctx->batch_pool);
    dev_free(ctx->merged_buf);
    dev_free(ctx->mlayer_q);
    dev_free(ctx->crot);
    dev_free(ctx->claim);
    dev_free(ctx->layer_buf);
    ffx_mlp_free(ctx);
    if (ctx->batch_exec) (void)hipGraphExecDestroy(ctx->batch_exec);
    if (ctx->batch_host) (void)hipHostFree(ctx->batch_host);
    if (ctx->mlayer_host) (void)hipHostFree(ctx->mlayer_host);
    for (hipEvent_t e : ctx->kev) (void)hipEventDestroy(e);
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->ctrl_host) (void)hipHostFree(ctx->ctrl_host);
    if (ctx->result_host) (void)hipHostFree(ctx->result_host);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return FFX_OK;
}

const char* ffx_last_error(const ffx_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

void ffx_terms_free(ffx_terms* t) {
    if (!t) return;
    free(t->x);
    free(t->z);
    free(t->c_re);
    free(t->c_im);
    t->x = t->z = nullptr;
    t->c_re = t->c_im = nullptr;
    t->n_terms = 0;
}

int ffx_set_kernel_timing(ffx_ctx* ctx, int every) {
    if (!ctx) return FFX_EINVAL;
    if (every < 0) return set_err(ctx, FFX_EINVAL, "kernel timing period must be >= 0");
    ctx->kernel_timing = every;
    return FFX_OK;
}

int ffx_profile_get(const ffx_ctx* ctx, ffx_profile* out) {
    if (!ctx || !out) return FFX_EINVAL;
    *out = ctx->prof;
    out->launches = ctx->launches;
    return FFX_OK;
}

}  // extern "C"

namespace {

// Argument checks shared by the Pauli entry points; K = 2 when any imaginary part is nonzero.
int check_pauli(ffx_ctx* ctx, const ffx_pauli_gate* gates, int64_t n_gates, int32_t n_words, const uint64_t* x,
                const uint64_t* z, const double* c_re, const double* c_im, int64_t n_terms, const ffx_options* opts,
                int* K_out) {
    if (!opts || n_gates < 0 || n_terms < 0 || (n_gates > 0 && !gates) ||
        (n_terms > 0 && (!x || !z || !c_re)))
        return set_err(ctx, FFX_EINVAL, "invalid arguments");
    if (n_words < 1 || n_words > FFX_MAX_WORDS) return set_err(ctx, FFX_EINVAL, "n_words must be in 1..4");
    if (n_terms > 0x3FFFFFFF) return set_err(ctx, FFX_ECAPACITY, "too many input terms");
    for (int64_t i = 0; i < n_gates; ++i) {
        const ffx_pauli_gate& g = gates[i];
        if (g.kind < FFX_ROT || g.kind > FFX_CZ) return set_err(ctx, FFX_EINVAL, "unknown gate kind");
        if (g.kind == FFX_ROT) {
            for (int w = n_words; w < FFX_MAX_WORDS; ++w)
                if (g.x[w] || g.z[w]) return set_err(ctx, FFX_EINVAL, "rotation axis exceeds n_words");
        } else {
            const int lim = 64 * n_words;
            const bool two = g.kind >= FFX_CNOT;
            if (g.q0 < 0 || g.q0 >= lim || (two && (g.q1 < 0 || g.q1 >= lim || g.q1 == g.q0)))
                return set_err(ctx, FFX_EINVAL, "Clifford qubit out of range");
        }
    }
    if (hipSetDevice(ctx->device) != hipSuccess) return set_err(ctx, FFX_ENODEV, "hipSetDevice failed");
    int K = 1;
    if (c_im)
        for (int64_t i = 0; i < n_terms; ++i)
            if (c_im[i] != 0.0) {
                K = 2;
                break;
            }
    const int prec = opts->precision;
    if (prec < FFX_FP64 || prec > FFX_FP32C) return set_err(ctx, FFX_EINVAL, "unknown precision");
    *K_out = K;
    return FFX_OK;
}

int check_majorana(ffx_ctx* ctx, const ffx_majorana_gate* gates, int64_t n_gates, int32_t n_words, const uint64_t* m,
                   const double* c_re, const double* c_im, int64_t n_terms, int* K_out) {
    if (n_gates < 0 || n_terms < 0 || (n_gates > 0 && !gates) || (n_terms > 0 && (!m || !c_re)))
        return set_err(ctx, FFX_EINVAL, "invalid arguments");
    if (n_words < 1 || n_words > FFX_MAX_WORDS) return set_err(ctx, FFX_EINVAL, "n_words must be in 1..4");
    if (n_terms > 0x3FFFFFFF) return set_err(ctx, FFX_ECAPACITY, "too many input terms");
    for (int64_t i = 0; i < n_gates; ++i)
        for (int w = n_words; w < FFX_MAX_WORDS; ++w)
            if (gates[i].m[w]) return set_err(ctx, FFX_EINVAL, "rotation axis exceeds n_words");
    if (hipSetDevice(ctx->device) != hipSuccess) return set_err(ctx, FFX_ENODEV, "hipSetDevice failed");
    // normalized coefficients d = c / i^r are real for every term iff K = 1 suffices
    int K = 1;
    for (int64_t i = 0; i < n_terms && K == 1; ++i) {
        const int r = mono_degree_mod4_r(m + (size_t)i * n_words, n_words);
        const double dim = r ? -c_re[i] : (c_im ? c_im[i] : 0.0);
        if (dim != 0.0) K = 2;
    }
    *K_out = K;
    return FFX_OK;
}

int check_prune(ffx_ctx* ctx, const ffx_prune_opts* po) {
    if (!po) return FFX_OK;
    if (!(po->clifford_tol >= 0.0) || po->clifford_tol >= 1.0)
        return set_err(ctx, FFX_EINVAL, "clifford_tol must be in [0, 1)");
    if (po->prune_period < 0) return set_err(ctx, FFX_EINVAL, "prune_period must be >= 0");
    if (po->obs_site < 0 || po->obs_site > 255) return set_err(ctx, FFX_EINVAL, "obs_site must be in 0..255");
    return FFX_OK;
}

int pauli_propagate_impl(ffx_ctx* ctx, const ffx_pauli_gate* gates, int64_t n_gates, int32_t n_words,
                         const uint64_t* x, const uint64_t* z, const double* c_re, const double* c_im, int64_t n_terms,
                         const ffx_options* opts, const ffx_prune_opts* po, ffx_terms* out,
                         const uint64_t* basis_state, double out_expect[2]) {
    if (!ctx) return FFX_EINVAL;
    ctx->err.clear();
    int K = 1;
    CHK(check_pauli(ctx, gates, n_gates, n_words, x, z, c_re, c_im, n_terms, opts, &K));
    CHK(check_prune(ctx, po));
    ctx->prof = ffx_profile{};
    ctx->launches = 0;
    const int prec = opts->precision;
    switch (n_words) {
        case 1: return dispatch_pauli<1>(ctx, gates, n_gates, x, z, c_re, c_im, n_terms, *opts, out, basis_state, out_expect, prec, K, po);
        case 2: return dispatch_pauli<2>(ctx, gates, n_gates, x, z, c_re, c_im, n_terms, *opts, out, basis_state, out_expect, prec, K, po);
        case 3: return dispatch_pauli<3>(ctx, gates, n_gates, x, z, c_re, c_im, n_terms, *opts, out, basis_state, out_expect, prec, K, po);
        default: return dispatch_pauli<4>(ctx, gates, n_gates, x, z, c_re, c_im, n_terms, *opts, out, basis_state, out_expect, prec, K, po);
    }
}

int majorana_propagate_impl(ffx_ctx* ctx, const ffx_majorana_gate* gates, int64_t n_gates, int32_t n_words,
                            const uint64_t* m, const double* c_re, const double* c_im, int64_t n_terms,
                            int32_t maxdegree, double mincoeff, int32_t precision, const ffx_prune_opts* po,
                            ffx_terms* out, const uint64_t* occ, double out_expect[2]) {
    if (!ctx) return FFX_EINVAL;
    ctx->err.clear();
    int K = 1;
    CHK(check_majorana(ctx, gates, n_gates, n_words, m, c_re, c_im, n_terms, &K));
    CHK(check_prune(ctx, po));
    ctx->prof = ffx_profile{};
    ctx->launches = 0;
    if (precision < FFX_FP64 || precision > FFX_FP32C) return set_err(ctx, FFX_EINVAL, "unknown precision");
    switch (n_words) {
        case 1: return dispatch_majorana<1>(ctx, gates, n_gates, m, c_re, c_im, n_terms, maxdegree, mincoeff, out, occ, out_expect, precision, K, po);
        case 2: return dispatch_majorana<2>(ctx, gates, n_gates, m, c_re, c_im, n_terms, maxdegree, mincoeff, out, occ, out_expect, precision, K, po);
        case 3: return dispatch_majorana<3>(ctx, gates, n_gates, m, c_re, c_im, n_terms, maxdegree, mincoeff, out, occ, out_expect, precision, K, po);
        default: return dispatch_majorana<4>(ctx, gates, n_gates, m, c_re, c_im, n_terms, maxdegree, mincoeff, out, occ, out_expect, precision, K, po);
    }
}

}  // namespace

namespace ffx_host {
int check_pauli_args(ffx_ctx* ctx, const ffx_pauli_gate* gates, int64_t n_gates, int32_t n_words, const uint64_t* x,
                     const uint64_t* z, const double* c_re, const double* c_im, int64_t n_terms,
                     const ffx_options* opts, int* K_out) {
    return check_pauli(ctx, gates, n_gates, n_words, x, z, c_re, c_im, n_terms, opts, K_out);
}
}  // namespace ffx_host

extern "C" {

int ffx_pauli_propagate(ffx_ctx* ctx, const ffx_pauli_gate* gates, int64_t n_gates, int32_t n_words,
                        const uint64_t* x, const uint64_t* z, const double* c_re, const double* c_im,
                        int64_t n_terms, const ffx_options* opts, ffx_terms* out, const uint64_t* basis_state,
                        double out_expect[2]) {
    return pauli_propagate_impl(ctx, gates, n_gates, n_words, x, z, c_re, c_im, n_terms, opts, nullptr, out,
                                basis_state, out_expect);
}

int ffx_pauli_propagate_pruned(ffx_ctx* ctx, const ffx_pauli_gate* gates, int64_t n_gates, int32_t n_words,
                               const uint64_t* x, const uint64_t* z, const double* c_re, const double* c_im,
                               int64_t n_terms, const ffx_options* opts, const ffx_prune_opts* po,
                               ffx_terms* out, const uint64_t* basis_state, double out_expect[2]) {
    return pauli_propagate_impl(ctx, gates, n_gates, n_words, x, z, c_re, c_im, n_terms, opts, po, out, basis_state,
                                out_expect);
}

int ffx_majorana_propagate_pruned(ffx_ctx* ctx, const ffx_majorana_gate* gates, int64_t n_gates, int32_t n_words,
                                  const uint64_t* m, const double* c_re, const double* c_im, int64_t n_terms,
                                  int32_t maxdegree, double mincoeff, int32_t precision, const ffx_prune_opts* po,
                                  ffx_terms* out, const uint64_t* occ, double out_expect[2]) {
    return majorana_propagate_impl(ctx, gates, n_gates, n_words, m, c_re, c_im, n_terms, maxdegree, mincoeff,
                                   precision, po, out, occ, out_expect);
}

int ffx_pauli_contrib_log(ffx_ctx* ctx, const ffx_pauli_gate* gates, int64_t n_gates, int32_t n_words,
                          const uint64_t* x, const uint64_t* z, const double* c_re, const double* c_im,
                          int64_t n_terms, const ffx_options* opts, const ffx_prune_opts* po,
                          const uint64_t* basis_state, int32_t log_every, int64_t max_rows, float* feats,
                          float* targets, int64_t* n_rows, double* sums, int32_t max_points, int32_t* n_points) {
    if (!ctx) return FFX_EINVAL;
    ctx->err.clear();
    if (!po || !sums || max_rows < 0 || (max_rows > 0 && (!feats || !targets)))
        return set_err(ctx, FFX_EINVAL, "invalid arguments");
    int K = 1;
    CHK(check_pauli(ctx, gates, n_gates, n_words, x, z, c_re, c_im, n_terms, opts, &K));
    CHK(check_prune(ctx, po));
    ctx->prof = ffx_profile{};
    ctx->launches = 0;
    switch (n_words) {
        case 1: return dispatch_log_pauli<1>(ctx, gates, n_gates, x, z, c_re, c_im, n_terms, *opts, *po, basis_state, log_every, max_rows, feats, targets, n_rows, sums, max_points, n_points, K);
        case 2: return dispatch_log_pauli<2>(ctx, gates, n_gates, x, z, c_re, c_im, n_terms, *opts, *po, basis_state, log_every, max_rows, feats, targets, n_rows, sums, max_points, n_points, K);
        case 3: return dispatch_log_pauli<3>(ctx, gates, n_gates, x, z, c_re, c_im, n_terms, *opts, *po, basis_state, log_every, max_rows, feats, targets, n_rows, sums, max_points, n_points, K);
        default: return dispatch_log_pauli<4>(ctx, gates, n_gates, x, z, c_re, c_im, n_terms, *opts, *po, basis_state, log_every, max_rows, feats, targets, n_rows, sums, max_points, n_points, K);
    }
}

int ffx_majorana_contrib_log(ffx_ctx* ctx, const ffx_majorana_gate* gates, int64_t n_gates, int32_t n_words,
                             const uint64_t* m, const double* c_re, const double* c_im, int64_t n_terms,
                             int32_t maxdegree, double mincoeff, const ffx_prune_opts* po, const uint64_t* occ,
                             int32_t log_every, int64_t max_rows, float* feats, float* targets, int64_t* n_rows,
                             double* sums, int32_t max_points, int32_t* n_points) {
    if (!ctx) return FFX_EINVAL;
    ctx->err.clear();
    if (!po || !sums || max_rows < 0 || (max_rows > 0 && (!feats || !targets)))
        return set_err(ctx, FFX_EINVAL, "invalid arguments");
    int K = 1;
    CHK(check_majorana(ctx, gates, n_gates, n_words, m, c_re, c_im, n_terms, &K));
    CHK(check_prune(ctx, po));
    ctx->prof = ffx_profile{};
    ctx->launches = 0;
    switch (n_words) {
        case 1: return dispatch_log_majorana<1>(ctx, gates, n_gates, m, c_re, c_im, n_terms, maxdegree, mincoeff, *po, occ, log_every, max_rows, feats, targets, n_rows, sums, max_points, n_points, K);
        case 2: return dispatch_log_majorana<2>(ctx, gates, n_gates, m, c_re, c_im, n_terms, maxdegree, mincoeff, *po, occ, log_every, max_rows, feats, targets, n_rows, sums, max_points, n_points, K);
        case 3: return dispatch_log_majorana<3>(ctx, gates, n_gates, m, c_re, c_im, n_terms, maxdegree, mincoeff, *po, occ, log_every, max_rows, feats, targets, n_rows, sums, max_points, n_points, K);
        default: return dispatch_log_majorana<4>(ctx, gates, n_gates, m, c_re, c_im, n_terms, maxdegree, mincoeff, *po, occ, log_every, max_rows, feats, targets, n_rows, sums, max_points, n_points, K);
    }
}

}  // extern "C"

namespace {

int check_batch(ffx_ctx* ctx, const ffx_instance* inst, int32_t n_inst, int32_t n_words, const ffx_options* opts,
                int* K_out) {
    if (!opts || n_inst < 0 || (n_inst > 0 && !inst)) return set_err(ctx, FFX_EINVAL, "invalid arguments");
    if (n_words < 1 || n_words > FFX_MAX_WORDS) return set_err(ctx, FFX_EINVAL, "n_words must be in 1..4");
    int K = 1;
    for (int32_t i = 0; i < n_inst; ++i) {
        const ffx_instance& in = inst[i];
        if (in.n_gates < 0 || in.n_terms < 0 || (in.n_gates > 0 && !in.gates) ||
            (in.n_terms > 0 && (!in.x || !in.z || !in.c_re)))
            return set_err(ctx, FFX_EINVAL, "invalid instance " + std::to_string(i));
        if (in.n_terms > 0x3FFFFFFF) return set_err(ctx, FFX_ECAPACITY, "too many input terms");
        for (int64_t g = 0; g < in.n_gates; ++g) {
            const ffx_pauli_gate& gt = in.gates[g];
            if (gt.kind < FFX_ROT || gt.kind > FFX_CZ) return set_err(ctx, FFX_EINVAL, "unknown gate kind");
            if (gt.kind == FFX_ROT) {
                for (int w = n_words; w < FFX_MAX_WORDS; ++w)
                    if (gt.x[w] || gt.z[w]) return set_err(ctx, FFX_EINVAL, "rotation axis exceeds n_words");
            } else {
                const int lim = 64 * n_words;
                const bool two = gt.kind >= FFX_CNOT;
                if (gt.q0 < 0 || gt.q0 >= lim || (two && (gt.q1 < 0 || gt.q1 >= lim || gt.q1 == gt.q0)))
                    return set_err(ctx, FFX_EINVAL, "Clifford qubit out of range");
            }
        }
        if (in.c_im)
            for (int64_t t = 0; t < in.n_terms && K == 1; ++t)
                if (in.c_im[t] != 0.0) K = 2;
    }
    if (opts->precision < FFX_FP64 || opts->precision > FFX_FP32C)
        return set_err(ctx, FFX_EINVAL, "unknown precision");
    if (hipSetDevice(ctx->device) != hipSuccess) return set_err(ctx, FFX_ENODEV, "hipSetDevice failed");
    *K_out = K;
    return FFX_OK;
}

int batch_impl(ffx_ctx* ctx, const ffx_instance* inst, int32_t n_inst, int32_t n_words, const ffx_options& o,
               int64_t slots, double* out_expect, int64_t* out_terms, int32_t* out_fallback, int K,
               const BatchConf* conf) {
    const int prec = o.precision;
    switch (n_words) {
        case 1: return dispatch_batch<1>(ctx, inst, n_inst, o, slots, out_expect, out_terms, out_fallback, prec, K, conf);
        case 2: return dispatch_batch<2>(ctx, inst, n_inst, o, slots, out_expect, out_terms, out_fallback, prec, K, conf);
        case 3: return dispatch_batch<3>(ctx, inst, n_inst, o, slots, out_expect, out_terms, out_fallback, prec, K, conf);
        default: return dispatch_batch<4>(ctx, inst, n_inst, o, slots, out_expect, out_terms, out_fallback, prec, K, conf);
    }
}

}  // namespace

extern "C" {

int ffx_pauli_propagate_batch(ffx_ctx* ctx, const ffx_instance* inst, int32_t n_inst, int32_t n_words,
                              const ffx_options* opts, int64_t slots, double* out_expect, int64_t* out_terms,
                              int32_t* out_fallback) {
    if (!ctx) return FFX_EINVAL;
    ctx->err.clear();
    if (slots < 0 || (n_inst > 0 && !out_expect)) return set_err(ctx, FFX_EINVAL, "invalid arguments");
    int K = 1;
    CHK(check_batch(ctx, inst, n_inst, n_words, opts, &K));
    ctx->prof = ffx_profile{};
    ctx->launches = 0;
    if (out_fallback) *out_fallback = 0;
    if (n_inst == 0) return FFX_OK;
    return batch_impl(ctx, inst, n_inst, n_words, *opts, slots, out_expect, out_terms, out_fallback, K, nullptr);
}

int ffx_batch_autotune(ffx_ctx* ctx, const ffx_instance* inst, int32_t n_inst, int32_t n_words,
                       const ffx_options* opts, int32_t sample, ffx_batch_tuning* out) {
    if (!ctx) return FFX_EINVAL;
    ctx->err.clear();
    if (n_inst <= 0 || !out) return set_err(ctx, FFX_EINVAL, "invalid arguments");
    int K = 1;
    CHK(check_batch(ctx, inst, n_inst, n_words, opts, &K));
    // timed batch: all instances (sample <= 0), or the sample with the longest circuits
    std::vector<int32_t> order((size_t)n_inst);
    for (int32_t i = 0; i < n_inst; ++i) order[(size_t)i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](int32_t a, int32_t b) { return inst[a].n_gates > inst[b].n_gates; });
    const int32_t m = sample > 0 ? std::min<int32_t>(n_inst, sample) : n_inst;
    std::vector<ffx_instance> sub((size_t)m);
    for (int32_t i = 0; i < m; ++i) sub[(size_t)i] = inst[order[(size_t)i]];
    std::vector<double> ex(2 * (size_t)m);
    const int blocks[2] = {256, 512};
    const int windows[2] = {3, 4};
    const int64_t budgets[2] = {8192, 32768};
    // candidates: the layer-group layout per compaction interval, general group size and
    // number of term stores cut by schedule length (1, 2 or 4, each on its own stream),
    // the window-step layout, then the one-workgroup layout
    std::vector<BatchConf> cands;
    for (int ce : {2, 4})
        for (int gm : {6, 8})
            for (int sp : {1, 2, 4}) cands.push_back(BatchConf{kBlockBatch, kWin, 32768, 1, ce, gm, sp});
    cands.push_back(BatchConf{kBlockBatch, kWin, 32768, 2, 0, 0, 1});
    for (int bl : blocks)
        for (int wi : windows)
            for (int64_t sl : budgets) cands.push_back(BatchConf{bl, wi, sl, 0, 0, 0, 1});
    double best_ms = 1e300;
    BatchConf best{};
    int32_t tried = 0;
    bool best_clean = false;
    for (const BatchConf& c : cands) {
        int32_t fb = 0;
        // first run builds (and captures); the second is timed
        CHK(batch_impl(ctx, sub.data(), m, n_words, *opts, 0, ex.data(), nullptr, &fb, K, &c));
        const double t0 = now_s();
        CHK(batch_impl(ctx, sub.data(), m, n_words, *opts, 0, ex.data(), nullptr, &fb, K, &c));
        const double ms = 1e3 * (now_s() - t0);
        ++tried;
        const bool clean = fb == 0;
        if ((clean && !best_clean) || (clean == best_clean && ms < best_ms)) {
            best_ms = ms;
            best = c;
            best_clean = clean;
        }
    }
    ffx_ctx::BatchCfg t;
    t.W = n_words;
    t.prec = opts->precision;
    t.K = K;
    t.block = best.block;
    t.window = best.window;
    t.slots = best.slots;
    t.merged = best.merged;
    t.compact_every = best.compact_every;
    t.group_max = best.group_max;
    t.split = best.split;
    bool replaced = false;
    for (auto& e : ctx->batch_tuned)
        if (e.W == t.W && e.prec == t.prec && e.K == t.K) {
            e = t;
            replaced = true;
        }
    if (!replaced) ctx->batch_tuned.push_back(t);
    out->block = best.block;
    out->window = best.window;
    out->slots = best.slots;
    out->merged = best.merged;
    out->compact_every = best.compact_every;
    out->group_max = best.group_max;
    out->split = best.split;
    out->ms = best_ms;
    out->candidates = tried;
    out->sample = m;
    ctx->prof = ffx_profile{};
    return FFX_OK;
}

int ffx_batch_graph_replays(const ffx_ctx* ctx, int64_t* out) {
    if (!ctx || !out) return FFX_EINVAL;
    *out = ctx->batch_graph_launches;
    return FFX_OK;
}

int ffx_majorana_propagate(ffx_ctx* ctx, const ffx_majorana_gate* gates, int64_t n_gates, int32_t n_words,
                           const uint64_t* m, const double* c_re, const double* c_im, int64_t n_terms,
                           int32_t maxdegree, double mincoeff, int32_t precision, ffx_terms* out,
                           const uint64_t* occ, double out_expect[2]) {
    return majorana_propagate_impl(ctx, gates, n_gates, n_words, m, c_re, c_im, n_terms, maxdegree, mincoeff,
                                   precision, nullptr, out, occ, out_expect);
}

int ffx_expect_last(ffx_ctx* ctx, const uint64_t* state, double out[2]) {
    if (!ctx || !out) return FFX_EINVAL;
    ctx->err.clear();
    if (ctx->last_kind == 0) return set_err(ctx, FFX_EINVAL, "no device-resident result");
    if (hipSetDevice(ctx->device) != hipSuccess) return set_err(ctx, FFX_ENODEV, "hipSetDevice failed");
    const int W = ctx->last_W, K = ctx->last_K, prec = ctx->last_prec;
    switch (W) {
        case 1: return dispatch_expect<1>(ctx, state, out, ctx->last_kind, prec, K);
        case 2: return dispatch_expect<2>(ctx, state, out, ctx->last_kind, prec, K);
        case 3: return dispatch_expect<3>(ctx, state, out, ctx->last_kind, prec, K);
        default: return dispatch_expect<4>(ctx, state, out, ctx->last_kind, prec, K);
    }
}

}  // extern "C"
