// ffx_engine.h -- host engine of libffamd.so: context, propagation drivers and
// their precision dispatch.  Included by ffx_lib.hip (C ABI) and by one
// translation unit per key width (ffx_w1.hip .. ffx_w4.hip), which instantiate
// the drivers so that the kernels compile in parallel.
//
// The host side restates the control flow of the reference drivers --
// reverse walk of the circuit, Clifford blocks, greedy commuting batches,
// truncation points (src/pauli/propagate.h:297-366) and the per-gate threshold of
// the Majorana driver (src/majorana_propagation.h:18-33) -- as a list of device
// steps, then runs every step as HIP kernels on one stream.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <type_traits>
#include <unordered_map>
#include <utility>
#include <thread>
#include <vector>

#include "../../include/fastfermion_amd.h"
#include "ffx_batch.h"
#include "ffx_layer.h"
#include "ffx_prune.h"
#include "ffx_dlayer.h"
#include "ffx_merged.h"
#include "ffx_mlayer.h"

using namespace ffx;

// Device state of a context's pruner model (allocated on first use).
struct MlpState {
    float* p = nullptr;    // fp32 master parameters (FFX_PRUNER_NPARAMS)
    float* g = nullptr;    // gradient
    float* m1 = nullptr;   // Adam moments
    float* m2 = nullptr;
    uint16_t* w1t = nullptr;  // bf16 W1^T [128][64]
    uint16_t* w2t = nullptr;  // bf16 W2^T [128][128]
    uint16_t* w2 = nullptr;   // bf16 W2 [128][128]
    float* w3r = nullptr;     // w3 rounded to bf16
    double* loss = nullptr;   // accumulated squared error
    int64_t step = 0;         // Adam steps taken
    bool ready = false;       // parameters set
    // training data and minibatch buffers
    uint16_t* X = nullptr;
    float* T = nullptr;
    size_t data_cap = 0;
    int32_t* idx = nullptr;
    size_t idx_cap = 0;
    uint16_t *H1 = nullptr, *DZ1 = nullptr, *DZ2 = nullptr;
    float* Y = nullptr;
    size_t mb_cap = 0;
    MlpDev dev() const { return MlpDev{w1t, w2t, w2, w3r, p + kOffB1, p + kOffB2, p + kOffB3}; }
};

struct ffx_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    // second context (own stream and buffers) for the part of a merged batch that runs
    // beside this one; created by the batch on first use, destroyed with this one
    ffx_ctx* aux = nullptr;
    std::string err;
    // term store
    uint64_t* key = nullptr;
    uint64_t* key2 = nullptr;
    void* coef = nullptr;
    void* coef2 = nullptr;
    uint64_t* index = nullptr;
    size_t cap = 0;      // slots
    size_t hslots = 0;   // index entries (power of two, >= 2*cap)
    size_t key_planes = 0, coef_planes = 0, coef_elem = 0;  // layout of the current store (words per key record,
                                                             // coefficient planes, bytes per coefficient word)
    Ctrl* ctrl = nullptr;
    Ctrl* ctrl_host = nullptr;  // pinned
    uint32_t* counts = nullptr;
    size_t counts_cap = 0;
    uint32_t* total = nullptr;
    double* partial = nullptr;
    double* result = nullptr;
    double* result_host = nullptr;  // pinned, 2 doubles
    void* staging = nullptr;
    size_t staging_bytes = 0;
    CliffGate* cliff = nullptr;
    size_t cliff_cap = 0;
    uint32_t* epochs = nullptr;
    uint32_t* topk_buf = nullptr;  // top-k select scratch: 256 bins, magnitude prefix, rank, hash prefix
    void* batch_pool = nullptr;    // batched instances: term stores, schedules, states, inputs
    size_t batch_pool_bytes = 0;
    void* merged_buf = nullptr;    // merged batch: windows or groups, schedule offsets, states, results
    size_t merged_buf_bytes = 0;
    void* mlayer_host = nullptr;   // merged batch in layer groups: pinned image of the device buffer
    size_t mlayer_host_bytes = 0;
    void* mlayer_q = nullptr;      // merged batch in layer groups: orbit leader lists, their counts
    size_t mlayer_q_bytes = 0;
    size_t mlayer_cap_hint = 0;    // merged batch in layer groups: store capacity of the last call
    size_t epochs_cap = 0;
    int n_cu = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int kernel_timing = 0;   // 0 off, N: time every N-th propagation launch
    uint64_t launch_seq = 0;  // propagation launches issued on this context
    std::vector<hipEvent_t> kev;  // event pairs around rotation launches
    // description of the device-resident result
    int last_kind = 0;  // 0 none, 1 Pauli, 2 Majorana
    int last_W = 0, last_K = 0, last_prec = 0;
    ffx_profile prof{};
    int64_t launches = 0;
    MlpState* mlp = nullptr;  // pruner model (ffx_pruner_*)
    // batched propagation: autotuned configurations, pinned host image, captured graph
    struct BatchCfg {
        int W = 0, prec = 0, K = 0;
        int block = 0, window = 0;
        int64_t slots = 0;
        int merged = 0;
        int compact_every = 0, group_max = 0;
        int split = -1;
    };
    std::vector<BatchCfg> batch_tuned;
    void* batch_host = nullptr;  // pinned
    size_t batch_host_bytes = 0;
    hipGraphExec_t batch_exec = nullptr;
    size_t batch_key[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // layout the captured graph was built for
    int64_t batch_graph_launches = 0;  // replays of a captured graph (profile)
    void* crot = nullptr;     // Clifford-rotation records of the pruned mode
    size_t crot_bytes = 0;
    uint64_t* claim = nullptr;  // orbit claim table of the layer kernel (hslots entries)
    size_t claim_slots = 0;
    uint32_t layer_stamp = 0;   // stamp of the last layer launch (24 bits)
    void* layer_buf = nullptr;  // layer records, qubit-position tables, per-rotation counters
    size_t layer_buf_bytes = 0;
    // allocation-free calls (ffx_pauli_propagate_noalloc): while `on`, device blocks are
    // carved from [base, base + size) of the device's fixed scratch array instead of
    // hipMalloc, and the term store has exactly `cap` slots
    struct Arena {
        char* base = nullptr;
        size_t size = 0, used = 0, cap = 0;
        bool on = false;
    } arena;
    // ... and the propagated terms go to caller-owned host arrays of `cap` terms
    struct Sink {
        uint64_t *x = nullptr, *z = nullptr;
        double *re = nullptr, *im = nullptr;
        int64_t cap = 0, n = 0;
        bool on = false;
    } sink;
};

// Releases the pruner model of a context (defined with the pruner kernels).
void ffx_mlp_free(ffx_ctx* ctx);

namespace ffx_host {

#define HIPCHK(expr)                                                                                   \
    do {                                                                                               \
        hipError_t _e = (expr);                                                                        \
        if (_e != hipSuccess) {                                                                        \
            ctx->err = std::string("HIP error: ") + hipGetErrorString(_e) + " at " #expr;              \
            return _e == hipErrorOutOfMemory ? FFX_ENOMEM : FFX_EHIP;                                  \
        }                                                                                              \
    } while (0)

#define CHK(expr)                   \
    do {                            \
        int _r = (expr);            \
        if (_r != FFX_OK) return _r; \
    } while (0)

inline double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

inline int set_err(ffx_ctx* ctx, int code, const std::string& msg) {
    ctx->err = msg;
    return code;
}

inline size_t pow2ceil(size_t v) {
    size_t p = 1;
    while (p < v) p <<= 1;
    return p;
}

// Device ranges of the fixed scratch arrays of allocation-free calls, one per
// device, registered when a device's array is first used; dev_free never
// releases memory inside them.
constexpr int kMaxDevices = 64;
inline std::atomic<uintptr_t> g_fixed_lo[kMaxDevices] = {};
inline std::atomic<uintptr_t> g_fixed_hi[kMaxDevices] = {};

inline bool in_fixed_scratch(const void* p) {
    const uintptr_t a = (uintptr_t)p;
    for (int d = 0; d < kMaxDevices; ++d)
        if (a >= g_fixed_lo[d].load(std::memory_order_relaxed) && a < g_fixed_hi[d].load(std::memory_order_relaxed))
            return true;
    return false;
}

template <class T>
int dev_alloc(ffx_ctx* ctx, T** p, size_t bytes) {
    *p = nullptr;
    if (bytes == 0) bytes = 256;
    if (ctx->arena.on) {
        const size_t off = (ctx->arena.used + 255) & ~(size_t)255;
        if (off > ctx->arena.size || bytes > ctx->arena.size - off)
            return set_err(ctx, FFX_ECAPACITY,
                           "allocation-free call: its scratch of " + std::to_string(ctx->arena.size) +
                               " bytes cannot hold another block of " + std::to_string(bytes) + " bytes");
        *p = (T*)(ctx->arena.base + off);
        ctx->arena.used = off + bytes;
        return FFX_OK;
    }
    hipError_t e = hipMalloc((void**)p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        return set_err(ctx, FFX_ENOMEM, "device allocation of " + std::to_string(bytes) + " bytes failed");
    }
    return FFX_OK;
}

template <class T>
void dev_free(T*& p) {
    if (p && !in_fixed_scratch((const void*)p)) (void)hipFree((void*)p);
    p = nullptr;
}

inline void free_store(ffx_ctx* ctx) {
    dev_free(ctx->key);
    dev_free(ctx->key2);
    dev_free(ctx->coef);
    dev_free(ctx->coef2);
    dev_free(ctx->index);
    ctx->cap = ctx->hslots = 0;
}

// Partner degree bound of the kernels: the caller's maxdegree compared literally
// (src/pauli/propagate.h:82, src/majorana_propagation.h:43; a negative bound drops
// every partner); FFX_NO_DEGREE_LIMIT and other values above any key's degree
// become 2^30, which no degree reaches.
inline int degree_limit(int32_t maxdegree) { return maxdegree > (1 << 30) ? (1 << 30) : maxdegree; }

// The device buffers a context allocates on demand (term store, counters,
// staging, schedule tables).  An allocation-free call swaps an empty set in,
// fills it from the scratch array, and swaps the context's own set back.
struct DevBuffers {
    uint64_t *key = nullptr, *key2 = nullptr, *index = nullptr;
    void *coef = nullptr, *coef2 = nullptr;
    size_t cap = 0, hslots = 0, key_planes = 0, coef_planes = 0, coef_elem = 0;
    Ctrl* ctrl = nullptr;
    uint32_t *counts = nullptr, *total = nullptr, *epochs = nullptr, *topk_buf = nullptr;
    size_t counts_cap = 0, epochs_cap = 0;
    double *partial = nullptr, *result = nullptr;
    void* staging = nullptr;
    size_t staging_bytes = 0;
    CliffGate* cliff = nullptr;
    size_t cliff_cap = 0;
    void* crot = nullptr;
    size_t crot_bytes = 0;
    uint64_t* claim = nullptr;
    size_t claim_slots = 0;
    uint32_t layer_stamp = 0;
    void* layer_buf = nullptr;
    size_t layer_buf_bytes = 0;
};

inline void swap_buffers(ffx_ctx* c, DevBuffers& b) {
    std::swap(c->key, b.key);
    std::swap(c->key2, b.key2);
    std::swap(c->index, b.index);
    std::swap(c->coef, b.coef);
    std::swap(c->coef2, b.coef2);
    std::swap(c->cap, b.cap);
    std::swap(c->hslots, b.hslots);
    std::swap(c->key_planes, b.key_planes);
    std::swap(c->coef_planes, b.coef_planes);
    std::swap(c->coef_elem, b.coef_elem);
    std::swap(c->ctrl, b.ctrl);
    std::swap(c->counts, b.counts);
    std::swap(c->total, b.total);
    std::swap(c->epochs, b.epochs);
    std::swap(c->topk_buf, b.topk_buf);
    std::swap(c->counts_cap, b.counts_cap);
    std::swap(c->epochs_cap, b.epochs_cap);
    std::swap(c->partial, b.partial);
    std::swap(c->result, b.result);
    std::swap(c->staging, b.staging);
    std::swap(c->staging_bytes, b.staging_bytes);
    std::swap(c->cliff, b.cliff);
    std::swap(c->cliff_cap, b.cliff_cap);
    std::swap(c->crot, b.crot);
    std::swap(c->crot_bytes, b.crot_bytes);
    std::swap(c->claim, b.claim);
    std::swap(c->claim_slots, b.claim_slots);
    std::swap(c->layer_stamp, b.layer_stamp);
    std::swap(c->layer_buf, b.layer_buf);
    std::swap(c->layer_buf_bytes, b.layer_buf_bytes);
}

inline unsigned grid_for(size_t n, size_t per_block, unsigned max_grid = 65535) {
    size_t g = (n + per_block - 1) / per_block;
    if (g < 1) g = 1;
    if (g > max_grid) g = max_grid;
    return (unsigned)g;
}

// One device step of a propagation.
struct Step {
    int kind = 0;        // 0 rotation, 1 Clifford block
    int64_t gate = 0;    // rotation: gate index
    int trunc = 0;       // rotation: truncate_all runs after it
    int rc = 0;          // rotation: rot_count argument of that truncate_all
    int64_t cb = 0, ce = 0;  // Clifford block [cb, ce)
};

// With truncation after every rotation, the device-decided compaction sequence
// (five launches that return at once when a quarter of the slots is not erased)
// is enqueued after every compact_every()-th truncation point (FFX_COMPACT_EVERY,
// default 256).  6x6 Fermi-Hubbard propagation (6960 rotations): checks every 16
// / 64 / 256 rotations run 16 compactions each and take 34.4 / 31.0 / 29.9 ms;
// the slots erased meanwhile cost less than the no-op checks.
inline int compact_every() {
    int k = 256;
    if (const char* ev = getenv("FFX_COMPACT_EVERY"))
        if (atoi(ev) > 0) k = atoi(ev);
    return k;
}

template <int W>
bool rot_commute(const ffx_pauli_gate& a, const ffx_pauli_gate& b) {
    int p1 = 0, p2 = 0;
    for (int w = 0; w < W; ++w) {
        p1 += __builtin_popcountll(a.x[w] & b.z[w]);
        p2 += __builtin_popcountll(a.z[w] & b.x[w]);
    }
    return (p1 & 1) == (p2 & 1);
}

// Schedule of src/pauli/propagate.h:318-366 (serial driver): the circuit is
// walked from the last gate to the first; runs of Clifford gates form blocks;
// rotations either truncate after every gate (batched = false) or are grouped by
// batch_commuting_gates (propagate.h:272-291) and truncate once per batch.
template <int W>
std::vector<Step> pauli_schedule(const ffx_pauli_gate* g, int64_t n, bool batched) {
    std::vector<Step> steps;
    int rot_count = 0;
    bool pending = false;
    int64_t cend = 0;
    std::vector<int64_t> buf;
    // rotations of the growing batch by qubit: a rotation can only fail to commute
    // with batch members that share a qubit with it, so it is checked against those
    // (the same batches as checking every member, without the quadratic cost)
    std::vector<std::vector<int32_t>> on_q(64 * W);
    std::vector<int> touched;
    auto support = [&](const ffx_pauli_gate& r, auto&& f) {
        for (int w = 0; w < W; ++w) {
            uint64_t m = r.x[w] | r.z[w];
            while (m) {
                f(64 * w + __builtin_ctzll(m));
                m &= m - 1;
            }
        }
    };
    auto flush = [&]() {
        if (buf.empty()) return;
        size_t s = 0;
        while (s < buf.size()) {
            for (int q : touched) on_q[q].clear();
            touched.clear();
            auto add = [&](size_t t) {
                support(g[buf[t]], [&](int q) {
                    if (on_q[q].empty()) touched.push_back(q);
                    on_q[q].push_back((int32_t)t);
                });
            };
            add(s);
            size_t e = s + 1;
            while (e < buf.size()) {
                bool ok = true;
                support(g[buf[e]], [&](int q) {
                    if (!ok) return;
                    for (int32_t t : on_q[q])
                        if (!rot_commute<W>(g[buf[e]], g[buf[t]])) {
                            ok = false;
                            return;
                        }
                });
                if (!ok) break;
                add(e);
                ++e;
            }
            for (size_t t = s; t < e; ++t) {
                Step st;
                st.kind = 0;
                st.gate = buf[t];
                steps.push_back(st);
            }
            rot_count += (int)(e - s);
            steps.back().trunc = 1;
            steps.back().rc = rot_count - 1;
            s = e;
        }
        buf.clear();
    };
    for (int64_t i = n - 1; i >= 0; --i) {
        if (g[i].kind != FFX_ROT) {
            if (batched) flush();
            if (!pending) {
                cend = i;
                pending = true;
            }
        } else {
            if (pending) {
                Step c;
                c.kind = 1;
                c.cb = i + 1;
                c.ce = cend + 1;
                steps.push_back(c);
                pending = false;
            }
            if (batched)
                buf.push_back(i);
            else {
                Step st;
                st.kind = 0;
                st.gate = i;
                st.trunc = 1;
                st.rc = rot_count;
                steps.push_back(st);
                rot_count++;
            }
        }
    }
    if (batched) flush();
    if (pending) {
        Step c;
        c.kind = 1;
        c.cb = 0;
        c.ce = cend + 1;
        steps.push_back(c);
    }
    return steps;
}

// Gate factors: one glibc sincos() call per angle, which is what GCC -O2 makes of
// the reference's cos(theta); sin(theta) pair (src/pauli/propagate.h:76-77,
// src/majorana_gates.h:86-88).  sincos and a separate sin can differ by one ulp.
inline void gate_trig(double theta, double& c, double& s) { ::sincos(theta, &s, &c); }

template <class CP>
typename CP::fac_t make_fac(double v);
template <>
inline F64::fac_t make_fac<F64>(double v) {
    return v;
}
template <>
inline F32::fac_t make_fac<F32>(double v) {
    return (float)v;
}
template <>
inline F32C::fac_t make_fac<F32C>(double v) {
    float h = (float)v;
    return F32C::fac_t{h, (float)(v - (double)h)};
}

// Engine over one term store; Alg = PauliAlg<W> or MajoranaAlg<W>.
template <class Alg, class CP, int K>
struct Engine {
    static constexpr int NP = Alg::NP;
    using store_t = typename CP::store_t;
    ffx_ctx* ctx;
    uint32_t n_host = 0;     // slots in use (exact after sync)
    uint32_t dead_host = 0;  // erased slots (exact after sync)

    explicit Engine(ffx_ctx* c) : ctx(c) {}

    Store<CP, K> store() const {
        Store<CP, K> s;
        s.key = ctx->key;
        s.coef = (store_t*)ctx->coef;
        s.coef2 = (store_t*)ctx->coef2;
        s.index = ctx->index;
        s.hmask = ctx->hslots - 1;
        s.cap = ctx->cap;
        s.ctrl = ctx->ctrl;
        return s;
    }

    size_t coef_planes() const { return (size_t)K * CP::NS; }

    int launch_check() {
        ctx->launches++;
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return set_err(ctx, FFX_EHIP, std::string("kernel launch: ") + hipGetErrorString(e));
        return FFX_OK;
    }

    // Index over slots [0, n_host).  The control block's n / n_dead are set from
    // the host values first: every later kernel of the step-wise paths reads them.
    int rebuild_index() {
        if (n_host > ctx->cap) return set_err(ctx, FFX_EHIP, "internal error: slot count exceeds capacity");
        HIPCHK(hipMemcpyAsync(&ctx->ctrl->n, &n_host, sizeof(uint32_t), hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(hipMemcpyAsync(&ctx->ctrl->n_dead, &dead_host, sizeof(uint32_t), hipMemcpyHostToDevice, ctx->stream));
        hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(ctx->hslots, kBlock, 8192)), dim3(kBlock), 0, ctx->stream,
                           ctx->index, ctx->hslots, kEmpty);
        CHK(launch_check());
        hipLaunchKernelGGL(k_index_build<NP>, dim3(grid_for(n_host, kBlock, 16384)), dim3(kBlock), 0, ctx->stream,
                           ctx->key, ctx->cap, ctx->index, ctx->hslots - 1, ctx->ctrl, n_host);
        return launch_check();
    }

    // Make room for `need` slots; existing slots [0, n_host) are preserved.
    // Store of at least `need` slots.  With `shrink`, a store that an earlier, larger
    // run left far larger than `need` is reallocated at this size: kernels that sweep
    // the whole index (its rebuild after every compaction) cost in proportion to the
    // capacity.
    int reserve(size_t need, bool shrink = false) {
        if (need > 0x7FFFFFF0ull) return set_err(ctx, FFX_ECAPACITY, "term count exceeds 2^31 slots");
        size_t ncap = pow2ceil(std::max<size_t>(need, 1024));
        const bool oversized = shrink && !ctx->arena.on && ctx->cap >= ((size_t)1 << 22) && ctx->cap >= 16 * ncap;
        const bool same_layout =
            ctx->key_planes == NP && ctx->coef_planes == coef_planes() && ctx->coef_elem == sizeof(store_t);
        if (need <= ctx->cap && same_layout && !oversized) return FFX_OK;
        // growth of the same layout doubles (amortised overflow restarts); a layout change
        // (key width, precision, real/complex) keeps the capacity the context already had
        if (!oversized) {
            if (same_layout && ncap < 2 * ctx->cap) ncap = std::max<size_t>(ncap, 2 * ctx->cap);
            if (!same_layout) ncap = std::max<size_t>(ncap, ctx->cap);
        }
        if (ctx->arena.on) {  // one store of the planned size; no growth
            if (need > ctx->arena.cap || ctx->cap != 0)
                return set_err(ctx, FFX_ECAPACITY,
                               "allocation-free call: the term store needs more than the " +
                                   std::to_string(ctx->arena.cap) + " slots its scratch holds");
            ncap = ctx->arena.cap;
        }
        uint64_t *nkey = nullptr, *nkey2 = nullptr, *nindex = nullptr;
        void *ncoef = nullptr, *ncoef2 = nullptr;
        const size_t cbytes = coef_planes() * ncap * sizeof(store_t);
        int rc = dev_alloc(ctx, &nkey, NP * ncap * sizeof(uint64_t));
        if (rc == FFX_OK) rc = dev_alloc(ctx, &nkey2, NP * ncap * sizeof(uint64_t));
        if (rc == FFX_OK) rc = dev_alloc(ctx, &ncoef, cbytes);
        if (rc == FFX_OK) rc = dev_alloc(ctx, &ncoef2, cbytes);
        if (rc == FFX_OK) rc = dev_alloc(ctx, &nindex, 2 * ncap * sizeof(uint64_t));
        if (rc != FFX_OK) {
            dev_free(nkey);
            dev_free(nkey2);
            dev_free(ncoef);
            dev_free(ncoef2);
            dev_free(nindex);
            return rc;
        }
        if (n_host > 0 && ctx->key && ctx->key_planes == NP && ctx->coef_planes == coef_planes()) {
            HIPCHK(hipMemcpyAsync(nkey, ctx->key, n_host * NP * sizeof(uint64_t), hipMemcpyDeviceToDevice,
                                  ctx->stream));
            for (size_t q = 0; q < coef_planes(); ++q)
                HIPCHK(hipMemcpyAsync((store_t*)ncoef + q * ncap, (store_t*)ctx->coef + q * ctx->cap,
                                      n_host * sizeof(store_t), hipMemcpyDeviceToDevice, ctx->stream));
        }
        HIPCHK(hipStreamSynchronize(ctx->stream));
        free_store(ctx);
        ctx->key = nkey;
        ctx->key2 = nkey2;
        ctx->coef = ncoef;
        ctx->coef2 = ncoef2;
        ctx->index = nindex;
        ctx->cap = ncap;
        ctx->hslots = 2 * ncap;
        ctx->key_planes = NP;
        ctx->coef_planes = coef_planes();
        ctx->coef_elem = sizeof(store_t);
        return rebuild_index();
    }

    int ensure_counts(size_t nb) {
        if (nb + 1 <= ctx->counts_cap) return FFX_OK;
        dev_free(ctx->counts);
        size_t c = pow2ceil(nb + 1);
        CHK(dev_alloc(ctx, &ctx->counts, c * sizeof(uint32_t)));
        ctx->counts_cap = c;
        return FFX_OK;
    }

    int sync_ctrl() {
        HIPCHK(hipMemcpyAsync(ctx->ctrl_host, ctx->ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        n_host = ctx->ctrl_host->n;
        dead_host = ctx->ctrl_host->n_dead;
        if (ctx->ctrl_host->overflow) return set_err(ctx, FFX_EHIP, "internal error: term store overflow");
        if (ctx->ctrl_host->index_full)
            return set_err(ctx, FFX_EHIP,
                           "internal error: hash index full (flags " + std::to_string(ctx->ctrl_host->index_full) + ")");
        return FFX_OK;
    }

    // Drop erased slots (order preserving) and rebuild the index.
    int compact() {
        if (n_host == 0) return FFX_OK;
        const size_t nb = (n_host + kTile - 1) / kTile;
        CHK(ensure_counts(nb));
        Store<CP, K> s = store();
        hipLaunchKernelGGL((k_count_live<CP, K>), dim3(nb), dim3(kBlock), 0, ctx->stream, s, ctx->counts);
        CHK(launch_check());
        hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(kBlock), 0, ctx->stream, ctx->counts, (uint32_t)nb,
                           ctx->total);
        CHK(launch_check());
        hipLaunchKernelGGL((k_compact<NP, CP, K>), dim3(nb), dim3(kBlock), 0, ctx->stream, s, ctx->counts,
                           ctx->key2, (store_t*)ctx->coef2);
        CHK(launch_check());
        uint32_t total = 0;
        HIPCHK(hipMemcpyAsync(&total, ctx->total, sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        std::swap(ctx->key, ctx->key2);
        std::swap(ctx->coef, ctx->coef2);
        n_host = total;
        dead_host = 0;
        HIPCHK(hipMemcpyAsync(&ctx->ctrl->n, &n_host, sizeof(uint32_t), hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(hipMemcpyAsync(&ctx->ctrl->n_dead, &dead_host, sizeof(uint32_t), hipMemcpyHostToDevice,
                              ctx->stream));
        return rebuild_index();
    }

    int upload(const uint64_t* hx, const uint64_t* hz, const double* re, const double* im, int64_t n0) {
        n_host = 0;
        dead_host = 0;
        CHK(reserve(std::max<size_t>(2 * (size_t)n0 + 64, 1024)));
        Ctrl c{};
        c.n = (uint32_t)n0;
        HIPCHK(hipMemcpyAsync(ctx->ctrl, &c, sizeof(Ctrl), hipMemcpyHostToDevice, ctx->stream));
        if (n0 > 0) {
            constexpr int W = Alg::W;
            const size_t kbytes = (size_t)n0 * W * sizeof(uint64_t);
            const size_t need = (Alg::kPauli ? 2 : 1) * kbytes + 2 * (size_t)n0 * sizeof(double);
            CHK(ensure_staging(need));
            char* sp = (char*)ctx->staging;
            uint64_t* dx = (uint64_t*)sp;
            uint64_t* dz = Alg::kPauli ? (uint64_t*)(sp + kbytes) : nullptr;
            double* dre = (double*)(sp + (Alg::kPauli ? 2 : 1) * kbytes);
            double* dim = dre + n0;
            HIPCHK(hipMemcpyAsync(dx, hx, kbytes, hipMemcpyHostToDevice, ctx->stream));
            if (Alg::kPauli) HIPCHK(hipMemcpyAsync(dz, hz, kbytes, hipMemcpyHostToDevice, ctx->stream));
            HIPCHK(hipMemcpyAsync(dre, re, n0 * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
            if (im)
                HIPCHK(hipMemcpyAsync(dim, im, n0 * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
            hipLaunchKernelGGL((k_unpack<Alg, CP, K>), dim3(grid_for(n0, kBlock, 16384)), dim3(kBlock), 0,
                               ctx->stream, store(), dx, dz, dre, im ? dim : nullptr, (uint32_t)n0);
            CHK(launch_check());
        }
        n_host = (uint32_t)n0;
        CHK(rebuild_index());
        hipLaunchKernelGGL(k_index_check<NP>, dim3(grid_for(n_host, kBlock, 16384)), dim3(kBlock), 0, ctx->stream,
                           ctx->key, ctx->cap, ctx->index, ctx->hslots - 1, ctx->ctrl);
        CHK(launch_check());
        CHK(sync_ctrl());
        if (ctx->ctrl_host->dup) return set_err(ctx, FFX_EINVAL, "observable has duplicate keys");
        return FFX_OK;
    }

    int ensure_staging(size_t bytes) {
        if (bytes <= ctx->staging_bytes) return FFX_OK;
        dev_free(ctx->staging);
        ctx->staging_bytes = 0;
        CHK(dev_alloc(ctx, &ctx->staging, bytes));
        ctx->staging_bytes = bytes;
        return FFX_OK;
    }

    int rotation(const typename Alg::Gate& g, const RotParams<CP>& rp) {
        CHK(reserve(2 * (size_t)n_host + 64));
        const unsigned grid = grid_for(n_host, kBlock, 16384);
        hipLaunchKernelGGL((k_rot<Alg, CP, K>), dim3(grid), dim3(kBlock), 0, ctx->stream, store(), g, rp);
        CHK(launch_check());
        std::swap(ctx->coef, ctx->coef2);
        return sync_ctrl();
    }

    int clifford(const ffx_pauli_gate* gates, int64_t cb, int64_t ce) {
        if constexpr (Alg::kPauli) {
            const size_t ng = (size_t)(ce - cb);
            if (ng > ctx->cliff_cap) {
                dev_free(ctx->cliff);
                CHK(dev_alloc(ctx, &ctx->cliff, ng * sizeof(CliffGate)));
                ctx->cliff_cap = ng;
            }
            std::vector<CliffGate> h(ng);
            for (size_t t = 0; t < ng; ++t) h[t] = CliffGate{gates[cb + t].kind, gates[cb + t].q0, gates[cb + t].q1, 0};
            HIPCHK(hipMemcpyAsync(ctx->cliff, h.data(), ng * sizeof(CliffGate), hipMemcpyHostToDevice, ctx->stream));
            hipLaunchKernelGGL((k_clifford<Alg::W, CP, K>), dim3(grid_for(n_host, kBlock, 16384)), dim3(kBlock), 0,
                               ctx->stream, store(), ctx->cliff, (int)ng);
            CHK(launch_check());
            CHK(rebuild_index());
            HIPCHK(hipStreamSynchronize(ctx->stream));
        }
        return FFX_OK;
    }

    int expect(const uint64_t* state, double out[2]) {
        StateWords sw{};
        if (state)
            for (int w = 0; w < Alg::W && w < FFX_MAX_WORDS; ++w) sw.s[w] = state[w];
        const unsigned nb = grid_for(n_host, kBlock, 1024);
        hipLaunchKernelGGL((k_expect_partial<Alg, CP, K>), dim3(nb), dim3(kBlock), 0, ctx->stream, store(), sw,
                           ctx->partial);
        CHK(launch_check());
        hipLaunchKernelGGL(k_expect_final, dim3(1), dim3(64), 0, ctx->stream, ctx->partial, (int)nb, ctx->result);
        CHK(launch_check());
        HIPCHK(hipMemcpyAsync(ctx->result_host, ctx->result, 2 * sizeof(double), hipMemcpyDeviceToHost,
                              ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        out[0] = ctx->result_host[0];
        out[1] = ctx->result_host[1];
        return FFX_OK;
    }

    int download(ffx_terms* out) {
        constexpr int W = Alg::W;
        std::memset(out, 0, sizeof(*out));
        out->n_words = W;
        const size_t nb = (n_host + kTile - 1) / kTile;
        uint32_t total = 0;
        if (n_host > 0) {
            CHK(ensure_counts(nb));
            Store<CP, K> s = store();
            hipLaunchKernelGGL((k_count_live<CP, K>), dim3(nb), dim3(kBlock), 0, ctx->stream, s, ctx->counts);
            CHK(launch_check());
            hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(kBlock), 0, ctx->stream, ctx->counts, (uint32_t)nb,
                               ctx->total);
            CHK(launch_check());
            HIPCHK(hipMemcpyAsync(&total, ctx->total, sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
            HIPCHK(hipStreamSynchronize(ctx->stream));
        }
        const size_t nt = total;
        const size_t kb = nt * W * sizeof(uint64_t);
        if (ctx->sink.on) return download_to_sink(nt, nb);
        out->x = (uint64_t*)malloc(std::max<size_t>(kb, 8));
        out->z = Alg::kPauli ? (uint64_t*)malloc(std::max<size_t>(kb, 8)) : nullptr;
        out->c_re = (double*)malloc(std::max<size_t>(nt * sizeof(double), 8));
        out->c_im = (double*)malloc(std::max<size_t>(nt * sizeof(double), 8));
        if (!out->x || (Alg::kPauli && !out->z) || !out->c_re || !out->c_im) {
            ffx_terms_free(out);
            return set_err(ctx, FFX_ENOMEM, "host allocation of the result failed");
        }
        out->n_terms = (int64_t)nt;
        if (nt == 0) return FFX_OK;
        const size_t need = (Alg::kPauli ? 2 : 1) * kb + 2 * nt * sizeof(double);
        CHK(ensure_staging(need));
        char* sp = (char*)ctx->staging;
        uint64_t* ox = (uint64_t*)sp;
        uint64_t* oz = Alg::kPauli ? (uint64_t*)(sp + kb) : nullptr;
        double* ore = (double*)(sp + (Alg::kPauli ? 2 : 1) * kb);
        double* oim = ore + nt;
        hipLaunchKernelGGL((k_pack<Alg, CP, K>), dim3(nb), dim3(kBlock), 0, ctx->stream, store(), ctx->counts, ox, oz,
                           ore, oim);
        CHK(launch_check());
        HIPCHK(hipMemcpyAsync(out->x, ox, kb, hipMemcpyDeviceToHost, ctx->stream));
        if (Alg::kPauli) HIPCHK(hipMemcpyAsync(out->z, oz, kb, hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipMemcpyAsync(out->c_re, ore, nt * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipMemcpyAsync(out->c_im, oim, nt * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        return FFX_OK;
    }

    // Live terms [0, nt) packed into caller-owned host arrays (allocation-free calls);
    // counts holds the per-tile live counts of the download's scan.
    int download_to_sink(size_t nt, size_t nb) {
        constexpr int W = Alg::W;
        ffx_ctx::Sink& sk = ctx->sink;
        sk.n = (int64_t)nt;
        if ((int64_t)nt > sk.cap)
            return set_err(ctx, FFX_ECAPACITY,
                           "the result has " + std::to_string(nt) + " terms; the output arrays hold " +
                               std::to_string(sk.cap));
        if (nt == 0) return FFX_OK;
        const size_t kb = nt * W * sizeof(uint64_t);
        CHK(ensure_staging((Alg::kPauli ? 2 : 1) * kb + 2 * nt * sizeof(double)));
        char* sp = (char*)ctx->staging;
        uint64_t* ox = (uint64_t*)sp;
        uint64_t* oz = Alg::kPauli ? (uint64_t*)(sp + kb) : nullptr;
        double* ore = (double*)(sp + (Alg::kPauli ? 2 : 1) * kb);
        double* oim = ore + nt;
        hipLaunchKernelGGL((k_pack<Alg, CP, K>), dim3(nb), dim3(kBlock), 0, ctx->stream, store(), ctx->counts, ox, oz,
                           ore, oim);
        CHK(launch_check());
        if (sk.x) HIPCHK(hipMemcpyAsync(sk.x, ox, kb, hipMemcpyDeviceToHost, ctx->stream));
        if (Alg::kPauli && sk.z) HIPCHK(hipMemcpyAsync(sk.z, oz, kb, hipMemcpyDeviceToHost, ctx->stream));
        if (sk.re) HIPCHK(hipMemcpyAsync(sk.re, ore, nt * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        if (sk.im) HIPCHK(hipMemcpyAsync(sk.im, oim, nt * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        return FFX_OK;
    }

    // ---- asynchronous engine: one launch per step, no host sync between steps ----
    struct AStep {
        int kind = 0;  // 0 rotation, 1 Clifford block, 2 run of Clifford rotations (pruned mode)
        typename Alg::Gate g{};
        double cs = 1.0, sn = 0.0;
        int trunc_thresh = 0, trunc_xw = 0;
        int compact_after = 0;  // run the (device-decided) compaction sequence after this step
        int topk_after = 0;     // top-k truncation after this step
        int topk_xw = 0;        // ... followed by the X/Y-weight cap
        int cb = 0, cn = 0;     // Clifford gates [cb, cb+cn) of the block list
        int rb = 0, rn = 0;     // kind 2: Clifford rotations [rb, rb+rn) of the record list
        int rcomm = 0;          // kind 2: their axes commute pairwise
        int prune_after = 0;    // learned pruning after this step
        FeatCtx fctx{};         // feature context of that pruning
    };
    uint32_t* ep_buf = nullptr;  // owned by ctx (ctx->epochs)
    int64_t st_updates = 0, st_outs = 0, st_peak = 0, st_rot = 0, st_compact = 0;

    int ensure_epochs(size_t ne) {
        if (ne <= ctx->epochs_cap) return FFX_OK;
        dev_free(ctx->epochs);
        ctx->epochs_cap = 0;
        const size_t c = std::max<size_t>(pow2ceil(ne), 64);  // stripe copies >= 1 KiB apart
        CHK(dev_alloc(ctx, &ctx->epochs, epoch_words(c) * sizeof(uint32_t)));
        ctx->epochs_cap = c;
        return FFX_OK;
    }

    // n, dead, app, dd (one word per epoch each), flags (64 words), sd (kWin per epoch,
    // kStripes copies), then the sums of the copies: kWin words and one 64-bit word
    // per epoch
    static size_t epoch_words(size_t c) { return (4 + (size_t)kStripes * kWin) * c + 64 + (kWin + 2) * c; }
    int32_t* sd_sum32() const { return (int32_t*)ctx->epochs + (4 + (size_t)kStripes * kWin) * ctx->epochs_cap + 64; }
    unsigned long long* sd_sum64() const {
        return (unsigned long long*)(ctx->epochs + (4 + (size_t)kStripes * kWin + kWin) * ctx->epochs_cap + 64);
    }

    Epochs epochs() const {
        Epochs ep;
        const size_t c = ctx->epochs_cap;
        ep.n = ctx->epochs;
        ep.dead = ctx->epochs + c;
        ep.app = ctx->epochs + 2 * c;
        ep.dd = (int32_t*)(ctx->epochs + 3 * c);
        ep.flags = ctx->epochs + 4 * c;
        ep.sd = (int32_t*)(ctx->epochs + 4 * c + 64);
        ep.sd_stride = (uint32_t)(kWin * c);
        return ep;
    }

    // Host-side test of the window conditions (pairwise commuting, independent axes).
    static bool axes_commute(const typename Alg::Gate& a, const typename Alg::Gate& b) {
        if constexpr (Alg::kPauli) {
            constexpr int W = Alg::W;
            int s = 0;
            for (int w = 0; w < W; ++w) s += __builtin_popcountll((a.p[w] & b.p[W + w]) ^ (a.p[W + w] & b.p[w]));
            return (s & 1) == 0;
        } else {
            int o = 0;
            for (int w = 0; w < NP; ++w) o += __builtin_popcountll(a.p[w] & b.p[w]);
            return ((a.deg * b.deg - o) & 1) == 0;
        }
    }

    // Length of the window starting at rotation step t: consecutive rotations,
    // no Clifford block or compaction inside, pairwise commuting, axes linearly
    // independent over GF(2) (reduced echelon basis with one pivot bit each).
    static size_t window_len(const std::vector<AStep>& steps, size_t t, int max_window) {
        uint64_t basis[kWin][NP];
        int pivot[kWin];
        int nb = 0;
        auto add = [&](const uint64_t* p) {
            uint64_t v[NP];
            for (int q = 0; q < NP; ++q) v[q] = p[q];
            for (int b = 0; b < nb; ++b)
                if ((v[pivot[b] >> 6] >> (pivot[b] & 63)) & 1)
                    for (int q = 0; q < NP; ++q) v[q] ^= basis[b][q];
            for (int q = 0; q < NP; ++q)
                if (v[q]) {
                    pivot[nb] = 64 * q + __builtin_ctzll(v[q]);
                    for (int r = 0; r < NP; ++r) basis[nb][r] = v[r];
                    ++nb;
                    return true;
                }
            return false;
        };
        if (max_window < 2 || !add(steps[t].g.p)) return 1;
        size_t k = 1;
        while ((int)k < max_window && t + k < steps.size()) {
            const AStep& nx = steps[t + k];
            if (steps[t + k - 1].compact_after || steps[t + k - 1].topk_after || steps[t + k - 1].prune_after ||
                nx.kind != 0)
                break;
            bool ok = true;
            for (size_t j = 0; j < k && ok; ++j) ok = axes_commute(steps[t + j].g, nx.g);
            if (!ok || !add(nx.g.p)) break;
            ++k;
        }
        return k;
    }

    // Single-qubit axis of a Pauli rotation: qubit and axis (1 X, 2 Z, 3 Y); false otherwise.
    static bool single_site(const typename Alg::Gate& g, int& site, int& axis) {
        if constexpr (Alg::kPauli) {
            constexpr int W = Alg::W;
            int cnt = 0;
            for (int w = 0; w < W; ++w) {
                const uint64_t sup = g.p[w] | g.p[W + w];
                cnt += __builtin_popcountll(sup);
                if (sup) {
                    site = 64 * w + __builtin_ctzll(sup);
                    axis = ((g.p[w] >> (site & 63)) & 1 ? 1 : 0) | ((g.p[W + w] >> (site & 63)) & 1 ? 2 : 0);
                }
            }
            return cnt == 1;
        } else {
            return false;
        }
    }

    // Length of the layer starting at rotation step t: consecutive rotations on
    // single qubits, all distinct, no compaction / top-k / pruning point inside;
    // 0 when shorter than kLayerMin.
    static size_t plan_layer(const std::vector<AStep>& steps, size_t t) {
        if (steps[t].kind != 0) return 0;
        uint64_t used[4] = {0, 0, 0, 0};
        size_t k = 0;
        while (t + k < steps.size() && k < (size_t)kLayerMax) {
            const AStep& a = steps[t + k];
            if (a.kind != 0) break;
            // compaction points inside a layer move to its end (housekeeping only)
            if (k > 0 && (steps[t + k - 1].topk_after || steps[t + k - 1].prune_after)) break;
            int site = 0, axis = 0;
            if (!single_site(a.g, site, axis) || site >= 256) break;
            if ((used[site >> 6] >> (site & 63)) & 1) break;
            used[site >> 6] |= 1ull << (site & 63);
            ++k;
        }
        return k >= (size_t)kLayerMin ? k : 0;
    }

    // Units of a rotation axis (qubits of a Pauli axis, indices of a Majorana string).
    static void axis_units(const typename Alg::Gate& g, uint64_t (&u)[4]) {
        constexpr int W = Alg::W;
        for (int w = 0; w < 4; ++w) u[w] = 0;
        for (int w = 0; w < W; ++w) u[w] = Alg::kPauli ? (g.p[w] | g.p[(W + w) % NP]) : g.p[w];
    }

    // Disjoint-support layer starting at rotation step t: consecutive rotations
    // whose axes have pairwise disjoint supports (Majorana strings of even
    // degree), at most kLayerMax; top-k and pruning points end it, compaction
    // points inside move to its end, and so does the start of a layer of
    // single-qubit rotations (whose own kernel is faster); 0 when shorter than
    // kLayerMin.
    static size_t plan_dlayer(const std::vector<AStep>& steps, size_t t) {
        uint64_t used[4] = {0, 0, 0, 0};
        size_t k = 0;
        while (t + k < steps.size() && k < (size_t)kLayerMax) {
            const AStep& a = steps[t + k];
            if (a.kind != 0) break;
            if (k > 0 && (steps[t + k - 1].topk_after || steps[t + k - 1].prune_after)) break;
            if (Alg::kPauli && plan_layer(steps, t + k)) break;
            uint64_t u[4];
            axis_units(a.g, u);
            int cnt = 0;
            bool overlap = false;
            for (int w = 0; w < 4; ++w) {
                cnt += __builtin_popcountll(u[w]);
                overlap = overlap || (u[w] & used[w]) != 0;
            }
            if (cnt == 0 || overlap || (!Alg::kPauli && (cnt & 1))) break;
            for (int w = 0; w < 4; ++w) used[w] |= u[w];
            ++k;
        }
        return k >= (size_t)kLayerMin ? k : 0;
    }

    // General layer starting at rotation step t: up to kLayerMaxA consecutive
    // rotations that commute pairwise with linearly independent axes (supports may
    // overlap); ends like a disjoint layer; 0 when shorter than kLayerMin.
    static size_t plan_glayer(const std::vector<AStep>& steps, size_t t) {
        uint64_t basis[kLayerMaxA][NP];
        int pivot[kLayerMaxA];
        int nb = 0;
        auto add = [&](const uint64_t* p) {
            uint64_t v[NP];
            for (int q = 0; q < NP; ++q) v[q] = p[q];
            for (int b = 0; b < nb; ++b)
                if ((v[pivot[b] >> 6] >> (pivot[b] & 63)) & 1)
                    for (int q = 0; q < NP; ++q) v[q] ^= basis[b][q];
            for (int q = 0; q < NP; ++q)
                if (v[q]) {
                    pivot[nb] = 64 * q + __builtin_ctzll(v[q]);
                    for (int r = 0; r < NP; ++r) basis[nb][r] = v[r];
                    ++nb;
                    return true;
                }
            return false;
        };
        size_t k = 0;
        while (t + k < steps.size() && k < (size_t)kLayerMaxA) {
            const AStep& a = steps[t + k];
            if (a.kind != 0) break;
            if (k > 0 && (steps[t + k - 1].topk_after || steps[t + k - 1].prune_after)) break;
            if (k > 0 && Alg::kPauli && plan_layer(steps, t + k)) break;
            bool ok = true;
            for (size_t j = 0; j < k && ok; ++j) ok = axes_commute(steps[t + j].g, a.g);
            if (!ok || !add(a.g.p)) break;
            ++k;
        }
        return k >= (size_t)kLayerMin ? k : 0;
    }

    // Dual vectors of a general layer's axes P_0..P_{nr-1} (linearly independent key
    // vectors, nr <= kLayerMaxA): w_j with parity(P_i & w_j) = [i == j], supported on
    // one pivot bit per axis.  For a key k, a_j = parity(k & w_j) are the coordinates of
    // k's component in span{P}; clearing those of the active axes gives one member of
    // k's orbit that every member agrees on.  False if the axes are dependent.
    static bool general_duals(const AStep* a, int nr, uint64_t (*dual)[NP]) {
        uint64_t v[kLayerMaxA][NP];
        int piv[kLayerMaxA];
        for (int b = 0; b < nr; ++b) {
            for (int q = 0; q < NP; ++q) v[b][q] = a[b].g.p[q];
            for (int c = 0; c < b; ++c)
                if ((v[b][piv[c] >> 6] >> (piv[c] & 63)) & 1)
                    for (int q = 0; q < NP; ++q) v[b][q] ^= v[c][q];
            piv[b] = -1;
            for (int q = 0; q < NP && piv[b] < 0; ++q)
                if (v[b][q]) piv[b] = 64 * q + __builtin_ctzll(v[b][q]);
            if (piv[b] < 0) return false;
        }
        // Q[i] (bit l) = bit piv[l] of P_i; Q is invertible over GF(2); R = Q^-1
        uint32_t Q[kLayerMaxA], R[kLayerMaxA];
        for (int i = 0; i < nr; ++i) {
            Q[i] = 0;
            R[i] = 1u << i;
            for (int l = 0; l < nr; ++l) Q[i] |= (uint32_t)((a[i].g.p[piv[l] >> 6] >> (piv[l] & 63)) & 1) << l;
        }
        for (int c = 0; c < nr; ++c) {
            int r = c;
            while (r < nr && !((Q[r] >> c) & 1)) ++r;
            if (r == nr) return false;
            std::swap(Q[r], Q[c]);
            std::swap(R[r], R[c]);
            for (int i = 0; i < nr; ++i)
                if (i != c && ((Q[i] >> c) & 1)) {
                    Q[i] ^= Q[c];
                    R[i] ^= R[c];
                }
        }
        // Q R_rows = I as row operations: R (rows) = Q^-1, so (Q^-1)[l][j] = bit j of R[l];
        // w_j = sum_l (Q^-1)[l][j] e_piv[l]  gives  parity(P_i & w_j) = (Q Q^-1)[i][j]
        for (int j = 0; j < nr; ++j) {
            for (int q = 0; q < NP; ++q) dual[j][q] = 0;
            for (int l = 0; l < nr; ++l)
                if ((R[l] >> j) & 1) dual[j][piv[l] >> 6] |= 1ull << (piv[l] & 63);
        }
        for (int i = 0; i < nr; ++i)
            for (int j = 0; j < nr; ++j) {
                int par = 0;
                for (int q = 0; q < NP; ++q) par ^= __builtin_popcountll(a[i].g.p[q] & dual[j][q]) & 1;
                if (par != (i == j ? 1 : 0)) return false;
            }
        return true;
    }

    TopK topk_scratch() const {
        TopK tk;
        tk.hist = ctx->topk_buf;
        tk.prefix = (unsigned long long*)(ctx->topk_buf + 256);
        tk.want = ctx->topk_buf + 258;
        tk.prefix2 = (unsigned long long*)(ctx->topk_buf + 260);
        return tk;
    }

    // Kernel argument of the window of rotation steps [t, t + kw).
    static Window<Alg, CP> make_window(const std::vector<AStep>& steps, size_t t, size_t kw, double mincoeff,
                                       int maxdeg, int max_xw) {
        Window<Alg, CP> w{};
        for (size_t j = 0; j < kw; ++j) {
            const AStep& b = steps[t + j];
            w.g[j] = b.g;
            w.cs[j] = make_fac<CP>(b.cs);
            w.sn[j] = make_fac<CP>(b.sn);
            w.thresh_mask |= b.trunc_thresh << j;
            w.xw_mask |= b.trunc_xw << j;
        }
        w.k = (int)kw;
        for (size_t j = 0; j < kw; ++j) {
            w.lp[j] = lin_hash<NP>(w.g[j].p);
            for (size_t i2 = 0; i2 < kw; ++i2) w.cross[j] |= (uint32_t)Alg::lin(w.g[i2].p, w.g[j]) << i2;
        }
        w.mincoeff = mincoeff;
        w.maxdeg = maxdeg;
        w.max_xw = max_xw;
        return w;
    }

    int run_async(const std::vector<AStep>& steps, const std::vector<CliffGate>& cliff, double mincoeff, int maxdeg,
                  int max_xw, uint32_t topk = 0, const std::vector<CRot<Alg, CP>>* crots = nullptr,
                  const PruneGeo* geo = nullptr, float log10_tau = 0.f, bool layers = false) {
        if (steps.empty()) return FFX_OK;
        if (crots && !crots->empty()) {
            const size_t bytes = crots->size() * sizeof(CRot<Alg, CP>);
            if (bytes > ctx->crot_bytes) {
                dev_free(ctx->crot);
                ctx->crot_bytes = 0;
                CHK(dev_alloc(ctx, &ctx->crot, bytes));
                ctx->crot_bytes = bytes;
            }
            HIPCHK(hipMemcpyAsync(ctx->crot, crots->data(), bytes, hipMemcpyHostToDevice, ctx->stream));
        }
        bool pruning = false;
        for (const AStep& a : steps) pruning = pruning || a.prune_after;
        if (pruning) {
            if (K != 1) return set_err(ctx, FFX_EINVAL, "the learned pruner needs real coefficients");
            if (!ctx->mlp || !ctx->mlp->ready || !geo)
                return set_err(ctx, FFX_EINVAL, "the learned pruner has no model (ffx_pruner_init / set_params)");
        }
        if (!cliff.empty()) {
            if (cliff.size() > ctx->cliff_cap) {
                dev_free(ctx->cliff);
                CHK(dev_alloc(ctx, &ctx->cliff, cliff.size() * sizeof(CliffGate)));
                ctx->cliff_cap = cliff.size();
            }
            HIPCHK(hipMemcpyAsync(ctx->cliff, cliff.data(), cliff.size() * sizeof(CliffGate), hipMemcpyHostToDevice,
                                  ctx->stream));
        }
        if (ctx->n_cu <= 0) {
            hipDeviceProp_t prop;
            HIPCHK(hipGetDeviceProperties(&prop, ctx->device));
            ctx->n_cu = prop.multiProcessorCount;
        }
        int rot_grid = 2 * ctx->n_cu;
        if (const char* ev = getenv("FFX_ROT_GRID"))
            if (atoi(ev) > 0) rot_grid = atoi(ev);
        // the disjoint / general layer kernel launches one resident round (4 blocks of 256
        // threads per CU): a round and a half left a tail of late blocks (kicked Ising
        // per-gate 8.9 -> 8.5 ms, Fermi-Hubbard 6.1 -> 6.0 ms).  The window and
        // single-qubit layer kernels keep 6 per CU: on 3.7e7 live terms (maxdegree 16, no
        // layers) 4 per CU took 1.63 s against 1.32 s, and the pruned step's RX layers
        // (~3e3 terms, bigger tiles on fewer blocks) 39.5 us against 33.5 us.
        int win_grid = 6 * ctx->n_cu;
        int layer_grid = 4 * ctx->n_cu;
        if (const char* ev = getenv("FFX_WIN_GRID"))
            if (atoi(ev) > 0) win_grid = layer_grid = atoi(ev);
        int max_window = kWin;
        if (const char* ev = getenv("FFX_WINDOW")) max_window = std::min(kWin, std::max(1, atoi(ev)));
        // layers of single-qubit rotations (one launch each), planned greedily from the
        // start; overflow restarts happen at launch boundaries, so the plan stays valid
        // layers: kind 1 (single-qubit Pauli rotations, ffx_layer.h) or kind 2 (disjoint
        // supports, ffx_dlayer.h); layer_off numbers their rotations for the counters,
        // rec_off their records (each kind in its own array)
        std::vector<uint32_t> layer_len(steps.size(), 0), layer_off(steps.size(), 0), rec_off(steps.size(), 0);
        std::vector<uint8_t> layer_kind(steps.size(), 0), layer_general(steps.size(), 0);
        size_t layer_bytes = 0, n_layer_rot = 0, n_rec1 = 0, n_rec2 = 0;
        if (layers) {
            const bool dlayers = getenv("FFX_NO_DLAYERS") == nullptr;  // kind 2 off: a cross-check
            for (size_t t = 0; t < steps.size();) {
                size_t k = 0;
                int kind = 0;
                if constexpr (Alg::kPauli) {
                    k = plan_layer(steps, t);
                    kind = k ? 1 : 0;
                }
                if (!k && dlayers) {
                    // the longer of a disjoint and a general layer (the disjoint one on ties)
                    const size_t kd = plan_dlayer(steps, t), kg = plan_glayer(steps, t);
                    k = std::max(kd, kg);
                    kind = k ? 2 : 0;
                    layer_general[t] = kg > kd ? 1 : 0;
                }
                if (k) {
                    layer_len[t] = (uint32_t)k;
                    layer_off[t] = (uint32_t)n_layer_rot;
                    layer_kind[t] = (uint8_t)kind;
                    rec_off[t] = (uint32_t)(kind == 1 ? n_rec1 : n_rec2);
                    (kind == 1 ? n_rec1 : n_rec2) += k;
                    n_layer_rot += k;
                    t += k;
                } else {
                    t += steps[t].kind == 0 ? window_len(steps, t, max_window) : 1;
                }
            }
        }
        size_t n_layers = 0;
        for (size_t t = 0; t < steps.size(); ++t) n_layers += layer_len[t] ? 1 : 0;
        // buffer: records (kind 1, then kind 2), position tables (256 int16 per layer),
        // counters (n_layer_rot int32)
        const size_t lrec2_off = (n_rec1 * sizeof(LayerGate<CP>) + 255) & ~(size_t)255;
        const size_t lrec_bytes = lrec2_off + n_rec2 * sizeof(DLayerGate<Alg, CP>);
        const size_t lpos_off = (lrec_bytes + 255) & ~(size_t)255;
        const size_t lsd_off = lpos_off + ((n_layers * 256 * sizeof(int16_t) + 255) & ~(size_t)255);
        const size_t lsd_stride = (n_layer_rot + 63) & ~(size_t)63;  // striped counter copies
        layer_bytes = lsd_off + (kStripes + 1) * lsd_stride * sizeof(int32_t);  // copies, then their sums
        constexpr size_t kLM = 3 * FFX_MAX_WORDS + 8;  // per layer: mx, mz, my (4 words each), tmask, xmask
        std::vector<uint64_t> layer_masks;
        if (n_layers) {
            if (layer_bytes > ctx->layer_buf_bytes) {
                dev_free(ctx->layer_buf);
                ctx->layer_buf_bytes = 0;
                CHK(dev_alloc(ctx, &ctx->layer_buf, layer_bytes));
                ctx->layer_buf_bytes = layer_bytes;
            }
            std::vector<char> hb(lsd_off, 0);
            LayerGate<CP>* rec = (LayerGate<CP>*)hb.data();
            DLayerGate<Alg, CP>* rec2 = (DLayerGate<Alg, CP>*)(hb.data() + lrec2_off);
            int16_t* pos = (int16_t*)(hb.data() + lpos_off);
            size_t li = 0;
            for (size_t t = 0; t < steps.size(); ++t) {
                if (!layer_len[t]) continue;
                int16_t* pt = pos + 256 * li;
                for (int q = 0; q < 256; ++q) pt[q] = -1;
                uint64_t m3[kLM] = {};
                if (layer_kind[t] == 2) {
                    // records with pivots; unit -> position table; union of the units
                    for (uint32_t j = 0; j < layer_len[t]; ++j) {
                        const AStep& a = steps[t + j];
                        DLayerGate<Alg, CP>& r = rec2[rec_off[t] + j];
                        r.g = a.g;
                        r.cs = make_fac<CP>(a.cs);
                        r.sn = make_fac<CP>(a.sn);
                        r.pivot = -1;
                        for (int q = 0; q < NP && r.pivot < 0; ++q)
                            if (a.g.p[q]) r.pivot = 64 * q + __builtin_ctzll(a.g.p[q]);
                        uint64_t u[4];
                        axis_units(a.g, u);
                        for (int w = 0; w < 4; ++w) {
                            m3[w] |= u[w];  // union of the units: words 0..3
                            for (uint64_t b = u[w]; b; b &= b - 1) pt[64 * w + __builtin_ctzll(b)] = (int16_t)j;
                        }
                        if (a.trunc_thresh) m3[3 * FFX_MAX_WORDS + (j >> 6)] |= 1ull << (j & 63);
                        if (a.trunc_xw) m3[3 * FFX_MAX_WORDS + 4 + (j >> 6)] |= 1ull << (j & 63);
                    }
                    if (layer_general[t]) {
                        // a general layer finds its rotations without the position table: its
                        // 512 bytes hold the axes' dual vectors instead
                        static_assert(kLayerMaxA * NP * sizeof(uint64_t) <= 256 * sizeof(int16_t), "dual vectors");
                        uint64_t dual[kLayerMaxA][NP];
                        if (!general_duals(&steps[t], (int)layer_len[t], dual))
                            return set_err(ctx, FFX_EHIP, "internal error: general layer with dependent axes");
                        std::memcpy(pt, dual, layer_len[t] * NP * sizeof(uint64_t));
                    }
                    layer_masks.insert(layer_masks.end(), m3, m3 + kLM);
                    ++li;
                    continue;
                }
                for (uint32_t j = 0; j < layer_len[t]; ++j) {
                    const AStep& a = steps[t + j];
                    int site = 0, axis = 0;
                    single_site(a.g, site, axis);
                    LayerGate<CP>& r = rec[rec_off[t] + j];
                    r.cs = make_fac<CP>(a.cs);
                    r.sn = make_fac<CP>(a.sn);
                    r.site = site;
                    r.axis = axis;
                    r.thresh = a.trunc_thresh;
                    r.xw = a.trunc_xw;
                    pt[site] = (int16_t)j;
                    const int which = axis == 1 ? 0 : (axis == 2 ? 1 : 2);
                    m3[which * FFX_MAX_WORDS + (site >> 6)] |= 1ull << (site & 63);
                    if (a.trunc_thresh) m3[3 * FFX_MAX_WORDS + (j >> 6)] |= 1ull << (j & 63);
                    if (a.trunc_xw) m3[3 * FFX_MAX_WORDS + 4 + (j >> 6)] |= 1ull << (j & 63);
                }
                layer_masks.insert(layer_masks.end(), m3, m3 + kLM);
                ++li;
            }
            HIPCHK(hipMemcpyAsync(ctx->layer_buf, hb.data(), lsd_off, hipMemcpyHostToDevice, ctx->stream));
            if (ctx->claim_slots != ctx->hslots || !ctx->claim) {
                dev_free(ctx->claim);
                ctx->claim_slots = 0;
                CHK(dev_alloc(ctx, &ctx->claim, ctx->hslots * sizeof(uint64_t)));
                ctx->claim_slots = ctx->hslots;
                HIPCHK(hipMemsetAsync(ctx->claim, 0, ctx->hslots * sizeof(uint64_t), ctx->stream));
                ctx->layer_stamp = 0;
            }
        }
        std::vector<size_t> layer_index(steps.size(), 0);
        {
            size_t li = 0;
            for (size_t t = 0; t < steps.size(); ++t)
                if (layer_len[t]) layer_index[t] = li++;
        }
        size_t s = 0;
        for (int attempt = 0; attempt < 64 && s < steps.size(); ++attempt) {
            // epochs needed from step s on
            size_t ne = 1;
            for (size_t t = s; t < steps.size(); ++t)
                ne += 1 + (steps[t].compact_after ? 1 : 0) + (steps[t].topk_after ? 1 : 0) +
                      (steps[t].prune_after ? 1 : 0);
            CHK(ensure_epochs(ne));
            CHK(ensure_counts(ctx->cap / kTile + 2));
            const Epochs ep = epochs();
            HIPCHK(hipMemsetAsync(ctx->epochs, 0, epoch_words(ctx->epochs_cap) * sizeof(uint32_t), ctx->stream));
            if (topk > 0) {
                if (!ctx->topk_buf) CHK(dev_alloc(ctx, &ctx->topk_buf, 264 * sizeof(uint32_t)));
                HIPCHK(hipMemsetAsync(ctx->topk_buf, 0, 264 * sizeof(uint32_t), ctx->stream));
            }
            uint32_t init[2] = {n_host, dead_host};
            HIPCHK(hipMemcpyAsync(ep.n, &init[0], sizeof(uint32_t), hipMemcpyHostToDevice, ctx->stream));
            HIPCHK(hipMemcpyAsync(ep.dead, &init[1], sizeof(uint32_t), hipMemcpyHostToDevice, ctx->stream));
            if (n_layers) {
                HIPCHK(hipMemsetAsync((char*)ctx->layer_buf + lsd_off, 0, kStripes * lsd_stride * sizeof(int32_t),
                                      ctx->stream));
                if (ctx->claim_slots != ctx->hslots) {
                    dev_free(ctx->claim);
                    ctx->claim_slots = 0;
                    CHK(dev_alloc(ctx, &ctx->claim, ctx->hslots * sizeof(uint64_t)));
                    ctx->claim_slots = ctx->hslots;
                    HIPCHK(hipMemsetAsync(ctx->claim, 0, ctx->hslots * sizeof(uint64_t), ctx->stream));
                    ctx->layer_stamp = 0;
                }
            }
            std::vector<int64_t> step_of(ne, -1);
            std::vector<int> is_rot(ne, 0), crot_len(ne, 0);
            std::vector<int64_t> layer_of(ne, -1);
            std::vector<std::pair<void*, void*>> planes_at(steps.size());
            const unsigned g_cap = grid_for(ctx->cap, kTile, 4096);
            const unsigned g_idx = grid_for(std::max(ctx->cap, ctx->hslots), kBlock, 2048);
            uint32_t e = 1;
            size_t nev = 0;
            if (ctx->kernel_timing) {
                size_t nrot = 0;
                for (size_t t = s; t < steps.size(); ++t) nrot += steps[t].kind == 0;
                while (ctx->kev.size() < 2 * nrot) {
                    hipEvent_t ev;
                    HIPCHK(hipEventCreate(&ev));
                    ctx->kev.push_back(ev);
                }
            }
            const double te0 = now_s();
            for (size_t t = s; t < steps.size();) {
                const AStep& a = steps[t];
                const bool is_layer = layer_len[t] != 0;
                const size_t kw = is_layer ? layer_len[t] : (a.kind == 0 ? window_len(steps, t, max_window) : 1);
                planes_at[t] = {ctx->coef, ctx->coef2};
                // one propagation launch in kernel_timing is bracketed by events, chosen by a
                // hash of the launch number (a fixed stride can alias with the periodic
                // mix of window and layer launches and bias the average)
                const bool timed = a.kind == 0 && ctx->kernel_timing > 0 &&
                                   (hmix(0x9E3779B97F4A7C15ull * ++ctx->launch_seq) % (uint64_t)ctx->kernel_timing) == 0;
                if (a.kind == 0) ctx->prof.run_launches += 1;
                if (timed) HIPCHK(hipEventRecord(ctx->kev[2 * nev], ctx->stream));
                if (is_layer && layer_kind[t] == 2) {
                    constexpr int W = Alg::W;
                    if (++ctx->layer_stamp >= (1u << 24)) {
                        HIPCHK(hipMemsetAsync(ctx->claim, 0, ctx->hslots * sizeof(uint64_t), ctx->stream));
                        ctx->layer_stamp = 1;
                    }
                    const size_t li = layer_index[t];
                    DLayer<W, CP, Alg> L{};
                    L.gates = (const DLayerGate<Alg, CP>*)((const char*)ctx->layer_buf + lrec2_off) + rec_off[t];
                    L.pos = (const int16_t*)((const char*)ctx->layer_buf + lpos_off) + 256 * li;
                    L.sd = (int32_t*)((char*)ctx->layer_buf + lsd_off) + layer_off[t];
                    L.sd_stride = (uint32_t)lsd_stride;
                    L.claim = ctx->claim;
                    const uint64_t* lm = layer_masks.data() + li * kLM;
                    for (int w = 0; w < W; ++w) L.umask[w] = lm[w];
                    for (int u = 0; u < 4; ++u) {
                        L.tmask[u] = lm[3 * FFX_MAX_WORDS + u];
                        L.xmask[u] = lm[3 * FFX_MAX_WORDS + 4 + u];
                    }
                    L.mincoeff = mincoeff;
                    L.nr = (int)kw;
                    L.maxdeg = maxdeg;
                    L.max_xw = max_xw;
                    L.stamp = ctx->layer_stamp;
                    L.general = layer_general[t];
                    L.spread = (L.tmask[0] | L.tmask[1] | L.tmask[2] | L.tmask[3]) != 0 && kw >= 16;
                    L.tflags = dlayer_tflags(L.tmask, L.xmask, L.nr);
                    if (L.tflags & kTruncAny)
                        hipLaunchKernelGGL((k_dlayer_async<Alg, CP, K, true>), dim3(layer_grid), dim3(kBlock), 0,
                                           ctx->stream, store(), (const store_t*)ctx->coef, (store_t*)ctx->coef2, L,
                                           ep, e);
                    else
                        hipLaunchKernelGGL((k_dlayer_async<Alg, CP, K, false>), dim3(layer_grid), dim3(kBlock), 0,
                                           ctx->stream, store(), (const store_t*)ctx->coef, (store_t*)ctx->coef2, L,
                                           ep, e);
                    CHK(launch_check());
                    if (timed) HIPCHK(hipEventRecord(ctx->kev[2 * nev++ + 1], ctx->stream));
                    std::swap(ctx->coef, ctx->coef2);
                    step_of[e] = (int64_t)t;
                    is_rot[e] = (int)kw;
                    layer_of[e] = (int64_t)layer_off[t];
                    ++e;
                } else if (is_layer) {
                    if constexpr (Alg::kPauli) {
                        constexpr int W = Alg::W;
                        if (++ctx->layer_stamp >= (1u << 24)) {
                            HIPCHK(hipMemsetAsync(ctx->claim, 0, ctx->hslots * sizeof(uint64_t), ctx->stream));
                            ctx->layer_stamp = 1;
                        }
                        const size_t li = layer_index[t];
                        Layer<W, CP> L{};
                        L.gates = (const LayerGate<CP>*)ctx->layer_buf + rec_off[t];
                        L.pos = (const int16_t*)((const char*)ctx->layer_buf + lpos_off) + 256 * li;
                        L.sd = (int32_t*)((char*)ctx->layer_buf + lsd_off) + layer_off[t];
                        L.sd_stride = (uint32_t)lsd_stride;
                        L.claim = ctx->claim;
                        const uint64_t* lm = layer_masks.data() + li * kLM;
                        for (int w = 0; w < W; ++w) {
                            L.mx[w] = lm[w];
                            L.mz[w] = lm[FFX_MAX_WORDS + w];
                            L.my[w] = lm[2 * FFX_MAX_WORDS + w];
                        }
                        for (int u = 0; u < 4; ++u) {
                            L.tmask[u] = lm[3 * FFX_MAX_WORDS + u];
                            L.xmask[u] = lm[3 * FFX_MAX_WORDS + 4 + u];
                        }
                        L.mincoeff = mincoeff;
                        L.nr = (int)kw;
                        L.maxdeg = maxdeg;
                        L.max_xw = max_xw;
                        L.stamp = ctx->layer_stamp;
                        hipLaunchKernelGGL((k_layer_async<Alg, CP, K>), dim3(win_grid), dim3(kBlock), 0, ctx->stream,
                                           store(), (const store_t*)ctx->coef, (store_t*)ctx->coef2, L, ep, e);
                        CHK(launch_check());
                        if (timed) HIPCHK(hipEventRecord(ctx->kev[2 * nev++ + 1], ctx->stream));
                        std::swap(ctx->coef, ctx->coef2);
                        step_of[e] = (int64_t)t;
                        is_rot[e] = (int)kw;
                        layer_of[e] = (int64_t)layer_off[t];
                        ++e;
                    }
                } else if (a.kind == 0 && kw > 1) {
                    const Window<Alg, CP> w = make_window(steps, t, kw, mincoeff, maxdeg, max_xw);
                    hipLaunchKernelGGL((k_win_async<Alg, CP, K>), dim3(win_grid), dim3(kBlockWin), 0, ctx->stream,
                                       store(), (const store_t*)ctx->coef, (store_t*)ctx->coef2, w, ep, e);
                    CHK(launch_check());
                    if (timed) HIPCHK(hipEventRecord(ctx->kev[2 * nev++ + 1], ctx->stream));
                    std::swap(ctx->coef, ctx->coef2);
                    step_of[e] = (int64_t)t;
                    is_rot[e] = (int)kw;
                    ++e;
                } else if (a.kind == 0) {
                    RotParams<CP> rp{};
                    rp.cs = make_fac<CP>(a.cs);
                    rp.sn = make_fac<CP>(a.sn);
                    rp.mincoeff = mincoeff;
                    rp.maxdeg = maxdeg;
                    rp.trunc_thresh = a.trunc_thresh;
                    rp.trunc_xw = a.trunc_xw;
                    rp.max_xw = max_xw;
                    hipLaunchKernelGGL((k_rot_async<Alg, CP, K>), dim3(rot_grid), dim3(kBlockRot), 0, ctx->stream,
                                       store(), (const store_t*)ctx->coef, (store_t*)ctx->coef2, a.g, rp, ep, e);
                    CHK(launch_check());
                    if (timed) HIPCHK(hipEventRecord(ctx->kev[2 * nev++ + 1], ctx->stream));
                    std::swap(ctx->coef, ctx->coef2);
                    step_of[e] = (int64_t)t;
                    is_rot[e] = 1;
                    ++e;
                } else if (a.kind == 2) {
                    hipLaunchKernelGGL((k_crot_async<Alg, CP, K>), dim3(g_idx), dim3(kBlock), 0, ctx->stream, store(),
                                       (const CRot<Alg, CP>*)ctx->crot + a.rb, a.rn, a.rcomm, mincoeff, maxdeg, max_xw,
                                       ep, e);
                    CHK(launch_check());
                    hipLaunchKernelGGL(k_index_build_async<NP>, dim3(g_idx), dim3(kBlock), 0, ctx->stream, ctx->key,
                                       ctx->cap, ctx->index, ctx->hslots - 1, ep, e, e, 0, &ctx->ctrl->index_full);
                    CHK(launch_check());
                    step_of[e] = (int64_t)t;
                    crot_len[e] = a.rn;
                    ++e;
                } else {
                    if constexpr (Alg::kPauli) {
                        hipLaunchKernelGGL((k_cliff_async<Alg::W, CP, K>), dim3(g_idx), dim3(kBlock), 0, ctx->stream,
                                           store(), (const CliffGate*)(ctx->cliff + a.cb), a.cn, ep, e);
                        CHK(launch_check());
                        hipLaunchKernelGGL(k_index_build_async<NP>, dim3(g_idx), dim3(kBlock), 0, ctx->stream, ctx->key,
                                           ctx->cap, ctx->index, ctx->hslots - 1, ep, e, e, 0,
                                           &ctx->ctrl->index_full);
                        CHK(launch_check());
                    }
                    step_of[e] = (int64_t)t;
                    ++e;
                }
                if (steps[t + kw - 1].prune_after) {
                    if constexpr (K == 1) {
                        const unsigned g_pr = (unsigned)std::max(1, 2 * ctx->n_cu);
                        hipLaunchKernelGGL((k_prune_async<Alg, CP>), dim3(g_pr), dim3(64 * kPruneWaves), 0,
                                           ctx->stream, store(), ep, e, ctx->mlp->dev(), *geo,
                                           steps[t + kw - 1].fctx, log10_tau);
                        CHK(launch_check());
                        step_of[e] = (int64_t)t;
                        ++e;
                    }
                }
                if (steps[t + kw - 1].topk_after) {
                    const uint32_t ep_prev = e - 1;
                    const TopK tk = topk_scratch();
                    const unsigned g_sel = grid_for(ctx->cap, 4 * kBlock, 1024);
                    // the magnitude of the topk-th largest term, then the key hash that ranks
                    // the terms at that magnitude
                    for (int phase = 0; phase < 2; ++phase)
                        for (int shift = 56; shift >= 0; shift -= 8) {
                            hipLaunchKernelGGL((k_topk_hist<NP, CP, K>), dim3(g_sel), dim3(kBlock), 0, ctx->stream,
                                               store(), ep, ep_prev, tk, topk, shift, phase);
                            CHK(launch_check());
                            hipLaunchKernelGGL(k_topk_pick, dim3(1), dim3(kBlock), 0, ctx->stream, ep, ep_prev, tk,
                                               topk, shift, phase);
                            CHK(launch_check());
                        }
                    hipLaunchKernelGGL((k_topk_ties<NP, CP, K>), dim3(g_sel), dim3(kBlock), 0, ctx->stream, store(),
                                       ep, ep_prev, tk, topk, ctx->counts);
                    CHK(launch_check());
                    hipLaunchKernelGGL(k_topk_tiescan, dim3(1), dim3(kBlock), 0, ctx->stream, ep, ep_prev, topk,
                                       ctx->counts);
                    CHK(launch_check());
                    hipLaunchKernelGGL((k_topk_apply<Alg, CP, K>), dim3(g_sel), dim3(kBlock), 0, ctx->stream, store(), ep,
                                       e, tk, topk, steps[t + kw - 1].topk_xw, max_xw, (const uint32_t*)ctx->counts);
                    CHK(launch_check());
                    step_of[e] = (int64_t)t;
                    ++e;
                }
                bool compact_here = steps[t + kw - 1].compact_after != 0;
                if (is_layer)
                    for (size_t j = 0; j < kw; ++j) compact_here = compact_here || steps[t + j].compact_after;
                if (compact_here) {
                    const uint32_t ep_prev = e - 1;
                    hipLaunchKernelGGL((k_cmp_count<CP, K>), dim3(g_cap), dim3(kBlock), 0, ctx->stream, store(), ep,
                                       ep_prev, ctx->counts);
                    CHK(launch_check());
                    hipLaunchKernelGGL(k_cmp_scan, dim3(1), dim3(kBlock), 0, ctx->stream, ep, ep_prev, ctx->counts);
                    CHK(launch_check());
                    hipLaunchKernelGGL((k_cmp_scatter<NP, CP, K>), dim3(g_cap), dim3(kBlock), 0, ctx->stream, store(),
                                       ctx->key2, (store_t*)ctx->coef2, ep, ep_prev, e, ctx->counts);
                    CHK(launch_check());
                    hipLaunchKernelGGL((k_cmp_copyback<NP, CP, K>), dim3(g_idx), dim3(kBlock), 0, ctx->stream, store(),
                                       ctx->key2, (const store_t*)ctx->coef2, ep, ep_prev, e);
                    CHK(launch_check());
                    hipLaunchKernelGGL(k_index_build_async<NP>, dim3(g_idx), dim3(kBlock), 0, ctx->stream, ctx->key,
                                       ctx->cap, ctx->index, ctx->hslots - 1, ep, ep_prev, e, 1,
                                       &ctx->ctrl->index_full);
                    CHK(launch_check());
                    step_of[e] = (int64_t)t;
                    ++e;
                }
                t += kw;
            }
            ctx->prof.seconds_enqueue += now_s() - te0;
            // one sync per run: epochs, flags
            std::vector<uint32_t> hn(e), hd(e), ha(e);
            std::vector<int32_t> hdd(e), hsd((size_t)kWin * e), hlsd(n_layer_rot);
            std::vector<unsigned long long> hsurv(e);
            uint32_t hf[2] = {0, 0};
            HIPCHK(hipMemcpyAsync(hn.data(), ep.n, e * sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
            HIPCHK(hipMemcpyAsync(hd.data(), ep.dead, e * sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
            HIPCHK(hipMemcpyAsync(ha.data(), ep.app, e * sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
            HIPCHK(hipMemcpyAsync(hdd.data(), ep.dd, e * sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
            HIPCHK(hipMemcpyAsync(hf, ep.flags, 2 * sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
            // the striped counter copies, summed on the device
            hipLaunchKernelGGL(k_sum_stripes, dim3(grid_for((size_t)kWin * e, kBlock, 64)), dim3(kBlock), 0,
                               ctx->stream, (const int32_t*)ep.sd, ep.sd_stride, (uint32_t)(kWin * e), (uint32_t)e,
                               sd_sum32(), sd_sum64(), kWin);
            HIPCHK(hipGetLastError());
            HIPCHK(hipMemcpyAsync(hsd.data(), sd_sum32(), hsd.size() * sizeof(int32_t), hipMemcpyDeviceToHost,
                                  ctx->stream));
            HIPCHK(hipMemcpyAsync(hsurv.data(), sd_sum64(), e * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                                  ctx->stream));
            if (n_layer_rot) {
                int32_t* lsum = (int32_t*)((char*)ctx->layer_buf + lsd_off) + kStripes * lsd_stride;
                hipLaunchKernelGGL(k_sum_stripes, dim3(grid_for(n_layer_rot, kBlock, 64)), dim3(kBlock), 0,
                                   ctx->stream, (const int32_t*)((char*)ctx->layer_buf + lsd_off),
                                   (uint32_t)lsd_stride, (uint32_t)n_layer_rot, 0u, lsum,
                                   (unsigned long long*)nullptr, kWin);
                HIPCHK(hipGetLastError());
                HIPCHK(hipMemcpyAsync(hlsd.data(), lsum, n_layer_rot * sizeof(int32_t), hipMemcpyDeviceToHost,
                                      ctx->stream));
            }
            HIPCHK(hipStreamSynchronize(ctx->stream));
            for (size_t q = 0; q < nev; ++q) {
                float ms = 0.f;
                HIPCHK(hipEventElapsedTime(&ms, ctx->kev[2 * q], ctx->kev[2 * q + 1]));
                ctx->prof.seconds_kernel += 1e-3 * ms;
                ctx->prof.timed_launches += 1;
            }
            const uint32_t e_end = hf[0] ? hf[1] : e;
            for (uint32_t q = 1; q < e_end; ++q) {
                if (crot_len[q]) {
                    // a run of Clifford rotations: the kernel sums, over the terms live at its
                    // start, the rotations each one survives; a term erased inside the run also
                    // entered the rotation that erased it
                    const uint64_t surv = hsurv[q];
                    const int64_t lout = (int64_t)(hn[q] + ha[q]) - (int64_t)((int32_t)hd[q] + hdd[q]);
                    st_rot += crot_len[q];
                    st_updates += (int64_t)surv + hdd[q];
                    st_outs += (int64_t)surv;
                    st_peak = std::max<int64_t>(st_peak, lout);
                    continue;
                }
                if (!is_rot[q]) {
                    if (hn[q] != hn[q - 1] + ha[q - 1]) st_compact++;
                    continue;
                }
                const int64_t live = (int64_t)(hn[q] + ha[q]) - (int64_t)((int32_t)hd[q] + hdd[q]);
                if (is_rot[q] == 1) {
                    st_rot++;
                    st_updates += (int64_t)(hn[q] - hd[q]);
                    st_outs += live;
                    st_peak = std::max<int64_t>(st_peak, live);
                    continue;
                }
                // window or layer: live count entering each of its rotations from the per-rotation changes
                int64_t l = (int64_t)hn[q] - (int64_t)hd[q];
                for (int j = 0; j < is_rot[q]; ++j) {
                    st_rot++;
                    st_updates += l;
                    l += layer_of[q] >= 0 ? hlsd[(size_t)layer_of[q] + j] : hsd[(size_t)kWin * q + j];
                    st_outs += l;
                    st_peak = std::max<int64_t>(st_peak, l);
                }
                if (l != live) return set_err(ctx, FFX_EHIP, "internal error: window live counts disagree");
            }
            if (!hf[0]) {
                n_host = hn[e - 1] + ha[e - 1];
                dead_host = (uint32_t)((int32_t)hd[e - 1] + hdd[e - 1]);
                s = steps.size();
                break;
            }
            // overflow at epoch hf[1]: restore the state at the start of that step and grow
            const uint32_t ef = hf[1];
            const size_t tf = (size_t)step_of[ef];
            ctx->coef = planes_at[tf].first;
            ctx->coef2 = planes_at[tf].second;
            n_host = hn[ef];
            dead_host = hd[ef];
            {
                Ctrl c{};
                c.n = n_host;
                c.n_dead = dead_host;
                HIPCHK(hipMemcpyAsync(ctx->ctrl, &c, sizeof(Ctrl), hipMemcpyHostToDevice, ctx->stream));
            }
            hipLaunchKernelGGL(k_index_rollback, dim3(grid_for(ctx->hslots, kBlock, 8192)), dim3(kBlock), 0, ctx->stream,
                               ctx->index, ctx->hslots, n_host);
            CHK(launch_check());
            CHK(reserve(2 * ctx->cap));
            s = tf;
        }
        if (s < steps.size()) return set_err(ctx, FFX_ECAPACITY, "term store kept overflowing");
        HIPCHK(hipMemcpyAsync(ctx->ctrl_host, ctx->ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        if (ctx->ctrl_host->index_full)
            return set_err(ctx, FFX_EHIP,
                           "internal error: hash index full (flags " + std::to_string(ctx->ctrl_host->index_full) + ")");
        Ctrl c{};
        c.n = n_host;
        c.n_dead = dead_host;
        HIPCHK(hipMemcpyAsync(ctx->ctrl, &c, sizeof(Ctrl), hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        *ctx->ctrl_host = c;
        return FFX_OK;
    }

    void finish_profile() {
        if (st_rot) {
            ctx->prof.n_gates = st_rot;
            ctx->prof.updates = st_updates;
            ctx->prof.outs = st_outs;
            ctx->prof.peak_terms = st_peak;
            ctx->prof.compactions = st_compact;
        } else {
            ctx->prof.n_gates = (int64_t)ctx->ctrl_host->n_rot;
            ctx->prof.updates = (int64_t)ctx->ctrl_host->updates;
            ctx->prof.outs = (int64_t)ctx->ctrl_host->outs;
            ctx->prof.peak_terms = (int64_t)ctx->ctrl_host->peak;
            ctx->prof.compactions = (int64_t)ctx->ctrl_host->n_compact;
        }
    }
};

// FFX_ENGINE=steps selects the one-launch-per-rotation engine with a host sync
// after every rotation (kept as a cross-check); the default is the asynchronous
// engine (one launch per step, one sync per propagation).
inline bool step_engine() {
    const char* e = getenv("FFX_ENGINE");
    return e && std::strcmp(e, "steps") == 0;
}

template <int W>
typename PauliAlg<W>::Gate pauli_gate_dev(const ffx_pauli_gate& g) {
    typename PauliAlg<W>::Gate d{};
    int pxz = 0;
    for (int w = 0; w < W; ++w) {
        d.p[w] = g.x[w];
        d.p[W + w] = g.z[w];
        pxz += __builtin_popcountll(g.x[w] & g.z[w]);
    }
    d.pxz = pxz;
    return d;
}

template <int W>
typename MajoranaAlg<W>::Gate majorana_gate_dev(const ffx_majorana_gate& g) {
    typename MajoranaAlg<W>::Gate d{};
    int deg = 0;
    for (int w = 0; w < W; ++w) {
        d.p[w] = g.m[w];
        deg += __builtin_popcountll(g.m[w]);
    }
    // suffix parity: bit t set iff an odd number of axis bits lie strictly above t
    int par = 0;
    for (int t = 64 * W - 1; t >= 0; --t) {
        if (par) d.sm[t >> 6] |= 1ull << (t & 63);
        if ((g.m[t >> 6] >> (t & 63)) & 1ull) par ^= 1;
    }
    d.deg = deg;
    d.r = g.r;
    return d;
}

// Site geometry of the term features from the pruned-mode options.
inline PruneGeo prune_geo(const ffx_prune_opts& po) {
    PruneGeo g{};
    std::memcpy(g.row, po.row, sizeof(g.row));
    std::memcpy(g.col, po.col, sizeof(g.col));
    const int s = (po.obs_site >= 0 && po.obs_site < 256) ? po.obs_site : 0;
    g.obs_row = po.row[s];
    g.obs_col = po.col[s];
    return g;
}

// Feature context after `done` of `nrot` rotations: remaining layers, mean |theta|
// of the remaining branching rotations (|cos theta| > 1e-6), fraction done.
struct FeatPlan {
    std::vector<double> suf_sum;
    std::vector<int> suf_cnt;
    size_t nrot = 0;
    int rot_per_layer = 1;
    template <class AStep>
    FeatPlan(const std::vector<AStep>& sd, int rpl) : rot_per_layer(rpl > 0 ? rpl : 1) {
        for (const AStep& a : sd) nrot += a.kind == 0;
        suf_sum.assign(nrot + 1, 0.0);
        suf_cnt.assign(nrot + 1, 0);
        size_t r = nrot;
        for (size_t t = sd.size(); t-- > 0;) {
            if (sd[t].kind != 0) continue;
            --r;
            const bool br = std::fabs(sd[t].cs) > 1e-6;
            suf_sum[r] = suf_sum[r + 1] + (br ? std::fabs(std::atan2(sd[t].sn, sd[t].cs)) : 0.0);
            suf_cnt[r] = suf_cnt[r + 1] + (br ? 1 : 0);
        }
    }
    FeatCtx at(size_t done) const {
        FeatCtx f{};
        f.rem_layers = (float)(nrot - done) / (float)rot_per_layer;
        f.mean_theta = suf_cnt[done] ? (float)(suf_sum[done] / suf_cnt[done]) : 0.f;
        f.frac_done = nrot ? (float)done / (float)nrot : 0.f;
        return f;
    }
};

// Pruned-mode plan over a list of device steps: rotations with |cos theta| <=
// clifford_tol become records of runs of Clifford rotations (consecutive ones
// share one kind-2 step); every prune_period-th rotation is followed by a
// learned-pruning epoch.
template <class Alg, class CP, int K>
void prune_plan(std::vector<typename Engine<Alg, CP, K>::AStep>& sd, std::vector<CRot<Alg, CP>>& crots,
                const ffx_prune_opts& po) {
    using AStep = typename Engine<Alg, CP, K>::AStep;
    const FeatPlan fp(sd, po.rot_per_layer);
    std::vector<AStep> out;
    out.reserve(sd.size());
    size_t done = 0, compacts = 0;
    size_t thin = 4;  // keep every thin-th compaction point
    if (const char* ev = getenv("FFX_PRUNE_COMPACT"))
        if (atoi(ev) > 0) thin = (size_t)atoi(ev);
    for (AStep a : sd) {
        if (a.kind != 0) {
            out.push_back(a);
            continue;
        }
        ++done;
        // pruned runs keep few terms: a quarter of the compaction points is enough
        if (a.compact_after) a.compact_after = (compacts++ % thin) == thin - 1 ? 1 : 0;
        const bool cliff = po.clifford_tol > 0 && std::fabs(a.cs) <= po.clifford_tol;
        if (cliff) {
            CRot<Alg, CP> cr{};
            cr.g = a.g;
            cr.sn = make_fac<CP>(a.sn);
            cr.thresh = a.trunc_thresh;
            cr.xw = a.trunc_xw;
            AStep* open = (!out.empty() && out.back().kind == 2 && !out.back().topk_after && !out.back().prune_after)
                              ? &out.back()
                              : nullptr;
            if (open) {
                open->rn++;
                open->compact_after |= a.compact_after;
                open->topk_after = a.topk_after;
                open->topk_xw = a.topk_xw;
            } else {
                AStep g2{};
                g2.kind = 2;
                g2.rb = (int)crots.size();
                g2.rn = 1;
                g2.compact_after = a.compact_after;
                g2.topk_after = a.topk_after;
                g2.topk_xw = a.topk_xw;
                out.push_back(g2);
            }
            crots.push_back(cr);
        } else {
            out.push_back(a);
        }
        if (po.prune_period > 0 && done % (size_t)po.prune_period == 0) {
            out.back().prune_after = 1;
            out.back().fctx = fp.at(done);
        }
    }
    // per run: the next truncation points of every record, and whether the axes
    // commute pairwise (all Z-type or all X-type Pauli axes; short runs are
    // checked pair by pair)
    for (AStep& a : out) {
        if (a.kind != 2) continue;
        CRot<Alg, CP>* r = crots.data() + a.rb;
        int nth = a.rn, nxw = a.rn;
        for (int t = a.rn - 1; t >= 0; --t) {
            if (r[t].thresh) nth = t;
            if (r[t].xw) nxw = t;
            r[t].next_th = nth;
            r[t].next_xw = nxw;
        }
        bool comm = false;
        if constexpr (Alg::kPauli) {
            constexpr int W = Alg::W;
            uint64_t xo = 0, zo = 0;
            for (int t = 0; t < a.rn; ++t)
                for (int w = 0; w < W; ++w) {
                    xo |= r[t].g.p[w];
                    zo |= r[t].g.p[W + w];
                }
            comm = xo == 0 || zo == 0;
        }
        if (!comm && a.rn <= 48) {
            comm = true;
            for (int i = 0; i < a.rn && comm; ++i)
                for (int j = i + 1; j < a.rn && comm; ++j)
                    comm = Engine<Alg, CP, K>::axes_commute(r[i].g, r[j].g);
        }
        a.rcomm = comm ? 1 : 0;
    }
    if (const char* ev = getenv("FFX_CROT_SERIAL"))  // rotation-by-rotation loop for every run (cross-check)
        if (atoi(ev) > 0)
            for (AStep& a : out) a.rcomm = 0;
    sd.swap(out);
}

// Device steps of the asynchronous engine for a Pauli schedule: rotation
// parameters, truncation points, top-k points, compaction points, Clifford blocks.
template <int W, class CP, int K>
void pauli_async_steps(const ffx_pauli_gate* gates, const std::vector<Step>& steps, const ffx_options& o,
                       std::vector<typename Engine<PauliAlg<W>, CP, K>::AStep>& sd, std::vector<CliffGate>& cl) {
    using AStep = typename Engine<PauliAlg<W>, CP, K>::AStep;
    sd.reserve(steps.size());
    int trunc_points = 0;
    const int every = compact_every();
    for (const Step& s : steps) {
        AStep d{};
        if (s.kind == 0) {
            const ffx_pauli_gate& g = gates[s.gate];
            d.g = pauli_gate_dev<W>(g);
            gate_trig(g.theta, d.cs, d.sn);
            d.kind = 0;
            d.trunc_thresh = (s.trunc && o.mincoeff > 0) ? 1 : 0;
            d.trunc_xw = (s.trunc && o.max_xweight >= 0 && o.xtrunc_period > 0 &&
                          (s.rc + 1) % o.xtrunc_period == 0)
                             ? 1
                             : 0;
            if (s.trunc && o.topk > 0) {
                // threshold, then top-k, then the X/Y cap (truncate.h:51-57)
                d.topk_after = 1;
                d.topk_xw = d.trunc_xw;
                d.trunc_xw = 0;
            }
            // batched runs compact at every batch end; per-gate runs at every compact_every()-th
            // truncation point (FFX_COMPACT_EVERY, default 256)
            if (s.trunc) d.compact_after = (o.batched || (++trunc_points % every) == 0) ? 1 : 0;
        } else {
            d.kind = 1;
            d.cb = (int)cl.size();
            d.cn = (int)(s.ce - s.cb);
            for (int64_t t = s.cb; t < s.ce; ++t) cl.push_back(CliffGate{gates[t].kind, gates[t].q0, gates[t].q1, 0});
        }
        sd.push_back(d);
    }
}

template <int W, class CP, int K>
int run_pauli(ffx_ctx* ctx, const ffx_pauli_gate* gates, int64_t n_gates, const uint64_t* x, const uint64_t* z,
              const double* re, const double* im, int64_t n_terms, const ffx_options& o, ffx_terms* out,
              const uint64_t* state, double* out_expect, const ffx_prune_opts* po = nullptr) {
    using Alg = PauliAlg<W>;
    Engine<Alg, CP, K> eng(ctx);
    const double t0 = now_s();
    CHK(eng.upload(x, z, re, im, n_terms));
    const double t1 = now_s();
    const std::vector<Step> steps = pauli_schedule<W>(gates, n_gates, o.batched != 0);
    const int maxdeg = degree_limit(o.maxdegree);
    if (!step_engine()) {
        using AStep = typename Engine<Alg, CP, K>::AStep;
        std::vector<AStep> sd;
        std::vector<CliffGate> cl;
        pauli_async_steps<W, CP, K>(gates, steps, o, sd, cl);
        std::vector<CRot<Alg, CP>> crots;
        PruneGeo geo{};
        if (po) {
            prune_plan<Alg, CP, K>(sd, crots, *po);
            geo = prune_geo(*po);
        }
        // layers need every term's degree <= kLayerMaxA: no Clifford blocks (they can raise degree),
        // partners capped by maxdegree, observable terms within the bound
        bool layers = cl.empty() && o.maxdegree <= kLayerMaxA;
        for (int64_t i = 0; i < n_terms && layers; ++i) {
            int d = 0;
            for (int w = 0; w < W; ++w) d += __builtin_popcountll(x[i * W + w] | z[i * W + w]);
            layers = d <= kLayerMaxA;
        }
        if (const char* ev = getenv("FFX_LAYERS"))
            if (std::strcmp(ev, "0") == 0) layers = false;
        CHK(eng.run_async(sd, cl, o.mincoeff, maxdeg, o.max_xweight, o.topk > 0 ? (uint32_t)o.topk : 0u, &crots,
                          &geo, po ? po->log10_tau : 0.f, layers));
    } else
    for (const Step& s : steps) {
        if (s.kind == 0) {
            const ffx_pauli_gate& g = gates[s.gate];
            RotParams<CP> rp{};
            double cs_, sn_;
            gate_trig(g.theta, cs_, sn_);
            rp.cs = make_fac<CP>(cs_);
            rp.sn = make_fac<CP>(sn_);
            rp.maxdeg = maxdeg;
            rp.mincoeff = o.mincoeff;
            rp.trunc_thresh = (s.trunc && o.mincoeff > 0) ? 1 : 0;
            rp.trunc_xw = (s.trunc && o.max_xweight >= 0 && o.xtrunc_period > 0 &&
                           (s.rc + 1) % o.xtrunc_period == 0 && o.topk <= 0)
                              ? 1
                              : 0;
            rp.max_xw = o.max_xweight;
            CHK(eng.rotation(pauli_gate_dev<W>(g), rp));
            if (s.trunc && o.topk > 0)
                return set_err(ctx, FFX_EINVAL, "topk truncation is not supported on the device path");
            if (s.trunc && eng.dead_host > 0 && (size_t)eng.dead_host * 4 > eng.n_host) CHK(eng.compact());
        } else {
            CHK(eng.clifford(gates, s.cb, s.ce));
        }
    }
    CHK(eng.sync_ctrl());
    eng.finish_profile();
    const double t2 = now_s();
    if (out_expect) CHK(eng.expect(state, out_expect));
    if (out) CHK(eng.download(out));
    const double t3 = now_s();
    ctx->prof.seconds_upload = t1 - t0;
    ctx->prof.seconds_propagate = t2 - t1;
    ctx->prof.seconds_download = t3 - t2;
    ctx->prof.seconds_total = t3 - t0;
    ctx->last_kind = 1;
    ctx->last_W = W;
    ctx->last_K = K;
    ctx->last_prec = std::is_same<CP, F64>::value ? FFX_FP64 : (std::is_same<CP, F32>::value ? FFX_FP32 : FFX_FP32C);
    return FFX_OK;
}

template <int W, class CP, int K>
int run_majorana(ffx_ctx* ctx, const ffx_majorana_gate* gates, int64_t n_gates, const uint64_t* m, const double* re,
                 const double* im, int64_t n_terms, int maxdegree, double mincoeff, ffx_terms* out,
                 const uint64_t* occ, double* out_expect, const ffx_prune_opts* po = nullptr) {
    using Alg = MajoranaAlg<W>;
    Engine<Alg, CP, K> eng(ctx);
    const double t0 = now_s();
    CHK(eng.upload(m, nullptr, re, im, n_terms));
    const double t1 = now_s();
    const int maxdeg = degree_limit(maxdegree);
    if (!step_engine()) {
        using AStep = typename Engine<Alg, CP, K>::AStep;
        std::vector<AStep> sd;
        sd.reserve((size_t)n_gates);
        int cnt = 0;
        const int every = compact_every();
        for (int64_t i = n_gates - 1; i >= 0; --i) {
            AStep d{};
            d.g = majorana_gate_dev<W>(gates[i]);
            gate_trig(gates[i].theta, d.cs, d.sn);
            d.kind = 0;
            d.trunc_thresh = mincoeff > 0 ? 1 : 0;
            d.compact_after = (mincoeff > 0 && (++cnt % every) == 0) ? 1 : 0;
            sd.push_back(d);
        }
        std::vector<CRot<Alg, CP>> crots;
        PruneGeo geo{};
        if (po) {
            prune_plan<Alg, CP, K>(sd, crots, *po);
            geo = prune_geo(*po);
        }
        // disjoint-support layers need every term's degree <= kLayerMaxA (as for Pauli layers)
        bool layers = maxdegree <= kLayerMaxA;
        for (int64_t i = 0; i < n_terms && layers; ++i) {
            int d = 0;
            for (int w = 0; w < W; ++w) d += __builtin_popcountll(m[i * W + w]);
            layers = d <= kLayerMaxA;
        }
        if (const char* ev = getenv("FFX_LAYERS"))
            if (std::strcmp(ev, "0") == 0) layers = false;
        CHK(eng.run_async(sd, {}, mincoeff, maxdeg, -1, 0u, &crots, &geo, po ? po->log10_tau : 0.f, layers));
    } else
    for (int64_t i = n_gates - 1; i >= 0; --i) {
        const ffx_majorana_gate& g = gates[i];
        RotParams<CP> rp{};
        double cs_, sn_;
        gate_trig(g.theta, cs_, sn_);
        rp.cs = make_fac<CP>(cs_);
        rp.sn = make_fac<CP>(sn_);
        rp.maxdeg = maxdeg;
        rp.mincoeff = mincoeff;
        rp.trunc_thresh = mincoeff > 0 ? 1 : 0;
        CHK(eng.rotation(majorana_gate_dev<W>(g), rp));
        if (eng.dead_host > 0 && (size_t)eng.dead_host * 4 > eng.n_host) CHK(eng.compact());
    }
    CHK(eng.sync_ctrl());
    eng.finish_profile();
    const double t2 = now_s();
    if (out_expect) CHK(eng.expect(occ, out_expect));
    if (out) CHK(eng.download(out));
    const double t3 = now_s();
    ctx->prof.seconds_upload = t1 - t0;
    ctx->prof.seconds_propagate = t2 - t1;
    ctx->prof.seconds_download = t3 - t2;
    ctx->prof.seconds_total = t3 - t0;
    ctx->last_kind = 2;
    ctx->last_W = W;
    ctx->last_K = K;
    ctx->last_prec = std::is_same<CP, F64>::value ? FFX_FP64 : (std::is_same<CP, F32>::value ? FFX_FP32 : FFX_FP32C);
    return FFX_OK;
}

template <class Alg, class CP, int K>
int expect_last_impl(ffx_ctx* ctx, const uint64_t* state, double* out) {
    Engine<Alg, CP, K> eng(ctx);
    CHK(eng.sync_ctrl());
    return eng.expect(state, out);
}

inline int mono_degree_mod4_r(const uint64_t* m, int W) {
    int d = 0;
    for (int w = 0; w < W; ++w) d += __builtin_popcountll(m[w]);
    return (d >> 1) & 1;
}

// Batched propagation: one workgroup per instance (k_batch_run); instances that
// are not eligible or outgrow their slot budget run on the device-wide engine.
// Configuration of a batched launch: threads per instance workgroup, maximum
// window length, slot budget per instance (0 in a field: the tuned or default value).
// term stores the layer-group layout cuts a batch into by schedule length (1..4), each
// on its own stream, by default
constexpr int kMSplitDefault = 4;

struct BatchConf {
    int block = 0;
    int window = 0;
    int64_t slots = 0;
    int merged = -1;         // 1: one term store, layer-group steps (ffx_mlayer.h); 2: one term store,
                             // window steps (ffx_merged.h); 0: one workgroup each (ffx_batch.h)
    int compact_every = 0;   // layer-group steps between compaction checks
    int group_max = 0;       // most commuting rotations with overlapping supports in a group
    int split = -1;          // layer-group layout: term stores (1..4) the batch is cut into by schedule
                             // length, each on its own stream (-1: default)
};

// Merged batch (ffx_merged.h): the instances elig[0..ne) in one term store, one
// launch per window step of the longest schedule, and the device-decided
// compaction every compact_every() / 32 steps (default 8: with truncation after
// every rotation of 512 kicked-Ising instances, every 64 / 16 / 8 / 4 / 1 steps
// take 250 / 188 / 181 / 183 / 209 ms).  `ran` stays false (nothing done) when the
// instance numbers do not fit above every qubit the instances use.
template <int W, class CP, int K>
int run_pauli_merged(ffx_ctx* ctx, const ffx_instance* inst, const std::vector<int32_t>& elig,
                     const std::vector<std::vector<BStep<PauliAlg<W>, CP>>>& sched, double* out_expect,
                     int64_t* out_terms, bool& ran) {
    using Alg = PauliAlg<W>;
    using Eng = Engine<Alg, CP, K>;
    using store_t = typename CP::store_t;
    using Win = Window<Alg, CP>;
    constexpr int NP = Alg::NP;
    ran = false;
    const size_t ne = elig.size();
    if (ne == 0) return FFX_OK;
    int idbits = 1;
    while (((size_t)1 << idbits) < ne) ++idbits;
    uint64_t used[W] = {};
    for (int32_t i : elig) {
        const ffx_instance& in = inst[i];
        for (int64_t g = 0; g < in.n_gates; ++g)
            for (int w = 0; w < W; ++w) used[w] |= in.gates[g].x[w] | in.gates[g].z[w];
        for (int64_t t = 0; t < in.n_terms; ++t)
            for (int w = 0; w < W; ++w) used[w] |= in.x[t * W + w] | in.z[t * W + w];
    }
    int top = 0;
    for (int w = 0; w < W; ++w)
        if (used[w]) top = 64 * w + 64 - __builtin_clzll(used[w]);
    if (top > 64 * W - idbits) return FFX_OK;
    const uint32_t idshift = (uint32_t)(64 - idbits);
    const uint64_t idmask = (1ull << idbits) - 1;
    const uint64_t idbits_word = idmask << idshift;
    // observable terms, the instance number in the top bits of the last Z word
    size_t nt = 0;
    for (int32_t i : elig) nt += (size_t)inst[i].n_terms;
    std::vector<uint64_t> hx(nt * W), hz(nt * W);
    std::vector<double> hre(nt), him(nt);
    std::vector<uint64_t> hstate(ne * W);
    std::vector<uint32_t> wbase(ne), wlen(ne);
    size_t nw = 0, o = 0;
    uint32_t maxlen = 0;
    int64_t n_rot = 0;
    for (size_t b = 0; b < ne; ++b) {
        const ffx_instance& in = inst[elig[b]];
        for (int64_t t = 0; t < in.n_terms; ++t, ++o) {
            for (int w = 0; w < W; ++w) {
                hx[o * W + w] = in.x[t * W + w];
                hz[o * W + w] = in.z[t * W + w];
            }
            hz[o * W + W - 1] |= (uint64_t)b << idshift;
            hre[o] = in.c_re[t];
            him[o] = (K == 2 && in.c_im) ? in.c_im[t] : 0.0;
        }
        for (int w = 0; w < W; ++w) hstate[b * W + w] = in.basis_state[w] & (w == W - 1 ? ~idbits_word : ~0ull);
        wbase[b] = (uint32_t)nw;
        wlen[b] = (uint32_t)sched[elig[b]].size();
        nw += wlen[b];
        maxlen = std::max(maxlen, wlen[b]);
    }
    std::vector<Win> hw(std::max<size_t>(nw, 1));
    for (size_t b = 0; b < ne; ++b)
        for (uint32_t q = 0; q < wlen[b]; ++q) {
            Win w = sched[elig[b]][q].w;
            if (w.maxdeg < (1 << 29)) w.maxdeg += __builtin_popcountll((unsigned long long)b);
            n_rot += w.k;
            hw[wbase[b] + q] = w;
        }
    // device buffer: windows, wbase, wlen, states, sums, counts, work stripes
    auto al = [](size_t v) { return (v + 255) & ~(size_t)255; };
    const size_t o_wbase = al(hw.size() * sizeof(Win));
    const size_t o_wlen = o_wbase + al(ne * 4);
    const size_t o_state = o_wlen + al(ne * 4);
    const size_t o_out = o_state + al(ne * W * 8);
    const size_t o_cnt = o_out + al(2 * ne * 8);
    const size_t o_work = o_cnt + al(ne * 4);
    const size_t total = o_work + al((size_t)kStripes * 16 * 8);
    if (total > ctx->merged_buf_bytes) {
        dev_free(ctx->merged_buf);
        ctx->merged_buf_bytes = 0;
        CHK(dev_alloc(ctx, &ctx->merged_buf, total));
        ctx->merged_buf_bytes = total;
    }
    char* mb = (char*)ctx->merged_buf;
    HIPCHK(hipMemcpyAsync(mb, hw.data(), hw.size() * sizeof(Win), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(mb + o_wbase, wbase.data(), ne * 4, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(mb + o_wlen, wlen.data(), ne * 4, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(mb + o_state, hstate.data(), ne * W * 8, hipMemcpyHostToDevice, ctx->stream));
    Merged<Alg, CP> M{};
    M.wins = (const Win*)mb;
    M.wbase = (const uint32_t*)(mb + o_wbase);
    M.wlen = (const uint32_t*)(mb + o_wlen);
    M.idshift = idshift;
    M.idmask = idmask;
    M.work = (unsigned long long*)(mb + o_work);
    M.states = (const uint64_t*)(mb + o_state);
    M.out = (double*)(mb + o_out);
    M.count = (unsigned int*)(mb + o_cnt);
    Eng eng(ctx);
    const uint32_t every = (uint32_t)std::max(1, compact_every() / 32);
    if (ctx->n_cu <= 0) {
        hipDeviceProp_t prop;
        HIPCHK(hipGetDeviceProperties(&prop, ctx->device));
        ctx->n_cu = prop.multiProcessorCount;
    }
    const unsigned win_grid = 6 * (unsigned)ctx->n_cu;
    size_t want = std::max<size_t>(4 * nt + 4096, 1 << 16);
    std::vector<uint32_t> hn, hd, ha;
    std::vector<int32_t> hdd;
    uint32_t e = 1;
    int n_compact = 0;
    for (int attempt = 0;; ++attempt) {
        if (attempt >= 24) return set_err(ctx, FFX_ECAPACITY, "merged batch: term store keeps overflowing");
        CHK(eng.upload(hx.data(), hz.data(), hre.data(), K == 2 ? him.data() : nullptr, (int64_t)nt));
        CHK(eng.reserve(want));
        const size_t n_ep = 2 + (size_t)maxlen + maxlen / every + 1;
        CHK(eng.ensure_epochs(n_ep));
        CHK(eng.ensure_counts(ctx->cap / kTile + 2));
        const Epochs ep = eng.epochs();
        HIPCHK(hipMemsetAsync(ctx->epochs, 0, Eng::epoch_words(ctx->epochs_cap) * sizeof(uint32_t), ctx->stream));
        HIPCHK(hipMemsetAsync(mb + o_work, 0, (size_t)kStripes * 16 * 8, ctx->stream));
        HIPCHK(hipMemsetAsync(mb + o_out, 0, 2 * ne * 8, ctx->stream));
        HIPCHK(hipMemsetAsync(mb + o_cnt, 0, ne * 4, ctx->stream));
        const uint32_t init[2] = {(uint32_t)nt, 0u};
        HIPCHK(hipMemcpyAsync(ep.n, &init[0], sizeof(uint32_t), hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(hipMemcpyAsync(ep.dead, &init[1], sizeof(uint32_t), hipMemcpyHostToDevice, ctx->stream));
        const unsigned g_cap = grid_for(ctx->cap, kTile, 4096);
        const unsigned g_idx = grid_for(std::max(ctx->cap, ctx->hslots), kBlock, 2048);
        e = 1;
        n_compact = 0;
        for (uint32_t s = 0; s < maxlen; ++s) {
            M.step = s;
            hipLaunchKernelGGL((k_mwin_async<Alg, CP, K>), dim3(win_grid), dim3(kBlockWin), 0, ctx->stream,
                               eng.store(), (const store_t*)ctx->coef, (store_t*)ctx->coef2, M, ep, e);
            CHK(eng.launch_check());
            std::swap(ctx->coef, ctx->coef2);
            ++e;
            if ((s + 1) % every == 0 && s + 1 < maxlen) {
                const uint32_t ep_prev = e - 1;
                hipLaunchKernelGGL((k_cmp_count<CP, K>), dim3(g_cap), dim3(kBlock), 0, ctx->stream, eng.store(), ep,
                                   ep_prev, ctx->counts);
                CHK(eng.launch_check());
                hipLaunchKernelGGL(k_cmp_scan, dim3(1), dim3(kBlock), 0, ctx->stream, ep, ep_prev, ctx->counts);
                CHK(eng.launch_check());
                hipLaunchKernelGGL((k_cmp_scatter<NP, CP, K>), dim3(g_cap), dim3(kBlock), 0, ctx->stream,
                                   eng.store(), ctx->key2, (store_t*)ctx->coef2, ep, ep_prev, e, ctx->counts);
                CHK(eng.launch_check());
                hipLaunchKernelGGL((k_cmp_copyback<NP, CP, K>), dim3(g_idx), dim3(kBlock), 0, ctx->stream,
                                   eng.store(), ctx->key2, (const store_t*)ctx->coef2, ep, ep_prev, e);
                CHK(eng.launch_check());
                hipLaunchKernelGGL(k_index_build_async<NP>, dim3(g_idx), dim3(kBlock), 0, ctx->stream, ctx->key,
                                   ctx->cap, ctx->index, ctx->hslots - 1, ep, ep_prev, e, 1, &ctx->ctrl->index_full);
                CHK(eng.launch_check());
                ++e;
                ++n_compact;
            }
        }
        hn.assign(e, 0);
        hd.assign(e, 0);
        ha.assign(e, 0);
        hdd.assign(e, 0);
        uint32_t hf[2] = {0, 0};
        HIPCHK(hipMemcpyAsync(hn.data(), ep.n, e * 4, hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipMemcpyAsync(hd.data(), ep.dead, e * 4, hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipMemcpyAsync(ha.data(), ep.app, e * 4, hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipMemcpyAsync(hdd.data(), ep.dd, e * 4, hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipMemcpyAsync(hf, ep.flags, 8, hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipMemcpyAsync(ctx->ctrl_host, ctx->ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        if (ctx->ctrl_host->index_full)
            return set_err(ctx, FFX_EHIP, "internal error: hash index full (flags " +
                                              std::to_string(ctx->ctrl_host->index_full) + ")");
        if (!hf[0]) break;
        want = 2 * ctx->cap;  // an append did not fit: a larger store, from the start
    }
    // slots and erased slots after the last launch, as the next launch would see them
    const uint32_t n_fin = hn[e - 1] + ha[e - 1];
    int64_t peak = 0;
    for (uint32_t q = 1; q < e; ++q)
        peak = std::max<int64_t>(peak, (int64_t)(hn[q] + ha[q]) - (int64_t)((int32_t)hd[q] + hdd[q]));
    if (n_fin > 0) {
        hipLaunchKernelGGL((k_mexpect<Alg, CP, K>), dim3(grid_for(n_fin, kBlock, 4096)), dim3(kBlock), 0, ctx->stream,
                           eng.store(), n_fin, (const uint64_t*)(mb + o_state), idshift, idmask, (double*)(mb + o_out),
                           (unsigned int*)(mb + o_cnt));
        CHK(eng.launch_check());
    }
    std::vector<double> hout(2 * ne);
    std::vector<uint32_t> hcnt(ne);
    std::vector<unsigned long long> hwork((size_t)kStripes * 16);
    HIPCHK(hipMemcpyAsync(hout.data(), mb + o_out, 2 * ne * 8, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipMemcpyAsync(hcnt.data(), mb + o_cnt, ne * 4, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipMemcpyAsync(hwork.data(), mb + o_work, hwork.size() * 8, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    for (size_t b = 0; b < ne; ++b) {
        out_expect[2 * elig[b]] = hout[2 * b];
        out_expect[2 * elig[b] + 1] = hout[2 * b + 1];
        if (out_terms) out_terms[elig[b]] = (int64_t)hcnt[b];
    }
    for (int c = 0; c < kStripes; ++c) {
        ctx->prof.updates += (int64_t)hwork[(size_t)c * 16];
        ctx->prof.outs += (int64_t)hwork[(size_t)c * 16 + 1];
    }
    ctx->prof.n_gates += n_rot;
    ctx->prof.peak_terms = std::max(ctx->prof.peak_terms, peak);
    ctx->prof.run_launches += maxlen;
    ctx->prof.compactions += n_compact;
    // the store holds every instance's terms: no device-resident result of one call
    ctx->last_kind = 0;
    ran = true;
    return FFX_OK;
}

// One instance's schedule cut into layer groups (ffx_mlayer.h): at each rotation
// the longer of a run on pairwise disjoint supports (at most kMGroupMax) and a
// run of pairwise commuting rotations with independent axes (at most
// kLayerMaxA; a single rotation is one), the disjoint run on ties.  Offsets in
// the records are local to the instance.  Groups with the same rotation records
// (the same layer in every Trotter step) share one copy of the records and of
// the position table.  False if a step is not a rotation.
template <int W, class CP>
struct MInstPlan {
    std::vector<MGroup<W>> groups;
    std::vector<DLayerGate<PauliAlg<W>, CP>> gates;
    std::vector<uint64_t> duals;  // 2W words per gate record: the dual vector in general groups, else 0
    std::vector<uint8_t> pos;
    int64_t n_rot = 0;  // rotations of the schedule (records are shared, so gates.size() is fewer)
};

template <int W, class CP, int K>
bool plan_mgroups(const std::vector<typename Engine<PauliAlg<W>, CP, K>::AStep>& sd, MInstPlan<W, CP>& P,
                  int group_max) {
    using Eng = Engine<PauliAlg<W>, CP, K>;
    constexpr int NP = 2 * W;
    size_t dmax = kMGroupMax;
    size_t gmax = group_max > 0 ? std::min<size_t>(kLayerMaxA, (size_t)group_max) : (size_t)kLayerMaxA;
    if (const char* ev = getenv("FFX_MGROUP_MAX"))
        if (atoi(ev) > 0) dmax = std::min<size_t>(kMGroupMax, (size_t)atoi(ev));
    if (const char* ev = getenv("FFX_MGENERAL_MAX"))
        if (atoi(ev) > 0) gmax = std::min<size_t>(kLayerMaxA, (size_t)atoi(ev));
    using Gate = DLayerGate<PauliAlg<W>, CP>;
    P.groups.clear();
    P.gates.clear();
    P.duals.clear();
    P.pos.clear();
    P.n_rot = (int64_t)sd.size();
    for (const auto& a : sd)
        if (a.kind != 0) return false;
    // first group seen with each hash of (records, kind); a hit with equal bytes
    // takes that group's records instead of a new copy
    std::unordered_map<uint64_t, uint32_t> seen;
    std::unordered_map<uint64_t, std::pair<size_t, size_t>> defseen;  // hash -> (first step, group)
    for (size_t t = 0; t < sd.size();) {
        // run on pairwise disjoint supports
        size_t kd = 0;
        {
            uint64_t used[4] = {0, 0, 0, 0};
            while (t + kd < sd.size() && kd < dmax) {
                uint64_t u[4];
                Eng::axis_units(sd[t + kd].g, u);
                int cnt = 0;
                bool overlap = false;
                for (int w = 0; w < 4; ++w) {
                    cnt += __builtin_popcountll(u[w]);
                    overlap = overlap || (u[w] & used[w]) != 0;
                }
                if (cnt == 0 || overlap) break;
                for (int w = 0; w < 4; ++w) used[w] |= u[w];
                ++kd;
            }
        }
        // run of pairwise commuting rotations with linearly independent axes (not needed
        // when the disjoint run already reaches gmax: the disjoint one wins ties)
        size_t kg = 0;
        if (!(kd >= 2 && kd >= gmax)) {
            uint64_t basis[kLayerMaxA][NP];
            int pivot[kLayerMaxA];
            int nb = 0;
            while (t + kg < sd.size() && kg < gmax) {
                const auto& a = sd[t + kg];
                bool ok = true;
                for (size_t j = 0; j < kg && ok; ++j) ok = Eng::axes_commute(sd[t + j].g, a.g);
                if (!ok) break;
                uint64_t v[NP];
                for (int q = 0; q < NP; ++q) v[q] = a.g.p[q];
                for (int b = 0; b < nb; ++b)
                    if ((v[pivot[b] >> 6] >> (pivot[b] & 63)) & 1)
                        for (int q = 0; q < NP; ++q) v[q] ^= basis[b][q];
                int q0 = -1;
                for (int q = 0; q < NP && q0 < 0; ++q)
                    if (v[q]) q0 = q;
                if (q0 < 0) break;  // dependent (or the identity)
                pivot[nb] = 64 * q0 + __builtin_ctzll(v[q0]);
                for (int q = 0; q < NP; ++q) basis[nb][q] = v[q];
                ++nb;
                ++kg;
            }
        }
        const bool general = !(kd >= 2 && kd >= kg);
        const size_t k = general ? std::max<size_t>(kg, 1) : kd;
        // a group with the same rotations (axes, angles, truncation points) and kind as an
        // earlier one (the same layer in another Trotter step) reuses its records as is
        uint64_t dh = 0xC2B2AE3D27D4EB4Full ^ (uint64_t)k ^ ((uint64_t)general << 40);
        auto mixin = [&](uint64_t v) {
            dh = (dh ^ v) * 0xFF51AFD7ED558CCDull;
            dh ^= dh >> 29;
        };
        for (size_t j = 0; j < k; ++j) {
            const auto& a = sd[t + j];
            for (int q = 0; q < NP; ++q) mixin(a.g.p[q]);
            uint64_t b;
            std::memcpy(&b, &a.cs, 8);
            mixin(b);
            std::memcpy(&b, &a.sn, 8);
            mixin(b ^ ((uint64_t)a.trunc_thresh << 62) ^ ((uint64_t)a.trunc_xw << 61));
        }
        {
            const auto it = defseen.find(dh);
            if (it != defseen.end()) {
                const size_t t0 = it->second.first;
                const MGroup<W>& E = P.groups[it->second.second];
                bool same = E.nr == (uint32_t)k && (E.general != 0) == general;
                for (size_t j = 0; j < k && same; ++j) {
                    const auto& a = sd[t + j];
                    const auto& b = sd[t0 + j];
                    for (int q = 0; q < NP; ++q) same = same && a.g.p[q] == b.g.p[q];
                    same = same && std::memcmp(&a.cs, &b.cs, 8) == 0 && std::memcmp(&a.sn, &b.sn, 8) == 0 &&
                           a.trunc_thresh == b.trunc_thresh && a.trunc_xw == b.trunc_xw;
                }
                if (same) {
                    P.groups.push_back(E);
                    t += k;
                    continue;
                }
            } else {
                defseen.emplace(dh, std::make_pair(t, P.groups.size()));
            }
        }
        MGroup<W> G{};
        G.gate_off = (uint32_t)P.gates.size();
        G.pos_off = (uint32_t)P.pos.size();
        G.nr = (uint32_t)k;
        G.general = general ? 1u : 0u;
        uint8_t upos[64 * W];
        for (size_t j = 0; j < k; ++j) {
            const auto& a = sd[t + j];
            DLayerGate<PauliAlg<W>, CP> r{};
            r.g = a.g;
            r.cs = make_fac<CP>(a.cs);
            r.sn = make_fac<CP>(a.sn);
            r.pivot = -1;
            for (int q = 0; q < NP && r.pivot < 0; ++q)
                if (a.g.p[q]) r.pivot = 64 * q + __builtin_ctzll(a.g.p[q]);
            if (a.trunc_thresh) G.tmask |= 1ull << j;
            if (a.trunc_xw) G.xmask |= 1ull << j;
            uint64_t u[4];
            Eng::axis_units(a.g, u);
            for (int w = 0; w < W; ++w) {
                G.umask[w] |= u[w];
                if (!general)
                    for (uint64_t b = u[w]; b; b &= b - 1) upos[64 * w + __builtin_ctzll(b)] = (uint8_t)j;
            }
            P.gates.push_back(r);
        }
        if (!general)
            for (int w = 0; w < W; ++w)
                for (uint64_t b = G.umask[w]; b; b &= b - 1) P.pos.push_back(upos[64 * w + __builtin_ctzll(b)]);
        {
            uint64_t dual[kLayerMaxA][NP] = {};
            if (general && !Eng::general_duals(&sd[t], (int)k, dual)) return false;
            for (size_t j = 0; j < k; ++j)
                for (int q = 0; q < NP; ++q) P.duals.push_back(general ? dual[j][q] : 0ull);
        }
        {
            // the position table is a function of the records and the kind, so equal
            // records of the same kind share it too
            const unsigned char* gb = (const unsigned char*)(P.gates.data() + G.gate_off);
            const size_t nb = k * sizeof(Gate);
            uint64_t h = 0x9E3779B97F4A7C15ull ^ (uint64_t)k ^ ((uint64_t)G.general << 32);
            for (size_t q = 0; q + 8 <= nb; q += 8) {
                uint64_t v;
                std::memcpy(&v, gb + q, 8);
                h = (h ^ v) * 0xFF51AFD7ED558CCDull;
                h ^= h >> 29;
            }
            const auto it = seen.find(h);
            if (it == seen.end()) {
                seen.emplace(h, (uint32_t)P.groups.size());
            } else {
                const MGroup<W>& E = P.groups[it->second];
                if (E.nr == G.nr && E.general == G.general &&
                    std::memcmp(P.gates.data() + E.gate_off, gb, nb) == 0) {
                    P.gates.resize(G.gate_off);
                    P.duals.resize((size_t)G.gate_off * NP);
                    P.pos.resize(G.pos_off);
                    G.gate_off = E.gate_off;
                    G.pos_off = E.pos_off;
                }
            }
        }
        P.groups.push_back(G);
        t += k;
    }
    return true;
}

// Merged batch in layer groups (ffx_mlayer.h): the instances elig[0..ne) in one
// term store, a claim and a replay launch per group step of the longest schedule,
// the device-decided compaction every `compact_every` steps (default 4;
// FFX_MCOMPACT overrides).  `ran` stays
// false (nothing done) when the instance numbers do not fit above every qubit
// the instances use (`used`: the union of their qubit masks).
template <int W, class CP, int K>
int run_pauli_mlayer(ffx_ctx* ctx, const ffx_instance* inst, const std::vector<int32_t>& elig,
                     const std::vector<MInstPlan<W, CP>>& plans, const uint64_t (&used)[W], const ffx_options& o,
                     int compact_every, double* out_expect, int64_t* out_terms, bool& ran) {
    using Alg = PauliAlg<W>;
    using Eng = Engine<Alg, CP, K>;
    using store_t = typename CP::store_t;
    using Gate = DLayerGate<Alg, CP>;
    constexpr int NP = Alg::NP;
    ran = false;
    const size_t ne = elig.size();
    if (ne == 0) return FFX_OK;
    int idbits = 1;
    while (((size_t)1 << idbits) < ne) ++idbits;
    int top = 0;
    for (int w = 0; w < W; ++w)
        if (used[w]) top = 64 * w + 64 - __builtin_clzll(used[w]);
    if (top > 64 * W - idbits) return FFX_OK;
    const uint32_t idshift = (uint32_t)(64 - idbits);
    const uint64_t idmask = (1ull << idbits) - 1;
    const uint64_t idbits_word = idmask << idshift;
    // sizes and per-instance offsets
    std::vector<uint32_t> gbase(ne), glen(ne), gate_base(ne), pos_base(ne);
    size_t ng = 0, ngate = 0, npos = 0, nt = 0;
    uint32_t maxlen = 0;
    int64_t n_rot = 0;
    for (size_t b = 0; b < ne; ++b) {
        const MInstPlan<W, CP>& P = plans[elig[b]];
        gbase[b] = (uint32_t)ng;
        glen[b] = (uint32_t)P.groups.size();
        gate_base[b] = (uint32_t)ngate;
        pos_base[b] = (uint32_t)npos;
        ng += P.groups.size();
        ngate += P.gates.size();
        npos += P.pos.size();
        nt += (size_t)inst[elig[b]].n_terms;
        maxlen = std::max(maxlen, glen[b]);
        n_rot += P.n_rot;
    }
    // device buffer: groups, gates, positions, gbase, glen, states | sums, counts, work stripes
    auto al = [](size_t v) { return (v + 255) & ~(size_t)255; };
    const size_t o_gate = al(std::max<size_t>(ng, 1) * sizeof(MGroup<W>));
    const size_t o_dual = o_gate + al(std::max<size_t>(ngate, 1) * sizeof(Gate));
    const size_t o_pos = o_dual + al(std::max<size_t>(ngate, 1) * NP * sizeof(uint64_t));
    const size_t o_gbase = o_pos + al(std::max<size_t>(npos, 1));
    const size_t o_glen = o_gbase + al(ne * 4);
    const size_t o_state = o_glen + al(ne * 4);
    const size_t o_out = o_state + al(ne * W * 8);
    const size_t o_cnt = o_out + al(2 * ne * 8);
    const size_t o_work = o_cnt + al(ne * 4);
    const size_t total = o_work + al((size_t)kStripes * 16 * 8);
    if (total > ctx->merged_buf_bytes) {
        dev_free(ctx->merged_buf);
        ctx->merged_buf_bytes = 0;
        CHK(dev_alloc(ctx, &ctx->merged_buf, total));
        ctx->merged_buf_bytes = total;
    }
    if (o_out > ctx->mlayer_host_bytes) {
        if (ctx->mlayer_host) (void)hipHostFree(ctx->mlayer_host);
        ctx->mlayer_host = nullptr;
        ctx->mlayer_host_bytes = 0;
        HIPCHK(hipHostMalloc(&ctx->mlayer_host, o_out, hipHostMallocDefault));
        ctx->mlayer_host_bytes = o_out;
    }
    char* H = (char*)ctx->mlayer_host;
    {
        // the pinned image, instance by instance on host threads (offsets made global)
        const int nth = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        std::atomic<size_t> next{0};
        auto fill = [&]() {
            for (size_t b = next++; b < ne; b = next++) {
                const MInstPlan<W, CP>& P = plans[elig[b]];
                MGroup<W>* hg = (MGroup<W>*)H + gbase[b];
                for (size_t q = 0; q < P.groups.size(); ++q) {
                    MGroup<W> G = P.groups[q];
                    G.gate_off += gate_base[b];
                    G.pos_off += pos_base[b];
                    hg[q] = G;
                }
                std::memcpy(H + o_gate + (size_t)gate_base[b] * sizeof(Gate), P.gates.data(),
                            P.gates.size() * sizeof(Gate));
                std::memcpy(H + o_dual + (size_t)gate_base[b] * NP * sizeof(uint64_t), P.duals.data(),
                            P.duals.size() * sizeof(uint64_t));
                if (!P.pos.empty()) std::memcpy(H + o_pos + pos_base[b], P.pos.data(), P.pos.size());
                const ffx_instance& in = inst[elig[b]];
                uint64_t* st = (uint64_t*)(H + o_state) + b * W;
                for (int w = 0; w < W; ++w) st[w] = in.basis_state[w] & (w == W - 1 ? ~idbits_word : ~0ull);
            }
        };
        std::vector<std::thread> pool;
        const int nuse = (int)std::min<size_t>((size_t)nth, ne);
        for (int t = 1; t < nuse; ++t) pool.emplace_back(fill);
        fill();
        for (auto& th : pool) th.join();
        std::memcpy(H + o_gbase, gbase.data(), ne * 4);
        std::memcpy(H + o_glen, glen.data(), ne * 4);
    }
    char* mb = (char*)ctx->merged_buf;
    HIPCHK(hipMemcpyAsync(mb, H, o_out, hipMemcpyHostToDevice, ctx->stream));
    // observable terms, the instance number in the top bits of the last Z word
    std::vector<uint64_t> hx(nt * W), hz(nt * W);
    std::vector<double> hre(nt), him(nt);
    {
        size_t q = 0;
        for (size_t b = 0; b < ne; ++b) {
            const ffx_instance& in = inst[elig[b]];
            for (int64_t t = 0; t < in.n_terms; ++t, ++q) {
                for (int w = 0; w < W; ++w) {
                    hx[q * W + w] = in.x[t * W + w];
                    hz[q * W + w] = in.z[t * W + w];
                }
                hz[q * W + W - 1] |= (uint64_t)b << idshift;
                hre[q] = in.c_re[t];
                him[q] = (K == 2 && in.c_im) ? in.c_im[t] : 0.0;
            }
        }
    }
    MLayer<Alg, CP> M{};
    M.groups = (const MGroup<W>*)mb;
    M.gates = (const Gate*)(mb + o_gate);
    M.duals = (const uint64_t*)(mb + o_dual);
    M.pos = (const uint8_t*)(mb + o_pos);
    M.gbase = (const uint32_t*)(mb + o_gbase);
    M.glen = (const uint32_t*)(mb + o_glen);
    M.states = (const uint64_t*)(mb + o_state);
    M.out = (double*)(mb + o_out);
    M.count = (unsigned int*)(mb + o_cnt);
    M.work = (unsigned long long*)(mb + o_work);
    M.idshift = idshift;
    M.idmask = idmask;
    M.mincoeff = o.mincoeff;
    M.maxdeg = degree_limit(o.maxdegree);
    M.max_xw = o.max_xweight;
    Eng eng(ctx);
    uint32_t every = compact_every > 0 ? (uint32_t)compact_every : 4u;
    if (const char* ev = getenv("FFX_MCOMPACT"))
        if (atoi(ev) > 0) every = (uint32_t)atoi(ev);
    if (ctx->n_cu <= 0) {
        hipDeviceProp_t prop;
        HIPCHK(hipGetDeviceProperties(&prop, ctx->device));
        ctx->n_cu = prop.multiProcessorCount;
    }
    unsigned grid = 6 * (unsigned)ctx->n_cu;
    if (const char* ev = getenv("FFX_WIN_GRID"))
        if (atoi(ev) > 0) grid = (unsigned)atoi(ev);
    // the replay launch: one resident round at its occupancy (3 blocks per CU for keys of
    // up to 3 words, else 2); the listed passes are spread over whatever grid runs them
    // (512-instance kicked-Ising batch: 77 -> 75 ms against the claim launch's 6 per CU)
    const unsigned rgrid = (Alg::NP <= 6 ? 3u : 2u) * (unsigned)ctx->n_cu;
    // the store: as large as the last call's (a batch of the same shape needs the same),
    // doubled from the start whenever an append does not fit
    size_t want = std::max<size_t>(std::max<size_t>(4 * nt + 4096, 1 << 16), ctx->mlayer_cap_hint);
    std::vector<uint32_t> hn, hd, ha;
    std::vector<int32_t> hdd;
    uint32_t e = 1;
    int n_compact = 0;
    for (int attempt = 0;; ++attempt) {
        if (attempt >= 24) return set_err(ctx, FFX_ECAPACITY, "merged batch: term store keeps overflowing");
        CHK(eng.upload(hx.data(), hz.data(), hre.data(), K == 2 ? him.data() : nullptr, (int64_t)nt));
        CHK(eng.reserve(want, true));
        if (ctx->claim_slots != ctx->hslots || !ctx->claim) {
            dev_free(ctx->claim);
            ctx->claim_slots = 0;
            CHK(dev_alloc(ctx, &ctx->claim, ctx->hslots * sizeof(uint64_t)));
            ctx->claim_slots = ctx->hslots;
            HIPCHK(hipMemsetAsync(ctx->claim, 0, ctx->hslots * sizeof(uint64_t), ctx->stream));
            ctx->layer_stamp = 0;
        }
        M.claim = ctx->claim;
        // orbit leader lists of the replay launches: a stripe's blocks hold at most
        // ceil(grid / kMStripes) * (slots per block) leaders of each orbit size
        {
            const size_t per_block = (ctx->cap + (size_t)grid * kBlock - 1) / ((size_t)grid * kBlock) * kBlock;
            M.ol_cap = (uint32_t)std::min<size_t>((grid + kMStripes - 1) / kMStripes * per_block, 0xFFFFFFFFu);
            const size_t lists = (size_t)kLayerMaxA * kMStripes;
            const size_t lb = (lists * M.ol_cap * 4 + 255) & ~(size_t)255;
            const size_t cb = lists * (size_t)maxlen * 4;
            if (lb + cb > ctx->mlayer_q_bytes) {
                dev_free(ctx->mlayer_q);
                ctx->mlayer_q_bytes = 0;
                CHK(dev_alloc(ctx, &ctx->mlayer_q, lb + cb));
                ctx->mlayer_q_bytes = lb + cb;
            }
            M.ol = (uint32_t*)ctx->mlayer_q;
            M.ol_cnt = (uint32_t*)((char*)ctx->mlayer_q + lb);
            HIPCHK(hipMemsetAsync(M.ol_cnt, 0, cb, ctx->stream));
        }
        const size_t n_ep = 2 + (size_t)maxlen + maxlen / every + 1;
        CHK(eng.ensure_epochs(n_ep));
        CHK(eng.ensure_counts(ctx->cap / kTile + 2));
        const Epochs ep = eng.epochs();
        HIPCHK(hipMemsetAsync(ctx->epochs, 0, Eng::epoch_words(ctx->epochs_cap) * sizeof(uint32_t), ctx->stream));
        HIPCHK(hipMemsetAsync(mb + o_work, 0, (size_t)kStripes * 16 * 8, ctx->stream));
        HIPCHK(hipMemsetAsync(mb + o_out, 0, 2 * ne * 8, ctx->stream));
        HIPCHK(hipMemsetAsync(mb + o_cnt, 0, ne * 4, ctx->stream));
        const uint32_t init[2] = {(uint32_t)nt, 0u};
        HIPCHK(hipMemcpyAsync(ep.n, &init[0], sizeof(uint32_t), hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(hipMemcpyAsync(ep.dead, &init[1], sizeof(uint32_t), hipMemcpyHostToDevice, ctx->stream));
        const unsigned g_cap = grid_for(ctx->cap, kTile, 4096);
        const unsigned g_idx = grid_for(std::max(ctx->cap, ctx->hslots), kBlock, 2048);
        e = 1;
        n_compact = 0;
        for (uint32_t s = 0; s < maxlen; ++s) {
            if (++ctx->layer_stamp >= (1u << 24)) {
                HIPCHK(hipMemsetAsync(ctx->claim, 0, ctx->hslots * sizeof(uint64_t), ctx->stream));
                ctx->layer_stamp = 1;
            }
            M.stamp = ctx->layer_stamp;
            M.step = s;
            hipLaunchKernelGGL((k_mlayer_claim<Alg, CP, K>), dim3(grid), dim3(kBlock), 0, ctx->stream, eng.store(),
                               (const store_t*)ctx->coef, (store_t*)ctx->coef2, M, ep, e);
            CHK(eng.launch_check());
            hipLaunchKernelGGL((k_mlayer_replay<Alg, CP, K>), dim3(rgrid), dim3(kBlock), 0, ctx->stream, eng.store(),
                               (const store_t*)ctx->coef, (store_t*)ctx->coef2, M, ep, e);
            CHK(eng.launch_check());
            std::swap(ctx->coef, ctx->coef2);
            ++e;
            if ((s + 1) % every == 0 && s + 1 < maxlen) {
                const uint32_t ep_prev = e - 1;
                hipLaunchKernelGGL((k_cmp_count<CP, K>), dim3(g_cap), dim3(kBlock), 0, ctx->stream, eng.store(), ep,
                                   ep_prev, ctx->counts);
                CHK(eng.launch_check());
                hipLaunchKernelGGL(k_cmp_scan, dim3(1), dim3(kBlock), 0, ctx->stream, ep, ep_prev, ctx->counts);
                CHK(eng.launch_check());
                hipLaunchKernelGGL((k_cmp_scatter<NP, CP, K>), dim3(g_cap), dim3(kBlock), 0, ctx->stream,
                                   eng.store(), ctx->key2, (store_t*)ctx->coef2, ep, ep_prev, e, ctx->counts);
                CHK(eng.launch_check());
                hipLaunchKernelGGL((k_cmp_copyback<NP, CP, K>), dim3(g_idx), dim3(kBlock), 0, ctx->stream,
                                   eng.store(), ctx->key2, (const store_t*)ctx->coef2, ep, ep_prev, e);
                CHK(eng.launch_check());
                hipLaunchKernelGGL(k_index_build_async<NP>, dim3(g_idx), dim3(kBlock), 0, ctx->stream, ctx->key,
                                   ctx->cap, ctx->index, ctx->hslots - 1, ep, ep_prev, e, 1, &ctx->ctrl->index_full);
                CHK(eng.launch_check());
                ++e;
                ++n_compact;
            }
        }
        hn.assign(e, 0);
        hd.assign(e, 0);
        ha.assign(e, 0);
        hdd.assign(e, 0);
        uint32_t hf[2] = {0, 0};
        HIPCHK(hipMemcpyAsync(hn.data(), ep.n, e * 4, hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipMemcpyAsync(hd.data(), ep.dead, e * 4, hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipMemcpyAsync(ha.data(), ep.app, e * 4, hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipMemcpyAsync(hdd.data(), ep.dd, e * 4, hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipMemcpyAsync(hf, ep.flags, 8, hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipMemcpyAsync(ctx->ctrl_host, ctx->ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        if (ctx->ctrl_host->index_full)
            return set_err(ctx, FFX_EHIP, "internal error: hash index full or layer degree bound broken (flags " +
                                              std::to_string(ctx->ctrl_host->index_full) + ")");
        if (!hf[0]) break;
        want = 2 * ctx->cap;  // an append did not fit: a larger store, from the start
    }
    // slots after the last launch, as the next launch would see them
    const uint32_t n_fin = hn[e - 1] + ha[e - 1];
    int64_t peak = 0;
    for (uint32_t q = 1; q < e; ++q)
        peak = std::max<int64_t>(peak, (int64_t)(hn[q] + ha[q]) - (int64_t)((int32_t)hd[q] + hdd[q]));
    if (n_fin > 0) {
        hipLaunchKernelGGL((k_mexpect<Alg, CP, K>), dim3(grid_for(n_fin, kBlock, 4096)), dim3(kBlock), 0, ctx->stream,
                           eng.store(), n_fin, (const uint64_t*)(mb + o_state), idshift, idmask, (double*)(mb + o_out),
                           (unsigned int*)(mb + o_cnt));
        CHK(eng.launch_check());
    }
    std::vector<double> hout(2 * ne);
    std::vector<uint32_t> hcnt(ne);
    std::vector<unsigned long long> hwork((size_t)kStripes * 16);
    HIPCHK(hipMemcpyAsync(hout.data(), mb + o_out, 2 * ne * 8, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipMemcpyAsync(hcnt.data(), mb + o_cnt, ne * 4, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipMemcpyAsync(hwork.data(), mb + o_work, hwork.size() * 8, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    for (size_t b = 0; b < ne; ++b) {
        out_expect[2 * elig[b]] = hout[2 * b];
        out_expect[2 * elig[b] + 1] = hout[2 * b + 1];
        if (out_terms) out_terms[elig[b]] = (int64_t)hcnt[b];
    }
    for (int c = 0; c < kStripes; ++c) {
        ctx->prof.updates += (int64_t)hwork[(size_t)c * 16];
        ctx->prof.outs += (int64_t)hwork[(size_t)c * 16 + 1];
    }
    ctx->prof.n_gates += n_rot;
    ctx->prof.peak_terms = std::max(ctx->prof.peak_terms, peak);
    ctx->prof.run_launches += maxlen;
    ctx->prof.compactions += n_compact;
    ctx->last_kind = 0;  // the store holds every instance's terms: no device-resident result of one call
    ctx->mlayer_cap_hint = ctx->cap;
    ran = true;
    return FFX_OK;
}

template <int W, class CP, int K>
int run_pauli_batch(ffx_ctx* ctx, const ffx_instance* inst, int32_t n_inst, const ffx_options& o, int64_t slots,
                    double* out_expect, int64_t* out_terms, int32_t* out_fallback, const BatchConf* conf = nullptr) {
    using Alg = PauliAlg<W>;
    using Eng = Engine<Alg, CP, K>;
    using AStep = typename Eng::AStep;
    using store_t = typename CP::store_t;
    constexpr int NP = Alg::NP;
    const double t0 = now_s();
    const int maxdeg = degree_limit(o.maxdegree);
    // configuration: explicit > autotuned for this (W, precision, K) > defaults
    const int prec_id = std::is_same<CP, F64>::value ? FFX_FP64 : (std::is_same<CP, F32>::value ? FFX_FP32 : FFX_FP32C);
    BatchConf cf{kBlockBatch, kWin, 32768, 1, 0, 0, kMSplitDefault};
    for (const auto& t : ctx->batch_tuned)
        if (t.W == W && t.prec == prec_id && t.K == K)
            cf = BatchConf{t.block, t.window, t.slots, t.merged, t.compact_every, t.group_max, t.split};
    if (conf) {
        if (conf->block) cf.block = conf->block;
        if (conf->window) cf.window = conf->window;
        if (conf->slots) cf.slots = conf->slots;
        if (conf->merged >= 0) cf.merged = conf->merged;
        if (conf->compact_every) cf.compact_every = conf->compact_every;
        if (conf->group_max) cf.group_max = conf->group_max;
        if (conf->split >= 0) cf.split = conf->split;
    }
    if (cf.split < 0) cf.split = kMSplitDefault;
    if (const char* ev = getenv("FFX_MSPLIT")) cf.split = atoi(ev);
    cf.split = std::max(1, std::min(4, cf.split));
    if (const char* ev = getenv("FFX_BATCH_MERGED")) cf.merged = std::max(0, std::min(2, atoi(ev)));
    if (slots > 0) cf.slots = slots;
    if (const char* ev = getenv("FFX_BATCH_BLOCK")) {
        const int v = atoi(ev);
        if (v == 64 || v == 128 || v == 256 || v == 512) cf.block = v;
    }
    if (cf.block != 64 && cf.block != 128 && cf.block != 256 && cf.block != 512)
        return set_err(ctx, FFX_EINVAL, "batch workgroup size must be 64, 128, 256 or 512");
    cf.window = std::max(1, std::min(kWin, cf.window));
    int64_t max_terms = 1;
    for (int32_t i = 0; i < n_inst; ++i) max_terms = std::max(max_terms, inst[i].n_terms);
    const size_t S = pow2ceil((size_t)std::max<int64_t>(std::max<int64_t>(cf.slots, 2 * max_terms), 1024));
    if (S > 0x40000000ull) return set_err(ctx, FFX_ECAPACITY, "slot budget too large");
    // layer groups in the merged layout (ffx_mlayer.h) need every live term's degree
    // <= kLayerMaxA: a degree limit that low and observables within it
    bool try_ml = cf.merged == 1 && getenv("FFX_NO_MLAYER") == nullptr && maxdeg <= kLayerMaxA;
    const bool merged_any = cf.merged == 1 || cf.merged == 2;
    for (int32_t i = 0; i < n_inst && try_ml; ++i)
        for (int64_t t = 0; t < inst[i].n_terms && try_ml; ++t) {
            int d = 0;
            for (int w = 0; w < W; ++w) d += __builtin_popcountll(inst[i].x[t * W + w] | inst[i].z[t * W + w]);
            try_ml = d <= kLayerMaxA;
        }
    // schedules: eligible instances get their layer groups or windows, the others the
    // device-wide engine
    std::vector<std::vector<BStep<Alg, CP>>> sched(n_inst);
    std::vector<MInstPlan<W, CP>> plans(try_ml ? n_inst : 0);
    std::vector<char> ok(n_inst, 0), done(n_inst, 0), have_windows(n_inst, 0);
    std::vector<double> est(n_inst, 0.0);
    std::vector<std::array<uint64_t, W>> used_of(n_inst);
    auto rotation_steps = [&](int32_t i, std::vector<AStep>& sd) {
        const std::vector<Step> steps = pauli_schedule<W>(inst[i].gates, inst[i].n_gates, o.batched != 0);
        std::vector<CliffGate> cl;
        pauli_async_steps<W, CP, K>(inst[i].gates, steps, o, sd, cl);
    };
    auto make_windows = [&](int32_t i, const std::vector<AStep>& sd) {
        sched[i].clear();
        for (size_t t = 0; t < sd.size();) {
            const size_t kw = Eng::window_len(sd, t, cf.window);
            BStep<Alg, CP> b{};
            b.w = Eng::make_window(sd, t, kw, o.mincoeff, maxdeg, o.max_xweight);
            sched[i].push_back(b);
            t += kw;
        }
        have_windows[i] = 1;
    };
    auto build = [&](int32_t i) {
        bool good = o.topk <= 0 && (size_t)inst[i].n_terms <= S / 2;
        for (int64_t g = 0; g < inst[i].n_gates && good; ++g) good = inst[i].gates[g].kind == FFX_ROT;
        if (!good) return;
        std::vector<AStep> sd;
        rotation_steps(i, sd);
        if (!try_ml || !plan_mgroups<W, CP, K>(sd, plans[i], cf.group_max)) make_windows(i, sd);
        // work estimate: a circuit's live terms grow with its depth, about
        // rotations / qubits, so its work grows as rotations^2 / qubits (on the
        // 512 kicked-Ising instances within about +-25% of the counted updates)
        uint64_t used[W] = {};
        for (int64_t g = 0; g < inst[i].n_gates; ++g)
            for (int w = 0; w < W; ++w) used[w] |= inst[i].gates[g].x[w] | inst[i].gates[g].z[w];
        int nq = 0;
        for (int w = 0; w < W; ++w) nq += __builtin_popcountll(used[w]);
        for (int64_t t = 0; t < inst[i].n_terms; ++t)
            for (int w = 0; w < W; ++w) used[w] |= inst[i].x[t * W + w] | inst[i].z[t * W + w];
        for (int w = 0; w < W; ++w) used_of[i][w] = used[w];
        est[i] = (double)sd.size() * (double)sd.size() / (double)std::max(1, nq);
        ok[i] = 1;
    };
    // the instances are independent: their schedules are built on host threads
    auto parallel = [&](const std::vector<int32_t>& which, auto&& fn) {
        const int nth = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        std::atomic<size_t> next{0};
        std::vector<std::thread> pool;
        for (int t = 0; t < std::min<int>(nth, (int)which.size()); ++t)
            pool.emplace_back([&]() {
                for (size_t q = next++; q < which.size(); q = next++) fn(which[q]);
            });
        for (auto& th : pool) th.join();
    };
    {
        std::vector<int32_t> all((size_t)n_inst);
        for (int32_t i = 0; i < n_inst; ++i) all[(size_t)i] = i;
        // longest circuits first: the threads take instances in this order, so the
        // big ones do not start last and leave the others idle
        std::stable_sort(all.begin(), all.end(),
                         [&](int32_t a, int32_t b) { return inst[a].n_gates > inst[b].n_gates; });
        parallel(all, build);
    }
    // largest estimated work first: workgroups are dispatched in launch order
    std::vector<int32_t> elig;
    for (int32_t i = 0; i < n_inst; ++i)
        if (ok[i]) elig.push_back(i);
    std::stable_sort(elig.begin(), elig.end(), [&](int32_t a, int32_t b) { return est[a] > est[b]; });
    bool merged_ran = false;
    bool all_grouped = try_ml && !elig.empty();
    for (int32_t i : elig) all_grouped = all_grouped && !have_windows[i];
    if (all_grouped) {
        // cut by schedule length into cf.split term stores, each on its own stream: the
        // short instances' heavy steps then overlap the long ones' latency-bound tails
        // (512-instance kicked-Ising batch: one store 75 ms; cut at 60 % of the longest
        // schedule 68 ms; at 45 and 70 % 69 ms; at 30, 55 and 80 % 65 ms)
        static const std::vector<int> kCuts[4] = {{}, {60}, {45, 70}, {30, 55, 80}};
        const std::vector<int>& cuts = kCuts[cf.split - 1];
        const int np = (int)cuts.size() + 1;
        std::vector<std::vector<int32_t>> part((size_t)np);
        size_t lmax = 0;
        for (int32_t i : elig) lmax = std::max(lmax, plans[i].groups.size());
        for (int32_t i : elig) {
            int h = 0;
            while (h < np - 1 && 100 * plans[i].groups.size() > (size_t)cuts[(size_t)h] * lmax) ++h;
            part[(size_t)h].push_back(i);
        }
        part.erase(std::remove_if(part.begin(), part.end(), [](const std::vector<int32_t>& v) { return v.empty(); }),
                   part.end());
        std::vector<std::array<uint64_t, W>> used(part.size());
        for (size_t h = 0; h < part.size(); ++h) {
            used[h].fill(0);
            for (int32_t i : part[h])
                for (int w = 0; w < W; ++w) used[h][w] |= used_of[i][w];
        }
        // contexts: this one for the first part, then its chain of second contexts
        std::vector<ffx_ctx*> cx(part.size(), ctx);
        for (size_t h = 1; h < part.size(); ++h) {
            if (!cx[h - 1]->aux) {
                const int rc = ffx_ctx_create(ctx->device, &cx[h - 1]->aux);
                if (rc != FFX_OK) return set_err(ctx, rc, "batch: second context for the long schedules");
            }
            cx[h] = cx[h - 1]->aux;
            cx[h]->prof = ffx_profile{};
        }
        std::vector<int> rcs(part.size(), FFX_OK);
        std::vector<char> ran(part.size(), 0);
        auto run_part = [&](size_t h) {
            uint64_t u[W];
            for (int w = 0; w < W; ++w) u[w] = used[h][w];
            bool r = false;
            rcs[h] = run_pauli_mlayer<W, CP, K>(cx[h], inst, part[h], plans, u, o, cf.compact_every, out_expect,
                                                out_terms, r);
            ran[h] = r ? 1 : 0;
        };
        std::vector<std::thread> side;
        for (size_t h = 1; h < part.size(); ++h)
            side.emplace_back([&, h]() {
                if (hipSetDevice(ctx->device) != hipSuccess) {
                    rcs[h] = FFX_EHIP;
                    return;
                }
                run_part(h);
            });
        run_part(0);
        for (auto& th : side) th.join();
        CHK(rcs[0]);
        merged_ran = ran[0] != 0;
        for (size_t h = 1; h < part.size(); ++h) {
            if (rcs[h] != FFX_OK) return set_err(ctx, rcs[h], "batch (long schedules): " + cx[h]->err);
            merged_ran = merged_ran && ran[h];
            ctx->prof.updates += cx[h]->prof.updates;
            ctx->prof.outs += cx[h]->prof.outs;
            ctx->prof.n_gates += cx[h]->prof.n_gates;
            ctx->prof.peak_terms += cx[h]->prof.peak_terms;
            ctx->prof.run_launches = std::max(ctx->prof.run_launches, cx[h]->prof.run_launches);
            ctx->prof.compactions += cx[h]->prof.compactions;
            ctx->prof.launches += cx[h]->prof.launches;
        }
    }
    if (!merged_ran) {
        // windows of the instances that only have layer groups so far
        std::vector<int32_t> need;
        for (int32_t i : elig)
            if (!have_windows[i]) need.push_back(i);
        parallel(need, [&](int32_t i) {
            std::vector<AStep> sd;
            rotation_steps(i, sd);
            make_windows(i, sd);
        });
    }
    if (!merged_ran && merged_any && !elig.empty())
        CHK((run_pauli_merged<W, CP, K>(ctx, inst, elig, sched, out_expect, out_terms, merged_ran)));
    if (merged_ran)
        for (int32_t i : elig) done[i] = 1;
    const size_t ne = merged_ran ? 0 : elig.size();
    int32_t fallback = n_inst - (int32_t)elig.size();
    if (ne > 0) {
        // one pool: per-instance stores, then the states, schedules and inputs
        const size_t hs = 2 * S;
        const size_t store_bytes = (2 * NP * S) * 8 + 2 * (size_t)K * CP::NS * S * sizeof(store_t) + hs * 8;
        size_t n_steps = 0, n_in = 0;
        for (int32_t i : elig) {
            n_steps += sched[i].size();
            n_in += (size_t)inst[i].n_terms;
        }
        auto al = [](size_t v) { return (v + 255) & ~(size_t)255; };
        const size_t off_states = al(ne * store_bytes);
        const size_t off_steps = off_states + al(ne * sizeof(InstState<CP>));
        const size_t off_x = off_steps + al(std::max<size_t>(n_steps, 1) * sizeof(BStep<Alg, CP>));
        const size_t off_z = off_x + al(std::max<size_t>(n_in, 1) * W * 8);
        const size_t off_re = off_z + al(std::max<size_t>(n_in, 1) * W * 8);
        const size_t off_im = off_re + al(std::max<size_t>(n_in, 1) * 8);
        const size_t off_off = off_im + al(std::max<size_t>(n_in, 1) * 8);
        const size_t total = off_off + al((ne + 1) * 8);
        if (total > ctx->batch_pool_bytes) {
            dev_free(ctx->batch_pool);
            ctx->batch_pool_bytes = 0;
            CHK(dev_alloc(ctx, &ctx->batch_pool, total));
            ctx->batch_pool_bytes = total;
        }
        char* pool = (char*)ctx->batch_pool;
        // host image of pool[off_states, total) in pinned memory: states, schedules, inputs, offsets
        const size_t img = total - off_states;
        if (img > ctx->batch_host_bytes) {
            if (ctx->batch_exec) {  // the captured graph copies from the old buffer
                (void)hipGraphExecDestroy(ctx->batch_exec);
                ctx->batch_exec = nullptr;
            }
            if (ctx->batch_host) (void)hipHostFree(ctx->batch_host);
            ctx->batch_host = nullptr;
            ctx->batch_host_bytes = 0;
            HIPCHK(hipHostMalloc(&ctx->batch_host, img, hipHostMallocDefault));
            ctx->batch_host_bytes = img;
        }
        char* H = (char*)ctx->batch_host;
        InstState<CP>* hst = (InstState<CP>*)H;
        BStep<Alg, CP>* hsteps = (BStep<Alg, CP>*)(H + (off_steps - off_states));
        uint64_t* hx = (uint64_t*)(H + (off_x - off_states));
        uint64_t* hz = (uint64_t*)(H + (off_z - off_states));
        double* hre = (double*)(H + (off_re - off_states));
        double* him = (double*)(H + (off_im - off_states));
        int64_t* hoff = (int64_t*)(H + (off_off - off_states));
        size_t ns_done = 0;
        hoff[0] = 0;
        for (size_t b = 0; b < ne; ++b) {
            const ffx_instance& in = inst[elig[b]];
            char* base = pool + b * store_bytes;
            InstState<CP> s = InstState<CP>{};
            s.key = (uint64_t*)base;
            s.key2 = s.key + NP * S;
            s.coef = (store_t*)(s.key2 + NP * S);
            s.coef2 = s.coef + (size_t)K * CP::NS * S;
            s.index = (uint64_t*)(s.coef2 + (size_t)K * CP::NS * S);
            s.hmask = hs - 1;
            s.cap = (uint32_t)S;
            s.n = (uint32_t)in.n_terms;
            s.steps = (const void*)(pool + off_steps + ns_done * sizeof(BStep<Alg, CP>));
            s.n_steps = (uint32_t)sched[elig[b]].size();
            for (int w = 0; w < FFX_MAX_WORDS_DEV; ++w) s.state[w] = in.basis_state[w];
            hst[b] = s;
            std::memcpy(hsteps + ns_done, sched[elig[b]].data(), sched[elig[b]].size() * sizeof(BStep<Alg, CP>));
            ns_done += sched[elig[b]].size();
            const int64_t o0 = hoff[b];
            for (int64_t t = 0; t < in.n_terms; ++t) {
                for (int w = 0; w < W; ++w) {
                    hx[(size_t)(o0 + t) * W + w] = in.x[t * W + w];
                    hz[(size_t)(o0 + t) * W + w] = in.z[t * W + w];
                }
                hre[o0 + t] = in.c_re[t];
                him[o0 + t] = in.c_im ? in.c_im[t] : 0.0;
            }
            hoff[b + 1] = o0 + in.n_terms;
        }
        InstState<CP>* dst = (InstState<CP>*)(pool + off_states);
        const unsigned bb = (unsigned)cf.block;
        // one captured graph per pool layout: upload, unpack, run, download of the states.
        // Calls with the same layout (angle sweeps, repeated batches) replay it.
        const size_t key[8] = {total, off_states, ne, (size_t)bb, (size_t)W, (size_t)prec_id, (size_t)K,
                               (size_t)(uintptr_t)pool};
        auto enqueue = [&]() -> hipError_t {
            hipError_t r = hipMemcpyAsync(pool + off_states, H, img, hipMemcpyHostToDevice, ctx->stream);
            if (r != hipSuccess) return r;
            hipLaunchKernelGGL((k_batch_unpack<W, CP, K>), dim3((unsigned)ne), dim3(kBlock), 0, ctx->stream, dst,
                               (const uint64_t*)(pool + off_x), (const uint64_t*)(pool + off_z),
                               (const double*)(pool + off_re), (const double*)(pool + off_im),
                               (const int64_t*)(pool + off_off));
            hipLaunchKernelGGL((k_batch_run<Alg, CP, K>), dim3((unsigned)ne), dim3(bb), 0, ctx->stream, dst);
            r = hipGetLastError();
            if (r != hipSuccess) return r;
            return hipMemcpyAsync(H, dst, ne * sizeof(InstState<CP>), hipMemcpyDeviceToHost, ctx->stream);
        };
        bool have = ctx->batch_exec && std::memcmp(key, ctx->batch_key, sizeof(key)) == 0;
        if (!have && !getenv("FFX_BATCH_NO_GRAPH")) {
            if (ctx->batch_exec) {
                (void)hipGraphExecDestroy(ctx->batch_exec);
                ctx->batch_exec = nullptr;
            }
            hipGraph_t graph = nullptr;
            if (hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal) == hipSuccess) {
                const hipError_t er = enqueue();
                const hipError_t ec = hipStreamEndCapture(ctx->stream, &graph);
                if (er == hipSuccess && ec == hipSuccess && graph &&
                    hipGraphInstantiate(&ctx->batch_exec, graph, nullptr, nullptr, 0) == hipSuccess) {
                    std::memcpy(ctx->batch_key, key, sizeof(key));
                    have = true;
                } else {
                    ctx->batch_exec = nullptr;
                }
                if (graph) (void)hipGraphDestroy(graph);
            }
            (void)hipGetLastError();  // a failed capture falls back to direct launches
        }
        if (ctx->kernel_timing) HIPCHK(hipEventRecord(ctx->ev0, ctx->stream));
        if (have) {
            HIPCHK(hipGraphLaunch(ctx->batch_exec, ctx->stream));
            ctx->batch_graph_launches++;
        } else {
            HIPCHK(enqueue());
        }
        if (ctx->kernel_timing) HIPCHK(hipEventRecord(ctx->ev1, ctx->stream));
        ctx->launches += 2;
        HIPCHK(hipStreamSynchronize(ctx->stream));
        std::vector<InstState<CP>> st(hst, hst + ne);
        if (ctx->kernel_timing) {
            float ms = 0.f;
            HIPCHK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
            ctx->prof.seconds_kernel += 1e-3 * ms;
            ctx->prof.timed_launches += 1;
        }
        ctx->prof.run_launches += 1;
        for (size_t b = 0; b < ne; ++b) {
            const int32_t i = elig[b];
            if (st[b].status != 0) {
                if (st[b].status == 2) return set_err(ctx, FFX_EHIP, "internal error: hash index full");
                continue;
            }
            out_expect[2 * i] = st[b].expect[0];
            out_expect[2 * i + 1] = st[b].expect[1];
            if (out_terms) out_terms[i] = (int64_t)st[b].n - (int64_t)st[b].dead;
            for (const auto& bs : sched[i]) ctx->prof.n_gates += bs.w.k;
            ctx->prof.updates += (int64_t)st[b].updates;
            ctx->prof.outs += (int64_t)st[b].outs;
            ctx->prof.peak_terms = std::max<int64_t>(ctx->prof.peak_terms, (int64_t)st[b].peak);
            done[i] = 1;
        }
    }
    // the rest on the device-wide engine, one instance at a time
    ffx_profile acc = ctx->prof;
    for (int32_t i = 0; i < n_inst; ++i) {
        if (done[i]) continue;
        if (ok[i]) ++fallback;
        const ffx_instance& in = inst[i];
        double ex[2] = {0.0, 0.0};
        ctx->prof = ffx_profile{};
        CHK((run_pauli<W, CP, K>(ctx, in.gates, in.n_gates, in.x, in.z, in.c_re, K == 2 ? in.c_im : nullptr,
                                 in.n_terms, o, nullptr, in.basis_state, ex)));
        out_expect[2 * i] = ex[0];
        out_expect[2 * i + 1] = ex[1];
        if (out_terms) out_terms[i] = (int64_t)ctx->ctrl_host->n - (int64_t)ctx->ctrl_host->n_dead;
        acc.n_gates += ctx->prof.n_gates;
        acc.updates += ctx->prof.updates;
        acc.outs += ctx->prof.outs;
        acc.peak_terms = std::max(acc.peak_terms, ctx->prof.peak_terms);
        acc.compactions += ctx->prof.compactions;
        acc.run_launches += ctx->prof.run_launches;
        acc.seconds_kernel += ctx->prof.seconds_kernel;
        acc.timed_launches += ctx->prof.timed_launches;
        acc.seconds_enqueue += ctx->prof.seconds_enqueue;
    }
    ctx->prof = acc;
    ctx->prof.seconds_total = now_s() - t0;
    ctx->prof.seconds_propagate = ctx->prof.seconds_total;
    ctx->last_kind = 0;  // the batch leaves no single device-resident result
    if (out_fallback) *out_fallback = fallback;
    return FFX_OK;
}

// Precision and coefficient-width dispatch for one key width W (instantiated in
// ffx_lib.hip; the drivers they call are instantiated by ffx_inst.hip).
template <int W>
int dispatch_pauli(ffx_ctx* ctx, const ffx_pauli_gate* gates, int64_t n_gates, const uint64_t* x, const uint64_t* z,
                   const double* re, const double* im, int64_t n_terms, const ffx_options& o, ffx_terms* out,
                   const uint64_t* state, double* out_expect, int prec, int K, const ffx_prune_opts* po) {
#define FFX_RP(CP_, K_) run_pauli<W, CP_, K_>(ctx, gates, n_gates, x, z, re, K_ == 2 ? im : nullptr, n_terms, o, out, \
                                              state, out_expect, po)
    if (prec == FFX_FP64) return K == 1 ? FFX_RP(F64, 1) : FFX_RP(F64, 2);
    if (prec == FFX_FP32) return K == 1 ? FFX_RP(F32, 1) : FFX_RP(F32, 2);
    return K == 1 ? FFX_RP(F32C, 1) : FFX_RP(F32C, 2);
#undef FFX_RP
}

template <int W>
int dispatch_majorana(ffx_ctx* ctx, const ffx_majorana_gate* gates, int64_t n_gates, const uint64_t* m,
                      const double* re, const double* im, int64_t n_terms, int maxdegree, double mincoeff,
                      ffx_terms* out, const uint64_t* occ, double* out_expect, int prec, int K,
                      const ffx_prune_opts* po) {
#define FFX_RM(CP_, K_) run_majorana<W, CP_, K_>(ctx, gates, n_gates, m, re, im, n_terms, maxdegree, mincoeff, out, \
                                                 occ, out_expect, po)
    if (prec == FFX_FP64) return K == 1 ? FFX_RM(F64, 1) : FFX_RM(F64, 2);
    if (prec == FFX_FP32) return K == 1 ? FFX_RM(F32, 1) : FFX_RM(F32, 2);
    return K == 1 ? FFX_RM(F32C, 1) : FFX_RM(F32C, 2);
#undef FFX_RM
}

// Term-contribution logs (FP64): forward propagation with one launch per
// rotation, no compaction, liveness bitmask after every rotation and the
// coefficients at the logged steps; then the transposed rotations from the last
// rotation to the first, extracting feature rows and influence targets at the
// logged steps.
template <class Alg, int K>
int run_contrib_log(ffx_ctx* ctx, std::vector<typename Engine<Alg, F64, K>::AStep>& sd, const uint64_t* hx,
                    const uint64_t* hz, const double* re, const double* im, int64_t n_terms, double mincoeff,
                    int maxdeg, int max_xw, const ffx_prune_opts& po, const uint64_t* state, int log_every,
                    int64_t max_rows, float* feats, float* targets, int64_t* n_rows, double* sums, int max_points,
                    int32_t* n_points) {
    using Eng = Engine<Alg, F64, K>;
    using AStep = typename Eng::AStep;
    for (AStep& a : sd) {
        if (a.kind != 0) return set_err(ctx, FFX_EINVAL, "contribution logs need a circuit of rotations only");
        if (a.topk_after) return set_err(ctx, FFX_EINVAL, "contribution logs do not support top-k truncation");
        a.compact_after = 0;
    }
    const size_t N = sd.size();
    if (log_every < 1) log_every = 1;
    std::vector<size_t> pts;
    for (size_t e = 0; e <= N; e += (size_t)log_every) pts.push_back(e);
    if ((int64_t)pts.size() > (int64_t)max_points) return set_err(ctx, FFX_EINVAL, "max_points is below the number of logged steps");
    if (max_rows < 0) return set_err(ctx, FFX_EINVAL, "max_rows must be >= 0");
    const FeatPlan fp(sd, po.rot_per_layer);
    const PruneGeo geo = prune_geo(po);
    StateWords sw{};
    if (state)
        for (int w = 0; w < FFX_MAX_WORDS; ++w) sw.s[w] = state[w];
    if (ctx->n_cu <= 0) {
        hipDeviceProp_t prop;
        HIPCHK(hipGetDeviceProperties(&prop, ctx->device));
        ctx->n_cu = prop.multiProcessorCount;
    }
    size_t want = std::max<size_t>(2 * (size_t)n_terms + 64, (size_t)1 << 16);
    for (int attempt = 0; attempt < 12; ++attempt) {
        // the history scales with the store's capacity: start from a store sized for
        // this log, not from whatever an earlier propagation on the context left
        if (ctx->cap > want) free_store(ctx);
        Eng eng(ctx);
        CHK(eng.upload(hx, hz, re, im, n_terms));
        CHK(eng.reserve(want));
        const size_t cap = ctx->cap, bw = cap / 64;
        const size_t snap_bytes = pts.size() * (size_t)K * cap * sizeof(double);
        const size_t bits_bytes = (N + 1) * bw * sizeof(uint64_t);
        if (snap_bytes + bits_bytes > ((size_t)64 << 30))
            return set_err(ctx, FFX_ENOMEM, "contribution log history exceeds 64 GiB; raise log_every");
        uint64_t* bits = nullptr;
        double *snap = nullptr, *va = nullptr, *vb = nullptr, *dsums = nullptr;
        float *dfeat = nullptr, *dtarg = nullptr;
        unsigned long long* counter = nullptr;
        auto release = [&]() {
            dev_free(bits);
            dev_free(snap);
            dev_free(va);
            dev_free(vb);
            dev_free(dsums);
            dev_free(dfeat);
            dev_free(dtarg);
            dev_free(counter);
        };
        int rc = dev_alloc(ctx, &bits, bits_bytes);
        if (rc == FFX_OK) rc = dev_alloc(ctx, &snap, snap_bytes);
        if (rc == FFX_OK) rc = dev_alloc(ctx, &va, cap * sizeof(double));
        if (rc == FFX_OK) rc = dev_alloc(ctx, &vb, cap * sizeof(double));
        if (rc == FFX_OK) rc = dev_alloc(ctx, &dsums, pts.size() * 2 * sizeof(double));
        if (rc == FFX_OK) rc = dev_alloc(ctx, &dfeat, std::max<size_t>((size_t)max_rows, 1) * kFeat * sizeof(float));
        if (rc == FFX_OK) rc = dev_alloc(ctx, &dtarg, std::max<size_t>((size_t)max_rows, 1) * sizeof(float));
        if (rc == FFX_OK) rc = dev_alloc(ctx, &counter, sizeof(unsigned long long));
        if (rc != FFX_OK) {
            release();
            return rc;
        }
        auto fail = [&](int code) {
            release();
            return code;
        };
#define FFX_LCHK(expr)                                                                          \
    do {                                                                                        \
        hipError_t _e = (expr);                                                                 \
        if (_e != hipSuccess) {                                                                 \
            ctx->err = std::string("HIP error: ") + hipGetErrorString(_e) + " at " #expr;       \
            return fail(FFX_EHIP);                                                              \
        }                                                                                       \
    } while (0)
        rc = eng.ensure_epochs(N + 1);
        if (rc != FFX_OK) return fail(rc);
        const Epochs ep = eng.epochs();
        FFX_LCHK(hipMemsetAsync(ctx->epochs, 0, Eng::epoch_words(ctx->epochs_cap) * sizeof(uint32_t), ctx->stream));
        const uint32_t init[2] = {eng.n_host, eng.dead_host};
        FFX_LCHK(hipMemcpyAsync(ep.n, &init[0], sizeof(uint32_t), hipMemcpyHostToDevice, ctx->stream));
        FFX_LCHK(hipMemcpyAsync(ep.dead, &init[1], sizeof(uint32_t), hipMemcpyHostToDevice, ctx->stream));
        const unsigned g_bits = grid_for(cap, kBlock, 4096);
        size_t pi = 0;
        auto snapshot = [&](size_t e) -> hipError_t {
            hipLaunchKernelGGL(k_live_bits<F64>, dim3(g_bits), dim3(kBlock), 0, ctx->stream,
                               (const double*)ctx->coef, ep, (uint32_t)e, (uint32_t)cap, bits + e * bw);
            if (pi < pts.size() && pts[pi] == e) {
                hipError_t r = hipMemcpyAsync(snap + pi * K * cap, ctx->coef, (size_t)K * cap * sizeof(double),
                                              hipMemcpyDeviceToDevice, ctx->stream);
                ++pi;
                return r;
            }
            return hipGetLastError();
        };
        FFX_LCHK(snapshot(0));
        for (size_t e = 1; e <= N; ++e) {
            const AStep& a = sd[e - 1];
            RotParams<F64> rp{};
            rp.cs = a.cs;
            rp.sn = a.sn;
            rp.mincoeff = mincoeff;
            rp.maxdeg = maxdeg;
            rp.trunc_thresh = a.trunc_thresh;
            rp.trunc_xw = a.trunc_xw;
            rp.max_xw = max_xw;
            hipLaunchKernelGGL((k_rot_async<Alg, F64, K>), dim3(2 * ctx->n_cu), dim3(kBlockRot), 0, ctx->stream,
                               eng.store(), (const double*)ctx->coef, (double*)ctx->coef2, a.g, rp, ep, (uint32_t)e);
            FFX_LCHK(hipGetLastError());
            std::swap(ctx->coef, ctx->coef2);
            FFX_LCHK(snapshot(e));
        }
        std::vector<uint32_t> hn(N + 1), ha(N + 1);
        uint32_t hf[2] = {0, 0};
        FFX_LCHK(hipMemcpyAsync(hn.data(), ep.n, (N + 1) * sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
        FFX_LCHK(hipMemcpyAsync(ha.data(), ep.app, (N + 1) * sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
        FFX_LCHK(hipMemcpyAsync(hf, ep.flags, 2 * sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
        FFX_LCHK(hipStreamSynchronize(ctx->stream));
        if (hf[0]) {  // the store overflowed: start again with twice the slots
            release();
            want = 2 * cap;
            continue;
        }
        // adjoint sweep
        const unsigned g_v = grid_for(cap, kBlock, 8192);
        FFX_LCHK(hipMemsetAsync(dsums, 0, pts.size() * 2 * sizeof(double), ctx->stream));
        FFX_LCHK(hipMemsetAsync(counter, 0, sizeof(unsigned long long), ctx->stream));
        hipLaunchKernelGGL(k_adj_init<Alg>, dim3(g_v), dim3(kBlock), 0, ctx->stream, ctx->key, cap, hn[N] + ha[N],
                           bits + N * bw, sw, va);
        FFX_LCHK(hipGetLastError());
        double *cur = va, *nxt = vb;
        int64_t p = (int64_t)pts.size() - 1;
        for (size_t e = N + 1; e-- > 0;) {
            if (p >= 0 && pts[(size_t)p] == e) {
                hipLaunchKernelGGL((k_log_extract<Alg, K>), dim3(grid_for(hn[e] + ha[e], kBlock, 2048)), dim3(kBlock),
                                   0, ctx->stream, ctx->key, cap, snap + (size_t)p * K * cap, hn[e] + ha[e],
                                   bits + e * bw, cur, geo, fp.at(e), (uint64_t)max_rows, counter, dfeat, dtarg,
                                   dsums + 2 * p);
                FFX_LCHK(hipGetLastError());
                --p;
            }
            if (e == 0) break;
            const AStep& a = sd[e - 1];
            hipLaunchKernelGGL(k_adjoint<Alg>, dim3(grid_for(hn[e], kBlock, 8192)), dim3(kBlock), 0, ctx->stream,
                               ctx->key, cap, ctx->index, ctx->hslots - 1, a.g, a.cs, a.sn, maxdeg, hn[e],
                               hn[e] + ha[e], bits + (e - 1) * bw, cur, nxt, &ctx->ctrl->index_full);
            FFX_LCHK(hipGetLastError());
            std::swap(cur, nxt);
        }
        unsigned long long cnt = 0;
        FFX_LCHK(hipMemcpyAsync(&cnt, counter, sizeof(cnt), hipMemcpyDeviceToHost, ctx->stream));
        FFX_LCHK(hipMemcpyAsync(sums, dsums, pts.size() * 2 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        FFX_LCHK(hipStreamSynchronize(ctx->stream));
        const size_t rows = std::min<size_t>((size_t)cnt, (size_t)max_rows);
        if (rows) {
            if (feats)
                FFX_LCHK(hipMemcpyAsync(feats, dfeat, rows * kFeat * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
            if (targets)
                FFX_LCHK(hipMemcpyAsync(targets, dtarg, rows * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
            FFX_LCHK(hipStreamSynchronize(ctx->stream));
        }
#undef FFX_LCHK
        release();
        if (n_rows) *n_rows = (int64_t)rows;
        if (n_points) *n_points = (int32_t)pts.size();
        CHK(eng.sync_ctrl());
        ctx->last_kind = 0;
        return FFX_OK;
    }
    return set_err(ctx, FFX_ECAPACITY, "term store kept overflowing");
}

template <int W>
int dispatch_log_pauli(ffx_ctx* ctx, const ffx_pauli_gate* gates, int64_t n_gates, const uint64_t* x, const uint64_t* z,
                       const double* re, const double* im, int64_t n_terms, const ffx_options& o,
                       const ffx_prune_opts& po, const uint64_t* state, int log_every, int64_t max_rows, float* feats,
                       float* targets, int64_t* n_rows, double* sums, int max_points, int32_t* n_points, int K) {
    const std::vector<Step> steps = pauli_schedule<W>(gates, n_gates, o.batched != 0);
    for (const Step& st : steps)
        if (st.kind != 0) return set_err(ctx, FFX_EINVAL, "contribution logs need a circuit of rotations only");
    const int maxdeg = degree_limit(o.maxdegree);
    std::vector<CliffGate> cl;
    if (K == 1) {
        std::vector<typename Engine<PauliAlg<W>, F64, 1>::AStep> sd;
        pauli_async_steps<W, F64, 1>(gates, steps, o, sd, cl);
        return run_contrib_log<PauliAlg<W>, 1>(ctx, sd, x, z, re, nullptr, n_terms, o.mincoeff, maxdeg, o.max_xweight, po,
                                               state, log_every, max_rows, feats, targets, n_rows, sums, max_points,
                                               n_points);
    }
    std::vector<typename Engine<PauliAlg<W>, F64, 2>::AStep> sd;
    pauli_async_steps<W, F64, 2>(gates, steps, o, sd, cl);
    return run_contrib_log<PauliAlg<W>, 2>(ctx, sd, x, z, re, im, n_terms, o.mincoeff, maxdeg, o.max_xweight, po, state,
                                           log_every, max_rows, feats, targets, n_rows, sums, max_points, n_points);
}

template <int W, int K>
void majorana_log_steps(const ffx_majorana_gate* gates, int64_t n_gates, double mincoeff,
                        std::vector<typename Engine<MajoranaAlg<W>, F64, K>::AStep>& sd) {
    for (int64_t i = n_gates - 1; i >= 0; --i) {
        typename Engine<MajoranaAlg<W>, F64, K>::AStep d{};
        d.g = majorana_gate_dev<W>(gates[i]);
        gate_trig(gates[i].theta, d.cs, d.sn);
        d.trunc_thresh = mincoeff > 0 ? 1 : 0;
        sd.push_back(d);
    }
}

template <int W>
int dispatch_log_majorana(ffx_ctx* ctx, const ffx_majorana_gate* gates, int64_t n_gates, const uint64_t* m,
                          const double* re, const double* im, int64_t n_terms, int maxdegree, double mincoeff,
                          const ffx_prune_opts& po, const uint64_t* occ, int log_every, int64_t max_rows,
                          float* feats, float* targets, int64_t* n_rows, double* sums, int max_points,
                          int32_t* n_points, int K) {
    const int maxdeg = degree_limit(maxdegree);
    if (K == 1) {
        std::vector<typename Engine<MajoranaAlg<W>, F64, 1>::AStep> sd;
        majorana_log_steps<W, 1>(gates, n_gates, mincoeff, sd);
        return run_contrib_log<MajoranaAlg<W>, 1>(ctx, sd, m, nullptr, re, im, n_terms, mincoeff, maxdeg, -1, po, occ,
                                                  log_every, max_rows, feats, targets, n_rows, sums, max_points,
                                                  n_points);
    }
    std::vector<typename Engine<MajoranaAlg<W>, F64, 2>::AStep> sd;
    majorana_log_steps<W, 2>(gates, n_gates, mincoeff, sd);
    return run_contrib_log<MajoranaAlg<W>, 2>(ctx, sd, m, nullptr, re, im, n_terms, mincoeff, maxdeg, -1, po, occ,
                                              log_every, max_rows, feats, targets, n_rows, sums, max_points, n_points);
}

template <int W>
int dispatch_batch(ffx_ctx* ctx, const ffx_instance* inst, int32_t n_inst, const ffx_options& o, int64_t slots,
                   double* out_expect, int64_t* out_terms, int32_t* out_fallback, int prec, int K,
                   const BatchConf* conf) {
#define FFX_RB(CP_, K_) run_pauli_batch<W, CP_, K_>(ctx, inst, n_inst, o, slots, out_expect, out_terms, out_fallback, conf)
    if (prec == FFX_FP64) return K == 1 ? FFX_RB(F64, 1) : FFX_RB(F64, 2);
    if (prec == FFX_FP32) return K == 1 ? FFX_RB(F32, 1) : FFX_RB(F32, 2);
    return K == 1 ? FFX_RB(F32C, 1) : FFX_RB(F32C, 2);
#undef FFX_RB
}

template <int W>
int dispatch_expect(ffx_ctx* ctx, const uint64_t* state, double* out, int kind, int prec, int K) {
    if (kind == 1) {
        using A = PauliAlg<W>;
        if (prec == FFX_FP64) return K == 1 ? expect_last_impl<A, F64, 1>(ctx, state, out) : expect_last_impl<A, F64, 2>(ctx, state, out);
        if (prec == FFX_FP32) return K == 1 ? expect_last_impl<A, F32, 1>(ctx, state, out) : expect_last_impl<A, F32, 2>(ctx, state, out);
        return K == 1 ? expect_last_impl<A, F32C, 1>(ctx, state, out) : expect_last_impl<A, F32C, 2>(ctx, state, out);
    }
    using M = MajoranaAlg<W>;
    if (prec == FFX_FP64) return K == 1 ? expect_last_impl<M, F64, 1>(ctx, state, out) : expect_last_impl<M, F64, 2>(ctx, state, out);
    if (prec == FFX_FP32) return K == 1 ? expect_last_impl<M, F32, 1>(ctx, state, out) : expect_last_impl<M, F32, 2>(ctx, state, out);
    return K == 1 ? expect_last_impl<M, F32C, 1>(ctx, state, out) : expect_last_impl<M, F32C, 2>(ctx, state, out);
}

// Explicit instantiations of the drivers: for one key width W_ and one coefficient
// arithmetic CP_ (both coefficient kinds) of each algebra, and the FP64
// contribution logs of one algebra.  EXT = extern declares them.  ffx_inst.hip is
// compiled once per (key width, arithmetic, algebra), so the kernels compile in
// parallel; ffx_lib.hip declares them all extern and instantiates the dispatchers.
#define FFX_PAULI_INST(EXT, W_, CP_)                                                                                \
    EXT template int run_pauli<W_, CP_, 1>(ffx_ctx*, const ffx_pauli_gate*, int64_t, const uint64_t*, const uint64_t*, \
                                           const double*, const double*, int64_t, const ffx_options&, ffx_terms*,     \
                                           const uint64_t*, double*, const ffx_prune_opts*);                          \
    EXT template int run_pauli<W_, CP_, 2>(ffx_ctx*, const ffx_pauli_gate*, int64_t, const uint64_t*, const uint64_t*, \
                                           const double*, const double*, int64_t, const ffx_options&, ffx_terms*,     \
                                           const uint64_t*, double*, const ffx_prune_opts*);                          \
    EXT template int run_pauli_batch<W_, CP_, 1>(ffx_ctx*, const ffx_instance*, int32_t, const ffx_options&, int64_t, \
                                                 double*, int64_t*, int32_t*, const BatchConf*);                      \
    EXT template int run_pauli_batch<W_, CP_, 2>(ffx_ctx*, const ffx_instance*, int32_t, const ffx_options&, int64_t, \
                                                 double*, int64_t*, int32_t*, const BatchConf*);                      \
    EXT template int expect_last_impl<PauliAlg<W_>, CP_, 1>(ffx_ctx*, const uint64_t*, double*);                      \
    EXT template int expect_last_impl<PauliAlg<W_>, CP_, 2>(ffx_ctx*, const uint64_t*, double*)

#define FFX_MAJORANA_INST(EXT, W_, CP_)                                                                             \
    EXT template int run_majorana<W_, CP_, 1>(ffx_ctx*, const ffx_majorana_gate*, int64_t, const uint64_t*,          \
                                              const double*, const double*, int64_t, int, double, ffx_terms*,        \
                                              const uint64_t*, double*, const ffx_prune_opts*);                      \
    EXT template int run_majorana<W_, CP_, 2>(ffx_ctx*, const ffx_majorana_gate*, int64_t, const uint64_t*,          \
                                              const double*, const double*, int64_t, int, double, ffx_terms*,        \
                                              const uint64_t*, double*, const ffx_prune_opts*);                      \
    EXT template int expect_last_impl<MajoranaAlg<W_>, CP_, 1>(ffx_ctx*, const uint64_t*, double*);                   \
    EXT template int expect_last_impl<MajoranaAlg<W_>, CP_, 2>(ffx_ctx*, const uint64_t*, double*)

#define FFX_LOG_INST(EXT, ALG_)                                                                                     \
    EXT template int run_contrib_log<ALG_, 1>(ffx_ctx*, std::vector<typename Engine<ALG_, F64, 1>::AStep>&,          \
                                              const uint64_t*, const uint64_t*, const double*, const double*, int64_t, \
                                              double, int, int, const ffx_prune_opts&, const uint64_t*, int, int64_t,  \
                                              float*, float*, int64_t*, double*, int, int32_t*);                      \
    EXT template int run_contrib_log<ALG_, 2>(ffx_ctx*, std::vector<typename Engine<ALG_, F64, 2>::AStep>&,          \
                                              const uint64_t*, const uint64_t*, const double*, const double*, int64_t, \
                                              double, int, int, const ffx_prune_opts&, const uint64_t*, int, int64_t,  \
                                              float*, float*, int64_t*, double*, int, int32_t*)

// Argument checks of the Pauli entry points (ffx_lib.hip); makes the context's
// device current and sets *K_out = 2 when any imaginary part is nonzero.
int check_pauli_args(ffx_ctx* ctx, const ffx_pauli_gate* gates, int64_t n_gates, int32_t n_words, const uint64_t* x,
                     const uint64_t* z, const double* c_re, const double* c_im, int64_t n_terms,
                     const ffx_options* opts, int* K_out);

#define FFX_EXTERN_WIDTH(W_)                                                                                         \
    FFX_PAULI_INST(extern, W_, F64);                                                                                 \
    FFX_PAULI_INST(extern, W_, F32);                                                                                 \
    FFX_PAULI_INST(extern, W_, F32C);                                                                                \
    FFX_MAJORANA_INST(extern, W_, F64);                                                                              \
    FFX_MAJORANA_INST(extern, W_, F32);                                                                              \
    FFX_MAJORANA_INST(extern, W_, F32C);                                                                             \
    FFX_LOG_INST(extern, PauliAlg<W_>);                                                                              \
    FFX_LOG_INST(extern, MajoranaAlg<W_>)

}  // namespace ffx_host
