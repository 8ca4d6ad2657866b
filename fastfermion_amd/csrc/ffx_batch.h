// ffx_batch.h -- many small propagations in one launch: one workgroup runs one
// whole circuit instance.
//
// Each instance has its own term store (keys, ping-pong coefficients, index) in
// HBM, sized for a slot budget.  The workgroup walks the instance's schedule of
// rotation windows (the same windows and per-slot orbit replay as the
// device-wide engine, win_tile), with LDS counters for the appends and workgroup
// barriers between windows instead of kernel boundaries, compacts its own store
// at truncation points, and finally reduces <b|O|b> for the instance's product
// state.  Instances whose terms outgrow the slot budget stop with a status, and
// the host reruns them on the device-wide engine, so every result is complete.
// This serves the batched circuits of one call (hundreds of uneven instances
// in a single launch) and small single circuits, whose per-rotation work is far
// below one launch's fixed cost.
#pragma once

#include "ffx_window.h"

namespace ffx {

constexpr int kBlockBatch = 512;

template <class Alg, class CP>
struct BStep {
    Window<Alg, CP> w;
};

template <class CP>
struct InstState {
    uint64_t* key;
    uint64_t* key2;
    typename CP::store_t* coef;
    typename CP::store_t* coef2;
    uint64_t* index;
    uint64_t hmask;
    uint32_t cap;
    uint32_t n;       // in: input terms; out: slots in use
    uint32_t dead;    // out: erased slots among them
    uint32_t status;  // out: 0 done, 1 slot budget exceeded, 2 index error
    const void* steps;  // BStep<Alg, CP>[n_steps]
    uint32_t n_steps;
    uint32_t pad;
    uint64_t state[FFX_MAX_WORDS_DEV];  // product state of the readout
    double expect[2];                   // out: <b|O|b>
    unsigned long long updates, outs, peak;  // out: work counters
    Ctrl ctrl;                                // index-error flag of the probes
};

template <class Alg, class CP, int K>
__global__ void __launch_bounds__(kBlockBatch) k_batch_run(InstState<CP>* insts) {
    constexpr int NP = Alg::NP;
    constexpr int NS = CP::NS;
    using store_t = typename CP::store_t;
    InstState<CP>* I = insts + blockIdx.x;
    __shared__ uint32_t s_app, s_flag, sbase, s_n, s_dead;
    const uint32_t bd = blockDim.x, tid = threadIdx.x;
    const size_t cap = I->cap;
    const uint64_t hmask = I->hmask;
    uint64_t* const index = I->index;
    uint64_t* key = I->key;
    uint64_t* key2 = I->key2;
    store_t* cin = I->coef;
    store_t* cout = I->coef2;
    uint32_t n = I->n, dead = 0;
    unsigned long long upd = 0, outs = 0, peak = 0;
    auto build_index = [&]() {
        for (uint64_t j = tid; j <= hmask; j += bd) index[j] = kEmpty;
        __syncthreads();
        for (uint32_t i = tid; i < n; i += bd) {
            uint64_t k[NP];
            key_load<NP>(key, i, k);
            index_insert(index, hmask, key_hash<NP>(k), i, &I->ctrl.index_full);
        }
        __syncthreads();
    };
    build_index();
    const BStep<Alg, CP>* steps = (const BStep<Alg, CP>*)I->steps;
    for (uint32_t s = 0; s < I->n_steps; ++s) {
        const Window<Alg, CP>& w = steps[s].w;
        if (tid == 0) {
            s_app = 0;
            s_flag = 0;
        }
        __syncthreads();
        Store<CP, K> st;
        st.key = key;
        st.coef = cin;
        st.coef2 = cout;
        st.index = index;
        st.hmask = hmask;
        st.cap = cap;
        st.ctrl = &I->ctrl;
        const AppendSink sink{&s_app, &s_flag, nullptr, 0};
        int dd = 0;
        int sdl[kWin] = {0, 0, 0, 0};
        for (uint32_t b0 = 0; b0 < n; b0 += bd)
            win_tile<Alg, CP, K>(st, cin, cout, w, n, b0, sink, &sbase, dd, sdl, nullptr, false,
                                 SlotLoad<NP, CP, K>{});
        int r[1 + kWin] = {dd, sdl[0], sdl[1], sdl[2], sdl[3]};
        block_sum_ints<1 + kWin>(r);
        if (tid == 0) {
            long long live = (long long)n - (long long)dead;
            for (int j = 0; j < w.k; ++j) {
                upd += (unsigned long long)live;
                live += r[1 + j];
                outs += (unsigned long long)live;
                if ((unsigned long long)live > peak) peak = (unsigned long long)live;
            }
            s_n = n + s_app;
            s_dead = (uint32_t)((int)dead + r[0]);
        }
        __syncthreads();
        if (s_flag) {
            if (tid == 0) I->status = 1;
            return;
        }
        n = s_n;
        dead = s_dead;
        store_t* t = cin;
        cin = cout;
        cout = t;
        // erased slots only appear at truncation points; compacting as soon as a
        // quarter of the slots, or any slot past half the budget, is erased keeps
        // the instance inside its slot budget
        if (dead > 0 && ((uint64_t)dead * 4 > n || (uint64_t)n * 2 > cap)) {
            // live slots of (key, cin) in order into (key2, cout), then swap and re-index
            uint32_t base = 0;
            for (uint32_t b0 = 0; b0 < n; b0 += bd) {
                const uint32_t i = b0 + tid;
                const bool live = i < n && !CP::dead(CP::load(cin, cap, i));
                uint32_t tot;
                const uint32_t pos = base + block_exclusive_scan(live ? 1u : 0u, &tot);
                if (live) {
                    key_copy<NP>(key2, pos, key, i);
#pragma unroll
                    for (int kk = 0; kk < K; ++kk)
                        CP::store(cout + (size_t)kk * NS * cap, cap, pos, CP::load(cin + (size_t)kk * NS * cap, cap, i));
                }
                base += tot;
            }
            uint64_t* tk = key;
            key = key2;
            key2 = tk;
            t = cin;
            cin = cout;
            cout = t;
            n = base;
            dead = 0;
            __syncthreads();
            build_index();
        }
        __syncthreads();
    }
    // <b|O|b>: per-thread partial sums, then a fixed-order tree in LDS
    double acc[2] = {0.0, 0.0};
    for (uint32_t i = tid; i < n; i += bd) {
        const typename CP::val_t c0 = CP::load(cin, cap, i);
        if (CP::dead(c0)) continue;
        uint64_t k[NP];
        key_load<NP>(key, i, k);
        const int f = state_factor<Alg>(k, I->state);
        if (f) {
            acc[0] += f * CP::to_double(c0);
            if constexpr (K == 2) acc[1] += f * CP::to_double(CP::load(cin + (size_t)NS * cap, cap, i));
        }
    }
    __shared__ double red[2][kBlockBatch];
    red[0][tid] = acc[0];
    red[1][tid] = acc[1];
    __syncthreads();
    for (uint32_t o = bd / 2; o > 0; o >>= 1) {
        if (tid < o) {
            red[0][tid] += red[0][tid + o];
            red[1][tid] += red[1][tid + o];
        }
        __syncthreads();
    }
    if (tid == 0) {
        I->expect[0] = red[0][0];
        I->expect[1] = red[1][0];
        I->n = n;
        I->dead = dead;
        I->key = key;
        I->coef = cin;
        I->updates = upd;
        I->outs = outs;
        I->peak = peak;
        I->status = I->ctrl.index_full ? 2u : 0u;
    }
}

// Input terms of every instance (term-major x/z words and coefficients,
// instance b owns terms [off[b], off[b+1])) into the instances' stores.
template <int W, class CP, int K>
__global__ void __launch_bounds__(kBlock) k_batch_unpack(InstState<CP>* insts, const uint64_t* x, const uint64_t* z,
                                                         const double* re, const double* im, const int64_t* off) {
    InstState<CP>* I = insts + blockIdx.x;
    const size_t cap = I->cap;
    const int64_t a = off[blockIdx.x], b = off[blockIdx.x + 1];
    for (int64_t t = a + threadIdx.x; t < b; t += blockDim.x) {
        const size_t j = (size_t)(t - a);
        uint64_t k[2 * W];
#pragma unroll
        for (int w = 0; w < W; ++w) {
            k[w] = x[t * W + w];
            k[W + w] = z[t * W + w];
        }
        key_store<2 * W>(I->key, j, k);
        CP::store(I->coef, cap, j, CP::from_double(re[t]));
        if constexpr (K == 2) CP::store(I->coef + (size_t)CP::NS * cap, cap, j, CP::from_double(im ? im[t] : 0.0));
    }
}

}  // namespace ffx
