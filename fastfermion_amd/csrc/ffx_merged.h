// ffx_merged.h -- the circuits of a batch propagated in one term store.
//
// A batch of small circuits run one workgroup per instance (ffx_batch.h) is
// bounded by its heaviest instance: a few instances with 100x the median work
// run alone on their compute units long after the rest are done.  Merged, every
// instance's terms live in one device-wide store and each launch applies to
// every slot the current window of the slot's own instance, so the whole GPU
// works on the heaviest instances and a batch costs about as many launches as
// its longest schedule has windows.
//
// Each term carries its instance number in the top bits of its last Z word,
// above every qubit the instance uses.  No rotation axis has bits there, so
// they change neither a term's anticommutation with an axis (they only pair
// with axis X bits), nor the sign of a product (|x & z| of the term and the
// axis's Z bits against the term's X bits), nor the X/Y weight; terms of
// different instances never share a key, and the key hash separates them.  The
// degree (popcount of x | z) counts them: every window of instance i carries
// maxdegree + popcount(i).  The windows are the device-wide engine's
// (ffx_window.h), replayed slot by slot as k_win_async does, read from the
// instance's schedule in HBM instead of the kernel arguments.  At the step where
// an instance's schedule ends, its live terms are read out into its sums and
// erased, so the compactions drop finished instances and later launches only
// carry the instances still running.
#pragma once

#include "ffx_window.h"

namespace ffx {

template <class Alg, class CP>
struct Merged {
    const Window<Alg, CP>* wins;  // every instance's windows, instance by instance
    const uint32_t* wbase;        // first window of instance i
    const uint32_t* wlen;         // windows of instance i
    uint32_t step;                // window step of this launch
    uint32_t idshift;             // instance number: bits [idshift, idshift + idbits) of the last key word
    uint64_t idmask;              // (1 << idbits) - 1
    unsigned long long* work;     // per stripe (kStripes copies, 16 words apart): [0] term-gate
                                  // updates, [1] live terms leaving the rotations
    const uint64_t* states;       // readout product state of instance i (W words)
    double* out;                  // <b_i|O_i|b_i> (2 per instance)
    unsigned int* count;          // live terms of instance i at its end
};

// Readout of one live term of instance `inst` into the instance's sums.
template <class Alg, class CP, int K>
__device__ __forceinline__ void merged_readout(const uint64_t (&k)[Alg::NP], const typename CP::val_t (&v)[K],
                                               uint32_t inst, const uint64_t* states, double* out,
                                               unsigned int* count) {
    atomicAdd(&count[inst], 1u);
    const int f = state_factor<Alg>(k, states + (size_t)Alg::W * inst);
    if (f) {
        atomicAdd(&out[2 * inst], f * CP::to_double(v[0]));
        if constexpr (K == 2) atomicAdd(&out[2 * inst + 1], f * CP::to_double(v[1]));
    }
}

template <class Alg>
__device__ __forceinline__ uint32_t merged_instance(const uint64_t (&k)[Alg::NP], uint32_t shift, uint64_t mask) {
    return (uint32_t)((k[Alg::NP - 1] >> shift) & mask);
}

// One tile of blockDim slots: as win_tile, with each slot's window taken from its
// instance's schedule, and the work (live terms entering and leaving each
// rotation of the orbit) counted by the orbit's leader.  Block-uniform.
template <class Alg, class CP, int K>
__device__ __forceinline__ void mwin_tile(const Store<CP, K>& st, const typename CP::store_t* cin,
                                          typename CP::store_t* cout, const Merged<Alg, CP>& M, uint32_t n,
                                          uint32_t b0, const AppendSink& sink, uint32_t* sbase, int& dd,
                                          unsigned long long& upd, unsigned long long& outs,
                                          AppendStage<Alg::NP, CP, K>* stg) {
    constexpr int NP = Alg::NP;
    using val_t = typename CP::val_t;
    using Win = Window<Alg, CP>;
    const size_t cap = st.cap;
    const uint32_t i = b0 + threadIdx.x;
    uint32_t fresh = 0;
    uint64_t k[NP];
    uint64_t lk = 0;
    val_t v[kOrbit][K];
    uint32_t inst = 0xFFFFFFFFu;
    if (i < n) {
        key_load<NP>(st.key, i, k);
#pragma unroll
        for (int kk = 0; kk < K; ++kk) v[0][kk] = CP::load(cin + (size_t)kk * CP::NS * cap, cap, i);
        inst = merged_instance<Alg>(k, M.idshift, M.idmask);
    }
    const Win* wp = M.wins;  // a valid window for the block-uniform append call
    if (i < n) {
        if (M.step >= M.wlen[inst]) {
            // the instance's schedule has ended (at this step: later its slots are all
            // erased): a live term goes into the instance's readout and is erased, so
            // that the next compaction drops it and later launches skip it
            if (!CP::dead(v[0][0])) {
                merged_readout<Alg, CP, K>(k, v[0], inst, M.states, M.out, M.count);
                ++dd;
            }
            CP::store(cout, cap, i, CP::make_dead());
#pragma unroll
            for (int kk = 1; kk < K; ++kk) CP::store(cout + (size_t)kk * CP::NS * cap, cap, i, CP::from_double(0.0));
        } else {
            wp = M.wins + M.wbase[inst] + M.step;
            const Win& w = *wp;
            const uint32_t amask = window_amask<Alg>(k, w.g, w.k);
            const uint32_t valid = subsets_of<Alg>(amask);
            uint32_t alive = CP::dead(v[0][0]) ? 0u : 1u, exist = 1;
            uint32_t slot[kOrbit];
            OrbitFacts f{0u, 0u, 0u};
            bool leader = true;
            if (amask) {
                lk = lin_hash<NP>(k);
                f = self_facts<Alg, CP>(k, w, amask);
                orbit_gather<Alg, CP, K>(st, cin, w, k, lk, valid, n, slot, v, exist, alive, f);
                uint32_t lead = i;
                sfor<kOrbit>([&](auto TT) {
                    constexpr int T = decltype(TT)::value;
                    if constexpr (T > 0)
                        if ((exist >> T) & 1) lead = min(lead, slot[T]);
                });
                leader = lead == i;
            }
            const uint32_t a0 = alive & 1u;
            int sdl[kWin] = {0, 0, 0, 0};
            const int live0 = __popc(alive);
            orbit_replay<Alg, CP, K>(w, k, amask, valid, f, v, alive, sdl, leader);
            if (leader) {
                // live members entering rotation s: live0 + changes at rotations < s
                const int nk = w.k;
                int live = live0;
#pragma unroll
                for (int s = 0; s < kWin; ++s) {
                    if (s < nk) {
                        upd += (unsigned long long)live;
                        live += sdl[s];
                        outs += (unsigned long long)live;
                    }
                }
            }
            put_value<CP, K>(cout, cap, i, alive & 1u, v[0]);
            dd += (int)a0 - (int)(alive & 1u);
            if (leader) fresh = alive & ~exist;
        }
    }
    orbit_stage<Alg, CP, K>(st, cout, *wp, sink, n, k, lk, v, fresh, *stg, sbase);
}

// One window step of every instance (device-wide, block-strided tiles).
template <class Alg, class CP, int K>
__global__ void __launch_bounds__(kBlockWin) k_mwin_async(Store<CP, K> st, const typename CP::store_t* cin,
                                                          typename CP::store_t* cout, Merged<Alg, CP> M, Epochs ep,
                                                          uint32_t e) {
    uint32_t n, dead;
    if (!step_enter(ep, e, n, dead)) return;
    const uint32_t begin = blockIdx.x * blockDim.x, step = gridDim.x * blockDim.x;
    if (begin >= n) return;  // block-uniform: nothing to do, nothing to report
    __shared__ uint32_t sbase;
    __shared__ AppendStage<Alg::NP, CP, K> stg;
    if (threadIdx.x == 0) stg.used = 0;
    __syncthreads();
    const AppendSink sink{&ep.app[e], &ep.flags[0], &ep.flags[1], e};
    int dd = 0;
    unsigned long long upd = 0, outs = 0;
    for (uint32_t b0 = begin; b0 < n; b0 += step)
        mwin_tile<Alg, CP, K>(st, cin, cout, M, n, b0, sink, &sbase, dd, upd, outs, &stg);
    stage_flush<Alg::NP, CP, K>(st, cout, sink, n, stg, &sbase);
    dd = block_sum_int(dd);
    // work counters: block sums, one atomic each into the block's stripe
    __shared__ unsigned long long swork[2][kBlockWin / 64];
    for (int o = 32; o > 0; o >>= 1) {
        upd += __shfl_xor(upd, o, 64);
        outs += __shfl_xor(outs, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        swork[0][threadIdx.x >> 6] = upd;
        swork[1][threadIdx.x >> 6] = outs;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long a = 0, b = 0;
        for (uint32_t wv = 0; wv < blockDim.x / 64; ++wv) {
            a += swork[0][wv];
            b += swork[1][wv];
        }
        unsigned long long* wk = M.work + (size_t)(blockIdx.x % kStripes) * 16;
        if (a) atomicAdd(&wk[0], a);
        if (b) atomicAdd(&wk[1], b);
        if (dd) atomicAdd(&ep.dd[e], dd);
    }
}

// Readout of the terms still live after the last launch (the instances whose
// schedule is the longest): per-slot atomics into the per-instance sums (their
// order, and so the last bits of the sums, depend on the schedule of the blocks).
template <class Alg, class CP, int K>
__global__ void __launch_bounds__(kBlock) k_mexpect(Store<CP, K> st, uint32_t n, const uint64_t* states,
                                                    uint32_t idshift, uint64_t idmask, double* out,
                                                    unsigned int* count) {
    const size_t cap = st.cap;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        typename CP::val_t v[K];
#pragma unroll
        for (int kk = 0; kk < K; ++kk) v[kk] = CP::load(st.coef + (size_t)kk * CP::NS * cap, cap, i);
        if (CP::dead(v[0])) continue;
        uint64_t k[Alg::NP];
        key_load<Alg::NP>(st.key, i, k);
        merged_readout<Alg, CP, K>(k, v, merged_instance<Alg>(k, idshift, idmask), states, out, count);
    }
}

}  // namespace ffx
