// ffx_async.h -- kernels of the asynchronous engine: one launch per step, no
// host synchronisation between steps.
//
// Kernel boundaries are the global barrier (on MI355X a launch gap costs about
// 2.5 us, against 4..70 us for an in-kernel grid barrier over 64..1024
// blocks).  The term count and erased count flow from step to step through
// per-epoch arrays in device memory: a step kernel for epoch e reads
//   n = n[e-1] + app[e-1],   dead = dead[e-1] + dd[e-1],
// records them in n[e], dead[e] (block 0, thread 0; read by later kernels
// only), and accumulates its own appends and erased-count change in app[e] and
// dd[e].  No counter is reset inside the run: the host zeroes app[] and dd[]
// once.  Appends are block-aggregated: one atomicAdd per block and loop
// iteration (a single contended address serialises at ~11 ns per atomic).
//
// A step whose appends would exceed the capacity sets flags[0] and records its
// epoch in flags[1]; every later kernel returns at entry, so the host can roll
// back that step's index entries, grow the store and resume from that step.
#pragma once

#include "ffx_kernels.h"

namespace ffx {

constexpr int kBlockRot = 1024;

// Counters that every block of a launch adds to, and that only the host reads
// back, are striped over kStripes copies one cache line or more apart (copy
// blockIdx % kStripes): atomics on one cache line serialize (about 11 ns each),
// so 1536 blocks adding five counters of one line cost ~90 us, striped ~3 us.
constexpr int kStripes = 32;

struct Epochs {
    uint32_t* n;     // slots in use at the start of the epoch's kernel
    uint32_t* dead;  // erased slots at the start
    uint32_t* app;   // appends of the epoch
    int32_t* dd;     // erased-count change of the epoch
    uint32_t* flags; // [0] overflow raised, [1] epoch that raised it
    int32_t* sd;     // windows: live-count change at each rotation of the epoch's window,
                     // kStripes copies sd_stride words apart (summed by the host)
    uint32_t sd_stride;
};

// Sums the kStripes copies of striped counters (stride words apart): out32[j]
// for j < len, and the 64-bit word at kWin * q of every copy into out64[q] for
// q < ne (the run-survival counts of Clifford-rotation runs).
static __global__ void k_sum_stripes(const int32_t* sd, uint32_t stride, uint32_t len, uint32_t ne,
                                     int32_t* out32, unsigned long long* out64, int win) {
    const uint32_t t0 = blockIdx.x * blockDim.x + threadIdx.x, nt = gridDim.x * blockDim.x;
    for (uint32_t j = t0; j < len; j += nt) {
        int32_t a = 0;
        for (int c = 0; c < kStripes; ++c) a += sd[(size_t)c * stride + j];
        out32[j] = a;
    }
    for (uint32_t q = t0; q < ne; q += nt) {
        unsigned long long a = 0;
        for (int c = 0; c < kStripes; ++c)
            a += *(const unsigned long long*)(sd + (size_t)c * stride + (size_t)win * q);
        out64[q] = a;
    }
}

// This block's copy of the striped per-rotation counters.
__device__ __forceinline__ int32_t* sd_stripe(const Epochs& ep, uint32_t copy) {
    return ep.sd + (size_t)(copy % kStripes) * ep.sd_stride;
}

__device__ __forceinline__ bool run_stopped(const Epochs& ep) { return atomic_load_u32(&ep.flags[0]) != 0; }

__device__ __forceinline__ void epoch_in(const Epochs& ep, uint32_t e, uint32_t& n, uint32_t& dead) {
    n = ep.n[e - 1] + ep.app[e - 1];
    dead = (uint32_t)((int32_t)ep.dead[e - 1] + ep.dd[e - 1]);
}

// Entry of a step kernel: the stop flag and the epoch's input counts are read
// in one round of independent loads.  Returns false if the run has stopped.
__device__ __forceinline__ bool step_enter(const Epochs& ep, uint32_t e, uint32_t& n, uint32_t& dead) {
    const uint32_t stop = atomic_load_u32(&ep.flags[0]);
    const uint32_t n0 = ep.n[e - 1], a0 = ep.app[e - 1], d0 = ep.dead[e - 1];
    const int32_t dd0 = ep.dd[e - 1];
    if (stop) return false;
    n = n0 + a0;
    dead = (uint32_t)((int32_t)d0 + dd0);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        ep.n[e] = n;
        ep.dead[e] = dead;
    }
    return true;
}

// One rotation.  Each block owns a contiguous chunk of slots.
template <class Alg, class CP, int K>
__global__ void __launch_bounds__(kBlockRot) k_rot_async(Store<CP, K> st, const typename CP::store_t* cin,
                                                         typename CP::store_t* cout, typename Alg::Gate g,
                                                         RotParams<CP> rp, Epochs ep, uint32_t e) {
    constexpr int NP = Alg::NP;
    using val_t = typename CP::val_t;
    uint32_t n, dead;
    if (!step_enter(ep, e, n, dead)) return;
    const uint32_t per = (n + gridDim.x - 1) / gridDim.x;
    const uint32_t chunk = (per + blockDim.x - 1) / blockDim.x * blockDim.x;
    const uint32_t begin = blockIdx.x * chunk;
    const uint32_t end = min(n, begin + chunk);
    if (begin >= end) return;  // block-uniform: nothing to do, nothing to report
    __shared__ uint32_t sbase;
    int dd = 0;
    for (uint32_t b0 = begin; b0 < end; b0 += blockDim.x) {
        const uint32_t i = b0 + threadIdx.x;
        uint64_t r[NP];
        val_t vapp[K];
        bool need_app = false;
        if (i < end) rot_slot<Alg, CP, K>(st, st.key, cin, cout, g, rp, n, i, need_app, r, vapp, dd);
        uint32_t tot;
        const uint32_t rank = block_exclusive_scan(need_app ? 1u : 0u, &tot);
        if (tot) {
            if (threadIdx.x == 0) sbase = atomicAdd(&ep.app[e], tot);
            __syncthreads();
            if (need_app) {
                const uint64_t slot = (uint64_t)n + sbase + rank;
                if (slot < st.cap) {
                    put_slot<NP, CP, K>(st, st.key, cout, slot, r, vapp);
                } else {
                    atomicOr(&ep.flags[0], 1u);
                    ep.flags[1] = e;
                }
            }
        }
    }
    dd = block_sum_int(dd);
    if (threadIdx.x == 0 && dd) atomicAdd(&ep.dd[e], dd);
}

// ---- compaction sequence (device-decided: runs when erased > n/4) ----
__device__ __forceinline__ bool compact_wanted(uint32_t n, uint32_t dead) {
    return dead > 0 && (uint64_t)dead * 4 > n;
}

template <class CP, int K>
__global__ void __launch_bounds__(kBlock) k_cmp_count(Store<CP, K> st, Epochs ep, uint32_t e_prev, uint32_t* counts) {
    if (run_stopped(ep)) return;
    const uint32_t n = ep.n[e_prev] + ep.app[e_prev];
    const uint32_t dead = (uint32_t)((int32_t)ep.dead[e_prev] + ep.dd[e_prev]);
    if (!compact_wanted(n, dead)) return;
    const uint32_t tiles = (n + kTile - 1) / kTile;
    for (uint32_t t = blockIdx.x; t < tiles; t += gridDim.x) {
        int c = 0;
#pragma unroll
        for (int it = 0; it < kItems; ++it) {
            const uint32_t i = t * kTile + it * kBlock + threadIdx.x;
            if (i < n && !CP::dead(st.ld(st.coef, 0, i))) ++c;
        }
        c = block_sum_int(c);
        if (threadIdx.x == 0) counts[t] = (uint32_t)c;
    }
}

static __global__ void __launch_bounds__(kBlock) k_cmp_scan(Epochs ep, uint32_t e_prev, uint32_t* counts) {
    if (run_stopped(ep)) return;
    const uint32_t n = ep.n[e_prev] + ep.app[e_prev];
    const uint32_t dead = (uint32_t)((int32_t)ep.dead[e_prev] + ep.dd[e_prev]);
    if (!compact_wanted(n, dead)) return;
    const uint32_t tiles = (n + kTile - 1) / kTile;
    uint32_t carry = 0;
    for (uint32_t b0 = 0; b0 < tiles; b0 += kBlock) {
        const uint32_t i = b0 + threadIdx.x;
        const uint32_t v = i < tiles ? counts[i] : 0u;
        uint32_t tot;
        const uint32_t ex = block_exclusive_scan(v, &tot);
        if (i < tiles) counts[i] = carry + ex;
        carry += tot;
        __syncthreads();
    }
    if (threadIdx.x == 0) counts[tiles] = carry;
}

// Scatter live slots into key2/coef2 and record the epoch's (n, dead).
template <int NP, class CP, int K>
__global__ void __launch_bounds__(kBlock) k_cmp_scatter(Store<CP, K> st, uint64_t* key2,
                                                        typename CP::store_t* coef2, Epochs ep, uint32_t e_prev,
                                                        uint32_t e, const uint32_t* counts) {
    if (run_stopped(ep)) return;
    const uint32_t n = ep.n[e_prev] + ep.app[e_prev];
    const uint32_t dead = (uint32_t)((int32_t)ep.dead[e_prev] + ep.dd[e_prev]);
    const bool want = compact_wanted(n, dead);
    const uint32_t tiles = (n + kTile - 1) / kTile;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        ep.n[e] = want ? counts[tiles] : n;
        ep.dead[e] = want ? 0u : dead;
    }
    if (!want) return;
    const size_t cap = st.cap;
    for (uint32_t t = blockIdx.x; t < tiles; t += gridDim.x) {
        uint32_t run = counts[t];
        for (int it = 0; it < kItems; ++it) {
            const uint32_t i = t * kTile + it * kBlock + threadIdx.x;
            bool live = false;
            if (i < n) live = !CP::dead(st.ld(st.coef, 0, i));
            uint32_t tot;
            const uint32_t pos = run + block_exclusive_scan(live ? 1u : 0u, &tot);
            if (live) {
                key_copy<NP>(key2, pos, st.key, i);
#pragma unroll
                for (int kk = 0; kk < K; ++kk)
                    CP::store(coef2 + (size_t)kk * CP::NS * cap, cap, pos, st.ld(st.coef, kk, i));
            }
            run += tot;
        }
    }
}

// Copy the compacted planes back, clear the index, then (next kernel) rebuild it.
template <int NP, class CP, int K>
__global__ void __launch_bounds__(kBlock) k_cmp_copyback(Store<CP, K> st, const uint64_t* key2,
                                                         const typename CP::store_t* coef2, Epochs ep,
                                                         uint32_t e_prev, uint32_t e) {
    if (run_stopped(ep)) return;
    const uint32_t n0 = ep.n[e_prev] + ep.app[e_prev];
    const uint32_t dead0 = (uint32_t)((int32_t)ep.dead[e_prev] + ep.dd[e_prev]);
    if (!compact_wanted(n0, dead0)) return;
    const uint32_t n = ep.n[e];
    const size_t cap = st.cap;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += stride) {
        key_copy<NP>(st.key, i, key2, i);
#pragma unroll
        for (int kk = 0; kk < K * CP::NS; ++kk) st.coef[kk * cap + i] = coef2[kk * cap + i];
    }
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i <= st.hmask; i += stride) st.index[i] = kEmpty;
}

template <int NP>
__global__ void __launch_bounds__(kBlock) k_index_build_async(const uint64_t* key, size_t cap, uint64_t* index,
                                                              uint64_t hmask, Epochs ep, uint32_t e_prev,
                                                              uint32_t e, int only_if_compacted, uint32_t* ierr) {
    if (run_stopped(ep)) return;
    if (only_if_compacted) {
        const uint32_t n0 = ep.n[e_prev] + ep.app[e_prev];
        const uint32_t dead0 = (uint32_t)((int32_t)ep.dead[e_prev] + ep.dd[e_prev]);
        if (!compact_wanted(n0, dead0)) return;
    }
    const uint32_t n = ep.n[e];
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        uint64_t k[NP];
        key_load<NP>(key, i, k);
        index_insert(index, hmask, key_hash<NP>(k), i, ierr);
    }
}

// Clifford block: remap keys in place (signs into the coefficients), clear the
// index; k_index_build_async rebuilds it afterwards.
template <int W, class CP, int K>
__global__ void __launch_bounds__(kBlock) k_cliff_async(Store<CP, K> st, const CliffGate* gates, int ng, Epochs ep,
                                                        uint32_t e) {
    if (run_stopped(ep)) return;
    uint32_t n, dead;
    epoch_in(ep, e, n, dead);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        ep.n[e] = n;
        ep.dead[e] = dead;
    }
    const size_t cap = st.cap;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        uint64_t kk2[2 * W];
        key_load<2 * W>(st.key, i, kk2);
        uint64_t x[W], z[W];
#pragma unroll
        for (int w = 0; w < W; ++w) {
            x[w] = kk2[w];
            z[w] = kk2[W + w];
        }
        int sgn = 1;
        for (int t = ng - 1; t >= 0; --t) {
            const CliffGate gt = gates[t];
            const int a = gt.q0, b = gt.q1;
            const bool xa = gbit(x, a), za = gbit(z, a);
            switch (gt.kind) {
                case 1:  // H: X <-> Z, Y -> -Y
                    sbit(x, a, za);
                    sbit(z, a, xa);
                    if (xa && za) sgn = -sgn;
                    break;
                case 2:  // S: X -> -Y, Y -> X
                    sbit(z, a, xa ^ za);
                    if (xa && !za) sgn = -sgn;
                    break;
                case 3: {  // CNOT(a -> b)
                    const bool xb = gbit(x, b), zb = gbit(z, b);
                    sbit(x, b, xa ^ xb);
                    sbit(z, a, za ^ zb);
                    if (xa && zb && !(za ^ xb)) sgn = -sgn;
                    break;
                }
                case 4: {  // SWAP
                    const bool xb = gbit(x, b), zb = gbit(z, b);
                    sbit(x, a, xb);
                    sbit(x, b, xa);
                    sbit(z, a, zb);
                    sbit(z, b, za);
                    break;
                }
                case 5: {  // CZ
                    const bool xb = gbit(x, b), zb = gbit(z, b);
                    if (xa && xb && (za ^ zb)) sgn = -sgn;
                    sbit(z, a, za ^ xb);
                    sbit(z, b, zb ^ xa);
                    break;
                }
                default:
                    break;
            }
        }
#pragma unroll
        for (int w = 0; w < W; ++w) {
            kk2[w] = x[w];
            kk2[W + w] = z[w];
        }
        key_store<2 * W>(st.key, i, kk2);
        if (sgn < 0 && !CP::dead(st.ld(st.coef, 0, i))) {
#pragma unroll
            for (int kk = 0; kk < K; ++kk) st.st(st.coef, kk, i, CP::neg(st.ld(st.coef, kk, i)));
        }
    }
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i <= st.hmask; i += stride) st.index[i] = kEmpty;
}

// ---- top-k truncation (src/pauli/truncate.h:21-39), device-side ----
// Keeps the topk largest |c| of the live terms.  A radix select over the bits of
// |c| (a non-negative double orders like its bit pattern) finds V, the magnitude
// of the topk-th largest term, 8 bits per pass from the top.  The reference
// erases the surplus terms at V in its map's iteration order, which is not
// reproducible; here the ties are ranked by a 64-bit hash of the key, with 8
// more passes over the terms at V selecting H, the hash of the last kept tie, so
// the survivors do not depend on the store's slot order or the grid.  Terms at
// (V, H) -- distinct keys with equal hashes, practically never -- are kept in
// slot order.  Then every live term below (V, H) is erased.
struct TopK {
    uint32_t* hist;              // 256 bins of the current digit
    unsigned long long* prefix;  // magnitude bits selected so far
    uint32_t* want;              // rank of the target among the terms matching the prefixes
    unsigned long long* prefix2; // key-hash bits selected so far among the terms at V
};

template <class CP, int K>
__device__ __forceinline__ uint64_t mag_bits(const Store<CP, K>& st, size_t i) {
    typename CP::val_t v[K];
#pragma unroll
    for (int kk = 0; kk < K; ++kk) v[kk] = st.ld(st.coef, kk, i);
    return (uint64_t)__double_as_longlong(magnitude<CP, K>(v));
}

__device__ __forceinline__ void state_after(const Epochs& ep, uint32_t e_prev, uint32_t& n, uint32_t& live) {
    n = ep.n[e_prev] + ep.app[e_prev];
    live = n - (uint32_t)((int32_t)ep.dead[e_prev] + ep.dd[e_prev]);
}

// Histogram of the digit at `shift` of the magnitude (phase 0) or, among the
// terms at V, of the key hash (phase 1), over the terms matching the selected
// higher digits.
template <int NP, class CP, int K>
__global__ void __launch_bounds__(kBlock) k_topk_hist(Store<CP, K> st, Epochs ep, uint32_t e_prev, TopK tk,
                                                      uint32_t topk, int shift, int phase) {
    if (run_stopped(ep)) return;
    uint32_t n, live;
    state_after(ep, e_prev, n, live);
    if (live <= topk) return;
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const uint64_t mask = shift >= 56 ? 0ull : (~0ull << (shift + 8));
    const uint64_t V = (uint64_t)*tk.prefix;
    const uint64_t pre = shift >= 56 ? 0ull : (uint64_t)(phase ? *tk.prefix2 : *tk.prefix);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        if (CP::dead(st.ld(st.coef, 0, i))) continue;
        uint64_t b = mag_bits<CP, K>(st, i);
        if (phase) {
            if (b != V) continue;
            uint64_t k[NP];
            key_load<NP>(st.key, i, k);
            b = key_hash<NP>(k);
        }
        if ((b & mask) == (pre & mask)) atomicAdd(&h[(b >> shift) & 255u], 1u);
    }
    __syncthreads();
    if (h[threadIdx.x]) atomicAdd(&tk.hist[threadIdx.x], h[threadIdx.x]);
}

// One block of 256 threads: choose the digit (largest first), update the prefix
// and the rank, clear the bins.
static __global__ void __launch_bounds__(kBlock) k_topk_pick(Epochs ep, uint32_t e_prev, TopK tk, uint32_t topk,
                                                      int shift, int phase) {
    if (run_stopped(ep)) return;
    uint32_t n, live;
    state_after(ep, e_prev, n, live);
    if (live <= topk) return;
    __shared__ uint32_t h[256];
    h[threadIdx.x] = tk.hist[threadIdx.x];
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t want = (shift >= 56 && !phase) ? topk - 1 : *tk.want;
        unsigned long long* pp = phase ? tk.prefix2 : tk.prefix;
        uint64_t pre = shift >= 56 ? 0ull : (uint64_t)*pp;
        uint32_t cum = 0;
        int digit = 0;
        for (int b = 255; b >= 0; --b) {
            if (cum + h[b] > want) {
                digit = b;
                break;
            }
            cum += h[b];
        }
        *pp = (unsigned long long)(pre | ((uint64_t)digit << shift));
        *tk.want = want - cum;
    }
    tk.hist[threadIdx.x] = 0;
}

// Terms at (V, H) per tile of kTile slots (counts[tile]).
template <int NP, class CP, int K>
__global__ void __launch_bounds__(kBlock) k_topk_ties(Store<CP, K> st, Epochs ep, uint32_t e_prev, TopK tk,
                                                      uint32_t topk, uint32_t* counts) {
    if (run_stopped(ep)) return;
    uint32_t n, live;
    state_after(ep, e_prev, n, live);
    if (live <= topk) return;
    const uint64_t V = (uint64_t)*tk.prefix, H = (uint64_t)*tk.prefix2;
    const uint32_t tiles = (n + kTile - 1) / kTile;
    for (uint32_t t = blockIdx.x; t < tiles; t += gridDim.x) {
        int c = 0;
#pragma unroll
        for (int j = 0; j < kItems; ++j) {
            const uint32_t i = t * kTile + j * kBlock + threadIdx.x;
            if (i < n && !CP::dead(st.ld(st.coef, 0, i)) && mag_bits<CP, K>(st, i) == V) {
                uint64_t k[NP];
                key_load<NP>(st.key, i, k);
                if (key_hash<NP>(k) == H) ++c;
            }
        }
        c = block_sum_int(c);
        if (threadIdx.x == 0) counts[t] = (uint32_t)c;
    }
}

// One block: counts[tile] -> terms at (V, H) in the tiles before it.
static __global__ void __launch_bounds__(kBlock) k_topk_tiescan(Epochs ep, uint32_t e_prev, uint32_t topk,
                                                                 uint32_t* counts) {
    if (run_stopped(ep)) return;
    uint32_t n, live;
    state_after(ep, e_prev, n, live);
    if (live <= topk) return;
    const uint32_t tiles = (n + kTile - 1) / kTile;
    uint32_t carry = 0;
    for (uint32_t b0 = 0; b0 < tiles; b0 += kBlock) {
        const uint32_t i = b0 + threadIdx.x;
        const uint32_t v = i < tiles ? counts[i] : 0u;
        uint32_t tot;
        const uint32_t ex = block_exclusive_scan(v, &tot);
        if (i < tiles) counts[i] = carry + ex;
        carry += tot;
        __syncthreads();
    }
}

// Epoch e: erase the live terms below (V, H) and those at (V, H) past the first
// `want` + 1 in slot order (ranked from the per-tile counts), then (xw) the
// X/Y-weight cap, which the reference applies after top-k (truncate.h:54-56).
template <class Alg, class CP, int K>
__global__ void __launch_bounds__(kBlock) k_topk_apply(Store<CP, K> st, Epochs ep, uint32_t e, TopK tk,
                                                       uint32_t topk, int xw, int max_xw, const uint32_t* counts) {
    constexpr int NP = Alg::NP;
    uint32_t n, dead;
    if (!step_enter(ep, e, n, dead)) return;
    const bool sel = n - dead > topk;
    if (!sel && !xw) return;
    const uint64_t V = sel ? (uint64_t)*tk.prefix : 0ull, H = sel ? (uint64_t)*tk.prefix2 : 0ull;
    const uint32_t keep_ties = sel ? *tk.want + 1u : 0u;
    const uint32_t tiles = (n + kTile - 1) / kTile;
    int dd = 0;
    for (uint32_t t = blockIdx.x; t < tiles; t += gridDim.x) {
        uint32_t rank0 = sel ? counts[t] : 0u;  // terms at (V, H) in the slots before this tile's
#pragma unroll
        for (int j = 0; j < kItems; ++j) {
            const uint32_t i = t * kTile + j * kBlock + threadIdx.x;
            const bool live = i < n && !CP::dead(st.ld(st.coef, 0, i));
            bool kill = false, tie = false, have_key = false;
            uint64_t k[NP];
            if (live && sel) {
                const uint64_t b = mag_bits<CP, K>(st, i);
                if (b < V) {
                    kill = true;
                } else if (b == V) {
                    key_load<NP>(st.key, i, k);
                    have_key = true;
                    const uint64_t hk = key_hash<NP>(k);
                    kill = hk < H;
                    tie = hk == H;
                }
            }
            uint32_t tot;
            const uint32_t r = rank0 + block_exclusive_scan(tie ? 1u : 0u, &tot);
            rank0 += tot;
            if (tie && r >= keep_ties) kill = true;
            if (live && !kill && xw) {
                if (!have_key) key_load<NP>(st.key, i, k);
                kill = Alg::xweight(k) > max_xw;
            }
            if (kill) {
                st.st(st.coef, 0, i, CP::make_dead());
#pragma unroll
                for (int kk = 1; kk < K; ++kk) st.st(st.coef, kk, i, CP::from_double(0.0));
                ++dd;
            }
        }
    }
    dd = block_sum_int(dd);
    if (threadIdx.x == 0 && dd) atomicAdd(&ep.dd[e], dd);
}

// Remove index entries of slots >= n (the appends of a step that overflowed).
static __global__ void k_index_rollback(uint64_t* index, size_t nslots, uint32_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < nslots; i += (size_t)gridDim.x * blockDim.x) {
        const uint64_t e = index[i];
        if (e != kEmpty && (uint32_t)e >= n) index[i] = kEmpty;
    }
}

}  // namespace ffx
