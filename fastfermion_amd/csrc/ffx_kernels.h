// ffx_kernels.h -- HIP kernels of the propagation engine (gfx950).
//
// One rotation = one k_rot launch over every slot of the term store:
//   * commuting and erased-commuting slots copy their coefficient to the output
//     plane (ping-pong coefficient buffers make the step race-free: every slot
//     only writes its own output value; lookups read the unchanged input plane);
//   * an anticommuting slot i looks up its partner key P*key_i in the hash index
//     and computes its own new coefficient cos*c_i + s*sin*c_partner (the partner
//     term only if it was live and key_i passes the degree filter);
//   * a live slot whose partner key is absent appends the partner as a new slot
//     (wave-aggregated atomic append, index insert by CAS);
//   * an erased slot whose live partner emits into it is revived in place.
// Truncation (|c| <= mincoeff, X/Y weight) is fused into the same pass.
#pragma once

#include "ffx_device.h"

namespace ffx {

constexpr int kBlock = 256;
constexpr int kMaxWaves = 16;  // helpers below support blocks of up to 1024 threads
constexpr int kItems = 4;
constexpr int kTile = kBlock * kItems;

template <class CP>
struct RotParams {
    typename CP::fac_t cs;    // cos(theta)
    typename CP::fac_t sn;    // sin(theta)
    double mincoeff;          // threshold (<= erases) when trunc_thresh
    int maxdeg;               // partner degree filter
    int trunc_thresh;
    int trunc_xw;
    int max_xw;
};

template <class CP, int K>
struct Store {
    uint64_t* key;
    typename CP::store_t* coef;   // input coefficient planes
    typename CP::store_t* coef2;  // output coefficient planes
    uint64_t* index;
    uint64_t hmask;
    size_t cap;
    Ctrl* ctrl;

    __device__ typename CP::val_t ld(const typename CP::store_t* c, int k, size_t i) const {
        return CP::load(c + (size_t)k * CP::NS * cap, cap, i);
    }
    __device__ void st(typename CP::store_t* c, int k, size_t i, typename CP::val_t v) const {
        CP::store(c + (size_t)k * CP::NS * cap, cap, i, v);
    }
};

template <class CP, int K>
__device__ __forceinline__ double magnitude(const typename CP::val_t (&v)[K]) {
    if constexpr (K == 1) {
        return fabs(CP::to_double(v[0]));
    } else {
        return hypot(CP::to_double(v[0]), CP::to_double(v[1]));
    }
}

__device__ __forceinline__ uint64_t lanemask_lt() {
    const int lane = __lane_id();
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// Block-wide sum of an int, result valid in thread 0.
__device__ __forceinline__ int block_sum_int(int v) {
    __shared__ int red[kMaxWaves];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    const int wid = threadIdx.x >> 6;
    if (__lane_id() == 0) red[wid] = v;
    __syncthreads();
    int t = 0;
    if (threadIdx.x == 0)
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += red[w];
    __syncthreads();
    return t;
}

// Block-wide sums of N ints at once (one barrier pair), results valid in thread 0.
template <int N>
__device__ __forceinline__ void block_sum_ints(int (&v)[N]) {
    __shared__ int red[kMaxWaves][N];
#pragma unroll
    for (int q = 0; q < N; ++q)
        for (int o = 32; o > 0; o >>= 1) v[q] += __shfl_down(v[q], o, 64);
    const int wid = threadIdx.x >> 6;
    if (__lane_id() == 0)
#pragma unroll
        for (int q = 0; q < N; ++q) red[wid][q] = v[q];
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < N; ++q) {
            int t = 0;
            for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += red[w][q];
            v[q] = t;
        }
    }
    __syncthreads();
}

// Block-wide exclusive scan of a uint32; returns the prefix, total in *total.
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* total) {
    __shared__ uint32_t wsum[kMaxWaves];
    const int lane = __lane_id();
    const int wid = threadIdx.x >> 6;
    uint32_t x = v;
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[wid] = x;
    __syncthreads();
    uint32_t off = 0, tot = 0;
    const int nw = blockDim.x >> 6;
    for (int w = 0; w < nw; ++w) {
        if (w < wid) off += wsum[w];
        tot += wsum[w];
    }
    __syncthreads();
    *total = tot;
    return off + x - v;
}

// The last block of a step folds the step's counters into the control block.
__device__ __forceinline__ void fold_step(Ctrl* ctrl, int block_dead_delta, bool is_rotation) {
    if (threadIdx.x == 0) {
        if (block_dead_delta) atomicAdd(&ctrl->dead_delta, block_dead_delta);
        __threadfence();
        const uint32_t d = atomicAdd(&ctrl->done, 1u);
        if (d == gridDim.x - 1) {
            __threadfence();
            const uint32_t n = atomic_load_u32(&ctrl->n);
            const uint32_t app = atomic_load_u32(&ctrl->n_app);
            const int32_t dd = atomic_load_i32(&ctrl->dead_delta);
            const uint32_t dead_in = atomic_load_u32(&ctrl->n_dead);
            const uint32_t n_out = n + app;
            const uint32_t dead_out = (uint32_t)((int32_t)dead_in + dd);
            if (is_rotation) {
                ctrl->updates += (unsigned long long)(n - dead_in);
                const unsigned long long live = (unsigned long long)(n_out - dead_out);
                ctrl->outs += live;
                if (live > ctrl->peak) ctrl->peak = live;
                ctrl->n_rot += 1;
            }
            ctrl->n = n_out;
            ctrl->n_dead = dead_out;
            ctrl->n_app = 0;
            ctrl->dead_delta = 0;
            ctrl->done = 0;
        }
    }
}

// ---------------------------------------------------------------------------
// One rotation: the per-slot work over [0, n), grid-stride from block `blk` of
// `nblk`.  Reads coefficients from cin, writes every slot's new value to cout,
// appends new partners at slots n + atomicAdd(app_ctr).  Returns this thread's
// change of the erased-slot count.
// ---------------------------------------------------------------------------
// Per-slot work of one rotation on slot i < n: writes the slot's new
// coefficient to cout and reports whether a new partner (key r, value vapp) must
// be appended, plus the slot's change of the erased count.
template <class Alg, class CP, int K>
__device__ __forceinline__ void rot_slot(const Store<CP, K>& st, const uint64_t* key,
                                         const typename CP::store_t* __restrict__ cin,
                                         typename CP::store_t* __restrict__ cout, const typename Alg::Gate& g,
                                         const RotParams<CP>& rp, uint32_t n, uint32_t i, bool& need_app,
                                         uint64_t (&r)[Alg::NP], typename CP::val_t (&vapp)[K], int& dead_delta) {
    constexpr int NP = Alg::NP;
    constexpr int NS = CP::NS;
    using val_t = typename CP::val_t;
    const size_t cap = st.cap;
    uint64_t k[NP];
    key_load<NP>(key, i, k);
    val_t c[K];
#pragma unroll
    for (int kk = 0; kk < K; ++kk) c[kk] = CP::load(cin + (size_t)kk * NS * cap, cap, i);
    const bool alive = !CP::dead(c[0]);
    val_t o[K];
#pragma unroll
    for (int kk = 0; kk < K; ++kk) o[kk] = c[kk];
    if (Alg::anticommutes(k, g)) {
#pragma unroll
        for (int q = 0; q < NP; ++q) r[q] = k[q] ^ g.p[q];
        const uint32_t j = probe<NP>(key, cap, st.index, st.hmask, r, n, &st.ctrl->index_full);
        val_t cj[K];
        bool j_alive = false;
        if (j != kNone) {
#pragma unroll
            for (int kk = 0; kk < K; ++kk) cj[kk] = CP::load(cin + (size_t)kk * NS * cap, cap, j);
            j_alive = !CP::dead(cj[0]);
        }
        const bool recv = j_alive && Alg::degree(k) <= rp.maxdeg;
        const int s_in = recv ? Alg::sign(r, k, g) : 1;
        if (alive) {
#pragma unroll
            for (int kk = 0; kk < K; ++kk) o[kk] = CP::mul(c[kk], rp.cs);
            if (recv) {
#pragma unroll
                for (int kk = 0; kk < K; ++kk) o[kk] = CP::add(o[kk], CP::mul_signed(cj[kk], rp.sn, s_in));
            }
            if (j == kNone && Alg::degree(r) <= rp.maxdeg) {
                const int s_out = Alg::sign(k, r, g);
#pragma unroll
                for (int kk = 0; kk < K; ++kk) vapp[kk] = CP::mul_signed(c[kk], rp.sn, s_out);
                need_app = true;
                if (rp.trunc_thresh && magnitude<CP, K>(vapp) <= rp.mincoeff) need_app = false;
                if (rp.trunc_xw && Alg::xweight(r) > rp.max_xw) need_app = false;
            }
        } else if (recv) {
            // revive an erased slot in place (0 + v in the reference)
#pragma unroll
            for (int kk = 0; kk < K; ++kk) o[kk] = CP::mul_signed(cj[kk], rp.sn, s_in);
        }
    }
    bool dead_out = CP::dead(o[0]);
    if (!dead_out && rp.trunc_thresh && magnitude<CP, K>(o) <= rp.mincoeff) dead_out = true;
    if (!dead_out && rp.trunc_xw && Alg::xweight(k) > rp.max_xw) dead_out = true;
    if (dead_out) {
        o[0] = CP::make_dead();
#pragma unroll
        for (int kk = 1; kk < K; ++kk) o[kk] = CP::from_double(0.0);
    }
#pragma unroll
    for (int kk = 0; kk < K; ++kk) CP::store(cout + (size_t)kk * NS * cap, cap, i, o[kk]);
    dead_delta += (alive && dead_out) ? 1 : 0;
    dead_delta -= (!alive && !dead_out) ? 1 : 0;
}

// Store a new partner at `slot` and index it.
template <int NP, class CP, int K>
__device__ __forceinline__ void put_slot(const Store<CP, K>& st, uint64_t* key, typename CP::store_t* cout,
                                         uint64_t slot, const uint64_t (&r)[NP],
                                         const typename CP::val_t (&vapp)[K]) {
    const size_t cap = st.cap;
    key_store<NP>(key, slot, r);
#pragma unroll
    for (int kk = 0; kk < K; ++kk) CP::store(cout + (size_t)kk * CP::NS * cap, cap, slot, vapp[kk]);
    index_insert(st.index, st.hmask, key_hash<NP>(r), (uint32_t)slot, &st.ctrl->index_full);
}

// Same with the key's index hash supplied by the caller.
template <int NP, class CP, int K>
__device__ __forceinline__ void put_slot_h(const Store<CP, K>& st, uint64_t* key, typename CP::store_t* cout,
                                           uint64_t slot, const uint64_t (&r)[NP],
                                           const typename CP::val_t (&vapp)[K], uint64_t h) {
    const size_t cap = st.cap;
    key_store<NP>(key, slot, r);
#pragma unroll
    for (int kk = 0; kk < K; ++kk) CP::store(cout + (size_t)kk * CP::NS * cap, cap, slot, vapp[kk]);
    index_insert(st.index, st.hmask, h, (uint32_t)slot, &st.ctrl->index_full);
}

// Grid-stride sweep of one rotation with wave-aggregated appends.
template <class Alg, class CP, int K>
__device__ __forceinline__ int rot_sweep(const Store<CP, K>& st, const uint64_t* key,
                                         const typename CP::store_t* __restrict__ cin,
                                         typename CP::store_t* __restrict__ cout, const typename Alg::Gate& g,
                                         const RotParams<CP>& rp, uint32_t n, uint32_t* app_ctr,
                                         uint32_t* overflow, uint32_t blk, uint32_t nblk) {
    constexpr int NP = Alg::NP;
    using val_t = typename CP::val_t;
    int dead_delta = 0;
    const uint32_t stride = nblk * blockDim.x;
    for (uint32_t base = blk * blockDim.x; base < n; base += stride) {
        const uint32_t i = base + threadIdx.x;
        uint64_t r[NP];
        val_t vapp[K];
        bool need_app = false;
        if (i < n) rot_slot<Alg, CP, K>(st, key, cin, cout, g, rp, n, i, need_app, r, vapp, dead_delta);
        const uint64_t m = __ballot(need_app);
        if (m) {
            const int leader = __ffsll((unsigned long long)m) - 1;
            uint32_t first = 0;
            if (__lane_id() == leader) first = atomicAdd(app_ctr, (uint32_t)__popcll(m));
            first = __shfl(first, leader, 64);
            if (need_app) {
                const uint64_t slot = (uint64_t)n + first + __popcll(m & lanemask_lt());
                if (slot < st.cap)
                    put_slot<NP, CP, K>(st, const_cast<uint64_t*>(key), cout, slot, r, vapp);
                else
                    atomicOr(overflow, 1u);
            }
        }
    }
    return dead_delta;
}

// One rotation as its own launch (step-by-step engine).
template <class Alg, class CP, int K>
__global__ void __launch_bounds__(kBlock) k_rot(Store<CP, K> st, typename Alg::Gate g, RotParams<CP> rp) {
    const uint32_t n = st.ctrl->n;
    const int dd = rot_sweep<Alg, CP, K>(st, st.key, st.coef, st.coef2, g, rp, n, &st.ctrl->n_app,
                                         &st.ctrl->overflow, blockIdx.x, gridDim.x);
    fold_step(st.ctrl, block_sum_int(dd), true);
}

// ---------------------------------------------------------------------------
// Clifford block on Pauli keys (src/pauli/gates.h:15-69, 273-301): every slot
// (live or erased) is remapped by the gates [begin, end) from last to first.
// ---------------------------------------------------------------------------
struct CliffGate {
    int32_t kind, q0, q1, pad;
};

__device__ __forceinline__ bool gbit(const uint64_t* v, int q) { return (v[q >> 6] >> (q & 63)) & 1ull; }
__device__ __forceinline__ void sbit(uint64_t* v, int q, bool b) {
    const uint64_t m = 1ull << (q & 63);
    v[q >> 6] = b ? (v[q >> 6] | m) : (v[q >> 6] & ~m);
}

template <int W, class CP, int K>
__global__ void __launch_bounds__(kBlock) k_clifford(Store<CP, K> st, const CliffGate* gates, int ng) {
    const uint32_t n = st.ctrl->n;
    const size_t cap = st.cap;
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
                case 1: {  // H: X<->Z, Y -> -Y
                    sbit(x, a, za);
                    sbit(z, a, xa);
                    if (xa && za) sgn = -sgn;
                    break;
                }
                case 2: {  // S: X -> -Y, Y -> X
                    sbit(z, a, xa ^ za);
                    if (xa && !za) sgn = -sgn;
                    break;
                }
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
        if (sgn < 0) {
            typename CP::val_t c0 = st.ld(st.coef, 0, i);
            if (!CP::dead(c0)) {
#pragma unroll
                for (int kk = 0; kk < K; ++kk) st.st(st.coef, kk, i, CP::neg(st.ld(st.coef, kk, i)));
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Index maintenance
// ---------------------------------------------------------------------------
static __global__ void k_fill_u64(uint64_t* p, size_t n, uint64_t v) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

template <int NP>
__global__ void __launch_bounds__(kBlock) k_index_build(const uint64_t* key, size_t cap, uint64_t* index,
                                                        uint64_t hmask, Ctrl* ctrl, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        uint64_t k[NP];
        key_load<NP>(key, i, k);
        index_insert(index, hmask, key_hash<NP>(k), i, &ctrl->index_full);
    }
}

// Every key must find itself first; a different slot means a duplicate key.
template <int NP>
__global__ void __launch_bounds__(kBlock) k_index_check(const uint64_t* key, size_t cap, const uint64_t* index,
                                                        uint64_t hmask, Ctrl* ctrl) {
    const uint32_t n = ctrl->n;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        uint64_t k[NP];
        key_load<NP>(key, i, k);
        if (probe<NP>(key, cap, index, hmask, k, n, &ctrl->index_full) != i) atomicOr(&ctrl->dup, 1u);
    }
}

// ---------------------------------------------------------------------------
// Compaction of live slots (order preserving)
// ---------------------------------------------------------------------------
template <class CP, int K>
__global__ void __launch_bounds__(kBlock) k_count_live(Store<CP, K> st, uint32_t* counts) {
    const uint32_t n = st.ctrl->n;
    const uint32_t base = blockIdx.x * kTile;
    int c = 0;
#pragma unroll
    for (int it = 0; it < kItems; ++it) {
        const uint32_t i = base + it * kBlock + threadIdx.x;
        if (i < n && !CP::dead(st.ld(st.coef, 0, i))) ++c;
    }
    c = block_sum_int(c);
    if (threadIdx.x == 0) counts[blockIdx.x] = (uint32_t)c;
}

// Single-block exclusive scan of counts[0..nb) in place; total to *total.
static __global__ void __launch_bounds__(kBlock) k_scan_counts(uint32_t* counts, uint32_t nb, uint32_t* total) {
    uint32_t carry = 0;
    for (uint32_t b0 = 0; b0 < nb; b0 += kBlock) {
        const uint32_t i = b0 + threadIdx.x;
        const uint32_t v = i < nb ? counts[i] : 0u;
        uint32_t tot;
        const uint32_t ex = block_exclusive_scan(v, &tot);
        if (i < nb) counts[i] = carry + ex;
        carry += tot;
        __syncthreads();
    }
    if (threadIdx.x == 0) *total = carry;
}

template <int NP, class CP, int K>
__global__ void __launch_bounds__(kBlock) k_compact(Store<CP, K> st, const uint32_t* offsets, uint64_t* key_dst,
                                                    typename CP::store_t* coef_dst) {
    const uint32_t n = st.ctrl->n;
    const uint32_t base = blockIdx.x * kTile;
    uint32_t run = offsets[blockIdx.x];
    const size_t cap = st.cap;
    for (int it = 0; it < kItems; ++it) {
        const uint32_t i = base + it * kBlock + threadIdx.x;
        bool live = false;
        if (i < n) live = !CP::dead(st.ld(st.coef, 0, i));
        uint32_t tot;
        const uint32_t pos = run + block_exclusive_scan(live ? 1u : 0u, &tot);
        if (live) {
            key_copy<NP>(key_dst, pos, st.key, i);
#pragma unroll
            for (int kk = 0; kk < K; ++kk)
                CP::store(coef_dst + (size_t)kk * CP::NS * cap, cap, pos, st.ld(st.coef, kk, i));
        }
        run += tot;
    }
}

// ---------------------------------------------------------------------------
// Upload / download conversion (host x/z word arrays <-> device key records)
// Majorana coefficients are stored normalized: d = c / i^{r}, r = 1 iff the
// string degree is 2 or 3 mod 4, which keeps Hermitian observables real.
// ---------------------------------------------------------------------------
template <class Alg, class CP, int K>
__global__ void __launch_bounds__(kBlock) k_unpack(Store<CP, K> st, const uint64_t* hx, const uint64_t* hz,
                                                   const double* re, const double* im, uint32_t n) {
    constexpr int W = Alg::W;
    const size_t cap = st.cap;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        uint64_t k[Alg::NP];
#pragma unroll
        for (int w = 0; w < W; ++w) {
            k[w] = hx[(size_t)i * W + w];
            if constexpr (Alg::kPauli) k[W + w] = hz[(size_t)i * W + w];
        }
        key_store<Alg::NP>(st.key, i, k);
        double a = re[i], b = im ? im[i] : 0.0;
        if constexpr (!Alg::kPauli) {
            if ((Alg::degree(k) >> 1) & 1) {  // d = c / i = (b, -a)
                const double t = a;
                a = b;
                b = -t;
            }
        }
        st.st(st.coef, 0, i, CP::from_double(a));
        if constexpr (K == 2) st.st(st.coef, 1, i, CP::from_double(b));
    }
}

template <class Alg, class CP, int K>
__global__ void __launch_bounds__(kBlock) k_pack(Store<CP, K> st, const uint32_t* offsets, uint64_t* ox,
                                                 uint64_t* oz, double* ore, double* oim) {
    constexpr int W = Alg::W;
    const uint32_t n = st.ctrl->n;
    const size_t cap = st.cap;
    const uint32_t base = blockIdx.x * kTile;
    uint32_t run = offsets[blockIdx.x];
    for (int it = 0; it < kItems; ++it) {
        const uint32_t i = base + it * kBlock + threadIdx.x;
        typename CP::val_t c0 = CP::from_double(0.0);
        bool live = false;
        if (i < n) {
            c0 = st.ld(st.coef, 0, i);
            live = !CP::dead(c0);
        }
        uint32_t tot;
        const uint32_t pos = run + block_exclusive_scan(live ? 1u : 0u, &tot);
        if (live) {
            uint64_t k[Alg::NP];
            key_load<Alg::NP>(st.key, i, k);
#pragma unroll
            for (int w = 0; w < W; ++w) {
                ox[(size_t)pos * W + w] = k[w];
                if constexpr (Alg::kPauli) oz[(size_t)pos * W + w] = k[W + w];
            }
            double a = CP::to_double(c0);
            double b = 0.0;
            if constexpr (K == 2) b = CP::to_double(st.ld(st.coef, 1, i));
            if constexpr (!Alg::kPauli) {
                if ((Alg::degree(k) >> 1) & 1) {  // c = i*d = (-b, a)
                    const double t = a;
                    a = -b;
                    b = t;
                }
            }
            ore[pos] = a;
            oim[pos] = b;
        }
        run += tot;
    }
}

// ---------------------------------------------------------------------------
// Product-state expectation: deterministic two-stage FP64 reduction
// ---------------------------------------------------------------------------
template <class Alg>
__device__ __forceinline__ int state_factor(const uint64_t (&k)[Alg::NP], const uint64_t* s) {
    constexpr int W = Alg::W;
    if constexpr (Alg::kPauli) {
        // <b|P|b> = (-1)^{|z & b|} if x = 0 (src/qubitproductstate.h:73-78)
        int par = 0;
#pragma unroll
        for (int w = 0; w < W; ++w) {
            if (k[w]) return 0;
            par += popc(k[W + w] & s[w]);
        }
        return (par & 1) ? -1 : 1;
    } else {
        // Fully paired strings {2j, 2j+1 : j in J} only (src/fockstate.h:95-104);
        // value i^{|J| + 2|b & J|}, times i^{r} from the normalization, r = |J| mod 2.
        int nj = 0, nb = 0;
#pragma unroll
        for (int w = 0; w < W; ++w) {
            const uint64_t ev = k[w] & 0x5555555555555555ull;
            const uint64_t od = (k[w] >> 1) & 0x5555555555555555ull;
            if (ev != od) return 0;
            uint64_t x = ev;  // compress even bits to the low half (mode bits)
            x = (x | (x >> 1)) & 0x3333333333333333ull;
            x = (x | (x >> 2)) & 0x0F0F0F0F0F0F0F0Full;
            x = (x | (x >> 4)) & 0x00FF00FF00FF00FFull;
            x = (x | (x >> 8)) & 0x0000FFFF0000FFFFull;
            x = (x | (x >> 16)) & 0x00000000FFFFFFFFull;
            nj += popc(x);
            const uint64_t bw = (s[w >> 1] >> (32 * (w & 1))) & 0xFFFFFFFFull;
            nb += popc(x & bw);
        }
        const int e = (nj & 1) + nj + 2 * nb;
        return ((e >> 1) & 1) ? -1 : 1;
    }
}

struct StateWords {
    uint64_t s[FFX_MAX_WORDS_DEV];
};

template <class Alg, class CP, int K>
__global__ void __launch_bounds__(kBlock) k_expect_partial(Store<CP, K> st, StateWords sw, double* partial) {
    const uint32_t n = st.ctrl->n;
    const size_t cap = st.cap;
    double acc[2] = {0.0, 0.0};
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const typename CP::val_t c0 = st.ld(st.coef, 0, i);
        if (CP::dead(c0)) continue;
        uint64_t k[Alg::NP];
        key_load<Alg::NP>(st.key, i, k);
        const int f = state_factor<Alg>(k, sw.s);
        if (f) {
            acc[0] += f * CP::to_double(c0);
            if constexpr (K == 2) acc[1] += f * CP::to_double(st.ld(st.coef, 1, i));
        }
    }
    __shared__ double red[2][kBlock];
    red[0][threadIdx.x] = acc[0];
    red[1][threadIdx.x] = acc[1];
    __syncthreads();
    for (int o = kBlock / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            red[0][threadIdx.x] += red[0][threadIdx.x + o];
            red[1][threadIdx.x] += red[1][threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = red[0][0];
        partial[2 * blockIdx.x + 1] = red[1][0];
    }
}

static __global__ void k_expect_final(const double* partial, int nb, double* out) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double a = 0.0, b = 0.0;
        for (int i = 0; i < nb; ++i) {
            a += partial[2 * i];
            b += partial[2 * i + 1];
        }
        out[0] = a;
        out[1] = b;
    }
}

}  // namespace ffx
