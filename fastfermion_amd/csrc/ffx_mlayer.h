// ffx_mlayer.h -- the merged batch layout (ffx_merged.h) with one layer group
// of every instance per launch.
//
// All instances' terms share one device-wide store, each term tagged with its
// instance number in the top bits of its last Z word (ffx_merged.h explains why
// the tag changes no anticommutation, sign or X/Y weight).  Each instance's
// schedule is cut into groups that the layer kernels can apply in one pass
// (ffx_dlayer.h): up to kMGroupMax consecutive rotations on pairwise disjoint
// supports, or up to kLayerMaxA consecutive rotations that commute pairwise with
// linearly independent axes (a single rotation is such a group).  A launch
// applies group `step` of its own instance to every slot: the slot's group
// record, gate records and unit -> position table are read from HBM (they are
// shared by all slots of an instance and stay in the caches), the orbit of the
// term under the group's active rotations is claimed by its canonical member,
// and one lane group replays the orbit with the one-rotation kernel's
// operations in layer order, exactly as k_dlayer_async does for one circuit.
// A batch then costs as many launches as its longest instance has groups, which
// for layered circuits is a small multiple of its number of layers.
//
// Work counters: every member of an orbit counts the group positions at which
// it is live entering the rotation (term-gate updates) and live after the
// rotation and its truncation (terms leaving), so the batch's counters equal
// the per-rotation sums of a rotation-by-rotation walk.
//
// Requires every live term's degree (without the instance tag) <= kLayerMaxA,
// as the other layer kernels (host check).
#pragma once

#include "ffx_dlayer.h"
#include "ffx_merged.h"

namespace ffx {

constexpr int kMGroupMax = 64;  // rotations per group: positions fit one mask word
constexpr int kMStage = 512;    // appends staged in LDS per replay block
constexpr int kMStripes = 32;   // leader lists per orbit size (one per block index mod 32)

template <int W>
struct MGroup {
    uint32_t gate_off;      // first gate record of the group
    uint32_t pos_off;       // disjoint groups: unit -> position bytes, one per unit of umask in bit order
    uint32_t nr;            // rotations (1..kMGroupMax)
    uint32_t general;       // 1: commuting rotations with independent axes, 0: disjoint supports
    uint64_t tmask, xmask;  // positions followed by the threshold / by the X/Y-weight cap
    uint64_t umask[W];      // union of the axes' units
};

template <class Alg, class CP>
struct MLayer {
    const MGroup<Alg::W>* groups;     // every instance's groups, instance by instance
    const uint32_t* gbase;            // first group of instance i
    const uint32_t* glen;             // groups of instance i
    const DLayerGate<Alg, CP>* gates;  // gate records, group by group
    const uint64_t* duals;            // general groups: NP-word dual vector of each gate record
    const uint8_t* pos;               // unit -> position bytes
    uint64_t* claim;                  // claim table (hmask + 1 entries, stamped)
    unsigned long long* work;         // per stripe (kStripes copies, 16 words apart): updates, outs
    const uint64_t* states;           // readout product state of instance i (W words)
    double* out;                      // <b_i|O_i|b_i> (2 per instance)
    unsigned int* count;              // live terms of instance i at its end
    uint64_t idmask;
    double mincoeff;
    uint32_t idshift;
    uint32_t stamp;  // this launch's claim stamp (24 bits, nonzero)
    uint32_t step;   // group step of this launch
    int maxdeg;      // degree limit without the instance tag
    int max_xw;
    // orbit leaders (slots) listed by the step's claim launch for its replay launch:
    // list (A, s) of orbits with A active axes from the blocks of stripe s (block
    // index mod kMStripes) holds ol_cap entries at ol + ((A - 1) kMStripes + s) ol_cap;
    // ol_cnt[kLayerMaxA kMStripes step + (A - 1) kMStripes + s] counts them (zeroed
    // per run)
    uint32_t* ol;
    uint32_t* ol_cnt;
    uint32_t ol_cap;
};

// Lowest set bit of m in positions [lo, hi), -1 if none (0 <= lo, hi <= 64).
__device__ __forceinline__ int first_in64(uint64_t m, int lo, int hi) {
    if (lo >= hi) return -1;
    uint64_t b = m >> lo;
    if (hi - lo < 64) b &= (1ull << (hi - lo)) - 1;
    return b ? lo + __builtin_ctzll(b) : -1;
}

// Packed positions (8 bits each, ascending) of the set bits of pm (at most 8).
__device__ __forceinline__ uint64_t pack64(uint64_t pm) {
    uint64_t ap = 0;
    int d = 0;
    while (pm) {
        ap |= (uint64_t)__builtin_ctzll(pm) << (8 * d);
        pm &= pm - 1;
        ++d;
    }
    return ap;
}

// Group positions whose rotations anticommute with k (a mask); *na their number.
template <class Alg, class CP>
__device__ __forceinline__ uint64_t mgroup_active(const uint64_t (&k)[Alg::NP], const MGroup<Alg::W>& G,
                                                  const DLayerGate<Alg, CP>* gt, const uint8_t* pos, int& na) {
    uint64_t pm = 0;
    na = 0;
    if (G.general) {
        // no unit of k under the group's axes: no rotation acts on it
        uint64_t touch = 0;
#pragma unroll
        for (int w = 0; w < Alg::W; ++w) touch |= key_units<Alg>(k, w) & G.umask[w];
        if (!touch) return 0;
        for (uint32_t j = 0; j < G.nr; ++j)
            if (Alg::anticommutes(k, gt[j].g)) {
                pm |= 1ull << j;
                ++na;
            }
        return pm;
    }
    uint64_t tested = 0;
    uint32_t rb = 0;
#pragma unroll
    for (int w = 0; w < Alg::W; ++w) {
        const uint64_t um = G.umask[w];
        uint64_t b = key_units<Alg>(k, w) & um;
        while (b) {
            const int t = __builtin_ctzll(b);
            b &= b - 1;
            const int p = pos[rb + (uint32_t)__builtin_popcountll(um & ((1ull << t) - 1))];
            if (!((tested >> p) & 1)) {
                tested |= 1ull << p;
                if (Alg::anticommutes(k, gt[p].g)) {
                    pm |= 1ull << p;
                    ++na;
                }
            }
        }
        rb += (uint32_t)__builtin_popcountll(um);
    }
    return pm;
}

// Canonical member of k's orbit: the key with the active axes' span coordinates
// cleared, read off with the group's dual vectors (general groups, as for one
// circuit's general layers), or with the private pivot bit of every active axis
// cleared (disjoint groups).
template <class Alg, class CP>
__device__ __forceinline__ void mgroup_canon(const uint64_t (&k)[Alg::NP], uint64_t pm, bool general,
                                             const DLayerGate<Alg, CP>* gt, const uint64_t* duals,
                                             uint64_t (&ck)[Alg::NP]) {
    constexpr int NP = Alg::NP;
    if (general) {
        dlayer_canon_dual<Alg, CP>(k, pm, (const uint64_t(*)[NP])duals, gt, ck);
        return;
    }
#pragma unroll
    for (int q = 0; q < NP; ++q) ck[q] = k[q];
    while (pm) {
        const DLayerGate<Alg, CP>& lg = gt[__builtin_ctzll(pm)];
        pm &= pm - 1;
        const int pv = lg.pivot;
        bool set = false;
#pragma unroll
        for (int q = 0; q < NP; ++q)
            if (q == (pv >> 6)) set = ((ck[q] >> (pv & 63)) & 1) != 0;
        if (set)
#pragma unroll
            for (int q = 0; q < NP; ++q) ck[q] ^= lg.g.p[q];
    }
}

// Orbits with A <= 6 active axes, 64 / 2^A per wave: lane group g = lane >> A
// replays orbit g (member T = lane & (2^A - 1)) when `on`; every lane of the wave
// calls it (the butterfly stages shuffle across the wave).
template <class Alg, class CP, int K, class Append>
__device__ __forceinline__ void mlayer_light(const Store<CP, K>& st, const typename CP::store_t* __restrict__ cin,
                                             typename CP::store_t* __restrict__ cout, const MLayer<Alg, CP>& M,
                                             const uint32_t n, const bool on, const int A,
                                             const uint64_t (&ck)[Alg::NP], const uint64_t ap, const int nr,
                                             const uint64_t tmask, const uint64_t xmask,
                                             const DLayerGate<Alg, CP>* gt, int& dd, unsigned long long& upd,
                                             unsigned long long& outs, Append&& append_wave) {
    constexpr int NP = Alg::NP;
    using val_t = typename CP::val_t;
    using Gate = DLayerGate<Alg, CP>;
    const size_t cap = st.cap;
    const int lane = __lane_id();
    const int T = lane & ((1 << A) - 1);
    const int Aon = on ? A : 0;
    uint64_t key[NP];
    dlayer_member<Alg, CP>(ck, ap, Aon, T, gt, key);
    const int mdeg = M.maxdeg + __builtin_popcountll((ck[NP - 1] >> M.idshift) & M.idmask);
    const bool recv_ok = Alg::degree(key) <= mdeg;
    bool exist = false, alive = false;
    uint32_t slot = kNone;
    val_t v[K];
#pragma unroll
    for (int kk = 0; kk < K; ++kk) v[kk] = CP::from_double(0.0);
    if (on) {
        slot = probe<NP>(st.key, cap, st.index, st.hmask, key, n, &st.ctrl->index_full);
        if (slot != kNone) {
            exist = true;
#pragma unroll
            for (int kk = 0; kk < K; ++kk) v[kk] = CP::load(cin + (size_t)kk * CP::NS * cap, cap, slot);
            alive = !CP::dead(v[0]);
        }
    }
    const bool alive0 = alive;
    int from = 0;                  // first position after which the member is live
    unsigned long long mouts = 0;  // positions after which the member is live
    // truncation points in [lo, hi) on this lane's member
    auto trunc1 = [&](int lo, int hi) {
        if (!alive) return;
        const int pt = first_in64(tmask, lo, hi), px = first_in64(xmask, lo, hi);
        const bool bt = pt >= 0 && magnitude<CP, K>(v) <= M.mincoeff;
        const bool bx = px >= 0 && Alg::xweight(key) > M.max_xw;
        int at = -1;
        if (bt && (!bx || pt <= px))
            at = pt;
        else if (bx)
            at = px;
        if (at >= 0) {
            alive = false;
            mouts += (unsigned long long)(at - from);
        }
    };
    trunc1(0, apos(ap, 0, Aon, nr));
    for (int d = 0; d < A; ++d) {
        const int pq = apos(ap, d, Aon, nr);
        const Gate& lg = gt[on ? pq : 0];
        val_t vu[K];
#pragma unroll
        for (int kk = 0; kk < K; ++kk) vu[kk] = shfl_xor_val<CP>(v[kk], 1 << d);
        const bool au = __shfl_xor((int)alive, 1 << d, 64) != 0;
        if (on) {
            const bool recv = au && recv_ok;
            if (alive || recv) {
                uint64_t ku[NP];
#pragma unroll
                for (int qq = 0; qq < NP; ++qq) ku[qq] = key[qq] ^ lg.g.p[qq];
                const int s_in = recv ? Alg::sign(ku, key, lg.g) : 1;
#pragma unroll
                for (int kk = 0; kk < K; ++kk) {
                    if (alive) {
                        val_t o = CP::mul(v[kk], lg.cs);
                        if (recv) o = CP::add(o, CP::mul_signed(vu[kk], lg.sn, s_in));
                        v[kk] = o;
                    } else {
                        v[kk] = CP::mul_signed(vu[kk], lg.sn, s_in);
                    }
                }
                if (!alive) from = pq;
                alive = true;
            }
            trunc1(pq, apos(ap, d + 1, Aon, nr));
        }
    }
    if (on) {
        if (alive) mouts += (unsigned long long)(nr - from);
        outs += mouts;
        upd += (unsigned long long)alive0 + mouts - (unsigned long long)alive;
        if (exist) {
            if (alive) {
#pragma unroll
                for (int kk = 0; kk < K; ++kk) CP::store(cout + (size_t)kk * CP::NS * cap, cap, slot, v[kk]);
            } else {
                CP::store(cout, cap, slot, CP::make_dead());
#pragma unroll
                for (int kk = 1; kk < K; ++kk)
                    CP::store(cout + (size_t)kk * CP::NS * cap, cap, slot, CP::from_double(0.0));
            }
            dd += (int)alive0 - (int)alive;
        }
    }
    append_wave(on && !exist && alive, key, v);
}

// Orbit of a listed leader, rebuilt from its key (a slot of an earlier launch):
// canonical key, packed active positions, their number, the group.
template <class Alg, class CP>
__device__ __forceinline__ void mlayer_rebuild(const uint64_t* keys, const MLayer<Alg, CP>& M, uint32_t leader,
                                               uint64_t (&ck)[Alg::NP], uint64_t& ap, int& a, MGroup<Alg::W>& G,
                                               const DLayerGate<Alg, CP>*& gt) {
    uint64_t kl[Alg::NP];
    key_load<Alg::NP>(keys, leader, kl);
    const uint32_t inst = merged_instance<Alg>(kl, M.idshift, M.idmask);
    G = M.groups[M.gbase[inst] + M.step];
    gt = M.gates + G.gate_off;
    const uint64_t pm = mgroup_active<Alg, CP>(kl, G, gt, M.pos + G.pos_off, a);
    mgroup_canon<Alg, CP>(kl, pm, G.general != 0, gt, M.duals + (size_t)G.gate_off * Alg::NP, ck);
    ap = pack64(pm);
}

// One orbit with 7 or 8 active axes on the calling wave (members T = lane + 64 m,
// m < 4): member lookups, the active rotations replayed in group order as
// butterfly stages (bits 0-5 by lane shuffles, 6-7 inside the lane), the group's
// truncation points between them, write-back, fresh live members to `append`
// (called by every lane of the wave), work counters.
template <class Alg, class CP, int K, class Append>
__device__ __forceinline__ void mlayer_heavy(const Store<CP, K>& st, const typename CP::store_t* __restrict__ cin,
                                             typename CP::store_t* __restrict__ cout, const MLayer<Alg, CP>& M,
                                             const uint32_t n, const uint64_t (&ck)[Alg::NP], const uint64_t ap,
                                             const int a, const MGroup<Alg::W>& G,
                                             const DLayerGate<Alg, CP>* gt, int& dd, unsigned long long& upd,
                                             unsigned long long& outs, Append&& append_wave) {
    constexpr int NP = Alg::NP;
    using val_t = typename CP::val_t;
    using Gate = DLayerGate<Alg, CP>;
    const size_t cap = st.cap;
    const int lane = __lane_id();
    auto store_dead = [&](uint32_t slot) {
        CP::store(cout, cap, slot, CP::make_dead());
#pragma unroll
        for (int kk = 1; kk < K; ++kk) CP::store(cout + (size_t)kk * CP::NS * cap, cap, slot, CP::from_double(0.0));
    };
    const int nr = (int)G.nr;
    const uint64_t tmask = G.tmask, xmask = G.xmask;
    const int mdeg = M.maxdeg + __builtin_popcountll((ck[NP - 1] >> M.idshift) & M.idmask);
    const int nm = 1 << a;
    // members T = lane + 64 m
    uint64_t key[4][NP];
    val_t v[4][K];
    uint32_t slot[4];
    bool alive[4], exist[4], recv_ok[4];
    int from[4];
    unsigned long long mouts[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        const int T = lane + 64 * m;
        exist[m] = alive[m] = false;
        slot[m] = kNone;
        from[m] = 0;
        mouts[m] = 0;
#pragma unroll
        for (int kk = 0; kk < K; ++kk) v[m][kk] = CP::from_double(0.0);
        dlayer_member<Alg, CP>(ck, ap, T < nm ? a : 0, T, gt, key[m]);
        recv_ok[m] = Alg::degree(key[m]) <= mdeg;
        if (T < nm) {
            slot[m] = probe<NP>(st.key, cap, st.index, st.hmask, key[m], n, &st.ctrl->index_full);
            if (slot[m] != kNone) {
                exist[m] = true;
#pragma unroll
                for (int kk = 0; kk < K; ++kk)
                    v[m][kk] = CP::load(cin + (size_t)kk * CP::NS * cap, cap, slot[m]);
                alive[m] = !CP::dead(v[m][0]);
            }
        }
    }
    bool alive0[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) alive0[m] = alive[m];
    // truncation points in group positions [lo, hi) applied to the orbit's current values
    auto truncate_range = [&](int lo, int hi) {
        const int pt = first_in64(tmask, lo, hi), px = first_in64(xmask, lo, hi);
        if (pt < 0 && px < 0) return;
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            if (!alive[m]) continue;
            const bool bt = pt >= 0 && magnitude<CP, K>(v[m]) <= M.mincoeff;
            const bool bx = px >= 0 && Alg::xweight(key[m]) > M.max_xw;
            int at = -1;
            if (bt && (!bx || pt <= px))
                at = pt;
            else if (bx)
                at = px;
            if (at >= 0) {
                alive[m] = false;
                mouts[m] += (unsigned long long)(at - from[m]);
            }
        }
    };
    truncate_range(0, apos(ap, 0, a, nr));
    for (int d = 0; d < a; ++d) {
        const int pq = apos(ap, d, a, nr);
        const Gate& lg = gt[pq];
        const int bitd = 1 << d;
        val_t vu[4][K];
        bool au[4];
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            if (bitd < 64) {
#pragma unroll
                for (int kk = 0; kk < K; ++kk) vu[m][kk] = shfl_xor_val<CP>(v[m][kk], bitd);
                au[m] = __shfl_xor((int)alive[m], bitd, 64) != 0;
            } else if (bitd == 64) {
#pragma unroll
                for (int kk = 0; kk < K; ++kk) vu[m][kk] = v[m ^ 1][kk];
                au[m] = alive[m ^ 1];
            } else {
#pragma unroll
                for (int kk = 0; kk < K; ++kk) vu[m][kk] = v[m ^ 2][kk];
                au[m] = alive[m ^ 2];
            }
        }
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            const int T = lane + 64 * m;
            if (T >= nm) continue;
            const bool recv = au[m] && recv_ok[m];
            if (alive[m] || recv) {
                uint64_t ku[NP];
#pragma unroll
                for (int qq = 0; qq < NP; ++qq) ku[qq] = key[m][qq] ^ lg.g.p[qq];
                const int s_in = recv ? Alg::sign(ku, key[m], lg.g) : 1;
#pragma unroll
                for (int kk = 0; kk < K; ++kk) {
                    if (alive[m]) {
                        val_t o = CP::mul(v[m][kk], lg.cs);
                        if (recv) o = CP::add(o, CP::mul_signed(vu[m][kk], lg.sn, s_in));
                        v[m][kk] = o;
                    } else {
                        v[m][kk] = CP::mul_signed(vu[m][kk], lg.sn, s_in);
                    }
                }
                if (!alive[m]) from[m] = pq;
                alive[m] = true;
            }
        }
        truncate_range(pq, apos(ap, d + 1, a, nr));
    }
    // write back existing members, stage fresh live ones, count the work
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        const int T = lane + 64 * m;
        if (T >= nm) continue;
        if (alive[m]) mouts[m] += (unsigned long long)(nr - from[m]);
        outs += mouts[m];
        upd += (unsigned long long)alive0[m] + mouts[m] - (unsigned long long)alive[m];
        if (exist[m]) {
            if (alive[m]) {
#pragma unroll
                for (int kk = 0; kk < K; ++kk)
                    CP::store(cout + (size_t)kk * CP::NS * cap, cap, slot[m], v[m][kk]);
            } else {
                store_dead(slot[m]);
            }
            dd += (int)alive0[m] - (int)alive[m];
        }
    }
#pragma unroll
    for (int m = 0; m < 4; ++m) append_wave(lane + 64 * m < nm && !exist[m] && alive[m], key[m], v[m]);
    }

// Claim launch of a group step: every slot reads out (finished instance), copies
// (no active rotation; the group's truncation points applied) or claims its orbit's
// canonical key; the first claimer lists itself as the orbit's leader, by orbit
// size, for the replay launch.
template <class Alg, class CP, int K>
__global__ void __launch_bounds__(kBlock) k_mlayer_claim(Store<CP, K> st, const typename CP::store_t* __restrict__ cin,
                                                         typename CP::store_t* __restrict__ cout, MLayer<Alg, CP> M,
                                                         Epochs ep, uint32_t e) {
    constexpr int NP = Alg::NP;
    using val_t = typename CP::val_t;
    using Gate = DLayerGate<Alg, CP>;
    __shared__ uint32_t slead[kLayerMaxA][kBlock];
    __shared__ uint32_t scount[kLayerMaxA], sbase[kLayerMaxA];
    __shared__ unsigned long long swork[2][kBlock / 64];
    uint32_t n, dead;
    if (!step_enter(ep, e, n, dead)) return;
    if ((uint64_t)blockIdx.x * blockDim.x >= n) return;  // block-uniform
    const size_t cap = st.cap;
    const int lane = __lane_id();
    const int wave = threadIdx.x >> 6, nwave = blockDim.x >> 6;
    const uint32_t stripe = blockIdx.x % kMStripes;
    uint32_t* const cnt = M.ol_cnt + (size_t)kLayerMaxA * kMStripes * M.step;
    int dd = 0;
    unsigned long long upd = 0, outs = 0;
    // live terms of finished instances into their sums: one atomic per wave and
    // instance (consecutive slots mostly belong to one instance), called by every lane
    auto readout_wave = [&](bool act, const uint64_t (&k)[NP], const val_t (&v)[K], uint32_t inst) {
        uint64_t rem = __ballot(act);
        if (!rem) return;
        double f0 = 0.0, f1 = 0.0;
        if (act) {
            const int f = state_factor<Alg>(k, M.states + (size_t)Alg::W * inst);
            f0 = f * CP::to_double(v[0]);
            if constexpr (K == 2) f1 = f * CP::to_double(v[1]);
        }
        while (rem) {
            const int first = __ffsll((long long)rem) - 1;
            const uint32_t li = (uint32_t)__shfl((int)inst, first, 64);
            const bool in = act && inst == li;
            const uint64_t m = __ballot(in);
            rem &= ~m;
            double s0 = in ? f0 : 0.0, s1 = in ? f1 : 0.0;
            for (int o = 32; o > 0; o >>= 1) {
                s0 += __shfl_xor(s0, o, 64);
                if constexpr (K == 2) s1 += __shfl_xor(s1, o, 64);
            }
            if (lane == first) {
                atomicAdd(&M.count[li], (unsigned int)__popcll(m));
                if (s0 != 0.0) atomicAdd(&M.out[2 * li], s0);
                if (K == 2 && s1 != 0.0) atomicAdd(&M.out[2 * li + 1], s1);
            }
        }
    };
    for (uint32_t b0 = blockIdx.x * blockDim.x; b0 < n; b0 += gridDim.x * blockDim.x) {
        if (threadIdx.x < kLayerMaxA) scount[threadIdx.x] = 0;
        __syncthreads();
        const uint32_t i = b0 + threadIdx.x;
        bool fin_live = false;
        val_t fin_v[K];
        uint64_t k[NP];
        uint32_t inst = 0;
        if (i < n) {
            key_load<NP>(st.key, i, k);
            inst = merged_instance<Alg>(k, M.idshift, M.idmask);
            if (M.step >= M.glen[inst]) {
                // the instance's schedule has ended: a live term goes into its readout and
                // is erased, so that the next compaction drops it
#pragma unroll
                for (int kk = 0; kk < K; ++kk) fin_v[kk] = CP::load(cin + (size_t)kk * CP::NS * cap, cap, i);
                if (!CP::dead(fin_v[0])) {
                    fin_live = true;
                    ++dd;
                }
                CP::store(cout, cap, i, CP::make_dead());
#pragma unroll
                for (int kk = 1; kk < K; ++kk) CP::store(cout + (size_t)kk * CP::NS * cap, cap, i, CP::from_double(0.0));
            } else {
                const uint32_t gi = M.gbase[inst] + M.step;
                const MGroup<Alg::W> G = M.groups[gi];
                const Gate* gt = M.gates + G.gate_off;
                const uint8_t* gpos = M.pos + G.pos_off;
                int a = 0;
                const uint64_t pm = mgroup_active<Alg, CP>(k, G, gt, gpos, a);
                if (a == 0) {
                    // untouched by every rotation of the group; its truncation points still
                    // apply (to an unchanged value: the first that erases decides)
                    val_t c[K];
#pragma unroll
                    for (int kk = 0; kk < K; ++kk) c[kk] = CP::load(cin + (size_t)kk * CP::NS * cap, cap, i);
                    if (!CP::dead(c[0])) {
                        const int nr = (int)G.nr;
                        int at = nr;
                        const int pt = first_in64(G.tmask, 0, nr), px = first_in64(G.xmask, 0, nr);
                        if (pt >= 0 && magnitude<CP, K>(c) <= M.mincoeff) at = pt;
                        if (px >= 0 && px < at && Alg::xweight(k) > M.max_xw) at = px;
                        // live entering positions 0..min(at, nr - 1), live after 0..at-1
                        upd += (unsigned long long)(at < nr ? at + 1 : nr);
                        outs += (unsigned long long)at;
                        if (at < nr) {
                            c[0] = CP::make_dead();
#pragma unroll
                            for (int kk = 1; kk < K; ++kk) c[kk] = CP::from_double(0.0);
                            ++dd;
                        }
                    }
#pragma unroll
                    for (int kk = 0; kk < K; ++kk) CP::store(cout + (size_t)kk * CP::NS * cap, cap, i, c[kk]);
                } else if (a > kLayerMaxA) {
                    // more active axes than a live term can have: its orbit holds no live
                    // member; copy it through (a live one is an internal error)
                    const val_t c0 = CP::load(cin, cap, i);
                    if (!CP::dead(c0)) atomicOr(&st.ctrl->index_full, 4u);
#pragma unroll
                    for (int kk = 0; kk < K; ++kk)
                        CP::store(cout + (size_t)kk * CP::NS * cap, cap, i,
                                  CP::load(cin + (size_t)kk * CP::NS * cap, cap, i));
                } else {
                    uint64_t ck[NP];
                    mgroup_canon<Alg, CP>(k, pm, G.general != 0, gt, M.duals + (size_t)G.gate_off * NP, ck);
                    const uint64_t h = hmix(lin_hash<NP>(ck));
                    const uint64_t tag8 = (h >> 56) & 0xFFull;
                    const uint64_t mine = ((uint64_t)M.stamp << 40) | (tag8 << 32) | (uint64_t)i;
                    uint64_t p = h & st.hmask;
                    bool leader = false;
                    for (uint64_t s = 0;; ++s) {
                        if (s > st.hmask) {
                            atomicOr(&st.ctrl->index_full, 2u);
                            break;
                        }
                        uint64_t e0 = M.claim[p];
                        if ((e0 >> 40) != M.stamp) {
                            const uint64_t old = (uint64_t)atomicCAS((unsigned long long*)&M.claim[p],
                                                                     (unsigned long long)e0, (unsigned long long)mine);
                            if (old == e0) {
                                leader = true;
                                break;
                            }
                            e0 = old;
                            if ((e0 >> 40) != M.stamp) continue;  // overwritten meanwhile: look again
                        }
                        if (((e0 >> 32) & 0xFFull) == tag8) {
                            // the claimant's key kj is in this orbit iff reducing it by this
                            // slot's active axes gives ck (the active set is an orbit
                            // invariant); a claimant of another instance never is (the tag is
                            // part of the key)
                            const uint32_t j = (uint32_t)e0;
                            uint64_t kj[NP];
                            key_load<NP>(st.key, j, kj);
                            if (merged_instance<Alg>(kj, M.idshift, M.idmask) == inst) {
                                uint64_t ckj[NP];
                                mgroup_canon<Alg, CP>(kj, pm, G.general != 0, gt, M.duals + (size_t)G.gate_off * NP, ckj);
                                bool eq = true;
#pragma unroll
                                for (int q = 0; q < NP; ++q) eq = eq && ckj[q] == ck[q];
                                if (eq) break;  // another slot leads this orbit
                            }
                        }
                        p = (p + 1) & st.hmask;
                    }
                    if (leader) slead[a - 1][atomicAdd(&scount[a - 1], 1u)] = i;
                }
            }
        }
        readout_wave(fin_live, k, fin_v, inst);
        __syncthreads();
        // the block's leaders into the lists of its stripe: one atomic per orbit size
        if (threadIdx.x < kLayerMaxA && scount[threadIdx.x])
            sbase[threadIdx.x] = atomicAdd(&cnt[threadIdx.x * kMStripes + stripe], scount[threadIdx.x]);
        __syncthreads();
        for (int A = 0; A < kLayerMaxA; ++A) {
            const uint32_t c = scount[A];
            uint32_t* lst = M.ol + ((size_t)A * kMStripes + stripe) * M.ol_cap;
            for (uint32_t r = threadIdx.x; r < c; r += blockDim.x) {
                const uint32_t at = sbase[A] + r;
                if (at < M.ol_cap)
                    lst[at] = slead[A][r];
                else
                    atomicOr(&st.ctrl->index_full, 16u);
            }
        }
        __syncthreads();
    }
    dd = block_sum_int(dd);
    for (int o = 32; o > 0; o >>= 1) {
        upd += __shfl_xor(upd, o, 64);
        outs += __shfl_xor(outs, o, 64);
    }
    if (lane == 0) {
        swork[0][wave] = upd;
        swork[1][wave] = outs;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long a = 0, b = 0;
        for (int wv = 0; wv < nwave; ++wv) {
            a += swork[0][wv];
            b += swork[1][wv];
        }
        unsigned long long* wk = M.work + (size_t)(blockIdx.x % kStripes) * 16;
        if (a) atomicAdd(&wk[0], a);
        if (b) atomicAdd(&wk[1], b);
        if (dd) atomicAdd(&ep.dd[e], dd);
    }
}

// Replay launch of a group step: the listed orbits spread evenly over the grid's
// waves.  The work is cut into passes -- 64 / 2^A orbits of A <= 6 active axes, or
// one orbit of 7 or 8 -- numbered over all lists (orbit size, then stripe), and
// wave w takes passes w, w + waves, ...  Same epoch as the claim launch: it reads
// the slots the claim launch saw (n), writes every member of its orbits and
// appends after the claim launch (which appends nothing).
template <class Alg, class CP, int K>
__global__ void __launch_bounds__(kBlock, Alg::NP <= 6 ? 3 : 2)
    k_mlayer_replay(Store<CP, K> st, const typename CP::store_t* __restrict__ cin,
                    typename CP::store_t* __restrict__ cout, MLayer<Alg, CP> M, Epochs ep, uint32_t e) {
    constexpr int NP = Alg::NP;
    using val_t = typename CP::val_t;
    constexpr int NL = kLayerMaxA * kMStripes;  // lists
    static_assert(NL == 4 * 64, "a wave holds the list counts four per lane");
    __shared__ uint64_t skey[kMStage][NP];
    __shared__ val_t sval[kMStage][K];
    __shared__ uint32_t s_cnt, s_base;
    __shared__ unsigned long long swork[2][kBlock / 64];
    if (run_stopped(ep)) return;
    const int lane = __lane_id();
    const int wave = threadIdx.x >> 6, nwave = blockDim.x >> 6;
    // passes of lists 4 lane .. 4 lane + 3 and their exclusive prefix over the wave
    const uint32_t* cnt = M.ol_cnt + (size_t)NL * M.step;
    uint32_t c4[4], p4[4], mine = 0;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int l = 4 * lane + u, A = l / kMStripes + 1;
        c4[u] = min(cnt[l], M.ol_cap);
        p4[u] = A <= 6 ? (c4[u] + (64u >> A) - 1) >> (6 - A) : c4[u];
        mine += p4[u];
    }
    uint32_t incl = mine;
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t t = (uint32_t)__shfl_up((int)incl, o, 64);
        if (lane >= o) incl += t;
    }
    const uint32_t total = (uint32_t)__shfl((int)incl, 63, 64);
    const uint32_t excl = incl - mine;
    if ((uint64_t)blockIdx.x * nwave >= total) return;  // block-uniform
    uint32_t n, dead;
    epoch_in(ep, e, n, dead);
    const size_t cap = st.cap;
    if (threadIdx.x == 0) s_cnt = 0;
    __syncthreads();
    int dd = 0;
    unsigned long long upd = 0, outs = 0;
    auto append_wave = [&](bool fresh, const uint64_t (&key)[NP], const val_t (&v)[K]) {
        const uint64_t fm = __ballot(fresh);
        if (!fm) return;
        const uint64_t lt = lanemask_lt();
        const int first = __ffsll((long long)fm) - 1;
        uint32_t base = 0;
        if (lane == first) base = atomicAdd(&s_cnt, (uint32_t)__popcll(fm));
        base = (uint32_t)__shfl((int)base, first, 64);
        const uint32_t r = base + (uint32_t)__popcll(fm & lt);
        if (fresh && r < (uint32_t)kMStage) {
#pragma unroll
            for (int q = 0; q < NP; ++q) skey[r][q] = key[q];
#pragma unroll
            for (int kk = 0; kk < K; ++kk) sval[r][kk] = v[kk];
        }
        const bool over = fresh && r >= (uint32_t)kMStage;
        const uint64_t om = __ballot(over);
        if (!om) return;
        const int ofirst = __ffsll((long long)om) - 1;
        uint32_t gb = 0;
        if (lane == ofirst) gb = atomicAdd(&ep.app[e], (uint32_t)__popcll(om));
        gb = (uint32_t)__shfl((int)gb, ofirst, 64);
        if (over) {
            const uint64_t at = (uint64_t)n + gb + (uint32_t)__popcll(om & lt);
            if (at < cap) {
                put_slot<NP, CP, K>(st, st.key, cout, at, key, v);
            } else {
                atomicOr(&ep.flags[0], 1u);
                ep.flags[1] = e;
            }
        }
    };
    const uint32_t stride = gridDim.x * (uint32_t)nwave;
    for (uint32_t pass = blockIdx.x * (uint32_t)nwave + (uint32_t)wave; pass < total; pass += stride) {
        // the list holding this pass: the lane whose four lists cover it
        int owner = -1, ul = 0;
        uint32_t q0 = 0;
        {
            uint32_t acc = excl;
            bool hit = false;
            int u_hit = 0;
            uint32_t qh = 0;
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                if (!hit && pass >= acc && pass < acc + p4[u]) {
                    hit = true;
                    u_hit = u;
                    qh = pass - acc;
                }
                acc += p4[u];
            }
            const uint64_t hm = __ballot(hit);
            owner = __ffsll((long long)hm) - 1;
            ul = __shfl(4 * lane + u_hit, owner, 64);
            q0 = (uint32_t)__shfl((int)qh, owner, 64);
        }
        const int A = ul / kMStripes + 1;
        const uint32_t c = (uint32_t)__shfl((int)c4[ul % 4], owner, 64);
        const uint32_t* lst = M.ol + (size_t)ul * M.ol_cap;
        if (A <= 6) {
            const uint32_t per = 64u >> A;
            const uint32_t o = q0 * per + ((uint32_t)lane >> A);
            const bool on = o < c;
            uint64_t ck[NP] = {};
            uint64_t ap = 0;
            int a = 0, nr = 0;
            uint64_t tmask = 0, xmask = 0;
            const DLayerGate<Alg, CP>* gt = M.gates;
            if (on) {
                MGroup<Alg::W> G;
                mlayer_rebuild<Alg, CP>(st.key, M, lst[o], ck, ap, a, G, gt);
                nr = (int)G.nr;
                tmask = G.tmask;
                xmask = G.xmask;
            }
            mlayer_light<Alg, CP, K>(st, cin, cout, M, n, on, A, ck, ap, nr, tmask, xmask, gt, dd, upd, outs,
                                     append_wave);
        } else {
            uint64_t ck[NP], ap = 0;
            int a = 0;
            MGroup<Alg::W> G;
            const DLayerGate<Alg, CP>* gt = nullptr;
            mlayer_rebuild<Alg, CP>(st.key, M, lst[q0], ck, ap, a, G, gt);
            mlayer_heavy<Alg, CP, K>(st, cin, cout, M, n, ck, ap, a, G, gt, dd, upd, outs, append_wave);
        }
    }
    __syncthreads();
    // one reservation per block for the staged appends
    const uint32_t staged = min(s_cnt, (uint32_t)kMStage);
    if (threadIdx.x == 0 && staged) s_base = atomicAdd(&ep.app[e], staged);
    __syncthreads();
    for (uint32_t r = threadIdx.x; r < staged; r += blockDim.x) {
        const uint64_t at = (uint64_t)n + s_base + r;
        if (at < cap) {
            uint64_t kr[NP];
            val_t vr[K];
#pragma unroll
            for (int q = 0; q < NP; ++q) kr[q] = skey[r][q];
#pragma unroll
            for (int kk = 0; kk < K; ++kk) vr[kk] = sval[r][kk];
            put_slot<NP, CP, K>(st, st.key, cout, at, kr, vr);
        } else {
            atomicOr(&ep.flags[0], 1u);
            ep.flags[1] = e;
        }
    }
    dd = block_sum_int(dd);
    for (int o = 32; o > 0; o >>= 1) {
        upd += __shfl_xor(upd, o, 64);
        outs += __shfl_xor(outs, o, 64);
    }
    if (lane == 0) {
        swork[0][wave] = upd;
        swork[1][wave] = outs;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long a = 0, b = 0;
        for (int wv = 0; wv < nwave; ++wv) {
            a += swork[0][wv];
            b += swork[1][wv];
        }
        unsigned long long* wk = M.work + (size_t)(blockIdx.x % kStripes) * 16;
        if (a) atomicAdd(&wk[0], a);
        if (b) atomicAdd(&wk[1], b);
        if (dd) atomicAdd(&ep.dd[e], dd);
    }
}

}  // namespace ffx
