// ffx_dlayer.h -- a layer of rotations with pairwise disjoint supports in one
// launch (Pauli axes on disjoint qubit sets, Majorana strings of even degree on
// disjoint index sets; e.g. one bond matching or the on-site terms of a
// Fermi-Hubbard Trotter step).
//
// Such rotations commute, and flipping a term q by one axis P_j changes neither
// its commutation with any other axis of the layer (their supports do not meet
// P_j's) nor with P_j itself (the axes have even overlap parity with themselves).
// So the set A(q) of layer rotations that anticommute with q is the same for
// every member of q's orbit q + span{P_j : j in A(q)}; each j in A(q) has a unit
// (qubit or Majorana index) in q's support, so |A| <= degree(q).  Every axis has
// a private pivot bit (the lowest bit of its key mask, set in no other axis of
// the layer): the orbit's canonical member is the one with the pivot bits of
// its active axes cleared.  Applying the layer gate by gate only mixes values
// inside an orbit, so one launch can do the whole layer exactly as the layer of
// single-qubit rotations does (ffx_layer.h): claim the canonical key, one lane
// group or wave per leader, butterfly replay of the active rotations in layer
// order with the one-rotation kernel's operations (per-member degree filter and
// sign), truncation points between stages, staged appends.
//
// General layers: at most kLayerMaxA (8) consecutive rotations that commute
// pairwise and have linearly independent axes, supports overlapping.  The set
// A(q) is again an orbit invariant (commutation is a bicharacter, so flipping q
// by P_j leaves q's commutation with every P_k unchanged when P_j and P_k
// commute), |A| <= 8 because the layer holds at most 8 rotations, and the 2^|A|
// members are distinct because the axes are independent.  With no private pivot
// bits, the canonical member clears the active axes' coordinates of the key's
// component in span{P}, read off with the axes' dual vectors (host-computed).
//
// Requires every live term's degree <= kLayerMaxA (host check as for layers).
#pragma once

#include "ffx_layer.h"

namespace ffx {

// appends staged in LDS per block iteration (more go straight to HBM); 256 keeps
// the kernel at 4 blocks of 256 threads per CU for keys of up to 4 words
constexpr int kDLayerStage = 256;
// consecutive slots per run of a block iteration's tile
constexpr int kDLayerRun = 4;

template <class Alg, class CP>
struct DLayerGate {
    typename Alg::Gate g;
    typename CP::fac_t cs, sn;
    int32_t pivot;  // private pivot bit (key bit index, word pivot / 64)
};

template <int W, class CP, class Alg>
struct DLayer {
    const DLayerGate<Alg, CP>* gates;  // nr records in layer order
    const int16_t* pos;                // unit -> position in the layer, -1: none (256 entries);
                                       // general layers: the axes' dual vectors (kLayerMaxA x NP words)
    int32_t* sd;                       // live-count change at each rotation (nr entries), kStripes copies
    uint32_t sd_stride;
    uint64_t* claim;                   // claim table (hmask + 1 entries)
    uint64_t umask[W];                 // units covered by the layer's axes
    uint64_t tmask[4], xmask[4];       // layer positions followed by the threshold / by the X/Y cap
    double mincoeff;
    int nr;
    int maxdeg;
    int max_xw;
    uint32_t stamp;  // this launch's stamp (24 bits, nonzero)
    int general;     // 1: at most kLayerMaxA commuting rotations with overlapping supports
    int spread;      // 1: small tiles taken in runs spread over the store (layers with threshold points)
    int tflags;      // kTruncAny | kTruncXw | kTruncAll (dlayer_tflags)
};

constexpr int kTruncAny = 1;  // some threshold or X/Y-cap point in the layer
constexpr int kTruncXw = 2;   // some X/Y-cap point
constexpr int kTruncAll = 4;  // a threshold point after every rotation

__host__ inline int dlayer_tflags(const uint64_t (&tmask)[4], const uint64_t (&xmask)[4], int nr) {
    bool any = false, xw = false, all = true;
    for (int u = 0; u < 4; ++u) {
        any = any || tmask[u] || xmask[u];
        xw = xw || xmask[u];
        const int lo = 64 * u, hi = nr < 64 * (u + 1) ? nr : 64 * (u + 1);
        const uint64_t want = hi <= lo ? 0ull : (hi - lo >= 64 ? ~0ull : ((1ull << (hi - lo)) - 1));
        all = all && (tmask[u] & want) == want;
    }
    return (any ? kTruncAny : 0) | (xw ? kTruncXw : 0) | (all ? kTruncAll : 0);
}

// Units of a key (qubits in the Pauli support, Majorana indices), word w.
template <class Alg>
__device__ __forceinline__ uint64_t key_units(const uint64_t (&k)[Alg::NP], int w) {
    if constexpr (Alg::kPauli) {
        return k[w] | k[Alg::W + w];
    } else {
        return k[w];
    }
}

// Positions of the layer rotations that anticommute with key k (a mask of
// layer positions); *na receives their number.
template <class Alg, class CP>
__device__ __forceinline__ void dlayer_active(const uint64_t (&k)[Alg::NP], const uint64_t* umask,
                                              const int16_t* spos, const DLayerGate<Alg, CP>* sg, uint64_t (&pm)[4],
                                              int& na) {
    constexpr int W = Alg::W;
    uint64_t cand[4] = {0, 0, 0, 0};
#pragma unroll
    for (int w = 0; w < W; ++w) {
        uint64_t b = key_units<Alg>(k, w) & umask[w];
        while (b) {
            const int t = __builtin_ctzll(b);
            b &= b - 1;
            const int p = spos[64 * w + t];
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (u == (p >> 6)) cand[u] |= 1ull << (p & 63);
        }
    }
    na = 0;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        pm[u] = 0;
        uint64_t b = cand[u];
        while (b) {
            const int t = __builtin_ctzll(b);
            b &= b - 1;
            if (Alg::anticommutes(k, sg[64 * u + t].g)) {
                pm[u] |= 1ull << t;
                ++na;
            }
        }
    }
}

// Positions of a general layer's rotations (nr <= kLayerMaxA) that anticommute with k.
// A Pauli axis anticommutes only with strings it shares a qubit with, so a Pauli key
// with no unit under the layer's axes (umask) has none.  (A Majorana axis of odd
// degree can anticommute without overlap.)
template <class Alg, class CP>
__device__ __forceinline__ void dlayer_active_all(const uint64_t (&k)[Alg::NP], const uint64_t* umask,
                                                  const DLayerGate<Alg, CP>* sg, int nr, uint64_t (&pm)[4], int& na) {
    pm[0] = pm[1] = pm[2] = pm[3] = 0;
    na = 0;
    if constexpr (Alg::kPauli) {
        uint64_t touch = 0;
#pragma unroll
        for (int w = 0; w < Alg::W; ++w) touch |= key_units<Alg>(k, w) & umask[w];
        if (!touch) return;
    }
    for (int j = 0; j < nr; ++j)
        if (Alg::anticommutes(k, sg[j].g)) {
            pm[0] |= 1ull << j;
            ++na;
        }
}

// Canonical member of a general layer's orbit from the axes' dual vectors
// (parity(P_i & w_j) = [i == j]): k with the span{P} coordinate a_j =
// parity(k & w_j) of every active axis j cleared, i.e. k ^ sum_{j active, a_j = 1} P_j.
// Flipping k by an active axis toggles only that axis's coordinate, so every
// member of the orbit gives the same key (A parities instead of a 2^A walk).
template <class Alg, class CP>
__device__ __forceinline__ void dlayer_canon_dual(const uint64_t (&k)[Alg::NP], uint64_t pm0,
                                                  const uint64_t (*dual)[Alg::NP], const DLayerGate<Alg, CP>* sg,
                                                  uint64_t (&ck)[Alg::NP]) {
    constexpr int NP = Alg::NP;
#pragma unroll
    for (int q = 0; q < NP; ++q) ck[q] = k[q];
    while (pm0) {
        const int j = __builtin_ctzll(pm0);
        pm0 &= pm0 - 1;
        int par = 0;
#pragma unroll
        for (int q = 0; q < NP; ++q) par ^= __builtin_popcountll(k[q] & dual[j][q]);
        if (par & 1)
#pragma unroll
            for (int q = 0; q < NP; ++q) ck[q] ^= sg[j].g.p[q];
    }
}

// Canonical member: the pivot bit of every active axis cleared.
template <class Alg, class CP>
__device__ __forceinline__ void dlayer_canon(const uint64_t (&k)[Alg::NP], const uint64_t (&pm)[4],
                                             const DLayerGate<Alg, CP>* sg, uint64_t (&ck)[Alg::NP]) {
    constexpr int NP = Alg::NP;
#pragma unroll
    for (int q = 0; q < NP; ++q) ck[q] = k[q];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        uint64_t b = pm[u];
        while (b) {
            const int t = __builtin_ctzll(b);
            b &= b - 1;
            const DLayerGate<Alg, CP>& lg = sg[64 * u + t];
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
}

// Active positions of an orbit (at most kLayerMaxA = 8, each < 256) packed in
// layer order, 8 bits each.
__device__ __forceinline__ uint64_t pack_positions(const uint64_t (&pm)[4]) {
    uint64_t ap = 0;
    int d = 0;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        uint64_t b = pm[u];
        while (b) {
            ap |= (uint64_t)(64 * u + __builtin_ctzll(b)) << (8 * d);
            b &= b - 1;
            ++d;
        }
    }
    return ap;
}

// d-th packed position, nr past the last of the a positions.
__device__ __forceinline__ int apos(uint64_t ap, int d, int a, int nr) {
    return d < a ? (int)((ap >> (8 * d)) & 0xFFu) : nr;
}

// Member T of the orbit of canonical key ck: the axes of the active positions
// whose bit d (d-th active position in layer order) is set in T.
template <class Alg, class CP>
__device__ __forceinline__ void dlayer_member(const uint64_t (&ck)[Alg::NP], uint64_t ap, int a, int T,
                                              const DLayerGate<Alg, CP>* sg, uint64_t (&key)[Alg::NP]) {
    constexpr int NP = Alg::NP;
#pragma unroll
    for (int q = 0; q < NP; ++q) key[q] = ck[q];
    for (int d = 0; d < a; ++d) {
        if ((T >> d) & 1) {
            const int pq = (int)((ap >> (8 * d)) & 0xFFu);
#pragma unroll
            for (int q = 0; q < NP; ++q) key[q] ^= sg[pq].g.p[q];
        }
    }
}

// TR: the layer has truncation points (kTruncAny); without, every per-member
// truncation check is compiled out.
template <class Alg, class CP, int K, bool TR>
__global__ void __launch_bounds__(kBlock, 4) k_dlayer_async(Store<CP, K> st, const typename CP::store_t* __restrict__ cin,
                                                         typename CP::store_t* __restrict__ cout,
                                                         DLayer<Alg::W, CP, Alg> L, Epochs ep, uint32_t e) {
    constexpr int NP = Alg::NP;
    using val_t = typename CP::val_t;
    __shared__ union {
        int16_t pos[256];                        // disjoint layers: unit -> position
        uint64_t dual[kLayerMaxA][NP];           // general layers: dual vectors of the axes
    } su;
    int16_t* const spos = su.pos;
    __shared__ DLayerGate<Alg, CP> sg[kLayerMax];
    __shared__ uint64_t skey[kDLayerStage][NP];
    __shared__ val_t sval[kDLayerStage][K];
    __shared__ int32_t ssd[kLayerMax];
    __shared__ uint32_t s_cnt, s_base;
    // leaders of the block iteration, by number of active axes (1..8): their keys and
    // packed active positions, so that any wave of the block can replay any leader's orbit
    __shared__ uint64_t slk[kBlock][NP];
    __shared__ uint64_t sap[kBlock];
    __shared__ uint8_t slist[kLayerMaxA + 1][kBlock];
    static_assert(kBlock <= 256, "leader lists hold 8-bit thread indices");
    __shared__ uint32_t scount[kLayerMaxA + 1];
    uint32_t n, dead;
    if (!step_enter(ep, e, n, dead)) return;
    // slots per block iteration: the fewest of 4, 16, 64 that fit every slot into the
    // grid, else one per thread (as in the single-qubit layer kernel)
    uint32_t tile = blockDim.x;
    for (uint32_t t = 4; t < blockDim.x; t <<= 2)
        if ((uint64_t)n <= (uint64_t)t * gridDim.x) {
            tile = t;
            break;
        }
    // layers of at least 16 rotations with threshold points (L.spread), when every
    // slot fits into one block iteration (tiles of at most 64 slots): the tile is
    // taken in runs of kDLayerRun consecutive slots spread over the store (run j of
    // block b at ((j * grid + b) * kDLayerRun)).  The fresh members of one orbit are
    // appended side by side and their orbits under the next large layer are often
    // all large; with truncation inside the layer such an orbit leaves few members
    // but costs its full size, and contiguous tiles would hand a few blocks most of
    // the layer's large orbits (6x6 Fermi-Hubbard, layers of 36-72 rotations,
    // threshold after every rotation: 8.3 ms with contiguous tiles, 7.2 ms with runs
    // of 16 slots, 6.8 ms with runs of 4 or 1).  Otherwise tiles stay contiguous: an
    // orbit's members then tend to share a block's cache lines (the kicked-Ising
    // layers of up to 10 rotations lose 4 % without threshold points and 11 % with
    // one after every rotation when spread).
    const uint32_t run = (!L.spread || tile == blockDim.x) ? tile
                                                           : (tile < (uint32_t)kDLayerRun ? tile : (uint32_t)kDLayerRun);
    const uint64_t step = (uint64_t)gridDim.x * tile;
    if ((uint64_t)blockIdx.x * run >= n) return;  // block-uniform
    for (int t = threadIdx.x; t < 256; t += blockDim.x) spos[t] = L.pos[t];  // a plain copy of either form
    for (int t = threadIdx.x; t < L.nr; t += blockDim.x) {
        sg[t] = L.gates[t];
        ssd[t] = 0;
    }
    const size_t cap = st.cap;
    const int lane = __lane_id();
    int dd = 0;
    __syncthreads();
    const int wave = threadIdx.x >> 6, nwave = blockDim.x >> 6;
    // a layer without truncation points (kernel-argument flags, so the tests are
    // uniform) skips every per-member truncation check; without X/Y caps, or with a
    // threshold after every rotation, the search of the position masks is skipped too
    constexpr bool tr_any = TR;
    const bool xm_any = (L.tflags & kTruncXw) != 0;
    const bool tm_all = (L.tflags & kTruncAll) != 0;
    // first threshold / X/Y-cap position in [lo, hi), -1: none
    auto first_t = [&](int lo, int hi) -> int { return tm_all ? (lo < hi ? lo : -1) : first_in(L.tmask, lo, hi); };
    auto first_x = [&](int lo, int hi) -> int { return xm_any ? first_in(L.xmask, lo, hi) : -1; };
    // stage a fresh live member for appending (one global reservation per block iteration)
    auto stage_append = [&](const uint64_t (&key)[NP], const val_t (&v)[K]) {
        const uint32_t r = atomicAdd(&s_cnt, 1u);
        if (r < (uint32_t)kDLayerStage) {
#pragma unroll
            for (int q = 0; q < NP; ++q) skey[r][q] = key[q];
#pragma unroll
            for (int kk = 0; kk < K; ++kk) sval[r][kk] = v[kk];
        } else {
            const uint64_t at = (uint64_t)n + atomicAdd(&ep.app[e], 1u);
            if (at < cap) {
                put_slot<NP, CP, K>(st, st.key, cout, at, key, v);
            } else {
                atomicOr(&ep.flags[0], 1u);
                ep.flags[1] = e;
            }
        }
    };
    for (uint64_t b0 = 0; b0 + (uint64_t)blockIdx.x * run < n; b0 += step) {
        if (threadIdx.x == 0) s_cnt = 0;
        if (threadIdx.x <= kLayerMaxA) scount[threadIdx.x] = 0;
        __syncthreads();
        // ---- phase 1: own slot: copy if inactive, else claim the orbit ----
        const uint64_t i64 =
            b0 + ((uint64_t)(threadIdx.x / run) * gridDim.x + blockIdx.x) * run + threadIdx.x % run;
        const uint32_t i = (uint32_t)i64;
        if (threadIdx.x < tile && i64 < n) {
            uint64_t k[NP], pm[4];
            key_load<NP>(st.key, i, k);
            int a = 0;
            if (L.general)
                dlayer_active_all<Alg, CP>(k, L.umask, sg, L.nr, pm, a);
            else
                dlayer_active<Alg, CP>(k, L.umask, spos, sg, pm, a);
            if (a == 0) {
                // untouched by every rotation of the layer; the truncation points inside
                // the layer still apply (to an unchanged value: once is enough)
                val_t c[K];
#pragma unroll
                for (int kk = 0; kk < K; ++kk) c[kk] = CP::load(cin + (size_t)kk * CP::NS * cap, cap, i);
                bool kill = false;
                int at = -1;
                if (tr_any && !CP::dead(c[0])) {
                    const int pt = first_t(0, L.nr), px = first_x(0, L.nr);
                    if (pt >= 0 && magnitude<CP, K>(c) <= L.mincoeff) {
                        kill = true;
                        at = pt;
                    }
                    if (px >= 0 && Alg::xweight(k) > L.max_xw && (!kill || px < at)) {
                        kill = true;
                        at = px;
                    }
                }
                if (kill) {
                    c[0] = CP::make_dead();
#pragma unroll
                    for (int kk = 1; kk < K; ++kk) c[kk] = CP::from_double(0.0);
                    ++dd;
                    atomicAdd(&ssd[at], -1);
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
                    CP::store(cout + (size_t)kk * CP::NS * cap, cap, i, CP::load(cin + (size_t)kk * CP::NS * cap, cap, i));
            } else {
                uint64_t ck[NP];
                if (L.general)
                    dlayer_canon_dual<Alg, CP>(k, pm[0], su.dual, sg, ck);
                else
                    dlayer_canon<Alg, CP>(k, pm, sg, ck);
                const uint64_t h = hmix(lin_hash<NP>(ck));
                const uint64_t tag8 = (h >> 56) & 0xFFull;
                const uint64_t mine = ((uint64_t)L.stamp << 40) | (tag8 << 32) | (uint64_t)i;
                uint64_t p = h & st.hmask;
                bool leader = false;
                for (uint64_t s = 0;; ++s) {
                    if (s > st.hmask) {
                        atomicOr(&st.ctrl->index_full, 2u);
                        break;
                    }
                    uint64_t e0 = L.claim[p];
                    if ((e0 >> 40) != L.stamp) {
                        const uint64_t old = (uint64_t)atomicCAS((unsigned long long*)&L.claim[p],
                                                                 (unsigned long long)e0, (unsigned long long)mine);
                        if (old == e0) {
                            leader = true;
                            break;
                        }
                        e0 = old;
                        if ((e0 >> 40) != L.stamp) continue;  // overwritten meanwhile: look again
                    }
                    if (((e0 >> 32) & 0xFFull) == tag8) {
                        // the claimant's key kj is in k's orbit iff reducing it by k's
                        // active axes gives k's canonical member (the active set is an
                        // orbit invariant, so kj's own is not needed)
                        const uint32_t j = (uint32_t)e0;
                        uint64_t kj[NP], ckj[NP];
                        key_load<NP>(st.key, j, kj);
                        if (L.general)
                            dlayer_canon_dual<Alg, CP>(kj, pm[0], su.dual, sg, ckj);
                        else
                            dlayer_canon<Alg, CP>(kj, pm, sg, ckj);
                        bool eq = true;
#pragma unroll
                        for (int q = 0; q < NP; ++q) eq = eq && ckj[q] == ck[q];
                        if (eq) break;  // another slot leads this orbit
                    }
                    p = (p + 1) & st.hmask;
                }
                if (leader) {
#pragma unroll
                    for (int q = 0; q < NP; ++q) slk[threadIdx.x][q] = ck[q];
                    sap[threadIdx.x] = pack_positions(pm);
                    slist[a][atomicAdd(&scount[a], 1u)] = (uint8_t)threadIdx.x;
                }
            }
        }
        __syncthreads();
        // ---- phase 2a: orbits with <= 6 active axes, 64 / 2^A of them per wave pass ----
        int poff = 0;
        for (int A = 1; A <= 6; ++A) {
            const int nlead = (int)scount[A], per = 64 >> A;
            const int npass = (nlead + per - 1) / per;
            const int pass0 = ((wave - poff) % nwave + nwave) % nwave;
            poff += npass;
            for (int pass = pass0; pass < npass; pass += nwave) {
                const int g = lane >> A, T = lane & ((1 << A) - 1);
                const int kth = pass * per + g;
                const bool on = kth < nlead;
                const int src = on ? (int)slist[A][kth] : 0;
                uint64_t ck[NP], key[NP];
#pragma unroll
                for (int q = 0; q < NP; ++q) ck[q] = on ? slk[src][q] : 0ull;
                const uint64_t ap = on ? sap[src] : 0ull;
                const int Aon = on ? A : 0;
                dlayer_member<Alg, CP>(ck, ap, Aon, T, sg, key);
                const bool recv_ok = Alg::degree(key) <= L.maxdeg;
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
                // truncation points in [lo, hi) on this lane's member
                auto trunc1 = [&](int lo, int hi) {
                    if (!tr_any) return;
                    if (!alive) return;
                    const int pt = first_t(lo, hi), px = first_x(lo, hi);
                    const bool bt = pt >= 0 && magnitude<CP, K>(v) <= L.mincoeff;
                    const bool bx = px >= 0 && Alg::xweight(key) > L.max_xw;
                    if (bt && (!bx || pt <= px)) {
                        alive = false;
                        atomicAdd(&ssd[pt], -1);
                    } else if (bx) {
                        alive = false;
                        atomicAdd(&ssd[px], -1);
                    }
                };
                trunc1(0, apos(ap, 0, Aon, L.nr));
                for (int d = 0; d < A; ++d) {
                    const int pq = apos(ap, d, Aon, L.nr);
                    const DLayerGate<Alg, CP>& lg = sg[on ? pq : 0];
                    val_t vu[K];
#pragma unroll
                    for (int kk = 0; kk < K; ++kk) vu[kk] = shfl_xor_val<CP>(v[kk], 1 << d);
                    const bool au = __shfl_xor((int)alive, 1 << d, 64) != 0;
                    if (on) {
                        const bool before = alive;
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
                            alive = true;
                        }
                        if (alive != before) atomicAdd(&ssd[pq], alive ? 1 : -1);
                        trunc1(pq, apos(ap, d + 1, Aon, L.nr));
                    }
                }
                if (on) {
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
                    } else if (alive) {
                        stage_append(key, v);
                    }
                }
            }
        }
        // ---- phase 2b: orbits with 7 or 8 active axes, one per wave pass ----
        const int n7 = (int)scount[7], n78 = n7 + (int)scount[8];
        for (int jb = ((wave - poff) % nwave + nwave) % nwave; jb < n78; jb += nwave) {
            const int src = jb < n7 ? (int)slist[7][jb] : (int)slist[8][jb - n7];
            const int a = jb < n7 ? 7 : 8;
            uint64_t ck[NP];
#pragma unroll
            for (int q = 0; q < NP; ++q) ck[q] = slk[src][q];
            const uint64_t ap = sap[src];
            const int nm = 1 << a;
            // members T = lane + 64 m
            uint64_t key[4][NP];
            val_t v[4][K];
            uint32_t slot[4];
            bool alive[4], exist[4], recv_ok[4];
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                const int T = lane + 64 * m;
                exist[m] = alive[m] = false;
                slot[m] = kNone;
#pragma unroll
                for (int kk = 0; kk < K; ++kk) v[m][kk] = CP::from_double(0.0);
                dlayer_member<Alg, CP>(ck, ap, T < nm ? a : 0, T, sg, key[m]);
                recv_ok[m] = Alg::degree(key[m]) <= L.maxdeg;
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
            // truncation points in layer positions [lo, hi) applied to the orbit's current
            // values; erasures are counted at the first such point that erases
            auto truncate_range = [&](int lo, int hi) {
                if (!tr_any) return;
                const int pt = first_t(lo, hi), px = first_x(lo, hi);
                if (pt < 0 && px < 0) return;
                int kt = 0, kx = 0;
#pragma unroll
                for (int m = 0; m < 4; ++m) {
                    if (!alive[m]) continue;
                    const bool bt = pt >= 0 && magnitude<CP, K>(v[m]) <= L.mincoeff;
                    const bool bx = px >= 0 && Alg::xweight(key[m]) > L.max_xw;
                    if (bt && (!bx || pt <= px)) {
                        alive[m] = false;
                        ++kt;
                    } else if (bx) {
                        alive[m] = false;
                        ++kx;
                    }
                }
                for (int o = 32; o > 0; o >>= 1) {
                    kt += __shfl_xor(kt, o, 64);
                    kx += __shfl_xor(kx, o, 64);
                }
                if (lane == 0) {
                    if (kt) atomicAdd(&ssd[pt], -kt);
                    if (kx) atomicAdd(&ssd[px], -kx);
                }
            };
            truncate_range(0, apos(ap, 0, a, L.nr));
            for (int d = 0; d < a; ++d) {
                const int pq = apos(ap, d, a, L.nr);
                const DLayerGate<Alg, CP>& lg = sg[pq];
                const int bitd = 1 << d;
                int before = 0, after = 0;
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
                    before += alive[m];
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
                        alive[m] = true;
                    }
                    after += alive[m];
                }
                // live-count change of the orbit from this rotation, then its truncation points
                int delta = after - before;
                for (int o = 32; o > 0; o >>= 1) delta += __shfl_xor(delta, o, 64);
                if (lane == 0 && delta) atomicAdd(&ssd[pq], delta);
                truncate_range(pq, apos(ap, d + 1, a, L.nr));
            }
            // write back existing members, stage fresh live ones
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                const int T = lane + 64 * m;
                if (T >= nm) continue;
                if (exist[m]) {
                    if (alive[m]) {
#pragma unroll
                        for (int kk = 0; kk < K; ++kk)
                            CP::store(cout + (size_t)kk * CP::NS * cap, cap, slot[m], v[m][kk]);
                    } else {
                        CP::store(cout, cap, slot[m], CP::make_dead());
#pragma unroll
                        for (int kk = 1; kk < K; ++kk)
                            CP::store(cout + (size_t)kk * CP::NS * cap, cap, slot[m], CP::from_double(0.0));
                    }
                    dd += (int)alive0[m] - (int)alive[m];
                } else if (alive[m]) {
                    stage_append(key[m], v[m]);
                }
            }
        }
        __syncthreads();
        // ---- phase 3: one reservation per block for the staged appends ----
        const uint32_t staged = min(s_cnt, (uint32_t)kDLayerStage);
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
        __syncthreads();
    }
    dd = block_sum_int(dd);
    if (threadIdx.x == 0 && dd) atomicAdd(&ep.dd[e], dd);
    int32_t* sd = L.sd + (size_t)(blockIdx.x % kStripes) * L.sd_stride;
    for (int t = threadIdx.x; t < L.nr; t += blockDim.x)
        if (ssd[t]) atomicAdd(&sd[t], ssd[t]);
}

}  // namespace ffx
