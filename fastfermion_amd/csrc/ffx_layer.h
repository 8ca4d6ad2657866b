// ffx_layer.h -- a layer of single-qubit rotations in one launch.
//
// A layer is a run of consecutive rotations whose axes are single-qubit Paulis
// on pairwise distinct qubits (the RX layer of a kicked-Ising step).  They
// commute, and for a term q the rotations that anticommute with it are those on
// its qubits whose letter the axis does not fix: the set A(q) of "active"
// qubits, |A| <= degree(q).  Rotating never changes a term's support, so the
// orbit of q under the layer is the 2^|A| strings that agree with q outside A
// and carry either of the two letters the axis toggles between on each active
// qubit; every member has the same A.  Applying the layer gate by gate (the
// reference's order, truncation after the gates that carry it) only mixes
// values inside an orbit.  So one launch can do the whole layer:
//
//   1. every slot computes its orbit's canonical key (each active qubit set to
//      a fixed one of its two letters) and claims it in a claim table stamped
//      with this launch's number; the first claimer leads the orbit;
//   2. one wavefront per leader: lane l holds orbit members l, l+64, l+128,
//      l+192 (member bit d toggles the d-th active qubit in layer order), looks
//      them up, then replays the active rotations in layer order as butterfly
//      stages (cross-lane shuffles for bits 0..5, in-lane pairs for bits 6, 7)
//      with exactly the per-rotation operations of the one-rotation kernel;
//   3. existing members get their new value; absent members that end up live
//      are staged in LDS and appended with one global atomic per block.
//
// Internal-error bits in Ctrl::index_full: 1 probe or insert found no free
// entry, 2 the claim table had no free entry, 4 an orbit exceeded kLayerMaxA.
// Requires every live term's degree <= kLayerMaxA (checked by the host: no
// Clifford blocks, maxdegree and observable degree <= kLayerMaxA), so an orbit
// with a live member has at most 256 members; erased slots beyond the bound
// (whose orbits hold no live member) are copied through.  FP64 results are bit-identical to one launch per rotation.
#pragma once

#include "ffx_window.h"

namespace ffx {

constexpr int kLayerMaxA = 8;      // active qubits per orbit (members <= 256)
constexpr int kLayerMax = 256;     // rotations per layer
constexpr int kLayerStage = 512;   // appends staged in LDS per block iteration
constexpr int kLayerMin = 5;       // shorter runs use windows

template <class CP>
struct LayerGate {
    typename CP::fac_t cs, sn;
    int32_t site;
    int32_t axis;    // 1: X, 2: Z, 3: Y
    int32_t thresh;  // threshold truncation after this rotation
    int32_t xw;      // X/Y-weight cap after this rotation
};

template <int W, class CP>
struct Layer {
    const LayerGate<CP>* gates;  // nr records in layer order
    const int16_t* pos;          // qubit -> position in the layer, -1: none (256 entries)
    int32_t* sd;                 // live-count change at each rotation of the layer (nr entries),
                                 // kStripes copies sd_stride words apart (summed by the host)
    uint32_t sd_stride;
    uint64_t* claim;             // claim table (hmask + 1 entries)
    uint64_t mx[W], mz[W], my[W];  // qubits whose axis is X / Z / Y
    uint64_t tmask[4], xmask[4];   // layer positions followed by the threshold / by the X/Y cap
    double mincoeff;
    int nr;
    int maxdeg;
    int max_xw;
    uint32_t stamp;  // this launch's stamp (24 bits, nonzero)
};

// Active qubits of a key: X axes act on Z/Y letters (z set), Z axes on X/Y (x
// set), Y axes on X/Z (x != z).
template <int W>
__device__ __forceinline__ void layer_active(const uint64_t (&k)[2 * W], const uint64_t (&mx)[W],
                                             const uint64_t (&mz)[W], const uint64_t (&my)[W], uint64_t (&am)[W]) {
#pragma unroll
    for (int w = 0; w < W; ++w) am[w] = (k[W + w] & mx[w]) | (k[w] & mz[w]) | ((k[w] ^ k[W + w]) & my[w]);
}

// Canonical member: X-axis active qubits get Z, Z-axis ones X, Y-axis ones X.
template <int W>
__device__ __forceinline__ void layer_canon(const uint64_t (&k)[2 * W], const uint64_t (&am)[W],
                                            const uint64_t (&mx)[W], const uint64_t (&mz)[W],
                                            const uint64_t (&my)[W], uint64_t (&ck)[2 * W]) {
#pragma unroll
    for (int w = 0; w < W; ++w) {
        const uint64_t ax = am[w] & mx[w], az = am[w] & mz[w], ay = am[w] & my[w];
        ck[w] = (k[w] & ~ax) | ay;
        ck[W + w] = k[W + w] & ~az & ~ay;
    }
}

// First position in [lo, hi) of a 256-position mask, or -1.
__device__ __forceinline__ int first_in(const uint64_t (&m)[4], int lo, int hi) {
    int r = -1;
#pragma unroll
    for (int u = 3; u >= 0; --u) {
        uint64_t b = m[u];
        const int a0 = 64 * u;
        if (lo > a0) b &= (lo - a0 >= 64) ? 0ull : (~0ull << (lo - a0));
        if (hi < a0 + 64) b &= (hi <= a0) ? 0ull : (~0ull >> (64 - (hi - a0)));
        if (b) r = a0 + __builtin_ctzll(b);
    }
    return r;
}

// Pop the lowest position of a 256-position mask (nr when empty).
__device__ __forceinline__ int pop_pos(uint64_t (&m)[4], int nr) {
    int r = nr;
#pragma unroll
    for (int u = 3; u >= 0; --u)
        if (m[u]) r = 64 * u + __builtin_ctzll(m[u]);
#pragma unroll
    for (int u = 0; u < 4; ++u)
        if (u == (r >> 6)) m[u] &= m[u] - 1;
    return r;
}

__device__ __forceinline__ int peek_pos(const uint64_t (&m)[4], int nr) {
    int r = nr;
#pragma unroll
    for (int u = 3; u >= 0; --u)
        if (m[u]) r = 64 * u + __builtin_ctzll(m[u]);
    return r;
}

template <class CP>
__device__ __forceinline__ typename CP::val_t shfl_xor_val(typename CP::val_t v, int mask) {
    if constexpr (std::is_same<CP, F32C>::value) {
        typename CP::val_t r;
        r.hi = __shfl_xor(v.hi, mask, 64);
        r.lo = __shfl_xor(v.lo, mask, 64);
        return r;
    } else {
        return __shfl_xor(v, mask, 64);
    }
}

template <class Alg, class CP, int K>
__global__ void __launch_bounds__(kBlock) k_layer_async(Store<CP, K> st, const typename CP::store_t* __restrict__ cin,
                                                        typename CP::store_t* __restrict__ cout,
                                                        Layer<Alg::W, CP> L, Epochs ep, uint32_t e) {
    constexpr int W = Alg::W, NP = Alg::NP;
    using val_t = typename CP::val_t;
    using Gate = typename Alg::Gate;
    __shared__ int16_t spos[256];
    __shared__ LayerGate<CP> sg[kLayerMax];
    __shared__ uint64_t skey[kLayerStage][NP];
    __shared__ val_t sval[kLayerStage][K];
    __shared__ int32_t ssd[kLayerMax];
    __shared__ uint32_t s_cnt, s_base;
    // leaders of the block iteration, by number of active qubits (1..8): their keys and
    // active masks, so that any wave of the block can replay any leader's orbit
    __shared__ uint64_t slk[kBlock][NP];
    __shared__ uint64_t slam[kBlock][W];
    __shared__ uint16_t slist[kLayerMaxA + 1][kBlock];
    __shared__ uint32_t scount[kLayerMaxA + 1];
    uint32_t n, dead;
    if (!step_enter(ep, e, n, dead)) return;
    // slots per block iteration: the fewest of 4, 16, 64 that fit every slot into the
    // grid (small term sets: a block's waves share the orbits of a few slots, and the
    // orbits spread over many blocks), else one per thread
    uint32_t tile = blockDim.x;
    for (uint32_t t = 4; t < blockDim.x; t <<= 2)
        if ((uint64_t)n <= (uint64_t)t * gridDim.x) {
            tile = t;
            break;
        }
    const uint32_t step = gridDim.x * tile;
    if (blockIdx.x * tile >= n) return;  // block-uniform
    for (int t = threadIdx.x; t < 256; t += blockDim.x) spos[t] = L.pos[t];
    for (int t = threadIdx.x; t < L.nr; t += blockDim.x) {
        sg[t] = L.gates[t];
        ssd[t] = 0;
    }
    const size_t cap = st.cap;
    const int lane = __lane_id();
    int dd = 0;
    __syncthreads();
    const int wave = threadIdx.x >> 6, nwave = blockDim.x >> 6;
    for (uint32_t b0 = blockIdx.x * tile; b0 < n; b0 += step) {
        if (threadIdx.x == 0) s_cnt = 0;
        if (threadIdx.x <= kLayerMaxA) scount[threadIdx.x] = 0;
        __syncthreads();
        // ---- phase 1: own slot: copy if inactive, else claim the orbit ----
        const uint32_t i = b0 + threadIdx.x;
        bool leader = false;
        uint64_t k[NP], am[W];
#pragma unroll
        for (int q = 0; q < NP; ++q) k[q] = 0;
#pragma unroll
        for (int w = 0; w < W; ++w) am[w] = 0;
        if (threadIdx.x < tile && i < n) {
            key_load<NP>(st.key, i, k);
            layer_active<W>(k, L.mx, L.mz, L.my, am);
            int a = 0;
#pragma unroll
            for (int w = 0; w < W; ++w) a += popc(am[w]);
            if (a == 0) {
                // untouched by every rotation of the layer; the truncation points inside
                // the layer still apply (to an unchanged value: once is enough)
                val_t c[K];
#pragma unroll
                for (int kk = 0; kk < K; ++kk) c[kk] = CP::load(cin + (size_t)kk * CP::NS * cap, cap, i);
                bool kill = false;
                int at = -1;
                if (!CP::dead(c[0])) {
                    const int pt = first_in(L.tmask, 0, L.nr), px = first_in(L.xmask, 0, L.nr);
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
                // more active qubits than a live term can have (erased slots keep their keys
                // through remaps that skip the degree filter): its whole orbit is erased and
                // stays so; copy it through
                const val_t c0 = CP::load(cin, cap, i);
                if (!CP::dead(c0)) atomicOr(&st.ctrl->index_full, 4u);
#pragma unroll
                for (int kk = 0; kk < K; ++kk)
                    CP::store(cout + (size_t)kk * CP::NS * cap, cap, i, CP::load(cin + (size_t)kk * CP::NS * cap, cap, i));
            } else {
                uint64_t ck[NP];
                layer_canon<W>(k, am, L.mx, L.mz, L.my, ck);
                const uint64_t h = hmix(lin_hash<NP>(ck));
                const uint64_t tag8 = (h >> 56) & 0xFFull;
                const uint64_t mine = ((uint64_t)L.stamp << 40) | (tag8 << 32) | (uint64_t)i;
                uint64_t p = h & st.hmask;
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
                        const uint32_t j = (uint32_t)e0;
                        uint64_t kj[NP], amj[W], ckj[NP];
                        key_load<NP>(st.key, j, kj);
                        layer_active<W>(kj, L.mx, L.mz, L.my, amj);
                        layer_canon<W>(kj, amj, L.mx, L.mz, L.my, ckj);
                        bool eq = true;
#pragma unroll
                        for (int q = 0; q < NP; ++q) eq = eq && ckj[q] == ck[q];
                        if (eq) break;  // another slot leads this orbit
                    }
                    p = (p + 1) & st.hmask;
                }
                if (leader) {
                    if (a > kLayerMaxA) {  // excluded by the host's degree bound; report, do not replay
                        atomicOr(&st.ctrl->index_full, 4u);
                    } else {
#pragma unroll
                        for (int q = 0; q < NP; ++q) slk[threadIdx.x][q] = k[q];
#pragma unroll
                        for (int w = 0; w < W; ++w) slam[threadIdx.x][w] = am[w];
                        slist[a][atomicAdd(&scount[a], 1u)] = (uint16_t)threadIdx.x;
                    }
                }
            }
        }
        __syncthreads();
        // ---- phase 2a: orbits with <= 6 active qubits, 64 / 2^A of them per wave pass ----
        // (lane group g = lane >> A holds one orbit, member T = lane & (2^A - 1)); the
        // passes of all A are numbered in one sequence that the waves take round robin
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
                uint64_t lk[NP], lam[W];
#pragma unroll
                for (int q = 0; q < NP; ++q) lk[q] = on ? slk[src][q] : 0ull;
#pragma unroll
                for (int w = 0; w < W; ++w) lam[w] = on ? slam[src][w] : 0ull;
                uint64_t pm[4] = {0, 0, 0, 0};
#pragma unroll
                for (int w = 0; w < W; ++w) {
                    uint64_t b = lam[w];
                    while (b) {
                        const int t = __builtin_ctzll(b);
                        b &= b - 1;
                        const int pq = spos[64 * w + t];
#pragma unroll
                        for (int u = 0; u < 4; ++u)
                            if (u == (pq >> 6)) pm[u] |= 1ull << (pq & 63);
                    }
                }
                uint64_t ck[NP], key[NP];
                layer_canon<W>(lk, lam, L.mx, L.mz, L.my, ck);
#pragma unroll
                for (int q = 0; q < NP; ++q) key[q] = ck[q];
                const bool recv_ok = Alg::degree(ck) <= L.maxdeg;
                bool exist = false, alive = false;
                uint32_t slot = kNone;
                val_t v[K];
#pragma unroll
                for (int kk = 0; kk < K; ++kk) v[kk] = CP::from_double(0.0);
                if (on) {
                    uint64_t pw[4] = {pm[0], pm[1], pm[2], pm[3]};
                    for (int d = 0; d < A; ++d) {
                        const int pq = pop_pos(pw, 0);
                        if ((T >> d) & 1) {
                            const int q = sg[pq].site, ax = sg[pq].axis;
                            const uint64_t bit = 1ull << (q & 63);
#pragma unroll
                            for (int w = 0; w < W; ++w) {
                                if (w == (q >> 6)) {
                                    if (ax & 1) key[w] ^= bit;
                                    if (ax & 2) key[W + w] ^= bit;
                                }
                            }
                        }
                    }
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
                    if (!alive) return;
                    const int pt = first_in(L.tmask, lo, hi), px = first_in(L.xmask, lo, hi);
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
                uint64_t pw[4] = {pm[0], pm[1], pm[2], pm[3]};
                trunc1(0, peek_pos(pw, L.nr));
                for (int d = 0; d < A; ++d) {
                    const int pq = pop_pos(pw, L.nr);
                    const LayerGate<CP>& lg = sg[on ? pq : 0];
                    Gate gt{};
                    const int q = lg.site;
#pragma unroll
                    for (int w = 0; w < W; ++w) {
                        if (w == (q >> 6)) {
                            if (lg.axis & 1) gt.p[w] = 1ull << (q & 63);
                            if (lg.axis & 2) gt.p[W + w] = 1ull << (q & 63);
                        }
                    }
                    gt.pxz = lg.axis == 3 ? 1 : 0;
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
                            for (int qq = 0; qq < NP; ++qq) ku[qq] = key[qq] ^ gt.p[qq];
                            const int s_in = recv ? Alg::sign(ku, key, gt) : 1;
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
                        trunc1(pq, peek_pos(pw, L.nr));
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
                        const uint32_t r = atomicAdd(&s_cnt, 1u);
                        if (r < (uint32_t)kLayerStage) {
#pragma unroll
                            for (int q2 = 0; q2 < NP; ++q2) skey[r][q2] = key[q2];
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
                    }
                }
            }
        }
        // ---- phase 2b: orbits with 7 or 8 active qubits, one per wave pass ----
        const int n7 = (int)scount[7], n78 = n7 + (int)scount[8];
        for (int jb = ((wave - poff) % nwave + nwave) % nwave; jb < n78; jb += nwave) {
            const int src = jb < n7 ? (int)slist[7][jb] : (int)slist[8][jb - n7];
            uint64_t lk[NP], lam[W];
#pragma unroll
            for (int q = 0; q < NP; ++q) lk[q] = slk[src][q];
#pragma unroll
            for (int w = 0; w < W; ++w) lam[w] = slam[src][w];
            // active qubits as a mask of their layer positions: popping its bits in
            // increasing order walks them in layer order (member bit d = d-th pop)
            uint64_t pm[4] = {0, 0, 0, 0};
            int a = 0;
#pragma unroll
            for (int w = 0; w < W; ++w) {
                uint64_t b = lam[w];
                while (b) {
                    const int t = __builtin_ctzll(b);
                    b &= b - 1;
                    const int pq = spos[64 * w + t];
#pragma unroll
                    for (int u = 0; u < 4; ++u)
                        if (u == (pq >> 6)) pm[u] |= 1ull << (pq & 63);
                    ++a;
                }
            }
            if (a > kLayerMaxA) {  // excluded by the host's degree bound; report, do not replay
                atomicOr(&st.ctrl->index_full, 4u);
                continue;
            }
            uint64_t ck[NP];
            layer_canon<W>(lk, lam, L.mx, L.mz, L.my, ck);
            const int deg = Alg::degree(ck);
            const int nm = 1 << a;
            // members T = lane + 64 m
            uint64_t key[4][NP];
            val_t v[4][K];
            uint32_t slot[4];
            bool alive[4], exist[4];
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                const int T = lane + 64 * m;
                exist[m] = alive[m] = false;
                slot[m] = kNone;
#pragma unroll
                for (int q = 0; q < NP; ++q) key[m][q] = ck[q];
#pragma unroll
                for (int kk = 0; kk < K; ++kk) v[m][kk] = CP::from_double(0.0);
                if (T < nm) {
                    uint64_t pw[4] = {pm[0], pm[1], pm[2], pm[3]};
                    for (int d = 0; d < a; ++d) {
                        const int pq = pop_pos(pw, 0);  // pw holds a positions: never empty here
                        if ((T >> d) & 1) {
                            const int q = sg[pq].site, ax = sg[pq].axis;
                            const uint64_t bit = 1ull << (q & 63);
#pragma unroll
                            for (int w = 0; w < W; ++w) {
                                if (w == (q >> 6)) {
                                    if (ax & 1) key[m][w] ^= bit;      // X or Y: toggles x
                                    if (ax & 2) key[m][W + w] ^= bit;  // Z or Y: toggles z
                                }
                            }
                        }
                    }
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
            const bool recv_ok = deg <= L.maxdeg;  // every member has the orbit's support
            // truncation points in layer positions [lo, hi) applied to the orbit's current
            // values; erasures are counted at the first such point that erases
            auto truncate_range = [&](int lo, int hi) {
                const int pt = first_in(L.tmask, lo, hi), px = first_in(L.xmask, lo, hi);
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
            uint64_t pw[4] = {pm[0], pm[1], pm[2], pm[3]};
            truncate_range(0, peek_pos(pw, L.nr));
            for (int d = 0; d < a; ++d) {
                const int pq = pop_pos(pw, L.nr);
                const LayerGate<CP>& lg = sg[pq];
                Gate g{};
                const int q = lg.site;
#pragma unroll
                for (int w = 0; w < W; ++w) {
                    if (w == (q >> 6)) {
                        if (lg.axis & 1) g.p[w] = 1ull << (q & 63);
                        if (lg.axis & 2) g.p[W + w] = 1ull << (q & 63);
                    }
                }
                g.pxz = lg.axis == 3 ? 1 : 0;
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
                    const bool recv = au[m] && recv_ok;
                    if (alive[m] || recv) {
                        uint64_t ku[NP];
#pragma unroll
                        for (int qq = 0; qq < NP; ++qq) ku[qq] = key[m][qq] ^ g.p[qq];
                        const int s_in = recv ? Alg::sign(ku, key[m], g) : 1;
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
                truncate_range(pq, peek_pos(pw, L.nr));
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
                    const uint32_t r = atomicAdd(&s_cnt, 1u);
                    if (r < (uint32_t)kLayerStage) {
#pragma unroll
                        for (int q = 0; q < NP; ++q) skey[r][q] = key[m][q];
#pragma unroll
                        for (int kk = 0; kk < K; ++kk) sval[r][kk] = v[m][kk];
                    } else {
                        // staging full: append directly
                        const uint64_t at = (uint64_t)n + atomicAdd(&ep.app[e], 1u);
                        if (at < cap) {
                            put_slot<NP, CP, K>(st, st.key, cout, at, key[m], v[m]);
                        } else {
                            atomicOr(&ep.flags[0], 1u);
                            ep.flags[1] = e;
                        }
                    }
                }
            }
        }
        __syncthreads();
        // ---- phase 3: one reservation per block for the staged appends ----
        const uint32_t staged = min(s_cnt, (uint32_t)kLayerStage);
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
