// ffx_window.h -- several consecutive rotations in one launch.
//
// A window is up to kWin consecutive rotations that pairwise commute and whose
// axes are linearly independent over GF(2).  Let A be the set of window gates a
// term q anticommutes with.  Commutation with a fixed axis is a GF(2)-linear
// function of the term's bits, and the window axes commute with each other, so
// every string of the orbit q ^ span{P_j : j in A} anticommutes with exactly
// the gates in A, and the orbit has 2^|A| distinct strings.  Applying the
// window gate by gate, as the reference does (src/pauli/propagate.h:71-89,
// src/majorana_propagation.h:71-96), therefore only mixes coefficients inside
// such an orbit, and the whole window can be replayed orbit by orbit with
// exactly the operations of the one-rotation kernel (c*cos + s*sin*c_partner,
// the partner-only degree filter, threshold and X/Y cap after the rotations
// that are truncation points): the result is bit-identical to one launch per
// rotation.
//
// Orbit members are indexed by the subset T of window positions whose axes are
// XORed into a key; all member indices are compile-time constants (sfor), so
// an orbit's values stay in registers.  Every slot looks up its whole orbit and
// replays it, then writes its own final value.  Member hashes come from the
// slot's own linear pre-hash and a per-window table, and the replay uses
// per-member degree and sign facts instead of the member keys.  The existing
// member with the smallest slot index (the orbit's leader) appends the absent
// members that end up live and records the orbit's live-count change at every
// rotation, so the host recovers the exact number of live terms entering each
// rotation.
#pragma once

#include <utility>

#include "ffx_async.h"

namespace ffx {

constexpr int kWin = 4;            // rotations per window
constexpr int kOrbit = 1 << kWin;  // orbit members per term
constexpr int kBlockWin = 256;

template <int... I, class F>
__device__ __forceinline__ void sfor_impl(std::integer_sequence<int, I...>, F&& f) {
    (f(std::integral_constant<int, I>{}), ...);
}
// f(integral_constant<int, 0>) ... f(integral_constant<int, N-1>)
template <int N, class F>
__device__ __forceinline__ void sfor(F&& f) {
    sfor_impl(std::make_integer_sequence<int, N>{}, f);
}

template <class Alg, class CP>
struct Window {
    typename Alg::Gate g[kWin];
    typename CP::fac_t cs[kWin], sn[kWin];
    double mincoeff;
    int k;            // rotations in the window (1..kWin)
    int thresh_mask;  // bit s: threshold truncation after rotation s
    int xw_mask;      // bit s: X/Y-weight cap after rotation s
    int maxdeg;
    int max_xw;
    uint64_t lp[kWin];     // lin_hash of each axis
    uint32_t cross[kWin];  // bit j of cross[s]: Alg::lin(P_j, gate s)
};

// XOR of lp[j] over j in T (compile-time subset): lin_hash(k ^ P_T) = lin_hash(k) ^ lp_of<T>.
template <int T, class Alg, class CP>
__device__ __forceinline__ uint64_t lp_of(const Window<Alg, CP>& w) {
    uint64_t h = 0;
    sfor<kWin>([&](auto JJ) {
        constexpr int J = decltype(JJ)::value;
        if constexpr ((T >> J) & 1) h ^= w.lp[J];
    });
    return h;
}

// Per-member facts the orbit replay needs instead of the member keys.
struct OrbitFacts {
    uint32_t rok;  // bit T: member T passes the degree filter
    uint32_t qm;   // bits 2T..2T+1: Alg::quad of member T
    uint32_t lin;  // bit s: Alg::lin(slot key, gate s)
};

// Key of orbit member T (compile-time subset): k ^ XOR_{j in T} P_j.
template <int T, class Alg, class CP>
__device__ __forceinline__ void orbit_key(const uint64_t (&k)[Alg::NP], const Window<Alg, CP>& w,
                                          uint64_t (&o)[Alg::NP]) {
#pragma unroll
    for (int q = 0; q < Alg::NP; ++q) o[q] = k[q];
    sfor<kWin>([&](auto JJ) {
        constexpr int J = decltype(JJ)::value;
        if constexpr ((T >> J) & 1) {
#pragma unroll
            for (int q = 0; q < Alg::NP; ++q) o[q] ^= w.g[J].p[q];
        }
    });
}

template <class Alg>
__device__ __forceinline__ uint32_t window_amask(const uint64_t (&k)[Alg::NP], const typename Alg::Gate (&g)[kWin],
                                                 int nk) {
    uint32_t a = 0;
    sfor<kWin>([&](auto JJ) {
        constexpr int J = decltype(JJ)::value;
        if (J < nk && Alg::anticommutes(k, g[J])) a |= 1u << J;
    });
    return a;
}

template <int T, class Alg>
__device__ __forceinline__ void member_facts(const uint64_t (&kt)[Alg::NP], int maxdeg, OrbitFacts& f) {
    if (Alg::degree(kt) <= maxdeg) f.rok |= 1u << T;
    f.qm |= (uint32_t)Alg::quad(kt) << (2 * T);
}

// Continues the lookup of kt (index hash h) past its home entry, which was
// examined and does not hold it, and loads the candidate's value together
// with its key words.  Returns the slot among [0, n), or kNone.
template <class Alg, class CP, int K>
__device__ __forceinline__ uint32_t probe_next(const Store<CP, K>& st, const typename CP::store_t* cin,
                                               const uint64_t (&kt)[Alg::NP], uint64_t h, uint32_t n,
                                               typename CP::val_t (&vt)[K]) {
    constexpr int NP = Alg::NP;
    const size_t cap = st.cap;
    const uint32_t tag = (uint32_t)(h >> 32);
    uint64_t p = h & st.hmask;
    for (uint64_t step = 1;; ++step) {
        if (step > st.hmask) {
            atomicOr(&st.ctrl->index_full, 1u);
            return kNone;
        }
        p = (p + 1) & st.hmask;
        const uint64_t e = st.index[p];
        if (e == kEmpty) return kNone;
        const uint32_t idx = (uint32_t)e;
        if ((uint32_t)(e >> 32) == tag && idx < n) {
#pragma unroll
            for (int kk = 0; kk < K; ++kk) vt[kk] = CP::load(cin + (size_t)kk * CP::NS * cap, cap, idx);
            uint64_t c[NP];
            key_load<NP>(st.key, idx, c);
            bool eq = true;
#pragma unroll
            for (int q = 0; q < NP; ++q) eq = eq && c[q] == kt[q];
            if (eq) return idx;
        }
    }
}

// Looks up the orbit members T != 0 (valid: subsets of the anticommuting set)
// among slots [0, n), loads their values and records their replay facts: home
// index entries for all members first (hashes from the slot's lin_hash lk),
// then the key checks, then the probes past the home entry for the rest.
template <class Alg, class CP, int K>
__device__ __forceinline__ void orbit_gather(const Store<CP, K>& st, const typename CP::store_t* cin,
                                             const Window<Alg, CP>& w, const uint64_t (&k)[Alg::NP], uint64_t lk,
                                             uint32_t valid, uint32_t n, uint32_t (&slot)[kOrbit],
                                             typename CP::val_t (&v)[kOrbit][K], uint32_t& exist, uint32_t& alive,
                                             OrbitFacts& f) {
    constexpr int NP = Alg::NP;
    const size_t cap = st.cap;
    uint64_t ent[kOrbit];
    uint32_t tag[kOrbit];
    sfor<kOrbit>([&](auto TT) {
        constexpr int T = decltype(TT)::value;
        if constexpr (T > 0) {
            slot[T] = kNone;
            if ((valid >> T) & 1) {
                const uint64_t h = hmix(lk ^ lp_of<T>(w));
                tag[T] = (uint32_t)(h >> 32);
                ent[T] = st.index[h & st.hmask];
                uint64_t kt[NP];
                orbit_key<T>(k, w, kt);
                member_facts<T, Alg>(kt, w.maxdeg, f);
            }
        }
    });
    uint32_t need_probe = 0, early = 0;
    sfor<kOrbit>([&](auto TT) {
        constexpr int T = decltype(TT)::value;
        if constexpr (T > 0) {
            if (((valid >> T) & 1) && ent[T] != kEmpty) {
                const uint32_t idx = (uint32_t)ent[T];
                if ((uint32_t)(ent[T] >> 32) == tag[T] && idx < n) {
                    uint64_t kt[NP];
                    orbit_key<T>(k, w, kt);
                    // the candidate's value travels with its key words (one round less)
#pragma unroll
                    for (int kk = 0; kk < K; ++kk) v[T][kk] = CP::load(cin + (size_t)kk * CP::NS * cap, cap, idx);
                    uint64_t kc[NP];
                    key_load<NP>(st.key, idx, kc);
                    bool eq = true;
#pragma unroll
                    for (int q = 0; q < NP; ++q) eq = eq && kc[q] == kt[q];
                    if (eq) {
                        slot[T] = idx;
                        early |= 1u << T;
                    } else {
                        need_probe |= 1u << T;
                    }
                } else {
                    need_probe |= 1u << T;
                }
            }
        }
    });
    if (need_probe) {
        sfor<kOrbit>([&](auto TT) {
            constexpr int T = decltype(TT)::value;
            if constexpr (T > 0) {
                if ((need_probe >> T) & 1) {
                    uint64_t kt[NP];
                    orbit_key<T>(k, w, kt);
                    slot[T] = probe_next<Alg, CP, K>(st, cin, kt, hmix(lk ^ lp_of<T>(w)), n, v[T]);
                    if (slot[T] != kNone) early |= 1u << T;
                }
            }
        });
    }
    sfor<kOrbit>([&](auto TT) {
        constexpr int T = decltype(TT)::value;
        if constexpr (T > 0) {
            if (((valid >> T) & 1) && slot[T] != kNone) {
                exist |= 1u << T;
                if (!((early >> T) & 1)) {
#pragma unroll
                    for (int kk = 0; kk < K; ++kk)
                        v[T][kk] = CP::load(cin + (size_t)kk * CP::NS * cap, cap, slot[T]);
                }
                if (!CP::dead(v[T][0])) alive |= 1u << T;
            }
        }
    });
}

// Replays the window's rotations on one orbit in window order, with the
// operations of the one-rotation kernel; adds the orbit's live-count change at
// each rotation to sdl when `count`.
template <class Alg, class CP, int K>
__device__ __forceinline__ void orbit_replay(const Window<Alg, CP>& w, const uint64_t (&k)[Alg::NP], uint32_t amask,
                                             uint32_t valid, const OrbitFacts& f,
                                             typename CP::val_t (&v)[kOrbit][K], uint32_t& alive, int (&sdl)[kWin],
                                             bool count) {
    constexpr int NP = Alg::NP;
    using val_t = typename CP::val_t;
    sfor<kWin>([&](auto SS) {
        constexpr int S = decltype(SS)::value;
        if (S < w.k) {
            const int before = __popc(alive);
            if ((amask >> S) & 1) {
                const typename CP::fac_t cs = w.cs[S], sn = w.sn[S];
                const uint32_t lin_s = (f.lin >> S) & 1;
                sfor<kOrbit>([&](auto TT) {
                    constexpr int T = decltype(TT)::value;
                    if constexpr (((T >> S) & 1) == 0) {
                        constexpr int U = T | (1 << S);
                        const bool aT = (alive >> T) & 1, aU = (alive >> U) & 1;
                        if (((valid >> U) & 1) && (aT || aU)) {
                            const bool rT = aU && ((f.rok >> T) & 1);
                            const bool rU = aT && ((f.rok >> U) & 1);
                            int sg = 1;  // factor sign of T -> U; that of U -> T is -sg
                            if (rT || rU) {
                                const uint32_t lin_t = lin_s ^ (__popc(w.cross[S] & (uint32_t)T) & 1u);
                                sg = Alg::sign_parts(w.g[S], (f.qm >> (2 * T)) & 3, (f.qm >> (2 * U)) & 3, lin_t);
                            }
                            val_t nT[K], nU[K];
#pragma unroll
                            for (int kk = 0; kk < K; ++kk) {
                                nT[kk] = v[T][kk];
                                nU[kk] = v[U][kk];
                                if (aT) {
                                    nT[kk] = CP::mul(v[T][kk], cs);
                                    if (rT) nT[kk] = CP::add(nT[kk], CP::mul_signed(v[U][kk], sn, -sg));
                                } else if (rT) {
                                    nT[kk] = CP::mul_signed(v[U][kk], sn, -sg);
                                }
                                if (aU) {
                                    nU[kk] = CP::mul(v[U][kk], cs);
                                    if (rU) nU[kk] = CP::add(nU[kk], CP::mul_signed(v[T][kk], sn, sg));
                                } else if (rU) {
                                    nU[kk] = CP::mul_signed(v[T][kk], sn, sg);
                                }
                            }
#pragma unroll
                            for (int kk = 0; kk < K; ++kk) {
                                v[T][kk] = nT[kk];
                                v[U][kk] = nU[kk];
                            }
                            if (rT) alive |= 1u << T;
                            if (rU) alive |= 1u << U;
                        }
                    }
                });
            }
            const bool th = (w.thresh_mask >> S) & 1, xw = (w.xw_mask >> S) & 1;
            if (th || xw) {
                sfor<kOrbit>([&](auto TT) {
                    constexpr int T = decltype(TT)::value;
                    if ((alive >> T) & 1) {
                        bool kill = th && magnitude<CP, K>(v[T]) <= w.mincoeff;
                        if (!kill && xw) {
                            uint64_t kt[NP];
                            orbit_key<T>(k, w, kt);
                            kill = Alg::xweight(kt) > w.max_xw;
                        }
                        if (kill) alive &= ~(1u << T);
                    }
                });
            }
            if (count) sdl[S] += __popc(alive) - before;
        }
    });
}

// Replay facts of the slot itself (member 0) and its lin bits.
template <class Alg, class CP>
__device__ __forceinline__ OrbitFacts self_facts(const uint64_t (&k)[Alg::NP], const Window<Alg, CP>& w,
                                                 uint32_t amask) {
    OrbitFacts f{0u, 0u, 0u};
    member_facts<0, Alg>(k, w.maxdeg, f);
    sfor<kWin>([&](auto SS) {
        constexpr int S = decltype(SS)::value;
        if ((amask >> S) & 1) f.lin |= (uint32_t)Alg::lin(k, w.g[S]) << S;
    });
    return f;
}

template <class CP, int K>
__device__ __forceinline__ void put_value(typename CP::store_t* cout, size_t cap, uint32_t s, bool live,
                                          const typename CP::val_t (&v)[K]) {
    if (live) {
#pragma unroll
        for (int kk = 0; kk < K; ++kk) CP::store(cout + (size_t)kk * CP::NS * cap, cap, s, v[kk]);
    } else {
        CP::store(cout, cap, s, CP::make_dead());
#pragma unroll
        for (int kk = 1; kk < K; ++kk) CP::store(cout + (size_t)kk * CP::NS * cap, cap, s, CP::from_double(0.0));
    }
}

template <class Alg>
__device__ __forceinline__ uint32_t subsets_of(uint32_t amask) {
    uint32_t valid = 0;
    sfor<kOrbit>([&](auto TT) {
        constexpr int T = decltype(TT)::value;
        if ((T & ~amask) == 0) valid |= 1u << T;
    });
    return valid;
}

// Where appends are counted and overflow is reported: the epoch arrays of the
// device-wide engine, or LDS counters of a one-block instance.
struct AppendSink {
    uint32_t* ctr;         // next free slot offset (added to n)
    uint32_t* flag;        // set to 1 when an append does not fit
    uint32_t* flag_epoch;  // receives `e` on overflow (may be null)
    uint32_t e;
};

// Appends the fresh orbit members (absent, live at the end) of the leaders of
// this block iteration: one atomic per block and iteration.
template <class Alg, class CP, int K>
__device__ __forceinline__ void orbit_append_ranked(const Store<CP, K>& st, typename CP::store_t* cout,
                                                    const Window<Alg, CP>& w, const AppendSink& sink, uint32_t n,
                                                    const uint64_t (&k)[Alg::NP], uint64_t lk,
                                                    const typename CP::val_t (&v)[kOrbit][K], uint32_t fresh,
                                                    uint32_t rank, uint32_t tot, uint32_t* sbase) {
    constexpr int NP = Alg::NP;
    const uint32_t napp = __popc(fresh);
    if (tot) {
        if (threadIdx.x == 0) *sbase = atomicAdd(sink.ctr, tot);
        __syncthreads();
        if (napp) {
            // store every fresh member, then claim all their home index entries at
            // once (independent CAS round trips in flight together); only the
            // members whose home entry was taken walk the probe sequence
            const uint64_t at0 = (uint64_t)n + *sbase + rank;
            const size_t cap = st.cap;
            uint64_t prev[kOrbit];
            sfor<kOrbit>([&](auto TT) {
                constexpr int T = decltype(TT)::value;
                if constexpr (T > 0) {
                    if ((fresh >> T) & 1) {
                        const uint64_t at = at0 + __popc(fresh & ((1u << T) - 1u));
                        if (at < cap) {
                            uint64_t kt[NP];
                            orbit_key<T>(k, w, kt);
                            key_store<NP>(st.key, at, kt);
#pragma unroll
                            for (int kk = 0; kk < K; ++kk)
                                CP::store(cout + (size_t)kk * CP::NS * cap, cap, at, v[T][kk]);
                            const uint64_t h = hmix(lk ^ lp_of<T>(w));
                            const uint64_t e = ((h >> 32) << 32) | at;
                            prev[T] = (uint64_t)atomicCAS((unsigned long long*)&st.index[h & st.hmask],
                                                          (unsigned long long)kEmpty, (unsigned long long)e);
                        } else {
                            prev[T] = kEmpty;
                            atomicOr(sink.flag, 1u);
                            if (sink.flag_epoch) *sink.flag_epoch = sink.e;
                        }
                    }
                }
            });
            uint32_t retry = 0;
            sfor<kOrbit>([&](auto TT) {
                constexpr int T = decltype(TT)::value;
                if constexpr (T > 0)
                    if (((fresh >> T) & 1) && prev[T] != kEmpty) retry |= 1u << T;
            });
            if (retry) {
                sfor<kOrbit>([&](auto TT) {
                    constexpr int T = decltype(TT)::value;
                    if constexpr (T > 0) {
                        if ((retry >> T) & 1) {
                            const uint64_t at = at0 + __popc(fresh & ((1u << T) - 1u));
                            index_insert(st.index, st.hmask, hmix(lk ^ lp_of<T>(w)), (uint32_t)at,
                                         &st.ctrl->index_full);
                        }
                    }
                });
            }
        }
    }
}

template <class Alg, class CP, int K>
__device__ __forceinline__ void orbit_append(const Store<CP, K>& st, typename CP::store_t* cout,
                                             const Window<Alg, CP>& w, const AppendSink& sink, uint32_t n,
                                             const uint64_t (&k)[Alg::NP], uint64_t lk,
                                             const typename CP::val_t (&v)[kOrbit][K], uint32_t fresh,
                                             uint32_t* sbase) {
    uint32_t tot;
    const uint32_t rank = block_exclusive_scan(__popc(fresh), &tot);
    orbit_append_ranked<Alg, CP, K>(st, cout, w, sink, n, k, lk, v, fresh, rank, tot, sbase);
}

// A block's appends gathered in LDS over all of its tiles and written out with
// one bump of the epoch's append counter (a counter every block bumps once per
// tile serializes: ~11 ns per bump, thousands of tiles per launch).
template <int NP, class CP, int K>
struct AppendStage {
    static constexpr int kEntry = 8 * NP + (int)sizeof(typename CP::val_t) * K;
    static constexpr int kFit = (24576 / kEntry) & ~63;
    static constexpr int kCap = kFit < 512 ? kFit : 512;
    uint64_t key[kCap][NP];
    typename CP::val_t val[kCap][K];
    uint32_t used;
};

// Writes the staged appends to slots [n + base, n + base + used) and indexes
// them.  Block-uniform.
template <int NP, class CP, int K>
__device__ __forceinline__ void stage_flush(const Store<CP, K>& st, typename CP::store_t* cout,
                                            const AppendSink& sink, uint32_t n, AppendStage<NP, CP, K>& stg,
                                            uint32_t* sbase) {
    const uint32_t used = stg.used;
    if (used == 0) return;
    if (threadIdx.x == 0) *sbase = atomicAdd(sink.ctr, used);
    __syncthreads();
    const uint64_t at0 = (uint64_t)n + *sbase;
    const size_t cap = st.cap;
    for (uint32_t r = threadIdx.x; r < used; r += blockDim.x) {
        const uint64_t at = at0 + r;
        if (at < cap) {
            uint64_t kt[NP];
#pragma unroll
            for (int q = 0; q < NP; ++q) kt[q] = stg.key[r][q];
            key_store<NP>(st.key, at, kt);
#pragma unroll
            for (int kk = 0; kk < K; ++kk) CP::store(cout + (size_t)kk * CP::NS * cap, cap, at, stg.val[r][kk]);
            index_insert(st.index, st.hmask, key_hash<NP>(kt), (uint32_t)at, &st.ctrl->index_full);
        } else {
            atomicOr(sink.flag, 1u);
            if (sink.flag_epoch) *sink.flag_epoch = sink.e;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) stg.used = 0;
    __syncthreads();
}

// Stages the fresh orbit members of this tile's leaders (flushing first when
// they do not fit; a tile with more than the stage holds appends directly).
// Block-uniform.
template <class Alg, class CP, int K>
__device__ __forceinline__ void orbit_stage(const Store<CP, K>& st, typename CP::store_t* cout,
                                            const Window<Alg, CP>& w, const AppendSink& sink, uint32_t n,
                                            const uint64_t (&k)[Alg::NP], uint64_t lk,
                                            const typename CP::val_t (&v)[kOrbit][K], uint32_t fresh,
                                            AppendStage<Alg::NP, CP, K>& stg, uint32_t* sbase) {
    constexpr int NP = Alg::NP;
    using Stage = AppendStage<NP, CP, K>;
    uint32_t tot;
    const uint32_t rank = block_exclusive_scan(__popc(fresh), &tot);
    if (tot == 0) return;
    if (stg.used + tot > (uint32_t)Stage::kCap) {
        stage_flush<NP, CP, K>(st, cout, sink, n, stg, sbase);
        if (tot > (uint32_t)Stage::kCap) {
            orbit_append_ranked<Alg, CP, K>(st, cout, w, sink, n, k, lk, v, fresh, rank, tot, sbase);
            return;
        }
    }
    const uint32_t u0 = stg.used;
    if (fresh) {
        uint32_t r = u0 + rank;
        sfor<kOrbit>([&](auto TT) {
            constexpr int T = decltype(TT)::value;
            if constexpr (T > 0) {
                if ((fresh >> T) & 1) {
                    uint64_t kt[NP];
                    orbit_key<T>(k, w, kt);
#pragma unroll
                    for (int q = 0; q < NP; ++q) stg.key[r][q] = kt[q];
#pragma unroll
                    for (int kk = 0; kk < K; ++kk) stg.val[r][kk] = v[T][kk];
                    ++r;
                }
            }
        });
    }
    __syncthreads();
    if (threadIdx.x == 0) stg.used = u0 + tot;
    __syncthreads();
}

// One tile of blockDim slots [b0, b0 + blockDim) of a window: every slot in
// [0, n) replays its orbit and writes its own value; the leaders append.
// Block-uniform (contains barriers).
// A slot's key and value, loaded ahead by the caller (passed by value).
template <int NP, class CP, int K>
struct SlotLoad {
    uint64_t k[NP];
    typename CP::val_t v[K];
};

// pre: the slot's key and value are already in `ld` (loaded by the caller
// together with other independent loads).
template <class Alg, class CP, int K, bool kStaged = false>
__device__ __forceinline__ void win_tile(const Store<CP, K>& st, const typename CP::store_t* cin,
                                         typename CP::store_t* cout, const Window<Alg, CP>& w, uint32_t n,
                                         uint32_t b0, const AppendSink& sink, uint32_t* sbase, int& dd,
                                         int (&sdl)[kWin], AppendStage<Alg::NP, CP, K>* stg, bool pre,
                                         SlotLoad<Alg::NP, CP, K> ld) {
    constexpr int NP = Alg::NP;
    using val_t = typename CP::val_t;
    const size_t cap = st.cap;
    const uint32_t i = b0 + threadIdx.x;
    uint32_t fresh = 0;
    uint64_t k[NP];
    uint64_t lk = 0;
    val_t v[kOrbit][K];
    if (i < n) {
        if (pre) {
#pragma unroll
            for (int q = 0; q < NP; ++q) k[q] = ld.k[q];
#pragma unroll
            for (int kk = 0; kk < K; ++kk) v[0][kk] = ld.v[kk];
        } else {
            key_load<NP>(st.key, i, k);
#pragma unroll
            for (int kk = 0; kk < K; ++kk) v[0][kk] = CP::load(cin + (size_t)kk * CP::NS * cap, cap, i);
        }
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
        orbit_replay<Alg, CP, K>(w, k, amask, valid, f, v, alive, sdl, leader);
        put_value<CP, K>(cout, cap, i, alive & 1u, v[0]);
        dd += (int)a0 - (int)(alive & 1u);
        if (leader) fresh = alive & ~exist;
    }
    if constexpr (kStaged)
        orbit_stage<Alg, CP, K>(st, cout, w, sink, n, k, lk, v, fresh, *stg, sbase);
    else
        orbit_append<Alg, CP, K>(st, cout, w, sink, n, k, lk, v, fresh, sbase);
}

__device__ __forceinline__ void window_report(const Epochs& ep, uint32_t e, int dd, const int (&sdl)[kWin]) {
    int r[1 + kWin] = {dd, sdl[0], sdl[1], sdl[2], sdl[3]};
    block_sum_ints<1 + kWin>(r);
    if (threadIdx.x == 0) {
        if (r[0]) atomicAdd(&ep.dd[e], r[0]);
        int32_t* sd = sd_stripe(ep, blockIdx.x);
#pragma unroll
        for (int s = 0; s < kWin; ++s)
            if (r[1 + s]) atomicAdd(&sd[(size_t)kWin * e + s], r[1 + s]);
    }
}


// Every slot replays its own orbit (device-wide engine, one window per launch).
template <class Alg, class CP, int K>
__global__ void __launch_bounds__(kBlockWin) k_win_async(Store<CP, K> st, const typename CP::store_t* cin,
                                                         typename CP::store_t* cout, Window<Alg, CP> w, Epochs ep,
                                                         uint32_t e) {
    // block-strided tiles of blockDim slots; in the first kPreBlocks blocks the
    // first tile's key and value are loaded together with the epoch counters
    // (slot < cap, so in bounds whatever n turns out to be; unused beyond n).
    // Small windows, whose launches are a chain of dependent loads, gain a round
    // trip; the bound keeps the unused loads of the blocks past n to ~1 MB.  The
    // values travel to the tile by value: a reference to a local array here made
    // the gfx950 backend emit a private-address check it then rejects ("operand has
    // incorrect register class").
    constexpr uint32_t kPreBlocks = 128;
    const uint32_t begin = blockIdx.x * blockDim.x, step = gridDim.x * blockDim.x;
    const uint32_t i0 = begin + threadIdx.x;
    const bool preload = blockIdx.x < kPreBlocks;
    SlotLoad<Alg::NP, CP, K> ld0{};
    if (preload && i0 < st.cap) {
        key_load<Alg::NP>(st.key, i0, ld0.k);
#pragma unroll
        for (int kk = 0; kk < K; ++kk) ld0.v[kk] = CP::load(cin + (size_t)kk * CP::NS * st.cap, st.cap, i0);
    }
    uint32_t n, dead;
    if (!step_enter(ep, e, n, dead)) return;
    if (begin >= n) return;  // block-uniform: nothing to do, nothing to report
    __shared__ uint32_t sbase;
    __shared__ AppendStage<Alg::NP, CP, K> stg;
    if (threadIdx.x == 0) stg.used = 0;
    __syncthreads();
    const AppendSink sink{&ep.app[e], &ep.flags[0], &ep.flags[1], e};
    int dd = 0;
    int sdl[kWin] = {0, 0, 0, 0};
    for (uint32_t b0 = begin; b0 < n; b0 += step)
        win_tile<Alg, CP, K, true>(st, cin, cout, w, n, b0, sink, &sbase, dd, sdl, &stg, preload && b0 == begin,
                                   ld0);
    stage_flush<Alg::NP, CP, K>(st, cout, sink, n, stg, &sbase);
    window_report(ep, e, dd, sdl);
}

}  // namespace ffx
