// ffx_device.h -- device-side building blocks of the propagation engine.
//
// Term store (per observable, capacity `cap` slots, structure of arrays):
//   key   : one record of NP uint64 words per slot, word p of slot i at
//           key[i*NP + p] (a lookup that verifies a key reads one 8*NP-byte
//           record, not NP scattered words; a wave streaming keys still reads
//           one contiguous 64*8*NP-byte run)
//           Pauli<W>: words 0..W-1 = xory, W..2W-1 = yorz
//           Majorana<W>: words 0..W-1 = occupation
//   coef  : K coefficient planes (K = 1 real, 2 complex); each plane holds
//           CP::NS storage slots per value (1 for FP64/FP32, 2 for FP32C)
//   index : open-addressing hash table, 64-bit entries (tag << 32 | slot)
// An erased term keeps its slot and its index entry; its coefficient holds a
// dead sentinel (a NaN payload that arithmetic never produces) and the slot is
// revived in place if a later partner lands on the same key.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ffx {

constexpr int FFX_MAX_WORDS_DEV = 4;
constexpr uint64_t kEmpty = ~0ull;
constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr uint64_t kDead64 = 0x7FF4DEAD0000DEADull;  // signalling-NaN payload
constexpr uint32_t kDead32 = 0x7FA0DEADu;            // signalling-NaN payload

// Per-observable control block in device memory.
struct Ctrl {
    uint32_t n;           // slots in use (live + erased) at the start of the next step
    uint32_t n_app;       // slots appended by the running step (step-by-step kernels)
    uint32_t n_dead;      // erased slots among [0, n)
    uint32_t done;        // finished blocks of the running kernel
    int32_t dead_delta;   // erased-count change of the running step
    uint32_t overflow;    // an append did not fit into cap
    uint32_t dup;         // duplicate keys found at upload
    uint32_t index_full;  // a probe or insert visited every index entry (internal error)
    unsigned long long updates;  // term-gate updates so far (sum of live terms entering a rotation)
    unsigned long long outs;     // sum of live terms leaving a rotation
    unsigned long long peak;     // largest live count after a rotation
    unsigned long long n_rot;    // rotations applied
    unsigned long long n_compact;  // compactions run
};

// Index hash = hmix(L(key)).  L is GF(2)-linear, L(a ^ b) = L(a) ^ L(b): word p
// contributes rotl(w, r1) ^ rotl(w, r2) ^ rotl(w, r3), an invertible map (odd
// number of rotations), with rotation triples that are not translates of one
// another, so keys differing in one word, or in one bit of each of two words,
// never share L.  Linearity lets a window derive the hash of every orbit
// string from the slot's own L and a per-window table (no per-string key
// hashing).  hmix is a bijective xorshift-multiply finalizer.
constexpr int kHashRot[8][3] = {{10, 21, 61}, {4, 26, 42}, {5, 35, 53}, {7, 24, 38},
                                {4, 33, 59}, {6, 16, 36}, {4, 28, 53}, {8, 37, 61}};

__host__ __device__ __forceinline__ uint64_t rotl64(uint64_t x, int r) { return (x << r) | (x >> (64 - r)); }

template <int NP>
__host__ __device__ __forceinline__ uint64_t lin_hash(const uint64_t* k) {
    uint64_t h = 0;
#pragma unroll
    for (int p = 0; p < NP; ++p)
        h ^= rotl64(k[p], kHashRot[p][0]) ^ rotl64(k[p], kHashRot[p][1]) ^ rotl64(k[p], kHashRot[p][2]);
    return h;
}

__host__ __device__ __forceinline__ uint64_t hmix(uint64_t h) {
    h ^= h >> 32;
    h *= 0xD6E8FEB86659FD93ull;
    h ^= h >> 32;
    h *= 0xD6E8FEB86659FD93ull;
    h ^= h >> 32;
    return h;
}

template <int NP>
__device__ __forceinline__ uint64_t key_hash(const uint64_t (&k)[NP]) {
    return hmix(lin_hash<NP>(k));
}

__device__ __forceinline__ int popc(uint64_t v) { return __popcll(v); }

// Key record of slot i.  Records of an even word count are 16-byte aligned
// (every key array starts 16-byte aligned) and move as 16-byte accesses.
template <int NP>
__device__ __forceinline__ void key_load(const uint64_t* __restrict__ key, size_t i, uint64_t (&k)[NP]) {
    const uint64_t* p = key + i * NP;
    if constexpr ((NP & 1) == 0) {
#pragma unroll
        for (int q = 0; q < NP; q += 2) {
            const ulonglong2 v = *reinterpret_cast<const ulonglong2*>(p + q);
            k[q] = v.x;
            k[q + 1] = v.y;
        }
    } else {
#pragma unroll
        for (int q = 0; q < NP; ++q) k[q] = p[q];
    }
}
template <int NP>
__device__ __forceinline__ void key_store(uint64_t* __restrict__ key, size_t i, const uint64_t (&k)[NP]) {
    uint64_t* p = key + i * NP;
    if constexpr ((NP & 1) == 0) {
#pragma unroll
        for (int q = 0; q < NP; q += 2) *reinterpret_cast<ulonglong2*>(p + q) = make_ulonglong2(k[q], k[q + 1]);
    } else {
#pragma unroll
        for (int q = 0; q < NP; ++q) p[q] = k[q];
    }
}
template <int NP>
__device__ __forceinline__ void key_copy(uint64_t* __restrict__ dst, size_t j, const uint64_t* __restrict__ src,
                                         size_t i) {
    uint64_t k[NP];
    key_load<NP>(src, i, k);
    key_store<NP>(dst, j, k);
}

// ---------------------------------------------------------------------------
// Coefficient policies
// ---------------------------------------------------------------------------

// IEEE double.  Operation order and rounding match the reference's
// std::complex<double> arithmetic term by term (kept: c*cos; partner: +-(c*sin);
// merge: a + b).  Requires -ffp-contract=off.
struct F64 {
    using store_t = double;
    using val_t = double;
    using fac_t = double;
    static constexpr int NS = 1;
    __device__ static val_t load(const store_t* p, size_t plane_stride, size_t i) { return p[i]; }
    __device__ static void store(store_t* p, size_t plane_stride, size_t i, val_t v) { p[i] = v; }
    __device__ static bool dead(val_t v) { return __double_as_longlong(v) == (long long)kDead64; }
    __device__ static val_t make_dead() { return __longlong_as_double((long long)kDead64); }
    __device__ static val_t mul(val_t a, fac_t f) { return a * f; }
    __device__ static val_t mul_signed(val_t a, fac_t f, int s) {
        val_t p = a * f;
        return s > 0 ? p : -p;
    }
    __device__ static val_t add(val_t a, val_t b) { return a + b; }
    __device__ static double to_double(val_t v) { return v; }
    __device__ static val_t from_double(double d) { return d; }
    __device__ static val_t neg(val_t v) { return -v; }
};

// IEEE single storage and arithmetic.
struct F32 {
    using store_t = float;
    using val_t = float;
    using fac_t = float;
    static constexpr int NS = 1;
    __device__ static val_t load(const store_t* p, size_t plane_stride, size_t i) { return p[i]; }
    __device__ static void store(store_t* p, size_t plane_stride, size_t i, val_t v) { p[i] = v; }
    __device__ static bool dead(val_t v) { return __float_as_uint(v) == kDead32; }
    __device__ static val_t make_dead() { return __uint_as_float(kDead32); }
    __device__ static val_t mul(val_t a, fac_t f) { return a * f; }
    __device__ static val_t mul_signed(val_t a, fac_t f, int s) {
        val_t p = a * f;
        return s > 0 ? p : -p;
    }
    __device__ static val_t add(val_t a, val_t b) { return a + b; }
    __device__ static double to_double(val_t v) { return (double)v; }
    __device__ static val_t from_double(double d) { return (float)d; }
    __device__ static val_t neg(val_t v) { return -v; }
};

// Float storage with a stored compensation term: value = hi + lo (float-float).
// Products use an FMA-exact TwoProduct, sums a TwoSum (Knuth), so every merge
// carries its rounding error forward instead of dropping it.
struct F32C {
    struct val_t {
        float hi, lo;
    };
    struct fac_t {
        float hi, lo;
    };
    using store_t = float;
    static constexpr int NS = 2;
    __device__ static val_t load(const store_t* p, size_t plane_stride, size_t i) {
        return val_t{p[i], p[plane_stride + i]};
    }
    __device__ static void store(store_t* p, size_t plane_stride, size_t i, val_t v) {
        p[i] = v.hi;
        p[plane_stride + i] = v.lo;
    }
    __device__ static bool dead(val_t v) { return __float_as_uint(v.hi) == kDead32; }
    __device__ static val_t make_dead() { return val_t{__uint_as_float(kDead32), 0.f}; }
    __device__ static val_t fast_two_sum(float a, float b) {
        float s = a + b;
        float e = b - (s - a);
        return val_t{s, e};
    }
    __device__ static val_t mul(val_t a, fac_t f) {
        float p = a.hi * f.hi;
        float e = fmaf(a.hi, f.hi, -p);
        e = fmaf(a.hi, f.lo, e);
        e = fmaf(a.lo, f.hi, e);
        return fast_two_sum(p, e);
    }
    __device__ static val_t mul_signed(val_t a, fac_t f, int s) {
        val_t p = mul(a, f);
        return s > 0 ? p : val_t{-p.hi, -p.lo};
    }
    __device__ static val_t add(val_t a, val_t b) {
        float s = a.hi + b.hi;
        float bb = s - a.hi;
        float e = (a.hi - (s - bb)) + (b.hi - bb);
        e += a.lo + b.lo;
        return fast_two_sum(s, e);
    }
    __device__ static double to_double(val_t v) { return (double)v.hi + (double)v.lo; }
    __device__ static val_t from_double(double d) {
        float h = (float)d;
        return val_t{h, (float)(d - (double)h)};
    }
    __device__ static val_t neg(val_t v) { return val_t{-v.hi, -v.lo}; }
};

// ---------------------------------------------------------------------------
// Algebra policies: anticommutation, partner, sign, degree
// ---------------------------------------------------------------------------

// Pauli strings with W words per mask (src/pauli/algebra.h:30-33).
template <int W_>
struct PauliAlg {
    static constexpr int W = W_;
    static constexpr int NP = 2 * W;
    static constexpr bool kPauli = true;
    struct Gate {
        uint64_t p[NP];  // axis: x words then z words
        int pxz;         // popcount(P.x & P.z)
    };
    // PauliString::commutes (src/pauli/algebra.h:209-211) negated: the symplectic
    // form popcount((x_Q & z_P) ^ (z_Q & x_P)) is odd.
    __device__ static bool anticommutes(const uint64_t (&k)[NP], const Gate& g) {
        int s = 0;
#pragma unroll
        for (int w = 0; w < W; ++w) s += popc((k[w] & g.p[W + w]) ^ (k[W + w] & g.p[w]));
        return s & 1;
    }
    // degree_total (src/pauli/algebra.h:178)
    __device__ static int degree(const uint64_t (&k)[NP]) {
        int d = 0;
#pragma unroll
        for (int w = 0; w < W; ++w) d += popc(k[w] | k[W + w]);
        return d;
    }
    // degree_x + degree_y (src/pauli/truncate.h:41-46)
    __device__ static int xweight(const uint64_t (&k)[NP]) {
        int d = 0;
#pragma unroll
        for (int w = 0; w < W; ++w) d += popc(k[w]);
        return d;
    }
    // Sign of the contribution of term q to its partner r = P*q: the reference
    // multiplies c_q by i*sin(theta)*i^jpow with
    // jpow = |P.x&P.z| + |q.x&q.z| + 2|P.z&q.x| - |r.x&r.z| (src/pauli/algebra.h:227-233,
    // src/pauli/propagate.h:79-84); jpow is odd for anticommuting q, so
    // i^(jpow+1) = +1 when jpow = 3 mod 4 and -1 when jpow = 1 mod 4.
    __device__ static int sign(const uint64_t (&q)[NP], const uint64_t (&r)[NP], const Gate& g) {
        int j = g.pxz;
#pragma unroll
        for (int w = 0; w < W; ++w) j += popc(q[w] & q[W + w]) + 2 * popc(g.p[W + w] & q[w]) - popc(r[w] & r[W + w]);
        return ((j & 3) == 3) ? 1 : -1;
    }
    // sign(q, r, g) factored for the window replay: it depends on q and r only
    // through quad(q), quad(r) (|x & z| mod 4) and lin(q, g) (|P.z & q.x| mod 2,
    // linear in q).
    __device__ static int quad(const uint64_t (&k)[NP]) {
        int s = 0;
#pragma unroll
        for (int w = 0; w < W; ++w) s += popc(k[w] & k[W + w]);
        return s & 3;
    }
    __host__ __device__ static int lin(const uint64_t* k, const Gate& g) {
        int s = 0;
        for (int w = 0; w < W; ++w) s += __builtin_popcountll(g.p[W + w] & k[w]);
        return s & 1;
    }
    __device__ static int sign_parts(const Gate& g, int quad_q, int quad_r, int lin_q) {
        const int j = g.pxz + quad_q + 2 * lin_q - quad_r;
        return ((j & 3) == 3) ? 1 : -1;
    }
};

// Majorana strings with W occupation words (src/majorana_algebra.h:30-36).
template <int W_>
struct MajoranaAlg {
    static constexpr int W = W_;
    static constexpr int NP = W;
    static constexpr bool kPauli = false;
    struct Gate {
        uint64_t p[NP];   // axis occupation words
        uint64_t sm[NP];  // bit t = parity of #{g in axis : g > t}
        int deg;          // |axis|
        int r;            // MROT::_r
    };
    __device__ static int degree(const uint64_t (&k)[NP]) {
        int d = 0;
#pragma unroll
        for (int w = 0; w < W; ++w) d += popc(k[w]);
        return d;
    }
    __device__ static int xweight(const uint64_t (&k)[NP]) { return 0; }
    // MajoranaString::commutes (src/majorana_algebra.h:79-84) negated:
    // |a||b| - |a&b| is odd.
    __device__ static bool anticommutes(const uint64_t (&k)[NP], const Gate& g) {
        int o = 0;
#pragma unroll
        for (int w = 0; w < W; ++w) o += popc(k[w] & g.p[w]);
        if (g.deg & 1) return ((degree(k) - o) & 1) != 0;
        return (o & 1) != 0;
    }
    // Sign of the contribution of q to r = m_S * q on normalized coefficients
    // d = c / i^{r(string)} (r(string) = 1 iff degree mod 4 in {2,3}).  The
    // reference multiplies c_q by i*i^{r(S)}*sin(theta)*(-1)^{merge_parity(q, S)}
    // (src/majorana_gates.h:86-99, src/majorana_algebra.h:441-444,
    // src/bits_utils.h:14-38); on d this is (-1)^{merge_parity} * i^e * sin with
    // e = r(q) + 1 + r(S) - r(r_), always even for anticommuting q.
    // merge_parity(q, S) = parity of #{(t in q, g in S): t < g} = parity(popcount(q & sm)).
    __device__ static int sign(const uint64_t (&q)[NP], const uint64_t (&r)[NP], const Gate& g) {
        int mp = 0;
#pragma unroll
        for (int w = 0; w < W; ++w) mp += popc(q[w] & g.sm[w]);
        const int rq = (degree(q) >> 1) & 1;
        const int rr = (degree(r) >> 1) & 1;
        const int e = rq + 1 + g.r - rr;
        int s = ((e & 3) == 0) ? 1 : -1;
        return (mp & 1) ? -s : s;
    }
    // sign(q, r, g) factored for the window replay: quad = r(string) =
    // (degree >> 1) & 1, lin(q, g) = merge parity |q & sm| mod 2 (linear in q).
    __device__ static int quad(const uint64_t (&k)[NP]) { return (degree(k) >> 1) & 1; }
    __host__ __device__ static int lin(const uint64_t* k, const Gate& g) {
        int s = 0;
        for (int w = 0; w < W; ++w) s += __builtin_popcountll(k[w] & g.sm[w]);
        return s & 1;
    }
    __device__ static int sign_parts(const Gate& g, int quad_q, int quad_r, int lin_q) {
        const int e = quad_q + 1 + g.r - quad_r;
        const int s = ((e & 3) == 0) ? 1 : -1;
        return lin_q ? -s : s;
    }
};

// ---------------------------------------------------------------------------
// Hash index
// ---------------------------------------------------------------------------

// Find slot index of key r among slots [0, n).  Entries appended by the running
// kernel have slot >= n and are skipped (their key words may not be visible yet;
// a running step never looks up a key it appends itself).
// The index always keeps at least half of its entries empty, so a probe ends
// long before it has visited hmask + 1 entries; the bound only turns an internal
// error into a reported one instead of an endless loop.
template <int NP>
__device__ __forceinline__ uint32_t probe(const uint64_t* __restrict__ key, size_t cap,
                                          const uint64_t* __restrict__ index, uint64_t hmask,
                                          const uint64_t (&r)[NP], uint32_t n, uint32_t* ierr = nullptr) {
    const uint64_t h = key_hash<NP>(r);
    const uint32_t tag = (uint32_t)(h >> 32);
    uint64_t p = h & hmask;
    for (uint64_t step = 0;; ++step) {
        if (step > hmask) {
            if (ierr) atomicOr(ierr, 1u);
            return kNone;
        }
        const uint64_t e = index[p];
        if (e == kEmpty) return kNone;
        const uint32_t idx = (uint32_t)e;
        if ((uint32_t)(e >> 32) == tag && idx < n) {
            uint64_t c[NP];
            key_load<NP>(key, idx, c);
            bool eq = true;
#pragma unroll
            for (int q = 0; q < NP; ++q) eq = eq && c[q] == r[q];
            if (eq) return idx;
        }
        p = (p + 1) & hmask;
    }
}

__device__ __forceinline__ void index_insert(uint64_t* index, uint64_t hmask, uint64_t h, uint32_t idx,
                                             uint32_t* ierr = nullptr) {
    const uint64_t e = ((h >> 32) << 32) | (uint64_t)idx;
    uint64_t p = h & hmask;
    for (uint64_t step = 0; step <= hmask; ++step) {
        if (atomicCAS((unsigned long long*)&index[p], (unsigned long long)kEmpty, (unsigned long long)e) ==
            (unsigned long long)kEmpty)
            return;
        p = (p + 1) & hmask;
    }
    if (ierr) atomicOr(ierr, 1u);
}

__device__ __forceinline__ uint32_t atomic_load_u32(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int32_t atomic_load_i32(const int32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

}  // namespace ffx
