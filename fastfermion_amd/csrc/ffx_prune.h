// ffx_prune.h -- device side of the pruned propagation mode and of the learned
// term pruner.
//
// Pruned mode (opt-in, never part of the reference semantics):
//   * Clifford rotations: a rotation with |cos theta| <= tol keeps only its sin
//     branch.  Every anticommuting term moves to the partner key P*q with the
//     factor sign*sin (the reference's kept branch c*cos, e.g. 6.1e-17*c for
//     theta = -pi/2, is dropped).  q -> P*q on anticommuting keys is a bijection,
//     so a run of such rotations is one in-place key remap per slot
//     (k_crot_async) followed by an index rebuild; no lookups, no appends.
//   * Learned pruning: every few rotations a small MLP scores each live term
//     from 64 features of its string; terms whose predicted contribution to the
//     final expectation is below a threshold are erased (k_prune_async).
//
// Contribution logs (training data for the pruner): the propagation is a
// linear map on coefficients once its truncation pattern is fixed, so the
// final value <b|O|b> = sum_q c_q(e) v_q(e) at every step e, where v(e) is
// obtained from the readout vector by the transposed rotations applied in the
// opposite order (k_adjoint).  v_q(e) is the influence of term q at step e;
// k_log_extract writes (features, log10|v|) rows and sum_q c_q v_q per logged
// step, which must equal the final expectation (an exact invariant).
//
// MLP: 64 -> 128 -> 128 -> 1, ReLU, bf16 weights and activations, fp32
// accumulation on MFMA (v_mfma_f32_16x16x32_bf16).  Layers are evaluated in
// transposed form (hidden x terms), so the accumulator layout of one layer is
// the B-operand layout of the next and no LDS round trip is needed between them.
#pragma once

#include "ffx_async.h"

namespace ffx {

// ---------------------------------------------------------------------------
// Clifford rotations
// ---------------------------------------------------------------------------
template <class Alg, class CP>
struct CRot {
    typename Alg::Gate g;
    typename CP::fac_t sn;  // sin(theta) (|cos theta| <= tol)
    int thresh;             // threshold after this rotation
    int xw;                 // X/Y-weight cap after this rotation
    int next_th;            // first rotation of the run at or after this one with a threshold (rn: none)
    int next_xw;            // same for the X/Y-weight cap
};

constexpr int kCRotLds = 256;  // Clifford-rotation records staged in LDS per pass

// A run of nr Clifford rotations on every slot (live and erased: the remap is a
// bijection on keys, so keys stay distinct), coefficients in place; the index
// is cleared and rebuilt by k_index_build_async afterwards.  The records of a
// run are read from LDS (one staging pass for runs of up to kCRotLds rotations).
//
// commuting: the axes of the run commute pairwise.  Then XORing one axis into a
// key does not change its commutation with another, so the rotations acting on
// a term are fixed by its key at the start of the run: they are found with
// independent tests (no dependent chain through the records), only they update
// the term, and between two of them coefficient and key are constant, so the
// truncation points there reduce to one test each (next_th / next_xw).  The
// result is the one of the rotation-by-rotation loop.
template <class Alg, class CP, int K>
__device__ __forceinline__ void crot_commuting(const CRot<Alg, CP>* rr, int nr, double mincoeff, int maxdeg,
                                               int max_xw, uint64_t (&k)[Alg::NP], typename CP::val_t (&c)[K],
                                               bool alive, int& kill_at) {
    constexpr int NP = Alg::NP;
    uint64_t k0[NP];
#pragma unroll
    for (int q = 0; q < NP; ++q) k0[q] = k[q];
    kill_at = nr;
    // truncation points in [a, b) with the current coefficient and key
    auto check = [&](int a, int b) {
        if (!alive || kill_at < nr || a >= b) return;
        const int th = rr[a].next_th, xp = rr[a].next_xw;
        int at = nr;
        if (th < b && magnitude<CP, K>(c) <= mincoeff) at = th;
        if (xp < b && xp < at && Alg::xweight(k) > max_xw) at = xp;
        kill_at = at;
    };
    int prev = 0;
    for (int c0 = 0; c0 < nr; c0 += 64) {
        const int cn = min(64, nr - c0);
        uint64_t m = 0;
        for (int j = 0; j < cn; ++j)
            if (Alg::anticommutes(k0, rr[c0 + j].g)) m |= 1ull << j;
        while (m) {
            const int t = c0 + __ffsll((long long)m) - 1;
            m &= m - 1;
            check(prev, t);
            const CRot<Alg, CP>& cr = rr[t];
            uint64_t r[NP];
#pragma unroll
            for (int q = 0; q < NP; ++q) r[q] = k[q] ^ cr.g.p[q];
            if (alive && kill_at == nr) {
                if (Alg::degree(r) > maxdeg) {
                    kill_at = t;
                } else {
                    const int sg = Alg::sign(k, r, cr.g);
#pragma unroll
                    for (int kk = 0; kk < K; ++kk) c[kk] = CP::mul_signed(c[kk], cr.sn, sg);
                }
            }
#pragma unroll
            for (int q = 0; q < NP; ++q) k[q] = r[q];
            prev = t;
        }
    }
    check(prev, nr);
}

template <class Alg, class CP, int K>
__global__ void __launch_bounds__(kBlock) k_crot_async(Store<CP, K> st, const CRot<Alg, CP>* __restrict__ rots,
                                                       int nr, int commuting, double mincoeff, int maxdeg, int max_xw,
                                                       Epochs ep, uint32_t e) {
    constexpr int NP = Alg::NP;
    using val_t = typename CP::val_t;
    __shared__ CRot<Alg, CP> srot[kCRotLds];
    uint32_t n, dead;
    if (!step_enter(ep, e, n, dead)) return;
    const size_t cap = st.cap;
    const bool staged = nr <= kCRotLds;
    if (staged) {
        for (int t = threadIdx.x; t < nr; t += blockDim.x) srot[t] = rots[t];
        __syncthreads();
    }
    const CRot<Alg, CP>* rr = staged ? srot : rots;
    int dd = 0;
    unsigned long long surv = 0;  // sum over live terms of the rotations they survive
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        uint64_t k[NP];
        key_load<NP>(st.key, i, k);
        val_t c[K];
#pragma unroll
        for (int kk = 0; kk < K; ++kk) c[kk] = st.ld(st.coef, kk, i);
        const bool alive = !CP::dead(c[0]);
        bool kill = false;
        if (commuting) {
            int kill_at;
            crot_commuting<Alg, CP, K>(rr, nr, mincoeff, maxdeg, max_xw, k, c, alive, kill_at);
            kill = kill_at < nr;
            if (alive) surv += (unsigned long long)kill_at;
        }
        for (int t = 0; t < (commuting ? 0 : nr); ++t) {
            const CRot<Alg, CP>& cr = rr[t];
            if (Alg::anticommutes(k, cr.g)) {
                uint64_t r[NP];
#pragma unroll
                for (int q = 0; q < NP; ++q) r[q] = k[q] ^ cr.g.p[q];
                if (alive && !kill) {
                    if (Alg::degree(r) > maxdeg) {
                        kill = true;
                    } else {
                        const int s = Alg::sign(k, r, cr.g);
#pragma unroll
                        for (int kk = 0; kk < K; ++kk) c[kk] = CP::mul_signed(c[kk], cr.sn, s);
                    }
                }
#pragma unroll
                for (int q = 0; q < NP; ++q) k[q] = r[q];
            }
            if (alive && !kill) {
                if (cr.thresh && magnitude<CP, K>(c) <= mincoeff) kill = true;
                if (cr.xw && Alg::xweight(k) > max_xw) kill = true;
                if (!kill) ++surv;
            }
        }
        key_store<NP>(st.key, i, k);
        if (alive) {
            if (kill) {
                st.st(st.coef, 0, i, CP::make_dead());
#pragma unroll
                for (int kk = 1; kk < K; ++kk) st.st(st.coef, kk, i, CP::from_double(0.0));
                ++dd;
            } else {
#pragma unroll
                for (int kk = 0; kk < K; ++kk) st.st(st.coef, kk, i, c[kk]);
            }
        }
    }
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i <= st.hmask; i += stride) st.index[i] = kEmpty;
    dd = block_sum_int(dd);
    // rotation-survival count of the block, 64-bit, into the epoch's sd words (read by the host)
    for (int o = 32; o > 0; o >>= 1) surv += __shfl_down(surv, o, 64);
    if (__lane_id() == 0 && surv)
        atomicAdd((unsigned long long*)(sd_stripe(ep, blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) +
                                        (size_t)kWin * e),
                  surv);
    if (threadIdx.x == 0 && dd) atomicAdd(&ep.dd[e], dd);
}

// ---------------------------------------------------------------------------
// bf16 helpers (round to nearest even; finite inputs)
// ---------------------------------------------------------------------------
typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

__host__ __device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u;
    __builtin_memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__host__ __device__ __forceinline__ float bf2f(uint16_t b) {
    const uint32_t u = (uint32_t)b << 16;
    float f;
    __builtin_memcpy(&f, &u, 4);
    return f;
}

// ---------------------------------------------------------------------------
// Term features (64 per term)
// ---------------------------------------------------------------------------
constexpr int kFeat = 64;
constexpr int kHid = 128;
constexpr int kNParams = kFeat * kHid + kHid + kHid * kHid + kHid + kHid + 1;  // 24961
// parameter layout (fp32 master copy): W1 [64][128] (in x out), b1, W2 [128][128], b2, w3, b3
constexpr int kOffW1 = 0, kOffB1 = kFeat * kHid, kOffW2 = kOffB1 + kHid, kOffB2 = kOffW2 + kHid * kHid,
              kOffW3 = kOffB2 + kHid, kOffB3 = kOffW3 + kHid;

// Site geometry of every qubit (Pauli) or Majorana index, and the observable's site.
struct PruneGeo {
    uint8_t row[256];
    uint8_t col[256];
    int obs_row, obs_col;
};

// Step context of the features: rotation layers still to apply, mean |theta| of
// the branching rotations still to apply, fraction of the circuit done.
struct FeatCtx {
    float rem_layers;
    float mean_theta;
    float frac_done;
};

// Add v to the bf16 value at p (LDS; the values stay exact small multiples of 1/4).
__device__ __forceinline__ void bf_add(uint16_t* p, float v) { *p = f2bf(bf2f(*p) + v); }

// Features of one term into a zeroed row of kFeat bf16 values:
//  0 degree/8; 1..3 Pauli #X/8, #Y/8, #Z/8 (Majorana: paired modes/8, unpaired/8, 0);
//  4 remaining layers/32; 5 mean theta; 6 log10|c|/8; 7 fraction done;
//  8, 9 row and column span/16; 10..25 row histogram/4; 26..41 column histogram/4;
//  42..57 row histogram of off-diagonal sites (Pauli X or Y, Majorana unpaired)/4;
//  58, 59 mean row/16, mean column/16; 60 diagonal string (readout can be nonzero);
//  61 distance of the support to the observable site/16; 62 0; 63 constant 1.
template <class Alg>
__device__ void term_features(const uint64_t (&k)[Alg::NP], float logc, const PruneGeo& geo, const FeatCtx& fc,
                              uint16_t* f) {
    constexpr int W = Alg::W;
    int deg = 0, n1 = 0, n2 = 0, n3 = 0;
    int rmin = 255, rmax = 0, cmin = 255, cmax = 0, rsum = 0, csum = 0, dmin = 1 << 20;
    auto site = [&](int q, bool offdiag) {
        const int r = geo.row[q], c = geo.col[q];
        const int rb = r < 15 ? r : 15, cb = c < 15 ? c : 15;
        bf_add(&f[10 + rb], 0.25f);
        bf_add(&f[26 + cb], 0.25f);
        if (offdiag) bf_add(&f[42 + rb], 0.25f);
        rmin = min(rmin, r);
        rmax = max(rmax, r);
        cmin = min(cmin, c);
        cmax = max(cmax, c);
        rsum += r;
        csum += c;
        dmin = min(dmin, abs(r - geo.obs_row) + abs(c - geo.obs_col));
        ++deg;
    };
    bool diag = true;
    if constexpr (Alg::kPauli) {
#pragma unroll
        for (int w = 0; w < W; ++w) {
            const uint64_t x = k[w], z = k[W + w];
            n1 += popc(x & ~z);
            n2 += popc(x & z);
            n3 += popc(z & ~x);
            if (x) diag = false;
            uint64_t b = x | z;
            while (b) {
                const int t = __builtin_ctzll(b);
                b &= b - 1;
                site(64 * w + t, (x >> t) & 1ull);
            }
        }
    } else {
#pragma unroll
        for (int w = 0; w < W; ++w) {
            const uint64_t m = k[w];
            const uint64_t pairs = m & (m >> 1) & 0x5555555555555555ull;  // both 2j and 2j+1
            const uint64_t paired = pairs | (pairs << 1);
            n1 += popc(pairs);
            n2 += popc(m & ~paired);
            if (m & ~paired) diag = false;
            uint64_t b = m;
            while (b) {
                const int t = __builtin_ctzll(b);
                b &= b - 1;
                site(64 * w + t, !((paired >> t) & 1ull));
            }
        }
    }
    f[0] = f2bf(deg * 0.125f);
    f[1] = f2bf(n1 * 0.125f);
    f[2] = f2bf(n2 * 0.125f);
    f[3] = f2bf(n3 * 0.125f);
    f[4] = f2bf(fc.rem_layers * (1.f / 32.f));
    f[5] = f2bf(fc.mean_theta);
    f[6] = f2bf(fmaxf(logc, -40.f) * 0.125f);
    f[7] = f2bf(fc.frac_done);
    if (deg > 0) {
        f[8] = f2bf((rmax - rmin) * (1.f / 16.f));
        f[9] = f2bf((cmax - cmin) * (1.f / 16.f));
        f[58] = f2bf((float)rsum / (float)deg * (1.f / 16.f));
        f[59] = f2bf((float)csum / (float)deg * (1.f / 16.f));
        f[61] = f2bf(dmin * (1.f / 16.f));
    }
    f[60] = f2bf(diag ? 1.f : 0.f);
    f[63] = f2bf(1.f);
}

// ---------------------------------------------------------------------------
// MLP on MFMA
// ---------------------------------------------------------------------------
// Device copies of the weights: bf16 W1^T [128][64], W2^T [128][128] (out x in),
// W2 [128][128] (in x out, backward), w3 rounded to bf16 (kept as float), fp32 biases.
struct MlpDev {
    const uint16_t* w1t;
    const uint16_t* w2t;
    const uint16_t* w2;
    const float* w3r;
    const float* b1;
    const float* b2;
    const float* b3;
};

// LDS row strides (bf16 elements), padded against bank conflicts
constexpr int kS1 = kFeat + 8;  // W1^T rows and feature rows
constexpr int kS2 = kHid + 8;   // W2^T / W2 rows

__device__ __forceinline__ v4s ld4(const uint16_t* p) { return *(const v4s*)p; }

typedef short v8s __attribute__((ext_vector_type(8)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

// One v_mfma_f32_16x16x32_bf16 (gfx950) over a 32-long slice of the
// contraction.  Operand slots 0-3 of lane group g (= lane / 16) carry slice
// positions 4g .. 4g+3 (the `lo` halves), slots 4-7 positions 16+4g .. 16+4g+3
// (`hi`).  A and B use the same slot map, so the instruction sums over all 32
// positions, whatever order the hardware adds them in.  In this map the two
// 16x16 accumulator tiles of a layer's outputs 32s .. 32s+15 and 32s+16 ..
// 32s+31 (output 16t + 4g + i in slot i of lane group g) are directly the B
// operand of slice s of the next layer.
__device__ __forceinline__ v4f mfma32(v4s a_lo, v4s a_hi, v4s b_lo, v4s b_hi, v4f acc) {
    const v8s a = __builtin_shufflevector(a_lo, a_hi, 0, 1, 2, 3, 4, 5, 6, 7);
    const v8s b = __builtin_shufflevector(b_lo, b_hi, 0, 1, 2, 3, 4, 5, 6, 7);
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(v8bf, a), __builtin_bit_cast(v8bf, b), acc, 0,
                                                   0, 0);
}

__device__ __forceinline__ v4s pack4(float a, float b, float c, float d) {
    v4s r;
    r[0] = (short)f2bf(a);
    r[1] = (short)f2bf(b);
    r[2] = (short)f2bf(c);
    r[3] = (short)f2bf(d);
    return r;
}

// Forward pass of 16 terms (one wave).  X: 16 feature rows (stride kS1, LDS);
// sw1t / sw2t: weights in LDS (strides kS1 / kS2); sb1, sb2, sw3: LDS floats.
// Lane l works on term l%16 and hidden units 16*mt + 4*(l/16) + i.
// Returns y of term l%16 in every lane; z1, z2 hold the pre-activations.
__device__ __forceinline__ float mlp16_forward(const uint16_t* sw1t, const uint16_t* sw2t, const float* sb1,
                                               const float* sb2, const float* sw3, float b3, const uint16_t* X,
                                               v4f (&z1)[8], v4f (&z2)[8]) {
    const int l = __lane_id();
    const int r = l & 15, g = l >> 4;
    // layer 1: z1^T (128 x 16) = W1^T (128 x 64) * X^T (64 x 16)
    v4s xb[4];
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) xb[ks] = ld4(X + r * kS1 + 16 * ks + 4 * g);
#pragma unroll
    for (int mt = 0; mt < 8; ++mt) {
        v4f acc = {0.f, 0.f, 0.f, 0.f};
        const uint16_t* w = sw1t + (16 * mt + r) * kS1;
#pragma unroll
        for (int s = 0; s < 2; ++s)
            acc = mfma32(ld4(w + 32 * s + 4 * g), ld4(w + 32 * s + 16 + 4 * g), xb[2 * s], xb[2 * s + 1], acc);
#pragma unroll
        for (int i = 0; i < 4; ++i) acc[i] += sb1[16 * mt + 4 * g + i];
        z1[mt] = acc;
    }
    v4s h1[8];
#pragma unroll
    for (int mt = 0; mt < 8; ++mt)
        h1[mt] = pack4(fmaxf(z1[mt][0], 0.f), fmaxf(z1[mt][1], 0.f), fmaxf(z1[mt][2], 0.f), fmaxf(z1[mt][3], 0.f));
    // layer 2: z2^T (128 x 16) = W2^T (128 x 128) * h1^T (128 x 16)
    float y = 0.f;
#pragma unroll
    for (int mt = 0; mt < 8; ++mt) {
        v4f acc = {0.f, 0.f, 0.f, 0.f};
        const uint16_t* w = sw2t + (16 * mt + r) * kS2;
#pragma unroll
        for (int s = 0; s < 4; ++s)
            acc = mfma32(ld4(w + 32 * s + 4 * g), ld4(w + 32 * s + 16 + 4 * g), h1[2 * s], h1[2 * s + 1], acc);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int o = 16 * mt + 4 * g + i;
            acc[i] += sb2[o];
            y += fmaxf(acc[i], 0.f) * sw3[o];
        }
        z2[mt] = acc;
    }
    // layer 3: sum over the 4 lane groups that hold term r
    y += __shfl_xor(y, 16, 64);
    y += __shfl_xor(y, 32, 64);
    return y + b3;
}

// Block-cooperative copy of the forward weights into LDS (padded strides).
// Copy a rows x cols bf16 matrix (cols a multiple of 8) into LDS rows of stride
// `stride` elements, 16 bytes per load.
__device__ __forceinline__ void stage_rows(const uint16_t* __restrict__ src, uint16_t* dst, int rows, int cols,
                                           int stride) {
    const int per_row = cols / 8;
    const uint4* s = (const uint4*)src;
    for (int t = threadIdx.x; t < rows * per_row; t += blockDim.x) {
        const int r = t / per_row, q = t % per_row;
        *(uint4*)(dst + r * stride + 8 * q) = s[t];
    }
}

__device__ __forceinline__ void stage_forward_weights(const MlpDev& m, uint16_t* sw1t, uint16_t* sw2t, float* sb1,
                                                      float* sb2, float* sw3) {
    stage_rows(m.w1t, sw1t, kHid, kFeat, kS1);
    stage_rows(m.w2t, sw2t, kHid, kHid, kS2);
    for (int t = threadIdx.x; t < kHid; t += blockDim.x) {
        sb1[t] = m.b1[t];
        sb2[t] = m.b2[t];
        sw3[t] = m.w3r[t];
    }
}

constexpr int kPruneWaves = 4;

// Learned pruning epoch e (real coefficients): every live term is scored by the
// MLP, y = predicted log10|v|/4; a term with log10|c| + 4y < log10_tau is erased.
template <class Alg, class CP>
__global__ void __launch_bounds__(64 * kPruneWaves) k_prune_async(Store<CP, 1> st, Epochs ep, uint32_t e, MlpDev m,
                                                                  PruneGeo geo, FeatCtx fc, float log10_tau) {
    constexpr int NP = Alg::NP;
    __shared__ __attribute__((aligned(16))) uint16_t sw1t[kHid * kS1];
    __shared__ __attribute__((aligned(16))) uint16_t sw2t[kHid * kS2];
    __shared__ __attribute__((aligned(16))) uint16_t sf[kPruneWaves][64 * kS1];
    __shared__ float sb1[kHid], sb2[kHid], sw3[kHid], sy[kPruneWaves][64];
    __shared__ PruneGeo sgeo;
    uint32_t n, dead;
    if (!step_enter(ep, e, n, dead)) return;
    const uint32_t per_block = 64 * kPruneWaves;
    if (blockIdx.x * per_block >= n) return;  // block-uniform
    stage_forward_weights(m, sw1t, sw2t, sb1, sb2, sw3);
    if (threadIdx.x == 0) sgeo = geo;
    const float b3 = m.b3[0];
    const int wave = threadIdx.x >> 6, l = __lane_id();
    uint16_t* F = sf[wave];
    const size_t cap = st.cap;
    int dd = 0;
    __syncthreads();
    for (uint32_t bb = blockIdx.x * per_block; bb < n; bb += gridDim.x * per_block) {
        const uint32_t i = bb + wave * 64 + l;
        bool live = false;
        float logc = 0.f;
        for (int t = 0; t < kS1; ++t) F[l * kS1 + t] = 0;
        if (i < n) {
            const typename CP::val_t c = st.ld(st.coef, 0, i);
            live = !CP::dead(c);
            if (live) {
                uint64_t k[NP];
                key_load<NP>(st.key, i, k);
                const double a = fabs(CP::to_double(c));
                logc = a > 0.0 ? (float)log10(a) : -300.f;
                term_features<Alg>(k, logc, sgeo, fc, F + l * kS1);
            }
        }
        __syncthreads();
#pragma unroll 1
        for (int t = 0; t < 4; ++t) {
            v4f z1[8], z2[8];
            const float y = mlp16_forward(sw1t, sw2t, sb1, sb2, sw3, b3, F + 16 * t * kS1, z1, z2);
            if (l < 16) sy[wave][16 * t + l] = y;
        }
        __syncthreads();
        if (live && logc + 4.f * sy[wave][l] < log10_tau) {
            st.st(st.coef, 0, i, CP::make_dead());
            ++dd;
        }
        __syncthreads();
    }
    dd = block_sum_int(dd);
    if (threadIdx.x == 0 && dd) atomicAdd(&ep.dd[e], dd);
}

// ---------------------------------------------------------------------------
// Contribution logs
// ---------------------------------------------------------------------------
// Liveness bitmask of slots [0, n) after epoch e (coefficients in coef).
template <class CP>
__global__ void __launch_bounds__(kBlock) k_live_bits(const typename CP::store_t* __restrict__ coef, Epochs ep,
                                                      uint32_t e, uint32_t cap, uint64_t* __restrict__ bits) {
    if (run_stopped(ep)) return;
    const uint32_t n = min(cap, ep.n[e] + ep.app[e]);
    const uint32_t words = (n + 63) / 64;
    const uint32_t wid = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t nw = (gridDim.x * blockDim.x) >> 6;
    for (uint32_t w = wid; w < words; w += nw) {
        const uint32_t i = 64 * w + __lane_id();
        bool live = false;
        if (i < n) live = !CP::dead(CP::load(coef, 0, i));
        const uint64_t m = __ballot(live);
        if (__lane_id() == 0) bits[w] = m;
    }
}

__device__ __forceinline__ bool bit_of(const uint64_t* bits, uint32_t i) { return (bits[i >> 6] >> (i & 63)) & 1ull; }

// Readout vector of the final state: v_q = <b|q|b> on live slots, 0 elsewhere.
template <class Alg>
__global__ void __launch_bounds__(kBlock) k_adj_init(const uint64_t* __restrict__ key, size_t cap, uint32_t n,
                                                     const uint64_t* __restrict__ bits, StateWords sw,
                                                     double* __restrict__ v) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double val = 0.0;
        if (bit_of(bits, i)) {
            uint64_t k[Alg::NP];
            key_load<Alg::NP>(key, i, k);
            val = (double)state_factor<Alg>(k, sw.s);
        }
        v[i] = val;
    }
}

// Transposed rotation: v_in(q) = [q live before] (cos^[q anticommutes] v_out(q)
//   + [anticommutes, deg(P q) <= maxdeg] sign(q -> P q) sin v_out(P q)),
// where v_out is zero on slots not live after the rotation.  Every key that ever
// existed keeps its index entry (no compaction, no Clifford blocks), so P q is
// found whenever it held a slot after the rotation (slot < n_out).
template <class Alg>
__global__ void __launch_bounds__(kBlock) k_adjoint(const uint64_t* __restrict__ key, size_t cap,
                                                    const uint64_t* __restrict__ index, uint64_t hmask,
                                                    typename Alg::Gate g, double cs, double sn, int maxdeg,
                                                    uint32_t n_in, uint32_t n_out, const uint64_t* __restrict__ bits_in,
                                                    const double* __restrict__ v_out, double* __restrict__ v_in,
                                                    uint32_t* ierr) {
    constexpr int NP = Alg::NP;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n_in; i += gridDim.x * blockDim.x) {
        double v = 0.0;
        if (bit_of(bits_in, i)) {
            uint64_t k[NP];
            key_load<NP>(key, i, k);
            v = v_out[i];
            if (Alg::anticommutes(k, g)) {
                v = v * cs;
                uint64_t r[NP];
#pragma unroll
                for (int q = 0; q < NP; ++q) r[q] = k[q] ^ g.p[q];
                if (Alg::degree(r) <= maxdeg) {
                    const uint32_t j = probe<NP>(key, cap, index, hmask, r, n_out, ierr);
                    if (j != kNone) {
                        const double t = v_out[j] * sn;
                        v = Alg::sign(k, r, g) > 0 ? v + t : v - t;
                    }
                }
            }
        }
        v_in[i] = v;
    }
}

// Logged state: for every live slot, its feature row and target
// clamp(log10(|v| + 1e-12), -12, 0)/4; sum_q c_q v_q into sums[0..1] (re, im).
template <class Alg, int K>
__global__ void __launch_bounds__(kBlock) k_log_extract(const uint64_t* __restrict__ key, size_t cap,
                                                        const double* __restrict__ coef, uint32_t n,
                                                        const uint64_t* __restrict__ bits,
                                                        const double* __restrict__ v, PruneGeo geo, FeatCtx fc,
                                                        uint64_t max_rows, unsigned long long* counter,
                                                        float* __restrict__ feats, float* __restrict__ targets,
                                                        double* __restrict__ sums) {
    constexpr int NP = Alg::NP;
    __shared__ __attribute__((aligned(16))) uint16_t sf[kBlock][kFeat];
    __shared__ PruneGeo sgeo;
    __shared__ double red[2][kBlock];
    if (threadIdx.x == 0) sgeo = geo;
    __syncthreads();
    double acc[2] = {0.0, 0.0};
    for (uint32_t b0 = blockIdx.x * blockDim.x; b0 < n; b0 += gridDim.x * blockDim.x) {
        const uint32_t i = b0 + threadIdx.x;
        if (i < n && bit_of(bits, i)) {
            uint64_t k[NP];
            key_load<NP>(key, i, k);
            const double cr = coef[i];
            const double ci = K == 2 ? coef[cap + i] : 0.0;
            const double vi = v[i];
            acc[0] += cr * vi;
            acc[1] += ci * vi;
            const double a = K == 2 ? hypot(cr, ci) : fabs(cr);
            const float logc = a > 0.0 ? (float)log10(a) : -300.f;
            uint16_t* f = sf[threadIdx.x];
            for (int t = 0; t < kFeat; ++t) f[t] = 0;
            term_features<Alg>(k, logc, sgeo, fc, f);
            const unsigned long long row = atomicAdd(counter, 1ull);
            if (row < max_rows) {
                for (int t = 0; t < kFeat; ++t) feats[row * kFeat + t] = bf2f(f[t]);
                const double lv = log10(fabs(vi) + 1e-12);
                targets[row] = (float)(fmin(fmax(lv, -12.0), 0.0) * 0.25);
            }
        }
    }
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
        atomicAdd(&sums[0], red[0][0]);
        atomicAdd(&sums[1], red[1][0]);
    }
}

}  // namespace ffx
