// ffx_mlp.hip -- training and inference of the learned term pruner on the GPU,
// and its C ABI (ffx_pruner_*).
//
// Model: MLP 64 -> 128 -> 128 -> 1 with ReLU, bf16 weights and activations,
// fp32 accumulation on MFMA (v_mfma_f32_16x16x32_bf16); fp32 master
// parameters and Adam moments.  One training step on a minibatch is
//   k_mlp_fb     forward and backward of every 16-row tile (one wave each):
//                activations h1, dz1, dz2 (bf16) for the weight gradients,
//                bias gradients, w3 gradient and the squared error directly;
//   k_wgrad x2   dW1 = X^T dz1 and dW2 = h1^T dz2 (MFMA over the minibatch rows);
//   k_adam       parameter update, then k_mlp_pack refreshes the bf16 copies.
// Loss: mean squared error against the target log10|v|/4 of the contribution
// logs, with predictions below their target weighted by under_weight >= 1 (a
// prediction that is too low can erase a term that matters; one that is too
// high only keeps a term).  Gradients are those of the bf16 model, applied to
// the fp32 masters.
#include "ffx_engine.h"

using namespace ffx_host;

namespace ffx {

constexpr int kMlpBlock = 256;  // 4 waves

// bf16 copies of the weights from the fp32 master vector.
static __global__ void k_mlp_pack(const float* __restrict__ p, uint16_t* __restrict__ w1t, uint16_t* __restrict__ w2t,
                                  uint16_t* __restrict__ w2, float* __restrict__ w3r) {
    const int stride = gridDim.x * blockDim.x;
    const int t0 = blockIdx.x * blockDim.x + threadIdx.x;
    for (int t = t0; t < kHid * kFeat; t += stride) {
        const int o = t / kFeat, i = t % kFeat;
        w1t[t] = f2bf(p[kOffW1 + i * kHid + o]);
    }
    for (int t = t0; t < kHid * kHid; t += stride) {
        const int o = t / kHid, i = t % kHid;
        w2t[t] = f2bf(p[kOffW2 + i * kHid + o]);
        w2[t] = f2bf(p[kOffW2 + t]);
    }
    for (int t = t0; t < kHid; t += stride) w3r[t] = bf2f(f2bf(p[kOffW3 + t]));
}

// Sum over the 16 lanes of a lane group (lanes with equal l/16).
__device__ __forceinline__ float group16_sum(float v) {
    v += __shfl_xor(v, 1, 64);
    v += __shfl_xor(v, 2, 64);
    v += __shfl_xor(v, 4, 64);
    v += __shfl_xor(v, 8, 64);
    return v;
}

// Forward (mode 0: predictions to Y) or forward + backward (mode 1) of nb rows;
// row r of the minibatch is row idx[r] (idx NULL: r) of X / T.
static __global__ void __launch_bounds__(kMlpBlock) k_mlp_fb(MlpDev m, const uint16_t* __restrict__ X,
                                                             const int32_t* __restrict__ idx,
                                                             const float* __restrict__ T, int nb, float inv_n,
                                                             float under_w, int mode, float* __restrict__ Y,
                                                             uint16_t* __restrict__ H1, uint16_t* __restrict__ DZ1,
                                                             uint16_t* __restrict__ DZ2, float* __restrict__ grad,
                                                             double* __restrict__ loss) {
    __shared__ __attribute__((aligned(16))) uint16_t sw1t[kHid * kS1];
    __shared__ __attribute__((aligned(16))) uint16_t sw2t[kHid * kS2];
    __shared__ __attribute__((aligned(16))) uint16_t sw2[kHid * kS2];
    __shared__ __attribute__((aligned(16))) uint16_t sx[4][16 * kS1];
    __shared__ float sb1[kHid], sb2[kHid], sw3[kHid];
    __shared__ float gb1[kHid], gb2[kHid], gw3[kHid];
    __shared__ float gb3;
    __shared__ double sloss;
    const int per_block = 64;
    if (blockIdx.x * per_block >= nb) return;  // block-uniform
    stage_forward_weights(m, sw1t, sw2t, sb1, sb2, sw3);
    if (mode) stage_rows(m.w2, sw2, kHid, kHid, kS2);
    for (int t = threadIdx.x; t < kHid; t += blockDim.x) gb1[t] = gb2[t] = gw3[t] = 0.f;
    if (threadIdx.x == 0) {
        gb3 = 0.f;
        sloss = 0.0;
    }
    const float b3 = m.b3[0];
    const int wave = threadIdx.x >> 6, l = __lane_id();
    const int r = l & 15, g = l >> 4;
    uint16_t* xs = sx[wave];
    __syncthreads();
    for (int base = blockIdx.x * per_block; base < nb; base += gridDim.x * per_block) {
        // stage the wave's 16 rows: lane l copies 16 features of row l/4
        {
            const int row = base + wave * 16 + (l >> 2), part = l & 3;
            uint16_t* dst = xs + (l >> 2) * kS1 + 16 * part;
            if (row < nb) {
                const int src = idx ? idx[row] : row;
                const v4s* s4 = (const v4s*)(X + (size_t)src * kFeat + 16 * part);
#pragma unroll
                for (int q = 0; q < 4; ++q) *(v4s*)(dst + 4 * q) = s4[q];
            } else {
#pragma unroll
                for (int q = 0; q < 16; ++q) dst[q] = 0;
            }
        }
        __syncthreads();
        v4f z1[8], z2[8];
        const float y = mlp16_forward(sw1t, sw2t, sb1, sb2, sw3, b3, xs, z1, z2);
        const int row = base + wave * 16 + r;
        const bool valid = row < nb;
        if (mode == 0) {
            if (g == 0 && valid) Y[row] = y;
        } else {
            const float t = valid ? T[idx ? idx[row] : row] : 0.f;
            const float err = valid ? y - t : 0.f;
            const float wgt = err < 0.f ? under_w : 1.f;  // predictions below the target weigh under_w
            if (g == 0 && valid) atomicAdd(&sloss, (double)wgt * (double)err * (double)err);
            const float dy = 2.f * wgt * err * inv_n;
            const float sdy = group16_sum(dy);  // all lanes: the shuffles need every lane active
            if (l == 0) atomicAdd(&gb3, sdy);
            // layer 3 and the ReLU of layer 2
            v4s dz2b[8];
#pragma unroll
            for (int mt = 0; mt < 8; ++mt) {
                float d[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int o = 16 * mt + 4 * g + i;
                    const float h2 = fmaxf(z2[mt][i], 0.f);
                    d[i] = z2[mt][i] > 0.f ? dy * sw3[o] : 0.f;
                    const float sd = group16_sum(d[i]);
                    const float sh = group16_sum(h2 * dy);
                    if (r == 0) {
                        atomicAdd(&gb2[o], sd);
                        atomicAdd(&gw3[o], sh);
                    }
                }
                dz2b[mt] = pack4(d[0], d[1], d[2], d[3]);
                if (valid) *(v4s*)(DZ2 + (size_t)row * kHid + 16 * mt + 4 * g) = dz2b[mt];
            }
            // dh1^T (128 x 16) = W2 (128 in x 128 out) * dz2^T (128 x 16); ReLU of layer 1
#pragma unroll
            for (int mt = 0; mt < 8; ++mt) {
                v4f acc = {0.f, 0.f, 0.f, 0.f};
                const uint16_t* w = sw2 + (16 * mt + r) * kS2;
#pragma unroll
                for (int s = 0; s < 4; ++s)
                    acc = mfma32(ld4(w + 32 * s + 4 * g), ld4(w + 32 * s + 16 + 4 * g), dz2b[2 * s], dz2b[2 * s + 1],
                                 acc);
                float d[4], h[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int o = 16 * mt + 4 * g + i;
                    d[i] = z1[mt][i] > 0.f ? acc[i] : 0.f;
                    h[i] = fmaxf(z1[mt][i], 0.f);
                    const float sd = group16_sum(d[i]);
                    if (r == 0) atomicAdd(&gb1[o], sd);
                }
                if (valid) {
                    *(v4s*)(DZ1 + (size_t)row * kHid + 16 * mt + 4 * g) = pack4(d[0], d[1], d[2], d[3]);
                    *(v4s*)(H1 + (size_t)row * kHid + 16 * mt + 4 * g) = pack4(h[0], h[1], h[2], h[3]);
                }
            }
        }
        __syncthreads();
    }
    if (mode) {
        for (int t = threadIdx.x; t < kHid; t += blockDim.x) {
            atomicAdd(&grad[kOffB1 + t], gb1[t]);
            atomicAdd(&grad[kOffB2 + t], gb2[t]);
            atomicAdd(&grad[kOffW3 + t], gw3[t]);
        }
        if (threadIdx.x == 0) {
            atomicAdd(&grad[kOffB3], gb3);
            atomicAdd(loss, sloss);
        }
    }
}

// out[a][b] += sum over rows of A[row][a] * B[row][b] (A rows gathered through aidx
// when given); a in {64, 128}, b = 128.  Block (x, y) takes rows [x * rows_per_block,
// ...) and the output rows [32y, 32y + 32): 2 x 8 tiles of 16x16, four per wave.  The
// block stages its row chunks (64 rows) of A's 32 columns and of B in LDS with
// 16-byte loads.
static __global__ void __launch_bounds__(kMlpBlock) k_wgrad(const uint16_t* __restrict__ A,
                                                            const int32_t* __restrict__ aidx, int a,
                                                            const uint16_t* __restrict__ B, int b, int nrows,
                                                            int rows_per_block, float* __restrict__ out) {
    constexpr int CH = 64;
    constexpr int sA = 32 + 8;
    __shared__ __attribute__((aligned(16))) uint16_t sa[CH * sA];
    __shared__ __attribute__((aligned(16))) uint16_t sb[CH * (kHid + 8)];
    const int sB = b + 8;
    const int tg = blockIdx.y;  // output rows [32 tg, 32 tg + 32)
    const int wave = threadIdx.x >> 6, l = __lane_id();
    const int r = l & 15, g = l >> 4;
    const int ia = wave >> 1;        // local 16-row tile of the output rows
    const int ib0 = 4 * (wave & 1);  // first of the wave's four 16-column tiles
    v4f acc[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[j] = v4f{0.f, 0.f, 0.f, 0.f};
    const int r0 = blockIdx.x * rows_per_block;
    const int r1 = min(nrows, r0 + rows_per_block);
    const int bq = b / 8;  // 16-byte pieces of a B row
    for (int c0 = r0; c0 < r1; c0 += CH) {
        {
            const int rr = threadIdx.x >> 2, q = threadIdx.x & 3, row = c0 + rr;
            uint4 v = {0u, 0u, 0u, 0u};
            if (row < r1) v = *(const uint4*)(A + (size_t)(aidx ? aidx[row] : row) * a + 32 * tg + 8 * q);
            *(uint4*)(sa + rr * sA + 8 * q) = v;
        }
        for (int t = threadIdx.x; t < CH * bq; t += blockDim.x) {
            const int rr = t / bq, q = t % bq, row = c0 + rr;
            uint4 v = {0u, 0u, 0u, 0u};
            if (row < r1) v = *(const uint4*)(B + (size_t)row * b + 8 * q);
            *(uint4*)(sb + rr * sB + 8 * q) = v;
        }
        __syncthreads();
#pragma unroll
        for (int s = 0; s < CH / 32; ++s) {
            v4s alo, ahi;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int rr = 32 * s + 4 * g + q;  // and rr + 16 for the hi halves
                alo[q] = (short)sa[rr * sA + 16 * ia + r];
                ahi[q] = (short)sa[(rr + 16) * sA + 16 * ia + r];
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                v4s blo, bhi;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int rr = 32 * s + 4 * g + q;
                    blo[q] = (short)sb[rr * sB + 16 * (ib0 + j) + r];
                    bhi[q] = (short)sb[(rr + 16) * sB + 16 * (ib0 + j) + r];
                }
                acc[j] = mfma32(alo, ahi, blo, bhi, acc[j]);
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int i = 0; i < 4; ++i)
            atomicAdd(&out[(size_t)(32 * tg + 16 * ia + 4 * g + i) * b + 16 * (ib0 + j) + r], acc[j][i]);
}

static __global__ void k_adam(float* __restrict__ p, const float* __restrict__ gr, float* __restrict__ m1,
                              float* __restrict__ m2, int n, float lr, float bc1, float bc2) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float gi = gr[i];
        const float a = 0.9f * m1[i] + 0.1f * gi;
        const float b = 0.999f * m2[i] + 0.001f * gi * gi;
        m1[i] = a;
        m2[i] = b;
        p[i] -= lr * (a / bc1) / (sqrtf(b / bc2) + 1e-8f);
    }
}

}  // namespace ffx

namespace {

int mlp_ensure(ffx_ctx* ctx) {
    if (ctx->mlp) return FFX_OK;
    MlpState* s = new (std::nothrow) MlpState();
    if (!s) return set_err(ctx, FFX_ENOMEM, "host allocation failed");
    ctx->mlp = s;
    const size_t np = kNParams;
    CHK(dev_alloc(ctx, &s->p, np * sizeof(float)));
    CHK(dev_alloc(ctx, &s->g, np * sizeof(float)));
    CHK(dev_alloc(ctx, &s->m1, np * sizeof(float)));
    CHK(dev_alloc(ctx, &s->m2, np * sizeof(float)));
    CHK(dev_alloc(ctx, &s->w1t, kHid * kFeat * sizeof(uint16_t)));
    CHK(dev_alloc(ctx, &s->w2t, kHid * kHid * sizeof(uint16_t)));
    CHK(dev_alloc(ctx, &s->w2, kHid * kHid * sizeof(uint16_t)));
    CHK(dev_alloc(ctx, &s->w3r, kHid * sizeof(float)));
    CHK(dev_alloc(ctx, &s->loss, sizeof(double)));
    return FFX_OK;
}

int mlp_pack(ffx_ctx* ctx) {
    MlpState* s = ctx->mlp;
    hipLaunchKernelGGL(k_mlp_pack, dim3(64), dim3(256), 0, ctx->stream, s->p, s->w1t, s->w2t, s->w2, s->w3r);
    HIPCHK(hipGetLastError());
    return FFX_OK;
}

int mlp_data(ffx_ctx* ctx, const float* feats, const float* targets, int64_t n) {
    MlpState* s = ctx->mlp;
    if ((size_t)n > s->data_cap) {
        dev_free(s->X);
        dev_free(s->T);
        s->data_cap = 0;
        CHK(dev_alloc(ctx, &s->X, (size_t)n * kFeat * sizeof(uint16_t)));
        CHK(dev_alloc(ctx, &s->T, (size_t)n * sizeof(float)));
        s->data_cap = (size_t)n;
    }
    std::vector<uint16_t> xb((size_t)n * kFeat);
    for (size_t i = 0; i < xb.size(); ++i) xb[i] = f2bf(feats[i]);
    HIPCHK(hipMemcpyAsync(s->X, xb.data(), xb.size() * sizeof(uint16_t), hipMemcpyHostToDevice, ctx->stream));
    if (targets) HIPCHK(hipMemcpyAsync(s->T, targets, (size_t)n * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));  // xb goes out of scope
    return FFX_OK;
}

int mlp_minibatch(ffx_ctx* ctx, size_t nb) {
    MlpState* s = ctx->mlp;
    if (nb <= s->mb_cap) return FFX_OK;
    dev_free(s->H1);
    dev_free(s->DZ1);
    dev_free(s->DZ2);
    dev_free(s->Y);
    s->mb_cap = 0;
    CHK(dev_alloc(ctx, &s->H1, nb * kHid * sizeof(uint16_t)));
    CHK(dev_alloc(ctx, &s->DZ1, nb * kHid * sizeof(uint16_t)));
    CHK(dev_alloc(ctx, &s->DZ2, nb * kHid * sizeof(uint16_t)));
    CHK(dev_alloc(ctx, &s->Y, nb * sizeof(float)));
    s->mb_cap = nb;
    return FFX_OK;
}

unsigned fb_grid(ffx_ctx* ctx, int64_t nb) {
    if (ctx->n_cu <= 0) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, ctx->device) == hipSuccess) ctx->n_cu = prop.multiProcessorCount;
    }
    const int64_t tiles = (nb + 63) / 64;
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>(tiles, 2 * std::max(ctx->n_cu, 1)));
}

// Forward + backward of rows [off, off + nb) of the uploaded data (through s->idx when shuffled).
int mlp_grad_step(ffx_ctx* ctx, const int32_t* idx, int64_t nb, float under_w) {
    MlpState* s = ctx->mlp;
    HIPCHK(hipMemsetAsync(s->g, 0, kNParams * sizeof(float), ctx->stream));
    hipLaunchKernelGGL(k_mlp_fb, dim3(fb_grid(ctx, nb)), dim3(kMlpBlock), 0, ctx->stream, s->dev(), s->X, idx, s->T,
                       (int)nb, 1.f / (float)nb, under_w, 1, s->Y, s->H1, s->DZ1, s->DZ2, s->g, s->loss);
    HIPCHK(hipGetLastError());
    const int rpb = 128;
    const unsigned gw = (unsigned)((nb + rpb - 1) / rpb);
    hipLaunchKernelGGL(k_wgrad, dim3(gw, kFeat / 32), dim3(kMlpBlock), 0, ctx->stream, s->X, idx, kFeat, s->DZ1, kHid,
                       (int)nb, rpb, s->g + kOffW1);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_wgrad, dim3(gw, kHid / 32), dim3(kMlpBlock), 0, ctx->stream, s->H1, (const int32_t*)nullptr,
                       kHid, s->DZ2, kHid, (int)nb, rpb, s->g + kOffW2);
    HIPCHK(hipGetLastError());
    return FFX_OK;
}

inline uint64_t splitmix(uint64_t& x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

}  // namespace

void ffx_mlp_free(ffx_ctx* ctx) {
    MlpState* s = ctx->mlp;
    if (!s) return;
    dev_free(s->p);
    dev_free(s->g);
    dev_free(s->m1);
    dev_free(s->m2);
    dev_free(s->w1t);
    dev_free(s->w2t);
    dev_free(s->w2);
    dev_free(s->w3r);
    dev_free(s->loss);
    dev_free(s->X);
    dev_free(s->T);
    dev_free(s->idx);
    dev_free(s->H1);
    dev_free(s->DZ1);
    dev_free(s->DZ2);
    dev_free(s->Y);
    delete s;
    ctx->mlp = nullptr;
}

extern "C" {

int ffx_pruner_set_params(ffx_ctx* ctx, const float* params, int64_t n) {
    if (!ctx) return FFX_EINVAL;
    ctx->err.clear();
    if (!params || n != kNParams) return set_err(ctx, FFX_EINVAL, "parameter vector must have FFX_PRUNER_NPARAMS floats");
    for (int64_t i = 0; i < n; ++i)
        if (!std::isfinite(params[i])) return set_err(ctx, FFX_EINVAL, "non-finite parameter");
    if (hipSetDevice(ctx->device) != hipSuccess) return set_err(ctx, FFX_ENODEV, "hipSetDevice failed");
    CHK(mlp_ensure(ctx));
    MlpState* s = ctx->mlp;
    HIPCHK(hipMemcpyAsync(s->p, params, kNParams * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemsetAsync(s->m1, 0, kNParams * sizeof(float), ctx->stream));
    HIPCHK(hipMemsetAsync(s->m2, 0, kNParams * sizeof(float), ctx->stream));
    s->step = 0;
    CHK(mlp_pack(ctx));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    s->ready = true;
    return FFX_OK;
}

int ffx_pruner_init(ffx_ctx* ctx, uint64_t seed) {
    if (!ctx) return FFX_EINVAL;
    std::vector<float> p(kNParams, 0.f);
    uint64_t st = seed;
    auto normal = [&]() {
        // Box-Muller on two SplitMix64 uniforms in (0, 1]
        const double u1 = ((splitmix(st) >> 11) + 1) * (1.0 / 9007199254740992.0);
        const double u2 = (splitmix(st) >> 11) * (1.0 / 9007199254740992.0);
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    };
    for (int i = 0; i < kFeat * kHid; ++i) p[kOffW1 + i] = (float)(normal() * std::sqrt(2.0 / kFeat));
    for (int i = 0; i < kHid * kHid; ++i) p[kOffW2 + i] = (float)(normal() * std::sqrt(2.0 / kHid));
    for (int i = 0; i < kHid; ++i) p[kOffW3 + i] = (float)(normal() * std::sqrt(1.0 / kHid));
    p[kOffB3] = -1.f;
    return ffx_pruner_set_params(ctx, p.data(), kNParams);
}

int ffx_pruner_get_params(ffx_ctx* ctx, float* params, int64_t n) {
    if (!ctx) return FFX_EINVAL;
    ctx->err.clear();
    if (!params || n != kNParams) return set_err(ctx, FFX_EINVAL, "parameter vector must have FFX_PRUNER_NPARAMS floats");
    if (!ctx->mlp || !ctx->mlp->ready) return set_err(ctx, FFX_EINVAL, "the pruner has no parameters yet");
    if (hipSetDevice(ctx->device) != hipSuccess) return set_err(ctx, FFX_ENODEV, "hipSetDevice failed");
    HIPCHK(hipMemcpyAsync(params, ctx->mlp->p, kNParams * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return FFX_OK;
}

int ffx_pruner_forward(ffx_ctx* ctx, const float* feats, int64_t n, float* out) {
    if (!ctx) return FFX_EINVAL;
    ctx->err.clear();
    if (n < 0 || (n > 0 && (!feats || !out)) || n > 0x7FFFFFFF) return set_err(ctx, FFX_EINVAL, "invalid arguments");
    if (!ctx->mlp || !ctx->mlp->ready) return set_err(ctx, FFX_EINVAL, "the pruner has no parameters yet");
    if (n == 0) return FFX_OK;
    if (hipSetDevice(ctx->device) != hipSuccess) return set_err(ctx, FFX_ENODEV, "hipSetDevice failed");
    CHK(mlp_data(ctx, feats, nullptr, n));
    CHK(mlp_minibatch(ctx, (size_t)n));
    MlpState* s = ctx->mlp;
    hipLaunchKernelGGL(k_mlp_fb, dim3(fb_grid(ctx, n)), dim3(kMlpBlock), 0, ctx->stream, s->dev(), s->X,
                       (const int32_t*)nullptr, (const float*)nullptr, (int)n, 0.f, 1.f, 0, s->Y, s->H1, s->DZ1, s->DZ2,
                       s->g, s->loss);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, s->Y, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return FFX_OK;
}

int ffx_pruner_loss_grad(ffx_ctx* ctx, const float* feats, const float* targets, int64_t n, float under_weight,
                         double* loss, float* grad) {
    if (!ctx) return FFX_EINVAL;
    ctx->err.clear();
    if (n <= 0 || !feats || !targets || n > 0x7FFFFFFF || !(under_weight > 0.f))
        return set_err(ctx, FFX_EINVAL, "invalid arguments");
    if (!ctx->mlp || !ctx->mlp->ready) return set_err(ctx, FFX_EINVAL, "the pruner has no parameters yet");
    if (hipSetDevice(ctx->device) != hipSuccess) return set_err(ctx, FFX_ENODEV, "hipSetDevice failed");
    CHK(mlp_data(ctx, feats, targets, n));
    CHK(mlp_minibatch(ctx, (size_t)n));
    MlpState* s = ctx->mlp;
    HIPCHK(hipMemsetAsync(s->loss, 0, sizeof(double), ctx->stream));
    CHK(mlp_grad_step(ctx, nullptr, n, under_weight));
    double l = 0.0;
    HIPCHK(hipMemcpyAsync(&l, s->loss, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    if (grad) HIPCHK(hipMemcpyAsync(grad, s->g, kNParams * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (loss) *loss = l / (double)n;
    return FFX_OK;
}

int ffx_pruner_train(ffx_ctx* ctx, const float* feats, const float* targets, int64_t n, int32_t epochs,
                     int32_t batch, float lr, float under_weight, uint64_t seed, double* epoch_loss) {
    if (!ctx) return FFX_EINVAL;
    ctx->err.clear();
    if (n <= 0 || !feats || !targets || epochs < 0 || batch < 16 || !(lr > 0.f) || n > 0x7FFFFFFF ||
        !(under_weight > 0.f))
        return set_err(ctx, FFX_EINVAL, "invalid arguments");
    if (!ctx->mlp || !ctx->mlp->ready) return set_err(ctx, FFX_EINVAL, "the pruner has no parameters yet");
    if (hipSetDevice(ctx->device) != hipSuccess) return set_err(ctx, FFX_ENODEV, "hipSetDevice failed");
    CHK(mlp_data(ctx, feats, targets, n));
    const int64_t nb = std::min<int64_t>(batch, n);
    CHK(mlp_minibatch(ctx, (size_t)nb));
    MlpState* s = ctx->mlp;
    if ((size_t)n > s->idx_cap) {
        dev_free(s->idx);
        s->idx_cap = 0;
        CHK(dev_alloc(ctx, &s->idx, (size_t)n * sizeof(int32_t)));
        s->idx_cap = (size_t)n;
    }
    std::vector<int32_t> perm((size_t)n);
    uint64_t st = seed;
    for (int32_t ep = 0; ep < epochs; ++ep) {
        for (int64_t i = 0; i < n; ++i) perm[(size_t)i] = (int32_t)i;
        for (int64_t i = n - 1; i > 0; --i) std::swap(perm[(size_t)i], perm[(size_t)(splitmix(st) % (uint64_t)(i + 1))]);
        HIPCHK(hipMemcpyAsync(s->idx, perm.data(), (size_t)n * sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(hipMemsetAsync(s->loss, 0, sizeof(double), ctx->stream));
        for (int64_t off = 0; off < n; off += nb) {
            const int64_t cur = std::min<int64_t>(nb, n - off);
            CHK(mlp_grad_step(ctx, s->idx + off, cur, under_weight));
            s->step++;
            const float bc1 = 1.f - std::pow(0.9f, (float)s->step);
            const float bc2 = 1.f - std::pow(0.999f, (float)s->step);
            hipLaunchKernelGGL(k_adam, dim3(64), dim3(256), 0, ctx->stream, s->p, s->g, s->m1, s->m2, kNParams, lr,
                               bc1, bc2);
            HIPCHK(hipGetLastError());
            CHK(mlp_pack(ctx));
        }
        double l = 0.0;
        HIPCHK(hipMemcpyAsync(&l, s->loss, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));  // perm is rewritten next epoch
        if (epoch_loss) epoch_loss[ep] = l / (double)n;
    }
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return FFX_OK;
}

}  // extern "C"
