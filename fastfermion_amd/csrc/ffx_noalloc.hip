// ffx_noalloc.hip -- ffx_pauli_propagate_noalloc: Pauli propagation that calls no
// device allocator.  Every device block of the call (term store, hash index, orbit
// claim table, counters, epoch arrays, schedule tables, staging, control block,
// partial sums) is carved from `scratch`, a fixed array that each device holds
// from the first such call until the process ends; the results go to caller
// arrays.  The kernels, schedule and results are those of ffx_pauli_propagate.
#include <mutex>

#include "ffx_engine.h"

using namespace ffx_host;

namespace ffx_host {
FFX_EXTERN_WIDTH(1);
FFX_EXTERN_WIDTH(2);
FFX_EXTERN_WIDTH(3);
FFX_EXTERN_WIDTH(4);
}  // namespace ffx_host

namespace {

constexpr size_t kScratchBytes = FFX_NOALLOC_SCRATCH_BYTES;

// The scratch array.  In host code its name is only a handle for
// hipGetSymbolAddress: kernels and copies get the device address from there.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mu;
char* g_scratch_addr[kMaxDevices] = {};

// Device address of the current device's scratch array, looked up once per device.
int scratch_address(ffx_ctx* ctx, char** out) {
    if (ctx->device < 0 || ctx->device >= kMaxDevices) return set_err(ctx, FFX_ENODEV, "device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    char*& a = g_scratch_addr[ctx->device];
    if (!a) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();
            return set_err(ctx, FFX_EHIP, std::string("hipGetSymbolAddress of the scratch array failed: ") +
                                              hipGetErrorString(e));
        }
        a = (char*)p;
        g_fixed_lo[ctx->device].store((uintptr_t)p, std::memory_order_relaxed);
        g_fixed_hi[ctx->device].store((uintptr_t)p + kScratchBytes, std::memory_order_relaxed);
    }
    *out = a;
    return FFX_OK;
}

// Term-store slots of an allocation-free call: the largest power of two whose
// per-slot blocks (keys and compaction keys, coefficient ping-pong planes, hash
// index and orbit claim entries, result staging, per-tile counts) fit beside the
// per-call blocks (epoch arrays for up to four device steps per gate, schedule
// tables, input staging, control and reduction blocks).
size_t plan_slots(int W, int prec, int K, int64_t n_gates, int64_t n_terms) {
    const size_t key_b = 16 * (size_t)W;
    const size_t coef_b = (size_t)K * (prec == FFX_FP32 ? 4 : 8);
    const size_t per_slot = 2 * key_b + 2 * coef_b + 16 + 16 + (key_b + 16) + 1;
    const size_t ne = pow2ceil(std::max<size_t>(4 * (size_t)n_gates + 8, 64));
    const size_t epoch_b = ((4 + (size_t)kStripes * kWin + kWin + 2) * ne + 64) * sizeof(uint32_t);
    const size_t fixed = epoch_b + 384 * (size_t)n_gates + (key_b + 16) * (size_t)n_terms + (4u << 20);
    if (fixed >= kScratchBytes) return 0;
    size_t cap = 1;
    while (2 * cap * per_slot <= kScratchBytes - fixed) cap *= 2;
    return cap;
}

}  // namespace

extern "C" int ffx_pauli_propagate_noalloc(ffx_ctx* ctx, const ffx_pauli_gate* gates, int64_t n_gates,
                                           int32_t n_words, const uint64_t* x, const uint64_t* z,
                                           const double* c_re, const double* c_im, int64_t n_terms,
                                           const ffx_options* opts, const uint64_t* basis_state, double out_expect[2],
                                           uint64_t* out_x, uint64_t* out_z, double* out_re, double* out_im,
                                           int64_t out_cap, int64_t* out_n) {
    if (!ctx) return FFX_EINVAL;
    ctx->err.clear();
    if (out_n) *out_n = 0;
    if (out_cap < 0 || (out_cap > 0 && (!out_x || !out_z || !out_re)))
        return set_err(ctx, FFX_EINVAL, "output arrays: out_x, out_z and out_re of out_cap terms (out_im optional)");
    int K = 1;
    CHK(check_pauli_args(ctx, gates, n_gates, n_words, x, z, c_re, c_im, n_terms, opts, &K));
    char* base = nullptr;
    CHK(scratch_address(ctx, &base));
    const size_t cap = plan_slots(n_words, opts->precision, K, n_gates, n_terms);
    if (cap < 1024 || cap < 2 * (size_t)n_terms + 64)
        return set_err(ctx, FFX_ECAPACITY,
                       "allocation-free call: " + std::to_string(n_terms) + " terms and " + std::to_string(n_gates) +
                           " gates exceed its scratch of " + std::to_string(kScratchBytes) + " bytes");
    // the context's own buffers step aside for the call; the call's come from the scratch array
    DevBuffers own;
    swap_buffers(ctx, own);
    const int timing = ctx->kernel_timing;
    ctx->kernel_timing = 0;  // event pairs are created on demand
    ctx->arena.base = base;
    ctx->arena.size = kScratchBytes;
    ctx->arena.used = 0;
    ctx->arena.cap = cap;
    ctx->arena.on = true;
    const bool want_terms = out_cap > 0;
    ctx->sink.x = out_x;
    ctx->sink.z = out_z;
    ctx->sink.re = out_re;
    ctx->sink.im = out_im;
    ctx->sink.cap = out_cap;
    ctx->sink.n = 0;
    ctx->sink.on = true;
    ctx->prof = ffx_profile{};
    ctx->launches = 0;
    int rc = dev_alloc(ctx, &ctx->ctrl, sizeof(Ctrl));
    if (rc == FFX_OK) rc = dev_alloc(ctx, &ctx->total, 64);
    if (rc == FFX_OK) rc = dev_alloc(ctx, &ctx->partial, 2 * 1024 * sizeof(double));
    if (rc == FFX_OK) rc = dev_alloc(ctx, &ctx->result, 64);
    if (rc == FFX_OK) {
        ffx_terms unused{};
        ffx_terms* out = want_terms ? &unused : nullptr;
        const int prec = opts->precision;
        switch (n_words) {
            case 1: rc = dispatch_pauli<1>(ctx, gates, n_gates, x, z, c_re, c_im, n_terms, *opts, out, basis_state, out_expect, prec, K, nullptr); break;
            case 2: rc = dispatch_pauli<2>(ctx, gates, n_gates, x, z, c_re, c_im, n_terms, *opts, out, basis_state, out_expect, prec, K, nullptr); break;
            case 3: rc = dispatch_pauli<3>(ctx, gates, n_gates, x, z, c_re, c_im, n_terms, *opts, out, basis_state, out_expect, prec, K, nullptr); break;
            default: rc = dispatch_pauli<4>(ctx, gates, n_gates, x, z, c_re, c_im, n_terms, *opts, out, basis_state, out_expect, prec, K, nullptr); break;
        }
        if (rc == FFX_OK && !want_terms) ctx->sink.n = (int64_t)ctx->ctrl_host->n - (int64_t)ctx->ctrl_host->n_dead;
    }
    if (out_n) *out_n = ctx->sink.n;
    // after an error, launches may still be queued that use the scratch array
    (void)hipStreamSynchronize(ctx->stream);
    ctx->arena = ffx_ctx::Arena{};
    ctx->sink = ffx_ctx::Sink{};
    ctx->kernel_timing = timing;
    swap_buffers(ctx, own);  // `own` now holds blocks of the scratch array: nothing to release
    ctx->last_kind = 0;      // no device-resident result for ffx_expect_last
    return rc;
}
