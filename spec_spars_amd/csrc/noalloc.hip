// ss_er_sample_noalloc: the fused effective-resistance / coin / sample /
// compaction pass over device-resident arrays, for callers that cannot
// allocate device memory.  The pass needs device memory beyond its inputs
// and outputs -- the ticket counter, the kept total, the error flag and one
// 64-bit look-back word per 256-edge tile -- and here it lives in a fixed
// array, `scratch`, that every GPU the library runs on holds until the
// process ends.  Every call on one GPU shares that GPU's copy: calls that
// overlap on one device (two streams, two threads, two contexts) race.
#include <algorithm>
#include <mutex>

#include "device.h"

namespace ss {
namespace {

// words [0, kHeadWords): ticket (low 32 bits of word 0), kept total (word 1),
// error flag (low 32 bits of word 2); then one word per tile
constexpr int64_t kHeadWords = 8;
constexpr int64_t kMaxTiles = (SS_NOALLOC_MAX_EDGES + 255) / 256;
constexpr int64_t kScratchWords = kHeadWords + kMaxTiles;

static __device__ __align__(256) uint64_t scratch[kScratchWords];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
uint64_t *g_scratch[kMaxDevices] = {};   // device address of `scratch`, per device

__global__ __launch_bounds__(256) void k_clear_words(int64_t count, uint64_t *p) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < count; i += (int64_t)gridDim.x * 256) p[i] = 0;
}

}  // namespace

// Once per device, with that device current (ss_context_create): the
// device address of this device's copy of `scratch`.  Host code uses that
// address, never the array's name, which on the host is only a stand-in.
int noalloc_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return fail(SS_ERR_INVALID, "device %d beyond %d", device, kMaxDevices);
    std::lock_guard<std::mutex> lock(g_mu);
    if (g_scratch[device]) return SS_OK;
    void *ptr = nullptr;
    SS_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) return fail(SS_ERR_HIP, "no device address for the sampling scratch array");
    g_scratch[device] = (uint64_t *)ptr;
    return SS_OK;
}

}  // namespace ss

using namespace ss;

extern "C" int ss_er_sample_noalloc(ss_context *ctx, const double *Z, int64_t n, int32_t k, const int32_t *u,
                                    const int32_t *v, const double *lval, int64_t m, double beta,
                                    uint64_t sampling_seed, const double *p_in, double *p_out, double *d_out,
                                    uint8_t *mask_out, int64_t *kept, int32_t *ku, int32_t *kv, double *kw) {
    if (!ctx) return fail(SS_ERR_INVALID, "null context");
    if (m > 0 && (!u || !v || !lval || !ku || !kv || !kw)) return fail(SS_ERR_INVALID, "null argument");
    if (!p_in && !Z) return fail(SS_ERR_INVALID, "need Z or p_in");
    if (n <= 0 || n >= INT32_MAX || k < 1 || m < 0) return fail(SS_ERR_INVALID, "bad sizes");
    if (m > SS_NOALLOC_MAX_EDGES)
        return fail(SS_ERR_INVALID, "m=%lld edges exceeds the allocation-free maximum of %lld", (long long)m,
                    (long long)SS_NOALLOC_MAX_EDGES);
    if (ctx->device < 0 || ctx->device >= kMaxDevices || !g_scratch[ctx->device])
        return fail(SS_ERR_INVALID, "the sampling scratch array of device %d is not set up", ctx->device);
    SS_HIP(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    uint64_t *sc = g_scratch[ctx->device];
    const int64_t ntiles = (m + 255) / 256;
    // the previous call's contents: the counters and this call's tile words
    // start at zero (a fill kernel, which graph capture records correctly)
    const int64_t words = kHeadWords + ntiles;
    hipLaunchKernelGGL(k_clear_words, dim3((unsigned)std::min<int64_t>((words + 255) / 256, 1024)), dim3(256), 0, s,
                       words, sc);
    SS_HIP(hipGetLastError());
    if (m) {
        ErSampleArgs<double> a;
        a.m = m;
        a.ntiles = ntiles;
        a.n = n;
        a.eu = u;
        a.ev = v;
        a.ew = lval;
        a.ew_is_lval = 1;
        a.Z = p_in ? nullptr : Z;
        a.k = k;
        a.beta = beta;
        a.ck = coin_key_from_u64(sampling_seed);
        a.p_in = p_in;
        a.p_out = p_out;
        a.d_out = p_in ? nullptr : d_out;
        a.mask = mask_out;
        a.ku = ku;
        a.kv = kv;
        a.kw = kw;
        a.ticket = (uint32_t *)sc;
        a.kept_total = (int64_t *)(sc + 1);
        a.err = (int *)(sc + 2);
        a.tile_state = sc + kHeadWords;
        a.epoch = 1;
        a.spin_limit = 1 << 22;
        launch_er_sample<double>(s, a);
        SS_HIP(hipGetLastError());
    }
    // kept total and error flag into the context's pinned words
    SS_HIP(hipMemcpyAsync(ctx->h_noalloc, sc + 1, 2 * sizeof(uint64_t), hipMemcpyDeviceToHost, s));
    SS_HIP(hipStreamSynchronize(s));
    const int err = (int)(uint32_t)ctx->h_noalloc[1];
    if (err == 2) return fail(SS_ERR_INVALID, "an edge is not u < v < n with L_uv < 0, or a probability is outside [0,1]");
    if (err) return fail(SS_ERR_TIMEOUT, "device wait exceeded its spin limit (compaction look-back)");
    if (kept) *kept = ctx->h_noalloc[0];
    return SS_OK;
}
