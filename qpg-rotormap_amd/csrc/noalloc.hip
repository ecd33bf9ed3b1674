// noalloc.hip -- rm_map_batch_noalloc: the fused encode -> GEMM -> top-k
// (-> rescore) of rm_map_batch without any device allocation, for callers
// that cannot allocate (code under HIP graph capture, tight serving loops,
// harnesses that forbid hipMalloc).  Its work buffers -- the phasor table,
// the encoded queries and the split partials -- live in `scratch`, one
// file-scope device array per GPU of kNoAllocScratchBytes, held until the
// process ends.  Every call on a GPU shares that GPU's array: calls that
// overlap on one GPU (two streams, two threads) race on it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>

#include "rm_internal.h"

namespace rm {

constexpr size_t kNoAllocScratchBytes = RM_NOALLOC_SCRATCH_BYTES;
constexpr int kNoAllocDevices = 64;
// the largest phasor table the entry point stages: 2*m*(k-s+1) int32
constexpr size_t kNoAllocTableInts = 1 << 18;

}  // namespace rm

static __device__ __align__(256) uint8_t scratch[rm::kNoAllocScratchBytes];

namespace rm {
namespace {

std::mutex g_mu;
// per device: the device address of `scratch` (hipGetSymbolAddress with that
// device current) and the (k, s, m) of the phasor table it holds
uint8_t *g_scratch[kNoAllocDevices] = {};
int g_tab_key[kNoAllocDevices][3] = {};
// host staging of a phasor table (no host allocation per call)
int32_t g_host_table[kNoAllocTableInts];

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

}  // namespace

// Once per device, with that device current (rm_ctx_create calls it): the
// device address of this GPU's copy of `scratch`.
int noalloc_setup(int device) {
    RM_CHECK(device >= 0 && device < kNoAllocDevices, RM_ERR_ARG, "device %d out of range", device);
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_scratch[device]) return RM_OK;
    void *p = nullptr;
    RM_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    RM_CHECK(p != nullptr, RM_ERR_HIP, "hipGetSymbolAddress returned NULL for the scratch array");
    g_scratch[device] = (uint8_t *)p;
    return RM_OK;
}

}  // namespace rm

using namespace rm;

extern "C" int rm_map_batch_noalloc(rm_ctx *ctx, const rm_index *idx, const uint32_t *words,
                                    int64_t n, int k, int s, int m, int c, int ktop, int kprime,
                                    int mode, float *vals, int32_t *locs, int64_t ldo,
                                    float *norms, void *stream) {
    RM_CHECK(ctx && idx && vals && locs, RM_ERR_ARG, "NULL argument");
    int rc;
    if ((rc = check_encoder(k, s, m, c, 0))) return rc;
    if ((rc = check_query(idx, ktop, kprime, mode))) return rc;
    const int64_t rdim = 2LL * m * (1LL << (2 * c));
    RM_CHECK(rdim == idx->rdim, RM_ERR_SHAPE, "encoder dim 2*m*4^c = %lld != index rdim %lld",
             (long long)rdim, (long long)idx->rdim);
    RM_CHECK(n >= 0 && ldo >= n, RM_ERR_SHAPE, "need n >= 0 and ldo >= n");
    if (n == 0) return RM_OK;
    RM_CHECK(words, RM_ERR_ARG, "words is NULL");
    const size_t tab_ints = (size_t)2 * m * (k - s + 1);
    RM_CHECK(tab_ints <= kNoAllocTableInts, RM_ERR_SHAPE,
             "phasor table of 2*m*(k-s+1) = %zu entries exceeds the no-allocation limit %zu",
             tab_ints, kNoAllocTableInts);
    const int KP = mode == RM_MODE_FP8_RESCORE ? kprime : ktop;
    const int elem = mode == RM_MODE_FP16 ? 1 : 0;
    // scratch layout: [ phasor table ][ queries ][ fp32 queries ][ partials x 2 ]
    const size_t tab_b = align256(tab_ints * 4);
    const size_t q_b = align256((size_t)n * rdim * (elem ? 2 : 1));
    const size_t q32_b = mode == RM_MODE_FP8_RESCORE ? align256((size_t)n * rdim * 4) : 0;
    const size_t fixed = tab_b + q_b + q32_b;
    const size_t min_part = align256(part_bytes(idx, n, KP, 1));
    RM_CHECK(fixed + 2 * min_part <= kNoAllocScratchBytes, RM_ERR_SHAPE,
             "%lld rows need %zu bytes of scratch; the no-allocation entry point holds %zu "
             "(about %lld rows of this configuration)",
             (long long)n, fixed + 2 * min_part, kNoAllocScratchBytes,
             (long long)((kNoAllocScratchBytes - tab_b) /
                         ((size_t)rdim * (elem ? 2 : 1) + (q32_b ? (size_t)rdim * 4 : 0) +
                          (size_t)8 * (KP + 1) + 1)));
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    RM_HIP(hipSetDevice(ctx->device));
    uint8_t *base;
    {
        std::lock_guard<std::mutex> g(g_mu);
        base = g_scratch[ctx->device];
    }
    RM_CHECK(base, RM_ERR_STATE, "no scratch address for device %d (context not set up)",
             ctx->device);
    hipStream_t st = pick_stream(ctx, stream);
    int32_t *tab = (int32_t *)base;
    uint8_t *q = base + tab_b;
    float *q32 = q32_b ? (float *)(base + tab_b + q_b) : nullptr;
    {
        // the table this device's scratch holds follows the last call's
        // (k, s, m); a new one is staged through a static host array
        std::lock_guard<std::mutex> g(g_mu);
        int *key = g_tab_key[ctx->device];
        if (key[0] != k || key[1] != s || key[2] != m) {
            build_phase_table(k, s, m, g_host_table);
            RM_HIP(hipStreamSynchronize(st));  // earlier calls on this stream read the old table
            RM_HIP(hipMemcpy(tab, g_host_table, tab_ints * 4, hipMemcpyHostToDevice));
            key[0] = k, key[1] = s, key[2] = m;
        }
    }
    EncOut o1{q, elem ? RM_F16 : RM_E4M3, rdim, 1}, o2{nullptr, 0, 0, 0};
    if (q32) o2 = EncOut{q32, RM_F32, rdim, 1};
    if ((rc = launch_encode(words, n, k, s, m, c, 0, tab, o1, o2, norms, nullptr, st))) return rc;
    // the context's plan, with at most the splits whose partials fit what is left
    Plan pl = plan_for(ctx, idx, n, mode, KP);
    const size_t part_avail = ((kNoAllocScratchBytes - fixed) / 2) & ~(size_t)255;
    pl.S = max_splits_within(idx, n, KP, pl.S, part_avail);
    const size_t pb = align256(part_bytes(idx, n, KP, pl.S));
    float *pv = (float *)(base + fixed);
    int32_t *pc = (int32_t *)(base + fixed + pb);
    return topk_core_bufs(ctx, idx, q, rdim * (elem ? 2 : 1), elem, q32, rdim, n, ktop, kprime, mode,
                          vals, locs, ldo, pl, pv, pc, pb, st);
}
