// capi.cpp -- the C ABI of librotormap_hip.so: argument checking, context /
// index lifetime, workspaces, and the host orchestration of the kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "rm_internal.h"

#define RM_VERSION "rotormap_hip 0.1.0 (gfx950)"

namespace rm {

static thread_local std::string g_err;

int fail(int code, const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}
void clear_error() { g_err.clear(); }

// NULL is the device's default (null) stream, as in the HIP runtime API; the
// context's own stream only carries the library's internal uploads.
hipStream_t pick_stream(rm_ctx *, void *stream) { return (hipStream_t)stream; }

int ws_get(rm_ctx *ctx, const char *name, size_t bytes, void **out) {
    rm_buf &b = ctx->ws[name];
    if (b.n < bytes) {
        if (b.p) {
            RM_HIP(hipDeviceSynchronize());
            RM_HIP(hipFree(b.p));
            b.p = nullptr;
            b.n = 0;
        }
        size_t want = std::max(bytes, (size_t)256);
        hipError_t e = hipMalloc(&b.p, want);
        if (e != hipSuccess) {
            b.p = nullptr;
            return fail(RM_ERR_NOMEM, "workspace '%s': hipMalloc(%zu) failed: %s", name, want,
                        hipGetErrorString(e));
        }
        b.n = want;
    }
    *out = b.p;
    return RM_OK;
}

int table_get(rm_ctx *ctx, int k, int s, int m, const int32_t **out) {
    auto key = std::make_tuple(k, s, m);
    auto it = ctx->tables.find(key);
    if (it != ctx->tables.end()) {
        *out = it->second;
        return RM_OK;
    }
    size_t cnt = (size_t)2 * m * (k - s + 1);
    std::vector<int32_t> h(cnt);
    build_phase_table(k, s, m, h.data());
    int32_t *d = nullptr;
    RM_HIP(hipMalloc(&d, cnt * 4));
    RM_HIP(hipMemcpy(d, h.data(), cnt * 4, hipMemcpyHostToDevice));
    ctx->tables[key] = d;
    ctx->table_bytes += cnt * 4;
    *out = d;
    return RM_OK;
}

// Column splits for M query rows over N index columns: as many (row tile,
// split) items as fit one wave of resident workgroups (one per CU: the fused
// kernel holds a CU's LDS) -- rounding up would leave a second, mostly idle
// wave -- never more splits than column tiles; the autotuned plan overrides
// per M bucket.
// Splits of a batch with no tuned plan: at least enough items to give every
// CU one (S >= CUs / R), then doubled until the query tiles worked on at once
// -- min(R, CUs / S) tiles of BR rows x row_bytes -- fit in half of the
// 256 MB MALL.  Each workgroup re-reads its query tile once per column tile;
// a working set past the MALL sends those re-reads to HBM (65,536 fp16
// queries of 4 KB: 256 MB at S = 1, 19 % slower than S = 2 at 3.1 M columns;
// the e4m3 batch, 128 MB, is fastest at S = 1).
int default_splits(rm_ctx *ctx, int64_t M, int64_t N, int bq, int64_t row_bytes) {
    const int64_t BR = bq;
    const int64_t R = cdiv(M, BR), tiles = cdiv(N, kBI), cus = ctx->num_cus;
    const int64_t cap = std::min<int64_t>(256, tiles);
    int64_t S = std::max<int64_t>(1, std::min<int64_t>(cus / R, cap));
    while (S < cap && std::min<int64_t>(R, cdiv(cus, S)) * BR * row_bytes > kQueryWorkingSet)
        S = std::min<int64_t>(2 * S, cap);
    return (int)S;
}

// column tiles a split should sweep before the built-in plan prefers a
// smaller query tile
constexpr int64_t kMinSplitTiles = 32;

static int bucket_of(int64_t M) {
    int b = 0;
    while (((int64_t)1 << b) < M) ++b;
    return b;  // ceil(log2 M)
}

// Query rows per tile without a tuned plan: the smallest ring tile that
// holds the batch up to 128 rows (a small batch wastes less MFMA work on
// padded query rows), then 256; lists longer than 32 have only the 128
// tile.  Above 128 rows the 256-row tile with one row tile and 256 splits
// beats two 128-row tiles (human scale: 1.94 / 2.07 / 2.29 ms at 160 / 192
// / 256 reads against 2.69 / 2.47 / 2.41 ms).  plan_for may step it down
// for short splits.
static int default_tile(int64_t M, int kp) {
    if (!ring_list(kp)) return 128;
    return M <= 64 ? 64 : M <= 128 ? 128 : 256;
}

// Tile plan for a batch: (query rows per tile, splits, query tiles per XCD
// group).  The built-in rule is default_tile + default_splits with one group
// over all query tiles; the autotuned plan (rm_autotune / rm_plan_set)
// overrides it per (mode, size bucket).  A tuned tile a list length cannot
// run falls back to that length's tile.
Plan plan_for(rm_ctx *ctx, const rm_index *idx, int64_t M, int mode, int kp) {
    const int64_t rb = idx->rdim * (mode == RM_MODE_FP16 ? 2 : 1);
    const int64_t tiles = cdiv(idx->N, kBI);
    Plan p;
    auto it = ctx->plan.find({mode, bucket_of(M)});
    if (it != ctx->plan.end()) {
        p = it->second;
        if (p.bq == 0 || !tile_rows_ok(kp, p.bq)) p.bq = default_tile(M, kp);
        if (p.S <= 0) p.S = default_splits(ctx, M, idx->N, p.bq, rb);
        p.S = (int)std::max<int64_t>(1, std::min<int64_t>(p.S, tiles));
        return p;
    }
    p.bq = default_tile(M, kp);
    p.S = default_splits(ctx, M, idx->N, p.bq, rb);
    // Each split warms its lists up from empty, so a split that sweeps only
    // a few column tiles spends most of its time in merges: below
    // kMinSplitTiles per split, halve the query tile (more row tiles, fewer
    // splits).  Measured at 100 Mbp (391 column tiles): 1,024 reads 0.52 ms
    // with 64-row tiles against 0.78 ms with 256 (6 tiles per split), 4,096
    // reads 1.31 against 1.51 ms with 128-row tiles.
    while (ring_list(kp) && p.bq > 64 && cdiv(tiles, p.S) < kMinSplitTiles) {
        p.bq /= 2;
        p.S = default_splits(ctx, M, idx->N, p.bq, rb);
    }
    // With splits and at least an XCD's worth of query tiles, group them so
    // that each XCD's resident workgroups (CUs / 8) hold whole groups of
    // CUs / 8 / S query tiles x S splits: those query tiles stay in the XCD's
    // L2 instead of being re-read from the MALL for every column tile.  One
    // group over all query tiles gave an XCD 32 tiles of one split (human
    // scale: 8,192 rows 45.6 ms against 41.2 ms grouped, 16,384 rows 87.9
    // against 81.0, 32,768 rows 171.5 against 165.3).
    const int64_t xcd_items = std::max(1, ctx->num_cus / 8);
    p.qg = (p.S > 1 && cdiv(M, p.bq) >= xcd_items) ? (int)std::max<int64_t>(1, xcd_items / p.S) : 0;
    return p;
}

// Rows per launch for a batch without a tuned plan, or 0 for one launch.
// Above 64 x CUs rows (16,384 on 256 CUs) a batch runs as equal launches of
// at most that many rows: each is one wave of CUs / 4 query tiles x 4 splits
// with L2-resident query tiles (grouped as above), and every launch starts
// all workgroups together.  One launch of a larger batch gives each
// workgroup several items, whose drifting starts mix query groups on an XCD
// (human scale, alternated: a 65,536-read step 319.0 ms as four launches
// against 334.9 ms as one, with the same rows).  A split must still sweep a
// long column range, or the launches add list warm-up (kChunkSplitTiles).
constexpr int64_t kChunkSplitTiles = 1024;
int64_t launch_rows(rm_ctx *ctx, const rm_index *idx, int64_t M, int mode, int kp) {
    const int64_t L = 64LL * ctx->num_cus;
    if (!ring_list(kp) || M <= L) return 0;
    if (ctx->plan.count({mode, bucket_of(M)})) return 0;  // a tuned plan runs as one launch
    const int64_t n = cdiv(M, L), rows = std::min<int64_t>(L, cdiv(cdiv(M, n), 256) * 256);
    const Plan pc = plan_for(ctx, idx, rows, mode, kp);
    if (cdiv(cdiv(idx->N, kBI), pc.S) < kChunkSplitTiles) return 0;
    return rows;
}

static int clamp_splits(const rm_index *idx, int S) {
    return (int)std::max<int64_t>(1, std::min<int64_t>({(int64_t)S, cdiv(idx->N, kBI), (int64_t)256}));
}

// split partials (M x S x KP) and, after them, M shared bounds
size_t part_bytes(const rm_index *idx, int64_t M, int KP, int S) {
    return ((size_t)M * clamp_splits(idx, S) * KP + (size_t)M) * 4;
}

int max_splits_within(const rm_index *idx, int64_t M, int KP, int S, size_t pbytes) {
    S = clamp_splits(idx, S);
    if (M <= 0) return S;
    const int64_t fit = ((int64_t)(pbytes / 4) - M) / (M * (int64_t)KP);
    return (int)std::max<int64_t>(1, std::min<int64_t>(S, fit));
}

int topk_core(rm_ctx *ctx, const rm_index *idx, const void *Q, int64_t ldq_bytes, int elem,
              const float *Q32, int64_t ldq32, int64_t M, int ktop, int kprime, int mode,
              float *vals, int32_t *locs, int64_t ldo, const Plan &pl, hipStream_t st) {
    const int KP = (mode == RM_MODE_FP8_RESCORE) ? kprime : ktop;
    if (M == 0) return RM_OK;
    const size_t pbytes = part_bytes(idx, M, KP, pl.S);
    void *pv = nullptr, *pc = nullptr;
    int rc;
    if ((rc = ws_get(ctx, "part_v", pbytes, &pv))) return rc;
    if ((rc = ws_get(ctx, "part_c", pbytes, &pc))) return rc;
    return topk_core_bufs(ctx, idx, Q, ldq_bytes, elem, Q32, ldq32, M, ktop, kprime, mode, vals,
                          locs, ldo, pl, (float *)pv, (int32_t *)pc, pbytes, st);
}

// A batch under the built-in launch rule: launch_rows, then plan_for per
// launch.  Rows are independent, so launches of consecutive row ranges give
// the rows of one launch.
int topk_batched(rm_ctx *ctx, const rm_index *idx, const void *Q, int64_t ldq_bytes, int elem,
                 const float *Q32, int64_t ldq32, int64_t M, int ktop, int kprime, int mode,
                 float *vals, int32_t *locs, int64_t ldo, hipStream_t st) {
    const int kp = mode == RM_MODE_FP8_RESCORE ? kprime : ktop;
    const int64_t rows = launch_rows(ctx, idx, M, mode, kp);
    if (rows == 0)
        return topk_core(ctx, idx, Q, ldq_bytes, elem, Q32, ldq32, M, ktop, kprime, mode, vals, locs,
                         ldo, plan_for(ctx, idx, M, mode, kp), st);
    const Plan pl = plan_for(ctx, idx, rows, mode, kp);
    for (int64_t r0 = 0; r0 < M; r0 += rows) {
        const int64_t Mc = std::min(rows, M - r0);
        int rc = topk_core(ctx, idx, (const uint8_t *)Q + r0 * ldq_bytes, ldq_bytes, elem,
                           Q32 ? Q32 + r0 * ldq32 : nullptr, ldq32, Mc, ktop, kprime, mode, vals + r0,
                           locs + r0, ldo, pl, st);
        if (rc) return rc;
    }
    return RM_OK;
}

int topk_core_bufs(rm_ctx *ctx, const rm_index *idx, const void *Q, int64_t ldq_bytes, int elem,
                   const float *Q32, int64_t ldq32, int64_t M, int ktop, int kprime, int mode,
                   float *vals, int32_t *locs, int64_t ldo, const Plan &pl, float *pv,
                   int32_t *pc, size_t pbytes, hipStream_t st) {
    const int KP = (mode == RM_MODE_FP8_RESCORE) ? kprime : ktop;
    const int64_t K = idx->rdim;
    const int esz = elem == 0 ? 1 : 2;
    const uint8_t *db = elem == 0 ? idx->b8 : (const uint8_t *)idx->b16;
    RM_CHECK(db, RM_ERR_STATE, "index has no %s copy (create it with RM_INDEX_FP16)",
             elem == 0 ? "e4m3" : "fp16");
    if (M == 0) return RM_OK;
    const int64_t tiles = cdiv(idx->N, kBI);
    const int S = clamp_splits(idx, pl.S);
    const int bq = pl.bq;
    RM_CHECK(tile_rows_ok(KP, bq), RM_ERR_ARG, "no fused kernel for %d queries per tile at k' = %d",
             bq, KP);
    RM_CHECK(part_bytes(idx, M, KP, S) <= pbytes, RM_ERR_STATE,
             "split partials need %zu bytes, workspace holds %zu", part_bytes(idx, M, KP, S), pbytes);
    int rc;
    GemmArgs a{};
    a.db = db;
    a.ldb = K * esz;
    a.q = (const uint8_t *)Q;
    a.ldq = ldq_bytes;
    a.M = M;
    a.BQ = bq;
    a.N = idx->N;
    a.Kb = (int)(K * esz);
    a.S = S;
    a.R = cdiv(M, bq);
    a.qg = (pl.qg <= 0 || pl.qg > a.R) ? (int)a.R : pl.qg;
    a.ngroups = (int)cdiv(a.R, a.qg);
    a.tiles = tiles;
    a.col_base = 0;
    a.part_v = (float *)pv;
    a.part_c = (int32_t *)pc;
    a.KP = KP;
    a.persist_grid = ctx->persistent ? ctx->num_cus : 0;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (ctx->prof) {
        RM_HIP(hipEventCreate(&e0));
        RM_HIP(hipEventCreate(&e1));
        RM_HIP(hipEventRecord(e0, st));
    }
    if ((rc = launch_gemm_topk(a, elem, st))) return rc;
    if (ctx->prof) {
        RM_HIP(hipEventRecord(e1, st));
        ctx->prof_ev.push_back({e0, e1});
        ctx->prof_flop += 2.0 * (double)M * (double)K * (double)idx->N;
    }
    const bool r16 = idx->rescore == RM_F16 && idx->b16;
    return launch_merge((const float *)pv, (const int32_t *)pc, S, KP, M, idx->N, ktop, mode, Q32,
                        ldq32, r16 ? (const void *)idx->b16 : (const void *)idx->b8,
                        r16 ? RM_F16 : RM_E4M3, K, (int)K, vals, locs, ldo, st);
}

int check_encoder(int k, int s, int m, int c, int normalize) {
    RM_CHECK(c >= 1 && c <= s && s <= 16, RM_ERR_ARG,
             "encoder needs 1 <= c <= s <= 16 (got s = %d, c = %d)", s, c);
    RM_CHECK(s <= k, RM_ERR_ARG, "k = %d < s = %d: no s-mer windows", k, s);
    RM_CHECK(m >= 1 && m <= 64, RM_ERR_ARG, "m must be in 1..64 (got %d)", m);
    RM_CHECK(2LL * m * (1LL << (2 * c)) <= 16384, RM_ERR_ARG, "2*m*4^c = %lld exceeds 16384",
             2LL * m * (1LL << (2 * c)));
    RM_CHECK(normalize >= 0 && normalize <= 3, RM_ERR_ARG, "normalize must be 0..3 (got %d)",
             normalize);
    return RM_OK;
}

static bool aligned16(const void *p) { return ((uintptr_t)p & 15u) == 0; }

}  // namespace rm

using namespace rm;

extern "C" {

const char *rm_version(void) { return RM_VERSION; }
const char *rm_last_error(void) { return g_err.c_str(); }

int rm_ctx_create(int device, rm_ctx **out) {
    RM_CHECK(out, RM_ERR_ARG, "out is NULL");
    int ndev = 0;
    RM_HIP(hipGetDeviceCount(&ndev));
    RM_CHECK(device >= 0 && device < ndev, RM_ERR_ARG, "device %d out of range (%d devices)",
             device, ndev);
    RM_HIP(hipSetDevice(device));
    int rc;
    if ((rc = noalloc_setup(device))) return rc;  // the scratch array's address on this device
    rm_ctx *c = new rm_ctx();
    c->device = device;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->num_cus = prop.multiProcessorCount;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return fail(RM_ERR_HIP, "hipStreamCreateWithFlags failed");
    }
    *out = c;
    return RM_OK;
}

int rm_ctx_destroy(rm_ctx *ctx) {
    if (!ctx) return RM_OK;
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    for (auto &kv : ctx->tables) hipFree(kv.second);
    for (auto &kv : ctx->ws)
        if (kv.second.p) hipFree(kv.second.p);
    hipStreamDestroy(ctx->stream);
    delete ctx;
    return RM_OK;
}

int rm_sync(rm_ctx *ctx, void *stream) {
    RM_CHECK(ctx, RM_ERR_ARG, "ctx is NULL");
    RM_HIP(hipStreamSynchronize(pick_stream(ctx, stream)));
    return RM_OK;
}

int64_t rm_ctx_bytes(rm_ctx *ctx) {
    if (!ctx) return 0;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    int64_t s = (int64_t)ctx->table_bytes;
    for (auto &kv : ctx->ws) s += (int64_t)kv.second.n;
    return s;
}

int rm_phase_table(int k, int s, int m, int32_t *host_out) {
    RM_CHECK(host_out, RM_ERR_ARG, "host_out is NULL");
    RM_CHECK(k >= 1 && s >= 1 && s <= k && m >= 1 && m <= 64, RM_ERR_ARG, "bad (k, s, m)");
    build_phase_table(k, s, m, host_out);
    return RM_OK;
}

int rm_encode_batch(rm_ctx *ctx, const uint32_t *words, int64_t n, int k, int s, int m, int c,
                    int normalize, void *out1, int dtype1, int64_t sf1, int64_t sd1, void *out2,
                    int dtype2, int64_t sf2, int64_t sd2, float *norms, int32_t *counts,
                    void *stream) {
    RM_CHECK(ctx, RM_ERR_ARG, "ctx is NULL");
    int rc;
    if ((rc = check_encoder(k, s, m, c, normalize))) return rc;
    RM_CHECK(n >= 0, RM_ERR_ARG, "n < 0");
    if (n == 0) return RM_OK;
    RM_CHECK(words && out1, RM_ERR_ARG, "words / out1 is NULL");
    RM_CHECK(dtype1 >= RM_F32 && dtype1 <= RM_E4M3, RM_ERR_ARG, "bad dtype1 %d", dtype1);
    RM_CHECK(!out2 || (dtype2 >= RM_F32 && dtype2 <= RM_E4M3), RM_ERR_ARG, "bad dtype2 %d", dtype2);
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    RM_HIP(hipSetDevice(ctx->device));
    const int32_t *tab;
    if ((rc = table_get(ctx, k, s, m, &tab))) return rc;
    EncOut o1{out1, dtype1, sf1, sd1}, o2{out2, dtype2, sf2, sd2};
    return launch_encode(words, n, k, s, m, c, normalize, tab, o1, o2, norms, counts,
                         pick_stream(ctx, stream));
}

int rm_revcomp_words(rm_ctx *ctx, uint32_t *rc, const uint32_t *words, int64_t n, int k,
                     void *stream) {
    RM_CHECK(ctx && rc && words, RM_ERR_ARG, "NULL argument");
    RM_CHECK(k >= 1 && n >= 0, RM_ERR_ARG, "bad k / n");
    RM_CHECK(rc != words, RM_ERR_ARG, "in-place reverse complement is not supported");
    return launch_revcomp(rc, words, n, k, pick_stream(ctx, stream));
}

int rm_extract_windows(rm_ctx *ctx, uint32_t *out, const uint32_t *rec_words,
                       const int64_t *starts, int64_t n, int k, void *stream) {
    RM_CHECK(ctx && out && rec_words && starts, RM_ERR_ARG, "NULL argument");
    RM_CHECK(k >= 1 && n >= 0, RM_ERR_ARG, "bad k / n");
    return launch_extract(out, rec_words, starts, n, k, pick_stream(ctx, stream));
}

int rm_fasta_reads(rm_ctx *ctx, const uint8_t *raw, int64_t n_bytes, int k, uint32_t *words,
                   int64_t max_reads, int64_t *hdr_off, int32_t *hdr_len, int64_t *n_reads,
                   void *stream) {
    RM_CHECK(ctx && n_reads, RM_ERR_ARG, "NULL argument");
    *n_reads = 0;
    RM_CHECK(k >= 1 && k <= 262144, RM_ERR_ARG, "k must be in 1..262144 (got %d)", k);
    RM_CHECK(n_bytes >= 0 && max_reads >= 0, RM_ERR_ARG, "negative size");
    RM_CHECK(n_bytes == 0 || raw, RM_ERR_ARG, "raw is NULL");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    RM_HIP(hipSetDevice(ctx->device));
    return fasta_reads(ctx, raw, n_bytes, k, words, max_reads, hdr_off, hdr_len, n_reads,
                       pick_stream(ctx, stream));
}

int rm_fasta_reads_host(rm_ctx *ctx, const uint8_t *raw, int64_t n_bytes, int k,
                        uint32_t *words, int64_t max_reads, int64_t *hdr_off, int32_t *hdr_len,
                        int64_t *n_reads) {
    RM_CHECK(ctx && n_reads, RM_ERR_ARG, "NULL argument");
    *n_reads = 0;
    RM_CHECK(k >= 1 && k <= 262144, RM_ERR_ARG, "k must be in 1..262144 (got %d)", k);
    RM_CHECK(n_bytes >= 0 && max_reads >= 0, RM_ERR_ARG, "negative size");
    RM_CHECK(n_bytes == 0 || raw, RM_ERR_ARG, "raw is NULL");
    if (n_bytes == 0) return RM_OK;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    RM_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const int64_t W = (k + 15) / 16;
    void *draw, *dw, *doff, *dlen;
    int rc;
    if ((rc = ws_get(ctx, "fah_raw", (size_t)n_bytes, &draw))) return rc;
    if ((rc = ws_get(ctx, "fah_words", (size_t)std::max<int64_t>(1, max_reads) * W * 4, &dw))) return rc;
    if ((rc = ws_get(ctx, "fah_off", (size_t)std::max<int64_t>(1, max_reads) * 8, &doff))) return rc;
    if ((rc = ws_get(ctx, "fah_len", (size_t)std::max<int64_t>(1, max_reads) * 4, &dlen))) return rc;
    RM_HIP(hipMemcpyAsync(draw, raw, (size_t)n_bytes, hipMemcpyHostToDevice, st));
    rc = fasta_reads(ctx, (const uint8_t *)draw, n_bytes, k, (uint32_t *)dw, max_reads,
                     (int64_t *)doff, (int32_t *)dlen, n_reads, st);
    if (rc) return rc;
    const int64_t n = *n_reads;
    if (n > 0 && words)
        RM_HIP(hipMemcpyAsync(words, dw, (size_t)n * W * 4, hipMemcpyDeviceToHost, st));
    if (n > 0 && hdr_off) RM_HIP(hipMemcpyAsync(hdr_off, doff, (size_t)n * 8, hipMemcpyDeviceToHost, st));
    if (n > 0 && hdr_len) RM_HIP(hipMemcpyAsync(hdr_len, dlen, (size_t)n * 4, hipMemcpyDeviceToHost, st));
    RM_HIP(hipStreamSynchronize(st));
    return RM_OK;
}

int rm_quantize_e4m3(rm_ctx *ctx, uint8_t *dst, const void *src, int src_dtype, int64_t n,
                     void *stream) {
    RM_CHECK(ctx && dst && src, RM_ERR_ARG, "NULL argument");
    RM_CHECK(src_dtype == RM_F32 || src_dtype == RM_F16, RM_ERR_ARG, "src_dtype must be F32 or F16");
    return launch_quantize(dst, src, src_dtype, n, pick_stream(ctx, stream));
}

int rm_dequantize_e4m3(rm_ctx *ctx, float *dst, const uint8_t *src, int64_t n, void *stream) {
    RM_CHECK(ctx && dst && src, RM_ERR_ARG, "NULL argument");
    return launch_dequantize(dst, src, n, pick_stream(ctx, stream));
}

int rm_gemm_chunk(rm_ctx *ctx, void *C, int c_dtype, int64_t ldc, const void *A,
                  int64_t a_stride_row, int64_t a_stride_k, const void *B, int64_t ldb, int ab_dtype,
                  int64_t M, int64_t wi, int64_t K, void *stream) {
    RM_CHECK(ctx && C && A && B, RM_ERR_ARG, "NULL argument");
    RM_CHECK(c_dtype == RM_F32 || c_dtype == RM_F16, RM_ERR_ARG, "c_dtype must be F32 or F16");
    RM_CHECK(ab_dtype == RM_E4M3 || ab_dtype == RM_F16, RM_ERR_ARG, "ab_dtype must be E4M3 or F16");
    RM_CHECK(M >= 0 && wi >= 0 && K > 0, RM_ERR_ARG, "bad M / wi / K");
    const int esz = ab_dtype == RM_E4M3 ? 1 : 2;
    RM_CHECK((K * esz) % kBK == 0, RM_ERR_SHAPE, "K*elemsize = %lld must be a multiple of 64",
             (long long)(K * esz));
    RM_CHECK(ldb >= K && (ldb * esz) % 16 == 0 && aligned16(B), RM_ERR_SHAPE,
             "B needs ldb >= K, 16-byte aligned columns");
    RM_CHECK(ldc >= M, RM_ERR_SHAPE, "ldc < M");
    if (M == 0 || wi == 0) return RM_OK;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    hipStream_t st = pick_stream(ctx, stream);
    const uint8_t *q = (const uint8_t *)A;
    int64_t ldq = a_stride_row * esz;
    if (!(a_stride_k == 1 && (ldq % 16) == 0 && aligned16(A))) {
        void *t;
        int rc;
        if ((rc = ws_get(ctx, "a_rowmajor", (size_t)(M * K * esz), &t))) return rc;
        if ((rc = launch_transpose_u8((uint8_t *)t, (const uint8_t *)A, M, K, a_stride_row,
                                      a_stride_k, esz, st)))
            return rc;
        q = (const uint8_t *)t;
        ldq = K * esz;
    }
    GemmArgs a{};
    a.db = (const uint8_t *)B;
    a.ldb = ldb * esz;
    a.q = q;
    a.ldq = ldq;
    a.M = M;
    a.BQ = 256;
    a.N = wi;
    a.Kb = (int)(K * esz);
    a.R = cdiv(M, 256);
    a.qg = (int)a.R;
    a.ngroups = 1;
    a.tiles = cdiv(wi, kBI);
    a.S = (int)a.tiles;
    a.C = C;
    a.c_dtype = c_dtype;
    a.ldc = ldc;
    a.persist_grid = ctx->persistent ? ctx->num_cus : 0;
    return launch_gemm_store(a, ab_dtype == RM_E4M3 ? 0 : 1, st);
}

int rm_rowtopk_merge(rm_ctx *ctx, float *D_val, int32_t *D_loc, int64_t ldd, const void *C,
                     int c_dtype, int64_t ldc, int64_t M, int64_t wi, int64_t offset, int k,
                     void *stream) {
    RM_CHECK(ctx && D_val && D_loc && C, RM_ERR_ARG, "NULL argument");
    RM_CHECK(c_dtype >= RM_F32 && c_dtype <= RM_E4M3, RM_ERR_ARG, "bad c_dtype");
    RM_CHECK(k >= 1 && k <= 64, RM_ERR_ARG, "k must be 1..64");
    RM_CHECK(wi >= 1, RM_ERR_ARG, "chunk must be non-empty");
    RM_CHECK(ldd >= M && ldc >= M, RM_ERR_SHAPE, "leading dimensions < M");
    RM_CHECK(offset >= 0 && offset + wi <= ((int64_t)1 << 31) - 1, RM_ERR_ARG,
             "column ids offset + wi must fit int32");
    if (M == 0) return RM_OK;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    const int G = rowtopk_segments(M, wi);
    void *pv, *pc;
    int rc;
    if ((rc = ws_get(ctx, "rowtopk_v", (size_t)M * G * k * 4, &pv))) return rc;
    if ((rc = ws_get(ctx, "rowtopk_c", (size_t)M * G * k * 4, &pc))) return rc;
    void *seed;
    if ((rc = ws_get(ctx, "rowtopk_seed", (size_t)M * 4, &seed))) return rc;
    return launch_rowtopk(D_val, D_loc, ldd, C, c_dtype, ldc, M, wi, offset, k, G, (float *)pv,
                          (int32_t *)pc, (int32_t *)seed, pick_stream(ctx, stream));
}

int rm_index_create(rm_ctx *ctx, const uint8_t *B, int64_t rdim, int64_t N, int flags,
                    rm_index **out) {
    RM_CHECK(ctx && B && out, RM_ERR_ARG, "NULL argument");
    RM_CHECK(rdim > 0 && rdim % 64 == 0, RM_ERR_SHAPE, "rdim = %lld must be a positive multiple of 64",
             (long long)rdim);
    RM_CHECK(N >= 1 && N < ((int64_t)1 << 31) - 1, RM_ERR_SHAPE, "N = %lld out of range",
             (long long)N);
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    RM_HIP(hipSetDevice(ctx->device));
    rm_index *ix = new rm_index();
    ix->ctx = ctx;
    ix->rdim = rdim;
    ix->N = N;
    size_t bytes = (size_t)rdim * N;
    if (hipMalloc(&ix->b8, bytes) != hipSuccess) {
        delete ix;
        return fail(RM_ERR_NOMEM, "index: hipMalloc(%zu) failed", bytes);
    }
    hipError_t e = hipMemcpyAsync(ix->b8, B, bytes,
                                  (flags & RM_INDEX_HOST) ? hipMemcpyHostToDevice
                                                          : hipMemcpyDeviceToDevice,
                                  ctx->stream);
    if (e == hipSuccess && (flags & RM_INDEX_FP16)) {
        if (hipMalloc(&ix->b16, bytes * 2) != hipSuccess) {
            hipFree(ix->b8);
            delete ix;
            return fail(RM_ERR_NOMEM, "index fp16 copy: hipMalloc(%zu) failed", bytes * 2);
        }
        int rc = launch_widen_f16(ix->b16, ix->b8, (int64_t)bytes, ctx->stream);
        if (rc) { hipFree(ix->b8); hipFree(ix->b16); delete ix; return rc; }
    }
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        hipFree(ix->b8);
        if (ix->b16) hipFree(ix->b16);
        delete ix;
        return fail(RM_ERR_HIP, "index upload failed: %s", hipGetErrorString(e));
    }
    *out = ix;
    return RM_OK;
}

int rm_index_create_f16(rm_ctx *ctx, const uint16_t *B16, int64_t rdim, int64_t N, int flags,
                        rm_index **out) {
    RM_CHECK(ctx && B16 && out, RM_ERR_ARG, "NULL argument");
    RM_CHECK(rdim > 0 && rdim % 64 == 0, RM_ERR_SHAPE, "rdim = %lld must be a positive multiple of 64",
             (long long)rdim);
    RM_CHECK(N >= 1 && N < ((int64_t)1 << 31) - 1, RM_ERR_SHAPE, "N = %lld out of range",
             (long long)N);
    RM_CHECK((flags & ~RM_INDEX_HOST) == 0, RM_ERR_ARG, "flags: only RM_INDEX_HOST applies");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    RM_HIP(hipSetDevice(ctx->device));
    rm_index *ix = new rm_index();
    ix->ctx = ctx;
    ix->rdim = rdim;
    ix->N = N;
    ix->verbatim16 = true;
    const size_t n = (size_t)rdim * N;
    if (hipMalloc(&ix->b16, n * 2) != hipSuccess || hipMalloc(&ix->b8, n) != hipSuccess) {
        if (ix->b16) hipFree(ix->b16);
        delete ix;
        return fail(RM_ERR_NOMEM, "fp16 index: hipMalloc(%zu + %zu) failed", n * 2, n);
    }
    hipError_t e = hipMemcpyAsync(ix->b16, B16, n * 2,
                                  (flags & RM_INDEX_HOST) ? hipMemcpyHostToDevice
                                                          : hipMemcpyDeviceToDevice,
                                  ctx->stream);
    // the e4m3 database of the same columns: e4m3(fp32(fp16)), the
    // quantisation index_to_f8 applies (RotorMap/index/load.jl:76-89)
    int rc = e == hipSuccess ? launch_quantize(ix->b8, ix->b16, RM_F16, (int64_t)n, ctx->stream) : 0;
    if (e == hipSuccess && rc == RM_OK) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess || rc != RM_OK) {
        hipFree(ix->b8);
        hipFree(ix->b16);
        delete ix;
        return rc ? rc : fail(RM_ERR_HIP, "fp16 index upload failed: %s", hipGetErrorString(e));
    }
    *out = ix;
    return RM_OK;
}

int rm_index_destroy(rm_index *idx) {
    if (!idx) return RM_OK;
    hipSetDevice(idx->ctx->device);
    hipDeviceSynchronize();
    if (idx->b8) hipFree(idx->b8);
    if (idx->b16) hipFree(idx->b16);
    delete idx;
    return RM_OK;
}

int64_t rm_index_cols(const rm_index *idx) { return idx ? idx->N : 0; }
int64_t rm_index_rdim(const rm_index *idx) { return idx ? idx->rdim : 0; }
const uint8_t *rm_index_e4m3(const rm_index *idx) { return idx ? idx->b8 : nullptr; }
const void *rm_index_f16(const rm_index *idx) { return idx ? idx->b16 : nullptr; }

int rm_index_set_rescore(rm_index *idx, int dtype) {
    RM_CHECK(idx, RM_ERR_ARG, "NULL index");
    RM_CHECK(dtype == RM_E4M3 || dtype == RM_F16, RM_ERR_ARG,
             "rescoring columns must be RM_E4M3 or RM_F16 (got %d)", dtype);
    RM_CHECK(dtype == RM_E4M3 || idx->b16, RM_ERR_STATE,
             "RM_F16 rescoring needs fp16 columns (rm_index_create_f16 or RM_INDEX_FP16)");
    std::lock_guard<std::recursive_mutex> lk(idx->ctx->mu);
    idx->rescore = dtype;
    return RM_OK;
}
int rm_index_rescore(const rm_index *idx) { return idx ? idx->rescore : -1; }

}  // extern "C"

namespace rm {

int check_query(const rm_index *idx, int ktop, int kprime, int mode) {
    RM_CHECK(mode == RM_MODE_FP8 || mode == RM_MODE_FP8_RESCORE || mode == RM_MODE_FP16,
             RM_ERR_ARG, "bad mode %d", mode);
    RM_CHECK(ktop >= 1 && ktop <= kMaxKP, RM_ERR_ARG, "ktop must be 1..%d (got %d)", kMaxKP, ktop);
    if (mode == RM_MODE_FP8_RESCORE)
        RM_CHECK(kprime >= ktop && kprime <= kMaxKP, RM_ERR_ARG,
                 "kprime must be in ktop..%d (got %d)", kMaxKP, kprime);
    RM_CHECK(ktop <= idx->N, RM_ERR_ARG, "ktop = %d > N = %lld", ktop, (long long)idx->N);
    if (mode == RM_MODE_FP16) RM_CHECK(idx->b16, RM_ERR_STATE, "index has no fp16 copy");
    return RM_OK;
}

}  // namespace rm

extern "C" {

int rm_topk_batch(rm_ctx *ctx, const rm_index *idx, const void *Q, int64_t ldq, const float *Q32,
                  int64_t ldq32, int64_t M, int ktop, int kprime, int mode, float *vals,
                  int32_t *locs, int64_t ldo, void *stream) {
    RM_CHECK(ctx && idx && vals && locs, RM_ERR_ARG, "NULL argument");
    int rc;
    if ((rc = check_query(idx, ktop, kprime, mode))) return rc;
    RM_CHECK(M >= 0 && ldo >= M, RM_ERR_SHAPE, "need M >= 0 and ldo >= M");
    if (M == 0) return RM_OK;
    RM_CHECK(Q, RM_ERR_ARG, "Q is NULL");
    const int elem = mode == RM_MODE_FP16 ? 1 : 0;
    const int64_t ldq_b = ldq * (elem ? 2 : 1);
    RM_CHECK(ldq >= idx->rdim && ldq_b % 16 == 0 && aligned16(Q), RM_ERR_SHAPE,
             "Q rows must be >= rdim elements, 16-byte aligned");
    if (mode == RM_MODE_FP8_RESCORE) RM_CHECK(Q32 && ldq32 >= idx->rdim, RM_ERR_ARG, "Q32 required");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    RM_HIP(hipSetDevice(ctx->device));
    return topk_batched(ctx, idx, Q, ldq_b, elem, Q32, ldq32, M, ktop, kprime, mode, vals, locs, ldo,
                        pick_stream(ctx, stream));
}

// encode n fragments straight into the query layouts the GEMM reads
}  // extern "C"

namespace rm {

// encode n fragments into the query layouts the GEMM reads: q (n x rdim,
// e4m3 or fp16) and, for rescoring, q32 (n x rdim fp32)
int encode_into(rm_ctx *ctx, const uint32_t *words, int64_t n, int k, int s, int m, int c,
                int mode, int64_t rdim, float *norms, void *q, float *q32, hipStream_t st) {
    int rc;
    const int32_t *tab;
    if ((rc = table_get(ctx, k, s, m, &tab))) return rc;
    const int elem = mode == RM_MODE_FP16 ? 1 : 0;
    EncOut o1{q, elem ? RM_F16 : RM_E4M3, rdim, 1}, o2{nullptr, 0, 0, 0};
    if (mode == RM_MODE_FP8_RESCORE) o2 = EncOut{q32, RM_F32, rdim, 1};
    return launch_encode(words, n, k, s, m, c, 0, tab, o1, o2, norms, nullptr, st);
}

static int encode_queries(rm_ctx *ctx, const uint32_t *words, int64_t n, int k, int s, int m,
                          int c, int mode, int64_t rdim, float *norms, void **q, float **q32,
                          hipStream_t st) {
    int rc;
    const int elem = mode == RM_MODE_FP16 ? 1 : 0;
    if ((rc = ws_get(ctx, "q", (size_t)n * rdim * (elem ? 2 : 1), q))) return rc;
    *q32 = nullptr;
    if (mode == RM_MODE_FP8_RESCORE) {
        void *p;
        if ((rc = ws_get(ctx, "q32", (size_t)n * rdim * 4, &p))) return rc;
        *q32 = (float *)p;
    }
    return encode_into(ctx, words, n, k, s, m, c, mode, rdim, norms, *q, *q32, st);
}

// encode -> fused GEMM + top-k -> merge of n fragments; one_launch keeps the
// whole batch in one fused launch, else the built-in launch rule applies
// (topk_batched)
static int map_batch_impl(rm_ctx *ctx, const rm_index *idx, const uint32_t *words, int64_t n,
                          int k, int s, int m, int c, int ktop, int kprime, int mode, float *vals,
                          int32_t *locs, int64_t ldo, float *norms, void *stream, bool one_launch) {
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
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    RM_HIP(hipSetDevice(ctx->device));
    hipStream_t st = pick_stream(ctx, stream);
    void *q;
    float *q32;
    if ((rc = encode_queries(ctx, words, n, k, s, m, c, mode, rdim, norms, &q, &q32, st))) return rc;
    const int elem = mode == RM_MODE_FP16 ? 1 : 0;
    if (one_launch)
        return topk_core(ctx, idx, q, rdim * (elem ? 2 : 1), elem, q32, rdim, n, ktop, kprime, mode,
                         vals, locs, ldo,
                         plan_for(ctx, idx, n, mode, mode == RM_MODE_FP8_RESCORE ? kprime : ktop), st);
    return topk_batched(ctx, idx, q, rdim * (elem ? 2 : 1), elem, q32, rdim, n, ktop, kprime, mode,
                        vals, locs, ldo, st);
}

}  // namespace rm

extern "C" {

int rm_map_batch(rm_ctx *ctx, const rm_index *idx, const uint32_t *words, int64_t n, int k,
                 int s, int m, int c, int ktop, int kprime, int mode, float *vals, int32_t *locs,
                 int64_t ldo, float *norms, void *stream) {
    return map_batch_impl(ctx, idx, words, n, k, s, m, c, ktop, kprime, mode, vals, locs, ldo, norms,
                          stream, false);
}

int rm_map_batches(rm_ctx *ctx, const rm_index *idx, const uint32_t *words, const int64_t *offs,
                   int nb, int k, int s, int m, int c, int ktop, int kprime, int mode, float *vals,
                   int32_t *locs, int64_t ldo, void *stream) {
    RM_CHECK(ctx && idx && offs && vals && locs, RM_ERR_ARG, "NULL argument");
    RM_CHECK(nb >= 0, RM_ERR_ARG, "nb < 0");
    if (nb == 0) return RM_OK;
    RM_CHECK(offs[0] == 0, RM_ERR_ARG, "offs[0] must be 0");
    for (int b = 0; b < nb; ++b)
        RM_CHECK(offs[b + 1] >= offs[b], RM_ERR_ARG, "offs must be non-decreasing");
    // rows are independent: the concatenation runs as one batch, whose
    // (row tile, split) work list spans every batch -- no per-batch launch,
    // no per-batch tail; the plan and the launch rule follow the total row
    // count
    return map_batch_impl(ctx, idx, words, offs[nb], k, s, m, c, ktop, kprime, mode, vals, locs, ldo,
                          nullptr, stream, false);
}

int rm_plan_get(rm_ctx *ctx, const rm_index *idx, int64_t M, int mode, int kp, int32_t *tile_rows,
                int32_t *splits, int32_t *qgroup) {
    RM_CHECK(ctx && idx && M >= 1, RM_ERR_ARG, "NULL argument or M < 1");
    RM_CHECK(kp >= 1 && kp <= kMaxKP, RM_ERR_ARG, "kp must be 1..%d", kMaxKP);
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    const Plan p = plan_for(ctx, idx, M, mode, kp);
    if (tile_rows) *tile_rows = p.bq;
    if (splits) *splits = p.S;
    if (qgroup) *qgroup = p.qg;
    return RM_OK;
}

int rm_launch_rows(rm_ctx *ctx, const rm_index *idx, int64_t M, int mode, int kp, int64_t *rows) {
    RM_CHECK(ctx && idx && rows && M >= 1, RM_ERR_ARG, "NULL argument or M < 1");
    RM_CHECK(kp >= 1 && kp <= kMaxKP, RM_ERR_ARG, "kp must be 1..%d", kMaxKP);
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    *rows = launch_rows(ctx, idx, M, mode, kp);
    return RM_OK;
}

int rm_plan_set(rm_ctx *ctx, int64_t M, int mode, int tile_rows, int splits, int qgroup) {
    RM_CHECK(ctx && M >= 1 && splits >= 0 && splits <= 256 && qgroup >= 0, RM_ERR_ARG,
             "bad plan entry");
    RM_CHECK(mode == RM_MODE_FP8 || mode == RM_MODE_FP8_RESCORE || mode == RM_MODE_FP16,
             RM_ERR_ARG, "bad mode %d", mode);
    RM_CHECK(tile_rows == 0 || tile_rows == 64 || tile_rows == 128 || tile_rows == 256, RM_ERR_ARG,
             "tile_rows must be 0 (default), 64, 128 or 256 (got %d)", tile_rows);
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    auto key = std::make_pair(mode, bucket_of(M));
    if (splits == 0 && tile_rows == 0) {
        ctx->plan.erase(key);
    } else {
        Plan p;
        p.bq = tile_rows;  // 0: the list length's default tile
        p.S = splits;      // 0: the default split rule for that tile
        p.qg = qgroup;
        ctx->plan[key] = p;
    }
    return RM_OK;
}

namespace {
// a pair of timing events, destroyed on every return path
struct EventPair {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    ~EventPair() {
        if (e0) hipEventDestroy(e0);
        if (e1) hipEventDestroy(e1);
    }
};
}  // namespace

int rm_autotune(rm_ctx *ctx, const rm_index *idx, const int64_t *bucket_rows, int nbuckets,
                int mode, int ktop, int kprime, int32_t *best_tile, int32_t *best_splits,
                int32_t *best_qgroup, float *best_ms) {
    RM_CHECK(ctx && idx && bucket_rows && nbuckets >= 0, RM_ERR_ARG, "NULL argument");
    int rc;
    if ((rc = check_query(idx, ktop, kprime, mode))) return rc;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    RM_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const int elem = mode == RM_MODE_FP16 ? 1 : 0;
    const int64_t rdim = idx->rdim;
    const int esz = elem ? 2 : 1;
    const int kp = mode == RM_MODE_FP8_RESCORE ? kprime : ktop;
    EventPair ev;
    RM_HIP(hipEventCreate(&ev.e0));
    RM_HIP(hipEventCreate(&ev.e1));
    for (int b = 0; b < nbuckets; ++b) {
        const int64_t M = bucket_rows[b];
        RM_CHECK(M >= 1, RM_ERR_ARG, "bucket rows must be >= 1");
        void *q, *q32 = nullptr, *vo, *lo;
        if ((rc = ws_get(ctx, "tune_q", (size_t)M * rdim * esz, &q))) return rc;
        if (mode == RM_MODE_FP8_RESCORE && (rc = ws_get(ctx, "tune_q32", (size_t)M * rdim * 4, &q32)))
            return rc;
        if ((rc = ws_get(ctx, "tune_v", (size_t)M * ktop * 4, &vo))) return rc;
        if ((rc = ws_get(ctx, "tune_l", (size_t)M * ktop * 4, &lo))) return rc;
        // tuning queries: index columns themselves (unit-energy embeddings
        // with one perfect match each, like reads), repeated when M > N
        const uint8_t *src = elem ? (const uint8_t *)idx->b16 : idx->b8;
        for (int64_t r = 0; r < M; r += idx->N) {
            const int64_t n = std::min<int64_t>(idx->N, M - r);
            RM_HIP(hipMemcpyAsync((uint8_t *)q + r * rdim * esz, src, (size_t)(n * rdim * esz),
                                  hipMemcpyDeviceToDevice, st));
        }
        if (q32 && (rc = launch_dequantize((float *)q32, (const uint8_t *)q, M * rdim, st))) return rc;
        // time each candidate on a prefix of the index sized to ~4e12 FLOP per
        // launch (a few ms) and at least 1 GiB, so that it streams from HBM
        // as the whole index does (a prefix inside the 256 MB MALL favours
        // plans that re-read the index more): the plan depends on the row
        // count, while the per-column work only scales the time (small
        // batches, HBM-bound, run over the whole index)
        rm_index sub = *idx;
        const int64_t ncap = std::max<int64_t>(cdiv(1LL << 30, rdim * esz),
                                               (int64_t)(4e12 / (2.0 * M * rdim)));
        sub.N = std::min<int64_t>(idx->N, ncap);
        const int64_t stiles = cdiv(sub.N, kBI);
        // a batch the built-in rule splits into launches (on the whole index)
        // keeps that rule: its launches take the tuned plan of their own
        // bucket, so no single-launch plan is stored for it
        ctx->plan.erase({mode, bucket_of(M)});
        if (const int64_t rows = launch_rows(ctx, idx, M, mode, kp)) {
            const Plan pc = plan_for(ctx, idx, rows, mode, kp);
            auto run = [&]() {
                return topk_batched(ctx, &sub, q, rdim * esz, elem, (float *)q32, rdim, M, ktop,
                                    kprime, mode, (float *)vo, (int32_t *)lo, M, st);
            };
            if ((rc = run())) return rc;  // warm-up
            RM_HIP(hipEventRecord(ev.e0, st));
            if ((rc = run())) return rc;
            RM_HIP(hipEventRecord(ev.e1, st));
            RM_HIP(hipEventSynchronize(ev.e1));
            float ms = 0;
            RM_HIP(hipEventElapsedTime(&ms, ev.e0, ev.e1));
            if (best_tile) best_tile[b] = pc.bq;
            if (best_splits) best_splits[b] = pc.S;
            if (best_qgroup) best_qgroup[b] = pc.qg;
            if (best_ms) best_ms[b] = ms / (float)cdiv(M, rows);  // per launch
            continue;
        }
        // tiles: every ring tile up to the first that holds the batch; splits
        // from a quarter of one wave of workgroups to four waves
        std::vector<Plan> cands;
        for (int bq : {64, 128, 256}) {
            if (!tile_rows_ok(kp, bq)) continue;
            if (bq > 64 && bq / 2 >= M) continue;
            // MFMA-bound from a few hundred rows: a small tile re-reads the index
            // (and runs more epilogue work per MFMA) without saving padding
            if (bq == 64 && M > 1024) continue;
            if (bq == 128 && M > 16384) continue;
            const int64_t R = cdiv(M, bq);
            const int64_t wave = std::max<int64_t>(1, ctx->num_cus / R);
            for (int S = 1; S <= 256 && S <= stiles; S *= 2) {
                if (4 * S < wave && 2 * S <= std::min<int64_t>(256, stiles)) continue;
                if (R * S > 4LL * ctx->num_cus && S > 1) break;
                for (int qg : {0, 8, 4, 2}) {
                    if (qg && qg >= R) continue;
                    Plan p;
                    p.bq = bq;
                    p.S = S;
                    p.qg = qg;
                    cands.push_back(p);
                }
            }
        }
        Plan best = cands.empty() ? plan_for(ctx, &sub, M, mode, kp) : cands[0];
        float bestT = 1e30f;
        for (const Plan &c : cands) {
            auto run = [&]() {
                return topk_core(ctx, &sub, q, rdim * esz, elem, (float *)q32, rdim, M, ktop, kprime,
                                 mode, (float *)vo, (int32_t *)lo, M, c, st);
            };
            if ((rc = run())) return rc;  // warm-up
            RM_HIP(hipEventRecord(ev.e0, st));
            const int reps = 2;
            for (int r = 0; r < reps; ++r)
                if ((rc = run())) return rc;
            RM_HIP(hipEventRecord(ev.e1, st));
            RM_HIP(hipEventSynchronize(ev.e1));
            float ms = 0;
            RM_HIP(hipEventElapsedTime(&ms, ev.e0, ev.e1));
            ms /= reps;
            if (ms < bestT) { bestT = ms; best = c; }
        }
        ctx->plan[{mode, bucket_of(M)}] = best;
        if (best_tile) best_tile[b] = best.bq;
        if (best_splits) best_splits[b] = best.S;
        if (best_qgroup) best_qgroup[b] = best.qg;
        if (best_ms) best_ms[b] = bestT;
    }
    return RM_OK;
}

int rm_set_persistent(rm_ctx *ctx, int on) {
    RM_CHECK(ctx, RM_ERR_ARG, "NULL context");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    ctx->persistent = on != 0;
    return RM_OK;
}

int rm_prof_enable(rm_ctx *ctx, int on) {
    RM_CHECK(ctx, RM_ERR_ARG, "ctx is NULL");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    ctx->prof = on != 0;
    return RM_OK;
}

int rm_prof_read(rm_ctx *ctx, double *kernel_ms, int64_t *launches, double *flop) {
    RM_CHECK(ctx, RM_ERR_ARG, "ctx is NULL");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    double ms = 0;
    for (auto &p : ctx->prof_ev) {
        RM_HIP(hipEventSynchronize(p.second));
        float t = 0;
        RM_HIP(hipEventElapsedTime(&t, p.first, p.second));
        ms += t;
        hipEventDestroy(p.first);
        hipEventDestroy(p.second);
    }
    if (kernel_ms) *kernel_ms = ms;
    if (launches) *launches = (int64_t)ctx->prof_ev.size();
    if (flop) *flop = ctx->prof_flop;
    ctx->prof_ev.clear();
    ctx->prof_flop = 0;
    return RM_OK;
}

int64_t rm_batch_bytes(int64_t M, int k, int rdim, int ktop, int kprime, int mode) {
    const int KP = (mode == RM_MODE_FP8_RESCORE) ? kprime : ktop;
    const int64_t W = (k + 15) / 16;
    int64_t b = M * W * 4;                                  // packed words
    b += M * rdim * (mode == RM_MODE_FP16 ? 2 : 1);         // encoded queries
    if (mode == RM_MODE_FP8_RESCORE) b += M * rdim * 4;     // fp32 rescoring copy
    b += M * 256LL * KP * 8 + M * 8LL;                      // split partials (<= 256 splits), bounds
    b += M * (int64_t)ktop * 8;                             // vals + locs
    return b;
}

int64_t rm_device_free_bytes(rm_ctx *ctx) {
    if (!ctx) return -1;
    size_t fr = 0, tot = 0;
    hipSetDevice(ctx->device);
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) return -1;
    return (int64_t)fr;
}

}  // extern "C"
