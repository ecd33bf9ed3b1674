// rm_internal.h -- shared host-side declarations of librotormap_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/rotormap_hip.h"

namespace rm {

// thread-local last error + status helper
int fail(int code, const char *fmt, ...);
void clear_error();

#define RM_HIP(x)                                                                      \
    do {                                                                               \
        hipError_t e_ = (x);                                                           \
        if (e_ != hipSuccess) return rm::fail(RM_ERR_HIP, "%s: %s", #x, hipGetErrorString(e_)); \
    } while (0)
#define RM_CHECK(cond, code, ...)                       \
    do {                                                \
        if (!(cond)) return rm::fail(code, __VA_ARGS__); \
    } while (0)

inline int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

// An HSA dispatch packet holds the grid size in work-items as 32 bits: a
// launch must stay below 2^32 work-items.  Element-wise kernels therefore
// walk their range with a grid stride over a grid of at most kStrideBlocks.
constexpr int64_t kMaxWorkItems = 0xffffffffLL;
constexpr int64_t kStrideBlocks = 1 << 20;
inline unsigned stride_grid(int64_t n, int block) {
    const int64_t g = cdiv(n, block);
    return (unsigned)(g < kStrideBlocks ? g : kStrideBlocks);
}

// ---- fused GEMM + top-k tile geometry (shared by host planning and kernels)
constexpr int kBI = 256;       // index columns per tile
constexpr int kBK = 64;        // K bytes per LDS stage
// Stage-buffer row layout: RM_LDS_SWZ 0 = 80-byte pitch (16-byte pad);
// 1 / 2 = 64-byte pitch with the 16-byte chunks XOR-swizzled by
// (row >> 2) & 3 / (row >> 1) & 3.
#ifndef RM_LDS_SWZ
#define RM_LDS_SWZ 1
#endif
constexpr int kSwz = RM_LDS_SWZ;
constexpr int kRow = kSwz ? 64 : 80;  // LDS row pitch in bytes
constexpr int kThreads = 256;  // 4 waves
constexpr int kMaxKP = 64;     // largest top-k' list

// Queries per tile of the fused kernel.  Lists of k' <= 32 run on the ring
// kernels with 256, 128 or 64 queries per tile (the tile plan picks one: a
// small batch wastes less MFMA work on padded query rows with a smaller
// tile); longer lists use the register-staged kernel with 128.
inline bool ring_list(int kp) { return kp <= 32; }
inline bool tile_rows_ok(int kp, int bq) {
    return ring_list(kp) ? (bq == 256 || bq == 128 || bq == 64) : bq == 128;
}
// query bytes the fused kernel's concurrent workgroups may re-read from the
// MALL (half of MI355X's 256 MB Infinity Cache)
constexpr int64_t kQueryWorkingSet = 128LL << 20;
// LDS layout of the fused kernel without the ring (br queries per tile, list
// length kp): [ stage buffers x2 ][ lists br x (kp+1) x 8 ][ candidate slots br x (capd+1)
// x 8 ][ tau br ][ slot counters br ][ 3 vote flags ].  The candidate slots
// persist across tiles (merged into the lists only when some query's slots
// overflow); capd takes what is left of the 160 KiB, at most 64.
constexpr size_t kLdsMax = 160 * 1024;
constexpr size_t stage_pair_bytes(int br) { return 2 * (size_t)(kBI + br) * kRow; }
constexpr size_t list_offset(int br) { return stage_pair_bytes(br); }
constexpr size_t cand_offset(int br, int kp) { return list_offset(br) + (size_t)br * (kp + 1) * 8; }
constexpr size_t tail_bytes(int br) { return (size_t)br * 8 + 16; }
constexpr int cand_slots(int br, int kp) {
    const size_t used = cand_offset(br, kp) + tail_bytes(br);
    const long long avail = used >= kLdsMax ? 0 : (long long)((kLdsMax - used) / ((size_t)br * 8)) - 1;
    return avail > 64 ? 64 : (int)avail;
}
constexpr size_t topk_lds_bytes_for(int br, int kp) {
    return cand_offset(br, kp) + (size_t)br * (cand_slots(br, kp) + 1) * 8 + tail_bytes(br);
}
// The 256-query tile (k' <= 32) streams K through a ring of kRing stage
// buffers filled by global -> LDS loads, and keeps each query's list in the
// split partials (global memory):
// [ ring kRing x (kBI + br) x kRow ][ candidate slots br x (capd+1) x 8 ][ tau
// br ][ slot counters br ][ 3 vote flags ][ shared floor br ].  The ring's load rows follow the
// 64-byte swizzled row layout.  During the epilogue ring buffers 2 and 3 are
// idle (the next tile's first stages may be landing in 0 and 1): a merge
// stages the lists there, br x 32 values in buffer 2 and as many columns in
// buffer 3, and the first tile's bound seed uses buffer 2.
constexpr int kRing = 4;
constexpr bool kRingOk = kSwz == 1;
inline bool uses_ring(int kp) { return kRingOk && ring_list(kp); }
constexpr size_t ring_bytes(int br) { return (size_t)kRing * (kBI + br) * kRow; }
constexpr size_t ring_cand_offset(int br) { return ring_bytes(br); }
// the ring layout's tail: tau, counters, flags, then the per-query floor
// from the bound the column splits of a query share
constexpr size_t ring_tail_bytes(int br) { return tail_bytes(br) + (size_t)br * 4; }
constexpr int ring_cand_slots(int br) {
    const size_t used = ring_cand_offset(br) + ring_tail_bytes(br);
    const long long avail = used >= kLdsMax ? 0 : (long long)((kLdsMax - used) / ((size_t)br * 8)) - 1;
    return avail > 64 ? 64 : (int)avail;
}
constexpr size_t ring_lds_bytes(int br) {
    return ring_cand_offset(br) + (size_t)br * (ring_cand_slots(br) + 1) * 8 + ring_tail_bytes(br);
}
constexpr size_t ring_buffer_bytes(int br) { return (size_t)(kBI + br) * kRow; }
// one array (values or columns) of the lists staged for a merge
constexpr size_t ring_staged_list_bytes(int br) { return (size_t)br * 32 * 4; }
// bq: queries per tile (tile_rows_ok); without the ring (k' > 32, or the
// unswizzled layouts) the register-staged kernel's tile of 128 or 256
inline int topk_cand_slots(int kp, int bq) {
    return uses_ring(kp) ? ring_cand_slots(bq) : cand_slots(ring_list(kp) ? 256 : 128, kp);
}
inline size_t topk_lds_bytes(int kp, int bq) {
    return uses_ring(kp) ? ring_lds_bytes(bq) : topk_lds_bytes_for(ring_list(kp) ? 256 : 128, kp);
}
// the C-writing kernel: the 256-query ring (or two register-staged stages)
inline size_t store_lds_bytes() { return kRingOk ? ring_bytes(256) : 2 * (size_t)(kBI + 256) * kRow; }

// Tile plan of one fused launch: query rows per tile, column splits and the
// query tiles grouped per XCD (0 = all).
struct Plan {
    int bq = 256;
    int S = 1;
    int qg = 0;
};

struct GemmArgs {
    const uint8_t *db;  // index rows (K contiguous bytes each), row pitch ldb bytes
    int64_t ldb;
    const uint8_t *q;   // query rows, row pitch ldq bytes
    int64_t ldq;
    int64_t M;          // query rows
    int BQ;             // query rows per tile (tile_rows_ok)
    int64_t N;          // index rows covered by this launch
    int Kb;             // K in bytes
    int S;              // column splits
    int64_t R;          // row tiles = ceil(M / rows per tile)
    int qg;             // row tiles per XCD group (items ordered group, split, tile)
    int ngroups;        // ceil(R / qg)
    int64_t tiles;      // column tiles = ceil(N / kBI)
    int64_t col_base;   // global column id of db row 0
    // top-k epilogue
    float *part_v;
    int32_t *part_c;
    int KP;
    int capd;           // candidate slots per query (cand_slots)
    // per-query lower bound of the k'-th score shared by a query's column
    // splits (order keys, M entries after the partials; ring layout with
    // S > 1), or null
    int32_t *bound;
    // store-C epilogue
    void *C;
    int c_dtype;
    int64_t ldc;
    // grid: items8 = item count rounded up to a multiple of 8 (set by the
    // launcher); persist_grid >= 8 caps the grid at that many workgroups
    // (rounded down to a multiple of 8), each walking items with the grid's
    // stride; 0 launches one workgroup per item
    int64_t items8;
    int persist_grid;
};

// launchers (gemm_topk.hip)
int launch_gemm_topk(const GemmArgs &a, int elem, hipStream_t st);
int launch_gemm_store(const GemmArgs &a, int elem, hipStream_t st);
// db: the columns the FP32 rescoring reads, db_dtype RM_E4M3 or RM_F16, ldb
// elements apart
int launch_merge(const float *part_v, const int32_t *part_c, int S, int KP, int64_t M, int64_t N,
                 int ktop, int mode, const float *q32, int64_t ldq32, const void *db,
                 int db_dtype, int64_t ldb, int K, float *vals, int32_t *locs, int64_t ldo,
                 hipStream_t st);
// the running row top-k of a C chunk in two passes: G column segments per
// chunk (rowtopk_segments), their lists in pv / pc (M x G x k each); seed (M
// ints, or NULL) holds the per-row bound of pass 0 for rows whose running
// list is not full
int rowtopk_segments(int64_t M, int64_t wi);
int launch_rowtopk(float *dv, int32_t *dl, int64_t ldd, const void *C, int c_dtype, int64_t ldc,
                   int64_t M, int64_t wi, int64_t offset, int k, int G, float *pv, int32_t *pc,
                   int32_t *seed, hipStream_t st);
int launch_transpose_u8(uint8_t *dst, const uint8_t *src, int64_t rows, int64_t cols,
                        int64_t s_row, int64_t s_col, int esize, hipStream_t st);

// launchers (encode.hip)
struct EncOut {
    void *p;
    int dtype;
    int64_t sf, sd;
};
int launch_encode(const uint32_t *words, int64_t n, int k, int s, int m, int c, int normalize,
                  const int32_t *table, EncOut o1, EncOut o2, float *norms, int32_t *counts,
                  hipStream_t st);
int launch_revcomp(uint32_t *rc, const uint32_t *w, int64_t n, int k, hipStream_t st);
int launch_extract(uint32_t *out, const uint32_t *rec, const int64_t *starts, int64_t n, int k,
                   hipStream_t st);
int launch_quantize(uint8_t *dst, const void *src, int src_dtype, int64_t n, hipStream_t st);
int launch_dequantize(float *dst, const uint8_t *src, int64_t n, hipStream_t st);
int launch_widen_f16(uint16_t *dst, const uint8_t *src, int64_t n, hipStream_t st);

void build_phase_table(int k, int s, int m, int32_t *tab);

// once per device with it current: the device address of the no-allocation
// entry point's scratch array (noalloc.hip)
int noalloc_setup(int device);

// FASTA bytes -> packed reads (fasta.hip)
int fasta_reads(rm_ctx *ctx, const uint8_t *raw, int64_t n, int k, uint32_t *words,
                int64_t max_reads, int64_t *hdr_off, int32_t *hdr_len, int64_t *n_reads,
                hipStream_t st);

}  // namespace rm

// growable device workspace slots
struct rm_buf {
    void *p = nullptr;
    size_t n = 0;
};

struct rm_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int num_cus = 256;
    std::recursive_mutex mu;
    std::map<std::tuple<int, int, int>, int32_t *> tables;  // (k, s, m) -> device phasor table
    size_t table_bytes = 0;
    std::map<std::string, rm_buf> ws;
    // tile plans: (mode, M bucket ceil(log2 M)) -> plan (rm_autotune, rm_plan_set)
    std::map<std::pair<int, int>, rm::Plan> plan;
    // kernel timing (rm_prof_enable): event pairs around fused launches
    bool persistent = true;  // fused launches use a grid of one workgroup per CU
    bool prof = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> prof_ev;
    double prof_flop = 0;
};

struct rm_index {
    rm_ctx *ctx = nullptr;
    uint8_t *b8 = nullptr;    // N x rdim e4m3 (column j contiguous)
    uint16_t *b16 = nullptr;  // N x rdim fp16: the widening, or the verbatim fp16 index
    int64_t rdim = 0, N = 0;
    bool verbatim16 = false;   // b16 holds the fp16 index itself (rm_index_create_f16)
    int rescore = RM_E4M3;     // columns the FP32 rescoring reads: RM_E4M3 (b8) or RM_F16 (b16)
};

namespace rm {
// workspace slot of at least `bytes` (grows; contents undefined)
int ws_get(rm_ctx *ctx, const char *name, size_t bytes, void **out);
int table_get(rm_ctx *ctx, int k, int s, int m, const int32_t **out);
hipStream_t pick_stream(rm_ctx *ctx, void *stream);
int default_splits(rm_ctx *ctx, int64_t M, int64_t N, int bq, int64_t row_bytes);
// the whole fused query path on already-encoded queries
int topk_core(rm_ctx *ctx, const rm_index *idx, const void *Q, int64_t ldq_bytes, int elem,
              const float *Q32, int64_t ldq32, int64_t M, int ktop, int kprime, int mode,
              float *vals, int32_t *locs, int64_t ldo, const Plan &pl, hipStream_t st);
// the same with caller-owned split partials (pv, pc of pbytes each)
int topk_core_bufs(rm_ctx *ctx, const rm_index *idx, const void *Q, int64_t ldq_bytes, int elem,
                   const float *Q32, int64_t ldq32, int64_t M, int ktop, int kprime, int mode,
                   float *vals, int32_t *locs, int64_t ldo, const Plan &pl, float *pv,
                   int32_t *pc, size_t pbytes, hipStream_t st);
// bytes of each split-partial array for M rows, list length KP, S splits
size_t part_bytes(const rm_index *idx, int64_t M, int KP, int S);
// the largest split count <= S whose partials for M rows fit pbytes (>= 1)
int max_splits_within(const rm_index *idx, int64_t M, int KP, int S, size_t pbytes);
// encode n fragments into caller-owned query buffers (q32 only for rescoring)
int encode_into(rm_ctx *ctx, const uint32_t *words, int64_t n, int k, int s, int m, int c,
                int mode, int64_t rdim, float *norms, void *q, float *q32, hipStream_t st);
int check_encoder(int k, int s, int m, int c, int normalize);
int check_query(const rm_index *idx, int ktop, int kprime, int mode);
// the tile plan of a batch of M rows (kp: the list length the fused kernel
// keeps, kprime for RM_MODE_FP8_RESCORE and ktop otherwise)
Plan plan_for(rm_ctx *ctx, const rm_index *idx, int64_t M, int mode, int kp);
// rows per fused launch under the built-in rule for a batch of M rows (0:
// one launch), and the query path over such launches
int64_t launch_rows(rm_ctx *ctx, const rm_index *idx, int64_t M, int mode, int kp);
int topk_batched(rm_ctx *ctx, const rm_index *idx, const void *Q, int64_t ldq_bytes, int elem,
                 const float *Q32, int64_t ldq32, int64_t M, int ktop, int kprime, int mode,
                 float *vals, int32_t *locs, int64_t ldo, hipStream_t st);
}  // namespace rm
