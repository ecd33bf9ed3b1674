// gemm_topk.hip -- reads x index GEMM on gfx950 MFMA with the per-row top-k
// fused into the epilogue (scores never reach HBM), the split merge with
// optional FP32 rescoring, and the two drop-in chunk kernels of the :hip
// engine (C-writing GEMM, running top-k merge of a C chunk).
//
// Reference path replaced: batch_gemm_topk! (RotorMap/search/engine_fp8.jl:
// 399-437) = cuBLASLt fp8 GEMM into an fp16 C chunk (RotorMap/gemm/
// fp8_lt.jl:128-168) + a separate seg_rowtopk_merge_kernel! pass over it
// (RotorMap/gemm/topk_kernels.jl:213-347), chunk by chunk.
//
// Layout: the index is N rows of K contiguous bytes (the reference's column-
// major `embeds8` (rdim, N) as is); queries are M rows of K contiguous bytes.
// The kernel computes C^T tiles: MFMA operand A = 32 index rows, operand B =
// 32 query rows, so each lane's 16 accumulators of a 32x32 block belong to
// ONE query (lane % 32) and 16 index columns -- the per-query top-k needs no
// cross-lane traffic to filter.
//
// Tile: 256 index rows x 256 queries per workgroup, 8 waves (two per SIMD)
// in 2 x 4, wave tile 128 index rows x 64 queries = 4 x 2 blocks of
// v_mfma_scale_f32_32x32x64_f8f6f4 (unit scales) per 64-byte K stage, or
// 2 x v_mfma_f32_32x32x16_f16 per block for the FP16 path; lists longer than
// 32 use 4 waves over 256 x 128 tiles.  K is streamed through LDS in 64-byte
// stages (64-byte rows, 16-byte chunks XOR-swizzled by (row >> 2) & 3,
// conflict-free for the fragment reads), one barrier per stage.  The
// 256-query tile fills a ring of four stage buffers with global -> LDS loads
// three stages ahead (no staging registers, no LDS store instructions); the
// 128-query tile and the C-writing kernel double-buffer through registers
// (global -> register loads two stages ahead, register -> LDS stores one
// stage ahead).  A workgroup owns one query tile and walks a contiguous range
// of index tiles (its column split); per query it keeps a top-k' list and a
// threshold tau = the list's k'-th value.  After each tile a lane flags
// values > tau (a query's 16 x 4 values are compared one by one only when
// their maximum beats tau); flagged values are appended to per-query LDS
// slots that persist across tiles and are merged into the lists (one thread
// per query, (score desc, column asc) order) only when some query's slots
// overflow.  With the ring the lists live in the split partials and are
// staged into the idle ring for a merge; otherwise they stay in LDS.  The
// grid is persistent by default (one workgroup per CU walking
// its XCD's items); items are ordered so that the concurrently resident
// workgroups of one XCD (dispatch is round-robin over the 8 XCDs) share
// index tiles through that XCD's L2.
#include "rm_device.h"
#include "rm_internal.h"

#include <climits>

namespace rm {

typedef int v8i __attribute__((ext_vector_type(8)));
typedef float v16f __attribute__((ext_vector_type(16)));
typedef _Float16 v8h __attribute__((ext_vector_type(8)));

// Tile shape: kBI = 256 index rows x BR queries over NT / 64 waves.  A wave
// owns 64 queries (BB = 2 blocks of 32) and 32 * AI index rows (AI blocks):
// the waves form WI index groups x WQ = BR / 64 query groups, wave w at index
// group w % WI and query group w / WI.  Every shape gives a query 16 groups
// of 16 accumulators (AI per lane x 2 lanes per wave x WI waves), which the
// first-tile bound relies on.
//   Shape<512, 256, true>   8 waves, 2 x 4, wave tile 128 x 64 (ring)
//   Shape<512, 128, true>   8 waves, 4 x 2, wave tile  64 x 64 (ring)
//   Shape<512,  64, true>   8 waves, 8 x 1, wave tile  32 x 64 (ring)
//   Shape<256, 128, false>  4 waves, 2 x 2, wave tile 128 x 64 (register staging)
//   Shape<512, 256, false>  the C-writing kernel (register staging)
template <int NT_, int BR_, bool RING_>
struct Shape {
    static constexpr int NT = NT_;               // threads
    static constexpr int NW = NT / 64;           // waves
    static constexpr int BR = BR_;               // queries per tile
    static constexpr int BB = 2;                 // 32-query blocks per wave
    static constexpr int WQ = BR / 64;           // query groups
    static constexpr int WI = NW / WQ;           // index groups
    static constexpr int AI = kBI / 32 / WI;     // 32-row index blocks per wave
    static constexpr bool RING = RING_;
    static constexpr int ROWS = kBI + BR;        // LDS rows per stage buffer
    static constexpr int CHUNKS = ROWS * 4 / NT; // 16-byte chunks per thread per stage
    static constexpr int STAGE_BYTES = ROWS * kRow;
    static_assert(WQ >= 1 && WI * WQ == NW && AI * 32 * WI == kBI, "tile shape");
    static_assert(2 * AI * WI == 16, "16 groups of 16 accumulators per query");
};
using ShapeTopk256 = Shape<512, 256, true>;
using ShapeTopk128 = Shape<512, 128, true>;
using ShapeTopk64 = Shape<512, 64, true>;
using ShapeList128 = Shape<256, 128, false>;
using ShapeStore = Shape<512, 256, kRingOk>;

template <class SH>
struct StageRegs {
    uint4 v[SH::CHUNKS];
};

// Buffer descriptors of a tile's operand rows: base at the tile's first row,
// num_records = the in-range bytes, so loads past the last row return zeros
// (no per-load branch; 32-bit offsets within the tile).
struct TileSrc {
    __amdgpu_buffer_rsrc_t db, q;
    int ldb, ldq;
};

__device__ __forceinline__ __amdgpu_buffer_rsrc_t rows_rsrc(const uint8_t *base, int64_t row0,
                                                            int64_t nrows, int64_t rows_max,
                                                            int64_t ld) {
    int64_t n = nrows - row0;
    n = n < 0 ? 0 : (n > rows_max ? rows_max : n);
    return __builtin_amdgcn_make_buffer_rsrc((void *)(base + row0 * ld), (short)0, (int)(n * ld),
                                             0x00020000);
}

// global -> registers for one 64-byte K stage: chunk j of thread tid covers
// row tid/4 + (NT/4)*j (index rows first, then query rows), bytes
// (tid%4)*16 .. +16 of the stage.
template <class SH>
__device__ __forceinline__ void g_load(StageRegs<SH> &st, const TileSrc &ts, int tid, int kb) {
    constexpr int RS = SH::NT / 4;  // rows per chunk step
    static_assert(kBI % RS == 0 && SH::BR % RS == 0, "chunk steps split index and query rows");
    const int r = tid >> 2, seg = (tid & 3) * 16;
    const int vdb = r * ts.ldb + seg, vq = r * ts.ldq + seg;
#pragma unroll
    for (int j = 0; j < SH::CHUNKS; ++j) {
        if (j < kBI / RS)
            st.v[j] = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(
                                                    ts.db, vdb, j * RS * ts.ldb + kb, 0));
        else
            st.v[j] = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(
                                                    ts.q, vq, (j - kBI / RS) * RS * ts.ldq + kb, 0));
    }
}

// physical 16-byte chunk of logical chunk c in row `row` of a stage buffer
__device__ __forceinline__ int phys_chunk(int row, int c) {
    if constexpr (kSwz == 1) return c ^ ((row >> 2) & 3);
    else if constexpr (kSwz == 2) return c ^ ((row >> 1) & 3);
    else return c;
}

template <class SH>
__device__ __forceinline__ void s_store(const StageRegs<SH> &st, uint8_t *sb, int tid) {
    constexpr int RS = SH::NT / 4;  // a multiple of 16: same swizzle per chunk j
    uint8_t *p = sb + (tid >> 2) * kRow + phys_chunk(tid >> 2, tid & 3) * 16;
#pragma unroll
    for (int j = 0; j < SH::CHUNKS; ++j) *(uint4 *)(p + j * RS * kRow) = st.v[j];
}

__device__ __forceinline__ v8i ld_v8i(const uint8_t *p) {
    uint4 x = *(const uint4 *)p, y = *(const uint4 *)(p + 16);
    v8i r;
    r[0] = (int)x.x; r[1] = (int)x.y; r[2] = (int)x.z; r[3] = (int)x.w;
    r[4] = (int)y.x; r[5] = (int)y.y; r[6] = (int)y.z; r[7] = (int)y.w;
    return r;
}

// One stage's operand fragments of a wave: lane l feeds row (l % 32) of each
// block with K bytes [(l/32)*32, +32) of the stage (fp8), or halves
// [kk*16 + (l/32)*8, +8) per 16-wide f16 step kk: the same K slice for both
// operands, so every product pairs matching K indices.
template <int AI>
struct Frags {
    uint4 a[AI][2];
    uint4 b[2][2];
};

// row = stage buffer + (32-aligned row base + lane % 32) * kRow: the swizzle
// depends on lane % 32 only
template <int ELEM>
__device__ __forceinline__ void ld_pair(uint4 (&d)[2], const uint8_t *row, int lane) {
    const int hl = lane >> 5, rl = lane & 31;
    const int c0 = ELEM == 0 ? 2 * hl : hl;          // first 16 bytes
    const int c1 = ELEM == 0 ? 2 * hl + 1 : 2 + hl;  // second 16 bytes
    d[0] = *(const uint4 *)(row + phys_chunk(rl, c0) * 16);
    d[1] = *(const uint4 *)(row + phys_chunk(rl, c1) * 16);
}

template <int ELEM, int AI>
__device__ __forceinline__ void ld_frags_b(Frags<AI> &f, const uint8_t *sb, int wr, int lane) {
    // wave wr owns queries [wr*64, (wr+1)*64)
#pragma unroll
    for (int bj = 0; bj < 2; ++bj)
        ld_pair<ELEM>(f.b[bj], sb + (kBI + wr * 64 + bj * 32 + (lane & 31)) * kRow, lane);
}

template <int ELEM, int AI, int A0, int A1>
__device__ __forceinline__ void ld_frags_a(Frags<AI> &f, const uint8_t *sb, int wi, int lane) {
#pragma unroll
    for (int ai = A0; ai < A1; ++ai)
        ld_pair<ELEM>(f.a[ai], sb + (wi * (32 * AI) + ai * 32 + (lane & 31)) * kRow, lane);
}

struct Bytes32 {
    uint4 lo, hi;
};

// FIRST: the tile's first stage, accumulating onto zero (no zeroing pass)
template <int ELEM, int AI, int A0, int A1, bool FIRST = false>
__device__ __forceinline__ void mma_part(v16f (&acc)[AI][2], const Frags<AI> &f) {
    const v16f zero = {};
#pragma unroll
    for (int ai = A0; ai < A1; ++ai)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj) {
            if constexpr (ELEM == 0) {
                const v8i fa = __builtin_bit_cast(v8i, (Bytes32{f.a[ai][0], f.a[ai][1]}));
                const v8i fb = __builtin_bit_cast(v8i, (Bytes32{f.b[bj][0], f.b[bj][1]}));
                acc[ai][bj] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(
                    fa, fb, FIRST ? zero : acc[ai][bj], 0, 0, 0, 0, 0, 0);
            } else {
#pragma unroll
                for (int kk = 0; kk < 2; ++kk)
                    acc[ai][bj] = __builtin_amdgcn_mfma_f32_32x32x16_f16(
                        __builtin_bit_cast(v8h, f.a[ai][kk]), __builtin_bit_cast(v8h, f.b[bj][kk]),
                        (FIRST && kk == 0) ? zero : acc[ai][bj], 0, 0, 0);
            }
        }
}

// Full-K product of one (256 index x BR query) tile into acc, one barrier per
// 64-byte stage.  After the barrier a wave reads the query fragments and the
// first two index fragments of stage s, stores stage s+1 (already in
// registers) into the other buffer and issues stage s+2's global loads; the
// last two index fragments are read while the first half's MFMAs run.
template <int ELEM, class SH>
__device__ __forceinline__ void tile_product(v16f (&acc)[SH::AI][2], const GemmArgs &a, int64_t i0,
                                             int64_t r0, uint8_t *sb0, uint8_t *sb1, int tid,
                                             int wi, int wr, int lane) {
    constexpr int AI = SH::AI, H = (AI + 1) / 2;
    TileSrc ts;
    ts.db = rows_rsrc(a.db, i0, a.N, kBI, a.ldb);
    ts.q = rows_rsrc(a.q, r0, a.M, SH::BR, a.ldq);
    ts.ldb = (int)a.ldb;
    ts.ldq = (int)a.ldq;
    const int nk = a.Kb / kBK;  // >= 1
    StageRegs<SH> g;
    Frags<AI> f;
    g_load<SH>(g, ts, tid, 0);
    s_store<SH>(g, sb0, tid);
    if (nk > 1) g_load<SH>(g, ts, tid, kBK);
    {  // stage 0 accumulates onto zero
        __syncthreads();
        ld_frags_b<ELEM, AI>(f, sb0, wr, lane);
        ld_frags_a<ELEM, AI, 0, H>(f, sb0, wi, lane);
        if (1 < nk) s_store<SH>(g, sb1, tid);
        if (2 < nk) g_load<SH>(g, ts, tid, 2 * kBK);
        mma_part<ELEM, AI, 0, H, true>(acc, f);
        ld_frags_a<ELEM, AI, H, AI>(f, sb0, wi, lane);
        mma_part<ELEM, AI, H, AI, true>(acc, f);
    }
    for (int s = 1; s < nk; ++s) {
        uint8_t *cur = (s & 1) ? sb1 : sb0;
        uint8_t *nxt = (s & 1) ? sb0 : sb1;
        __syncthreads();
        ld_frags_b<ELEM, AI>(f, cur, wr, lane);
        ld_frags_a<ELEM, AI, 0, H>(f, cur, wi, lane);
        if (s + 1 < nk) s_store<SH>(g, nxt, tid);
        if (s + 2 < nk) g_load<SH>(g, ts, tid, (s + 2) * kBK);
        mma_part<ELEM, AI, 0, H>(acc, f);
        ld_frags_a<ELEM, AI, H, AI>(f, cur, wi, lane);
        mma_part<ELEM, AI, H, AI>(acc, f);
    }
}

// ---------------------------------------------------------------------------
// K loop of the 8-wave ring tiles (256, 128 or 64 queries) fed by global ->
// LDS loads into a ring of kRing stage buffers (kRing - 1 stages ahead).  No
// staging registers and no LDS store instructions: waves 0-3 load the
// stage's 256 index rows (64 each, 4 x 16-row pieces of 1 KB), waves 4-7 its
// BR query rows (BR / 4 each, BR / 64 pieces); lane l writes LDS bytes l*16
// of its piece and fetches the logical chunk that the 64-byte swizzle puts
// there.  A stage waits for its own loads (vmcnt counts this wave's loads
// per stage) before the barrier that publishes them, and reuses the buffer
// of the stage before it.
// ---------------------------------------------------------------------------
static_assert(!kRingOk || kSwz == 1, "ring rows use the 64-byte swizzle");

struct RingSrc {
    __amdgpu_buffer_rsrc_t r;  // this wave's operand
    int voff;                  // lane offset of its first 16-row piece
    int step16;                // bytes between pieces (16 rows)
    int lds_off;               // LDS offset of the wave's rows in a stage buffer
    int wid;                   // wave id (uniform)
};

// loads per stage of an index-row wave (64 rows in 16-row pieces)
constexpr int kRingLoads = 4;
static_assert((kRing - 2) * kRingLoads == 8, "ring_wait's index waves wait with vmcnt(8)");

template <class SH>
__device__ __forceinline__ RingSrc ring_src(const TileSrc &ts, int tid) {
    static_assert(SH::NT == 512 && SH::BR % 64 == 0, "ring tiles: 8 waves, 16-row query pieces");
    const int wid = __builtin_amdgcn_readfirstlane(tid >> 6);  // uniform: no waterfall
    const int lane = tid & 63;
    // LDS row (16-aligned piece base + lane/4) has ((row >> 2) & 3) = (lane >> 4) & 3
    const int c = (lane & 3) ^ ((lane >> 4) & 3);
    constexpr int QR = SH::BR / 4;  // query rows per query wave
    const bool iw = wid < 4;         // an index-row wave
    const int ld = iw ? ts.ldb : ts.ldq;
    RingSrc d;
    d.wid = wid;
    d.r = iw ? ts.db : ts.q;
    d.voff = ((wid & 3) * (iw ? 64 : QR) + (lane >> 2)) * ld + c * 16;
    d.step16 = 16 * ld;
    d.lds_off = (iw ? wid * 64 : kBI + (wid & 3) * QR) * kRow;
    return d;
}

template <class SH>
__device__ __forceinline__ void ring_load(const RingSrc &d, uint8_t *sb, int kb) {
    constexpr int QL = SH::BR / 64;  // loads per stage of a query wave
#pragma unroll
    for (int j = 0; j < kRingLoads; ++j)
        if (j < QL || d.wid < 4)  // wave-uniform
            __builtin_amdgcn_raw_ptr_buffer_load_lds(
                d.r, (__attribute__((address_space(3))) void *)(sb + d.lds_off + j * 16 * kRow), 16,
                d.voff, kb + j * d.step16, 0, 0);
}

// Wait until the stage issued kRing - 1 stages ago has landed in this wave's
// rows (two younger stages stay in flight: vmcnt = 2 x the wave's loads per
// stage), then the barrier that publishes it.  Written out so that the wait
// stays at that count instead of the vmcnt(0) a fence would add.
template <class SH>
__device__ __forceinline__ void ring_wait_asm(int wid) {
    if constexpr (SH::BR == 256) {
        asm volatile("s_waitcnt vmcnt(8) lgkmcnt(0)\n\ts_barrier" ::: "memory");
    } else if (wid < 4) {
        asm volatile("s_waitcnt vmcnt(8) lgkmcnt(0)\n\ts_barrier" ::: "memory");
    } else if constexpr (SH::BR == 128) {
        asm volatile("s_waitcnt vmcnt(4) lgkmcnt(0)\n\ts_barrier" ::: "memory");
    } else {
        static_assert(SH::BR == 64, "ring tiles of 256, 128 or 64 queries");
        asm volatile("s_waitcnt vmcnt(2) lgkmcnt(0)\n\ts_barrier" ::: "memory");
    }
}

// The sched_barriers keep the machine scheduler from moving any instruction
// across the wait and barrier: the fragment reads carry alias scopes apart
// from the loads (below), and the scheduler was seen to sink such reads past
// a barrier asm, after which their buffer is refilled.
template <class SH>
__device__ __forceinline__ void ring_wait(int wid) {
    static_assert(kRing == 4, "the waits below assume two younger stages in flight");
    __builtin_amdgcn_sched_barrier(0);
    ring_wait_asm<SH>(wid);
    __builtin_amdgcn_sched_barrier(0);
}

// Waves 4-7 of a ring tile (the second wave on each SIMD) can run one stage
// behind waves 0-3 (RM_LAG): in the step of stage s a lagging wave first runs
// the MFMAs of stage s-1 on fragments it read in the previous step, then
// reads stage s's fragments; waves 0-3 read first and compute after.  Within
// a step the two waves of a SIMD then overlap one's LDS reads with the
// other's MFMAs, where in lockstep both read right after the barrier and the
// MFMA pipe idles.  Same registers (the fragments live across the barrier),
// same barriers and buffer lifetimes: every read of a stage still completes
// before the barrier that precedes the refill of its buffer.  Off by
// default: alternated runs on one MI355X measured it neutral at the
// human-scale index as one 65,536-read launch (332.6 / 332.6 ms against
// 330.2 / 334.8 ms), 2.7 % slower at 100 Mbp (13.11 against 12.75 ms), and,
// with L2-resident query tiles, 16,384-read launches at 80.0 / 80.2 ms
// against 77.9 / 77.8 ms (100 Mbp 12.52 against 12.11 ms).
#ifndef RM_LAG
#define RM_LAG 0
#endif
// RM_SETPRIO: raise the wave priority while it issues a block of MFMAs.
// Off: 16,384-read launches at human scale 78.8 / 79.1 ms against 77.9 /
// 77.8 ms, 100 Mbp 12.30 against 12.11 ms (alternated, one MI355X).
#ifndef RM_SETPRIO
#define RM_SETPRIO 0
#endif
#ifndef RM_LAG16
#define RM_LAG16 0
#endif
template <int ELEM>
constexpr bool kLag = ELEM == 0 ? RM_LAG != 0 : RM_LAG16 != 0;

template <int ELEM, int AI>
__device__ __forceinline__ void zero_acc(v16f (&acc)[AI][2]) {
    const v16f zero = {};
#pragma unroll
    for (int ai = 0; ai < AI; ++ai)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj) acc[ai][bj] = zero;
}

// A lagging wave's tile tail: the MFMAs of the tile's last stage.
template <int ELEM, class SH, bool LAG>
__device__ __forceinline__ void ring_lag_tail(v16f (&acc)[SH::AI][2], const Frags<SH::AI> &f) {
    if constexpr (LAG) mma_part<ELEM, SH::AI, 0, SH::AI, false>(acc, f);
}

// One stage: wait until stage s (issued kRing - 1 stages ago) has landed in
// every wave's rows, then load stage s + kRing - 1 into the buffer stage s - 1
// left and run stage s's MFMAs (a lagging wave: stage s-1's, then read stage
// s).  The restrict pointers give the fragment reads an alias scope apart
// from the loads, so the compiler does not wait for the loads in flight
// before each read.  LOAD false: no load (the streamed loop's last stage).
template <int ELEM, class SH, bool LAG, bool FIRST, bool LOAD = true>
__device__ __forceinline__ void ring_stage(v16f (&acc)[SH::AI][2], Frags<SH::AI> &f,
                                           const RingSrc &d, uint8_t *__restrict__ bw,
                                           const uint8_t *__restrict__ br, int kb, int wi, int wr,
                                           int lane) {
    constexpr int AI = SH::AI, H = (AI + 1) / 2;
    ring_wait<SH>(d.wid);
    if constexpr (LAG) {
        if constexpr (LOAD) ring_load<SH>(d, bw, kb);
        if constexpr (FIRST) {
            zero_acc<ELEM, AI>(acc);
        } else {
            if (RM_SETPRIO) __builtin_amdgcn_s_setprio(1);
            mma_part<ELEM, AI, 0, AI, false>(acc, f);
            if (RM_SETPRIO) __builtin_amdgcn_s_setprio(0);
        }
        ld_frags_b<ELEM, AI>(f, br, wr, lane);
        ld_frags_a<ELEM, AI, 0, AI>(f, br, wi, lane);
        return;
    }
    ld_frags_b<ELEM, AI>(f, br, wr, lane);
    ld_frags_a<ELEM, AI, 0, H>(f, br, wi, lane);
    if constexpr (LOAD) ring_load<SH>(d, bw, kb);
    if (RM_SETPRIO) __builtin_amdgcn_s_setprio(1);
    mma_part<ELEM, AI, 0, H, FIRST>(acc, f);
    if (RM_SETPRIO) __builtin_amdgcn_s_setprio(0);
    ld_frags_a<ELEM, AI, H, AI>(f, br, wi, lane);
    if (RM_SETPRIO) __builtin_amdgcn_s_setprio(1);
    mma_part<ELEM, AI, H, AI, FIRST>(acc, f);
    if (RM_SETPRIO) __builtin_amdgcn_s_setprio(0);
}

template <int ELEM, class SH, bool LAG>
__device__ __forceinline__ void tile_product_ring(v16f (&acc)[SH::AI][2], const GemmArgs &a,
                                                  int64_t i0, int64_t r0, uint8_t *ring, int tid,
                                                  int wi, int wr, int lane) {
    TileSrc ts;
    ts.db = rows_rsrc(a.db, i0, a.N, kBI, a.ldb);
    ts.q = rows_rsrc(a.q, r0, a.M, SH::BR, a.ldq);
    ts.ldb = (int)a.ldb;
    ts.ldq = (int)a.ldq;
    const RingSrc d = ring_src<SH>(ts, tid);
    const int nk = a.Kb / kBK;  // >= 1
    constexpr int SB = SH::STAGE_BYTES;
    uint8_t *b0 = ring, *b1 = ring + SB, *b2 = ring + 2 * SB, *b3 = ring + 3 * SB;
    // stage s loads stage s + kRing - 1 (past the end: the last stage again,
    // into a buffer no later stage reads)
    auto kl = [&](int s) { return min(s + kRing - 1, nk - 1) * kBK; };
    Frags<SH::AI> f;
    // the callers' last barrier ordered every earlier use of the ring
    ring_load<SH>(d, b0, 0);
    ring_load<SH>(d, b1, min(1, nk - 1) * kBK);
    ring_load<SH>(d, b2, min(2, nk - 1) * kBK);
    ring_stage<ELEM, SH, LAG, true>(acc, f, d, b3, b0, kl(0), wi, wr, lane);
    int s = 1;
    for (; s + 3 < nk; s += 4) {
        ring_stage<ELEM, SH, LAG, false>(acc, f, d, b0, b1, kl(s), wi, wr, lane);
        ring_stage<ELEM, SH, LAG, false>(acc, f, d, b1, b2, kl(s + 1), wi, wr, lane);
        ring_stage<ELEM, SH, LAG, false>(acc, f, d, b2, b3, kl(s + 2), wi, wr, lane);
        ring_stage<ELEM, SH, LAG, false>(acc, f, d, b3, b0, kl(s + 3), wi, wr, lane);
    }
    if (s < nk) ring_stage<ELEM, SH, LAG, false>(acc, f, d, b0, b1, kl(s), wi, wr, lane);
    if (s + 1 < nk) ring_stage<ELEM, SH, LAG, false>(acc, f, d, b1, b2, kl(s + 1), wi, wr, lane);
    if (s + 2 < nk) ring_stage<ELEM, SH, LAG, false>(acc, f, d, b2, b3, kl(s + 2), wi, wr, lane);
    // every load of the tile (the clamped ones too) has landed and every
    // fragment read is done before the epilogue reuses the ring
    __builtin_amdgcn_sched_barrier(0);
    asm volatile("s_waitcnt vmcnt(0) lgkmcnt(0)\n\ts_barrier" ::: "memory");
    __builtin_amdgcn_sched_barrier(0);
    ring_lag_tail<ELEM, SH, LAG>(acc, f);
}

// The same K loop as one stream over the tiles of an item (K a multiple of
// 4 stages): the last stages of a tile load the next tile's stages 0 and 1,
// and the next tile loads its stage 2 on entry, so no tile waits for its
// first stage and no load is wasted.  During the epilogue the next tile's
// stages 0-1 may be in flight into buffers 0-1; buffers 2-3 are idle (the
// epilogue's scratch).  The last tile of an item loads its last stage again
// in place of the next tile's (into buffers 0-1, which the next item's
// first tile refills after the item's closing barrier) so that every stage
// keeps two younger stages in flight for ring_wait.
template <int ELEM, class SH, bool LAG>
__device__ __forceinline__ void tile_product_ring_stream(v16f (&acc)[SH::AI][2], const GemmArgs &a,
                                                         int64_t i0, int64_t r0, uint8_t *ring,
                                                         int tid, int wi, int wr, int lane,
                                                         bool first, bool has_next) {
    TileSrc ts;
    ts.db = rows_rsrc(a.db, i0, a.N, kBI, a.ldb);
    ts.q = rows_rsrc(a.q, r0, a.M, SH::BR, a.ldq);
    ts.ldb = (int)a.ldb;
    ts.ldq = (int)a.ldq;
    const RingSrc d = ring_src<SH>(ts, tid);
    const int nk = a.Kb / kBK;  // a multiple of 4
    RingSrc dn = d;             // next tile (or this one again)
    int kn = (nk - 1) * kBK;    // next tile's stage 0 (or this tile's last)
    if (has_next) {
        ts.db = rows_rsrc(a.db, i0 + kBI, a.N, kBI, a.ldb);
        dn = ring_src<SH>(ts, tid);
        kn = 0;
    }
    constexpr int SB = SH::STAGE_BYTES;
    uint8_t *b0 = ring, *b1 = ring + SB, *b2 = ring + 2 * SB, *b3 = ring + 3 * SB;
    Frags<SH::AI> f;
    // the callers' last barrier ordered every earlier use of buffers 2-3
    // (and of 0-1 on an item's first tile; this wave's loads into its rows of
    // 0-1 from the previous item have landed)
    if (first) {
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        ring_load<SH>(d, b0, 0);
        ring_load<SH>(d, b1, kBK);
    }
    ring_load<SH>(d, b2, 2 * kBK);
    ring_stage<ELEM, SH, LAG, true>(acc, f, d, b3, b0, 3 * kBK, wi, wr, lane);
    for (int s = 1; s + 3 < nk; s += 4) {
        ring_stage<ELEM, SH, LAG, false>(acc, f, d, b0, b1, (s + 3) * kBK, wi, wr, lane);
        ring_stage<ELEM, SH, LAG, false>(acc, f, d, b1, b2, (s + 4) * kBK, wi, wr, lane);
        ring_stage<ELEM, SH, LAG, false>(acc, f, d, b2, b3, (s + 5) * kBK, wi, wr, lane);
        ring_stage<ELEM, SH, LAG, false>(acc, f, d, b3, b0, (s + 6) * kBK, wi, wr, lane);
    }
    // stages nk-3 .. nk-1 (buffers 1, 2, 3) load the next tile's 0 and 1
    ring_stage<ELEM, SH, LAG, false>(acc, f, dn, b0, b1, kn, wi, wr, lane);
    ring_stage<ELEM, SH, LAG, false>(acc, f, dn, b1, b2, kn + (has_next ? kBK : 0), wi, wr, lane);
    ring_stage<ELEM, SH, LAG, false, false>(acc, f, dn, b2, b3, 0, wi, wr, lane);
    // every fragment read is done before the epilogue uses buffers 2-3
    __builtin_amdgcn_sched_barrier(0);
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
    __builtin_amdgcn_sched_barrier(0);
    ring_lag_tail<ELEM, SH, LAG>(acc, f);
}

// Float <-> int key with the same order (signed int compare): negative
// floats get their magnitude bits flipped.
__device__ __forceinline__ int order_key(float f) {
    const int b = __float_as_int(f);
    return b >= 0 ? b : b ^ 0x7fffffff;
}
__device__ __forceinline__ float order_key_inv(int k) {
    return __int_as_float(k >= 0 ? k : k ^ 0x7fffffff);
}

// Block-wide OR of a predicate with one barrier: waves with a true lane set
// flag slot `k % 3`; thread 0 clears the next slot for a later vote.
__device__ __forceinline__ bool block_any(bool p, int *flags, int &k) {
    const int slot = k % 3;
    if (threadIdx.x == 0) flags[(slot + 1) % 3] = 0;
    if (__builtin_amdgcn_read_exec() && __ballot(p) && (threadIdx.x & 63) == __builtin_ffsll(__ballot(p)) - 1)
        atomicOr(&flags[slot], 1);
    __syncthreads();
    ++k;
    return flags[slot] != 0;
}

// Accumulator i of block (ai, bj) at lane l: index row
// wi*32*AI + ai*32 + 8*(i/4) + 4*(l/32) + (i%4), query wr*64 + bj*32 + l%32
// (the 32x32 MFMA output layout).
template <int AI>
__device__ __forceinline__ int acc_row(int wi, int ai, int i, int lane) {
    return wi * (32 * AI) + ai * 32 + 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3);
}

// Decode a workgroup's (row tile, split) from the launch.  Dispatch is
// round-robin over the 8 XCDs, so workgroup g runs on XCD g % 8; the logical
// item L = (g % 8) * (G / 8) + g / 8 makes an XCD's concurrently resident
// workgroups consecutive in L.  Items are ordered (query group, split,
// query tile in group): the resident workgroups of an XCD then cover qg query
// tiles x S_xcd splits, so each query tile and each index tile is fetched
// into that XCD's L2 once and shared.
__device__ __forceinline__ bool decode_item(const GemmArgs &a, int64_t g, int64_t G, int64_t &rt,
                                            int &split) {
    const int64_t L = (g & 7) * (G >> 3) + (g >> 3);
    const int64_t per = (int64_t)a.qg * a.S;
    const int64_t grp = L / per, rem = L % per;
    split = (int)(rem / a.qg);
    rt = grp * a.qg + rem % a.qg;
    return grp * per < (int64_t)a.ngroups * per && rt < a.R;
}

template <int ELEM, class SH>
__device__ __forceinline__ void store_item(const GemmArgs &a, int64_t rt, int split) {
    extern __shared__ __align__(16) uint8_t smem[];
    constexpr int AI = SH::AI;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6, wi = wid % SH::WI,
              wr = wid / SH::WI;
    uint8_t *sb0 = smem, *sb1 = smem + SH::STAGE_BYTES;
    // the ring base is smem plus a wave-uniform opaque zero (keeps the
    // fragment reads' alias scopes, as in the top-k kernel)
    int zero = 0;
    asm volatile("" : "+s"(zero));
    uint8_t *const ring = smem + zero;
    const int64_t r0 = rt * SH::BR;
    const int64_t t0 = (int64_t)split * a.tiles / a.S, t1 = (int64_t)(split + 1) * a.tiles / a.S;
    v16f acc[AI][2];
    for (int64_t t = t0; t < t1; ++t) {
        const int64_t i0 = t * kBI;
        // the ring's K loop ends with every load landed and every fragment
        // read done (a barrier), so the next tile may refill it at once
        if constexpr (SH::RING && kRingOk)
            tile_product_ring<ELEM, SH, false>(acc, a, i0, r0, ring, tid, wi, wr, lane);
        else
            tile_product<ELEM, SH>(acc, a, i0, r0, sb0, sb1, tid, wi, wr, lane);
#pragma unroll
        for (int bj = 0; bj < 2; ++bj) {
            const int64_t q = r0 + wr * 64 + bj * 32 + (lane & 31);
            if (q >= a.M) continue;
#pragma unroll
            for (int ai = 0; ai < AI; ++ai)
#pragma unroll
                for (int i = 0; i < 16; ++i) {
                    const int64_t col = i0 + acc_row<AI>(wi, ai, i, lane);
                    if (col < a.N) {
                        const int64_t o = q + col * a.ldc;
                        if (a.c_dtype == RM_F32) ((float *)a.C)[o] = acc[ai][bj][i];
                        else ((uint16_t *)a.C)[o] = f32_to_f16_bits(acc[ai][bj][i]);
                    }
                }
        }
        if constexpr (!(SH::RING && kRingOk)) __syncthreads();  // the next prologue refills the stages
    }
}

// Hardware slot `it` walks the padded item range with the grid's stride: a
// grid of one workgroup per item runs each item once; a persistent grid
// (a multiple of 8 workgroups, at most one per CU) keeps every workgroup on
// its XCD's item range, since the stride preserves it % 8.
template <int ELEM, class SH>
__global__ __launch_bounds__(SH::NT, 1) void gemm_store_kernel(GemmArgs a) {
    for (int64_t it = blockIdx.x; it < a.items8; it += gridDim.x) {
        int64_t rt;
        int split;
        if (!decode_item(a, it, a.items8, rt, split)) continue;
        store_item<ELEM, SH>(a, rt, split);
    }
}

// max(a, b, c) of accumulators in one v_max3_f32.  fmaxf first quiets each
// operand it cannot prove canonical (an MFMA result) with v_max_f32 x, x, x,
// which tripled the per-tile filter's maximum.  The accumulators are never
// signalling NaNs, and v_max3_f32 returns the larger operand of a quiet NaN
// and a number, as fmaxf does.
__device__ __forceinline__ float max3_acc(float a, float b, float c) {
    float r;
    asm("v_max3_f32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}

// The median of a, b, c (v_med3_f32), for the same accumulators.
__device__ __forceinline__ float med3_acc(float a, float b, float c) {
    float r;
    asm("v_med3_f32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}

// An opaque copy of v.  Addresses computed from it stay in the rarely taken
// paths that use them, instead of being hoisted out of the tile loop and
// held in registers (or spilled) across the K loop.
__device__ __forceinline__ int opaque(int v) {
    asm volatile("" : "+v"(v));
    return v;
}

// A wave-uniform opaque copy of v (an SGPR).  The ring base is smem plus an
// opaque zero: with the constant smem, interprocedural constant propagation
// would substitute it into a stage's restrict buffer parameter wherever every
// call passes the same buffer, and the fragment reads would lose the alias
// scope that keeps the compiler from waiting for every LDS load in flight.
__device__ __forceinline__ int opaque_uniform(int v) {
    asm volatile("" : "+s"(v));
    return v;
}

// LDS index of entry j of query q's list: rows of LP = k' + 1 entries, or
// (ROT, lists staged in the ring, k' <= 32) rows of 32 entries rotated by q.
// A wave's access is then free of bank conflicts both with one lane per entry
// of two consecutive queries (staging, the lane merge) and with one lane per
// query at the same j.
template <bool ROT>
struct ListAt {
    int LP;
    __device__ __forceinline__ int operator()(int q, int j) const {
        return ROT ? q * 32 + ((j + q) & 31) : q * LP + j;
    }
};

// Merge every query's candidate slots into its sorted list (one thread per
// query), reset the slot counters and raise tau to the list's k'-th value.
template <bool ROT>
__device__ __forceinline__ void merge_slots(float *lv, int32_t *lc, const float *cv,
                                            const int32_t *cc, float *tauv, int *cnt, int BR,
                                            int KP, ListAt<ROT> at, int CP, int capd, int tid,
                                            int nt) {
    for (int q = opaque(tid); q < BR; q += nt) {
        const int nc = min(cnt[q], capd);
        if (nc == 0) continue;
        for (int j = 0; j < nc; ++j) {
            const float v = cv[q * CP + j];
            const int32_t c = cc[q * CP + j];
            if (!beats(v, c, lv[at(q, KP - 1)], lc[at(q, KP - 1)])) continue;
            int pos = KP - 1;
            while (pos > 0 && beats(v, c, lv[at(q, pos - 1)], lc[at(q, pos - 1)])) {
                lv[at(q, pos)] = lv[at(q, pos - 1)];
                lc[at(q, pos)] = lc[at(q, pos - 1)];
                --pos;
            }
            lv[at(q, pos)] = v;
            lc[at(q, pos)] = c;
        }
        cnt[q] = 0;
        tauv[q] = lv[at(q, KP - 1)];
    }
}

// The value of lane l - 1 (a whole-wave DPP shift right by one lane; lane 0
// keeps its own value).
__device__ __forceinline__ int lane_up(int x) {
    return __builtin_amdgcn_update_dpp(x, x, 0x138 /* wave_shr:1 */, 0xf, 0xf, false);
}

// OR of x over the wave's 64 lanes, returned wave-uniform (a row_shr scan
// within each row of 16 lanes, then row_bcast:15 and row_bcast:31 carry the
// rows' ORs up to lane 63).  Every lane must be active.
__device__ __forceinline__ uint32_t wave_or32(uint32_t x) {
    x |= __builtin_amdgcn_update_dpp(0u, x, 0x111 /* row_shr:1 */, 0xf, 0xf, false);
    x |= __builtin_amdgcn_update_dpp(0u, x, 0x112 /* row_shr:2 */, 0xf, 0xf, false);
    x |= __builtin_amdgcn_update_dpp(0u, x, 0x114 /* row_shr:4 */, 0xf, 0xf, false);
    x |= __builtin_amdgcn_update_dpp(0u, x, 0x118 /* row_shr:8 */, 0xf, 0xf, false);
    x |= __builtin_amdgcn_update_dpp(0u, x, 0x142 /* row_bcast:15 */, 0xa, 0xf, false);
    x |= __builtin_amdgcn_update_dpp(0u, x, 0x143 /* row_bcast:31 */, 0xc, 0xf, false);
    return __builtin_amdgcn_readlane(x, 63);
}
__device__ __forceinline__ uint64_t wave_or64(uint64_t x) {
    return (uint64_t)wave_or32((uint32_t)(x >> 32)) << 32 | wave_or32((uint32_t)x);
}

// (value, column) as one signed 64-bit key whose order is beats(): the
// value's order key above, INT_MAX - column below (columns are >= 0).  The
// key keeps the value's bits; it orders -0 below +0, which beats() ties, but
// no score is -0: MFMA sums start from +0, and an exact zero sum rounds to +0.
__device__ __forceinline__ int64_t merge_key(float v, int32_t c) {
    return (int64_t)((uint64_t)(uint32_t)order_key(v) << 32 | (uint32_t)(INT_MAX - c));
}
__device__ __forceinline__ float merge_key_value(int64_t k) { return order_key_inv((int)(k >> 32)); }
__device__ __forceinline__ int32_t merge_key_col(int64_t k) { return INT_MAX - (int32_t)(uint32_t)k; }

// The same merge as merge_slots for lists staged as rotated 32-entry rows
// (k' <= 32), with 32 lanes per query: lane j holds entry j as a merge key.
// A candidate's place p is the number of the half-wave's 32 keys above it (a
// ballot), and lanes j > p take lane j - 1's key: a sorted insertion into the
// half-wave's 32 keys (the shift into lane 32 from lane 31 is never taken:
// p >= 0).  Lanes k' and up hold the list's tail past k' (or fillers), which
// is never stored, so no candidate needs a p < k' test.  A half past its last
// candidate inserts fillers ((-inf, INT_MAX), the key of an empty entry):
// above no key, so p = the half's real entries, and the lanes from p on hold
// fillers already.  Keys are distinct, so the lists equal those of one-by-one
// insertion.  All NT threads take part, two queries per wave.
template <int BR, int NT>
__device__ __forceinline__ void merge_slots_lanes(float *lv, int32_t *lc, const float *cv,
                                                  const int32_t *cc, float *tauv, int *cnt,
                                                  int KP, ListAt<true> at, int CP, int capd,
                                                  int tid) {
    const int t = opaque(tid), lane = t & 63, j = lane & 31, half = lane >> 5;
    const int64_t filler = merge_key(-INFINITY, INT_MAX);
    for (int q0 = (t >> 6) * 2; q0 < BR; q0 += NT / 32) {
        const int q = q0 + half;
        const int nc = min(cnt[q], capd);
        const int ncmax = max(__builtin_amdgcn_readlane(nc, 0), __builtin_amdgcn_readlane(nc, 32));
        if (ncmax == 0) continue;
        int64_t e = filler;
        if (j < KP) e = merge_key(lv[at(q, j)], lc[at(q, j)]);
        // candidates are read G at a time, so a merge step waits on LDS once
        // per G candidates
        constexpr int G = 4;
        const int cb = q * CP, cl = max(nc - 1, 0);
        for (int i0 = 0; i0 < ncmax; i0 += G) {
            int64_t xs[G];
#pragma unroll
            for (int u = 0; u < G; ++u) {
                const int ci = cb + min(i0 + u, cl);
                const int64_t k = merge_key(cv[ci], cc[ci]);
                xs[u] = i0 + u < nc ? k : filler;
            }
            // all G keys are formed here: the compiler would otherwise sink
            // each read past the early exits into its own step and wait there
            // (the 256-query tile, at its register limit, ran 1 % slower with
            // the 8 registers this holds)
            static_assert(G == 4, "one operand per key");
            if constexpr (BR < 256)
                asm volatile("" : "+v"(xs[0]), "+v"(xs[1]), "+v"(xs[2]), "+v"(xs[3]));
#pragma unroll
            for (int u = 0; u < G; ++u) {
                if (i0 + u >= ncmax) break;  // wave-uniform
                const int64_t x = xs[u];
                const uint64_t m = __ballot(e > x);
                const int p = __popc((uint32_t)(m >> (32 * half)));
                const int64_t es = (int64_t)((uint64_t)(uint32_t)lane_up((int)(e >> 32)) << 32 |
                                             (uint32_t)lane_up((int)e));
                e = j == p ? x : (j > p ? es : e);
            }
        }
        // a query without candidates keeps its list and tau (the first-tile
        // seed may sit above its still empty list)
        if (nc > 0) {
            if (j < KP) {
                lv[at(q, j)] = merge_key_value(e);
                lc[at(q, j)] = merge_key_col(e);
            }
            if (j == KP - 1) tauv[q] = merge_key_value(e);
            if (j == 0) cnt[q] = 0;
        }
    }
}

// Ring layout: a (query, split) list lives in the split partials between
// merges.  For a merge it is staged into the two ring buffers that are idle
// during the epilogue: from the partials once a merge has stored it, else
// empty.  All threads take part, 32 lanes per query (lane = entry), so each
// wave moves two contiguous list rows per access; the barriers around a merge
// order a store before any later load of the same list.
//
// Stored lists come back by LDS DMA, so a wave has all its loads in flight
// before one wait (loads through registers would wait on each query's, and
// the 256-query tile has no registers to spare for more in flight).  One
// 4-byte load per lane fills two rotated rows, queries 2p and 2p + 1: lane l
// writes position l & 31 of query q = 2p + (l >> 5), i.e. entry
// ((l & 31) - q) & 31.  Entries at or past k' are never read (the lane merge
// and store_lists read j < k'), so those lanes load entry 0 again.  A pair's
// descriptor starts at query 2p's list and ends after the last in-range
// query's, so a query past M reads zeros.
template <int BR, int NT, bool ROT>
__device__ __forceinline__ void stage_lists(float *lv, int32_t *lc, const GemmArgs &a, int64_t r0,
                                            int split, int KP, ListAt<ROT> at, bool stored,
                                            int tid) {
    static_assert(ROT, "staged lists use the rotated 32-entry rows");
    if (stored) {
        constexpr int NW = NT / 64;
        const int wid = __builtin_amdgcn_readfirstlane(tid >> 6);
        const int lane = tid & 63, half = lane >> 5;
        const int64_t qs = (int64_t)a.S * KP;  // entries from one query's list to the next's
        for (int p = wid; p < BR / 2; p += NW) {
            const int64_t qa = r0 + 2 * p;  // wave-uniform
            if (qa >= a.M) break;
            const int q = 2 * p + half, j = ((lane & 31) - q) & 31;
            const int voff = (int)((half * qs + (j < KP ? j : 0)) * 4);
            const int nrec = (int)(((qa + 1 < a.M ? qs : 0) + KP) * 4);
            const int64_t e0 = (qa * a.S + split) * (int64_t)KP;
            __builtin_amdgcn_raw_ptr_buffer_load_lds(
                __builtin_amdgcn_make_buffer_rsrc((void *)(a.part_v + e0), (short)0, nrec, 0x00020000),
                (__attribute__((address_space(3))) void *)(lv + p * 64), 4, voff, 0, 0, 0);
            __builtin_amdgcn_raw_ptr_buffer_load_lds(
                __builtin_amdgcn_make_buffer_rsrc((void *)(a.part_c + e0), (short)0, nrec, 0x00020000),
                (__attribute__((address_space(3))) void *)(lc + p * 64), 4, voff, 0, 0, 0);
        }
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    } else {
        const int t = opaque(tid), j = t & 31;
        for (int q = t >> 5; q < BR; q += NT / 32) {
            if (j >= KP) continue;
            lv[at(q, j)] = -INFINITY;
            lc[at(q, j)] = INT_MAX;
        }
    }
    __syncthreads();
}

template <int BR, int NT, bool ROT>
__device__ __forceinline__ void store_lists(const float *lv, const int32_t *lc, const GemmArgs &a,
                                            int64_t r0, int split, int KP, ListAt<ROT> at,
                                            int tid) {
    const int t = opaque(tid);
    for (int q = t >> 5; q < BR; q += NT / 32) {
        if (r0 + q >= a.M) continue;
        const int64_t base = ((r0 + q) * a.S + split) * (int64_t)KP;
        for (int j = t & 31; j < KP; j += 32) {
            a.part_v[base + j] = lv[at(q, j)];
            a.part_c[base + j] = lc[at(q, j)];
        }
    }
}

// After a merge: a full list's k'-th value is a lower bound of its query's
// global k'-th value (the list holds k' values at or above it); raise the
// query's shared bound to it.  Run by the thread that merged the query.
template <int BR, int NT>
__device__ __forceinline__ void publish_bounds(const GemmArgs &a, const float *tauv, int64_t r0,
                                               int tid) {
    for (int q = opaque(tid); q < BR; q += NT)
        if (r0 + q < a.M && tauv[q] > -INFINITY) atomicMax(&a.bound[r0 + q], order_key(tauv[q]));
}

// floor = the float just below the query's shared bound (values at or above
// the bound pass a strict compare); -inf before any split has published.
template <int BR, int NT>
__device__ __forceinline__ void load_floor(const GemmArgs &a, float *floorv, int64_t r0, int tid) {
    for (int q = opaque(tid); q < BR; q += NT) {
        float f = -INFINITY;
        if (r0 + q < a.M)
            f = nextafterf(order_key_inv(__hip_atomic_load(&a.bound[r0 + q], __ATOMIC_RELAXED,
                                                           __HIP_MEMORY_SCOPE_AGENT)),
                           -INFINITY);
        floorv[q] = f;
    }
}

template <int ELEM, class SH>
__device__ __forceinline__ void topk_item(const GemmArgs &a, int64_t rt, int split) {
    constexpr int BR = SH::BR;
    constexpr int NT = SH::NT;
    constexpr int AI = SH::AI, BB = SH::BB;
    extern __shared__ __align__(16) uint8_t smem[];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6, wi = wid % SH::WI,
              wr = wid / SH::WI;
    const int KP = a.KP, LP = KP + 1, capd = a.capd, CP = capd + 1;
    static_assert(2 * SH::STAGE_BYTES == stage_pair_bytes(BR), "stage layout");
    // RING: the 8-wave tiles stream K through the ring and keep the lists in
    // the split partials; the epilogue's scratch (first-tile seed, lists
    // staged for a merge: BR x 32 values in ring buffer 2, columns in 3) sits
    // in the ring buffers no load is in flight to
    constexpr bool RING = kRingOk && SH::RING;
    constexpr int SB = SH::STAGE_BYTES;
    static_assert(!RING || kRing * SB == ring_bytes(BR), "ring layout");
    static_assert(!RING || BR * 32 * 4 <= SB, "staged lists fit one ring buffer each");
    uint8_t *sb0 = smem, *sb1 = smem + SB;
    uint8_t *scratch = RING ? smem + 2 * SB : sb0;
    float *lv = (float *)(RING ? scratch : smem + list_offset(BR));  // lists
    int32_t *lc = (int32_t *)(RING ? scratch + SB : (uint8_t *)(lv + BR * LP));
    const ListAt<RING> at{LP};
    float *cv = (float *)(smem + (RING ? ring_cand_offset(BR) : cand_offset(BR, KP)));  // BR x (capd+1)
    int32_t *cc = (int32_t *)(cv + BR * CP);
    float *tauv = (float *)(cc + BR * CP);               // BR
    int *cnt = (int *)(tauv + BR);                       // BR
    int *flags = cnt + BR;                               // 3 vote slots
    // ring layout with column splits: a per-query floor from the bound the
    // splits of a query share (any split's k'-th value bounds the query's
    // global k'-th value from below); values below it cannot reach the
    // final top-k', so the filter keeps only values >= the bound
    float *floorv = RING ? (float *)((uint8_t *)tauv + tail_bytes(BR)) : nullptr;  // BR
    const bool share = RING && a.bound != nullptr;

    const int64_t r0 = rt * BR;
    const int64_t t0 = (int64_t)split * a.tiles / a.S, t1 = (int64_t)(split + 1) * a.tiles / a.S;

    if constexpr (!RING)
        for (int i = tid; i < BR * LP; i += NT) { lv[i] = -INFINITY; lc[i] = INT_MAX; }
    // RING: `stored` = the lists in the partials hold this item's merges so
    // far (else they are empty); `staged` = they sit in LDS for this tile
    bool stored = false, staged = false;
    for (int i = tid; i < BR; i += NT) { tauv[i] = -INFINITY; cnt[i] = 0; }
    if (share) load_floor<BR, NT>(a, floorv, r0, tid);
    if (tid < 3) flags[tid] = 0;
    int vote = 0;
    // (the first tile's barriers order this before any use)

    int ql[BB];
    bool qv[BB];
#pragma unroll
    for (int bj = 0; bj < BB; ++bj) {
        ql[bj] = wr * 32 * BB + bj * 32 + (lane & 31);  // this lane's queries
        qv[bj] = r0 + ql[bj] < a.M;
    }
    // accumulator (ai, i) of this lane sits at tile column rowbase + RC(ai, i)
    const int rowbase = wi * (32 * AI) + 4 * (lane >> 5);
#define RC(ai, i) ((ai) * 32 + 8 * ((i) >> 2) + ((i) & 3))
    v16f acc[AI][BB];
    uint8_t *const ring = smem + (RING ? opaque_uniform(0) : 0);
    for (int64_t t = t0; t < t1; ++t) {
        const int64_t i0 = t * kBI;
        if constexpr (RING) {
            // waves 4-7 run one stage behind (RM_LAG): a separate copy of the
            // K loop per role, a wave-uniform branch
            const bool lag = kLag<ELEM> && __builtin_amdgcn_readfirstlane(tid >> 6) >= 4;
            if ((a.Kb / kBK) % 4 == 0) {
                if (lag)
                    tile_product_ring_stream<ELEM, SH, kLag<ELEM>>(acc, a, i0, r0, ring, tid, wi, wr,
                                                                    lane, t == t0, t + 1 < t1);
                else
                    tile_product_ring_stream<ELEM, SH, false>(acc, a, i0, r0, ring, tid, wi, wr,
                                                              lane, t == t0, t + 1 < t1);
            } else {
                if (lag)
                    tile_product_ring<ELEM, SH, kLag<ELEM>>(acc, a, i0, r0, ring, tid, wi, wr, lane);
                else
                    tile_product_ring<ELEM, SH, false>(acc, a, i0, r0, ring, tid, wi, wr, lane);
            }
        } else
            tile_product<ELEM, SH>(acc, a, i0, r0, sb0, sb1, tid, wi, wr, lane);
        // columns of this tile valid for this lane: RC(ai, i) < lim
        const int lim = (int)min<int64_t>(a.N - i0 - rowbase, (int64_t)kBI);
        const int32_t colbase = (int32_t)(a.col_base + i0 + rowbase);
        const bool full = i0 + kBI <= a.N;  // uniform: all columns in range
        if (!full) {
            // columns past N (rows read as zeros) score -inf: never above tau,
            // and they can only lower the first-tile bound
#pragma unroll
            for (int ai = 0; ai < AI; ++ai)
#pragma unroll
                for (int bj = 0; bj < BB; ++bj)
#pragma unroll
                    for (int i = 0; i < 16; ++i)
                        if (RC(ai, i) >= lim) acc[ai][bj][i] = -INFINITY;
        }

        // --- first tile of the item: seed tau with a lower bound of each
        // query's k'-th value, so the first tiles do not append every value.
        // A lane's values of a query form AI groups of 16 (ai), and a query
        // has 16 such groups over its 2 * WI lanes (2 per wave); a group's
        // second largest has 2 values at or above it, so the smallest second
        // maximum has >= 32 values at or above it (the smallest maximum:
        // >= 16).  Values equal to the bound must stay candidates: tau is the
        // next float below.
        if (t == t0 && KP <= 32) {
            int key[BB];
#pragma unroll
            for (int bj = 0; bj < BB; ++bj) {
                float bound = INFINITY;
#pragma unroll
                for (int ai = 0; ai < AI; ++ai) {
                    float m1 = -INFINITY, m2 = -INFINITY;
#pragma unroll
                    for (int i = 0; i < 16; ++i) {
                        // m1 >= m2: the new second largest is the median
                        const float v = acc[ai][bj][i];
                        m2 = med3_acc(m1, m2, v);
                        m1 = max3_acc(m1, v, v);
                    }
                    bound = fminf(bound, KP <= 16 ? m1 : m2);
                }
                key[bj] = order_key(bound);
            }
            int *seed = (int *)scratch;  // idle until the next tile
            __syncthreads();
            for (int i = tid; i < BR; i += NT) seed[i] = INT_MAX;
            __syncthreads();
#pragma unroll
            for (int bj = 0; bj < BB; ++bj)
                if (qv[bj]) atomicMin(&seed[ql[bj]], key[bj]);
            __syncthreads();
            for (int i = tid; i < BR; i += NT) {
                const int k = seed[i];
                const float f = order_key_inv(k);
                if (k != INT_MAX && f > -INFINITY) tauv[i] = nextafterf(f, -INFINITY);
            }
            __syncthreads();
        }

        // --- filter: a query's values are compared one by one only when its
        // tile maximum beats tau.  Strictly above tau: a later column can
        // only beat an equal list entry by value, never by the column rule.
        uint64_t pend[BB];
#pragma unroll
        for (int bj = 0; bj < BB; ++bj) {
            pend[bj] = 0;
            float tau = tauv[ql[bj]];
            if (share) tau = fmaxf(tau, floorv[ql[bj]]);  // floor: strictly above = at or above the bound
            float mx = -INFINITY;
#pragma unroll
            for (int ai = 0; ai < AI; ++ai)
#pragma unroll
                for (int i = 0; i < 16; i += 2)
                    mx = max3_acc(mx, acc[ai][bj][i], acc[ai][bj][i + 1]);
            if (qv[bj] && mx > tau) {
#pragma unroll
                for (int ai = 0; ai < AI; ++ai)
#pragma unroll
                    for (int i = 0; i < 16; ++i)
                        if (acc[ai][bj][i] > tau) pend[bj] |= 1ull << (ai * 16 + i);
            }
        }
        // One vote per tile: lanes append without a prior vote (a slot is
        // reserved with an atomic, and merges only run after the vote below)
        for (;;) {
            // --- append flagged values to the query's slots.  A position
            // (ai, i) no lane of the wave flags is skipped on the wave-uniform
            // OR of the flags, with a scalar test instead of a per-lane one
#pragma unroll
            for (int bj = 0; bj < BB; ++bj) {
                if (!__ballot(pend[bj] != 0)) continue;  // wave-uniform
                const uint64_t any_lane = wave_or64(pend[bj]);
                const int q = ql[bj];
                // one atomic reserves a slot per flagged value; the first
                // capd - slot of them (in (ai, i) order) are written, the
                // rest stay pending for the retry after the merge
                int slot = capd;
                if (pend[bj]) slot = atomicAdd(&cnt[q], __popcll(pend[bj]));
#pragma unroll
                for (int ai = 0; ai < AI; ++ai) {
                    if (!((any_lane >> (ai * 16)) & 0xffffull)) continue;
#pragma unroll
                    for (int i = 0; i < 16; ++i) {
                        const uint64_t bit = 1ull << (ai * 16 + i);
                        if (!(any_lane & bit)) continue;  // wave-uniform
                        if ((pend[bj] & bit) && slot < capd) {
                            cv[q * CP + slot] = acc[ai][bj][i];
                            cc[q * CP + slot] = colbase + RC(ai, i);
                            pend[bj] &= ~bit;
                            ++slot;
                        }
                    }
                }
            }
            uint64_t any = 0;
#pragma unroll
            for (int bj = 0; bj < BB; ++bj) any |= pend[bj];
            // slots of every query still hold what they took: merge only when
            // some value found its query's slots full
            if (!block_any(any != 0ull, flags, vote)) break;
            if constexpr (RING) {
                if (!staged) stage_lists<BR, NT>(lv, lc, a, r0, split, KP, at, stored, tid);
                staged = true;
            }
            if constexpr (RING)
                merge_slots_lanes<BR, NT>(lv, lc, cv, cc, tauv, cnt, KP, at, CP, capd, tid);
            else
                merge_slots(lv, lc, cv, cc, tauv, cnt, BR, KP, at, CP, capd, tid, NT);
            __syncthreads();
            if (share) publish_bounds<BR, NT>(a, tauv, r0, tid);
            // --- overflowed values retry against the raised tau (>=: an equal
            // value of this tile may still win on its lower column)
#pragma unroll
            for (int bj = 0; bj < BB; ++bj) {
                if (!pend[bj]) continue;
                const float tau = tauv[ql[bj]];
                const float fl = share ? floorv[ql[bj]] : -INFINITY;
#pragma unroll
                for (int ai = 0; ai < AI; ++ai) {
                    if (!((pend[bj] >> (ai * 16)) & 0xffffull)) continue;
#pragma unroll
                    for (int i = 0; i < 16; ++i) {
                        const uint64_t bit = 1ull << (ai * 16 + i);
                        if ((pend[bj] & bit) && !(acc[ai][bj][i] >= tau && acc[ai][bj][i] > fl))
                            pend[bj] &= ~bit;
                    }
                }
            }
        }
        if constexpr (RING) {
            // the next tile's loads refill the ring: lists go back first
            if (staged) {
                store_lists<BR, NT>(lv, lc, a, r0, split, KP, at, tid);
                stored = true;
                staged = false;
                __syncthreads();
            }
        }
        // every tile: the floor follows the shared bound (the next tile's
        // barriers order the write before the next filter).  Splits of mid-size
        // batches sweep few tiles, so the sooner a split sees the others'
        // bounds the fewer of its warm-up values it appends and merges
        // (256 / 1,024 reads at human scale: 2.70 / 6.70 ms against 2.86-2.93
        // / 7.12-7.14 ms refreshing every 4 tiles)
        if (share) load_floor<BR, NT>(a, floorv, r0, tid);
    }
#undef RC
    __syncthreads();
    if constexpr (RING) {
        stage_lists<BR, NT>(lv, lc, a, r0, split, KP, at, stored, tid);
        merge_slots_lanes<BR, NT>(lv, lc, cv, cc, tauv, cnt, KP, at, CP, capd, tid);
    } else {
        merge_slots(lv, lc, cv, cc, tauv, cnt, BR, KP, at, CP, capd, tid, NT);
    }
    __syncthreads();
    if (share) publish_bounds<BR, NT>(a, tauv, r0, tid);
    store_lists<BR, NT>(lv, lc, a, r0, split, KP, at, tid);
}

template <int ELEM, class SH>
__global__ __launch_bounds__(SH::NT, 1) void gemm_topk_kernel(GemmArgs a) {
    for (int64_t it = blockIdx.x; it < a.items8; it += gridDim.x) {
        int64_t rt;
        int split;
        if (!decode_item(a, it, a.items8, rt, split)) continue;
        topk_item<ELEM, SH>(a, rt, split);
        __syncthreads();  // the next item re-initialises LDS (and refills the ring)
    }
}

static int launch_common(const GemmArgs &a0, const void *fn, size_t lds, int nt, hipStream_t st) {
    GemmArgs a = a0;
    const int64_t items = (int64_t)a.ngroups * a.qg * a.S;
    if (items == 0) return RM_OK;
    a.items8 = cdiv(items, 8) * 8;
    int64_t G = a.items8;
    if (a.persist_grid >= 8) G = std::min<int64_t>(G, a.persist_grid / 8 * 8);
    RM_CHECK(G * nt <= kMaxWorkItems, RM_ERR_SHAPE, "grid too large (%lld workgroups)", (long long)G);
    RM_CHECK(lds <= 160 * 1024, RM_ERR_SHAPE, "fused kernel needs %zu B of LDS", lds);
    RM_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    void *args[] = {(void *)&a};
    RM_HIP(hipLaunchKernel(fn, dim3((unsigned)G), dim3(nt), args, lds, st));
    return RM_OK;
}

// Kernel shapes: 8 waves (two per SIMD) over tiles of 256 index rows x 256,
// 128 or 64 queries, fed by the load ring; lists longer than 32 use 4 waves
// x 128 x 64 (128 queries per tile) with register staging.

// order key of -inf: the shared bounds' starting value
constexpr int32_t kNegInfKey = (int32_t)(0xff800000u ^ 0x7fffffffu);

__global__ void fill_i32_kernel(int32_t *p, int64_t n, int32_t v) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

template <class SH>
static const void *topk_fn(int elem) {
    return elem == 0 ? (const void *)gemm_topk_kernel<0, SH> : (const void *)gemm_topk_kernel<1, SH>;
}

// k' <= 32: the ring kernel of bq queries per tile (with the unswizzled
// LDS layouts: the register-staged kernel of 256)
template <bool RING_OK>
static const void *short_list_fn(int bq, int elem) {
    if constexpr (RING_OK)
        return bq == 256 ? topk_fn<ShapeTopk256>(elem)
               : bq == 128 ? topk_fn<ShapeTopk128>(elem) : topk_fn<ShapeTopk64>(elem);
    else
        return topk_fn<Shape<512, 256, false>>(elem);
}

int launch_gemm_topk(const GemmArgs &a0, int elem, hipStream_t st) {
    GemmArgs a = a0;
    RM_CHECK(tile_rows_ok(a.KP, a.BQ), RM_ERR_ARG, "no fused kernel for %d queries per tile at k' = %d",
             a.BQ, a.KP);
    RM_CHECK(a.R == cdiv(a.M, a.BQ), RM_ERR_ARG, "row tiles %lld != ceil(M / %d)", (long long)a.R,
             a.BQ);
    a.capd = topk_cand_slots(a.KP, a.BQ);
    // list staging addresses two queries' lists with one 32-bit byte offset
    RM_CHECK(((int64_t)a.S + 1) * a.KP * 4 < INT32_MAX, RM_ERR_ARG, "%d splits at k' = %d", a.S, a.KP);
    a.bound = nullptr;
    if (uses_ring(a.KP) && a.S > 1 && a.M > 0) {
        // the bounds follow the partials (part_bytes holds M more entries);
        // a fill kernel rather than a memset, so a captured graph replays it
        a.bound = (int32_t *)(a.part_v + (size_t)a.M * a.S * a.KP);
        fill_i32_kernel<<<dim3(stride_grid(a.M, 256)), dim3(256), 0, st>>>(a.bound, a.M, kNegInfKey);
        RM_HIP(hipGetLastError());
    }
    RM_CHECK(a.capd >= 1, RM_ERR_SHAPE, "no LDS left for candidate slots (k' = %d)", a.KP);
    const void *fn;
    int nt = 512;
    if (!ring_list(a.KP)) {
        fn = topk_fn<ShapeList128>(elem);
        nt = ShapeList128::NT;
    } else {
        RM_CHECK(kRingOk || a.BQ == 256, RM_ERR_ARG, "unswizzled LDS layouts run 256 queries per tile");
        fn = short_list_fn<kRingOk>(a.BQ, elem);
    }
    return launch_common(a, fn, topk_lds_bytes(a.KP, a.BQ), nt, st);
}
int launch_gemm_store(const GemmArgs &a, int elem, hipStream_t st) {
    const void *fn;
    fn = elem == 0 ? (const void *)gemm_store_kernel<0, ShapeStore>
                   : (const void *)gemm_store_kernel<1, ShapeStore>;
    return launch_common(a, fn, store_lds_bytes(), ShapeStore::NT, st);
}

// ---------------------------------------------------------------------------
// Split merge (+ FP32 rescoring): one wave per query.  The S partial lists of
// a query are each sorted; a k'-step tournament over the list heads yields
// the exact top-k' (a column lives in exactly one split).  RM_MODE_FP8_RESCORE
// then recomputes each candidate's score as the FP32 dot of the
// pre-quantisation query with the dequantised e4m3 column in a fixed order
// (lane l owns K elements [l*C, (l+1)*C), sequential multiply-then-add, xor
// butterfly) and re-ranks under (score desc, column asc).
// ---------------------------------------------------------------------------
// rescoring column element: CE = RM_E4M3 (the e4m3 database) or RM_F16 (the
// verbatim fp16 columns of an fp16 index); ldb in elements
template <int CE>
__device__ __forceinline__ float col_elem(const void *db, int64_t o) {
    if constexpr (CE == RM_F16) return f16_bits_to_f32(((const uint16_t *)db)[o]);
    else return e4m3_to_f32(((const uint8_t *)db)[o]);
}

template <int CE>
__global__ __launch_bounds__(64) void merge_kernel(const float *__restrict__ pv,
                                                   const int32_t *__restrict__ pc, int S, int KP,
                                                   int64_t M, int64_t N, int ktop, int mode,
                                                   const float *__restrict__ q32, int64_t ldq32,
                                                   const void *__restrict__ db, int64_t ldb, int K,
                                                   float *__restrict__ vals,
                                                   int32_t *__restrict__ locs, int64_t ldo) {
    constexpr int ESZ = CE == RM_F16 ? 2 : 1;  // bytes per column element
    const uint8_t *__restrict__ dbb = (const uint8_t *)db;
    __shared__ float mv[kMaxKP];
    __shared__ int32_t mc[kMaxKP];
    __shared__ float sc[kMaxKP];
    extern __shared__ float qs[];  // K floats (rescoring only)
    const int64_t r = blockIdx.x;
    const int lane = threadIdx.x;
    if (r >= M) return;
    const int64_t base = r * (int64_t)S * KP;
    int hp[4] = {0, 0, 0, 0};
    int j = 0;
    for (; j < KP; ++j) {
        float bv = -INFINITY;
        int32_t bc = INT_MAX;
        int bu = -1;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int s = lane + 64 * u;
            if (s < S && hp[u] < KP) {
                const float v = pv[base + (int64_t)s * KP + hp[u]];
                const int32_t c = pc[base + (int64_t)s * KP + hp[u]];
                if (beats(v, c, bv, bc)) { bv = v; bc = c; bu = u; }
            }
        }
        float wv = bv;
        int32_t wc = bc;
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            const float ov = __shfl_xor(wv, off, 64);
            const int32_t oc = __shfl_xor(wc, off, 64);
            if (beats(ov, oc, wv, wc)) { wv = ov; wc = oc; }
        }
        if (wc == INT_MAX) break;  // only empty slots remain
        if (bu >= 0 && bc == wc) hp[bu]++;  // columns are unique across splits
        // a column id outside [0, N) cannot come from the fused kernel; it is
        // kept out of the gathers below
        if (lane == 0) { mv[j] = wv; mc[j] = (wc >= 0 && wc < N) ? wc : INT_MAX; }
    }
    for (int t = j + lane; t < KP; t += 64) { mv[t] = -INFINITY; mc[t] = INT_MAX; }
    __syncthreads();

    if (mode == RM_MODE_FP8_RESCORE) {
        const float *qr = q32 + r * ldq32;
        const int C = (K + 63) / 64;  // elements per lane: [lane*C, lane*C + C)
        const int e_lo = lane * C;
        if (C % 16 == 0 && K % 64 == 0 && KP <= 32) {
            // all k' candidates at once: per 16-element piece of the lane's
            // slice, the query piece is loaded once and the 16-byte pieces of
            // the k' columns are issued before use; each candidate's sum
            // still runs over the lane's elements in increasing order
            float acc[32];
#pragma unroll
            for (int t = 0; t < 32; ++t) acc[t] = 0.0f;
            for (int e0 = 0; e0 < C; e0 += 16) {
                float qv[16];
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const float4 x = *(const float4 *)(qr + e_lo + e0 + 4 * j);
                    qv[4 * j] = x.x; qv[4 * j + 1] = x.y; qv[4 * j + 2] = x.z; qv[4 * j + 3] = x.w;
                }
#pragma unroll
                for (int t0 = 0; t0 < 32; t0 += 8) {
                    if (t0 >= KP) break;
                    uint4 cv[8][ESZ];  // 16 elements of each of 8 candidates
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        const int32_t c = (t0 + u < KP) ? mc[t0 + u] : INT_MAX;
#pragma unroll
                        for (int h = 0; h < ESZ; ++h)
                            cv[u][h] = c == INT_MAX
                                           ? make_uint4(0, 0, 0, 0)
                                           : *(const uint4 *)(dbb + ((int64_t)c * ldb + e_lo + e0) * ESZ +
                                                              16 * h);
                    }
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        uint32_t w[4 * ESZ];
#pragma unroll
                        for (int h = 0; h < ESZ; ++h) {
                            w[4 * h] = cv[u][h].x; w[4 * h + 1] = cv[u][h].y;
                            w[4 * h + 2] = cv[u][h].z; w[4 * h + 3] = cv[u][h].w;
                        }
#pragma unroll
                        for (int j = 0; j < 16; ++j) {
                            float d;
                            if constexpr (CE == RM_F16)
                                d = f16_bits_to_f32((uint16_t)(w[j >> 1] >> (16 * (j & 1))));
                            else
                                d = e4m3_byte_to_f32(w[j >> 2], j & 3);
                            const float p = qv[j] * d;
                            acc[t0 + u] = acc[t0 + u] + p;
                        }
                    }
                }
            }
#pragma unroll
            for (int t = 0; t < 32; ++t) {
                if (t >= KP) break;
                float v = acc[t];
#pragma unroll
                for (int off = 32; off >= 1; off >>= 1) v = v + __shfl_xor(v, off, 64);
                if (lane == 0) sc[t] = (mc[t] == INT_MAX) ? -INFINITY : v;
            }
        } else {
            for (int t = lane; t < K; t += 64) qs[t] = qr[t];
            __syncthreads();
            for (int t = 0; t < KP; ++t) {
                const int32_t c = mc[t];
                if (c == INT_MAX) { if (lane == 0) sc[t] = -INFINITY; continue; }
                const int64_t col = (int64_t)c * ldb;
                float acc = 0.0f;
                for (int e = e_lo; e < e_lo + C && e < K; ++e) {
                    const float p = qs[e] * col_elem<CE>(db, col + e);
                    acc = acc + p;
                }
#pragma unroll
                for (int off = 32; off >= 1; off >>= 1) acc = acc + __shfl_xor(acc, off, 64);
                if (lane == 0) sc[t] = acc;
            }
        }
        __syncthreads();
        if (lane < KP) {
            const float v = sc[lane];
            const int32_t c = mc[lane];
            int rank = 0;  // equal entries (empty slots) are ordered by position
            for (int t = 0; t < KP; ++t)
                rank += (beats(sc[t], mc[t], v, c) || (t < lane && sc[t] == v && mc[t] == c)) ? 1 : 0;
            if (rank < ktop) {
                vals[r + rank * ldo] = v;
                locs[r + rank * ldo] = (c == INT_MAX) ? 0 : c + 1;
            }
        }
    } else {
        if (lane < ktop) {
            vals[r + lane * ldo] = mv[lane];
            locs[r + lane * ldo] = (mc[lane] == INT_MAX) ? 0 : mc[lane] + 1;
        }
    }
}

int launch_merge(const float *part_v, const int32_t *part_c, int S, int KP, int64_t M, int64_t N,
                 int ktop, int mode, const float *q32, int64_t ldq32, const void *db,
                 int db_dtype, int64_t ldb, int K, float *vals, int32_t *locs, int64_t ldo,
                 hipStream_t st) {
    if (M == 0) return RM_OK;
    RM_CHECK(S <= 256, RM_ERR_ARG, "at most 256 column splits (got %d)", S);
    RM_CHECK(M * 64 <= kMaxWorkItems, RM_ERR_SHAPE, "%lld rows exceed one launch", (long long)M);
    RM_CHECK(db_dtype == RM_E4M3 || db_dtype == RM_F16, RM_ERR_ARG, "rescoring columns: e4m3 or fp16");
    size_t lds = (mode == RM_MODE_FP8_RESCORE) ? (size_t)K * 4 : 0;
    if (db_dtype == RM_F16 && mode == RM_MODE_FP8_RESCORE)
        merge_kernel<RM_F16><<<dim3((unsigned)M), dim3(64), lds, st>>>(
            part_v, part_c, S, KP, M, N, ktop, mode, q32, ldq32, db, ldb, K, vals, locs, ldo);
    else
        merge_kernel<RM_E4M3><<<dim3((unsigned)M), dim3(64), lds, st>>>(
            part_v, part_c, S, KP, M, N, ktop, mode, q32, ldq32, db, ldb, K, vals, locs, ldo);
    RM_HIP(hipGetLastError());
    return RM_OK;
}

// ---------------------------------------------------------------------------
// Drop-in running top-k over a C chunk (column-major, rows = queries), in two
// passes.  Pass 1: the chunk's columns are cut into G segments; a wave owns
// 64 rows of one segment (lane = row: one coalesced 64-row load per column,
// eight columns in flight) and keeps each row's top-k of its segment in LDS,
// admitting only values that beat the k-th entry of the running list (when
// full) and of the segment's own list.  Pass 2: one wave per row merges the G
// segment lists and the running list (earlier columns) by a k-step
// tournament under the total order (value desc, column asc), so the result
// equals one sequential scan.  Lists of the segments go through a workspace
// of M x G x k (value, column) pairs.  Before pass 1, rows whose running list
// is not yet full (the first chunk) get a chunk-level admission bound
// (rowtopk_bound_kernel).
// ---------------------------------------------------------------------------
template <int CDT>
__device__ __forceinline__ float load_c(const void *C, int64_t o) {
    if constexpr (CDT == RM_F32) return ((const float *)C)[o];
    else if constexpr (CDT == RM_F16) return f16_bits_to_f32(((const uint16_t *)C)[o]);
    else return e4m3_to_f32(((const uint8_t *)C)[o]);
}

// Pass 0, for rows whose running list is not full yet (the first chunk of a
// batch): cut the chunk's columns into k groups; each group's maximum is a
// distinct value of the chunk, so the smallest of the k maxima has at least k
// values at or above it and bounds the k-th value of (running list + chunk)
// from below.  Pass 1 then admits only values >= it (with column INT_MAX as
// the threshold's column, equal values pass), which skips the insertions of
// the segments' warm-up.  One wave per (64 rows, group); the bound is the
// atomicMin of order keys over the groups (seed preset to INT_MAX).
template <int CDT>
__global__ __launch_bounds__(64) void rowtopk_bound_kernel(const void *C, int64_t ldc, int64_t M,
                                                           int64_t wi, int k, const int32_t *dl,
                                                           int64_t ldd, int32_t *seed) {
    constexpr int U = 8;
    const int64_t row = (int64_t)blockIdx.x * 64 + threadIdx.x;
    const int g = blockIdx.y;
    if (row >= M || dl[row + (int64_t)(k - 1) * ldd] > 0) return;
    const int64_t j0 = wi * g / k, j1 = wi * (g + 1) / k;
    float gm = -INFINITY;
    int64_t j = j0;
    for (; j + U <= j1; j += U) {
        float v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = load_c<CDT>(C, row + (j + u) * ldc);
#pragma unroll
        for (int u = 0; u < U; ++u) gm = fmaxf(gm, v[u]);
    }
    for (; j < j1; ++j) gm = fmaxf(gm, load_c<CDT>(C, row + j * ldc));
    atomicMin(&seed[row], order_key(gm));
}

template <int CDT>
__global__ __launch_bounds__(64) void rowtopk_seg_kernel(const void *C, int64_t ldc,
                                                         int64_t M, int64_t wi, int64_t offset,
                                                         int k, int G, const float *dv,
                                                         const int32_t *dl, int64_t ldd,
                                                         const int32_t *seed, float *pv,
                                                         int32_t *pc) {
    constexpr int U = 8;  // columns per step: their loads are issued together
    extern __shared__ float lsh[];
    const int t = threadIdx.x;
    const int64_t row = (int64_t)blockIdx.x * 64 + t;
    const int g = blockIdx.y;
    const int64_t j0 = (int64_t)g * wi / G, j1 = (int64_t)(g + 1) * wi / G;
    // a lane's list: k values then k columns, 2k + 1 words apart (an odd
    // stride puts the 64 lanes' entries e in distinct LDS banks)
    float *L = lsh + t * (2 * k + 1);
    int32_t *Lc = (int32_t *)(L + k);
    for (int e = 0; e < k; ++e) { L[e] = -INFINITY; Lc[e] = INT_MAX; }
    if (row >= M) return;
    // Admission threshold: the running list's k-th entry when the list is
    // full (a column that does not beat it cannot reach the merged top-k),
    // else the segment list's own k-th entry once that fills.
    float tv = -INFINITY;
    int32_t tc = INT_MAX;
    {
        const int32_t l = dl[row + (int64_t)(k - 1) * ldd];
        if (l > 0) {
            tv = dv[row + (int64_t)(k - 1) * ldd];
            tc = l - 1;
        }
    }
    // no full running list yet (the first chunk): the chunk-level bound of
    // pass 0, when there is one
    if (tc == INT_MAX && seed != nullptr) {
        const int32_t key = seed[row];
        if (key != INT_MAX) tv = order_key_inv(key);
    }
    auto insert = [&](float v, int32_t c) {
        int pos = k - 1;
        while (pos > 0 && beats(v, c, L[pos - 1], Lc[pos - 1])) {
            L[pos] = L[pos - 1];
            Lc[pos] = Lc[pos - 1];
            --pos;
        }
        L[pos] = v;
        Lc[pos] = c;
        if (beats(L[k - 1], Lc[k - 1], tv, tc)) {
            tv = L[k - 1];
            tc = Lc[k - 1];
        }
    };
    const int64_t base = row + j0 * ldc;
    int64_t j = j0;
    for (; j + U <= j1; j += U) {
        float v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = load_c<CDT>(C, base + (j - j0 + u) * ldc);
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int32_t c = (int32_t)(offset + j + u);
            if (beats(v[u], c, tv, tc)) insert(v[u], c);
        }
    }
    for (; j < j1; ++j) {
        const float v = load_c<CDT>(C, base + (j - j0) * ldc);
        const int32_t c = (int32_t)(offset + j);
        if (beats(v, c, tv, tc)) insert(v, c);
    }
    const int64_t o = ((int64_t)g * M + row) * k;
    for (int e = 0; e < k; ++e) { pv[o + e] = L[e]; pc[o + e] = Lc[e]; }
}

// One wave per row: the running list (D_val / D_loc, 1-based ids, 0 =
// empty) and the G segment lists, each sorted; lane u < G + 1 holds the head
// of list u; k steps of a wave-wide argmax under the total order.
__global__ __launch_bounds__(64) void rowtopk_final_kernel(float *dv, int32_t *dl, int64_t ldd,
                                                           int64_t M, int k, int G,
                                                           const float *pv, const int32_t *pc) {
    __shared__ float ov[kMaxKP];
    __shared__ int32_t oc[kMaxKP];
    const int64_t row = blockIdx.x;
    const int lane = threadIdx.x;
    if (row >= M) return;
    int hp = 0;  // this lane's list position
    auto head = [&](float &v, int32_t &c) {
        v = -INFINITY;
        c = INT_MAX;
        if (hp >= k) return;
        if (lane == 0) {
            const int32_t l = dl[row + (int64_t)hp * ldd];
            v = dv[row + (int64_t)hp * ldd];
            c = l > 0 ? l - 1 : INT_MAX;
        } else if (lane <= G) {
            const int64_t o = ((int64_t)(lane - 1) * M + row) * k + hp;
            v = pv[o];
            c = pc[o];
        }
    };
    float hv;
    int32_t hc;
    head(hv, hc);
    for (int e = 0; e < k; ++e) {
        float wv = hv;
        int32_t wc = hc;
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            const float xv = __shfl_xor(wv, off, 64);
            const int32_t xc = __shfl_xor(wc, off, 64);
            if (beats(xv, xc, wv, wc)) { wv = xv; wc = xc; }
        }
        if (lane == 0) { ov[e] = wv; oc[e] = wc; }
        if (wc != INT_MAX && hc == wc && hv == wv) {  // columns are unique: one lane advances
            ++hp;
            head(hv, hc);
        }
    }
    __syncthreads();
    for (int e = lane; e < k; e += 64) {
        dv[row + (int64_t)e * ldd] = ov[e];
        dl[row + (int64_t)e * ldd] = oc[e] == INT_MAX ? 0 : oc[e] + 1;
    }
}

// segments per chunk: about 8192 waves in pass 1 (enough loads in flight to
// stream C from HBM), at least 256 columns each
int rowtopk_segments(int64_t M, int64_t wi) {
    const int64_t rb = cdiv(M, 64);
    int64_t G = cdiv(8192, rb);
    G = std::min<int64_t>(G, std::max<int64_t>(1, wi / 256));
    return (int)std::max<int64_t>(1, std::min<int64_t>(G, 63));
}

int launch_rowtopk(float *dv, int32_t *dl, int64_t ldd, const void *C, int c_dtype, int64_t ldc,
                   int64_t M, int64_t wi, int64_t offset, int k, int G, float *pv, int32_t *pc,
                   int32_t *seed, hipStream_t st) {
    if (M == 0) return RM_OK;
    RM_CHECK(G >= 1 && G <= 63, RM_ERR_ARG, "segments must be 1..63 (got %d)", G);
    // work-items per launch: the final pass M x 64, pass 1 (64 rows per wave)
    // x G segments, pass 0 x k groups
    RM_CHECK(cdiv(M, 64) <= 0x7fffffffLL && M <= 0x7fffffffLL && M * 64 <= kMaxWorkItems &&
                 cdiv(M, 64) * 64 * std::max(G, k) <= kMaxWorkItems,
             RM_ERR_SHAPE, "%lld rows exceed one launch", (long long)M);
    const size_t lds = (size_t)64 * (2 * k + 1) * 4;
    // pass 0 only where it can bind: k groups of at least 2 columns
    if (seed != nullptr && wi >= 2 * (int64_t)k) {
        fill_i32_kernel<<<dim3(stride_grid(M, 256)), dim3(256), 0, st>>>(seed, M, INT_MAX);
        RM_HIP(hipGetLastError());
        const dim3 g0((unsigned)cdiv(M, 64), (unsigned)k);
        if (c_dtype == RM_F32)
            rowtopk_bound_kernel<RM_F32><<<g0, dim3(64), 0, st>>>(C, ldc, M, wi, k, dl, ldd, seed);
        else if (c_dtype == RM_F16)
            rowtopk_bound_kernel<RM_F16><<<g0, dim3(64), 0, st>>>(C, ldc, M, wi, k, dl, ldd, seed);
        else
            rowtopk_bound_kernel<RM_E4M3><<<g0, dim3(64), 0, st>>>(C, ldc, M, wi, k, dl, ldd, seed);
        RM_HIP(hipGetLastError());
    } else {
        seed = nullptr;
    }
    const dim3 grid((unsigned)cdiv(M, 64), (unsigned)G);
    if (c_dtype == RM_F32)
        rowtopk_seg_kernel<RM_F32><<<grid, dim3(64), lds, st>>>(C, ldc, M, wi, offset, k, G, dv, dl,
                                                               ldd, seed, pv, pc);
    else if (c_dtype == RM_F16)
        rowtopk_seg_kernel<RM_F16><<<grid, dim3(64), lds, st>>>(C, ldc, M, wi, offset, k, G, dv, dl,
                                                               ldd, seed, pv, pc);
    else
        rowtopk_seg_kernel<RM_E4M3><<<grid, dim3(64), lds, st>>>(C, ldc, M, wi, offset, k, G, dv,
                                                                dl, ldd, seed, pv, pc);
    RM_HIP(hipGetLastError());
    rowtopk_final_kernel<<<dim3((unsigned)M), dim3(64), 0, st>>>(dv, dl, ldd, M, k, G, pv, pc);
    RM_HIP(hipGetLastError());
    return RM_OK;
}

// dst[r*cols + c] = src[r*s_row + c*s_col] for esize-byte elements.
__global__ void transpose_kernel(uint8_t *dst, const uint8_t *src, int64_t rows, int64_t cols,
                                 int64_t s_row, int64_t s_col, int esize) {
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < rows * cols;
         g += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = g / cols, c = g % cols;
        const uint8_t *s = src + (r * s_row + c * s_col) * esize;
        uint8_t *d = dst + g * esize;
        for (int b = 0; b < esize; ++b) d[b] = s[b];
    }
}

int launch_transpose_u8(uint8_t *dst, const uint8_t *src, int64_t rows, int64_t cols, int64_t s_row,
                        int64_t s_col, int esize, hipStream_t st) {
    const int64_t n = rows * cols;
    if (n == 0) return RM_OK;
    transpose_kernel<<<dim3(stride_grid(n, 256)), dim3(256), 0, st>>>(dst, src, rows, cols, s_row,
                                                                         s_col, esize);
    RM_HIP(hipGetLastError());
    return RM_OK;
}

}  // namespace rm
