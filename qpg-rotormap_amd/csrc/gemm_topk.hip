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
// Tile: 256 index rows x 128 queries per workgroup (4 waves in 2x2, wave
// tile 128 x 64 = 4 x 2 blocks of v_mfma_f32_32x32x64_f8f6f4, or 2 x
// v_mfma_f32_32x32x16_f16 per 64-byte stage for the FP16 path), K streamed
// through double-buffered LDS in 64-byte stages (80-byte row pitch: the
// ds_read_b128 fragment reads of 8 consecutive rows hit 8 distinct 4-bank
// groups).  A workgroup owns one query tile and walks a contiguous range of
// index tiles (its column split); per query it keeps a top-k' list in LDS
// and a threshold tau = the list's k'-th value.  After each tile a lane
// flags values > tau; only flagged values (rare once tau has risen) are
// appended to per-query LDS candidate slots and merged by one thread per
// query under the (score desc, column asc) order.  Workgroups are mapped so
// that the 32 concurrently resident workgroups of one XCD (dispatch is
// round-robin over the 8 XCDs) share a column split: each index tile is
// fetched into that XCD's L2 once and reused by all its query tiles.
#include "rm_device.h"
#include "rm_internal.h"

#include <climits>

namespace rm {

typedef int v8i __attribute__((ext_vector_type(8)));
typedef float v16f __attribute__((ext_vector_type(16)));
typedef _Float16 v8h __attribute__((ext_vector_type(8)));

struct StageRegs {
    uint4 v[6];
};

// global -> registers for one 64-byte K stage: index rows 0..255 (4 x 16 B
// chunks each, 1024 chunks) and query rows 0..127 (512 chunks); out-of-range
// rows read as zeros.
__device__ __forceinline__ void g_load(StageRegs &st, const GemmArgs &a, int64_t i0, int64_t r0,
                                       int kb, int tid) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        int c = tid + 256 * j;
        int row = c >> 2, seg = c & 3;
        int64_t gi = i0 + row;
        uint4 z = make_uint4(0, 0, 0, 0);
        if (gi < a.N) z = *(const uint4 *)(a.db + gi * a.ldb + kb + seg * 16);
        st.v[j] = z;
    }
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        int c = tid + 256 * j;
        int row = c >> 2, seg = c & 3;
        int64_t gr = r0 + row;
        uint4 z = make_uint4(0, 0, 0, 0);
        if (gr < a.M) z = *(const uint4 *)(a.q + gr * a.ldq + kb + seg * 16);
        st.v[4 + j] = z;
    }
}

__device__ __forceinline__ void s_store(const StageRegs &st, uint8_t *sb, int tid) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        int c = tid + 256 * j;
        int row = c >> 2, seg = c & 3;
        *(uint4 *)(sb + row * kRow + seg * 16) = st.v[j];
    }
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        int c = tid + 256 * j;
        int row = c >> 2, seg = c & 3;
        *(uint4 *)(sb + (kBI + row) * kRow + seg * 16) = st.v[4 + j];
    }
}

__device__ __forceinline__ v8i ld_v8i(const uint8_t *p) {
    uint4 x = *(const uint4 *)p, y = *(const uint4 *)(p + 16);
    v8i r;
    r[0] = (int)x.x; r[1] = (int)x.y; r[2] = (int)x.z; r[3] = (int)x.w;
    r[4] = (int)y.x; r[5] = (int)y.y; r[6] = (int)y.z; r[7] = (int)y.w;
    return r;
}
__device__ __forceinline__ v8h ld_v8h(const uint8_t *p) {
    uint4 x = *(const uint4 *)p;
    return __builtin_bit_cast(v8h, x);
}

// One 64-byte K stage of MFMAs for this wave's 4 x 2 blocks.  Lane l feeds
// row (l % 32) of its block with K bytes [(l/32)*32, +32) of the stage (fp8)
// or, per 16-wide f16 step kk, halves [kk*16 + (l/32)*8, +8): the same K
// slice for both operands, so every product pairs matching K indices.
template <int ELEM>
__device__ __forceinline__ void mma_stage(v16f (&acc)[4][2], const uint8_t *sb, int wi, int wr,
                                          int lane) {
    const int rl = lane & 31, hl = lane >> 5;
    if constexpr (ELEM == 0) {
        v8i fa[4], fb[2];
#pragma unroll
        for (int ai = 0; ai < 4; ++ai)
            fa[ai] = ld_v8i(sb + (wi * 128 + ai * 32 + rl) * kRow + hl * 32);
#pragma unroll
        for (int bj = 0; bj < 2; ++bj)
            fb[bj] = ld_v8i(sb + (kBI + wr * 64 + bj * 32 + rl) * kRow + hl * 32);
#pragma unroll
        for (int ai = 0; ai < 4; ++ai)
#pragma unroll
            for (int bj = 0; bj < 2; ++bj)
                acc[ai][bj] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(
                    fa[ai], fb[bj], acc[ai][bj], 0, 0, 0, 0, 0, 0);
    } else {
#pragma unroll
        for (int kk = 0; kk < 2; ++kk) {
            v8h fa[4], fb[2];
#pragma unroll
            for (int ai = 0; ai < 4; ++ai)
                fa[ai] = ld_v8h(sb + (wi * 128 + ai * 32 + rl) * kRow + kk * 32 + hl * 16);
#pragma unroll
            for (int bj = 0; bj < 2; ++bj)
                fb[bj] = ld_v8h(sb + (kBI + wr * 64 + bj * 32 + rl) * kRow + kk * 32 + hl * 16);
#pragma unroll
            for (int ai = 0; ai < 4; ++ai)
#pragma unroll
                for (int bj = 0; bj < 2; ++bj)
                    acc[ai][bj] = __builtin_amdgcn_mfma_f32_32x32x16_f16(fa[ai], fb[bj],
                                                                         acc[ai][bj], 0, 0, 0);
        }
    }
}

// Full-K product of one (256 index x 128 query) tile into acc.
template <int ELEM>
__device__ __forceinline__ void tile_product(v16f (&acc)[4][2], const GemmArgs &a, int64_t i0,
                                             int64_t r0, uint8_t *sb0, uint8_t *sb1, int tid,
                                             int wi, int wr, int lane) {
#pragma unroll
    for (int ai = 0; ai < 4; ++ai)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj)
#pragma unroll
            for (int i = 0; i < 16; ++i) acc[ai][bj][i] = 0.0f;
    const int nk = a.Kb / kBK;
    StageRegs st;
    g_load(st, a, i0, r0, 0, tid);
    s_store(st, sb0, tid);
    __syncthreads();
    for (int ks = 0; ks < nk; ++ks) {
        uint8_t *cur = (ks & 1) ? sb1 : sb0;
        uint8_t *nxt = (ks & 1) ? sb0 : sb1;
        const bool more = ks + 1 < nk;
        if (more) g_load(st, a, i0, r0, (ks + 1) * kBK, tid);
        mma_stage<ELEM>(acc, cur, wi, wr, lane);
        if (more) s_store(st, nxt, tid);
        __syncthreads();
    }
}

// Accumulator i of block (ai, bj) at lane l: index row
// wi*128 + ai*32 + 8*(i/4) + 4*(l/32) + (i%4), query wr*64 + bj*32 + l%32
// (the 32x32 MFMA output layout).
__device__ __forceinline__ int acc_row(int wi, int ai, int i, int lane) {
    return wi * 128 + ai * 32 + 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3);
}

// Decode a workgroup's (row tile, split) from the launch: XCD-grouped grid.
__device__ __forceinline__ bool decode_item(const GemmArgs &a, int64_t &rt, int &split) {
    const int64_t G = gridDim.x, g = blockIdx.x;
    const int64_t L = (g & 7) * (G >> 3) + (g >> 3);
    if (L >= a.R * a.S) return false;
    split = (int)(L / a.R);
    rt = L % a.R;
    return true;
}

template <int ELEM>
__global__ __launch_bounds__(256, 1) void gemm_store_kernel(GemmArgs a) {
    extern __shared__ __align__(16) uint8_t smem[];
    int64_t rt;
    int split;
    if (!decode_item(a, rt, split)) return;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6, wi = wid & 1, wr = wid >> 1;
    uint8_t *sb0 = smem, *sb1 = smem + (kBI + kBR) * kRow;
    const int64_t r0 = rt * kBR;
    const int64_t t0 = (int64_t)split * a.tiles / a.S, t1 = (int64_t)(split + 1) * a.tiles / a.S;
    v16f acc[4][2];
    for (int64_t t = t0; t < t1; ++t) {
        const int64_t i0 = t * kBI;
        tile_product<ELEM>(acc, a, i0, r0, sb0, sb1, tid, wi, wr, lane);
#pragma unroll
        for (int bj = 0; bj < 2; ++bj) {
            const int64_t q = r0 + wr * 64 + bj * 32 + (lane & 31);
            if (q >= a.M) continue;
#pragma unroll
            for (int ai = 0; ai < 4; ++ai)
#pragma unroll
                for (int i = 0; i < 16; ++i) {
                    const int64_t col = i0 + acc_row(wi, ai, i, lane);
                    if (col < a.N) {
                        const int64_t o = q + col * a.ldc;
                        if (a.c_dtype == RM_F32) ((float *)a.C)[o] = acc[ai][bj][i];
                        else ((uint16_t *)a.C)[o] = f32_to_f16_bits(acc[ai][bj][i]);
                    }
                }
        }
    }
}

template <int ELEM>
__global__ __launch_bounds__(256, 1) void gemm_topk_kernel(GemmArgs a) {
    extern __shared__ __align__(16) uint8_t smem[];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6, wi = wid & 1, wr = wid >> 1;
    const int KP = a.KP, LP = KP + 1, CP = kCap + 1;
    uint8_t *sb0 = smem, *sb1 = smem + (kBI + kBR) * kRow;
    float *lv = (float *)(smem + 2 * (kBI + kBR) * kRow);  // kBR x (KP+1)
    int32_t *lc = (int32_t *)(lv + kBR * LP);
    float *cv = (float *)(lc + kBR * LP);                  // kBR x (kCap+1)
    int32_t *cc = (int32_t *)(cv + kBR * CP);
    float *tauv = (float *)(cc + kBR * CP);                // kBR
    int *cnt = (int *)(tauv + kBR);                        // kBR

    int64_t rt;
    int split;
    if (!decode_item(a, rt, split)) return;
    const int64_t r0 = rt * kBR;
    const int64_t t0 = (int64_t)split * a.tiles / a.S, t1 = (int64_t)(split + 1) * a.tiles / a.S;

    for (int i = tid; i < kBR * LP; i += 256) { lv[i] = -INFINITY; lc[i] = INT_MAX; }
    if (tid < kBR) { tauv[tid] = -INFINITY; cnt[tid] = 0; }
    // (the first tile's barriers order this before any use)

    const int ql0 = wr * 64 + (lane & 31), ql1 = ql0 + 32;  // this lane's two queries
    const bool qv0 = r0 + ql0 < a.M, qv1 = r0 + ql1 < a.M;
    v16f acc[4][2];
    for (int64_t t = t0; t < t1; ++t) {
        const int64_t i0 = t * kBI;
        tile_product<ELEM>(acc, a, i0, r0, sb0, sb1, tid, wi, wr, lane);

        // --- filter: flag values strictly above tau (a later column can only
        // beat an equal list entry by value, never by the column tie-break)
        float tau0 = tauv[ql0], tau1 = tauv[ql1];
        uint64_t pend0 = 0, pend1 = 0;
#pragma unroll
        for (int ai = 0; ai < 4; ++ai)
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                const bool ok = i0 + acc_row(wi, ai, i, lane) < a.N;
                const uint64_t bit = 1ull << (ai * 16 + i);
                if (ok && qv0 && acc[ai][0][i] > tau0) pend0 |= bit;
                if (ok && qv1 && acc[ai][1][i] > tau1) pend1 |= bit;
            }
        while (__syncthreads_or((pend0 | pend1) != 0ull)) {
            // --- append flagged values to the per-query candidate slots
#pragma unroll
            for (int ai = 0; ai < 4; ++ai)
#pragma unroll
                for (int i = 0; i < 16; ++i) {
                    const uint64_t bit = 1ull << (ai * 16 + i);
                    const int32_t col = (int32_t)(a.col_base + i0 + acc_row(wi, ai, i, lane));
                    if (pend0 & bit) {
                        int slot = atomicAdd(&cnt[ql0], 1);
                        if (slot < kCap) {
                            cv[ql0 * CP + slot] = acc[ai][0][i];
                            cc[ql0 * CP + slot] = col;
                            pend0 &= ~bit;
                        }
                    }
                    if (pend1 & bit) {
                        int slot = atomicAdd(&cnt[ql1], 1);
                        if (slot < kCap) {
                            cv[ql1 * CP + slot] = acc[ai][1][i];
                            cc[ql1 * CP + slot] = col;
                            pend1 &= ~bit;
                        }
                    }
                }
            __syncthreads();
            // --- merge: one thread per query inserts its candidates
            if (tid < kBR) {
                const int nc = min(cnt[tid], kCap);
                float *L = lv + tid * LP;
                int32_t *Lc = lc + tid * LP;
                for (int j = 0; j < nc; ++j) {
                    const float v = cv[tid * CP + j];
                    const int32_t c = cc[tid * CP + j];
                    if (!beats(v, c, L[KP - 1], Lc[KP - 1])) continue;
                    int pos = KP - 1;
                    while (pos > 0 && beats(v, c, L[pos - 1], Lc[pos - 1])) {
                        L[pos] = L[pos - 1];
                        Lc[pos] = Lc[pos - 1];
                        --pos;
                    }
                    L[pos] = v;
                    Lc[pos] = c;
                }
                cnt[tid] = 0;
                tauv[tid] = L[KP - 1];
            }
            __syncthreads();
            // --- overflowed values retry against the raised tau (>=: an equal
            // value of this tile may still win on its lower column)
            tau0 = tauv[ql0];
            tau1 = tauv[ql1];
            if (pend0 | pend1) {
#pragma unroll
                for (int ai = 0; ai < 4; ++ai)
#pragma unroll
                    for (int i = 0; i < 16; ++i) {
                        const uint64_t bit = 1ull << (ai * 16 + i);
                        if ((pend0 & bit) && !(acc[ai][0][i] >= tau0)) pend0 &= ~bit;
                        if ((pend1 & bit) && !(acc[ai][1][i] >= tau1)) pend1 &= ~bit;
                    }
            }
        }
    }
    __syncthreads();
    if (tid < kBR && r0 + tid < a.M) {
        const int64_t base = ((r0 + tid) * a.S + split) * (int64_t)KP;
        for (int j = 0; j < KP; ++j) {
            a.part_v[base + j] = lv[tid * LP + j];
            a.part_c[base + j] = lc[tid * LP + j];
        }
    }
}

static int launch_common(const GemmArgs &a, const void *fn, size_t lds, hipStream_t st) {
    int64_t items = a.R * a.S;
    if (items == 0) return RM_OK;
    int64_t G = cdiv(items, 8) * 8;
    RM_CHECK(G < (int64_t)1 << 31, RM_ERR_SHAPE, "grid too large (%lld workgroups)", (long long)G);
    RM_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    void *args[] = {(void *)&a};
    RM_HIP(hipLaunchKernel(fn, dim3((unsigned)G), dim3(kThreads), args, lds, st));
    return RM_OK;
}

int launch_gemm_topk(const GemmArgs &a, int elem, hipStream_t st) {
    const void *fn = elem == 0 ? (const void *)gemm_topk_kernel<0> : (const void *)gemm_topk_kernel<1>;
    return launch_common(a, fn, topk_lds_bytes(a.KP), st);
}
int launch_gemm_store(const GemmArgs &a, int elem, hipStream_t st) {
    const void *fn = elem == 0 ? (const void *)gemm_store_kernel<0> : (const void *)gemm_store_kernel<1>;
    return launch_common(a, fn, store_lds_bytes(), st);
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
__global__ __launch_bounds__(64) void merge_kernel(const float *__restrict__ pv,
                                                   const int32_t *__restrict__ pc, int S, int KP,
                                                   int64_t M, int ktop, int mode,
                                                   const float *__restrict__ q32, int64_t ldq32,
                                                   const uint8_t *__restrict__ db, int64_t ldb, int K,
                                                   float *__restrict__ vals,
                                                   int32_t *__restrict__ locs, int64_t ldo) {
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
        if (lane == 0) { mv[j] = wv; mc[j] = wc; }
    }
    for (int t = j + lane; t < KP; t += 64) { mv[t] = -INFINITY; mc[t] = INT_MAX; }
    __syncthreads();

    if (mode == RM_MODE_FP8_RESCORE) {
        const float *qr = q32 + r * ldq32;
        for (int t = lane; t < K; t += 64) qs[t] = qr[t];
        __syncthreads();
        const int C = (K + 63) / 64;
        for (int t = 0; t < KP; ++t) {
            const int32_t c = mc[t];
            if (c == INT_MAX) { if (lane == 0) sc[t] = -INFINITY; continue; }
            const uint8_t *col = db + (int64_t)c * ldb;
            float acc = 0.0f;
            for (int e = lane * C; e < (lane + 1) * C && e < K; ++e) {
                const float p = qs[e] * e4m3_to_f32(col[e]);
                acc = acc + p;
            }
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) acc = acc + __shfl_xor(acc, off, 64);
            if (lane == 0) sc[t] = acc;
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

int launch_merge(const float *part_v, const int32_t *part_c, int S, int KP, int64_t M, int ktop,
                 int mode, const float *q32, int64_t ldq32, const uint8_t *db, int64_t ldb, int K,
                 float *vals, int32_t *locs, int64_t ldo, const int64_t *, hipStream_t st) {
    if (M == 0) return RM_OK;
    RM_CHECK(S <= 256, RM_ERR_ARG, "at most 256 column splits (got %d)", S);
    size_t lds = (mode == RM_MODE_FP8_RESCORE) ? (size_t)K * 4 : 0;
    merge_kernel<<<dim3((unsigned)M), dim3(64), lds, st>>>(part_v, part_c, S, KP, M, ktop, mode,
                                                           q32, ldq32, db, ldb, K, vals, locs, ldo);
    RM_HIP(hipGetLastError());
    return RM_OK;
}

// ---------------------------------------------------------------------------
// Drop-in running top-k over a C chunk (column-major, one thread per row):
// the running list (earlier columns) merged with the chunk under the total
// order.  Lists live in LDS, 64 rows per block.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void rowtopk_kernel(float *dv, int32_t *dl, int64_t ldd,
                                                     const void *C, int cdt, int64_t ldc, int64_t M,
                                                     int64_t wi, int64_t offset, int k) {
    extern __shared__ float lsh[];
    const int t = threadIdx.x;
    const int64_t row = (int64_t)blockIdx.x * 64 + t;
    if (row >= M) return;
    float *L = lsh + t * (2 * k);
    int32_t *Lc = (int32_t *)(L + k);
    for (int j = 0; j < k; ++j) {
        const int32_t l = dl[row + j * ldd];
        L[j] = dv[row + j * ldd];
        Lc[j] = l > 0 ? l - 1 : INT_MAX;
    }
    for (int64_t j = 0; j < wi; ++j) {
        const int64_t o = row + j * ldc;
        float v;
        if (cdt == RM_F32) v = ((const float *)C)[o];
        else if (cdt == RM_F16) v = f16_bits_to_f32(((const uint16_t *)C)[o]);
        else v = e4m3_to_f32(((const uint8_t *)C)[o]);
        const int32_t c = (int32_t)(offset + j);
        if (!beats(v, c, L[k - 1], Lc[k - 1])) continue;
        int pos = k - 1;
        while (pos > 0 && beats(v, c, L[pos - 1], Lc[pos - 1])) {
            L[pos] = L[pos - 1];
            Lc[pos] = Lc[pos - 1];
            --pos;
        }
        L[pos] = v;
        Lc[pos] = c;
    }
    for (int j = 0; j < k; ++j) {
        dv[row + j * ldd] = L[j];
        dl[row + j * ldd] = Lc[j] == INT_MAX ? 0 : Lc[j] + 1;
    }
}

int launch_rowtopk(float *dv, int32_t *dl, int64_t ldd, const void *C, int c_dtype, int64_t ldc,
                   int64_t M, int64_t wi, int64_t offset, int k, hipStream_t st) {
    if (M == 0) return RM_OK;
    size_t lds = (size_t)64 * 2 * k * 4;
    rowtopk_kernel<<<dim3((unsigned)cdiv(M, 64)), dim3(64), lds, st>>>(dv, dl, ldd, C, c_dtype, ldc,
                                                                       M, wi, offset, k);
    RM_HIP(hipGetLastError());
    return RM_OK;
}

// dst[r*cols + c] = src[r*s_row + c*s_col] for esize-byte elements.
__global__ void transpose_kernel(uint8_t *dst, const uint8_t *src, int64_t rows, int64_t cols,
                                 int64_t s_row, int64_t s_col, int esize) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= rows * cols) return;
    const int64_t r = g / cols, c = g % cols;
    const uint8_t *s = src + (r * s_row + c * s_col) * esize;
    uint8_t *d = dst + g * esize;
    for (int b = 0; b < esize; ++b) d[b] = s[b];
}

int launch_transpose_u8(uint8_t *dst, const uint8_t *src, int64_t rows, int64_t cols, int64_t s_row,
                        int64_t s_col, int esize, hipStream_t st) {
    const int64_t n = rows * cols;
    if (n == 0) return RM_OK;
    transpose_kernel<<<dim3((unsigned)cdiv(n, 256)), dim3(256), 0, st>>>(dst, src, rows, cols, s_row,
                                                                         s_col, esize);
    RM_HIP(hipGetLastError());
    return RM_OK;
}

}  // namespace rm
