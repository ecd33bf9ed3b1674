/*
 * rotormap_hip.h -- C ABI of librotormap_hip.so, the MI355X (gfx950) mapping
 * hot path of RotorMap: RoPE s-mer-histogram encoder -> reads x index GEMM
 * (FP8 e4m3 or FP16 on MFMA, FP32 accumulate) -> fused streaming per-row
 * top-k -> FP32 rescoring.
 *
 * Conventions (all entry points):
 *   - plain pointers and sizes; no C++ or framework types cross the boundary;
 *   - every function returns an int status (RM_OK = 0); rm_last_error() gives
 *     the calling thread's last message;
 *   - pointers marked (device) are HIP device allocations owned by the caller
 *     and never freed by the library; (host) pointers are ordinary memory;
 *   - `stream` is a hipStream_t passed as void* (NULL = the default stream);
 *     calls are stream ordered: they return after enqueueing unless stated;
 *   - index column ids are 1-based int32 in every output, 0 = empty slot,
 *     value -inf for an empty slot (RotorMap/gemm/topk_kernels.jl:56-60);
 *   - ties in every top-k are broken by (score desc, column asc).
 */
#ifndef ROTORMAP_HIP_H
#define ROTORMAP_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------- */
#define RM_OK 0
#define RM_ERR_ARG 1   /* null pointer / parameter out of range            */
#define RM_ERR_SHAPE 2 /* inconsistent shapes or alignment                 */
#define RM_ERR_NOMEM 3 /* device allocation failed / admission refused     */
#define RM_ERR_HIP 4   /* HIP runtime error                                */
#define RM_ERR_STATE 5 /* object in the wrong state (stopped, cancelled)   */

/* ---- element types ------------------------------------------------------ */
#define RM_F32 0
#define RM_F16 1
#define RM_E4M3 2 /* OCP FP8 E4M3FN, round to nearest even, satfinite */

/* ---- scoring modes ------------------------------------------------------ */
#define RM_MODE_FP8 0         /* e4m3 GEMM, fused top-k, no rescoring          */
#define RM_MODE_FP8_RESCORE 1 /* e4m3 GEMM, fused top-k', FP32 rescore -> top-k */
#define RM_MODE_FP16 2        /* fp16 GEMM (fp32 accumulate), fused top-k       */

typedef struct rm_ctx rm_ctx;
typedef struct rm_index rm_index;
typedef struct rm_server rm_server;

/* ---- library / context --------------------------------------------------- */
const char *rm_version(void);
const char *rm_last_error(void);
/* Context on HIP device `device`: owns a stream, a workspace, the phasor
 * tables and the tile-plan cache.  Replaces the process-global cuBLASLt
 * handle/workspace of RotorMap/gemm/fp8_lt.jl:47-93 (cublaslt_init). */
int rm_ctx_create(int device, rm_ctx **out);
int rm_ctx_destroy(rm_ctx *ctx);
int rm_sync(rm_ctx *ctx, void *stream);
/* Device memory the context holds right now (workspace + tables), bytes. */
int64_t rm_ctx_bytes(rm_ctx *ctx);

/* ---- encoder ---------------------------------------------------------------
 * Replaces encode_frag_real_batch! (RotorMap/encode/kernel.jl:298-325) and its
 * kernel kernel_rope_frag_real (:86-290).  words (device): n fragments x
 * W = ceil(k/16) uint32, fragment contiguous, base j in word j>>4 at bit
 * 2*(j&15).  Output element (fragment f, dim d) of out1/out2 is written at
 * f*stride_frag + d*stride_dim (elements): the reference layout (n, 2D)
 * column-major is stride_frag = 1, stride_dim = n; the K-contiguous layout
 * the GEMM reads is stride_frag = 2D, stride_dim = 1.  out2 may be NULL.
 * norms (device, may be NULL): (m, n) column-major fp32, the layout of
 * embeds_norms.  counts (device, may be NULL): (n, 4^c) int32 row-major, the
 * integer s-mer key histogram (windows per compact key, before G zeroing).
 * Requires 1 <= c <= s <= 16, s <= k, 2*m*4^c <= 16384, 0 <= normalize <= 3. */
int rm_encode_batch(rm_ctx *ctx, const uint32_t *words, int64_t n, int k, int s, int m, int c,
                    int normalize, void *out1, int dtype1, int64_t stride_frag1,
                    int64_t stride_dim1, void *out2, int dtype2, int64_t stride_frag2,
                    int64_t stride_dim2, float *norms, int32_t *counts, void *stream);
/* The phasor table the encoder uses (host copy): 2*m*(k-s+1) int32. */
int rm_phase_table(int k, int s, int m, int32_t *host_out);

/* Reverse complement of n fragments of k bases (device -> device).
 * Replaces revcomp_frag_words! (RotorMap/index/build.jl:448-476). */
int rm_revcomp_words(rm_ctx *ctx, uint32_t *rc, const uint32_t *words, int64_t n, int k,
                     void *stream);

/* Slice n k-base windows out of a packed record (device): window i starts at
 * 0-based base starts[i] (device int64); rec_words must hold ceil(len/16)+1
 * words (zero guard word).  Replaces _window_words! (RotorMap/index/build.jl:
 * 256-270) on the device. */
int rm_extract_windows(rm_ctx *ctx, uint32_t *out, const uint32_t *rec_words,
                       const int64_t *starts, int64_t n, int k, void *stream);

/* ---- conversions -------------------------------------------------------- */
/* dst[i] = e4m3_satfinite(src[i]) for n elements, src fp32 or fp16 (device).
 * Replaces f32_to_f8! (RotorMap/gemm/fp8_convert.jl:100-103) and the query
 * quantisation of upload_fp16_as_f8! (RotorMap/search/engine_fp8.jl:355-371). */
int rm_quantize_e4m3(rm_ctx *ctx, uint8_t *dst, const void *src, int src_dtype, int64_t n,
                     void *stream);
/* dst[i] = float(e4m3 src[i]) (device); f8_to_f32! (fp8_convert.jl:105-108). */
int rm_dequantize_e4m3(rm_ctx *ctx, float *dst, const uint8_t *src, int64_t n, void *stream);

/* ---- GEMM chunk (the :hip engine's plug point) ---------------------------
 * C (M x wi) = A (M x K) * B (K x wi), fp32 accumulate, C column-major with
 * leading dimension ldc, c_dtype RM_F32 or RM_F16.  A element (r, k) is at
 * A[r*a_stride_row + k*a_stride_k] (the reference's a8 is a_stride_row = 1,
 * a_stride_k = rows_cap); B column j is K contiguous elements at B + j*ldb.
 * ab_dtype RM_E4M3 or RM_F16 (both operands).  K*elemsize % 64 == 0.
 * Replaces fp8_gemm_lt! (RotorMap/gemm/fp8_lt.jl:128-168), reached from
 * _gemm_chunk! (RotorMap/search/engine_fp8.jl:379-387). */
int rm_gemm_chunk(rm_ctx *ctx, void *C, int c_dtype, int64_t ldc, const void *A,
                  int64_t a_stride_row, int64_t a_stride_k, const void *B, int64_t ldb,
                  int ab_dtype, int64_t M, int64_t wi, int64_t K, void *stream);

/* Running per-row top-k merge of a C chunk (M x wi, column-major, ldc;
 * c_dtype RM_F32 / RM_F16 / RM_E4M3 bytes) into D_val / D_loc (rows_cap x k,
 * column-major, ldd), fresh columns offset by `offset` (1-based ids
 * offset+1 .. offset+wi).  Replaces launch_seg_rowtopk_merge!
 * (RotorMap/gemm/topk_kernels.jl:331-347). */
int rm_rowtopk_merge(rm_ctx *ctx, float *D_val, int32_t *D_loc, int64_t ldd, const void *C,
                     int c_dtype, int64_t ldc, int64_t M, int64_t wi, int64_t offset, int k,
                     void *stream);

/* ---- resident index -------------------------------------------------------
 * B: rdim x N e4m3, column-major (column j = rdim contiguous bytes), the
 * `embeds8` matrix of an indexreal.fp8.v1 .bin (RotorMap/index/load.jl:49-73,
 * index_database :156-161).  flags: RM_INDEX_HOST (B is a host pointer, the
 * library uploads it once) and RM_INDEX_FP16 (also keep the exact fp16
 * widening for RM_MODE_FP16, RotorMap/index/load.jl:130-143). */
#define RM_INDEX_HOST 1
#define RM_INDEX_FP16 2
int rm_index_create(rm_ctx *ctx, const uint8_t *B, int64_t rdim, int64_t N, int flags,
                    rm_index **out);
/* The classic fp16 index ("indexreal.v1"): B16 rdim x N binary16,
 * column-major like B, kept verbatim as the RM_MODE_FP16 database -- the
 * reference-accuracy pipeline's index_to_f16 (RotorMap/index/load.jl:96-116,
 * RotorMap/legacy/test/e2ehuman16.jl:4-16) -- with the e4m3 database of the
 * same columns, e4m3(fp32(fp16)), quantised on the device for the FP8 modes
 * (index_to_f8, RotorMap/index/load.jl:76-89).  flags: RM_INDEX_HOST. */
int rm_index_create_f16(rm_ctx *ctx, const uint16_t *B16, int64_t rdim, int64_t N, int flags,
                        rm_index **out);
int rm_index_destroy(rm_index *idx);
int64_t rm_index_cols(const rm_index *idx);
int64_t rm_index_rdim(const rm_index *idx);
/* Device pointers of the resident copies (fp16 NULL unless RM_INDEX_FP16). */
const uint8_t *rm_index_e4m3(const rm_index *idx);
const void *rm_index_f16(const rm_index *idx);
/* Columns the FP32 rescoring of RM_MODE_FP8_RESCORE reads: RM_E4M3 (the
 * default: the e4m3 database the candidates were ranked on) or RM_F16 (the
 * index's fp16 columns; for an index from rm_index_create_f16 these are the
 * verbatim fp16 columns, so the final scores and ranking carry the fp16
 * index's precision while the k' candidates still come from the FP8 GEMM).
 * RM_ERR_STATE if the index holds no fp16 columns.  Same fixed FP32 order
 * as the e4m3 rescoring.  No counterpart in the reference, which has no
 * rescoring stage.  rm_index_rescore returns the current setting. */
int rm_index_set_rescore(rm_index *idx, int dtype);
int rm_index_rescore(const rm_index *idx);

/* ---- fused GEMM + top-k over the resident index ---------------------------
 * Replaces batch_gemm_topk! (RotorMap/search/engine_fp8.jl:399-437; fp16
 * twin RotorMap/search/engine_fp16.jl:255-292): one launch of the fused
 * MFMA GEMM + per-row top-k over all N columns (scores never reach HBM),
 * then the split merge and (RM_MODE_FP8_RESCORE) the FP32 rescoring.
 * Q (device): M x rdim, row stride ldq elements, RM_E4M3 for the FP8 modes,
 * RM_F16 for RM_MODE_FP16.  Q32 (device, RM_MODE_FP8_RESCORE only): the
 * pre-quantisation fp32 queries, row stride ldq32.  kprime: candidates kept
 * per read before rescoring (ktop <= kprime <= 64; ignored otherwise).
 * vals / locs (device): (M x ktop) column-major with leading dimension ldo. */
int rm_topk_batch(rm_ctx *ctx, const rm_index *idx, const void *Q, int64_t ldq,
                  const float *Q32, int64_t ldq32, int64_t M, int ktop, int kprime, int mode,
                  float *vals, int32_t *locs, int64_t ldo, void *stream);

/* Fused encode -> GEMM -> top-k (-> rescore) of n packed fragments (device
 * words, as rm_encode_batch) against the resident index; norms as
 * rm_encode_batch (may be NULL).  The end-to-end hot loop of topk_flow /
 * rope_topk_stream (RotorMap/search/engine_fp8.jl:498-594). */
int rm_map_batch(rm_ctx *ctx, const rm_index *idx, const uint32_t *words, int64_t n, int k,
                 int s, int m, int c, int ktop, int kprime, int mode, float *vals,
                 int32_t *locs, int64_t ldo, float *norms, void *stream);

/* rm_map_batch without any allocation.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments and results as rm_map_batch (bit for bit); its work buffers --
 * the phasor table, the encoded queries and the split partials -- live in a
 * static device array of RM_NOALLOC_SCRATCH_BYTES per GPU, held until the
 * process ends, and it calls no hipMalloc / hipFree.  A batch whose buffers
 * do not fit returns RM_ERR_SHAPE (at rdim 2048 in RM_MODE_FP8_RESCORE about
 * 6,300 rows), as does a phasor table past 2^18 entries (2*m*(k-s+1)); the
 * split count is lowered to what fits (results do not depend on it).  Every
 * call on one GPU shares that GPU's array: calls that overlap on a GPU (two
 * streams or threads) race.  The first call with a new (k, s, m) stages its
 * table and synchronizes `stream`. */
#define RM_NOALLOC_SCRATCH_BYTES (64u << 20)
int rm_map_batch_noalloc(rm_ctx *ctx, const rm_index *idx, const uint32_t *words, int64_t n,
                         int k, int s, int m, int c, int ktop, int kprime, int mode, float *vals,
                         int32_t *locs, int64_t ldo, float *norms, void *stream);

/* Uneven batches in one call: batch b holds rows [offs[b], offs[b+1]) of
 * the concatenated words (host offs, nb+1 entries); results land at the same
 * rows of vals/locs.  Rows are independent, so the batches are encoded
 * together and the persistent kernel walks one work list of (row tile,
 * column split) items across all of them, with the tile shape and split
 * count the context's tile plan gives the total row count (rm_autotune);
 * a large total runs as the launches rm_launch_rows gives it.  Each row's
 * result equals that of rm_map_batch on its batch alone. */
int rm_map_batches(rm_ctx *ctx, const rm_index *idx, const uint32_t *words,
                   const int64_t *offs, int nb, int k, int s, int m, int c, int ktop, int kprime,
                   int mode, float *vals, int32_t *locs, int64_t ldo, void *stream);

/* ---- tile plan / autotuner ------------------------------------------------
 * A plan is (query rows per tile 64 / 128 / 256, column splits S, query
 * tiles grouped per XCD qg) per (mode, batch-size bucket ceil(log2 M)); the
 * reference's fixed chunk width w and rows_cap (RotorMap/search/
 * engine_fp8.jl:309-337) are what it replaces.  Without a tuned entry a
 * batch takes 64 rows up to 64, 128 up to 128, else 256, and the
 * built-in split rule, stepping the tile down while a split would sweep
 * fewer than 32 column tiles of the index; lists longer than 32 (kp > 32)
 * always run 128-row tiles.  With splits and at least CUs / 8 query tiles
 * the query tiles are grouped CUs / 8 / S per XCD.  Results never depend on
 * the plan.
 * rm_plan_get: the plan one launch of M rows gets; kp is the list length the
 * fused kernel keeps (kprime for RM_MODE_FP8_RESCORE, else ktop). */
int rm_plan_get(rm_ctx *ctx, const rm_index *idx, int64_t M, int mode, int kp, int32_t *tile_rows,
                int32_t *splits, int32_t *qgroup);
/* Rows per fused launch that rm_topk_batch, rm_map_batch and rm_map_batches
 * (for the total row count) use for a batch of M rows (0: one launch).  Without a tuned plan for M's bucket, a batch
 * above 64 x CUs rows on a large index (each of the launches' splits
 * sweeping at least 1,024 column tiles) runs as equal launches of at most
 * that many rows, each with rm_plan_get's plan for that many rows: one wave
 * of workgroups with its query tiles resident in the XCDs' L2.  No
 * counterpart in the
 * reference, whose engine cuts batches at its own rows_cap
 * (RotorMap/search/engine_fp8.jl:309-337). */
int rm_launch_rows(rm_ctx *ctx, const rm_index *idx, int64_t M, int mode, int kp, int64_t *rows);
/* Pin the plan for batches in M's bucket: tile_rows 0 (the default tile),
 * 64, 128 or 256; splits 0 (the default rule) .. 256; qgroup 0 (all) or the
 * query tiles grouped per XCD.  tile_rows = splits = 0 clears the entry. */
int rm_plan_set(rm_ctx *ctx, int64_t M, int mode, int tile_rows, int splits, int qgroup);
/* Autotuner: times the candidate plans of each bucket on `idx` (tile rows
 * up to the first that holds the bucket, split counts around one wave of
 * workgroups, XCD groupings; tuning queries: the index's own columns, at
 * this ktop / kprime) and keeps the fastest in the context.  Host arrays of
 * nbuckets entries: bucket row counts in; chosen tile rows, S, qg and the
 * mean ms per launch out (any may be NULL). */
int rm_autotune(rm_ctx *ctx, const rm_index *idx, const int64_t *bucket_rows, int nbuckets,
                int mode, int ktop, int kprime, int32_t *best_tile, int32_t *best_splits,
                int32_t *best_qgroup, float *best_ms);

/* Launch shape of the fused kernels (no counterpart in the reference): on
 * (the default), a grid of one workgroup per CU walks the (row tile, split)
 * items of the launch -- one persistent launch however uneven the rows of
 * the batches concatenated into it; off, one workgroup per item.  Results
 * are identical either way. */
int rm_set_persistent(rm_ctx *ctx, int on);
/* ---- kernel timing ---------------------------------------------------------
 * When enabled, the context records a pair of HIP events around every launch
 * of the fused GEMM + top-k kernel, on the stream it is launched on.
 * rm_prof_read waits for them and returns the summed kernel time (ms), the
 * launch count and the summed algorithmic FLOP (2*M*K*N per launch), then
 * clears the record.  No counterpart in the reference (it times with host
 * clocks, RotorMap/common/util.jl:20-78). */
int rm_prof_enable(rm_ctx *ctx, int on);
int rm_prof_read(rm_ctx *ctx, double *kernel_ms, int64_t *launches, double *flop);

/* ---- admission model -------------------------------------------------------
 * Device bytes a batch of M rows needs beyond the resident index (queries,
 * fp32 rescoring copies, split partials, outputs); the resident-index rule
 * of RotorMap/search/engine_compact.jl:195-242 restated for the fused path.
 * It takes no plan, so the split partials are counted for the largest split
 * count (256): a worst-case upper bound, far above what a large batch uses
 * (a 65,536-row batch runs one split).  rm_server_create admits by the
 * actual plans instead. */
int64_t rm_batch_bytes(int64_t M, int k, int rdim, int ktop, int kprime, int mode);
/* Free device memory reported by the HIP runtime, bytes. */
int64_t rm_device_free_bytes(rm_ctx *ctx);

/* ---- serving loop (continuous batching on one GPU) ------------------------
 * A server owns worker streams over a resident index.  Requests (host packed
 * words) are admitted when their rows fit the free row slots and the
 * free-HBM byte budget; admitted requests are packed into the next launch
 * (requests that arrive while launches are in flight join later launches,
 * so work from different requests overlaps on the device).  No counterpart
 * in the reference, whose "server" is a file watcher
 * (RotorMap/legacy/src/RotorMap.jl:233). */
int rm_server_create(rm_ctx *ctx, const rm_index *idx, int k, int s, int m, int c, int ktop,
                     int kprime, int mode, int64_t max_rows_per_launch, int n_streams,
                     int64_t hbm_budget_bytes, rm_server **out);
/* Enqueue a request of n fragments (host words, copied).  Returns a request
 * id >= 0 through *req_id. */
int rm_server_submit(rm_server *srv, const uint32_t *host_words, int64_t n, int64_t *req_id);
/* Enqueue without copying: the caller keeps host_words valid and unchanged
 * until rm_server_wait has returned for this id (pinned memory gives the
 * fastest uploads). */
int rm_server_submit_ref(rm_server *srv, const uint32_t *host_words, int64_t n, int64_t *req_id);
/* Poll / wait for a finished request: copies (n x ktop) row-major vals/locs
 * to host buffers.  timeout_ms < 0 waits forever; returns RM_ERR_STATE if
 * not finished within the timeout.  t_* (may be NULL): submit, admit and
 * finish times in seconds on the server's monotonic clock.  A failed request
 * returns only once none of its rows is in flight (its submit_ref words are
 * then free to reuse).  A request id has one waiter: the call that returns
 * releases it. */
int rm_server_wait(rm_server *srv, int64_t req_id, int timeout_ms, float *host_vals,
                   int32_t *host_locs, double *t_submit, double *t_admit, double *t_finish);
/* Cancel a request that has not been admitted yet. */
int rm_server_cancel(rm_server *srv, int64_t req_id);
int rm_server_stats(rm_server *srv, int64_t *launches, int64_t *rows_done, int64_t *max_inflight);
/* Order in which the dispatcher fills a launch with queued rows.
 * RM_SCHED_FIFO: arrival order (a request's rows before any later request's).
 * RM_SCHED_SRPT: the oldest queued request takes up to a quarter of the
 * launch (it keeps progressing however many small requests arrive), then
 * fewest remaining rows first, ties in arrival order, so a small request that
 * arrives behind a large one rides the next launch.  Every
 * launch is still filled to the row capacity; results do not depend on the
 * order.  Takes effect from the next launch.  No counterpart in the reference
 * (it has no serving loop). */
#define RM_SCHED_FIFO 0
#define RM_SCHED_SRPT 1
int rm_server_set_policy(rm_server *srv, int policy);
/* Low-load mode (0 = off, the default): while at most backlog_rows rows wait
 * unlaunched, (1) one launch is in flight at a time, (2) a launch holds whole
 * requests only, in the policy's order, while they fit in launch_rows rows,
 * and (3) a first request larger than that gets a launch_rows slice.  The
 * fused kernel holds every CU while it runs, so a request that arrives
 * during a large one otherwise waits for the running launch, the launch
 * queued behind it and its own launch padded with the large request's rows.
 * Once a backlog builds the server returns to full, padded launches with
 * every slot in flight, where throughput is highest.  Results do not depend
 * on it.  Takes effect from the next launch.  No counterpart in the
 * reference. */
int rm_server_set_lowload(rm_server *srv, int64_t launch_rows, int64_t backlog_rows);
/* Seconds on the server's monotonic clock (the clock of the t_* stamps). */
double rm_server_now(rm_server *srv);
/* Stop the loop: requests not yet launched end as cancelled (or failed when
 * some of their rows were launched); launches in flight finish first. */
int rm_server_destroy(rm_server *srv);

/* ---- synthetic inputs (host, deterministic in their seeds) ----------------
 * Record r's bases drawn i.i.d. with P(G)=P(C)=gc/2, P(A)=P(T)=(1-gc)/2 from
 * a counter-based hash of (seed, r, position), packed as forward words with
 * a zero guard word: ceil(len/16)+1 words. */
int rm_synth_record(uint64_t seed, int64_t rec, int64_t len, double gc, uint32_t *words);
/* Set bases [lo, hi) of a packed record to code 2: N runs as the reader maps
 * them (RotorMap/fasta/reader.jl:21-28, N -> G). */
int rm_mark_n(uint32_t *words, int64_t lo, int64_t hi);
/* FASTA bytes (device, n_bytes) -> packed reads on the GPU, the parse of
 * fasta_reads (RotorMap/fasta/reader.jl:321-350 with _LUT3 :21-28, record
 * split _next_rec3 :61-76, header / trim / skip _part_task3 :204-300, packing
 * _emit3! :163-197): records start at a '>' at a line start; bytes before
 * the first record are ignored; the header is '>' .. the byte before the
 * first '\n' (one trailing '\r' dropped); every other byte but '\n' / '\r'
 * is a base (A/a 0, C/c 1, G/g 2, T/t 3, anything else 2); records with >= k
 * bases are trimmed to k and kept, in file order, shorter ones skipped.
 * words (device): max_reads x ceil(k/16) uint32 rows; hdr_off / hdr_len
 * (device, may be NULL): each kept read's header as a byte range of raw.
 * *n_reads (host) = reads kept; RM_ERR_SHAPE (with *n_reads set) when more
 * than max_reads are kept.  Synchronises `stream`.  raw may be split between
 * calls only at a line-start '>'. */
int rm_fasta_reads(rm_ctx *ctx, const uint8_t *raw, int64_t n_bytes, int k, uint32_t *words,
                   int64_t max_reads, int64_t *hdr_off, int32_t *hdr_len, int64_t *n_reads,
                   void *stream);
/* rm_fasta_reads on host buffers (raw, words, hdr_off, hdr_len all host; the
 * library stages them through its workspace): for callers without device
 * arrays, such as a Julia process without a GPU array package. */
int rm_fasta_reads_host(rm_ctx *ctx, const uint8_t *raw, int64_t n_bytes, int k,
                        uint32_t *words, int64_t max_reads, int64_t *hdr_off, int32_t *hdr_len,
                        int64_t *n_reads);
/* Host twin of rm_extract_windows (host pointers). */
int rm_extract_windows_host(uint32_t *out, const uint32_t *rec_words, const int64_t *starts,
                            int64_t n, int k);
/* Mutate n packed k-base reads in place.  model 0: substitutions only,
 * ceil(k*err) draws with replacement, new base (b + 1 + r) mod 4, r in 0..2;
 * model 1: the reference's mutate (RotorMap/common/dna.jl:120-183):
 * n_err = ceil(k*err), n_ins = n_del = floor(n_err/3), n_sub the rest,
 * length kept.  err[i] per read; read i draws from hash(seed, read_ids[i]). */
int rm_mutate_words(uint32_t *words, int64_t n, int k, const double *err, uint64_t seed,
                    const int64_t *read_ids, int model);

#ifdef __cplusplus
}
#endif
#endif /* ROTORMAP_HIP_H */
