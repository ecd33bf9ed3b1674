// fasta.hip -- FASTA bytes -> packed 2-bit reads on the GPU: the record
// split, header capture, byte -> code mapping, line-terminator removal,
// trim-to-k and forward packing of fasta_reads (RotorMap/fasta/reader.jl:
// 21-28 _LUT3, 61-76 _next_rec3, 204-300 _part_task3, 163-197 _emit3!).
//
// Rules (the reference's, restated): a record starts at a '>' at a line start
// (byte 0 or preceded by '\n' / '\r'); bytes before the first record are
// ignored; the header is the bytes from '>' up to the first '\n' with one
// trailing '\r' dropped (a record whose header line has no '\n' yields
// nothing); every other byte of the record except '\n' and '\r' is a base:
// A/a 0, C/c 1, G/g 2, T/t 3, anything else 2; a record with >= k bases is
// trimmed to its first k and packed (base j in word j >> 4 at bit 2 (j & 15),
// zero tail), shorter records are skipped.  Kept reads come out in file
// order.
//
// Passes (all integer, HBM-bound): (1) per 4 KiB block, count record starts;
// (2) exclusive scan of the block counts (one workgroup); (3) scatter the
// record start offsets in order; (4) one workgroup per record: find the
// header end, then walk the sequence in 4 KiB chunks -- each thread counts
// the bases of its 16 bytes, a block scan gives every byte its base index,
// and codes are OR-ed into the read's words in LDS until k bases; (5) scan of
// the kept flags and (6) compaction of kept reads and their headers.
#include "rm_device.h"
#include "rm_internal.h"

namespace rm {

constexpr int kFaThreads = 256;
constexpr int kFaBytes = 16;                        // bytes per thread per step
constexpr int kFaBlock = kFaThreads * kFaBytes;     // 4 KiB per workgroup step

__device__ __forceinline__ bool is_eol(uint8_t b) { return b == '\n' || b == '\r'; }

__device__ __forceinline__ uint32_t base_code(uint8_t b) {
    switch (b) {
    case 'A': case 'a': return 0;
    case 'C': case 'c': return 1;
    case 'T': case 't': return 3;
    default: return 2;  // G, g, N and every other byte
    }
}

__device__ __forceinline__ bool rec_start(const uint8_t *raw, int64_t i) {
    return raw[i] == '>' && (i == 0 || is_eol(raw[i - 1]));
}

// Block-wide exclusive scan of one int per thread (256 threads); returns the
// exclusive prefix and writes the block total to *total.
__device__ __forceinline__ int block_scan(int v, int *sh, int *total) {
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    int x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63) sh[w] = x;
    __syncthreads();
    int base = 0;
    for (int j = 0; j < w; ++j) base += sh[j];
    const int tot = sh[0] + sh[1] + sh[2] + sh[3];
    __syncthreads();
    *total = tot;
    return base + x - v;
}

// (1) record starts per 4 KiB block
__global__ __launch_bounds__(kFaThreads) void fa_count_kernel(const uint8_t *__restrict__ raw,
                                                              int64_t n, int *__restrict__ bcnt) {
    __shared__ int sh[4];
    const int64_t i0 = (int64_t)blockIdx.x * kFaBlock + threadIdx.x * kFaBytes;
    int c = 0;
    for (int j = 0; j < kFaBytes; ++j)
        if (i0 + j < n && rec_start(raw, i0 + j)) ++c;
    int tot;
    block_scan(c, sh, &tot);
    if (threadIdx.x == 0) bcnt[blockIdx.x] = tot;
}

// (2) / (5) exclusive scan of n ints in place by one workgroup; the total
// goes to *total
__global__ __launch_bounds__(kFaThreads) void fa_scan_kernel(int *__restrict__ a, int64_t n,
                                                             int64_t *__restrict__ total) {
    __shared__ int sh[4];
    int64_t carry = 0;
    for (int64_t b = 0; b < n; b += kFaThreads) {
        const int64_t i = b + threadIdx.x;
        const int v = i < n ? a[i] : 0;
        int tot;
        const int ex = block_scan(v, sh, &tot);
        if (i < n) a[i] = (int)(carry + ex);
        carry += tot;
    }
    if (threadIdx.x == 0) *total = carry;
}

// (3) record start offsets in file order
__global__ __launch_bounds__(kFaThreads) void fa_scatter_kernel(const uint8_t *__restrict__ raw,
                                                                int64_t n,
                                                                const int *__restrict__ boff,
                                                                int64_t *__restrict__ starts) {
    __shared__ int sh[4];
    const int64_t i0 = (int64_t)blockIdx.x * kFaBlock + threadIdx.x * kFaBytes;
    int c = 0;
    for (int j = 0; j < kFaBytes; ++j)
        if (i0 + j < n && rec_start(raw, i0 + j)) ++c;
    int tot;
    int o = boff[blockIdx.x] + block_scan(c, sh, &tot);
    for (int j = 0; j < kFaBytes; ++j)
        if (i0 + j < n && rec_start(raw, i0 + j)) starts[o++] = i0 + j;
}

// (4) one workgroup per record: header end, base count and packed words
__global__ __launch_bounds__(kFaThreads) void fa_pack_kernel(
    const uint8_t *__restrict__ raw, int64_t n, const int64_t *__restrict__ starts, int64_t nrec,
    int k, int W, uint32_t *__restrict__ stage, int *__restrict__ keep,
    int64_t *__restrict__ hoff, int32_t *__restrict__ hlen) {
    extern __shared__ uint32_t fw[];  // W words of the read
    __shared__ int sh[4];
    __shared__ int64_t hmin;
    const int tid = threadIdx.x;
    const int64_t r = blockIdx.x;
    const int64_t s = starts[r], e = r + 1 < nrec ? starts[r + 1] : n;
    for (int i = tid; i < W; i += kFaThreads) fw[i] = 0u;
    if (tid == 0) hmin = e;
    __syncthreads();
    // header: the first '\n' in [s, e)
    for (int64_t b = s; b < e; b += kFaThreads) {
        const int64_t i = b + tid;
        if (i < e && raw[i] == '\n') atomicMin((unsigned long long *)&hmin, (unsigned long long)i);
        __syncthreads();
        if (hmin < e) break;
        __syncthreads();
    }
    const int64_t h = hmin;
    int64_t bases = 0;
    if (h < e) {
        // aligned 16-byte loads from the 16-byte boundary at or below the
        // sequence start; bytes outside [h + 1, e) count as line ends.  An
        // aligned 16-byte block overlapping the buffer never leaves its page.
        const uint8_t *seq = raw + h + 1, *end = raw + e;
        const uint8_t *a0 = (const uint8_t *)((uintptr_t)seq & ~(uintptr_t)15);
        for (const uint8_t *b = a0; b < end && bases < k; b += kFaBlock) {
            const uint8_t *p = b + tid * kFaBytes;
            uint4 v = make_uint4(0u, 0u, 0u, 0u);
            if (p < end) v = *(const uint4 *)p;
            const uint32_t vw[4] = {v.x, v.y, v.z, v.w};
            uint32_t base_mask = 0u;  // bit j: byte j is a base
#pragma unroll
            for (int j = 0; j < kFaBytes; ++j) {
                const uint8_t by = (uint8_t)(vw[j >> 2] >> (8 * (j & 3)));
                if (p + j >= seq && p + j < end && !is_eol(by)) base_mask |= 1u << j;
            }
            int tot;
            int64_t pos = bases + block_scan(__popc(base_mask), sh, &tot);
            // this thread's bases land in at most two words: OR them locally
            uint32_t wv = 0u;
            int64_t wi = -1;
#pragma unroll
            for (int j = 0; j < kFaBytes; ++j) {
                if (!((base_mask >> j) & 1u)) continue;
                if (pos < k) {
                    const int64_t wj = pos >> 4;
                    if (wj != wi) {
                        if (wi >= 0) atomicOr(&fw[wi], wv);
                        wi = wj;
                        wv = 0u;
                    }
                    wv |= base_code((uint8_t)(vw[j >> 2] >> (8 * (j & 3)))) << (2 * (pos & 15));
                }
                ++pos;
            }
            if (wi >= 0) atomicOr(&fw[wi], wv);
            bases += tot;
        }
    }
    __syncthreads();
    const bool kept = bases >= k;
    uint32_t *out = stage + r * W;
    for (int i = tid; i < W; i += kFaThreads) out[i] = kept ? fw[i] : 0u;
    if (tid == 0) {
        keep[r] = kept ? 1 : 0;
        int64_t hend = h - 1;
        if (h < e && hend >= s && raw[hend] == '\r') --hend;
        hoff[r] = s;
        hlen[r] = (int32_t)(h < e ? hend - s + 1 : 0);
    }
}

// (6) kept reads (and headers) to their compacted rows
__global__ void fa_compact_kernel(const uint32_t *__restrict__ stage, const int *__restrict__ keep,
                                  const int *__restrict__ kidx, int64_t nrec, int W,
                                  const int64_t *__restrict__ hoff, const int32_t *__restrict__ hlen,
                                  uint32_t *__restrict__ words, int64_t *__restrict__ ooff,
                                  int32_t *__restrict__ olen) {
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < nrec * W;
         g += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = g / W;
        if (!keep[r]) continue;
        const int64_t o = kidx[r];
        const int w = (int)(g % W);
        words[o * W + w] = stage[g];
        if (w == 0) {
            if (ooff) ooff[o] = hoff[r];
            if (olen) olen[o] = hlen[r];
        }
    }
}

int fasta_reads(rm_ctx *ctx, const uint8_t *raw, int64_t n, int k, uint32_t *words,
                int64_t max_reads, int64_t *hdr_off, int32_t *hdr_len, int64_t *n_reads,
                hipStream_t st) {
    *n_reads = 0;
    if (n == 0) return RM_OK;
    const int W = (k + 15) / 16;
    const int64_t nb = cdiv(n, kFaBlock);
    RM_CHECK(nb * kFaThreads <= kMaxWorkItems && nb < (1LL << 31), RM_ERR_SHAPE,
             "%lld bytes exceed one parse", (long long)n);
    void *p;
    int rc;
    if ((rc = ws_get(ctx, "fa_bcnt", (size_t)nb * 4 + 64, &p))) return rc;
    int *bcnt = (int *)p;
    if ((rc = ws_get(ctx, "fa_tot", 64, &p))) return rc;
    int64_t *dtot = (int64_t *)p;
    fa_count_kernel<<<dim3((unsigned)nb), dim3(kFaThreads), 0, st>>>(raw, n, bcnt);
    fa_scan_kernel<<<1, kFaThreads, 0, st>>>(bcnt, nb, dtot);
    RM_HIP(hipGetLastError());
    int64_t nrec = 0;
    RM_HIP(hipMemcpyAsync(&nrec, dtot, 8, hipMemcpyDeviceToHost, st));
    RM_HIP(hipStreamSynchronize(st));
    if (nrec == 0) return RM_OK;
    RM_CHECK(nrec * kFaThreads <= kMaxWorkItems, RM_ERR_SHAPE, "%lld records exceed one parse",
             (long long)nrec);
    if ((rc = ws_get(ctx, "fa_starts", (size_t)nrec * 8, &p))) return rc;
    int64_t *starts = (int64_t *)p;
    if ((rc = ws_get(ctx, "fa_stage", (size_t)nrec * W * 4, &p))) return rc;
    uint32_t *stage = (uint32_t *)p;
    if ((rc = ws_get(ctx, "fa_keep", (size_t)nrec * 8, &p))) return rc;
    int *keep = (int *)p, *kidx = keep + nrec;
    if ((rc = ws_get(ctx, "fa_hoff", (size_t)nrec * 12, &p))) return rc;
    int64_t *hoff = (int64_t *)p;
    int32_t *hlen = (int32_t *)(hoff + nrec);
    fa_scatter_kernel<<<dim3((unsigned)nb), dim3(kFaThreads), 0, st>>>(raw, n, bcnt, starts);
    const size_t lds = (size_t)W * 4;
    RM_CHECK(lds <= 64 * 1024, RM_ERR_ARG, "k = %d too large for the packer", k);
    RM_HIP(hipFuncSetAttribute((const void *)fa_pack_kernel,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    fa_pack_kernel<<<dim3((unsigned)nrec), dim3(kFaThreads), lds, st>>>(raw, n, starts, nrec, k, W,
                                                                       stage, keep, hoff, hlen);
    RM_HIP(hipMemcpyAsync(kidx, keep, (size_t)nrec * 4, hipMemcpyDeviceToDevice, st));
    fa_scan_kernel<<<1, kFaThreads, 0, st>>>(kidx, nrec, dtot);
    RM_HIP(hipGetLastError());
    int64_t nkeep = 0;
    RM_HIP(hipMemcpyAsync(&nkeep, dtot, 8, hipMemcpyDeviceToHost, st));
    RM_HIP(hipStreamSynchronize(st));
    *n_reads = nkeep;
    RM_CHECK(nkeep <= max_reads, RM_ERR_SHAPE, "%lld reads kept, output holds %lld",
             (long long)nkeep, (long long)max_reads);
    if (nkeep == 0) return RM_OK;
    RM_CHECK(words, RM_ERR_ARG, "words is NULL");
    fa_compact_kernel<<<dim3(stride_grid(nrec * W, 256)), dim3(256), 0, st>>>(
        stage, keep, kidx, nrec, W, hoff, hlen, words, hdr_off, hdr_len);
    RM_HIP(hipGetLastError());
    return RM_OK;
}

}  // namespace rm
