// encode.hip -- the RoPE s-mer-histogram encoder for gfx950, plus the small
// word kernels around it (reverse complement, window slicing, conversions).
//
// Semantics follow RotorMap/encode/kernel.jl:86-290 (kernel_rope_frag_real)
// and its CPU golden twin RotorMap/encode/reference.jl:37-105:
//   window p (0-based, p < nw = k-s+1) -> little-endian s-mer field of bases
//   [p, p+s); compact key = lower ^ ((upper * 0x9E3779B1) & (4^c-1));
//   bin = key*m + idm gets the phasor exp(+i*theta_idm*p),
//   theta_idm = f*2pi/k, f = 1 (m == 1) else 2*idm/(m-1) + 1;
//   the G-homopolymer key 2*sum_{i<c}4^i is zeroed; normalize modes 0-3;
//   output [Re(b_0..b_{D-1}) ; Im(...)], D = m*4^c.
//
// MI355X design: one workgroup (4 wave64) per fragment; the fragment's words
// are staged in LDS and each lane unpacks its window's s-mer from two staged
// words; each window adds an exact integer phasor (2^30 fixed point, from a
// per-(k,s,m) table in HBM/L2) into int64 LDS bins with ds_add_u64, and a
// wave whose 64 windows share one key (a ballot) reduces across lanes and
// adds once.  Integer addition is associative, so the histogram -- and
// everything computed from it in a fixed order afterwards -- is bitwise
// deterministic regardless of wave scheduling (the reference's fp32 LDS
// atomics are not).  Normalisation runs in FP32, as the reference kernel's
// (RotorMap/encode/kernel.jl:189-279), with one correctly rounded operation
// at a time in a fixed reduction order (wave 0: 64 lane partials + halving
// tree), so outputs are reproducible bit for bit on the CPU; no step uses
// FP64.
#include "rm_device.h"
#include "rm_internal.h"

#include <cmath>

namespace rm {

constexpr float kInvS30f = 1.0f / 1073741824.0f;  // 2^-30
constexpr float kEps52 = 2.220446049250313e-16f;    // 2^-52 (Julia's eps()), exact in fp32

// Host: the exact phasor table (int32 pairs, idm-major).  The phase fraction
// f*p/k is formed exactly in integers before the FP64 cos/sin.
void build_phase_table(int k, int s, int m, int32_t *tab) {
    int nw = k - s + 1;
    for (int idm = 0; idm < m; ++idm) {
        int64_t mul = (m == 1) ? 1 : (int64_t)(2 * idm + m - 1);
        int64_t den = (m == 1) ? (int64_t)k : (int64_t)(m - 1) * (int64_t)k;
        for (int p = 0; p < nw; ++p) {
            int64_t num = (mul * (int64_t)p) % den;
            double ang = (2.0 * M_PI) * ((double)num / (double)den);
            tab[2 * ((size_t)idm * nw + p) + 0] = (int32_t)llround(cos(ang) * 1073741824.0);
            tab[2 * ((size_t)idm * nw + p) + 1] = (int32_t)llround(sin(ang) * 1073741824.0);
        }
    }
}

__device__ __forceinline__ float wave_tree_sum(float v) {
    // lanes l < off add lane l+off; lane 0 ends with the halving-tree sum
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        float o = __shfl_down(v, off, 64);
        v = v + o;
    }
    return v;
}

__global__ __launch_bounds__(256) void rope_encode_kernel(
    const uint32_t *__restrict__ words, int64_t n, int k, int s, int m, int c, int normalize,
    const int2 *__restrict__ table, EncOut o1, EncOut o2, float *__restrict__ norms,
    int32_t *__restrict__ counts) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int64_t f = blockIdx.x;
    if (f >= n) return;
    const int t = threadIdx.x;
    const int nk = 1 << (2 * c);
    const int D = m * nk;
    const int W = (k + 15) >> 4;
    const int nw = k - s + 1;

    long long *acc = (long long *)smem;                  // 2D int64, interleaved re/im
    float *cp = (float *)(acc + 2 * D);                  // 64 per-copy energies
    float *nr = cp + 64;                                 // 64 per-copy norms
    uint32_t *wsh = (uint32_t *)(nr + 64);               // W + 1 words
    int32_t *cnt = (int32_t *)(wsh + W + 2);             // 4^c counts (when requested)

    const uint32_t *fw = words + f * (int64_t)W;
    for (int i = t; i < W; i += 256) wsh[i] = fw[i];
    if (t == 0) wsh[W] = 0u;
    for (int i = t; i < 2 * D; i += 256) acc[i] = 0;
    if (counts)
        for (int i = t; i < nk; i += 256) cnt[i] = 0;
    __syncthreads();

    const uint32_t bmask = (s >= 16) ? 0xffffffffu : ((1u << (2 * s)) - 1u);
    const uint32_t cbmask = (c >= 16) ? 0xffffffffu : ((1u << (2 * c)) - 1u);
    const int csh = 2 * c;
    for (int p = t; p < nw; p += 256) {
        const int wo = p >> 4;
        const uint32_t sh = (uint32_t)(p & 15) << 1;
        const uint32_t w0 = wsh[wo], w1 = wsh[wo + 1];
        const uint32_t v = sh ? ((w0 >> sh) | (w1 << (32u - sh))) : w0;
        const uint32_t smer = v & bmask;
        const uint32_t lower = smer & cbmask;
        const uint32_t upper = (c >= 16) ? 0u : (smer >> csh);
        const uint32_t key = lower ^ ((upper * 0x9E3779B1u) & cbmask);
        // A ballot asks whether all 64 windows of the wave share one key (runs
        // of N, which read as G, and other homopolymer stretches): then the
        // wave sums its phasors across lanes and one lane adds the sums, where
        // 64 atomics to one LDS address would serialise.  Integer sums: the
        // result is the same either way.
        const bool uni = __ballot(1) == ~0ull &&
                         __ballot(key == (uint32_t)__builtin_amdgcn_readfirstlane(key)) == ~0ull;
        if (uni) {
            const int lane = t & 63;
            if (counts && lane == 0) atomicAdd(&cnt[key], 64);
            for (int idm = 0; idm < m; ++idm) {
                const int2 ph = table[(int64_t)idm * nw + p];
                long long x = ph.x, y = ph.y;
#pragma unroll
                for (int off = 32; off >= 1; off >>= 1) {
                    x += __shfl_xor(x, off, 64);
                    y += __shfl_xor(y, off, 64);
                }
                if (lane == 0) {
                    const int b = (int)key * m + idm;
                    atomicAdd((unsigned long long *)&acc[2 * b], (unsigned long long)x);
                    atomicAdd((unsigned long long *)&acc[2 * b + 1], (unsigned long long)y);
                }
            }
            continue;
        }
        if (counts) atomicAdd(&cnt[key], 1);
        for (int idm = 0; idm < m; ++idm) {
            const int2 ph = table[(int64_t)idm * nw + p];
            const int b = (int)key * m + idm;
            atomicAdd((unsigned long long *)&acc[2 * b], (unsigned long long)(long long)ph.x);
            atomicAdd((unsigned long long *)&acc[2 * b + 1], (unsigned long long)(long long)ph.y);
        }
    }
    __syncthreads();

    // G-homopolymer bins; counts are written before zeroing (raw histogram)
    if (counts)
        for (int i = t; i < nk; i += 256) counts[f * nk + i] = cnt[i];
    int allones = 0;
    for (int i = 0; i < c; ++i) allones += 1 << (2 * i);
    const int gkey = 2 * allones;
    if (t < m) { acc[2 * (gkey * m + t)] = 0; acc[2 * (gkey * m + t) + 1] = 0; }
    __syncthreads();

    // fixed point -> fp32 in place: bin b's (re, im) go to the low 4 bytes of
    // its own two 8-byte slots (no other thread reads them), each the
    // correctly rounded int64 -> fp32 conversion scaled by 2^-30 (exact)
    float *h = (float *)acc;  // re of bin b at h[4b], im at h[4b + 2]
    for (int b = t; b < D; b += 256) {
        float x = (float)acc[2 * b] * kInvS30f;
        float y = (float)acc[2 * b + 1] * kInvS30f;
        if (normalize == 2 || normalize == 3) {
            const float t1 = x * x, t2 = y * y;
            const float a = __builtin_sqrtf(t1 + t2);
            const float d = (normalize == 2) ? (a + kEps52) : __builtin_sqrtf(a + kEps52);
            x = x / d;
            y = y / d;
        }
        h[4 * b] = x;
        h[4 * b + 2] = y;
    }
    __syncthreads();

    if (t < 64) {  // wave 0: per-copy energies in the fixed order
        for (int idm = 0; idm < m; ++idm) {
            float part = 0.0f;
            for (int j = t; idm + (int64_t)j * m < D; j += 64) {
                const int b = idm + j * m;
                const float x = h[4 * b], y = h[4 * b + 2];
                const float t1 = x * x, t2 = y * y;
                part = part + (t1 + t2);
            }
            part = wave_tree_sum(part);
            if (t == 0) cp[idm] = part;
        }
        if (t == 0) {
            float *nf = norms ? norms + f * m : nullptr;
            if (normalize == 0) {
                float tot = cp[0];
                for (int idm = 1; idm < m; ++idm) tot = tot + cp[idm];
                const float nrm = __builtin_sqrtf(tot) + 1e-16f;
                for (int idm = 0; idm < m; ++idm) nr[idm] = nrm;
                if (nf) {
                    nf[0] = nrm;
                    for (int idm = 1; idm < m; ++idm) nf[idm] = cp[idm];
                }
            } else {
                const float add = (normalize == 1) ? 1e-7f : kEps52;
                for (int idm = 0; idm < m; ++idm) {
                    nr[idm] = __builtin_sqrtf(cp[idm]) + add;
                    if (nf) nf[idm] = nr[idm];
                }
            }
        }
    }
    __syncthreads();

    for (int b = t; b < D; b += 256) {
        const float nv = nr[b % m];
        const float vr = h[4 * b] / nv;
        const float vi = h[4 * b + 2] / nv;
        store_elem(o1.p, o1.dtype, f * o1.sf + (int64_t)b * o1.sd, vr);
        store_elem(o1.p, o1.dtype, f * o1.sf + (int64_t)(b + D) * o1.sd, vi);
        if (o2.p) {
            store_elem(o2.p, o2.dtype, f * o2.sf + (int64_t)b * o2.sd, vr);
            store_elem(o2.p, o2.dtype, f * o2.sf + (int64_t)(b + D) * o2.sd, vi);
        }
    }
}

int launch_encode(const uint32_t *words, int64_t n, int k, int s, int m, int c, int normalize,
                  const int32_t *table, EncOut o1, EncOut o2, float *norms, int32_t *counts,
                  hipStream_t st) {
    if (n == 0) return RM_OK;
    const int D = m * (1 << (2 * c));
    const int W = (k + 15) / 16;
    size_t lds = (size_t)2 * D * 8 + 128 * 4 + (size_t)(W + 2) * 4 + (counts ? (size_t)(1 << (2 * c)) * 4 : 0);
    RM_CHECK(lds <= 160 * 1024, RM_ERR_SHAPE,
             "encoder needs %zu B of LDS (> 160 KiB): k = %d, D = %d too large", lds, k, D);
    RM_HIP(hipFuncSetAttribute((const void *)rope_encode_kernel,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    // one workgroup of 256 per fragment: launches of < 2^32 work-items
    const int64_t chunk = kMaxWorkItems / 256;
    auto esz = [](int dt) { return dt == RM_F32 ? 4 : dt == RM_F16 ? 2 : 1; };
    for (int64_t f0 = 0; f0 < n; f0 += chunk) {
        const int64_t nf = std::min(chunk, n - f0);
        EncOut a1 = o1, a2 = o2;
        a1.p = (uint8_t *)o1.p + f0 * o1.sf * esz(o1.dtype);
        if (o2.p) a2.p = (uint8_t *)o2.p + f0 * o2.sf * esz(o2.dtype);
        rope_encode_kernel<<<dim3((unsigned)nf), dim3(256), lds, st>>>(
            words + f0 * W, nf, k, s, m, c, normalize, (const int2 *)table, a1, a2,
            norms ? norms + f0 * m : nullptr, counts ? counts + f0 * ((int64_t)1 << (2 * c)) : nullptr);
        RM_HIP(hipGetLastError());
    }
    return RM_OK;
}

// Reverse the sixteen 2-bit lanes of a word (lane contents kept).
__device__ __forceinline__ uint32_t lanerev32(uint32_t x) {
    x = ((x & 0x33333333u) << 2) | ((x >> 2) & 0x33333333u);
    x = ((x & 0x0f0f0f0fu) << 4) | ((x >> 4) & 0x0f0f0f0fu);
    x = ((x & 0x00ff00ffu) << 8) | ((x >> 8) & 0x00ff00ffu);
    return (x << 16) | (x >> 16);
}

// rc base j = 3 - fwd base (k-1-j).  Complementing a code is ~ on its lane;
// left-aligning the k complemented bases to the end of the W-word stream and
// reading it backwards word by word (lane-reversed) yields the rc words with
// a zero tail (the aligned shift pushes the padding out).
__global__ void revcomp_kernel(uint32_t *__restrict__ rc, const uint32_t *__restrict__ w,
                               int64_t n, int W, uint32_t sh) {
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < n * W;
         g += (int64_t)gridDim.x * blockDim.x) {
        const int64_t fr = g / W;
        const int p = (int)(g % W);
        const int i = W - 1 - p;  // aligned source word
        const uint32_t *fw = w + fr * W;
        const uint32_t x = ~fw[i];
        uint32_t v = x;
        if (sh) v = (x << sh) | (i > 0 ? (~fw[i - 1] >> (32u - sh)) : 0u);
        rc[g] = lanerev32(v);
    }
}

int launch_revcomp(uint32_t *rc, const uint32_t *w, int64_t n, int k, hipStream_t st) {
    int W = (k + 15) / 16;
    int64_t tot = n * W;
    if (tot == 0) return RM_OK;
    uint32_t sh = (uint32_t)(32 * W - 2 * k);
    revcomp_kernel<<<dim3(stride_grid(tot, 256)), dim3(256), 0, st>>>(rc, w, n, W, sh);
    RM_HIP(hipGetLastError());
    return RM_OK;
}

// Window i: bases [starts[i], starts[i]+k) of the packed record, re-packed
// from bit 0 with the tail masked to zero.
__global__ void extract_kernel(uint32_t *__restrict__ out, const uint32_t *__restrict__ rec,
                               const int64_t *__restrict__ starts, int64_t n, int k, int W) {
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < n * W;
         g += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = g / W;
        const int w = (int)(g % W);
        const int64_t b = 2 * starts[i] + 32 * (int64_t)w;
        const int64_t q = b >> 5;
        const uint32_t sh = (uint32_t)(b & 31);
        uint32_t v = sh ? ((rec[q] >> sh) | (rec[q + 1] << (32u - sh))) : rec[q];
        const int r = k & 15;
        if (w == W - 1 && r) v &= (1u << (2 * r)) - 1u;
        out[g] = v;
    }
}

int launch_extract(uint32_t *out, const uint32_t *rec, const int64_t *starts, int64_t n, int k,
                   hipStream_t st) {
    int W = (k + 15) / 16;
    int64_t tot = n * W;
    if (tot == 0) return RM_OK;
    extract_kernel<<<dim3(stride_grid(tot, 256)), dim3(256), 0, st>>>(out, rec, starts, n, k, W);
    RM_HIP(hipGetLastError());
    return RM_OK;
}

__global__ void quantize_kernel(uint8_t *__restrict__ dst, const void *__restrict__ src, int sd,
                                int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const float v = (sd == RM_F32) ? ((const float *)src)[i]
                                       : f16_bits_to_f32(((const uint16_t *)src)[i]);
        dst[i] = f32_to_e4m3(v);
    }
}
__global__ void dequantize_kernel(float *__restrict__ dst, const uint8_t *__restrict__ src, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = e4m3_byte_to_f32(src[i], 0);
}
__global__ void widen_kernel(uint16_t *__restrict__ dst, const uint8_t *__restrict__ src, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = f32_to_f16_bits(e4m3_to_f32(src[i]));
}

int launch_quantize(uint8_t *dst, const void *src, int sd, int64_t n, hipStream_t st) {
    if (n == 0) return RM_OK;
    quantize_kernel<<<dim3(stride_grid(n, 256)), dim3(256), 0, st>>>(dst, src, sd, n);
    RM_HIP(hipGetLastError());
    return RM_OK;
}
int launch_dequantize(float *dst, const uint8_t *src, int64_t n, hipStream_t st) {
    if (n == 0) return RM_OK;
    dequantize_kernel<<<dim3(stride_grid(n, 256)), dim3(256), 0, st>>>(dst, src, n);
    RM_HIP(hipGetLastError());
    return RM_OK;
}
int launch_widen_f16(uint16_t *dst, const uint8_t *src, int64_t n, hipStream_t st) {
    if (n == 0) return RM_OK;
    widen_kernel<<<dim3(stride_grid(n, 256)), dim3(256), 0, st>>>(dst, src, n);
    RM_HIP(hipGetLastError());
    return RM_OK;
}

}  // namespace rm
