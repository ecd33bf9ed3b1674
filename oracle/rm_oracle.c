/*
 * rm_oracle.c -- CPU restatement of RotorMap's mapping hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.  The
 * product (librotormap_hip.so) never links, loads or calls it.
 *
 * Two encoders live here:
 *   rmo_ref_rope_frag_real  the FP64 golden encoder, restated line by line
 *                           from RotorMap/encode/reference.jl:37-105
 *                           (cis(theta*p) in FP64, no recurrence);
 *   rmo_rope_fixed          the deterministic fixed-point encoder the HIP
 *                           kernel implements (integer phasor table, int64
 *                           bin sums, FP32 normalisation in a fixed order).
 *                           The HIP kernel must match it bit for bit; it is
 *                           itself pinned to rmo_ref_rope_frag_real by
 *                           tolerance and to closed forms by the tests.
 * and the rest of the path: fp16 / OCP e4m3 conversions, the FP64 GEMM of
 * dequantised operands, the total-order top-k, the fixed-order FP32 rescoring
 * dot, reverse complement, the s-mer key counts.
 *
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off, no fast-math: every
 * floating operation below is one IEEE-754 correctly rounded operation).
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define RMO_S30 1073741824.0 /* 2^30: fixed-point scale of the phasor table */

/* ------------------------------------------------------------------------ */
/* Fragment words <-> base codes                                             */
/* ------------------------------------------------------------------------ */

/* Base j (0-based) sits in word j>>4 at bit 2*(j&15).
 * RotorMap/encode/reference.jl:29-35 (_frag_codes). */
void rmo_frag_codes(const uint32_t *words, int k, uint8_t *codes) {
    for (int j = 0; j < k; ++j)
        codes[j] = (uint8_t)((words[j >> 4] >> (2 * (j & 15))) & 3u);
}

/* Codes -> forward 2-bit words, zero-padded tail.
 * RotorMap/common/harness.jl:112-118 (_ref_forward_pack). */
void rmo_forward_pack(const uint8_t *codes, int L, uint32_t *words) {
    int W = (L + 15) / 16;
    memset(words, 0, sizeof(uint32_t) * (size_t)W);
    for (int j = 0; j < L; ++j)
        words[j >> 4] |= (uint32_t)(codes[j] & 3u) << (2 * (j & 15));
}

/* Byte -> code: A/a=0 C/c=1 G/g=2 T/t=3, every other byte 2 (kept in place).
 * RotorMap/fasta/reader.jl:21-28 (_LUT3). */
uint8_t rmo_byte_code(uint8_t b) {
    switch (b) {
    case 'A': case 'a': return 0;
    case 'C': case 'c': return 1;
    case 'G': case 'g': return 2;
    case 'T': case 't': return 3;
    default: return 2;
    }
}

/* FASTA parse of one byte buffer, restating fasta_reads on a single part
 * (RotorMap/fasta/reader.jl): record starts = '>' at a line start (byte 0 or
 * after '\n' / '\r', _next_rec3 :61-76); bytes before the first record are
 * ignored (:247-250); header = '>' .. the byte before the first '\n', one
 * trailing '\r' dropped (:256-261); sequence bytes = the rest of the record
 * minus '\n' / '\r', mapped by _LUT3 (:21-28, _collect3! :105-160); a record
 * with >= k bases is trimmed to k and emitted (:266-270, _emit3! :163-197).
 * A header line with no '\n' before the next record (a bare '\r' line end)
 * yields no read and the parse goes on with the next record (the reference
 * ends its byte range there, so what it drops depends on the thread split).
 * Writes up to max_reads reads (words: W = ceil(k/16) per read) and their
 * header byte ranges; returns the number of reads kept. */
static int rmo_is_eol(uint8_t b) { return b == '\n' || b == '\r'; }

static int64_t rmo_next_rec(const uint8_t *raw, int64_t from, int64_t n) {
    for (int64_t p = from; p < n; ++p)
        if (raw[p] == '>' && (p == 0 || rmo_is_eol(raw[p - 1]))) return p;
    return n;
}

int64_t rmo_fasta_reads(const uint8_t *raw, int64_t n, int k, uint32_t *words, int64_t max_reads,
                        int64_t *hoff, int32_t *hlen) {
    const int W = (k + 15) / 16;
    uint8_t *codes = (uint8_t *)malloc((size_t)k);
    int64_t cnt = 0;
    int64_t s = rmo_next_rec(raw, 0, n);
    while (s < n) {
        const int64_t e = rmo_next_rec(raw, s + 1, n);
        int64_t h = s;
        while (h < e && raw[h] != '\n') ++h;
        if (h < e) {
            int64_t hend = h - 1;
            if (hend >= s && raw[hend] == '\r') --hend;
            int b = 0;
            for (int64_t i = h + 1; i < e && b < k; ++i)
                if (!rmo_is_eol(raw[i])) codes[b++] = rmo_byte_code(raw[i]);
            if (b >= k) {
                if (cnt < max_reads) {
                    rmo_forward_pack(codes, k, words + cnt * W);
                    hoff[cnt] = s;
                    hlen[cnt] = (int32_t)(hend - s + 1);
                }
                ++cnt;
            }
        }
        s = e;
    }
    free(codes);
    return cnt;
}

/* Reverse complement of a k-base fragment, codes -> codes.
 * RotorMap/index/build.jl:415-422 (_revcomp_words_ref): rc[j] = 3 - codes[k-1-j]. */
void rmo_revcomp_codes(const uint8_t *codes, int k, uint8_t *rc) {
    for (int j = 0; j < k; ++j) rc[j] = (uint8_t)(3u - codes[k - 1 - j]);
}

/* ------------------------------------------------------------------------ */
/* s-mer field and compact key                                               */
/* ------------------------------------------------------------------------ */

/* Compact key (0-based csmer) of the window at 0-based p.
 * RotorMap/encode/reference.jl:43-50: little-endian s-mer field,
 * lower ^ ((upper * 0x9E3779B1 mod 2^32) & (4^c-1)). */
static uint32_t rmo_key(const uint8_t *codes, int p, int s, int c) {
    uint32_t smer = 0;
    for (int j = 0; j < s; ++j) smer |= (uint32_t)codes[p + j] << (2 * j);
    uint32_t cbmask = (c >= 16) ? 0xffffffffu : ((1u << (2 * c)) - 1u);
    uint32_t lower = smer & cbmask;
    uint32_t upper = (c >= 16) ? 0u : (smer >> (2 * c));
    return lower ^ ((upper * 0x9E3779B1u) & cbmask);
}

/* Integer s-mer histogram: counts[csmer] = windows whose compact key is csmer
 * (before any G-bin zeroing; the bin bookkeeping of reference.jl:43-55). */
void rmo_smer_counts(const uint8_t *codes, int k, int s, int c, int32_t *counts) {
    int nbin = 1 << (2 * c);
    memset(counts, 0, sizeof(int32_t) * (size_t)nbin);
    int nw = k - s + 1;
    for (int p = 0; p < nw; ++p) counts[rmo_key(codes, p, s, c)]++;
}

/* 0-based compact key of the zeroed G-homopolymer bin: 2 * sum_{i<c} 4^i.
 * RotorMap/encode/reference.jl:56-60. */
static int rmo_gbin(int c) {
    int allones = 0;
    for (int i = 0; i < c; ++i) allones += 1 << (2 * i);
    return 2 * allones;
}

/* ------------------------------------------------------------------------ */
/* FP64 golden encoder (the reference's CPU golden path)                     */
/* ------------------------------------------------------------------------ */

/* RotorMap/encode/reference.jl:37-105 (_ref_rope_frag :37-93 + _ref_rope_frag_real :96-105).
 * out[2D] = [Re(b_0..b_{D-1}) ; Im(...)], norms[m].  theta = f * (2pi/k),
 * f = 1 (m == 1) else 2*idm/(m-1) + 1 (0-based idm); hist[b] += cis(theta*p). */
void rmo_ref_rope_frag_real(const uint8_t *codes, int k, int s, int m, int c,
                            int normalize, double *out, double *norms) {
    int nb4 = 1 << (2 * c);
    int D = m * nb4;
    double *hr = (double *)calloc((size_t)D, sizeof(double));
    double *hi = (double *)calloc((size_t)D, sizeof(double));
    int nw = k - s + 1;
    for (int p = 0; p < nw; ++p) {
        int cs = (int)rmo_key(codes, p, s, c);
        for (int idm = 0; idm < m; ++idm) {
            double f = (m == 1) ? 1.0 : 2.0 * idm / (double)(m - 1) + 1.0;
            double th = f * (2.0 * M_PI / (double)k);
            double a = th * (double)p;
            hr[cs * m + idm] += cos(a);
            hi[cs * m + idm] += sin(a);
        }
    }
    int g = rmo_gbin(c);
    for (int idm = 0; idm < m; ++idm) { hr[g * m + idm] = 0.0; hi[g * m + idm] = 0.0; }
    for (int i = 0; i < m; ++i) norms[i] = 0.0;
    const double eps = 2.220446049250313e-16; /* Julia eps() = eps(Float64) */
    if (normalize == 0) {
        double acc = 0.0;
        for (int b = 0; b < D; ++b) acc += hr[b] * hr[b] + hi[b] * hi[b];
        double nrm = sqrt(acc) + 1e-16;
        for (int b = 0; b < D; ++b) { hr[b] /= nrm; hi[b] /= nrm; }
        norms[0] = nrm;
    } else {
        if (normalize == 2 || normalize == 3) {
            for (int b = 0; b < D; ++b) {
                double a = sqrt(hr[b] * hr[b] + hi[b] * hi[b]);
                double d = (normalize == 2) ? (a + eps) : sqrt(a + eps);
                hr[b] /= d; hi[b] /= d;
            }
        }
        double add = (normalize == 1) ? 1e-7 : eps;
        for (int idm = 0; idm < m; ++idm) {
            double acc = 0.0;
            for (int b = idm; b < D; b += m) acc += hr[b] * hr[b] + hi[b] * hi[b];
            double nrm = sqrt(acc) + add;
            for (int b = idm; b < D; b += m) { hr[b] /= nrm; hi[b] /= nrm; }
            norms[idm] = nrm;
        }
    }
    for (int b = 0; b < D; ++b) { out[b] = hr[b]; out[D + b] = hi[b]; }
    free(hr); free(hi);
}

/* ------------------------------------------------------------------------ */
/* Deterministic fixed-point encoder (the HIP kernel's specification)        */
/* ------------------------------------------------------------------------ */

/* Phasor table: tab[2*(idm*nw + p) + {0,1}] = round(2^30 * {cos,sin}(2pi*q)),
 * q = frac(f*p/k) formed exactly in integers: f*p/k = (2*idm + m-1)*p /
 * ((m-1)*k) for m > 1, p/k for m == 1 (the phase of reference.jl:52-53). */
void rmo_phase_table(int k, int s, int m, int32_t *tab) {
    int nw = k - s + 1;
    for (int idm = 0; idm < m; ++idm) {
        int64_t mul = (m == 1) ? 1 : (int64_t)(2 * idm + m - 1);
        int64_t den = (m == 1) ? (int64_t)k : (int64_t)(m - 1) * (int64_t)k;
        for (int p = 0; p < nw; ++p) {
            int64_t num = (mul * (int64_t)p) % den;
            double ang = (2.0 * M_PI) * ((double)num / (double)den);
            tab[2 * ((size_t)idm * nw + p) + 0] = (int32_t)llround(cos(ang) * RMO_S30);
            tab[2 * ((size_t)idm * nw + p) + 1] = (int32_t)llround(sin(ang) * RMO_S30);
        }
    }
}

/* Sum of e[idx(j)] in the kernel's fixed order, in fp32: 64 lane partials
 * (lane l adds items l, l+64, ... sequentially) folded by a halving tree
 * (off = 32..1). */
static float rmo_tree_sum_f32(const float *e, int D, int m, int idm) {
    float part[64];
    for (int l = 0; l < 64; ++l) {
        float acc = 0.0f;
        for (int j = l; idm + (int64_t)j * m < D; j += 64) acc = acc + e[idm + j * m];
        part[l] = acc;
    }
    for (int off = 32; off >= 1; off >>= 1)
        for (int l = 0; l < off; ++l) part[l] = part[l] + part[l + off];
    return part[0];
}

/* The encoder the HIP kernel implements, bit for bit: integer bin sums of
 * the phasor table (exact, any order), then FP32 normalisation with one
 * correctly rounded operation at a time, the reference kernel's precision
 * (RotorMap/encode/kernel.jl:189-279) with its epsilons (1e-16 for mode 0,
 * 1e-7 for mode 1, eps() = 2^-52 for modes 2-3): bin value = fp32(sum) *
 * 2^-30; energies per copy in the fixed tree order; norms and division in
 * fp32.  out (2D, the fp32 results widened to double) holds the normalised
 * values before storage rounding; norms[m] fp32 (widened).  acc_out
 * (optional, 2D int64, interleaved re/im) returns the bin sums. */
void rmo_rope_fixed(const uint8_t *codes, int k, int s, int m, int c, int normalize,
                    const int32_t *tab, double *out, double *norms, int64_t *acc_out) {
    int nb4 = 1 << (2 * c);
    int D = m * nb4;
    int nw = k - s + 1;
    int64_t *ar = (int64_t *)calloc((size_t)D, sizeof(int64_t));
    int64_t *ai = (int64_t *)calloc((size_t)D, sizeof(int64_t));
    for (int p = 0; p < nw; ++p) {
        int cs = (int)rmo_key(codes, p, s, c);
        for (int idm = 0; idm < m; ++idm) {
            ar[cs * m + idm] += tab[2 * ((size_t)idm * nw + p) + 0];
            ai[cs * m + idm] += tab[2 * ((size_t)idm * nw + p) + 1];
        }
    }
    int g = rmo_gbin(c);
    for (int idm = 0; idm < m; ++idm) { ar[g * m + idm] = 0; ai[g * m + idm] = 0; }
    if (acc_out)
        for (int b = 0; b < D; ++b) { acc_out[2 * b] = ar[b]; acc_out[2 * b + 1] = ai[b]; }
    float *x = (float *)malloc(sizeof(float) * (size_t)D);
    float *y = (float *)malloc(sizeof(float) * (size_t)D);
    float *e = (float *)malloc(sizeof(float) * (size_t)D);
    const float inv = 1.0f / 1073741824.0f; /* 2^-30 */
    for (int b = 0; b < D; ++b) {
        x[b] = (float)ar[b] * inv; /* int64 -> fp32: correctly rounded (RNE) */
        y[b] = (float)ai[b] * inv;
    }
    const float eps = 2.220446049250313e-16f; /* 2^-52, exact in fp32 */
    if (normalize == 2 || normalize == 3) {
        for (int b = 0; b < D; ++b) {
            float t1 = x[b] * x[b], t2 = y[b] * y[b];
            float a = sqrtf(t1 + t2);
            float d = (normalize == 2) ? (a + eps) : sqrtf(a + eps);
            x[b] = x[b] / d; y[b] = y[b] / d;
        }
    }
    for (int b = 0; b < D; ++b) { float t1 = x[b] * x[b], t2 = y[b] * y[b]; e[b] = t1 + t2; }
    float cp[64];
    for (int idm = 0; idm < m; ++idm) cp[idm] = rmo_tree_sum_f32(e, D, m, idm);
    if (normalize == 0) {
        float tot = cp[0];
        for (int idm = 1; idm < m; ++idm) tot = tot + cp[idm];
        float nrm = sqrtf(tot) + 1e-16f;
        norms[0] = nrm;
        for (int idm = 1; idm < m; ++idm) norms[idm] = cp[idm];
        for (int b = 0; b < D; ++b) { out[b] = x[b] / nrm; out[D + b] = y[b] / nrm; }
    } else {
        float add = (normalize == 1) ? 1e-7f : eps;
        float nr[64];
        for (int idm = 0; idm < m; ++idm) { nr[idm] = sqrtf(cp[idm]) + add; norms[idm] = nr[idm]; }
        for (int b = 0; b < D; ++b) {
            out[b] = x[b] / nr[b % m];
            out[D + b] = y[b] / nr[b % m];
        }
    }
    free(ar); free(ai); free(x); free(y); free(e);
}

/* ------------------------------------------------------------------------ */
/* Storage formats: IEEE binary16 (RNE) and OCP FP8 E4M3FN (RNE, satfinite)   */
/* ------------------------------------------------------------------------ */

static uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

/* fp32 -> binary16 bits, round to nearest even (IEEE 754-2008, 5.3). */
uint16_t rmo_f32_to_f16(float f) {
    uint32_t u = f2u(f);
    uint16_t sign = (uint16_t)((u >> 16) & 0x8000u);
    uint32_t a = u & 0x7fffffffu;
    if (a > 0x7f800000u) return (uint16_t)(sign | 0x7e00u);      /* NaN */
    if (a >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u);     /* >= 65520 -> inf */
    if (a < 0x38800000u) {                                       /* below 2^-14: subnormal */
        /* value / 2^-24 rounded to nearest even integer */
        float v = u2f(a) * 16777216.0f;                          /* exact (power of two) */
        float r = rintf(v);
        return (uint16_t)(sign | (uint16_t)r);
    }
    uint32_t e = (a >> 23) - 127 + 15;
    uint32_t mant = a & 0x7fffffu;
    uint32_t h = (e << 10) | (mant >> 13);
    uint32_t rem = mant & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h += 1;   /* carry may bump exponent */
    return (uint16_t)(sign | h);
}

float rmo_f16_to_f32(uint16_t h) {
    uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
    uint32_t e = (h >> 10) & 0x1fu, mant = h & 0x3ffu;
    if (e == 0) {
        float v = (float)mant * (1.0f / 16777216.0f);
        return (h & 0x8000u) ? -v : v;
    }
    if (e == 31) return u2f(sign | 0x7f800000u | (mant << 13));
    return u2f(sign | ((e - 15 + 127) << 23) | (mant << 13));
}

/* fp32 -> e4m3fn: bias 7, max 448 (S.1111.110), NaN S.1111.111, no inf.
 * Round to nearest even; |x| > 448 saturates to 448 (the satfinite rule the
 * query side uses, RotorMap/gemm/fp8_ptx.jl:165).  Written from the OCP
 * 8-bit floating point specification (v1.0, section 4.1). */
uint8_t rmo_f32_to_e4m3(float f) {
    uint32_t u = f2u(f);
    uint8_t sign = (uint8_t)((u >> 24) & 0x80u);
    float a = fabsf(f);
    if (a != a) return (uint8_t)(sign | 0x7fu);
    if (a > 448.0f) return (uint8_t)(sign | 0x7eu);
    if (a < 0.015625f) {                       /* below 2^-6: multiples of 2^-9 */
        float q = rintf(a * 512.0f);            /* exact scaling; 0..8 */
        return (uint8_t)(sign | (uint8_t)q);    /* q == 8 is the min normal code */
    }
    int e = (int)((f2u(a) >> 23) & 0xffu) - 127; /* -6 .. 8 */
    float mant = a / ldexpf(1.0f, e);           /* [1, 2), exact */
    float q = rintf((mant - 1.0f) * 8.0f);      /* 0 .. 8 */
    int qi = (int)q;
    if (qi == 8) { e += 1; qi = 0; }
    int code = ((e + 7) << 3) | qi;
    if (code > 0x7e) code = 0x7e;
    return (uint8_t)(sign | (uint8_t)code);
}

float rmo_e4m3_to_f32(uint8_t b) {
    float sgn = (b & 0x80u) ? -1.0f : 1.0f;
    int e = (b >> 3) & 0xf, m = b & 7;
    if (e == 15 && m == 7) return u2f((b & 0x80u) ? 0xffc00000u : 0x7fc00000u);
    if (e == 0) return sgn * ldexpf((float)m, -9);
    return sgn * ldexpf(1.0f + (float)m / 8.0f, e - 7);
}

/* The storage chain both sides of the GEMM share (RotorMap/index/build.jl:730,
 * RotorMap/search/engine_fp8.jl:365-368): e4m3(fp32(fp16(fp32(x)))). */
uint8_t rmo_f64_to_e4m3_chain(double v) {
    float f = (float)v;
    return rmo_f32_to_e4m3(rmo_f16_to_f32(rmo_f32_to_f16(f)));
}

/* ------------------------------------------------------------------------ */
/* GEMM (FP64 of dequantised operands) and top-k                             */
/* ------------------------------------------------------------------------ */

/* C[r*N + j] = sum_k deq(Q[r*K + k]) * deq(B[j*K + k]) in FP64: the score
 * model of RotorMap/search/engine_fp8.jl:873-876 (Bh' * hrefq). */
void rmo_gemm_e4m3_f64(const uint8_t *Q, const uint8_t *B, int64_t M, int64_t N, int64_t K,
                       double *C) {
    float lut[256];
    for (int i = 0; i < 256; ++i) lut[i] = rmo_e4m3_to_f32((uint8_t)i);
    double *qr = (double *)malloc(sizeof(double) * (size_t)K);
    for (int64_t r = 0; r < M; ++r) {
        for (int64_t k = 0; k < K; ++k) qr[k] = lut[Q[r * K + k]];
        for (int64_t j = 0; j < N; ++j) {
            const uint8_t *bj = B + j * K;
            double acc = 0.0;
            for (int64_t k = 0; k < K; ++k) acc += qr[k] * (double)lut[bj[k]];
            C[r * N + j] = acc;
        }
    }
    free(qr);
}

/* Same for binary16 operands (the FP16 path). */
void rmo_gemm_f16_f64(const uint16_t *Q, const uint16_t *B, int64_t M, int64_t N, int64_t K,
                      double *C) {
    double *qr = (double *)malloc(sizeof(double) * (size_t)K);
    double *br = (double *)malloc(sizeof(double) * (size_t)K);
    for (int64_t r = 0; r < M; ++r) {
        for (int64_t k = 0; k < K; ++k) qr[k] = rmo_f16_to_f32(Q[r * K + k]);
        for (int64_t j = 0; j < N; ++j) {
            for (int64_t k = 0; k < K; ++k) br[k] = rmo_f16_to_f32(B[j * K + k]);
            double acc = 0.0;
            for (int64_t k = 0; k < K; ++k) acc += qr[k] * br[k];
            C[r * N + j] = acc;
        }
    }
    free(qr); free(br);
}

/* (value desc, column asc): a strict total order on (v, col) pairs. */
static int rmo_beats(float va, int32_t ca, float vb, int32_t cb) {
    return va > vb || (va == vb && ca < cb);
}

/* Per row of a row-major (M x N) fp32 score matrix: the top-k under the total
 * order, as 1-based columns (0 = empty, value -inf), the output convention of
 * RotorMap/gemm/topk_kernels.jl:56-60,141-160. */
void rmo_topk_rows_f32(const float *C, int64_t M, int64_t N, int64_t ldc, int k,
                       float *vals, int32_t *locs) {
    for (int64_t r = 0; r < M; ++r) {
        float *v = vals + r * k;
        int32_t *l = locs + r * k;
        for (int j = 0; j < k; ++j) { v[j] = -INFINITY; l[j] = 0; }
        int fill = 0;
        for (int64_t j = 0; j < N; ++j) {
            float x = C[r * ldc + j];
            int32_t cj = (int32_t)(j + 1);
            if (fill == k && !rmo_beats(x, cj, v[k - 1], l[k - 1])) continue;
            int pos = (fill < k) ? fill : k - 1;
            while (pos > 0 && rmo_beats(x, cj, v[pos - 1], l[pos - 1])) {
                v[pos] = v[pos - 1]; l[pos] = l[pos - 1]; --pos;
            }
            v[pos] = x; l[pos] = cj;
            if (fill < k) ++fill;
        }
    }
}

/* The rescoring dot in its fixed FP32 order: 64 lanes, lane l owns elements
 * [l*C, (l+1)*C) (C = ceil(K/64)), each a sequential sum of q*deq(b) products
 * (multiply, then add: two roundings), folded by an xor butterfly (off = 32..1,
 * s[l] = s[l] + s[l^off]).  Returns the value every lane ends with.  The
 * column b holds K e4m3 bytes (elem 0) or K binary16 values (elem 1: the
 * verbatim fp16 columns of an fp16 index). */
float rmo_rescore_dot_t(const float *q, const void *b, int elem, int K) {
    float s[64], t[64];
    int C = (K + 63) / 64;
    for (int l = 0; l < 64; ++l) {
        float acc = 0.0f;
        for (int j = l * C; j < (l + 1) * C && j < K; ++j) {
            float d = elem == 1 ? rmo_f16_to_f32(((const uint16_t *)b)[j])
                                : rmo_e4m3_to_f32(((const uint8_t *)b)[j]);
            float p = q[j] * d;
            acc = acc + p;
        }
        s[l] = acc;
    }
    for (int off = 32; off >= 1; off >>= 1) {
        for (int l = 0; l < 64; ++l) t[l] = s[l] + s[l ^ off];
        memcpy(s, t, sizeof(s));
    }
    return s[0];
}

float rmo_rescore_dot(const float *q, const uint8_t *b, int K) {
    return rmo_rescore_dot_t(q, b, 0, K);
}

/* Rescore a candidate list: scores[j] = rmo_rescore_dot_t(q, column col_j - 1
 * of B (N x K elements of type elem, column-contiguous)), then order by
 * (score desc, col asc) and keep the first ktop.  Empty candidates (col 0)
 * stay empty. */
void rmo_rescore_rank_t(const float *q, const void *B, int elem, int K, const int32_t *cand,
                        int nc, int ktop, float *vals, int32_t *locs) {
    float sc[1024];
    int32_t cc[1024];
    const size_t esz = elem == 1 ? 2 : 1;
    int n = 0;
    for (int j = 0; j < nc && n < 1024; ++j) {
        if (cand[j] <= 0) continue;
        sc[n] = rmo_rescore_dot_t(q, (const uint8_t *)B + (size_t)(cand[j] - 1) * K * esz, elem, K);
        cc[n] = cand[j];
        ++n;
    }
    for (int j = 0; j < ktop; ++j) { vals[j] = -INFINITY; locs[j] = 0; }
    for (int i = 0; i < n; ++i) {
        int rank = 0; /* candidates are distinct columns: ranks are distinct */
        for (int j = 0; j < n; ++j) rank += rmo_beats(sc[j], cc[j], sc[i], cc[i]);
        if (rank < ktop) { vals[rank] = sc[i]; locs[rank] = cc[i]; }
    }
}

void rmo_rescore_rank(const float *q, const uint8_t *B, int K, const int32_t *cand, int nc,
                      int ktop, float *vals, int32_t *locs) {
    rmo_rescore_rank_t(q, B, 0, K, cand, nc, ktop, vals, locs);
}

/* ------------------------------------------------------------------------ */
/* The reference's CPU golden mapping path, threaded (cpu_baseline only)      */
/* ------------------------------------------------------------------------ */

/* Per read: FP64 golden encoder (reference.jl:37-105) -> the query storage
 * chain e4m3(fp32(fp16(.))) (engine_fp8.jl:365-368) -> FP64 dot products
 * against every dequantised e4m3 column (engine_fp8.jl:873-876) -> top-ktop
 * by (score desc, column asc) (engine_fp8.jl:875).  OpenMP over reads. */
void rmo_cpu_map(const uint32_t *words, int64_t n, int k, int s, int m, int c,
                 const uint8_t *B, int64_t N, int K, int ktop, float *vals, int32_t *locs,
                 int nthreads) {
    float lut[256];
    for (int i = 0; i < 256; ++i) lut[i] = rmo_e4m3_to_f32((uint8_t)i);
    const int W = (k + 15) / 16;
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 1)
    for (int64_t r = 0; r < n; ++r) {
        uint8_t *codes = (uint8_t *)malloc((size_t)k);
        double *emb = (double *)malloc(sizeof(double) * (size_t)K);
        double *q = (double *)malloc(sizeof(double) * (size_t)K);
        double nr[64];
        rmo_frag_codes(words + r * W, k, codes);
        rmo_ref_rope_frag_real(codes, k, s, m, c, 0, emb, nr);
        for (int j = 0; j < K; ++j) q[j] = lut[rmo_f64_to_e4m3_chain(emb[j])];
        float *v = vals + r * ktop;
        int32_t *l = locs + r * ktop;
        for (int j = 0; j < ktop; ++j) { v[j] = -INFINITY; l[j] = 0; }
        int fill = 0;
        for (int64_t col = 0; col < N; ++col) {
            const uint8_t *b = B + col * K;
            double acc = 0.0;
            for (int j = 0; j < K; ++j) acc += q[j] * (double)lut[b[j]];
            float x = (float)acc;
            int32_t cj = (int32_t)(col + 1);
            if (fill == ktop && !rmo_beats(x, cj, v[ktop - 1], l[ktop - 1])) continue;
            int pos = (fill < ktop) ? fill : ktop - 1;
            while (pos > 0 && rmo_beats(x, cj, v[pos - 1], l[pos - 1])) {
                v[pos] = v[pos - 1]; l[pos] = l[pos - 1]; --pos;
            }
            v[pos] = x; l[pos] = cj;
            if (fill < ktop) ++fill;
        }
        free(codes); free(emb); free(q);
    }
}

/* ------------------------------------------------------------------------ */
/* Score models of the three GPU modes (tests and the bench's score error)   */
/* ------------------------------------------------------------------------ */

/* The golden FP64 encoder's output of one read (words: ceil(k/16) uint32)
 * rounded through a query storage chain:
 *   0  e4m3(fp32(fp16(fp32(x))))  the FP8 query (engine_fp8.jl:365-368,
 *                                 the hrefq of engine_fp8.jl:873)
 *   1  fp32(x)                    the rescoring query
 *   2  fp16(fp32(x))              the FP16 query (engine_fp16.jl upload)
 *   3  x                          the unquantised FP64 embedding
 * normalize = 0, the mapping flow's mode (RotorMap/index/load.jl:57-59). */
void rmo_golden_query(const uint32_t *words, int k, int s, int m, int c, int chain, double *q) {
    const int K = 2 * m * (1 << (2 * c));
    uint8_t *codes = (uint8_t *)malloc((size_t)k);
    double nr[64];
    rmo_frag_codes(words, k, codes);
    rmo_ref_rope_frag_real(codes, k, s, m, c, 0, q, nr);
    for (int j = 0; j < K; ++j) {
        const float f = (float)q[j];
        if (chain == 0) q[j] = rmo_e4m3_to_f32(rmo_f64_to_e4m3_chain(q[j]));
        else if (chain == 1) q[j] = f;
        else if (chain == 2) q[j] = rmo_f16_to_f32(rmo_f32_to_f16(f));
    }
    free(codes);
}

/* out[j] = sum_k q[k] * B[cols[j]-1][k] in FP64 over the stored column
 * (elem 0: e4m3 bytes, 1: binary16 bits), sequential in k; column id 0 (an
 * empty slot) gives NAN. */
void rmo_dots_f64(const double *q, const void *B, int elem, int64_t K, const int32_t *cols, int n,
                  double *out) {
    for (int j = 0; j < n; ++j) {
        if (cols[j] <= 0) { out[j] = NAN; continue; }
        const int64_t o = (int64_t)(cols[j] - 1) * K;
        double acc = 0.0;
        if (elem == 0) {
            const uint8_t *b = (const uint8_t *)B + o;
            for (int64_t t = 0; t < K; ++t) acc += q[t] * (double)rmo_e4m3_to_f32(b[t]);
        } else {
            const uint16_t *b = (const uint16_t *)B + o;
            for (int64_t t = 0; t < K; ++t) acc += q[t] * (double)rmo_f16_to_f32(b[t]);
        }
        out[j] = acc;
    }
}

/* Insert (x, col) into a sorted list (v, vd, l) of ktop entries holding fill. */
static void rmo_list_insert(float *v, double *vd, int32_t *l, int ktop, int *fill, float x,
                            double xd, int32_t cj) {
    if (*fill == ktop && !rmo_beats(x, cj, v[ktop - 1], l[ktop - 1])) return;
    int pos = (*fill < ktop) ? *fill : ktop - 1;
    while (pos > 0 && rmo_beats(x, cj, v[pos - 1], l[pos - 1])) {
        v[pos] = v[pos - 1]; vd[pos] = vd[pos - 1]; l[pos] = l[pos - 1]; --pos;
    }
    v[pos] = x; vd[pos] = xd; l[pos] = cj;
    if (*fill < ktop) ++*fill;
}

/* Full-scan FP64 top-ktop of n queries (row-major n x K doubles) over all N
 * stored columns (elem as rmo_dots_f64), ranked by (float(score) desc,
 * column asc) like rmo_cpu_map; vals are the FP64 scores.  OpenMP over
 * column blocks (each column is read once for all queries); the per-block
 * lists merge into the same result in any order, since (score, column)
 * keys are distinct. */
void rmo_scan_topk_f64(const double *Q, int64_t n, const void *B, int elem, int64_t N, int64_t K,
                       int ktop, double *vals, int32_t *locs, int nthreads) {
    float lut[256];
    for (int i = 0; i < 256; ++i) lut[i] = rmo_e4m3_to_f32((uint8_t)i);
    if (nthreads < 1) nthreads = 1;
    const int64_t per = (int64_t)n * ktop;
    float *tv = (float *)malloc(sizeof(float) * (size_t)(per * nthreads));
    double *tvd = (double *)malloc(sizeof(double) * (size_t)(per * nthreads));
    int32_t *tl = (int32_t *)malloc(sizeof(int32_t) * (size_t)(per * nthreads));
    int *tf = (int *)calloc((size_t)(n * nthreads), sizeof(int));
#pragma omp parallel num_threads(nthreads)
    {
#ifdef _OPENMP
        const int t = omp_get_thread_num(), nt = omp_get_num_threads();
#else
        const int t = 0, nt = 1;
#endif
        float *v = tv + per * t;
        double *vd = tvd + per * t;
        int32_t *l = tl + per * t;
        int *fill = tf + n * t;
        double *bcol = (double *)malloc(sizeof(double) * (size_t)K);
        for (int64_t col = t; col < N; col += nt) {
            if (elem == 0) {
                const uint8_t *b = (const uint8_t *)B + col * K;
                for (int64_t x = 0; x < K; ++x) bcol[x] = lut[b[x]];
            } else {
                const uint16_t *b = (const uint16_t *)B + col * K;
                for (int64_t x = 0; x < K; ++x) bcol[x] = rmo_f16_to_f32(b[x]);
            }
            for (int64_t r = 0; r < n; ++r) {
                const double *q = Q + r * K;
                double acc = 0.0;
                for (int64_t x = 0; x < K; ++x) acc += q[x] * bcol[x];
                rmo_list_insert(v + r * ktop, vd + r * ktop, l + r * ktop, ktop, fill + r,
                                (float)acc, acc, (int32_t)(col + 1));
            }
        }
        free(bcol);
    }
    for (int64_t r = 0; r < n; ++r) {
        float *v = (float *)malloc(sizeof(float) * (size_t)ktop);
        double *vd = vals + r * ktop;
        int32_t *l = locs + r * ktop;
        int fill = 0;
        for (int j = 0; j < ktop; ++j) { v[j] = -INFINITY; vd[j] = -INFINITY; l[j] = 0; }
        for (int t = 0; t < nthreads; ++t)
            for (int j = 0; j < tf[n * t + r]; ++j) {
                const int64_t e = per * t + r * ktop + j;
                rmo_list_insert(v, vd, l, ktop, &fill, tv[e], tvd[e], tl[e]);
            }
        free(v);
    }
    free(tv); free(tvd); free(tl); free(tf);
}
