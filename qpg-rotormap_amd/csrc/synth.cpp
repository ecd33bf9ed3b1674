// synth.cpp -- deterministic synthetic inputs on the host: reference records,
// N runs, window slicing and the read error models.  Every random draw is a
// counter-based hash of (seed, stream, counter), so any record or read can be
// regenerated alone, in any order, on any number of threads.
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/rotormap_hip.h"

namespace {

inline uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
inline uint64_t draw(uint64_t seed, uint64_t stream, uint64_t ctr) {
    return mix64(mix64(seed ^ mix64(stream + 0x632BE59BD9B4E019ull)) + ctr);
}
inline double unit(uint64_t h) { return (double)(h >> 11) * (1.0 / 9007199254740992.0); }

struct Rng {  // sequential draws of one (seed, stream)
    uint64_t seed, stream, ctr = 0;
    uint64_t next() { return draw(seed, stream, ctr++); }
    double u01() { return unit(next()); }          // [0, 1)
    int64_t below(int64_t n) { return (int64_t)(u01() * (double)n); }
};

inline uint8_t get_code(const uint32_t *w, int64_t j) { return (w[j >> 4] >> (2 * (j & 15))) & 3u; }

}  // namespace

extern "C" {

int rm_synth_record(uint64_t seed, int64_t rec, int64_t len, double gc, uint32_t *words) {
    if (!words || len < 0 || !(gc >= 0.0 && gc <= 1.0)) return RM_ERR_ARG;
    const int64_t W = (len + 15) / 16;
    const double pc = gc * 0.5, pg = gc, pa = gc + (1.0 - gc) * 0.5;
#pragma omp parallel for schedule(static)
    for (int64_t w = 0; w < W; ++w) {
        uint32_t v = 0;
        for (int j = 0; j < 16; ++j) {
            const int64_t p = w * 16 + j;
            if (p >= len) break;
            const double u = unit(draw(seed, (uint64_t)rec, (uint64_t)p));
            const uint32_t code = u < pc ? 1u : (u < pg ? 2u : (u < pa ? 0u : 3u));
            v |= code << (2 * j);
        }
        words[w] = v;
    }
    words[W] = 0u;  // zero guard word for funnel reads
    return RM_OK;
}

int rm_mark_n(uint32_t *words, int64_t lo, int64_t hi) {
    if (!words || lo < 0 || hi < lo) return RM_ERR_ARG;
    for (int64_t p = lo; p < hi; ++p) {
        uint32_t &w = words[p >> 4];
        const int sh = 2 * (int)(p & 15);
        w = (w & ~(3u << sh)) | (2u << sh);
    }
    return RM_OK;
}

int rm_extract_windows_host(uint32_t *out, const uint32_t *rec, const int64_t *starts, int64_t n,
                            int k) {
    if (!out || !rec || !starts || k < 1 || n < 0) return RM_ERR_ARG;
    const int W = (k + 15) / 16, r = k & 15;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        for (int w = 0; w < W; ++w) {
            const int64_t b = 2 * starts[i] + 32 * (int64_t)w;
            const int64_t q = b >> 5;
            const uint32_t sh = (uint32_t)(b & 31);
            uint32_t v = sh ? ((rec[q] >> sh) | (rec[q + 1] << (32u - sh))) : rec[q];
            if (w == W - 1 && r) v &= (1u << (2 * r)) - 1u;
            out[i * W + w] = v;
        }
    }
    return RM_OK;
}

// model 0: substitutions only; model 1: RotorMap/common/dna.jl:120-183.
int rm_mutate_words(uint32_t *words, int64_t n, int k, const double *err, uint64_t seed,
                    const int64_t *read_ids, int model) {
    if (!words || !err || k < 1 || n < 0 || (model != 0 && model != 1)) return RM_ERR_ARG;
    const int W = (k + 15) / 16;
    int bad = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(| : bad)
    for (int64_t i = 0; i < n; ++i) {
        const double e = err[i];
        if (!(e >= 0.0 && e <= 1.0)) { bad = 1; continue; }
        if (e == 0.0) continue;
        uint32_t *w = words + i * W;
        std::vector<uint8_t> orig(k), mut;
        for (int j = 0; j < k; ++j) orig[j] = get_code(w, j);
        Rng rng{seed, (uint64_t)(read_ids ? read_ids[i] : i)};
        const int n_err = (int)std::ceil((double)k * e);
        const int n_ins = model == 1 ? n_err / 3 : 0;
        const int n_del = n_ins;
        const int n_sub = n_err - n_ins - n_del;
        // substitutions: positions with replacement, a different base each
        std::vector<int> is(n_sub);
        for (int t = 0; t < n_sub; ++t) is[t] = (int)std::min<int64_t>(rng.below(k), k - 1);
        std::sort(is.begin(), is.end());
        mut = orig;
        for (int t = 0; t < n_sub; ++t) mut[is[t]] = (uint8_t)((orig[is[t]] + 1 + rng.below(3)) & 3);
        if (model == 1 && n_del > 0) {
            // deletions: sorted, made strictly increasing (dna.jl:146-155)
            std::vector<int> id(n_del);
            for (int t = 0; t < n_del; ++t) id[t] = (int)std::min<int64_t>(rng.below(k - t), k - t - 1);
            std::sort(id.begin(), id.end());
            for (int t = 1; t < n_del; ++t) id[t] = std::max(id[t], id[t - 1] + 1);
            std::vector<uint8_t> kept;
            kept.reserve(k);
            int q = 0;
            for (int j = 0; j < k; ++j) {
                if (q < n_del && id[q] == j) { ++q; continue; }
                kept.push_back(mut[j]);
            }
            // insertions: sorted positions in the shortened read, random bases
            const int L = (int)kept.size();
            std::vector<int> ii(n_ins);
            for (int t = 0; t < n_ins; ++t) ii[t] = (int)std::min<int64_t>(rng.below(L + 1), L);
            std::sort(ii.begin(), ii.end());
            std::vector<uint8_t> out;
            out.reserve(k);
            int u = 0;
            for (int j = 0; j <= L; ++j) {
                while (u < n_ins && ii[u] == j) { out.push_back((uint8_t)rng.below(4)); ++u; }
                if (j < L) out.push_back(kept[j]);
            }
            mut.assign(out.begin(), out.begin() + std::min<size_t>(out.size(), (size_t)k));
            while ((int)mut.size() < k) mut.push_back(0);
        }
        std::memset(w, 0, sizeof(uint32_t) * W);
        for (int j = 0; j < k; ++j) w[j >> 4] |= (uint32_t)mut[j] << (2 * (j & 15));
    }
    return bad ? RM_ERR_ARG : RM_OK;
}

}  // extern "C"
