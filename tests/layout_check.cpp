// Host-side check of the fused kernel's LDS layout (rm_internal.h): for every
// list length and stage-buffer layout the candidate slots, the lists and the
// per-query words are disjoint and inside the LDS the launch requests.
#include <cstdio>

#include "../qpg-rotormap_amd/csrc/rm_internal.h"

int main() {
    int bad = 0;
    for (int kp = 1; kp <= rm::kMaxKP; ++kp) {
        const int br = rm::rows_per_tile(kp);
        const size_t total = rm::topk_lds_bytes(kp);
        const int capd = rm::cand_slots(br, kp);
        // regions in kernel order: stage buffers, lists, candidate slots, tau,
        // counters, 3 flags
        const size_t lists_end = rm::list_offset(br) + (size_t)br * (kp + 1) * 8;
        const size_t cand_end = rm::cand_offset(br, kp) + (size_t)br * (capd + 1) * 8;
        const size_t end = cand_end + (size_t)br * 8 + 3 * 4;
        if (rm::list_offset(br) < rm::stage_pair_bytes(br) || rm::cand_offset(br, kp) < lists_end ||
            capd < 1 || end > total || total > rm::kLdsMax) {
            std::printf("kp=%d br=%d capd=%d: lists@%zu cands@%zu end %zu total %zu\n", kp, br, capd,
                        rm::list_offset(br), rm::cand_offset(br, kp), end, total);
            ++bad;
        }
    }
    std::printf("layout %s (kRow %d)\n", bad ? "BAD" : "ok", rm::kRow);
    return bad != 0;
}
