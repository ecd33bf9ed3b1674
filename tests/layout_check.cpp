// Host-side check of the fused kernel's LDS layout (rm_internal.h): for every
// list length and stage-buffer layout the candidate slots, the lists and the
// per-query words are disjoint and inside the LDS the launch requests (for
// the ring layout: the staged lists fit the ring).
#include <cstdio>

#include "../qpg-rotormap_amd/csrc/rm_internal.h"

int main() {
    int bad = 0;
    for (int kp = 1; kp <= rm::kMaxKP; ++kp)
      for (int br : {64, 128, 256}) {
        if (!rm::tile_rows_ok(kp, br)) continue;
        if (rm::uses_ring(kp)) {
            // ring layout: the ring, then candidate slots, tau, counters, 3
            // flags; a merge stages the lists' values and columns in ring
            // buffers 2 and 3 (32 entries per query), the first tile its seed
            // words in buffer 2
            const size_t total = rm::topk_lds_bytes(kp, br);
            const int capd = rm::topk_cand_slots(kp, br);
            const size_t cand_end = rm::ring_cand_offset(br) + (size_t)br * (capd + 1) * 8;
            const size_t end = cand_end + rm::tail_bytes(br) + (size_t)br * 4;  // + floor
            if (capd < 1 || rm::ring_cand_offset(br) < rm::ring_bytes(br) || end > total ||
                total > rm::kLdsMax || kp > 32 || rm::kRing < 4 ||
                rm::ring_staged_list_bytes(br) > rm::ring_buffer_bytes(br) ||
                (size_t)br * 4 > rm::ring_buffer_bytes(br) ||
                rm::kRing * rm::ring_buffer_bytes(br) != rm::ring_bytes(br)) {
                std::printf("ring kp=%d br=%d capd=%d: cands@%zu end %zu total %zu\n", kp, br, capd,
                            rm::ring_cand_offset(br), end, total);
                ++bad;
            }
            continue;
        }
        if (rm::ring_list(kp) && br != 256) continue;  // unswizzled layouts: one tile of 256
        const size_t total = rm::topk_lds_bytes(kp, br);
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
