// Host check of the batched stencil assembly's slot rule (batch.hip,
// assemble_vertex_stencil): on a topology with the stencil assembly form,
// vertex i's spring at assembly slot k (ainc tables) is the pair at stencil
// position p with p's presence bit set in asm_mask, and k equals the number
// of the vertex's springs at positions before p (presence bits plus
// second-spring bits), plus one for a pair's second spring.  Reads
// nx, ny, n_springs and the oc_spring records from the file named on the
// command line; prints "ok", "no stencil assembly form" or the first mismatch.
#include <bit>
#include <cstdio>
#include <string>
#include <vector>

#include "oc_internal.hpp"

int main(int argc, char** argv) {
    if (argc != 2) return std::printf("usage: stencil_slots file\n"), 2;
    FILE* f = std::fopen(argv[1], "rb");
    if (!f) return std::printf("cannot open %s\n", argv[1]), 2;
    int32_t hdr[3];
    if (std::fread(hdr, 4, 3, f) != 3) return std::printf("short header\n"), 2;
    std::vector<oc_spring> sp(hdr[2]);
    if (std::fread(sp.data(), sizeof(oc_spring), sp.size(), f) != sp.size()) return std::printf("short file\n"), 2;
    std::fclose(f);
    oc::GridTopology t;
    std::string err;
    if (oc::build_grid_topology(hdr[0], hdr[1], sp.data(), hdr[2], &t, &err) != OC_OK)
        return std::printf("topology: %s\n", err.c_str()), 1;
    if (!t.stencil_asm) return std::printf("no stencil assembly form\n"), 0;
    for (int i = 0; i < t.N; i++) {
        const uint32_t mi = t.asm_mask[i];
        int seen = 0, prev_pos = -1;
        for (int k = 0; k < t.wa; k++) {
            const int s = t.ainc_spring[(size_t)k * t.N + i];
            if (s < 0) break;
            const uint16_t m = t.ainc_meta[(size_t)k * t.N + i];
            int pos = -1;
            for (int q = 0; q < oc::kStencil; q++)
                if (oc::slot_di(m) == oc::stencil_di(q) && oc::slot_dj(m) == oc::stencil_dj(q)) pos = q;
            if (pos < 0 || !((mi >> pos) & 1u))
                return std::printf("vertex %d slot %d: position %d not in the mask\n", i, k, pos), 1;
            const uint32_t below = (1u << pos) - 1u;
            const int first = std::popcount(mi & 0xFFFu & below) + std::popcount((mi >> 16) & below);
            const bool second = pos == prev_pos;
            if (second && !((mi >> (16 + pos)) & 1u))
                return std::printf("vertex %d slot %d: second spring without its bit\n", i, k), 1;
            if (k != first + (second ? 1 : 0))
                return std::printf("vertex %d slot %d: rule gives %d\n", i, k, first + (second ? 1 : 0)), 1;
            // the first spring of a pair is the one with the lower id
            if (second && s < t.ainc_spring[(size_t)(k - 1) * t.N + i])
                return std::printf("vertex %d slot %d: springs out of id order\n", i, k), 1;
            prev_pos = pos;
            seen++;
        }
        const int total = std::popcount(mi & 0xFFFu) + std::popcount((mi >> 16) & 0xFFFu);
        if (seen != total) return std::printf("vertex %d: %d springs, mask counts %d\n", i, seen, total), 1;
    }
    std::printf("ok\n");
    return 0;
}
