// Host walk-through of the batched spring-once assembly's bands and rings
// (batch.hip, assemble_once): a cloth swept in bands of T vertices, each
// owned direction's results kept in a ring of ring_len(o, nx) entries.  For
// every band, direction and vertex, the entry a partner reads (the band's
// result column, or the ring slot of an owner in an earlier band) must hold
// that pair's owner, and every LDS index must stay inside the workgroup's
// allocation (12 floats per thread, then ring_floats(nx) per cloth).  Reads
// nx, ny, n_springs and the oc_spring records from the file named first,
// the threads per cloth T and the block size NT after it; prints "ok" or the
// first mismatch.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "oc_internal.hpp"

int main(int argc, char** argv) {
    if (argc != 4) return std::printf("usage: band_ring file T NT\n"), 2;
    const int T = std::atoi(argv[2]), NT = std::atoi(argv[3]);
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
    if (!t.stencil_asm) return std::printf("no stencil assembly form\n"), 1;
    const int nx = t.nx, Ni = t.N, G = NT / T;
    const long lds = 12L * NT + (long)G * oc::ring_floats(nx);
    // the last cloth of the workgroup (largest ring offset)
    const int g = G - 1;
    std::vector<long> res(NT, -1);                          // owner id per result column
    std::vector<long> ring(oc::ring_base(oc::kOwned, nx), -1);  // owner id per ring entry
    const int nb = (Ni + T - 1) / T;
    for (int band = 0; band < nb; band++) {
        const int v0 = band * T;
        for (int o = 0; o < oc::kOwned; o++) {
            const int ko = oc::owned_k(o), kp = oc::owned_partner_k(o), off = oc::own_offset(o, nx);
            const int R = oc::ring_len(o, nx), rb = oc::ring_base(o, nx);
            if (R < off) return std::printf("o %d: ring %d shorter than offset %d\n", o, R, off), 1;
            for (int lt = 0; lt < T; lt++) {  // evaluate
                const int vv = v0 + lt, tid = g * T + lt;
                if (vv < Ni && ((t.asm_mask[vv] >> ko) & 1u)) res[tid] = vv;
            }
            for (int lt = 0; lt < T; lt++) {  // accumulate: the partner's read
                const int vv = v0 + lt, tid = g * T + lt;
                if (vv >= Ni || !((t.asm_mask[vv] >> kp) & 1u)) continue;
                const int jv = vv - off;
                long got, idx;
                if (jv >= v0) {
                    idx = 11L * NT + (tid - off);  // component 11 of the column
                    if (tid - off < g * T || tid - off >= NT) return std::printf("column out of the cloth\n"), 1;
                    got = res[tid - off];
                } else {
                    const int e = jv & (R - 1);
                    idx = 12L * NT + (long)g * oc::ring_floats(nx) + 12L * rb + 11L * R + e;
                    got = ring[rb + e];
                }
                if (idx < 0 || idx >= lds) return std::printf("band %d o %d vv %d: LDS index %ld of %ld\n", band, o, vv, idx, lds), 1;
                if (got != jv)
                    return std::printf("band %d o %d vertex %d: read owner %ld, expected %d\n", band, o, vv, got, jv), 1;
            }
            for (int lt = 0; lt < T; lt++) {  // keep the band's last R owners
                const int vv = v0 + lt, tid = g * T + lt;
                if (vv < Ni && ((t.asm_mask[vv] >> ko) & 1u) && vv >= v0 + T - R) ring[rb + (vv & (R - 1))] = res[tid];
            }
        }
    }
    std::printf("ok\n");
    return 0;
}
