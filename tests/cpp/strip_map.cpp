// Host check of the spring-once assembly's strip tables (oc_internal.hpp):
// for every owned direction and every vertex of a tile, the LDS column a
// partner reads is the one its pair's owner wrote, and every strip owner
// lies in the tile's 2-wide halo with its partner inside the tile; the
// strips of all directions, laid out one after another, map back to their
// direction and index, and for the kernel's tile fit one round of the block.  Prints
// "ok" or the first mismatch.
#include <cstdio>

#include "oc_internal.hpp"

template <int TX, int TY>
int check(bool kernel_tile) {
    constexpr int H = oc::kHalo;
    if (oc::strip_max<TX, TY>() > TX * TY) return std::printf("strip larger than a block\n"), 1;
    if (kernel_tile && oc::strip_total<TX, TY>() > TX * TY)
        return std::printf("strips larger than a block\n"), 1;
    // the strips of all directions laid out one after another
    int total = 0;
    for (int o = 0; o < oc::kOwned; o++) {
        if (oc::strip_base<TX, TY>(o) != total) return std::printf("o %d: base\n", o), 1;
        total += oc::strip_count<TX, TY>(o);
    }
    if (total != oc::strip_total<TX, TY>()) return std::printf("total %d\n", total), 1;
    for (int t = 0; t < total; t++) {
        int o = -1, s = -1;
        oc::strip_of_column<TX, TY>(t, o, s);
        if (o < 0 || o >= oc::kOwned || s < 0 || s >= oc::strip_count<TX, TY>(o) ||
            oc::strip_base<TX, TY>(o) + s != t)
            return std::printf("column %d: o %d s %d\n", t, o, s), 1;
    }
    for (int o = 0; o < oc::kOwned; o++) {
        const int dx = oc::own_dx(o), dy = oc::own_dy(o), n = oc::strip_count<TX, TY>(o);
        if (n > oc::strip_max<TX, TY>()) return std::printf("o %d: count %d > max\n", o, n), 1;
        // owner of every strip column: outside the tile, in the halo, partner inside
        for (int s = 0; s < n; s++) {
            int wx, wy;
            oc::strip_owner<TX, TY>(o, s, wx, wy);
            const bool in = wx >= 0 && wx < TX && wy >= 0 && wy < TY;
            const int px = wx + dx, py = wy + dy;
            const bool pin = px >= 0 && px < TX && py >= 0 && py < TY;
            if (in || !pin || wx < -H || wx >= TX + H || wy < -H || wy >= TY + H)
                return std::printf("o %d s %d: owner (%d,%d) partner (%d,%d)\n", o, s, wx, wy, px, py), 1;
            if (oc::strip_index<TX, TY>(o, wx, wy) != s)
                return std::printf("o %d s %d: index %d\n", o, s, oc::strip_index<TX, TY>(o, wx, wy)), 1;
        }
        // every partner in the tile whose owner is outside finds that owner's column
        int seen = 0;
        for (int ly = 0; ly < TY; ly++)
            for (int lx = 0; lx < TX; lx++) {
                const int wx = lx - dx, wy = ly - dy;
                if (wx >= 0 && wx < TX && wy >= 0 && wy < TY) continue;
                const int s = oc::strip_index<TX, TY>(o, wx, wy);
                int ox, oy;
                if (s < 0 || s >= n) return std::printf("o %d (%d,%d): index %d of %d\n", o, lx, ly, s, n), 1;
                oc::strip_owner<TX, TY>(o, s, ox, oy);
                if (ox != wx || oy != wy)
                    return std::printf("o %d (%d,%d): owner (%d,%d) vs (%d,%d)\n", o, lx, ly, ox, oy, wx, wy), 1;
                seen++;
            }
        if (seen != n) return std::printf("o %d: %d partners outside, %d columns\n", o, seen, n), 1;
    }
    return 0;
}

int main() {
    if (check<oc::kTileX, oc::kTileY>(true) || check<64, 8>(false) || check<32, 8>(false) ||
        check<32, 16>(false) || check<16, 16>(false) || check<128, 4>(false) || check<8, 8>(false))
        return 1;
    std::printf("ok\n");
    return 0;
}
