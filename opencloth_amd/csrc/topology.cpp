// Host-side topology: spring list -> unique pairs, block planes and the
// per-vertex slot tables that the HIP kernels gather from.
//
// The reference scatters each spring's force into both endpoints in list
// order (OpenCloth_ImplicitEuler/.../main.cpp:658-681).  Here every vertex
// gathers its incident springs instead, in the same relative order, so the
// per-vertex sums match the reference's scatter while no two threads write
// the same vertex (no atomics, deterministic).
#include <algorithm>
#include <cstdlib>
#include <numeric>

#include "oc_internal.hpp"

namespace oc {

int build_grid_topology(int nx, int ny, const oc_spring* springs, int n_springs,
                        GridTopology* out, std::string* err) {
    if (nx < 1 || ny < 1 || (int64_t)nx * ny > (1 << 28)) {
        *err = "grid size out of range";
        return OC_ERR_INVALID_ARGUMENT;
    }
    if (n_springs < 0 || (n_springs > 0 && !springs)) {
        *err = "invalid spring list";
        return OC_ERR_INVALID_ARGUMENT;
    }
    if (n_springs > (1 << 29)) {  // the incidence offsets below count 2 per spring in int
        *err = "too many springs";
        return OC_ERR_INVALID_ARGUMENT;
    }
    GridTopology& t = *out;
    t = GridTopology();
    t.nx = nx;
    t.ny = ny;
    const int N = nx * ny;
    t.N = N;
    t.S = n_springs;
    const int S = n_springs;

    for (int s = 0; s < S; s++) {
        int a = springs[s].p1, b = springs[s].p2;
        if (a < 0 || a >= N || b < 0 || b >= N || a == b) {
            *err = "spring " + std::to_string(s) + " has an invalid endpoint";
            return OC_ERR_INVALID_ARGUMENT;
        }
        int ddi = (b % nx) - (a % nx), ddj = (b / nx) - (a / nx);
        if (std::abs(ddi) > kHalo || std::abs(ddj) > kHalo) {
            *err = "spring " + std::to_string(s) + " spans more than " + std::to_string(kHalo) +
                   " grid rows/columns";
            return OC_ERR_UNSUPPORTED_TOPOLOGY;
        }
    }

    // unique unordered pairs, numbered by first spring in list order
    std::vector<std::pair<int64_t, int>> key(S);
    for (int s = 0; s < S; s++) {
        int a = springs[s].p1, b = springs[s].p2;
        int lo = std::min(a, b), hi = std::max(a, b);
        key[s] = {(int64_t)lo * N + hi, s};
    }
    std::sort(key.begin(), key.end());
    std::vector<int> grp_first;
    std::vector<int> spring_grp(S);
    for (int k = 0; k < S; k++) {
        if (k == 0 || key[k].first != key[k - 1].first) grp_first.push_back(key[k].second);
        spring_grp[key[k].second] = (int)grp_first.size() - 1;
    }
    const int P = (int)grp_first.size();
    std::vector<int> order(P);
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.begin(), order.end(),
              [&](int x, int y) { return grp_first[x] < grp_first[y]; });
    std::vector<int> grp_to_pair(P);
    for (int p = 0; p < P; p++) grp_to_pair[order[p]] = p;
    t.P = P;
    t.pair_a.assign(P, 0);
    t.pair_b.assign(P, 0);
    t.pair_plane.assign(P, 0);
    t.spring_pair.assign(S, 0);
    for (int s = 0; s < S; s++) {
        int p = grp_to_pair[spring_grp[s]];
        t.spring_pair[s] = p;
        t.pair_a[p] = std::min(springs[s].p1, springs[s].p2);
        t.pair_b[p] = std::max(springs[s].p1, springs[s].p2);
    }

    // planes: one per distinct offset of a pair's upper vertex from its lower
    for (int p = 0; p < P; p++) {
        int a = t.pair_a[p], b = t.pair_b[p];
        int ddi = (b % nx) - (a % nx), ddj = (b / nx) - (a / nx);
        int plane = -1;
        for (int k = 0; k < t.n_planes; k++)
            if (t.di[k] == ddi && t.dj[k] == ddj) plane = k;
        if (plane < 0) {
            if (t.n_planes == kMaxPlanes) {
                *err = "too many distinct spring directions";
                return OC_ERR_UNSUPPORTED_TOPOLOGY;
            }
            plane = t.n_planes++;
            t.di[plane] = ddi;
            t.dj[plane] = ddj;
        }
        t.pair_plane[p] = plane;
    }

    // per-vertex incidence grouped by pair (pairs by number, springs by id)
    std::vector<int> ptr(N + 1, 0);
    for (int s = 0; s < S; s++) {
        ptr[springs[s].p1 + 1]++;
        ptr[springs[s].p2 + 1]++;
    }
    for (int i = 0; i < N; i++) ptr[i + 1] += ptr[i];
    std::vector<int> inc(2 * (size_t)S);
    {
        std::vector<int> fill(N, 0);
        for (int s = 0; s < S; s++) {
            int a = springs[s].p1, b = springs[s].p2;
            inc[ptr[a] + fill[a]++] = s;
            inc[ptr[b] + fill[b]++] = s;
        }
    }
    int wa = 0;
    for (int i = 0; i < N; i++) {
        auto b0 = inc.begin() + ptr[i], b1 = inc.begin() + ptr[i + 1];
        std::sort(b0, b1, [&](int x, int y) {
            int px = t.spring_pair[x], py = t.spring_pair[y];
            return px != py ? px < py : x < y;
        });
        wa = std::max(wa, ptr[i + 1] - ptr[i]);
    }
    if (wa > kMaxAssemblyWidth) {
        *err = "a vertex has more than " + std::to_string(kMaxAssemblyWidth) + " springs";
        return OC_ERR_UNSUPPORTED_TOPOLOGY;
    }
    t.wa = wa;
    t.ainc_spring.assign((size_t)wa * N, -1);
    t.ainc_meta.assign((size_t)wa * N, kEmptySlot);
    std::vector<int> npair(N, 0);
    int wp = 0;
    for (int i = 0; i < N; i++) {
        int cnt = ptr[i + 1] - ptr[i];
        int np = 0;
        for (int k = 0; k < cnt; k++) {
            int s = inc[ptr[i] + k];
            int p = t.spring_pair[s];
            bool last = (k + 1 == cnt) || t.spring_pair[inc[ptr[i] + k + 1]] != p;
            int other = springs[s].p1 == i ? springs[s].p2 : springs[s].p1;
            int ddi = (other % nx) - (i % nx), ddj = (other / nx) - (i / nx);
            uint16_t m = (uint16_t)((ddi + 8) | ((ddj + 8) << 4));
            if (springs[s].p2 == i) m |= 1u << 8;
            m |= (uint16_t)(t.pair_plane[p] << 9);
            if (last) m |= 1u << 13;
            if (t.pair_a[p] == i) m |= 1u << 14;
            t.ainc_spring[(size_t)k * N + i] = s;
            t.ainc_meta[(size_t)k * N + i] = m;
            if (last) np++;
        }
        npair[i] = np;
        wp = std::max(wp, np);
    }
    if (wp > kMaxPairWidth) {
        *err = "a vertex has more than " + std::to_string(kMaxPairWidth) + " neighbours";
        return OC_ERR_UNSUPPORTED_TOPOLOGY;
    }
    t.wp = wp;
    t.pair_meta.assign((size_t)wp * N, kEmptySlot);
    for (int i = 0; i < N; i++) {
        int cnt = ptr[i + 1] - ptr[i];
        int slot = 0;
        for (int k = 0; k < cnt; k++) {
            int s = inc[ptr[i] + k];
            int p = t.spring_pair[s];
            bool last = (k + 1 == cnt) || t.spring_pair[inc[ptr[i] + k + 1]] != p;
            if (!last) continue;
            int other = t.pair_a[p] == i ? t.pair_b[p] : t.pair_a[p];
            int ddi = (other % nx) - (i % nx), ddj = (other / nx) - (i / nx);
            uint16_t m = (uint16_t)((ddi + 8) | ((ddj + 8) << 4) | (t.pair_plane[p] << 8));
            if (t.pair_a[p] == i) m |= 1u << 12;
            t.pair_meta[(size_t)slot * N + i] = m;
            slot++;
        }
    }

    // stencil form: every vertex's pair slots at increasing kStencil positions
    t.stencil = nx >= 2 && wp <= kStencil;
    if (t.stencil) {
        t.pair_mask.assign(N, 0);
        for (int k = 0; k < kStencil; k++) {
            // plane of the pair whose upper vertex is at the owner's offset
            const int udi = stencil_owner(k) ? stencil_di(k) : -stencil_di(k);
            const int udj = stencil_owner(k) ? stencil_dj(k) : -stencil_dj(k);
            t.stencil_plane[k] = 0;
            for (int q = 0; q < t.n_planes; q++)
                if (t.di[q] == udi && t.dj[q] == udj) t.stencil_plane[k] = q;
        }
    }
    for (int i = 0; i < N && t.stencil; i++) {
        int last = -1;
        for (int slot = 0; slot < wp; slot++) {
            const uint16_t m = t.pair_meta[(size_t)slot * N + i];
            if (m == kEmptySlot) break;
            int pos = -1;
            for (int k = 0; k < kStencil; k++)
                if (slot_di(m) == stencil_di(k) && slot_dj(m) == stencil_dj(k)) pos = k;
            if (pos <= last) {
                t.stencil = false;
                break;
            }
            t.pair_mask[i] |= (uint16_t)(1u << pos);
            last = pos;
        }
    }
    if (!t.stencil) t.pair_mask.clear();

    // stencil assembly form: each pair at most two springs; the owner (the
    // pair's lower vertex) holds the constants of its pairs in owner planes
    t.stencil_asm = t.stencil;
    if (t.stencil_asm) {
        t.asm_mask.assign(N, 0);
        t.spring_cst.assign(S, -1);
        std::vector<int> count(P, 0);
        for (int s = 0; s < S && t.stencil_asm; s++)
            if (++count[t.spring_pair[s]] > 2) t.stencil_asm = false;
        std::vector<int> seen(P, 0);
        for (int s = 0; s < S && t.stencil_asm; s++) {  // springs in id order
            const int p = t.spring_pair[s];
            const int a = t.pair_a[p], b = t.pair_b[p];
            const int ddi = (b % nx) - (a % nx), ddj = (b / nx) - (a / nx);
            int o = -1;
            for (int q = 0; q < kOwned; q++)
                if (stencil_di(owned_k(q)) == ddi && stencil_dj(owned_k(q)) == ddj) o = q;
            if (o < 0) {
                t.stencil_asm = false;
                break;
            }
            const int m = seen[p]++;
            t.spring_cst[s] = (int32_t)(((size_t)(m * kOwned + o) * 3) * N + a);
            t.asm_mask[a] |= 1u << owned_k(o);
            t.asm_mask[b] |= 1u << owned_partner_k(o);
            if (m == 1) {
                t.asm_mask[a] |= 1u << (16 + owned_k(o));
                t.asm_mask[b] |= 1u << (16 + owned_partner_k(o));
            }
        }
        // the masks' presence bits are the SpMV's stencil mask
        for (int i = 0; i < N && t.stencil_asm; i++)
            if ((t.asm_mask[i] & 0xFFFu) != t.pair_mask[i]) t.stencil_asm = false;
    }
    if (!t.stencil_asm) {
        t.asm_mask.clear();
        t.spring_cst.clear();
    }
    return OC_OK;
}

}  // namespace oc
