// Persistent batched implicit step (gfx950, wave64).
//
// One launch steps every cloth of a batch.  Workgroups pull work items (packs
// of similar-sized cloths, largest first) from a queue; the threads of a
// workgroup are split evenly over the pack's cloths and each cloth's
// assembly, filtered CG and update run inside its workgroup, so no step needs
// a grid-wide barrier.  Cloths that meet their stopping test are masked out
// while the rest of the pack iterates.  The loop exits when the queue is
// drained, which every workgroup reaches.
//
// Per cloth the arithmetic is the single-cloth step's (same per-vertex
// gathers, pair order and expressions, optional Jacobi preconditioner).
// Spring constants come from slot-major copies and, when every cloth has
// the stencil form, the SpMV takes its neighbours from a 12-bit mask.
// Reductions: thread t of a cloth's T threads sums vertices t, t+T, ... in
// order, then the 64-lane wave tree, then the cloth's waves in order.  A
// thread's first kResidentPerThread vertices keep their CG vectors (d, r, q
// and the solution) in LDS through the solve; the rest stream from memory.
#include "oc_device.hpp"

namespace oc {

namespace {

using namespace dev;

// Sum over the T threads serving each cloth; valid where threadIdx.x % T == 0.
template <int NW>
__device__ inline float seg_sum(float t, int T, float* ws) {
    t = wave_sum(t);
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = t;
    __syncthreads();
    float p = 0.0f;
    if (threadIdx.x % T == 0) {
        const int w0 = threadIdx.x >> 6, nw = T >> 6;
        p = ws[w0];
        for (int k = 1; k < nw; k++) p = p + ws[w0 + k];
    }
    __syncthreads();
    return p;
}

// Jacobi preconditioner entry 1/diag(A), 1 where the diagonal is not
// positive and finite (as in the single-cloth kernels).
__device__ inline float jacobi_inv(float dg) {
    return (dg > 0.0f && isfinite(dg)) ? 1.0f / dg : 1.0f;
}

// CG start of vertex i from its assembled sums: b = h (f + h K v),
// r = S (b - A dv), the Jacobi entries and d = z = P^-1 r (z = r without the
// preconditioner); returns r.z.  keep (a vertex whose CG state stays in
// LDS): r, z and the solution's start go there, and r is not written to
// memory (d is: the neighbours' SpMV reads it).
__device__ inline float cg_start(const DevBatch& c, const StepScalars& s, int i, bool pinned, V3 acc, V3 kv,
                                 V3 adv, V3 dg, V3 dvi, V3* keep) {
    const int N = c.Ntot;
    const float h = s.h;
    V3 b;
    b.x = pinned ? 0.0f : h * (acc.x + h * kv.x);
    b.y = pinned ? 0.0f : h * (acc.y + h * kv.y);
    b.z = pinned ? 0.0f : h * (acc.z + h * kv.z);
    st3(c.b, N, i, b);
    const V3 r = pinned ? V3{0.0f, 0.0f, 0.0f} : sub3(b, adv);
    V3 z = r;
    if (s.flags & OC_FLAG_JACOBI) {
        const V3 pv = {jacobi_inv(dg.x), jacobi_inv(dg.y), jacobi_inv(dg.z)};
        st3(c.pinv, N, i, pv);
        z = {pv.x * r.x, pv.y * r.y, pv.z * r.z};
    }
    st3(c.d, N, i, z);
    if (keep) {
        keep[0] = r;
        keep[1] = z;
        keep[2] = dvi;
    } else {
        st3(c.r, N, i, r);
    }
    return dot3(r, z);
}

// Slot k of vertex i: spring id and meta, the other end's x, v (dv for a
// warm start) and the spring's constants (from the slot-major copies, so a
// wave reads them coalesced instead of gathering by spring id); sp < 0 past
// the last slot.
struct BSpringIn {
    int sp;
    uint16_t m;
    V3 x, v, dv;
    float L0, ks, kd;
};

__device__ inline BSpringIn fetch_bspring(const DevBatch& c, const StepScalars& s, int i, int nx, int k) {
    BSpringIn r{};
    const int N = c.Ntot;
    r.sp = c.ainc_s[(size_t)k * N + i];
    if (r.sp >= 0) {
        r.m = c.ainc_m[(size_t)k * N + i];
        const int j = i + slot_dj(r.m) * nx + slot_di(r.m);
        r.x = ld3(c.x, N, j);
        r.v = ld3(c.v, N, j);
        r.dv = s.warm ? ld3(c.dv, N, j) : V3{0.0f, 0.0f, 0.0f};
        const size_t plane = (size_t)c.wa * N;
        const size_t o = (size_t)k * N + i;
        r.L0 = c.slot_c[o];
        r.ks = c.slot_c[plane + o];
        r.kd = c.slot_c[2 * plane + o];
    }
    return r;
}

// Assembly of vertex i (global index) of a cloth with row length nx:
// pair blocks it owns, rhs b, CG start r = S(b - A dv), d = z = P^-1 r (z = r
// without the Jacobi preconditioner); returns r.z.
__device__ inline float assemble_vertex(const DevBatch& c, const StepScalars& s, int i, int nx,
                                       V3* keep = nullptr) {
    const int N = c.Ntot;
    const float h = s.h, hh = s.hh;
    const bool pinned = c.pin[i] != 0;
    V3 xi = ld3(c.x, N, i), vi = ld3(c.v, N, i);
    V3 dvi = s.warm ? ld3(c.dv, N, i) : V3{0.0f, 0.0f, 0.0f};
    if (!s.warm) st3(c.dv, N, i, dvi);
    V3 acc = {0.0f, 0.0f, 0.0f}, kv = {0.0f, 0.0f, 0.0f};
    if (!pinned) acc = {0.0f + s.mass * s.gx, 0.0f + s.mass * s.gy, 0.0f + s.mass * s.gz};
    if (!pinned && s.damp != 0.0f)
        acc = {acc.x + s.damp * vi.x, acc.y + s.damp * vi.y, acc.z + s.damp * vi.z};
    V3 adv = {s.md * dvi.x, s.md * dvi.y, s.md * dvi.z};
    V3 dg = {s.md, s.md, s.md};  // diag(A) = m - h c - sum of the pair blocks' diagonals
    float C[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    // Springs are evaluated from this vertex's side (x1 = x_i), which is
    // bitwise the negated p1-side force and K w; the next spring's data is
    // loaded before the current one is evaluated.
    BSpringIn nxt{};
    nxt.sp = -1;
    if (c.wa > 0) nxt = fetch_bspring(c, s, i, nx, 0);
    for (int k = 0; k < c.wa; k++) {
        const BSpringIn in = nxt;
        if (in.sp < 0) break;
        if (k + 1 < c.wa) nxt = fetch_bspring(c, s, i, nx, k + 1);
        const uint16_t m = in.m;
        V3 f;
        float K[6], P[6];
        spring_terms(xi, in.x, vi, in.v, in.L0, in.ks, in.kd, f, K, P);
#pragma unroll
        for (int q = 0; q < 6; q++) C[q] += h * (in.kd * P[q]) + hh * K[q];
        acc = add3(acc, f);
        kv = add3(kv, symmv(K, sub3(vi, in.v)));
        if ((m >> 13) & 1u) {
            if ((m >> 14) & 1u) {
                const int plane = (m >> 9) & 15;
#pragma unroll
                for (int q = 0; q < 6; q++) c.blocks[(size_t)(plane * 6 + q) * N + i] = C[q];
            }
            adv = add3(adv, symmv(C, sub3(in.dv, dvi)));
            dg = {dg.x - C[0], dg.y - C[1], dg.z - C[2]};
#pragma unroll
            for (int q = 0; q < 6; q++) C[q] = 0.0f;
        }
    }
    return cg_start(c, s, i, pinned, acc, kv, adv, dg, dvi, keep);
}

// Stencil form of assemble_vertex, for batches whose cloths all have the
// stencil assembly form: the vertex's pairs in stencil order from its mask
// (bit k: a pair at position k; bit 16 + k: that pair's second spring) and
// neighbour offsets from the position.  A pair's first spring takes its
// constants from the owner planes (c.cst) at the pair's lower vertex, so the
// partner's read hits the cache line its owner's thread loads; a second
// spring takes them from its assembly slot, which is the count of the
// vertex's springs at earlier positions, plus one.  No slot table is read
// and no load waits on another, so the next position's loads are issued
// before the current one is evaluated.  The pairs, springs and expressions
// are those of assemble_vertex, in the same order (the slot order is the
// stencil order on these topologies).
struct StencilIn {
    V3 x, v, dv;
    float L0, ks, kd;
};

__device__ inline StencilIn fetch_stencil(const DevBatch& c, const StepScalars& s, int i, int nx,
                                          int k, bool have) {
    StencilIn r{};
    if (have) {
        const int N = c.Ntot;
        const int j = i + stencil_dj(k) * nx + stencil_di(k);
        r.x = ld3(c.x, N, j);
        r.v = ld3(c.v, N, j);
        r.dv = s.warm ? ld3(c.dv, N, j) : V3{0.0f, 0.0f, 0.0f};
        const size_t o = (size_t)owned_index(k) * 3 * N + (stencil_owner(k) ? i : j);
        r.L0 = c.cst[o];
        r.ks = c.cst[o + N];
        r.kd = c.cst[o + 2 * (size_t)N];
    }
    return r;
}

__device__ inline float assemble_vertex_stencil(const DevBatch& c, const StepScalars& s, int i, int nx,
                                               V3* keep = nullptr) {
    const int N = c.Ntot;
    const float h = s.h, hh = s.hh;
    const bool pinned = c.pin[i] != 0;
    const uint32_t mi = c.asm_mask[i];
    V3 xi = ld3(c.x, N, i), vi = ld3(c.v, N, i);
    V3 dvi = s.warm ? ld3(c.dv, N, i) : V3{0.0f, 0.0f, 0.0f};
    if (!s.warm) st3(c.dv, N, i, dvi);
    V3 acc = {0.0f, 0.0f, 0.0f}, kv = {0.0f, 0.0f, 0.0f};
    if (!pinned) acc = {0.0f + s.mass * s.gx, 0.0f + s.mass * s.gy, 0.0f + s.mass * s.gz};
    if (!pinned && s.damp != 0.0f)
        acc = {acc.x + s.damp * vi.x, acc.y + s.damp * vi.y, acc.z + s.damp * vi.z};
    V3 adv = {s.md * dvi.x, s.md * dvi.y, s.md * dvi.z};
    V3 dg = {s.md, s.md, s.md};
    const size_t plane = (size_t)c.wa * N;
    int slot = 0;
    StencilIn nxt = fetch_stencil(c, s, i, nx, 0, mi & 1u);
#pragma unroll 1
    for (int k = 0; k < kStencil; k++) {
        const StencilIn in = nxt;
        const bool here = (mi >> k) & 1u;
        const bool second = (mi >> (16 + k)) & 1u;
        const int nsp = here ? (second ? 2 : 1) : 0;
        if (k + 1 < kStencil) nxt = fetch_stencil(c, s, i, nx, k + 1, (mi >> (k + 1)) & 1u);
        if (!here) continue;
        float C[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
        float L0 = in.L0, ks = in.ks, kd = in.kd;
        for (int m = 0; m < nsp; m++) {
            if (m == 1) {
                const size_t o = (size_t)(slot + 1) * N + i;
                L0 = c.slot_c[o];
                ks = c.slot_c[plane + o];
                kd = c.slot_c[2 * plane + o];
            }
            V3 f;
            float K[6], P[6];
            spring_terms(xi, in.x, vi, in.v, L0, ks, kd, f, K, P);
#pragma unroll
            for (int q = 0; q < 6; q++) C[q] += h * (kd * P[q]) + hh * K[q];
            acc = add3(acc, f);
            kv = add3(kv, symmv(K, sub3(vi, in.v)));
        }
        slot += nsp;
        if (stencil_owner(k)) {
            const int pl = c.stencil_plane[k];
#pragma unroll
            for (int q = 0; q < 6; q++) c.blocks[(size_t)(pl * 6 + q) * N + i] = C[q];
        }
        adv = add3(adv, symmv(C, sub3(in.dv, dvi)));
        dg = {dg.x - C[0], dg.y - C[1], dg.z - C[2]};
    }
    return cg_start(c, s, i, pinned, acc, kv, adv, dg, dvi, keep);
}

// z = P^-1 r of vertex i (r itself without the Jacobi preconditioner).
__device__ inline V3 precond(const DevBatch& c, const StepScalars& s, int i, V3 r) {
    if (!(s.flags & OC_FLAG_JACOBI)) return r;
    const V3 pv = ld3(c.pinv, c.Ntot, i);
    return {pv.x * r.x, pv.y * r.y, pv.z * r.z};
}

// Pair slot data for the SpMV: meta, the block (6 floats) and d of the
// other end.
struct PairIn {
    uint16_t m;
    float C[6];
    V3 dj;
};

__device__ inline PairIn fetch_pair(const DevBatch& c, int i, int nx, uint16_t m) {
    PairIn r{};
    r.m = m;
    if (m != kEmptySlot) {
        const int N = c.Ntot;
        const int j = i + slot_dj(m) * nx + slot_di(m);
        const int blk = ((m >> 12) & 1) ? i : j;
        const float* bp = c.blocks + (size_t)((m >> 8) & 15) * 6 * N + blk;
#pragma unroll
        for (int q = 0; q < 6; q++) r.C[q] = bp[(size_t)q * N];
        r.dj = ld3(c.d, N, j);
    }
    return r;
}

// Stencil form of spmv_vertex (every cloth's pairs in the kStencil order,
// one plane table): a 12-bit mask replaces the slot metas; the pairs are
// summed in the same order.
__device__ inline V3 spmv_stencil(const DevBatch& c, int i, int nx, V3 di, V3 acc) {
    const size_t N = c.Ntot;
    const uint32_t mask = c.pair_mask[i];
#pragma unroll
    for (int k = 0; k < kStencil; k++) {
        if (!((mask >> k) & 1u)) continue;
        const int j = i + stencil_dj(k) * nx + stencil_di(k);
        const int blk = stencil_owner(k) ? i : j;
        const float* bp = c.blocks + (size_t)c.stencil_plane[k] * 6 * N + blk;
        float C[6];
#pragma unroll
        for (int q = 0; q < 6; q++) C[q] = bp[(size_t)q * N];
        acc = add3(acc, symmv(C, sub3(ld3(c.d, (int)N, j), di)));
    }
    return acc;
}

// q_i = S (A d)_i into q; returns d_i . q_i.  Pair metas are loaded two
// slots ahead and block data one slot ahead of the slot being summed.
__device__ inline float spmv_core(const DevBatch& c, const StepScalars& s, int i, int nx, V3& q) {
    const int N = c.Ntot;
    const V3 di = ld3(c.d, N, i);
    V3 acc = scl3(s.md, di);
    if (c.stencil) {
        acc = spmv_stencil(c, i, nx, di, acc);
        q = c.pin[i] ? V3{0.0f, 0.0f, 0.0f} : acc;
        return dot3(di, q);
    }
    const uint16_t m0 = c.wp > 0 ? c.pair_m[i] : kEmptySlot;
    uint16_t m1 = c.wp > 1 ? c.pair_m[(size_t)N + i] : kEmptySlot;
    PairIn cur = fetch_pair(c, i, nx, m0);
    for (int sl = 0; sl < c.wp; sl++) {
        if (cur.m == kEmptySlot) break;
        const PairIn nxt = fetch_pair(c, i, nx, m1);
        m1 = (sl + 2 < c.wp) ? c.pair_m[(size_t)(sl + 2) * N + i] : kEmptySlot;
        acc = add3(acc, symmv(cur.C, sub3(cur.dj, di)));
        cur = nxt;
    }
    q = c.pin[i] ? V3{0.0f, 0.0f, 0.0f} : acc;
    return dot3(di, q);
}

// spmv_core with q_i to memory.
__device__ inline float spmv_vertex(const DevBatch& c, const StepScalars& s, int i, int nx) {
    V3 q;
    const float t = spmv_core(c, s, i, nx, q);
    st3(c.q, c.Ntot, i, q);
    return t;
}

// Adjoint start for vertex i: gvt = gv' + h gx', r = S gvt, d = P^-1 r (the
// solution buffer, c.dv, was zeroed by the cold assembly); returns r.(P^-1 r).
__device__ inline float adjoint_start_vertex(const DevBatch& c, const StepScalars& s,
                                             const BatchAdjBufs& a, int i, V3* keep = nullptr) {
    const int N = c.Ntot;
    const V3 gx = ld3(a.gx_out, N, i), gv = ld3(a.gv_out, N, i);
    const V3 g = {gv.x + s.h * gx.x, gv.y + s.h * gx.y, gv.z + s.h * gx.z};
    st3(a.gvt, N, i, g);
    const V3 r = c.pin[i] ? V3{0.0f, 0.0f, 0.0f} : g;
    const V3 z = precond(c, s, i, r);  // the preconditioner of the cold assembly
    st3(c.d, N, i, z);
    if (keep) {  // as in cg_start; the solve starts from mu = 0
        keep[0] = r;
        keep[1] = z;
        keep[2] = V3{0.0f, 0.0f, 0.0f};
    } else {
        st3(c.r, N, i, r);
    }
    return dot3(r, z);
}

// Spring-once assembly of the workgroup's cloths; every thread of the
// workgroup calls it (it has barriers).  A cloth's T threads sweep it in
// bands of T consecutive vertices.  In each band, direction by direction
// (the six owned offsets), every vertex evaluates the pair it owns in that
// direction (it is the pair's lower vertex) into LDS; after a barrier each
// vertex adds, in stencil order, the pair where it is the partner (the
// owner's force and K w negated: swapping a spring's ends negates them
// exactly and leaves the block term; the sums start at +0, where the sign
// of a zero term cannot matter) and its own pair.  Owners in earlier bands
// are read from a ring per direction that keeps that direction's latest
// results (ring_len >= the direction's offset, so an owner's entry is not
// overwritten before its partner's band reads it).  A pair's second spring
// is evaluated by each end itself.  The sums are assemble_vertex_stencil's,
// term by term in the same order.  Returns this thread's sum of r.z (ADJ:
// of the adjoint start's r.z) over its vertices in band order.
template <int NT, bool ADJ>
__device__ float assemble_once(const DevBatch& c, const StepScalars& s, const BatchAdjBufs& adj,
                               float* lds, int T, int g, int lt, bool act, int voff, int nx, int Ni,
                               int Nmax, int ringf) {
    const int N = c.Ntot, tid = threadIdx.x;
    const float h = s.h, hh = s.hh;
    const size_t plane = (size_t)c.wa * N;
    float* res = lds;  // res[q * NT + tid]: this direction's owner results
    float* ring = lds + 12 * NT + g * ringf;
    float t = 0.0f;
    const int nb = (Nmax + T - 1) / T;
    for (int band = 0; band < nb; band++) {
        const int v0 = band * T, vv = v0 + lt;
        const bool here = act && vv < Ni;
        const int i = voff + vv;
        uint32_t mi = 0u;
        bool pinned = true;
        V3 xi = {0.0f, 0.0f, 0.0f}, vi = xi, dvi = xi, acc = xi, kv = xi, adv = xi, dg = xi;
        if (here) {
            mi = c.asm_mask[i];
            pinned = c.pin[i] != 0;
            xi = ld3(c.x, N, i);
            vi = ld3(c.v, N, i);
            if (s.warm) dvi = ld3(c.dv, N, i);
            else st3(c.dv, N, i, dvi);
            if (!pinned) acc = {0.0f + s.mass * s.gx, 0.0f + s.mass * s.gy, 0.0f + s.mass * s.gz};
            if (!pinned && s.damp != 0.0f)
                acc = {acc.x + s.damp * vi.x, acc.y + s.damp * vi.y, acc.z + s.damp * vi.z};
            adv = {s.md * dvi.x, s.md * dvi.y, s.md * dvi.z};
            dg = {s.md, s.md, s.md};
        }
        float C[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
        auto add_spring = [&](const SpringOut& r) {
#pragma unroll
            for (int q = 0; q < 6; q++) C[q] += r.c[q];
            acc = {acc.x + r.f.x, acc.y + r.f.y, acc.z + r.f.z};
            kv = {kv.x + r.kw.x, kv.y + r.kw.y, kv.z + r.kw.z};
        };
        // assembly slot of the first spring at stencil position k
        auto slot_of = [&](int k) {
            const uint32_t below = (1u << k) - 1u;
            return __popc(mi & 0xFFFu & below) + __popc((mi >> 16) & below);
        };
        // the pair's second spring, from this vertex's side
        auto add_second = [&](int k, int j) {
            const size_t o = (size_t)(slot_of(k) + 1) * N + i;
            add_spring(eval_spring(xi, ld3(c.x, N, j), vi, ld3(c.v, N, j), c.slot_c[o], c.slot_c[plane + o],
                                   c.slot_c[2 * plane + o], h, hh));
        };
        // the pair is complete: its block (when owned), A dv and diag(A) terms
        auto finish_pair = [&](int k, int j) {
            if (stencil_owner(k)) {
                const int pl = c.stencil_plane[k];
#pragma unroll
                for (int q = 0; q < 6; q++) c.blocks[(size_t)(pl * 6 + q) * N + i] = C[q];
            }
            const V3 dvj = s.warm ? ld3(c.dv, N, j) : V3{0.0f, 0.0f, 0.0f};
            adv = add3(adv, symmv(C, sub3(dvj, dvi)));
            dg = {dg.x - C[0], dg.y - C[1], dg.z - C[2]};
#pragma unroll
            for (int q = 0; q < 6; q++) C[q] = 0.0f;
        };
        auto load = [&](const float* src, int stride) {
            SpringOut e;
            e.f = {src[0], src[stride], src[2 * stride]};
            e.kw = {src[3 * stride], src[4 * stride], src[5 * stride]};
#pragma unroll
            for (int q = 0; q < 6; q++) e.c[q] = src[(6 + q) * stride];
            return e;
        };
        SpringOut held{};  // this vertex's k = 7 pair, added after k = 5 and k = 6
#pragma unroll 1
        for (int o = 0; o < kOwned; o++) {
            const int ko = owned_k(o), kp = owned_partner_k(o), off = own_offset(o, nx);
            const int R = ring_len(o, nx);
            float* rg = ring + 12 * ring_base(o, nx);
            // ---- evaluate this vertex's pair in direction o
            if (here && ((mi >> ko) & 1u)) {
                const int j = i + off;
                const size_t so = (size_t)slot_of(ko) * N + i;
                const SpringOut e = eval_spring(xi, ld3(c.x, N, j), vi, ld3(c.v, N, j), c.slot_c[so],
                                                c.slot_c[plane + so], c.slot_c[2 * plane + so], h, hh);
                res[0 * NT + tid] = e.f.x;
                res[1 * NT + tid] = e.f.y;
                res[2 * NT + tid] = e.f.z;
                res[3 * NT + tid] = e.kw.x;
                res[4 * NT + tid] = e.kw.y;
                res[5 * NT + tid] = e.kw.z;
#pragma unroll
                for (int q = 0; q < 6; q++) res[(6 + q) * NT + tid] = e.c[q];
            }
            __syncthreads();
            // ---- accumulate, in this vertex's stencil order
            if (here && ((mi >> kp) & 1u)) {  // partner of the pair owned by vertex vv - off
                const int jv = vv - off;
                SpringOut e = jv >= v0 ? load(res + (tid - off), NT) : load(rg + (jv & (R - 1)), R);
                e.f = {0.0f - e.f.x, 0.0f - e.f.y, 0.0f - e.f.z};
                e.kw = {0.0f - e.kw.x, 0.0f - e.kw.y, 0.0f - e.kw.z};
                add_spring(e);
                if ((mi >> (16 + kp)) & 1u) add_second(kp, i - off);
                finish_pair(kp, i - off);
            }
            if (o == 2) {
                if (here && ((mi >> ko) & 1u)) held = load(res + tid, NT);  // k = 7 follows k = 5, 6
            } else if (here) {
#pragma unroll
                for (int u = 0; u < 2; u++) {  // k = 6, then the held k = 7; else just ko
                    if (u == 1 && o != 3) break;
                    const int oo = o == 3 ? (u == 0 ? 3 : 2) : o;
                    const int kk = owned_k(oo);
                    if (!((mi >> kk) & 1u)) continue;
                    add_spring(u == 0 ? load(res + tid, NT) : held);
                    const int j = i + own_offset(oo, nx);
                    if ((mi >> (16 + kk)) & 1u) add_second(kk, j);
                    finish_pair(kk, j);
                }
            }
            __syncthreads();
            // ---- keep this direction's results of the band's last R owners
            if (here && ((mi >> ko) & 1u) && vv >= v0 + T - R) {
                const int e = vv & (R - 1);
#pragma unroll
                for (int q = 0; q < 12; q++) rg[q * R + e] = res[q * NT + tid];
            }
        }
        if (here) {
            V3 b;
            b.x = pinned ? 0.0f : h * (acc.x + h * kv.x);
            b.y = pinned ? 0.0f : h * (acc.y + h * kv.y);
            b.z = pinned ? 0.0f : h * (acc.z + h * kv.z);
            st3(c.b, N, i, b);
            V3 r = pinned ? V3{0.0f, 0.0f, 0.0f} : sub3(b, adv);
            st3(c.r, N, i, r);
            float rr;
            if (s.flags & OC_FLAG_JACOBI) {
                const V3 pv = {jacobi_inv(dg.x), jacobi_inv(dg.y), jacobi_inv(dg.z)};
                st3(c.pinv, N, i, pv);
                const V3 z = {pv.x * r.x, pv.y * r.y, pv.z * r.z};
                st3(c.d, N, i, z);
                rr = dot3(r, z);
            } else {
                st3(c.d, N, i, r);
                rr = dot3(r, r);
            }
            if constexpr (ADJ)
                t += adjoint_start_vertex(c, s, adj, i);
            else
                t += rr;
        }
    }
    return t;
}

// Adjoint gather for vertex i (same as the single-cloth k_adj_gather):
// c.x, c.v the pre-step state, c.dv the adjoint solve mu, a.dvsol the step's
// solution.
__device__ inline void adjoint_gather_vertex(const DevBatch& c, const StepScalars& s,
                                             const BatchAdjBufs& a, int i, int nx) {
    const int N = c.Ntot;
    const V3 xi = ld3(c.x, N, i), vi = ld3(c.v, N, i), mi = ld3(c.dv, N, i),
             pi = ld3(a.dvsol, N, i);
    const V3 wi = add3(vi, pi);
    V3 gx = ld3(a.gx_out, N, i), gv = ld3(a.gvt, N, i);
    if (s.damp != 0.0f && !c.pin[i])  // the damping force c v_i: dJ/dv_i += h c mu_i
        gv = {gv.x + s.h * (s.damp * mi.x), gv.y + s.h * (s.damp * mi.y), gv.z + s.h * (s.damp * mi.z)};
    for (int k = 0; k < c.wa; k++) {
        const int sp = c.ainc_s[(size_t)k * N + i];
        if (sp < 0) break;
        const uint16_t m = c.ainc_m[(size_t)k * N + i];
        const int j = i + slot_dj(m) * nx + slot_di(m);
        const V3 xj = ld3(c.x, N, j), vj = ld3(c.v, N, j), mj = ld3(c.dv, N, j),
                 pj = ld3(a.dvsol, N, j);
        const V3 wj = add3(vj, pj);
        const bool second = (m >> 8) & 1u;
        const size_t plane = (size_t)c.wa * N, o = (size_t)k * N + i;
        const float L0 = c.slot_c[o], ks = c.slot_c[plane + o], kd = c.slot_c[2 * plane + o];
        float ge[3], gvv[3], gk;
        if (second)
            spring_adjoint(xj, xi, vj, vi, mj, mi, wj, wi, pj, pi, L0, ks, kd, s.h,
                           s.hh, ge, gvv, gk);
        else
            spring_adjoint(xi, xj, vi, vj, mi, mj, wi, wj, pi, pj, L0, ks, kd, s.h,
                           s.hh, ge, gvv, gk);
        if (second) {
            gx = {gx.x - ge[0], gx.y - ge[1], gx.z - ge[2]};
            gv = {gv.x - gvv[0], gv.y - gvv[1], gv.z - gvv[2]};
        } else {
            gx = {gx.x + ge[0], gx.y + ge[1], gx.z + ge[2]};
            gv = {gv.x + gvv[0], gv.y + gvv[1], gv.z + gvv[2]};
            a.gks[sp] = a.gks[sp] + gk;
        }
    }
    st3(a.gx_in, N, i, gx);
    st3(a.gv_in, N, i, gv);
}

// ADJ = false: forward step.  ADJ = true: one step's adjoint at the state in
// c.x/c.v (assembly without warm start, CG for mu in c.dv, gather).
// ONCE: the spring-once assembly (c.asm_once), a separate instantiation so
// the default kernel carries none of its registers.
template <int NT, bool ADJ, bool ONCE>
__global__ __launch_bounds__(NT, 4) void k_batch_step(DevBatch c, StepScalars s, BatchAdjBufs adj) {
    constexpr int NW = NT / 64;
    constexpr int RV = ONCE ? 0 : kResidentPerThread;
    __shared__ float ws[NW];
    __shared__ float sd[NW], sd_old[NW], sd0[NW], ssig[NW];
    __shared__ int sdone[NW], sit[NW], sst[NW];
    __shared__ int sitem;
    extern __shared__ float dyn_lds[];  // spring-once assembly: lds_floats; else the CG's resident state
    const int tid = threadIdx.x;
    const int N = c.Ntot;
    for (;;) {
        if (tid == 0) sitem = atomicAdd(c.queue, 1);
        __syncthreads();
        const int item = sitem;
        __syncthreads();
        if (item >= c.n_items) return;
        const int T = c.item_T[item], G = c.item_count[item];
        const int g = tid / T, lt = tid % T;
        const bool act = g < G;
        int ins = 0, voff = 0, nx = 1, Ni = 0;
        if (act) {
            ins = c.order[c.item_inst[item] + g];
            voff = c.inst_voff[ins];
            nx = c.inst_nx[ins];
            Ni = c.inst_n[ins];
        }

        // ---- this thread's first RV vertices keep d, r, q and the solution
        // in LDS (slot u of the thread) for the whole solve: only d goes to
        // memory each iteration (the neighbours' SpMV reads it there).  Same
        // arithmetic in the same order as the vertices kept in memory.
        constexpr int RP = RV * NT;  // floats per LDS plane: d, r, solution, q (xyz each)
        float* const rs = dyn_lds;
        auto rld = [&](int pl, int u) {
            const int o = u * NT + tid;
            return V3{rs[(3 * pl) * RP + o], rs[(3 * pl + 1) * RP + o], rs[(3 * pl + 2) * RP + o]};
        };
        auto rst = [&](int pl, int u, V3 v) {
            const int o = u * NT + tid;
            rs[(3 * pl) * RP + o] = v.x;
            rs[(3 * pl + 1) * RP + o] = v.y;
            rs[(3 * pl + 2) * RP + o] = v.z;
        };

        // ---- assembly and CG start
        float t = 0.0f;
        if constexpr (ONCE) {
            // the item's longest sweep and widest rings
            int Nmax = 0, nxm = 1;
            for (int gg = 0; gg < G; gg++) {
                const int in = c.order[c.item_inst[item] + gg];
                Nmax = max(Nmax, c.inst_n[in]);
                nxm = max(nxm, c.inst_nx[in]);
            }
            t = assemble_once<NT, ADJ>(c, s, adj, dyn_lds, T, g, lt, act, voff, nx, Ni, Nmax,
                                       ring_floats(nxm));
        } else {
            // the thread's first RV vertices keep their CG state in LDS
            int u = 0;
            for (int vv = lt; vv < Ni; vv += T, u++) {
                V3 keep[3];
                V3* const kp = u < RV ? keep : nullptr;
                const float rr = c.stencil_asm ? assemble_vertex_stencil(c, s, voff + vv, nx, ADJ ? nullptr : kp)
                                               : assemble_vertex(c, s, voff + vv, nx, ADJ ? nullptr : kp);
                if constexpr (ADJ)
                    t += adjoint_start_vertex(c, s, adj, voff + vv, kp);
                else
                    t += rr;
                if (kp) {
                    rst(0, u, keep[1]);
                    rst(1, u, keep[0]);
                    rst(2, u, keep[2]);
                }
            }
        }
        float tot = seg_sum<NW>(t, T, ws);
        if (act && lt == 0) {
            sd[g] = tot;
            sd0[g] = tot;
            sd_old[g] = 0.0f;
            sdone[g] = 0;
            sit[g] = 0;
            sst[g] = 0;
        }
        __syncthreads();


        // ---- filtered CG, converged cloths masked
        for (int k = 0; k < s.max_it; k++) {
            if (act && lt == 0 && !sdone[g]) {
                const float thresh = s.tol_mode == 1 ? s.tol2 * sd0[g] : s.tol2;
                if (s.tol_mode != 2 && !(sd[g] > thresh)) {
                    sdone[g] = 1;
                    sit[g] = k;
                }
            }
            __syncthreads();
            bool all = true;
            for (int gg = 0; gg < G; gg++) all = all && sdone[gg];
            if (all) break;
            const bool live = act && !sdone[g];
            if (live && k > 0) {
                const float beta = sd_old[g] != 0.0f ? sd[g] / sd_old[g] : 0.0f;
#pragma unroll 1
                for (int u = 0; u < RV; u++) {
                    const int vv = lt + u * T;
                    if (vv < Ni) {
                        const int i = voff + vv;
                        const V3 z = precond(c, s, i, rld(1, u)), dp = rld(0, u);
                        const V3 dn = {z.x + beta * dp.x, z.y + beta * dp.y, z.z + beta * dp.z};
                        rst(0, u, dn);
                        st3(c.d, N, i, dn);
                    }
                }
                for (int vv = lt + RV * T; vv < Ni; vv += T) {
                    const int i = voff + vv;
                    V3 z = precond(c, s, i, ld3(c.r, N, i)), dp = ld3(c.d, N, i);
                    st3(c.d, N, i, V3{z.x + beta * dp.x, z.y + beta * dp.y, z.z + beta * dp.z});
                }
            }
            __syncthreads();
            t = 0.0f;
            if (live) {
#pragma unroll 1
                for (int u = 0; u < RV; u++) {
                    const int vv = lt + u * T;
                    if (vv < Ni) {
                        V3 q;
                        t += spmv_core(c, s, voff + vv, nx, q);
                        rst(3, u, q);
                    }
                }
                for (int vv = lt + RV * T; vv < Ni; vv += T) t += spmv_vertex(c, s, voff + vv, nx);
            }
            tot = seg_sum<NW>(t, T, ws);
            if (act && lt == 0) ssig[g] = tot;
            __syncthreads();
            bool bd = false;
            float alpha = 0.0f;
            if (live) {
                const float sigma = ssig[g];
                if (sigma < 0.0f || !isfinite(sigma))
                    bd = true;
                else
                    alpha = sigma == 0.0f ? 0.0f : sd[g] / sigma;
            }
            __syncthreads();
            if (live && bd && lt == 0) {
                sdone[g] = 1;
                sit[g] = k;
                sst[g] |= OC_SOLVE_BREAKDOWN;
            }
            t = 0.0f;
            if (live && !bd) {
#pragma unroll 1
                for (int u = 0; u < RV; u++) {
                    const int vv = lt + u * T;
                    if (vv < Ni) {
                        const int i = voff + vv;
                        const V3 d = rld(0, u), q = rld(3, u);
                        V3 x = rld(2, u), r = rld(1, u);
                        x = {x.x + alpha * d.x, x.y + alpha * d.y, x.z + alpha * d.z};
                        r = {r.x - alpha * q.x, r.y - alpha * q.y, r.z - alpha * q.z};
                        rst(2, u, x);
                        rst(1, u, r);
                        t += dot3(r, precond(c, s, i, r));
                    }
                }
                for (int vv = lt + RV * T; vv < Ni; vv += T) {
                    const int i = voff + vv;
                    V3 d = ld3(c.d, N, i), q = ld3(c.q, N, i), x = ld3(c.dv, N, i),
                       r = ld3(c.r, N, i);
                    x = {x.x + alpha * d.x, x.y + alpha * d.y, x.z + alpha * d.z};
                    r = {r.x - alpha * q.x, r.y - alpha * q.y, r.z - alpha * q.z};
                    st3(c.dv, N, i, x);
                    st3(c.r, N, i, r);
                    t += dot3(r, precond(c, s, i, r));
                }
            }
            tot = seg_sum<NW>(t, T, ws);
            if (live && !bd && lt == 0) {
                sd_old[g] = sd[g];
                sd[g] = tot;
            }
            __syncthreads();
        }

        // ---- the resident vertices' solution back to memory (the residual
        // is not read after the solve); the forward update below reads it
        // from LDS
        if constexpr (RV > 0) {
            if (act) {
#pragma unroll 1
                for (int u = 0; u < RV; u++) {
                    const int vv = lt + u * T;
                    if (vv < Ni) st3(c.dv, N, voff + vv, rld(2, u));
                }
            }
        }

        // ---- status and update
        if (act && lt == 0) {
            const int it = sdone[g] ? sit[g] : s.max_it;
            const float delta = sd[g];
            const float thresh = s.tol_mode == 1 ? s.tol2 * sd0[g] : s.tol2;
            int st = sst[g];
            if (s.tol_mode != 2 && delta > thresh && !(st & OC_SOLVE_BREAKDOWN))
                st |= OC_SOLVE_NOT_CONVERGED;
            if (!isfinite(delta)) st |= OC_SOLVE_NONFINITE;
            sit[g] = it;
            sst[g] = st;
        }
        __syncthreads();
        if constexpr (ADJ) {
            for (int vv = lt; vv < Ni; vv += T) adjoint_gather_vertex(c, s, adj, voff + vv, nx);
        } else {
            bool bad = false;
            for (int vv = lt; vv < Ni; vv += T) {
                const int i = voff + vv;
                V3 v = ld3(c.v, N, i), dv = ld3(c.dv, N, i), x = ld3(c.x, N, i);
                v = add3(v, dv);
                x = {x.x + s.h * v.x, x.y + s.h * v.y, x.z + s.h * v.z};
                if ((s.flags & OC_FLAG_GROUND_CLAMP) && x.y < 0.0f) x.y = 0.0f;
                st3(c.v, N, i, v);
                st3(c.x, N, i, x);
                bad = bad || !(isfinite(x.x) && isfinite(x.y) && isfinite(x.z) && isfinite(v.x) &&
                               isfinite(v.y) && isfinite(v.z));
            }
            if (bad) atomicOr(&sst[g], OC_SOLVE_NONFINITE);
        }
        __syncthreads();
        if (act && lt == 0) {
            c.inst_iters[ins] = sit[g];
            c.inst_status[ins] = sst[g];
            c.inst_sticky[ins] |= sst[g];  // one workgroup per cloth and launch
        }
        __syncthreads();
    }
}

// Owner-plane constants from the slot-major copies: the first spring of the
// pair vertex i owns in direction o sits at its assembly slot (the count of
// the vertex's springs at earlier stencil positions).
__global__ __launch_bounds__(256) void k_owner_constants(DevBatch c) {
    const int N = c.Ntot;
    const size_t plane = (size_t)c.wa * N;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < N; i += gridDim.x * 256) {
        const uint32_t mi = c.asm_mask[i];
#pragma unroll
        for (int o = 0; o < kOwned; o++) {
            const int k = owned_k(o);
            const uint32_t below = (1u << k) - 1u;
            const bool own = (mi >> k) & 1u;
            // (a vertex with this pair has a spring at that slot, so slot < wa)
            const int slot = min(__popc(mi & 0xFFFu & below) + __popc((mi >> 16) & below), c.wa - 1);
#pragma unroll
            for (int q = 0; q < 3; q++) {
                float v = 0.0f;
                if (own) v = c.slot_c[q * plane + (size_t)slot * N + i];
                c.cst[(size_t)(o * 3 + q) * N + i] = v;
            }
        }
    }
}

__global__ void k_batch_reset(int* queue) {
    if (threadIdx.x == 0) *queue = 0;
}

__global__ __launch_bounds__(256) void k_slot_constants(DevBatch c, int q0, int nq) {
    const size_t n = (size_t)c.wa * c.Ntot;
    const float* src[3] = {c.L0, c.ks, c.kd};
    for (size_t o = (size_t)blockIdx.x * 256 + threadIdx.x; o < n; o += (size_t)gridDim.x * 256) {
        const int sp = c.ainc_s[o];
        for (int q = q0; q < q0 + nq; q++) c.slot_c[q * n + o] = sp >= 0 ? src[q][sp] : 0.0f;
    }
}

}  // namespace

hipError_t launch_batch_slot_constants(const DevBatch& c, int q0, int nq, hipStream_t st) {
    if (q0 < 0 || nq < 0 || q0 + nq > 3) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_slot_constants, dim3(4096), dim3(256), 0, st, c, q0, nq);
    return hipGetLastError();
}

hipError_t launch_batch_owner_constants(const DevBatch& c, hipStream_t st) {
    if (!c.stencil_asm) return hipSuccess;
    hipLaunchKernelGGL(k_owner_constants, dim3(4096), dim3(256), 0, st, c);
    return hipGetLastError();
}

hipError_t launch_batch_reset(const DevBatch& c, hipStream_t st) {
    hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(64), 0, st, c.queue);
    return hipGetLastError();
}

template <bool ADJ>
static hipError_t launch_batch(const DevBatch& c, const StepScalars& s, const BatchAdjBufs& a,
                               int block_threads, int grid, hipStream_t st) {
    if (c.asm_once && (c.lds_floats < 12 * block_threads || c.lds_floats > kBatchLdsFloats))
        return hipErrorInvalidValue;
    const size_t lds = sizeof(float) * (size_t)(c.asm_once ? c.lds_floats : batch_resident_floats(block_threads));
    switch (block_threads * 2 + (c.asm_once ? 1 : 0)) {
        case 512:
            hipLaunchKernelGGL((k_batch_step<256, ADJ, false>), dim3(grid), dim3(256), lds, st, c, s, a);
            break;
        case 513:
            hipLaunchKernelGGL((k_batch_step<256, ADJ, true>), dim3(grid), dim3(256), lds, st, c, s, a);
            break;
        case 1024:
            hipLaunchKernelGGL((k_batch_step<512, ADJ, false>), dim3(grid), dim3(512), lds, st, c, s, a);
            break;
        case 1025:
            hipLaunchKernelGGL((k_batch_step<512, ADJ, true>), dim3(grid), dim3(512), lds, st, c, s, a);
            break;
        case 2048:
            hipLaunchKernelGGL((k_batch_step<1024, ADJ, false>), dim3(grid), dim3(1024), lds, st, c, s, a);
            break;
        case 2049:
            hipLaunchKernelGGL((k_batch_step<1024, ADJ, true>), dim3(grid), dim3(1024), lds, st, c, s, a);
            break;
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_batch_step(const DevBatch& c, const StepScalars& s, int block_threads, int grid,
                             hipStream_t st) {
    return launch_batch<false>(c, s, BatchAdjBufs{}, block_threads, grid, st);
}

hipError_t launch_batch_adjoint(const DevBatch& c, const StepScalars& s, const BatchAdjBufs& a,
                                int block_threads, int grid, hipStream_t st) {
    return launch_batch<true>(c, s, a, block_threads, grid, st);
}

}  // namespace oc
