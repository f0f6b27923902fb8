// HIP kernels of the implicit mass-spring step (gfx950, wave64).
//
// One step (OpenCloth_ImplicitEuler/.../main.cpp:774-779 without the
// post-step passes) is
//   assemble  : per-vertex gather of spring forces and 3x3 Jacobian blocks,
//               rhs b = h(f + h df/dx v), pair blocks C, CG start r = S(b - A dv)
//   cg_spmv   : d = r + beta d on the tile and its halo in LDS, q = S A d,
//               partial d.q                                    (x iterations)
//   cg_axpy   : dv += alpha d, r -= alpha q, partial r.r        (x iterations)
//   cg_finish : iteration count and status
//   update    : v += dv, x += h v
//
// Layout: vertex vectors are SoA planes (x[c*N + i]).  Pair blocks live in
// planes: the block of the pair (a, b = a + offset_p) is stored at
// blocks[(p*6 + comp)*N + a], so neighbouring threads read neighbouring
// addresses for every stencil direction.  The diagonal block
// A_ii = m I - sum_j C_ij is never stored: (A d)_i = m d_i + sum_j C_ij (d_j - d_i).
//
// Arithmetic follows the reference's GLM expression order with no FMA
// contraction (built with -ffp-contract=off), so per-spring forces and blocks
// equal the CPU oracle bit for bit.  Reductions use one fixed tree: a 64-lane
// shuffle tree per wave, the waves of a tile in order, then a lane-strided
// pass over tile partials followed by the same lane tree.  Partials are
// indexed by tile, so the block-to-tile mapping does not change any sum.
//
// Every CG kernel issues all loads that do not depend on the previous
// reduction (its own vectors, pair blocks, the partials) before waiting, so
// the cross-tile reduction overlaps the memory traffic.
#include <hip/hip_ext.h>

#include <algorithm>

#include "oc_device.hpp"

namespace oc {

namespace {

constexpr int H = kHalo;
using namespace dev;


// Jacobi preconditioner entry 1/diag(A); a non-positive or non-finite
// diagonal (A can be indefinite under strong compression) falls back to 1.
__device__ inline float jacobi_inv(float dg) {
    return (dg > 0.0f && isfinite(dg)) ? 1.0f / dg : 1.0f;
}

// ---------------------------------------------------------------------------
// forces (ComputeForces, IE:645-682): gather in spring order, pins get zero
// ---------------------------------------------------------------------------
template <int TX, int TY>
__global__ __launch_bounds__(TX* TY) void k_forces(DevCloth c, StepScalars s) {
    const int lx = threadIdx.x % TX, ly = threadIdx.x / TX;
    const int t = tile_of_block(blockIdx.x, c.nt);
    const int bx = t % c.ntx, by = t / c.ntx;
    const int gi = bx * TX + lx, gj = by * TY + ly;
    if (gi >= c.nx || gj >= c.ny) return;
    const int N = c.N, i = gj * c.nx + gi;
    V3 acc = {0.0f, 0.0f, 0.0f};
    if (!c.pin[i]) {
        V3 xi = ld3(c.x, N, i), vi = ld3(c.v, N, i);
        acc = {0.0f + s.mass * s.gx, 0.0f + s.mass * s.gy, 0.0f + s.mass * s.gz};
        if (s.damp != 0.0f) acc = {acc.x + s.damp * vi.x, acc.y + s.damp * vi.y, acc.z + s.damp * vi.z};
        for (int k = 0; k < c.wa; k++) {
            const int sp = c.ainc_s[k * N + i];
            if (sp < 0) break;
            const uint16_t m = c.ainc_m[k * N + i];
            const int j = i + slot_dj(m) * c.nx + slot_di(m);
            V3 xj = ld3(c.x, N, j), vj = ld3(c.v, N, j);
            const bool second = (m >> 8) & 1u;
            V3 f = second ? spring_force(xj, xi, vj, vi, c.L0[sp], c.ks[sp], c.kd[sp])
                          : spring_force(xi, xj, vi, vj, c.L0[sp], c.ks[sp], c.kd[sp]);
            if (second)
                acc = {acc.x - f.x, acc.y - f.y, acc.z - f.z};
            else
                acc = {acc.x + f.x, acc.y + f.y, acc.z + f.z};
        }
    }
    st3(c.f, N, i, acc);
}

// Iteration count and status of the solve, by all threads of one block
// (uniform control flow: partial_sum synchronises the block).  Bits are
// OR-ed in, so update blocks may flag non-finite values concurrently.
__device__ inline void cg_finish_block(const DevCloth& c, const StepScalars& s, float* bc) {
    const int done = c.ctl[0];
    int it;
    float delta;
    if (done) {
        it = c.ctl[1];
        delta = c.scal[it];
    } else {
        it = s.max_it;
        delta = partial_sum(c.del_part + (size_t)it * c.nt, c.nt, bc);
    }
    // r0.r0: written by the first spmv phase, or equal to delta without iterations
    const float delta0 = s.max_it > 0 ? c.scal[0] : delta;
    if (threadIdx.x == 0) {
        const float thresh = s.tol_mode == 1 ? s.tol2 * delta0 : s.tol2;
        const int st = c.ctl[2];
        int extra = 0;
        if (s.tol_mode != 2 && delta > thresh && !(st & OC_SOLVE_BREAKDOWN))
            extra |= OC_SOLVE_NOT_CONVERGED;
        if (!isfinite(delta)) extra |= OC_SOLVE_NONFINITE;
        c.ctl[1] = it;
        atomicOr(c.ctl + 2, extra);
        atomicOr(c.ctl + 3, st | extra);
    }
}

// ---------------------------------------------------------------------------
// assembly + CG start
// ---------------------------------------------------------------------------
// Constants of one incident spring (gathered by spring id).
struct SpringC {
    float L0, ks, kd;
};

template <int WA>
__device__ inline SpringC fetch_constants(const DevCloth& c, const int (&sps)[WA], int k) {
    SpringC r{0.0f, 0.0f, 0.0f};
    const int sp = sps[k];
    if (sp >= 0) {
        r.L0 = ldo(c.L0, (uint32_t)sp);
        r.ks = ldo(c.ks, (uint32_t)sp);
        r.kd = ldo(c.kd, (uint32_t)sp);
    }
    return r;
}

// Tile + halo of x, v and (warm start) dv in LDS.
template <int TX, int TY>
struct AssembleShared {
    static constexpr int WX = TX + 2 * H, WY = TY + 2 * H;
    float wx[3][WY][WX];
    float wv[3][WY][WX];
    float wdv[3][WY][WX];
    float ws[TX * TY / 64];
    float bc;
};

// UPD: the window's x and v are the previous step's state before its
// update; apply the update (k_update's arithmetic) on the fly and write the
// tile's updated x, v to xo, vo (other blocks still read x and v).  Only
// with a warm start: a cold start writes dv = 0 here, which other blocks
// would still be reading.
template <int TX, int TY, int WA, bool UPD>
__global__ __launch_bounds__(TX* TY) void k_assemble(DevCloth c, StepScalars s) {
    constexpr int NT = TX * TY;
    constexpr int WX = TX + 2 * H, WY = TY + 2 * H;
    constexpr int NWIN = (WX * WY + NT - 1) / NT;
    __shared__ AssembleShared<TX, TY> sh;
    const int lx = threadIdx.x % TX, ly = threadIdx.x / TX;
    const int tile = tile_of_block(blockIdx.x, c.nt);
    const int bx = tile % c.ntx, by = tile / c.ntx;
    const int gi = bx * TX + lx, gj = by * TY + ly;
    const bool own = gi < c.nx && gj < c.ny;
    const int N = c.N, i = gj * c.nx + gi;
    if (blockIdx.x == 0) {
        // the previous solve's iteration count and status bits go into the
        // sticky word before the control block is cleared (with UPD that
        // solve had no update launch to record them)
        if (UPD) cg_finish_block(c, s, &sh.bc);
        if (threadIdx.x == 0) {
            c.ctl[0] = 0;
            c.ctl[1] = 0;
            c.ctl[2] = 0;
        }
    }
    // own slot tables and the first springs' constants, then the window
    int sps[WA];
    uint16_t ms[WA];
#pragma unroll
    for (int k = 0; k < WA; k++) {
        sps[k] = (own && k < c.wa) ? ldo(c.ainc_s, (uint32_t)(k * N + i)) : -1;
        ms[k] = (own && k < c.wa) ? ldo(c.ainc_m, (uint32_t)(k * N + i)) : kEmptySlot;
    }
    constexpr int kPrefetch = 4;
    SpringC pre[kPrefetch];
#pragma unroll
    for (int k = 0; k < kPrefetch; k++) pre[k] = fetch_constants(c, sps, k);
    const int x0 = bx * TX - H, y0 = by * TY - H;
#pragma unroll
    for (int n = 0; n < NWIN; n++) {
        const int idx = threadIdx.x + n * NT;
        if (idx < WX * WY) {
            const int wx = idx % WX, wy = idx / WX;
            const int gx = x0 + wx, gy = y0 + wy;
            V3 xw = {0.0f, 0.0f, 0.0f}, vw = xw, dw = xw;
            if (gx >= 0 && gx < c.nx && gy >= 0 && gy < c.ny) {
                const int g = gy * c.nx + gx;
                xw = ld3o(c.x, N, g);
                vw = ld3o(c.v, N, g);
                if (UPD) {
                    const V3 dp = ld3o(c.dv, N, g);
                    vw = {vw.x + dp.x, vw.y + dp.y, vw.z + dp.z};
                    xw = {xw.x + s.h * vw.x, xw.y + s.h * vw.y, xw.z + s.h * vw.z};
                    if ((s.flags & OC_FLAG_GROUND_CLAMP) && xw.y < 0.0f) xw.y = 0.0f;
                    if (s.warm) dw = dp;
                } else if (s.warm) {
                    dw = ld3o(c.dv, N, g);
                }
            }
            sh.wx[0][wy][wx] = xw.x;
            sh.wx[1][wy][wx] = xw.y;
            sh.wx[2][wy][wx] = xw.z;
            sh.wv[0][wy][wx] = vw.x;
            sh.wv[1][wy][wx] = vw.y;
            sh.wv[2][wy][wx] = vw.z;
            sh.wdv[0][wy][wx] = dw.x;
            sh.wdv[1][wy][wx] = dw.y;
            sh.wdv[2][wy][wx] = dw.z;
        }
    }
    __syncthreads();

    float t = 0.0f;
    if (own) {
        const float h = s.h, hh = s.hh;
        const bool pinned = ldo(c.pin, (uint32_t)i) != 0;
        const int cy = ly + H, cx = lx + H;
        const V3 xi = {sh.wx[0][cy][cx], sh.wx[1][cy][cx], sh.wx[2][cy][cx]};
        const V3 vi = {sh.wv[0][cy][cx], sh.wv[1][cy][cx], sh.wv[2][cy][cx]};
        const V3 dvi = {sh.wdv[0][cy][cx], sh.wdv[1][cy][cx], sh.wdv[2][cy][cx]};
        if (!s.warm) st3o(c.dv, N, i, dvi);
        if (UPD) {
            st3o(c.xo, N, i, xi);
            st3o(c.vo, N, i, vi);
            // k_update's check of the updated state; the bit belongs to the
            // previous step, whose status is already in the sticky word
            if (!(isfinite(xi.x) && isfinite(xi.y) && isfinite(xi.z) && isfinite(vi.x) &&
                  isfinite(vi.y) && isfinite(vi.z)))
                atomicOr(c.ctl + 3, OC_SOLVE_NONFINITE);
        }
        V3 acc = {0.0f, 0.0f, 0.0f}, kv = {0.0f, 0.0f, 0.0f};
        if (!pinned) acc = {0.0f + s.mass * s.gx, 0.0f + s.mass * s.gy, 0.0f + s.mass * s.gz};
        if (!pinned && s.damp != 0.0f)
            acc = {acc.x + s.damp * vi.x, acc.y + s.damp * vi.y, acc.z + s.damp * vi.z};
        V3 adv = {s.md * dvi.x, s.md * dvi.y, s.md * dvi.z};
        V3 dg = {s.md, s.md, s.md};  // diag(A) = m - h c - sum of the pair blocks' diagonals
        float C[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
        // Every spring is evaluated from this vertex's side (x1 = x_i): when
        // i is the spring's p2 that negates dp, f and K(v_i - v_j) exactly
        // and leaves K and P unchanged, so adding them equals subtracting
        // the p1-side values bit for bit.  Neighbours come from the LDS
        // window; spring constants are loaded kPrefetch springs ahead.
#pragma unroll
        for (int k = 0; k < WA; k++) {
            const int sp = sps[k];
            if (sp < 0) break;
            const uint16_t m = ms[k];
            const SpringC in = pre[0];
#pragma unroll
            for (int q = 0; q + 1 < kPrefetch; q++) pre[q] = pre[q + 1];
            if (k + kPrefetch < WA) pre[kPrefetch - 1] = fetch_constants(c, sps, k + kPrefetch);
            const int wy = cy + slot_dj(m), wx = cx + slot_di(m);
            const V3 xj = {sh.wx[0][wy][wx], sh.wx[1][wy][wx], sh.wx[2][wy][wx]};
            const V3 vj = {sh.wv[0][wy][wx], sh.wv[1][wy][wx], sh.wv[2][wy][wx]};
            V3 f;
            float K[6], P[6];
            spring_terms(xi, xj, vi, vj, in.L0, in.ks, in.kd, f, K, P);
#pragma unroll
            for (int q = 0; q < 6; q++) C[q] += h * (in.kd * P[q]) + hh * K[q];
            const V3 kw = symmv(K, V3{vi.x - vj.x, vi.y - vj.y, vi.z - vj.z});
            acc = {acc.x + f.x, acc.y + f.y, acc.z + f.z};
            kv = {kv.x + kw.x, kv.y + kw.y, kv.z + kw.z};
            if ((m >> 13) & 1u) {  // last spring of this pair
                if ((m >> 14) & 1u) {
                    const int plane = (m >> 9) & 15;
#pragma unroll
                    for (int q = 0; q < 6; q++)
                        sto(c.blocks + (size_t)q * N, (uint32_t)(plane * 6 * N + i), C[q]);
                }
                const V3 dvj = {sh.wdv[0][wy][wx], sh.wdv[1][wy][wx], sh.wdv[2][wy][wx]};
                V3 wd = {dvj.x - dvi.x, dvj.y - dvi.y, dvj.z - dvi.z};
                V3 a = symmv(C, wd);
                adv = {adv.x + a.x, adv.y + a.y, adv.z + a.z};
                dg = {dg.x - C[0], dg.y - C[1], dg.z - C[2]};
#pragma unroll
                for (int q = 0; q < 6; q++) C[q] = 0.0f;
            }
        }
        V3 b;
        b.x = pinned ? 0.0f : h * (acc.x + h * kv.x);
        b.y = pinned ? 0.0f : h * (acc.y + h * kv.y);
        b.z = pinned ? 0.0f : h * (acc.z + h * kv.z);
        st3o(c.b, N, i, b);
        V3 r = pinned ? V3{0.0f, 0.0f, 0.0f} : V3{b.x - adv.x, b.y - adv.y, b.z - adv.z};
        st3o(c.r, N, i, r);
        if (s.flags & OC_FLAG_JACOBI) {
            const V3 pv = {jacobi_inv(dg.x), jacobi_inv(dg.y), jacobi_inv(dg.z)};
            st3o(c.pinv, N, i, pv);
            const V3 z = {pv.x * r.x, pv.y * r.y, pv.z * r.z};
            st3o(c.s, N, i, z);
            st3o(c.d0, N, i, z);
            t = dot3(r, z);
        } else {
            st3o(c.d0, N, i, r);
            t = dot3(r, r);
        }
    }
    float p = tile_sum<NT>(t, sh.ws);
    if (threadIdx.x == 0) c.del_part[tile] = p;
}

// ---------------------------------------------------------------------------
// assembly, stencil assembly form: every spring evaluated once per tile
// ---------------------------------------------------------------------------
// A pair's owner is its lower-index vertex; the partner sits at one of six
// offsets (dx, dy) from it (owned direction o, owned_k).  The assembly visits
// the owned directions one after another.  In each, the block evaluates the
// springs owned by its tile's vertices and by the owners just outside the
// tile whose partner is inside (the direction's "strip"), once each, from
// the owner's side; the results (force, K (v_owner - v_partner) and the
// pair-block term) go to LDS.  Then every vertex adds, in its own stencil
// order, the pair it is partner of (the owner's results negated) and the
// pair it owns.  Negating the owner's force and K w equals evaluating from
// the partner's side (dp, f and K w change sign exactly, K and P are even in
// dp) up to the sign of zero, and these sums start at +0, where a signed
// zero cannot change a sum; so the results are the per-end kernel's bits.
// A second spring of a pair (the reference's duplicated last bends) is
// evaluated by each end itself.

template <int TX, int TY>
struct AssembleStShared {
    static constexpr int WX = TX + 2 * H, WY = TY + 2 * H;
    // columns 0 .. TX TY - 1: this direction's pairs owned in the tile;
    // then the strips of every direction, all evaluated before the first
    static constexpr int NR = TX * TY + strip_total<TX, TY>();
    float wx[3][WY][WX];
    float wv[3][WY][WX];
    float wdv[3][WY][WX];
    uint32_t wm[WY][WX];
    float res[12][NR];  // per evaluated pair: f, K w, block term
    float ws[TX * TY / 64];
    float bc;
};

// UPD as in k_assemble.
template <int TX, int TY, bool UPD>
__global__ __launch_bounds__(TX* TY, 4) void k_assemble_st(DevCloth c, StepScalars s) {
    constexpr int NT = TX * TY;
    using SH = AssembleStShared<TX, TY>;
    constexpr int WX = SH::WX, WY = SH::WY;
    constexpr int NWIN = (WX * WY + NT - 1) / NT;
    static_assert(strip_total<TX, TY>() <= NT, "the strips must fit one round of the block");
    __shared__ SH sh;
    const int lx = threadIdx.x % TX, ly = threadIdx.x / TX;
    const int tile = tile_of_block(blockIdx.x, c.nt);
    const int bx = tile % c.ntx, by = tile / c.ntx;
    const int gi = bx * TX + lx, gj = by * TY + ly;
    const bool own = gi < c.nx && gj < c.ny;
    const int N = c.N, nx = c.nx, i = gj * nx + gi;
    if (blockIdx.x == 0) {
        if (UPD) cg_finish_block(c, s, &sh.bc);
        if (threadIdx.x == 0) {
            c.ctl[0] = 0;
            c.ctl[1] = 0;
            c.ctl[2] = 0;
        }
    }
    auto own_cst = [&](int m, int o, int q, int g) {
        return ldo(c.cst, (uint32_t)(((m * kOwned + o) * 3 + q) * N + g));
    };
    // this thread's strip column (owner outside the tile, partner inside):
    // the strips of all directions are evaluated in one pass, packed into the
    // first waves, so no phase below waits on a wave with two evaluations
    int so = 0, swx = 0, swy = 0;
    bool sin = false;
    float sL0 = 0.0f, sks = 0.0f, skd = 0.0f;
    if ((int)threadIdx.x < strip_total<TX, TY>()) {
        int ss;
        strip_of_column<TX, TY>(threadIdx.x, so, ss);
        strip_owner<TX, TY>(so, ss, swx, swy);
        const int gx = bx * TX + swx, gy = by * TY + swy;
        if (gx >= 0 && gx < nx && gy >= 0 && gy < c.ny) {
            const int g = gy * nx + gx;
            sin = true;
            sL0 = own_cst(0, so, 0, g);
            sks = own_cst(0, so, 1, g);
            skd = own_cst(0, so, 2, g);
        }
    }
    // the window: x, v, dv and the stencil masks of the tile and its halo
    const int x0 = bx * TX - H, y0 = by * TY - H;
#pragma unroll
    for (int n = 0; n < NWIN; n++) {
        const int idx = threadIdx.x + n * NT;
        if (idx < WX * WY) {
            const int wx = idx % WX, wy = idx / WX;
            const int gx = x0 + wx, gy = y0 + wy;
            V3 xw = {0.0f, 0.0f, 0.0f}, vw = xw, dw = xw;
            uint32_t mw = 0;
            if (gx >= 0 && gx < nx && gy >= 0 && gy < c.ny) {
                const int g = gy * nx + gx;
                mw = ldo(c.asm_mask, (uint32_t)g);
                xw = ld3o(c.x, N, g);
                vw = ld3o(c.v, N, g);
                if (UPD) {
                    const V3 dp = ld3o(c.dv, N, g);
                    vw = {vw.x + dp.x, vw.y + dp.y, vw.z + dp.z};
                    xw = {xw.x + s.h * vw.x, xw.y + s.h * vw.y, xw.z + s.h * vw.z};
                    if ((s.flags & OC_FLAG_GROUND_CLAMP) && xw.y < 0.0f) xw.y = 0.0f;
                    dw = dp;
                } else if (s.warm) {
                    dw = ld3o(c.dv, N, g);
                }
            }
            sh.wx[0][wy][wx] = xw.x;
            sh.wx[1][wy][wx] = xw.y;
            sh.wx[2][wy][wx] = xw.z;
            sh.wv[0][wy][wx] = vw.x;
            sh.wv[1][wy][wx] = vw.y;
            sh.wv[2][wy][wx] = vw.z;
            sh.wdv[0][wy][wx] = dw.x;
            sh.wdv[1][wy][wx] = dw.y;
            sh.wdv[2][wy][wx] = dw.z;
            sh.wm[wy][wx] = mw;
        }
    }
    // the first direction's constants of this vertex's pair
    float oL0 = 0.0f, oks = 0.0f, okd = 0.0f;
    auto prefetch = [&](int o) {
        if (own) {
            oL0 = own_cst(0, o, 0, i);
            oks = own_cst(0, o, 1, i);
            okd = own_cst(0, o, 2, i);
        }
    };
    prefetch(0);
    __syncthreads();

    const float h = s.h, hh = s.hh;
    const int cx = lx + H, cy = ly + H;
    const V3 xi = {sh.wx[0][cy][cx], sh.wx[1][cy][cx], sh.wx[2][cy][cx]};
    const V3 vi = {sh.wv[0][cy][cx], sh.wv[1][cy][cx], sh.wv[2][cy][cx]};
    const V3 dvi = {sh.wdv[0][cy][cx], sh.wdv[1][cy][cx], sh.wdv[2][cy][cx]};
    const uint32_t mi = own ? sh.wm[cy][cx] : 0u;
    bool pinned = true;
    V3 acc = {0.0f, 0.0f, 0.0f}, kv = {0.0f, 0.0f, 0.0f};
    V3 adv = {0.0f, 0.0f, 0.0f}, dg = {0.0f, 0.0f, 0.0f};
    if (own) {
        pinned = ldo(c.pin, (uint32_t)i) != 0;
        if (!s.warm) st3o(c.dv, N, i, dvi);
        if (UPD) {
            st3o(c.xo, N, i, xi);
            st3o(c.vo, N, i, vi);
            if (!(isfinite(xi.x) && isfinite(xi.y) && isfinite(xi.z) && isfinite(vi.x) &&
                  isfinite(vi.y) && isfinite(vi.z)))
                atomicOr(c.ctl + 3, OC_SOLVE_NONFINITE);
        }
        if (!pinned) acc = {0.0f + s.mass * s.gx, 0.0f + s.mass * s.gy, 0.0f + s.mass * s.gz};
        if (!pinned && s.damp != 0.0f)
            acc = {acc.x + s.damp * vi.x, acc.y + s.damp * vi.y, acc.z + s.damp * vi.z};
        adv = {s.md * dvi.x, s.md * dvi.y, s.md * dvi.z};
        dg = {s.md, s.md, s.md};
    }
    float C[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    auto window3 = [&](float (&w)[3][WY][WX], int px, int py) {
        return V3{w[0][py][px], w[1][py][px], w[2][py][px]};
    };
    // Add one spring's terms (from this vertex's side) to the pair's sums.
    auto add_spring = [&](const SpringOut& r) {
#pragma unroll
        for (int q = 0; q < 6; q++) C[q] += r.c[q];
        acc = {acc.x + r.f.x, acc.y + r.f.y, acc.z + r.f.z};
        kv = {kv.x + r.kw.x, kv.y + r.kw.y, kv.z + r.kw.z};
    };
    // The second spring of the pair with the neighbour at window (px, py),
    // from this vertex's side (constants of the owner g at multiplicity 1).
    auto add_second = [&](int o, int g, int px, int py) {
        const SpringOut r = eval_spring(xi, window3(sh.wx, px, py), vi, window3(sh.wv, px, py),
                                        own_cst(1, o, 0, g), own_cst(1, o, 1, g), own_cst(1, o, 2, g),
                                        h, hh);
        add_spring(r);
    };
    // The pair is complete: its block (when owned), A dv and diag(A) terms.
    auto finish_pair = [&](int px, int py, int k_store) {
        if (k_store >= 0) {
            const int plane = c.stencil_plane[k_store];
#pragma unroll
            for (int q = 0; q < 6; q++)
                sto(c.blocks + (size_t)q * N, (uint32_t)(plane * 6 * N + i), C[q]);
        }
        const V3 dvj = window3(sh.wdv, px, py);
        const V3 a = symmv(C, V3{dvj.x - dvi.x, dvj.y - dvi.y, dvj.z - dvi.z});
        adv = {adv.x + a.x, adv.y + a.y, adv.z + a.z};
        dg = {dg.x - C[0], dg.y - C[1], dg.z - C[2]};
#pragma unroll
        for (int q = 0; q < 6; q++) C[q] = 0.0f;
    };

    // The result stored at LDS column r (the owner's side).
    auto stored = [&](int r) {
        SpringOut e;
        e.f = {sh.res[0][r], sh.res[1][r], sh.res[2][r]};
        e.kw = {sh.res[3][r], sh.res[4][r], sh.res[5][r]};
#pragma unroll
        for (int q = 0; q < 6; q++) e.c[q] = sh.res[6 + q][r];
        return e;
    };
    // ---- the strips: column NT + t holds strip column t's pair (the owner's side)
    if (sin && ((sh.wm[swy + H][swx + H] >> owned_k(so)) & 1u)) {
        const int px = swx + H, py = swy + H, dx = own_dx(so), dy = own_dy(so);
        const SpringOut e = eval_spring(window3(sh.wx, px, py), window3(sh.wx, px + dx, py + dy),
                                        window3(sh.wv, px, py), window3(sh.wv, px + dx, py + dy),
                                        sL0, sks, skd, h, hh);
        const int r = NT + threadIdx.x;
        sh.res[0][r] = e.f.x;
        sh.res[1][r] = e.f.y;
        sh.res[2][r] = e.f.z;
        sh.res[3][r] = e.kw.x;
        sh.res[4][r] = e.kw.y;
        sh.res[5][r] = e.kw.z;
#pragma unroll
        for (int q = 0; q < 6; q++) sh.res[6 + q][r] = e.c[q];
    }
    SpringOut held{};  // this vertex's k = 7 pair, added after k = 5 and k = 6
#pragma unroll 1
    for (int o = 0; o < kOwned; o++) {
        const int dx = own_dx(o), dy = own_dy(o);
        const int ko = owned_k(o), kp = owned_partner_k(o);
        // ---- evaluate this vertex's pair in direction o
        if (own && ((mi >> ko) & 1u)) {
            const SpringOut e = eval_spring(xi, window3(sh.wx, cx + dx, cy + dy), vi,
                                            window3(sh.wv, cx + dx, cy + dy), oL0, oks, okd, h, hh);
            const int r = threadIdx.x;
            sh.res[0][r] = e.f.x;
            sh.res[1][r] = e.f.y;
            sh.res[2][r] = e.f.z;
            sh.res[3][r] = e.kw.x;
            sh.res[4][r] = e.kw.y;
            sh.res[5][r] = e.kw.z;
#pragma unroll
            for (int q = 0; q < 6; q++) sh.res[6 + q][r] = e.c[q];
        }
        if (o + 1 < kOwned) prefetch(o + 1);
        __syncthreads();
        // ---- accumulate, in this vertex's stencil order
        if (own && ((mi >> kp) & 1u)) {  // partner of the pair owned at -(dx, dy)
            const int wx = lx - dx, wy = ly - dy;
            const int r = (wx >= 0 && wx < TX && wy >= 0 && wy < TY)
                              ? wy * TX + wx
                              : NT + strip_base<TX, TY>(o) + strip_index<TX, TY>(o, wx, wy);
            SpringOut e = stored(r);
            e.f = {0.0f - e.f.x, 0.0f - e.f.y, 0.0f - e.f.z};
            e.kw = {0.0f - e.kw.x, 0.0f - e.kw.y, 0.0f - e.kw.z};
            add_spring(e);
            if ((mi >> (16 + kp)) & 1u) add_second(o, i - dy * nx - dx, cx - dx, cy - dy);
            finish_pair(cx - dx, cy - dy, -1);
        }
        if (o == 2) {
            if (own && ((mi >> ko) & 1u)) held = stored(threadIdx.x);  // k = 7 follows k = 5, 6
        } else if (own) {
#pragma unroll
            for (int t = 0; t < 2; t++) {  // k = 6, then the held k = 7; else just ko
                const int oo = o == 3 ? (t == 0 ? 3 : 2) : o;
                if (t == 1 && o != 3) break;
                const int kk = owned_k(oo), ddx = own_dx(oo), ddy = own_dy(oo);
                if (!((mi >> kk) & 1u)) continue;
                add_spring(t == 0 ? stored(threadIdx.x) : held);
                if ((mi >> (16 + kk)) & 1u) add_second(oo, i, cx + ddx, cy + ddy);
                finish_pair(cx + ddx, cy + ddy, kk);
            }
        }
        __syncthreads();
    }

    float t = 0.0f;
    if (own) {
        V3 b;
        b.x = pinned ? 0.0f : h * (acc.x + h * kv.x);
        b.y = pinned ? 0.0f : h * (acc.y + h * kv.y);
        b.z = pinned ? 0.0f : h * (acc.z + h * kv.z);
        st3o(c.b, N, i, b);
        V3 r = pinned ? V3{0.0f, 0.0f, 0.0f} : V3{b.x - adv.x, b.y - adv.y, b.z - adv.z};
        st3o(c.r, N, i, r);
        if (s.flags & OC_FLAG_JACOBI) {
            const V3 pv = {jacobi_inv(dg.x), jacobi_inv(dg.y), jacobi_inv(dg.z)};
            st3o(c.pinv, N, i, pv);
            const V3 z = {pv.x * r.x, pv.y * r.y, pv.z * r.z};
            st3o(c.s, N, i, z);
            st3o(c.d0, N, i, z);
            t = dot3(r, z);
        } else {
            st3o(c.d0, N, i, r);
            t = dot3(r, r);
        }
    }
    float p = tile_sum<NT>(t, sh.ws);
    if (threadIdx.x == 0) c.del_part[tile] = p;
}

// ---------------------------------------------------------------------------
// CG iteration k, part 1: direction update on the tile + halo, q = S A d
// ---------------------------------------------------------------------------
template <int TX, int TY>
struct SpmvShared {
    static constexpr int WX = TX + 2 * H, WY = TY + 2 * H;
    float wd[3][WY][WX];
    float ws[TX * TY / 64];
    float bc;
};

// One block's share of CG iteration k, part 1.  read_done: take the stop
// flag of earlier launches from c.ctl[0] (multi-kernel form).  Returns true
// when the solve has stopped (every block decides alike).  STENCIL: the
// cloth's pairs follow the kStencil order (c.stencil), so a 12-bit mask per
// vertex replaces the WP slot metas and every slot's offset, owner and plane
// are constants; the sums run in the same pair order either way.
template <int TX, int TY, int WP, bool STENCIL = false>
__device__ inline bool spmv_phase(const DevCloth& c, const StepScalars& s, int k, bool read_done,
                                  SpmvShared<TX, TY>& sh, int vb) {
    constexpr int NT = TX * TY;
    constexpr int WX = TX + 2 * H, WY = TY + 2 * H;
    constexpr int NWIN = (WX * WY + NT - 1) / NT;
    const int nt = c.nt, N = c.N;
    const int tile = tile_of_block(vb, nt);
    const int bx = tile % c.ntx, by = tile / c.ntx;
    const int lx = threadIdx.x % TX, ly = threadIdx.x / TX;
    const int gi = bx * TX + lx, gj = by * TY + ly;
    const bool own = gi < c.nx && gj < c.ny;
    const int i = gj * c.nx + gi;

    // ---- loads that do not depend on beta (the stencil mask first: the
    // block loads wait for it)
    uint32_t mask = 0;
    if constexpr (STENCIL) mask = own ? ldo(c.pair_mask, (uint32_t)i) : 0u;
    const int done = read_done ? c.ctl[0] : 0;
    const float delta0 = c.scal[0];
    const float delta_old = k > 0 ? c.scal[k - 1] : 0.0f;
    float pv[kPartialRegs];
    if (threadIdx.x < 64) partial_load(c.del_part + (size_t)k * nt, nt, pv);
    const float* dprev = (k & 1) ? c.d0 : c.d1;
    float* dcur = (k & 1) ? c.d1 : c.d0;
    const float* zsrc = (s.flags & OC_FLAG_JACOBI) ? c.s : c.r;  // d = z + beta d
    const int x0 = bx * TX - H, y0 = by * TY - H;
    V3 wr[NWIN], wp[NWIN];
#pragma unroll
    for (int n = 0; n < NWIN; n++) {
        const int idx = threadIdx.x + n * NT;
        const int gx = x0 + idx % WX, gy = y0 + idx / WX;
        wr[n] = wp[n] = V3{0.0f, 0.0f, 0.0f};
        if (idx < WX * WY && gx >= 0 && gx < c.nx && gy >= 0 && gy < c.ny) {
            const int g = gy * c.nx + gx;
            if (k == 0) {
                wr[n] = ld3o(c.d0, N, g);
            } else {
                wr[n] = ld3o(zsrc, N, g);
                wp[n] = ld3o(dprev, N, g);
            }
        }
    }
    constexpr int NS = STENCIL ? kStencil : WP;
    uint16_t m[NS];
    float C[NS][6];
    if constexpr (STENCIL) {
#pragma unroll
        for (int k = 0; k < kStencil; k++) {
            const int di = stencil_di(k), dj = stencil_dj(k);
            const uint32_t blk = stencil_owner(k) ? (uint32_t)i : (uint32_t)(i + dj * c.nx + di);
            const uint32_t off = blk + (uint32_t)c.stencil_plane[k] * 6u * (uint32_t)N;
            if ((mask >> k) & 1u) {
#pragma unroll
                for (int q = 0; q < 6; q++) C[k][q] = ldo(c.blocks, off + (uint32_t)(q * N));
            } else {
#pragma unroll
                for (int q = 0; q < 6; q++) C[k][q] = 0.0f;
            }
        }
    } else {
#pragma unroll
        for (int sl = 0; sl < WP; sl++) {
            m[sl] = (own && sl < c.wp) ? ldo(c.pair_m, (uint32_t)(sl * N + i)) : kEmptySlot;
            if (m[sl] != kEmptySlot) {
                const int j = i + slot_dj(m[sl]) * c.nx + slot_di(m[sl]);
                const int blk = ((m[sl] >> 12) & 1) ? i : j;
                const uint32_t b0 = (uint32_t)((m[sl] >> 8) & 15) * 6u * (uint32_t)N + (uint32_t)blk;
#pragma unroll
                for (int q = 0; q < 6; q++) C[sl][q] = ldo(c.blocks + (size_t)q * N, b0);
            } else {
#pragma unroll
                for (int q = 0; q < 6; q++) C[sl][q] = 0.0f;
            }
        }
    }
    const bool pinned = own ? ldo(c.pin, (uint32_t)i) != 0 : true;

    // ---- r.r of this iteration, stopping test, beta
    if (threadIdx.x < 64) {
        const float sum = partial_finish(c.del_part + (size_t)k * nt, nt, pv);
        if (threadIdx.x == 0) sh.bc = sum;
    }
    __syncthreads();
    if (done) return true;
    const float delta = sh.bc;
    const float thresh = s.tol_mode == 1 ? s.tol2 * (k == 0 ? delta : delta0) : s.tol2;
    const bool stop = s.tol_mode != 2 && !(delta > thresh);
    if (vb == 0 && threadIdx.x == 0) {
        c.scal[k] = delta;
        if (stop) {
            c.ctl[0] = 1;
            c.ctl[1] = k;
        }
    }
    if (stop) return true;
    // beta = 0 once r.r has underflowed to 0 (fixed iteration counts)
    const float beta = (k > 0 && delta_old != 0.0f) ? delta / delta_old : 0.0f;

    // ---- d = r + beta d on tile + halo into LDS; the tile's own d to memory
#pragma unroll
    for (int n = 0; n < NWIN; n++) {
        const int idx = threadIdx.x + n * NT;
        if (idx < WX * WY) {
            const int wx = idx % WX, wy = idx / WX;
            V3 d = wr[n];
            if (k > 0) {
                d = {wr[n].x + beta * wp[n].x, wr[n].y + beta * wp[n].y, wr[n].z + beta * wp[n].z};
                const int gx = x0 + wx, gy = y0 + wy;
                if (wx >= H && wx < H + TX && wy >= H && wy < H + TY && gx < c.nx && gy < c.ny)
                    st3o(dcur, N, gy * c.nx + gx, d);
            }
            sh.wd[0][wy][wx] = d.x;
            sh.wd[1][wy][wx] = d.y;
            sh.wd[2][wy][wx] = d.z;
        }
    }
    __syncthreads();

    // ---- q = S A d
    float t = 0.0f;
    if (own) {
        const V3 di = {sh.wd[0][ly + H][lx + H], sh.wd[1][ly + H][lx + H], sh.wd[2][ly + H][lx + H]};
        V3 acc = {s.md * di.x, s.md * di.y, s.md * di.z};
        if constexpr (STENCIL) {
#pragma unroll
            for (int k = 0; k < kStencil; k++) {
                if (!((mask >> k) & 1u)) continue;
                const int wy = ly + H + stencil_dj(k), wx = lx + H + stencil_di(k);
                V3 w = {sh.wd[0][wy][wx] - di.x, sh.wd[1][wy][wx] - di.y, sh.wd[2][wy][wx] - di.z};
                V3 a = symmv(C[k], w);
                acc = {acc.x + a.x, acc.y + a.y, acc.z + a.z};
            }
        } else {
#pragma unroll
            for (int sl = 0; sl < WP; sl++) {
                if (m[sl] == kEmptySlot) break;
                const int wy = ly + H + slot_dj(m[sl]), wx = lx + H + slot_di(m[sl]);
                V3 w = {sh.wd[0][wy][wx] - di.x, sh.wd[1][wy][wx] - di.y, sh.wd[2][wy][wx] - di.z};
                V3 a = symmv(C[sl], w);
                acc = {acc.x + a.x, acc.y + a.y, acc.z + a.z};
            }
        }
        V3 q = pinned ? V3{0.0f, 0.0f, 0.0f} : acc;
        st3o(c.q, N, i, q);
        t = dot3(di, q);
    }
    float p = tile_sum<NT>(t, sh.ws);
    if (threadIdx.x == 0) c.sig_part[(size_t)k * nt + tile] = p;
    return false;
}

template <int TX, int TY, int WP, bool STENCIL>
__global__ __launch_bounds__(TX* TY) void k_cg_spmv(DevCloth c, StepScalars s, int k) {
    __shared__ SpmvShared<TX, TY> sh;
    spmv_phase<TX, TY, WP, STENCIL>(c, s, k, true, sh, blockIdx.x);
}

// ---------------------------------------------------------------------------
// single-reduction CG (OC_FLAG_CG_SINGLE_REDUCTION): one launch per iteration
// ---------------------------------------------------------------------------
// The Chronopoulos-Gear recurrence (s = 1): with u = P^-1 r and w = S A u the
// two dots gamma = r.u and delta = w.u of an iteration come from the same
// launch, so each iteration needs one reduction point, and one launch:
//   beta = gamma_k / gamma_{k-1},  alpha = gamma_k / (delta_k - beta gamma_k / alpha_{k-1})
//   p = u + beta p,  s = w + beta s  (s = S A p),  dv += alpha p,  r -= alpha s,
//   u = P^-1 r,  w = S A u,  partials of r.u and w.u.
// r, s and w are read on the tile's halo by neighbouring blocks while the
// launch writes their next values, so each has two buffers used by iteration
// parity; p and dv are only read on the tile itself.  The start (w0 = S A u0,
// delta_0) is the first launch of the standard SpMV, whose q is w0.  The
// arithmetic and guards are the oracle's cg_core_cg1, step for step.
template <int TX, int TY>
struct Cg1Shared {
    static constexpr int WX = TX + 2 * H, WY = TY + 2 * H;
    float wu[3][WY][WX];  // u_{k+1} on the tile and halo
    float rt[3][TY][TX];  // r_{k+1} of the tile
    float ws[TX * TY / 64];
    float ws2[TX * TY / 64];
    float bc, bc2;
};

// Buffers of iteration k: r_k, w_k, s_{k-1} are read, r_{k+1}, w_{k+1}, s_k
// written (parity k & 1).
struct Cg1Bufs {
    const float* r_in;
    float* r_out;
    const float* w_in;
    float* w_out;
    const float* s_in;
    float* s_out;
};
__device__ inline Cg1Bufs cg1_bufs(const DevCloth& c, int k) {
    Cg1Bufs b;
    const bool odd = (k & 1) != 0;
    b.r_in = odd ? c.r2 : c.r;
    b.r_out = odd ? c.r : c.r2;
    b.w_in = odd ? c.q2 : c.q;
    b.w_out = odd ? c.q : c.q2;
    b.s_in = odd ? c.s1 : c.s2;
    b.s_out = odd ? c.s2 : c.s1;
    return b;
}

template <int TX, int TY, int WP, bool STENCIL>
__global__ __launch_bounds__(TX* TY) void k_cg1(DevCloth c, StepScalars s, int k) {
    constexpr int NT = TX * TY;
    constexpr int WX = TX + 2 * H, WY = TY + 2 * H;
    constexpr int NWIN = (WX * WY + NT - 1) / NT;
    constexpr int NP = 8;  // partials per lane loaded up front, per dot
    __shared__ Cg1Shared<TX, TY> sh;
    const int nt = c.nt, N = c.N;
    const int tile = tile_of_block(blockIdx.x, nt);
    const int bx = tile % c.ntx, by = tile / c.ntx;
    const int lx = threadIdx.x % TX, ly = threadIdx.x / TX;
    const int gi = bx * TX + lx, gj = by * TY + ly;
    const bool own = gi < c.nx && gj < c.ny;
    const int i = gj * c.nx + gi;
    const bool jac = (s.flags & OC_FLAG_JACOBI) != 0;
    const Cg1Bufs bf = cg1_bufs(c, k);

    // ---- loads that do not depend on the reduction
    const int done = c.ctl[0];
    const float gamma_old = k > 0 ? c.scal[k - 1] : 0.0f;
    const float alpha_old = k > 0 ? c.scal[c.cg1_alpha + k - 1] : 0.0f;
    const float gamma0 = c.scal[0];
    float pg[NP], pd[NP];
    if (threadIdx.x < 64) {
#pragma unroll
        for (int j = 0; j < NP; j++) {
            const int t = threadIdx.x + 64 * j;
            pg[j] = t < nt ? c.del_part[(size_t)k * nt + t] : 0.0f;
            pd[j] = t < nt ? c.sig_part[(size_t)k * nt + t] : 0.0f;
        }
    }
    const int x0 = bx * TX - H, y0 = by * TY - H;
    V3 wr[NWIN], ww[NWIN], wsv[NWIN], wpv[NWIN];
#pragma unroll
    for (int n = 0; n < NWIN; n++) {
        const int idx = threadIdx.x + n * NT;
        const int gx = x0 + idx % WX, gy = y0 + idx / WX;
        wr[n] = ww[n] = wsv[n] = wpv[n] = V3{0.0f, 0.0f, 0.0f};
        if (idx < WX * WY && gx >= 0 && gx < c.nx && gy >= 0 && gy < c.ny) {
            const int g = gy * c.nx + gx;
            wr[n] = ld3o(bf.r_in, N, g);
            ww[n] = ld3o(bf.w_in, N, g);
            if (k > 0) wsv[n] = ld3o(bf.s_in, N, g);
            if (jac) wpv[n] = ld3o(c.pinv, N, g);
        }
    }
    V3 pold = {0.0f, 0.0f, 0.0f}, x = pold, rk = pold, pv = pold;
    if (own) {
        if (k > 0) pold = ld3o(c.d1, N, i);
        x = ld3o(c.dv, N, i);
        rk = ld3o(bf.r_in, N, i);
        if (jac) pv = ld3o(c.pinv, N, i);
    }
    uint32_t mask = 0;
    if constexpr (STENCIL) mask = own ? ldo(c.pair_mask, (uint32_t)i) : 0u;
    const bool pinned = own ? ldo(c.pin, (uint32_t)i) != 0 : true;

    // ---- gamma_k, delta_k (the same lane tree as the other launches)
    if (threadIdx.x < 64) {
        float sg = 0.0f, sd = 0.0f;
#pragma unroll
        for (int j = 0; j < NP; j++) {
            sg += pg[j];
            sd += pd[j];
        }
        for (int t = threadIdx.x + 64 * NP; t < nt; t += 64) {
            sg += c.del_part[(size_t)k * nt + t];
            sd += c.sig_part[(size_t)k * nt + t];
        }
        sg = wave_sum(sg);
        sd = wave_sum(sd);
        if (threadIdx.x == 0) {
            sh.bc = sg;
            sh.bc2 = sd;
        }
    }
    __syncthreads();
    if (done) return;
    const float gamma = sh.bc, delta = sh.bc2;
    const float thresh = s.tol_mode == 1 ? s.tol2 * (k == 0 ? gamma : gamma0) : s.tol2;
    const bool stop = s.tol_mode != 2 && !(gamma > thresh);
    const float beta = (k > 0 && gamma_old != 0.0f) ? gamma / gamma_old : 0.0f;
    float den = delta;
    if (beta != 0.0f && alpha_old != 0.0f) {
        float t = beta * gamma;
        t = t / alpha_old;
        den = delta - t;
    }
    const bool brk = !stop && (den < 0.0f || !isfinite(den));
    const float alpha = den == 0.0f ? 0.0f : gamma / den;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        c.scal[k] = gamma;
        c.scal[c.cg1_alpha + k] = alpha;
        if (stop || brk) {
            c.ctl[0] = 1;
            c.ctl[1] = k;
            if (brk) c.ctl[2] |= OC_SOLVE_BREAKDOWN;
        }
    }
    if (stop || brk) return;

    // ---- pair blocks of this vertex (the SpMV below)
    constexpr int NS = STENCIL ? kStencil : WP;
    uint16_t m[NS];
    float C[NS][6];
    if constexpr (STENCIL) {
#pragma unroll
        for (int q = 0; q < kStencil; q++) {
            const int di = stencil_di(q), dj = stencil_dj(q);
            const uint32_t blk = stencil_owner(q) ? (uint32_t)i : (uint32_t)(i + dj * c.nx + di);
            const uint32_t off = blk + (uint32_t)c.stencil_plane[q] * 6u * (uint32_t)N;
#pragma unroll
            for (int e = 0; e < 6; e++) C[q][e] = ((mask >> q) & 1u) ? ldo(c.blocks, off + (uint32_t)(e * N)) : 0.0f;
        }
    } else {
#pragma unroll
        for (int sl = 0; sl < WP; sl++) {
            m[sl] = (own && sl < c.wp) ? ldo(c.pair_m, (uint32_t)(sl * N + i)) : kEmptySlot;
            if (m[sl] != kEmptySlot) {
                const int j = i + slot_dj(m[sl]) * c.nx + slot_di(m[sl]);
                const int blk = ((m[sl] >> 12) & 1) ? i : j;
                const uint32_t b0 = (uint32_t)((m[sl] >> 8) & 15) * 6u * (uint32_t)N + (uint32_t)blk;
#pragma unroll
                for (int e = 0; e < 6; e++) C[sl][e] = ldo(c.blocks + (size_t)e * N, b0);
            } else {
#pragma unroll
                for (int e = 0; e < 6; e++) C[sl][e] = 0.0f;
            }
        }
    }

    // ---- s_k, r_{k+1}, u_{k+1} on the tile + halo; the tile's own values out
#pragma unroll
    for (int n = 0; n < NWIN; n++) {
        const int idx = threadIdx.x + n * NT;
        if (idx < WX * WY) {
            const int wx = idx % WX, wy = idx / WX;
            const V3 sk = k == 0 ? ww[n] : V3{ww[n].x + beta * wsv[n].x, ww[n].y + beta * wsv[n].y,
                                              ww[n].z + beta * wsv[n].z};
            const V3 rn = {wr[n].x - alpha * sk.x, wr[n].y - alpha * sk.y, wr[n].z - alpha * sk.z};
            const V3 un = jac ? V3{wpv[n].x * rn.x, wpv[n].y * rn.y, wpv[n].z * rn.z} : rn;
            const int gx = x0 + wx, gy = y0 + wy;
            if (wx >= H && wx < H + TX && wy >= H && wy < H + TY && gx < c.nx && gy < c.ny) {
                const int g = gy * c.nx + gx;
                st3o(bf.s_out, N, g, sk);
                st3o(bf.r_out, N, g, rn);
            }
            if (wx >= H && wx < H + TX && wy >= H && wy < H + TY) {
                sh.rt[0][wy - H][wx - H] = rn.x;
                sh.rt[1][wy - H][wx - H] = rn.y;
                sh.rt[2][wy - H][wx - H] = rn.z;
            }
            sh.wu[0][wy][wx] = un.x;
            sh.wu[1][wy][wx] = un.y;
            sh.wu[2][wy][wx] = un.z;
        }
    }
    // this vertex: p_k = u_k + beta p_{k-1}, dv += alpha p_k
    if (own) {
        const V3 uk = jac ? V3{pv.x * rk.x, pv.y * rk.y, pv.z * rk.z} : rk;
        const V3 pk = k == 0 ? uk : V3{uk.x + beta * pold.x, uk.y + beta * pold.y, uk.z + beta * pold.z};
        st3o(c.d1, N, i, pk);
        st3o(c.dv, N, i, V3{x.x + alpha * pk.x, x.y + alpha * pk.y, x.z + alpha * pk.z});
    }
    __syncthreads();

    // ---- w_{k+1} = S A u_{k+1}; partials of r.u and w.u
    float tg = 0.0f, td = 0.0f;
    if (own) {
        const V3 ui = {sh.wu[0][ly + H][lx + H], sh.wu[1][ly + H][lx + H], sh.wu[2][ly + H][lx + H]};
        V3 acc = {s.md * ui.x, s.md * ui.y, s.md * ui.z};
        if constexpr (STENCIL) {
#pragma unroll
            for (int q = 0; q < kStencil; q++) {
                if (!((mask >> q) & 1u)) continue;
                const int wy = ly + H + stencil_dj(q), wx = lx + H + stencil_di(q);
                V3 w = {sh.wu[0][wy][wx] - ui.x, sh.wu[1][wy][wx] - ui.y, sh.wu[2][wy][wx] - ui.z};
                V3 a = symmv(C[q], w);
                acc = {acc.x + a.x, acc.y + a.y, acc.z + a.z};
            }
        } else {
#pragma unroll
            for (int sl = 0; sl < WP; sl++) {
                if (m[sl] == kEmptySlot) break;
                const int wy = ly + H + slot_dj(m[sl]), wx = lx + H + slot_di(m[sl]);
                V3 w = {sh.wu[0][wy][wx] - ui.x, sh.wu[1][wy][wx] - ui.y, sh.wu[2][wy][wx] - ui.z};
                V3 a = symmv(C[sl], w);
                acc = {acc.x + a.x, acc.y + a.y, acc.z + a.z};
            }
        }
        const V3 wn = pinned ? V3{0.0f, 0.0f, 0.0f} : acc;
        st3o(bf.w_out, N, i, wn);
        const V3 rn = {sh.rt[0][ly][lx], sh.rt[1][ly][lx], sh.rt[2][ly][lx]};
        tg = dot3(rn, ui);
        td = dot3(wn, ui);
    }
    const float pgs = tile_sum<NT>(tg, sh.ws);
    const float pds = tile_sum<NT>(td, sh.ws2);
    if (threadIdx.x == 0) {
        c.del_part[(size_t)(k + 1) * nt + tile] = pgs;
        c.sig_part[(size_t)(k + 1) * nt + tile] = pds;
    }
}

// CG iteration k, part 2: dv += alpha d, r -= alpha q, partial r.r.  Returns
// true when the solve has stopped (breakdown or an earlier stop).
template <int TX, int TY>
__device__ inline bool axpy_phase(const DevCloth& c, const StepScalars& s, int k, bool read_done,
                                  float* ws, float& bc, int vb) {
    constexpr int NT = TX * TY;
    const int nt = c.nt, N = c.N;
    const int tile = tile_of_block(vb, nt);
    const int bx = tile % c.ntx, by = tile / c.ntx;
    const int lx = threadIdx.x % TX, ly = threadIdx.x / TX;
    const int gi = bx * TX + lx, gj = by * TY + ly;
    const bool own = gi < c.nx && gj < c.ny;
    const int i = gj * c.nx + gi;

    // ---- loads that do not depend on alpha
    const int done = read_done ? c.ctl[0] : 0;
    const float delta = c.scal[k];
    float pv[kPartialRegs];
    if (threadIdx.x < 64) partial_load(c.sig_part + (size_t)k * nt, nt, pv);
    const float* dcur = (k & 1) ? c.d1 : c.d0;
    const bool jac = (s.flags & OC_FLAG_JACOBI) != 0;
    V3 d = {0.0f, 0.0f, 0.0f}, q = d, x = d, r = d, jv = d;
    if (own) {
        d = ld3o(dcur, N, i);
        q = ld3o(c.q, N, i);
        x = ld3o(c.dv, N, i);
        r = ld3o(c.r, N, i);
        if (jac) jv = ld3o(c.pinv, N, i);
    }
    if (threadIdx.x < 64) {
        const float sum = partial_finish(c.sig_part + (size_t)k * nt, nt, pv);
        if (threadIdx.x == 0) bc = sum;
    }
    __syncthreads();
    if (done) return true;
    const float sigma = bc;
    if (sigma < 0.0f || !isfinite(sigma)) {
        if (vb == 0 && threadIdx.x == 0) {
            c.ctl[0] = 1;
            c.ctl[1] = k;
            c.ctl[2] |= OC_SOLVE_BREAKDOWN;
        }
        return true;
    }
    // sigma == 0 only after r (hence d) underflowed: no update, no 0/0
    const float alpha = sigma == 0.0f ? 0.0f : delta / sigma;
    float t = 0.0f;
    if (own) {
        x = {x.x + alpha * d.x, x.y + alpha * d.y, x.z + alpha * d.z};
        r = {r.x - alpha * q.x, r.y - alpha * q.y, r.z - alpha * q.z};
        st3o(c.dv, N, i, x);
        st3o(c.r, N, i, r);
        if (jac) {
            const V3 z = {jv.x * r.x, jv.y * r.y, jv.z * r.z};
            st3o(c.s, N, i, z);
            t = dot3(r, z);
        } else {
            t = dot3(r, r);
        }
    }
    float p = tile_sum<NT>(t, ws);
    if (threadIdx.x == 0) c.del_part[(size_t)(k + 1) * nt + tile] = p;
    return false;
}

template <int TX, int TY>
__global__ __launch_bounds__(TX* TY) void k_cg_axpy(DevCloth c, StepScalars s, int k) {
    __shared__ float ws[TX * TY / 64];
    __shared__ float bc;
    axpy_phase<TX, TY>(c, s, k, true, ws, bc, blockIdx.x);
}

// Grid barrier of the persistent solve: every block adds 1 to *bar and waits
// until it reaches target.  The add is a release (the block's writes reach
// the device-coherent level first); the wait polls with relaxed agent-scope
// loads, which read the coherent value without invalidating caches, and a
// single acquire fence follows.  The wait is bounded; on time-out the block
// sets ctl[4] and the caller ends the kernel, so every wave exits.
__device__ inline bool grid_barrier(unsigned* bar, unsigned target, int* err) {
    __syncthreads();
    __shared__ int ok;
    if (threadIdx.x == 0) {
        __hip_atomic_fetch_add(bar, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        int good = 1;
        unsigned spins = 0;
        while (__hip_atomic_load(bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            __builtin_amdgcn_s_sleep(1);
            if (++spins > (1u << 24)) {
                good = 0;
                __hip_atomic_fetch_or(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                break;
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        ok = good;
    }
    __syncthreads();
    return ok != 0;
}

// All CG iterations in one cooperative launch: the spmv and axpy phases of
// the multi-kernel form, separated by grid barriers instead of kernel
// boundaries.  Each resident block serves tiles blockIdx.x, +gridDim.x, ...
// with the same per-tile code.  Every block owns at least one tile and so
// takes the same stop decisions as the others: all blocks pass the same
// barriers.
template <int TX, int TY, int WP>
__global__ __launch_bounds__(TX* TY) void k_cg_persistent(DevCloth c, StepScalars s, unsigned* bar) {
    __shared__ SpmvShared<TX, TY> sh;
    __shared__ float ws[TX * TY / 64];
    __shared__ float bc;
    const unsigned nb = gridDim.x;
    unsigned target = 0;
    for (int k = 0; k < s.max_it; k++) {
        bool stop = false;
        for (int vb = blockIdx.x; vb < c.nt && !stop; vb += gridDim.x)
            stop = spmv_phase<TX, TY, WP>(c, s, k, false, sh, vb);
        if (stop) break;
        target += nb;
        if (!grid_barrier(bar, target, c.ctl + 4)) break;
        for (int vb = blockIdx.x; vb < c.nt && !stop; vb += gridDim.x)
            stop = axpy_phase<TX, TY>(c, s, k, false, ws, bc, vb);
        if (stop) break;
        target += nb;
        if (!grid_barrier(bar, target, c.ctl + 4)) break;
    }
}

__global__ void k_reset_barrier(unsigned* bar) {
    if (threadIdx.x == 0) *bar = 0u;
}

// Iteration count and status of the solve (one 64-thread block).
__global__ void k_cg_finish(DevCloth c, StepScalars s) {
    __shared__ float bc;
    cg_finish_block(c, s, &bc);
}

// v += dv, x += h v, optional ground clamp (IE:693-699).  FIN: block 0 also
// records the solve's iteration count and status (the step's last launch).
template <bool FIN>
__global__ __launch_bounds__(256) void k_update(DevCloth c, StepScalars s) {
    if (FIN && blockIdx.x == 0) {
        __shared__ float bc;
        cg_finish_block(c, s, &bc);
    }
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= c.N) return;
    const int N = c.N;
    V3 v = ld3o(c.v, N, i), dv = ld3o(c.dv, N, i), x = ld3o(c.x, N, i);
    v = {v.x + dv.x, v.y + dv.y, v.z + dv.z};
    x = {x.x + s.h * v.x, x.y + s.h * v.y, x.z + s.h * v.z};
    if ((s.flags & OC_FLAG_GROUND_CLAMP) && x.y < 0.0f) x.y = 0.0f;
    st3o(c.v, N, i, v);
    st3o(c.x, N, i, x);
    if (!(isfinite(x.x) && isfinite(x.y) && isfinite(x.z) && isfinite(v.x) && isfinite(v.y) &&
          isfinite(v.z))) {
        atomicOr(c.ctl + 2, OC_SOLVE_NONFINITE);
        atomicOr(c.ctl + 3, OC_SOLVE_NONFINITE);
    }
}

// ---------------------------------------------------------------------------
// adjoint of one step
// ---------------------------------------------------------------------------
// Adjoint solve start: gvt = gv' + h gx', r = d = S gvt, mu = 0, partial r.r
template <int TX, int TY>
__global__ __launch_bounds__(TX* TY) void k_adj_init(DevCloth c, StepScalars s, AdjointBufs a) {
    constexpr int NT = TX * TY;
    __shared__ float ws[NT / 64];
    const int lx = threadIdx.x % TX, ly = threadIdx.x / TX;
    const int tile = tile_of_block(blockIdx.x, c.nt);
    const int bx = tile % c.ntx, by = tile / c.ntx;
    const int gi = bx * TX + lx, gj = by * TY + ly;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        c.ctl[0] = 0;
        c.ctl[1] = 0;
        c.ctl[2] = 0;
    }
    float t = 0.0f;
    if (gi < c.nx && gj < c.ny) {
        const int N = c.N, i = gj * c.nx + gi;
        V3 gx = ld3(a.gx_out, N, i), gv = ld3(a.gv_out, N, i);
        V3 g = {gv.x + s.h * gx.x, gv.y + s.h * gx.y, gv.z + s.h * gx.z};
        st3(a.gvt, N, i, g);
        V3 r = c.pin[i] ? V3{0.0f, 0.0f, 0.0f} : g;
        st3(c.r, N, i, r);
        st3(c.dv, N, i, V3{0.0f, 0.0f, 0.0f});
        if (s.flags & OC_FLAG_JACOBI) {  // pinv from the re-assembly
            const V3 pv = ld3(c.pinv, N, i);
            const V3 z = {pv.x * r.x, pv.y * r.y, pv.z * r.z};
            st3(c.s, N, i, z);
            st3(c.d0, N, i, z);
            t = dot3(r, z);
        } else {
            st3(c.d0, N, i, r);
            t = dot3(r, r);
        }
    }
    float p = tile_sum<NT>(t, ws);
    if (threadIdx.x == 0) c.del_part[tile] = p;
}

// gx_in = gx' + dJ/dx, gv_in = gvt + dJ/dv, gks += dJ/dks (added by each
// spring's p1).  c.x, c.v: pre-step state; c.dv: the step's solution.
template <int TX, int TY>
__global__ __launch_bounds__(TX* TY) void k_adj_gather(DevCloth c, StepScalars s, AdjointBufs a) {
    const int lx = threadIdx.x % TX, ly = threadIdx.x / TX;
    const int tile = tile_of_block(blockIdx.x, c.nt);
    const int bx = tile % c.ntx, by = tile / c.ntx;
    const int gi = bx * TX + lx, gj = by * TY + ly;
    if (gi >= c.nx || gj >= c.ny) return;
    const int N = c.N, i = gj * c.nx + gi;
    const V3 xi = ld3(c.x, N, i), vi = ld3(c.v, N, i), mi = ld3(a.mu, N, i), pi = ld3(c.dv, N, i);
    const V3 wi = add3(vi, pi);
    V3 gx = ld3(a.gx_out, N, i), gv = ld3(a.gvt, N, i);
    if (s.damp != 0.0f && !c.pin[i])  // the damping force c v_i: dJ/dv_i += h c mu_i
        gv = {gv.x + s.h * (s.damp * mi.x), gv.y + s.h * (s.damp * mi.y), gv.z + s.h * (s.damp * mi.z)};
    for (int k = 0; k < c.wa; k++) {
        const int sp = c.ainc_s[k * N + i];
        if (sp < 0) break;
        const uint16_t m = c.ainc_m[k * N + i];
        const int j = i + slot_dj(m) * c.nx + slot_di(m);
        const V3 xj = ld3(c.x, N, j), vj = ld3(c.v, N, j), mj = ld3(a.mu, N, j), pj = ld3(c.dv, N, j);
        const V3 wj = add3(vj, pj);
        const bool second = (m >> 8) & 1u;
        float ge[3], gvv[3], gk;
        if (second)
            spring_adjoint(xj, xi, vj, vi, mj, mi, wj, wi, pj, pi, c.L0[sp], c.ks[sp], c.kd[sp], s.h,
                           s.hh, ge, gvv, gk);
        else
            spring_adjoint(xi, xj, vi, vj, mi, mj, wi, wj, pi, pj, c.L0[sp], c.ks[sp], c.kd[sp], s.h,
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

thread_local LaunchProfile* t_prof = nullptr;

// Launch with the installed profile's next event pair, if any (see
// LaunchProfile).
template <typename F, typename... Args>
hipError_t launch_k(int kind, F kernel, dim3 grid, dim3 block, hipStream_t st, Args... args) {
    LaunchProfile* p = t_prof;
    if (p && p->n < p->cap) {
        const int k = p->n++;
        p->kind[k] = kind;
        hipExtLaunchKernelGGL(kernel, grid, block, 0, st, p->start[k], p->stop[k], 0, args...);
    } else {
        hipLaunchKernelGGL(kernel, grid, block, 0, st, args...);
    }
    return hipGetLastError();
}

}  // namespace

void set_launch_profile(LaunchProfile* p) { t_prof = p; }

hipError_t launch_adj_init(const DevCloth& c, const StepScalars& s, const AdjointBufs& a,
                           hipStream_t st) {
    hipLaunchKernelGGL((k_adj_init<kTileX, kTileY>), dim3(c.nt), dim3(kTileX * kTileY), 0, st, c, s,
                       a);
    return hipGetLastError();
}

hipError_t launch_adj_gather(const DevCloth& c, const StepScalars& s, const AdjointBufs& a,
                             hipStream_t st) {
    hipLaunchKernelGGL((k_adj_gather<kTileX, kTileY>), dim3(c.nt), dim3(kTileX * kTileY), 0, st, c,
                       s, a);
    return hipGetLastError();
}

hipError_t launch_forces(const DevCloth& c, const StepScalars& s, hipStream_t st) {
    hipLaunchKernelGGL((k_forces<kTileX, kTileY>), dim3(c.nt), dim3(kTileX * kTileY), 0, st, c, s);
    return hipGetLastError();
}

hipError_t launch_assemble(const DevCloth& c, const StepScalars& s, hipStream_t st, bool fused_update) {
    // A fused update needs a warm start (a cold assembly zeroes dv, which
    // neighbouring blocks still read for the update) and at least one CG
    // iteration (the previous solve's status is taken from partials that a
    // solve without iterations leaves where this assembly writes its own).
    if (fused_update && (!s.warm || s.max_it < 1)) return hipErrorInvalidValue;
    const dim3 grid(c.nt), block(kTileX * kTileY);
    if (c.stencil_asm)
        return fused_update ? launch_k(0, k_assemble_st<kTileX, kTileY, true>, grid, block, st, c, s)
                            : launch_k(0, k_assemble_st<kTileX, kTileY, false>, grid, block, st, c, s);
    if (c.wa <= 16)
        return fused_update ? launch_k(0, k_assemble<kTileX, kTileY, 16, true>, grid, block, st, c, s)
                            : launch_k(0, k_assemble<kTileX, kTileY, 16, false>, grid, block, st, c, s);
    return fused_update
               ? launch_k(0, k_assemble<kTileX, kTileY, kMaxAssemblyWidth, true>, grid, block, st, c, s)
               : launch_k(0, k_assemble<kTileX, kTileY, kMaxAssemblyWidth, false>, grid, block, st, c, s);
}

// cst[spring_cst[sp] + q*N] = (L0, ks, kd)[q][sp] for q in [q0, q0 + nq).
__global__ __launch_bounds__(256) void k_scatter_cst(DevCloth c, const int32_t* spring_cst, int S, int q0,
                                                     int nq) {
    const float* src[3] = {c.L0, c.ks, c.kd};
    float* dst = const_cast<float*>(c.cst);
    for (int sp = blockIdx.x * 256 + threadIdx.x; sp < S; sp += gridDim.x * 256) {
        const size_t base = (size_t)spring_cst[sp];
        for (int q = q0; q < q0 + nq; q++) dst[base + (size_t)q * c.N] = src[q][sp];
    }
}

hipError_t launch_scatter_cst(const DevCloth& c, const int32_t* spring_cst, int S, int q0, int nq,
                              hipStream_t st) {
    if (q0 < 0 || nq < 0 || q0 + nq > 3) return hipErrorInvalidValue;
    if (S <= 0 || nq == 0) return hipSuccess;
    const int grid = std::min((S + 255) / 256, 4096);
    hipLaunchKernelGGL(k_scatter_cst, dim3(grid), dim3(256), 0, st, c, spring_cst, S, q0, nq);
    return hipGetLastError();
}

hipError_t launch_cg_spmv(const DevCloth& c, const StepScalars& s, int k, hipStream_t st) {
    const dim3 grid(c.nt), block(kTileX * kTileY);
    if (c.stencil) return launch_k(1, k_cg_spmv<kTileX, kTileY, kStencil, true>, grid, block, st, c, s, k);
    if (c.wp <= 12) return launch_k(1, k_cg_spmv<kTileX, kTileY, 12, false>, grid, block, st, c, s, k);
    return launch_k(1, k_cg_spmv<kTileX, kTileY, kMaxPairWidth, false>, grid, block, st, c, s, k);
}

hipError_t launch_cg_axpy(const DevCloth& c, const StepScalars& s, int k, hipStream_t st) {
    return launch_k(2, k_cg_axpy<kTileX, kTileY>, dim3(c.nt), dim3(kTileX * kTileY), st, c, s, k);
}

hipError_t launch_cg_finish(const DevCloth& c, const StepScalars& s, hipStream_t st) {
    return launch_k(3, k_cg_finish, dim3(1), dim3(64), st, c, s);
}

int persistent_cg_grid(const DevCloth& c) {
    if (c.wp > 12) return 0;
    int dev = 0, coop = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev) != hipSuccess || !coop)
        return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_cg_persistent<kTileX, kTileY, 12>,
                                                     kTileX * kTileY, 0) != hipSuccess)
        return 0;
    // Every block must own at least one tile: a block without one would not
    // see the stop decision and would wait at a barrier the others have
    // left.  Near-equal tile counts per block; a multiple of 8 blocks keeps
    // every tile a block serves on the block's XCD (tile_of_block).
    const int resident = per_cu * cus;
    if (resident < 1) return 0;
    const int per_block = (c.nt + resident - 1) / resident;
    int grid = (c.nt + per_block - 1) / per_block;
    if (grid >= 8) grid = grid / 8 * 8;
    return std::min(grid, c.nt);
}

hipError_t launch_cg(const DevCloth& c, const StepScalars& s, hipStream_t st, bool finish) {
    hipError_t e;
    if (c.persistent && s.max_it > 0 && !(s.flags & OC_FLAG_CG_SINGLE_REDUCTION)) {
        if (c.persistent > c.nt) return hipErrorInvalidValue;  // a block without a tile
        unsigned* bar = reinterpret_cast<unsigned*>(c.ctl + 6);
        hipLaunchKernelGGL(k_reset_barrier, dim3(1), dim3(64), 0, st, bar);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        DevCloth ca = c;
        StepScalars sa = s;
        void* args[] = {(void*)&ca, (void*)&sa, (void*)&bar};
        e = hipLaunchCooperativeKernel(
            reinterpret_cast<const void*>(k_cg_persistent<kTileX, kTileY, 12>), dim3(c.persistent),
            dim3(kTileX * kTileY), args, 0, st);
        if (e != hipSuccess) return e;
        return finish ? launch_cg_finish(c, s, st) : hipSuccess;
    }
    if (s.flags & OC_FLAG_CG_SINGLE_REDUCTION) {
        // the start: w0 = S A u0 and delta_0 (the first standard SpMV), then
        // one launch per iteration
        if (s.max_it > 0 && (e = launch_cg_spmv(c, s, 0, st)) != hipSuccess) return e;
        const dim3 grid(c.nt), block(kTileX * kTileY);
        for (int k = 0; k < s.max_it; k++) {
            if (c.stencil)
                e = launch_k(1, k_cg1<kTileX, kTileY, kStencil, true>, grid, block, st, c, s, k);
            else if (c.wp <= 12)
                e = launch_k(1, k_cg1<kTileX, kTileY, 12, false>, grid, block, st, c, s, k);
            else
                e = launch_k(1, k_cg1<kTileX, kTileY, kMaxPairWidth, false>, grid, block, st, c, s, k);
            if (e != hipSuccess) return e;
        }
        return finish ? launch_cg_finish(c, s, st) : hipSuccess;
    }
    for (int k = 0; k < s.max_it; k++) {
        if ((e = launch_cg_spmv(c, s, k, st)) != hipSuccess) return e;
        if ((e = launch_cg_axpy(c, s, k, st)) != hipSuccess) return e;
    }
    return finish ? launch_cg_finish(c, s, st) : hipSuccess;
}

hipError_t launch_update(const DevCloth& c, const StepScalars& s, hipStream_t st, bool finish) {
    const dim3 grid((c.N + 255) / 256), block(256);
    return finish ? launch_k(4, k_update<true>, grid, block, st, c, s)
                  : launch_k(4, k_update<false>, grid, block, st, c, s);
}

}  // namespace oc
