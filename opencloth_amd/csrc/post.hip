// Post-step passes of StepPhysics and the position-Verlet step (gfx950).
//
//   k_ellipsoid : EllipsoidCollision (IE:727-751, VL:509-533), per vertex
//   k_provot    : ApplyProvotDynamicInverse (IE:702-726), per-vertex gather
//                 of the springs' velocity corrections in list order (the
//                 reference only reads X there, so the gather equals its
//                 sequential scatter bit for bit)
//   k_verlet    : ComputeForces(dt) + IntegrateVerlet (VL:428-484) fused per
//                 vertex on 32x8 tiles, whose positions and velocities (with
//                 a 2-wide halo) are staged in LDS; reads x, x_last and
//                 writes the next positions into a third buffer, so no
//                 thread sees a half-updated neighbour
//
// Expressions follow the reference's GLM order, built with -ffp-contract=off.
#include "oc_device.hpp"

namespace oc {

namespace {

using namespace dev;

__global__ __launch_bounds__(256) void k_ellipsoid(float* x, float* w, int N, oc_ellipsoid e,
                                                   int mode) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= N) return;
    const V3 p = ld3(x, N, i);
    float d0[3];
#pragma unroll
    for (int r = 0; r < 3; r++) {
        const float t = e.inverse[0 * 4 + r] * p.x + e.inverse[1 * 4 + r] * p.y +
                        e.inverse[2 * 4 + r] * p.z + e.inverse[3 * 4 + r] * 1.0f;
        d0[r] = t - e.center[r];
    }
    const float dist = sqrtf(d0[0] * d0[0] + d0[1] * d0[1] + d0[2] * d0[2]);
    if (!(dist < 1.0f)) return;
#pragma unroll
    for (int c = 0; c < 3; c++) d0[c] = ((e.radius - dist) * d0[c]) / dist;
    float delta[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        float t[3] = {e.transform[0 * 4 + a], e.transform[1 * 4 + a], e.transform[2 * 4 + a]};
        const float tt = t[0] * t[0] + t[1] * t[1] + t[2] * t[2];
#pragma unroll
        for (int c = 0; c < 3; c++) t[c] = t[c] / tt;
        delta[a] = d0[0] * t[0] + d0[1] * t[1] + d0[2] * t[2];
    }
    const V3 q = {p.x + delta[0], p.y + delta[1], p.z + delta[2]};
    st3(x, N, i, q);
    st3(w, N, i, mode == 1 ? q : V3{0.0f, 0.0f, 0.0f});
}

// Velocity correction of one spring for its endpoints (IE:706-714).
__device__ inline bool provot_delta(V3 x1, V3 x2, float rest, float slack, V3& dP) {
    V3 dp = sub3(x1, x2);
    const float sq = dp.x * dp.x + dp.y * dp.y + dp.z * dp.z;
    float dist = sqrtf(sq);
    const float thr = rest * slack;
    if (!(dist > thr)) return false;
    dist -= thr;
    dist /= 2.0f;
    const float inv = 1.0f / sqrtf(sq);
    dP = {(dp.x * inv) * dist, (dp.y * inv) * dist, (dp.z * inv) * dist};
    return true;
}

// Every slot's loads (slot table, then the other end's position and pin and
// the spring's rest length) are issued before the first correction is
// computed: the pass is one short chain per vertex rather than one load
// round trip per spring.  The corrections are then added in slot order.
template <int WA>
__global__ __launch_bounds__(256) void k_provot(DevCloth c, float slack) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= c.N) return;
    const int N = c.N;
    int sps[WA];
    uint16_t ms[WA];
#pragma unroll
    for (int k = 0; k < WA; k++) {
        sps[k] = k < c.wa ? c.ainc_s[k * N + i] : -1;
        ms[k] = k < c.wa ? c.ainc_m[k * N + i] : kEmptySlot;
    }
    const V3 xi = ld3(c.x, N, i);
    V3 v = ld3(c.v, N, i);
    const bool pin_i = c.pin[i] != 0;
    V3 xs[WA];
    float rest[WA];
    bool pins[WA];
#pragma unroll
    for (int k = 0; k < WA; k++) {
        xs[k] = V3{0.0f, 0.0f, 0.0f};
        rest[k] = 0.0f;
        pins[k] = false;
        if (sps[k] >= 0) {
            const int j = i + slot_dj(ms[k]) * c.nx + slot_di(ms[k]);
            xs[k] = ld3(c.x, N, j);
            rest[k] = c.L0[sps[k]];
            pins[k] = c.pin[j] != 0;
        }
    }
    bool touched = false;
#pragma unroll
    for (int k = 0; k < WA; k++) {
        if (sps[k] < 0) continue;  // (the slots are packed: the rest are empty too)
        const bool second = (ms[k] >> 8) & 1u;
        V3 dP;
        if (!(second ? provot_delta(xs[k], xi, rest[k], slack, dP)
                     : provot_delta(xi, xs[k], rest[k], slack, dP)))
            continue;
        if (!second) {  // i is p1: V[p1] -= dP unless p1 is pinned
            if (!pin_i) {
                v = sub3(v, dP);
                touched = true;
            }
        } else {  // i is p2: V[p2] += dP when p1 is pinned or p2 is free
            if (pins[k] || !pin_i) {
                v = add3(v, dP);
                touched = true;
            }
        }
    }
    if (touched) st3(c.v, N, i, v);
}

// Verlet step on a 2D tile with a kHalo-wide halo staged in LDS: positions
// and Verlet velocities v = (x - x_last)/dt of the window are computed once
// (the same division per vertex as evaluating them per spring end).
template <int TX, int TY>
struct VerletShared {
    static constexpr int WX = TX + 2 * kHalo, WY = TY + 2 * kHalo;
    float wx[3][WY][WX];
    float wv[3][WY][WX];
};

// Force on the tile's vertex (VL:448-484): gravity on free vertices, damping
// on all, springs in list order (pinned vertices: damping only).  Springs are
// evaluated from this vertex's side (x1 = x_i), which is bitwise the negated
// p1-side force, so it is always added.  FORCES: write the force; otherwise
// integrate (VL:428-444) into the next positions.
// STENCIL: the topology has the stencil assembly form; a vertex's springs
// come from its ownership mask in stencil order (the slot order on such
// topologies) and their constants from the owner planes at the pair's lower
// vertex, so no slot table or spring-id gather is read.
template <int TX, int TY, int WA, bool FORCES, bool STENCIL>
__global__ __launch_bounds__(TX* TY) void k_verlet(DevVerlet c, VerletScalars s) {
    constexpr int NT = TX * TY, H = kHalo;
    constexpr int WX = TX + 2 * H, WY = TY + 2 * H;
    constexpr int NWIN = (WX * WY + NT - 1) / NT;
    __shared__ VerletShared<TX, TY> sh;
    const int ntx = (c.nx + TX - 1) / TX, nt = ntx * ((c.ny + TY - 1) / TY);
    const int tile = tile_of_block(blockIdx.x, nt);
    const int bx = tile % ntx, by = tile / ntx;
    const int lx = threadIdx.x % TX, ly = threadIdx.x / TX;
    const int gi = bx * TX + lx, gj = by * TY + ly;
    const bool own = gi < c.nx && gj < c.ny;
    const int N = c.N, i = gj * c.nx + gi;
    int sps[WA];
    uint16_t ms[WA];
    uint32_t mi = 0u;
    if constexpr (STENCIL) {
        mi = own ? c.asm_mask[i] : 0u;
    } else {
#pragma unroll
        for (int k = 0; k < WA; k++) {
            sps[k] = (own && k < c.wa) ? c.ainc_s[k * N + i] : -1;
            ms[k] = (own && k < c.wa) ? c.ainc_m[k * N + i] : kEmptySlot;
        }
    }
    const int x0 = bx * TX - H, y0 = by * TY - H;
#pragma unroll
    for (int n = 0; n < NWIN; n++) {
        const int idx = threadIdx.x + n * NT;
        if (idx < WX * WY) {
            const int wx = idx % WX, wy = idx / WX;
            const int gx = x0 + wx, gy = y0 + wy;
            V3 xw = {0.0f, 0.0f, 0.0f}, vw = xw;
            if (gx >= 0 && gx < c.nx && gy >= 0 && gy < c.ny) {
                const int g = gy * c.nx + gx;
                xw = ld3(c.x, N, g);
                const V3 xl = ld3(c.xl, N, g);
                vw = {(xw.x - xl.x) / s.dt, (xw.y - xl.y) / s.dt, (xw.z - xl.z) / s.dt};
            }
            sh.wx[0][wy][wx] = xw.x;
            sh.wx[1][wy][wx] = xw.y;
            sh.wx[2][wy][wx] = xw.z;
            sh.wv[0][wy][wx] = vw.x;
            sh.wv[1][wy][wx] = vw.y;
            sh.wv[2][wy][wx] = vw.z;
        }
    }
    __syncthreads();
    if (!own) return;
    const int cy = ly + H, cx = lx + H;
    const V3 xi = {sh.wx[0][cy][cx], sh.wx[1][cy][cx], sh.wx[2][cy][cx]};
    const V3 vi = {sh.wv[0][cy][cx], sh.wv[1][cy][cx], sh.wv[2][cy][cx]};
    const bool pinned = c.pin[i] != 0;
    V3 F = {0.0f, 0.0f, 0.0f};
    if (!pinned) F = {F.x + s.gx * s.mass, F.y + s.gy * s.mass, F.z + s.gz * s.mass};
    F = {F.x + s.damping * vi.x, F.y + s.damping * vi.y, F.z + s.damping * vi.z};
    if (!pinned && STENCIL) {
#pragma unroll
        for (int k = 0; k < kStencil; k++) {
            if (!((mi >> k) & 1u)) continue;
            const int wy = cy + stencil_dj(k), wx = cx + stencil_di(k);
            const V3 xj = {sh.wx[0][wy][wx], sh.wx[1][wy][wx], sh.wx[2][wy][wx]};
            const V3 vj = {sh.wv[0][wy][wx], sh.wv[1][wy][wx], sh.wv[2][wy][wx]};
            const int g = stencil_owner(k) ? i : i + stencil_dj(k) * c.nx + stencil_di(k);
            const float* cp = c.cst + (size_t)owned_index(k) * 3 * N + g;
            F = add3(F, spring_force(xi, xj, vi, vj, cp[0], cp[N], cp[2 * (size_t)N]));
            if ((mi >> (16 + k)) & 1u) {  // the pair's second spring (multiplicity 1 planes)
                const float* c2 = cp + (size_t)kOwned * 3 * N;
                F = add3(F, spring_force(xi, xj, vi, vj, c2[0], c2[N], c2[2 * (size_t)N]));
            }
        }
    } else if (!pinned) {
#pragma unroll
        for (int k = 0; k < WA; k++) {
            const int sp = sps[k];
            if (sp < 0) break;
            const int wy = cy + slot_dj(ms[k]), wx = cx + slot_di(ms[k]);
            const V3 xj = {sh.wx[0][wy][wx], sh.wx[1][wy][wx], sh.wx[2][wy][wx]};
            const V3 vj = {sh.wv[0][wy][wx], sh.wv[1][wy][wx], sh.wv[2][wy][wx]};
            F = add3(F, spring_force(xi, xj, vi, vj, c.L0[sp], c.ks[sp], c.kd[sp]));
        }
    }
    if (FORCES) {
        st3(c.f, N, i, F);
        return;
    }
    const float c2 = (s.dt * s.dt) / s.mass;
    const V3 xl = ld3(c.xl, N, i);
    V3 xn = {xi.x + (xi.x - xl.x) + c2 * F.x, xi.y + (xi.y - xl.y) + c2 * F.y,
             xi.z + (xi.z - xl.z) + c2 * F.z};
    if ((s.flags & OC_FLAG_GROUND_CLAMP) && xn.y < 0.0f) xn.y = 0.0f;
    st3(c.xn, N, i, xn);
}

// The Verlet step has no reduction, so its tile shape is free: 32x8 measured
// faster than the implicit kernels' 64x8 at 512x512 (15.5 against 17.4 us).
constexpr int kVerletTileX = 32, kVerletTileY = 8;

template <bool FORCES>
hipError_t launch_verlet(const DevVerlet& c, const VerletScalars& s, hipStream_t st) {
    constexpr int TX = kVerletTileX, TY = kVerletTileY;
    const int nt = ((c.nx + TX - 1) / TX) * ((c.ny + TY - 1) / TY);
    if (c.stencil_asm)
        hipLaunchKernelGGL((k_verlet<TX, TY, 1, FORCES, true>), dim3(nt), dim3(TX * TY), 0, st, c, s);
    else if (c.wa <= 16)
        hipLaunchKernelGGL((k_verlet<TX, TY, 16, FORCES, false>), dim3(nt), dim3(TX * TY), 0, st, c, s);
    else
        hipLaunchKernelGGL((k_verlet<TX, TY, kMaxAssemblyWidth, FORCES, false>), dim3(nt), dim3(TX * TY), 0,
                           st, c, s);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_ellipsoid(float* x, float* w, int N, const oc_ellipsoid& e, int mode,
                            hipStream_t st) {
    hipLaunchKernelGGL(k_ellipsoid, dim3((N + 255) / 256), dim3(256), 0, st, x, w, N, e, mode);
    return hipGetLastError();
}

hipError_t launch_provot(const DevCloth& c, float slack, hipStream_t st) {
    if (c.wa <= 16)
        hipLaunchKernelGGL(k_provot<16>, dim3((c.N + 255) / 256), dim3(256), 0, st, c, slack);
    else
        hipLaunchKernelGGL(k_provot<kMaxAssemblyWidth>, dim3((c.N + 255) / 256), dim3(256), 0, st, c, slack);
    return hipGetLastError();
}

hipError_t launch_verlet_forces(const DevVerlet& c, const VerletScalars& s, hipStream_t st) {
    return launch_verlet<true>(c, s, st);
}

hipError_t launch_verlet_step(const DevVerlet& c, const VerletScalars& s, hipStream_t st) {
    return launch_verlet<false>(c, s, st);
}

}  // namespace oc
