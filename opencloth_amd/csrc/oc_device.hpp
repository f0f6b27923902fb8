// Device-side helpers shared by the single-cloth and batched kernels:
// vec3 arithmetic in the reference's GLM expression order, the per-spring
// force/Jacobian, and the fixed-order reductions.  Built with
// -ffp-contract=off, so every expression rounds exactly as written.
#pragma once

#include "oc_internal.hpp"

namespace oc {
namespace dev {

constexpr int kPartialRegs = 16;  // tile partials per lane loaded up front

// Blocks are dispatched round-robin over the 8 XCDs; each XCD gets a
// contiguous band of tiles so that halo and shared-data re-reads hit its L2.
__device__ inline int tile_of_block(int b, int nt) {
    if ((nt & 7) == 0) return (b & 7) * (nt >> 3) + (b >> 3);
    return b;
}

struct V3 {
    float x, y, z;
};

__device__ inline V3 ld3(const float* p, int N, int i) { return {p[i], p[N + i], p[2 * N + i]}; }
__device__ inline void st3(float* p, int N, int i, V3 v) {
    p[i] = v.x;
    p[N + i] = v.y;
    p[2 * N + i] = v.z;
}

// Element idx of a device array, addressed by a 32-bit byte offset from the
// array's base.  With a uniform base this is the scalar-base form of a global
// load (no 64-bit address arithmetic per element).  Callers guarantee that
// the array is below 4 GiB (checked when a cloth is created).
template <typename T>
__device__ inline T ldo(const T* p, uint32_t idx) {
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(p) + idx * (uint32_t)sizeof(T));
}
template <typename T>
__device__ inline void sto(T* p, uint32_t idx, T v) {
    *reinterpret_cast<T*>(reinterpret_cast<char*>(p) + idx * (uint32_t)sizeof(T)) = v;
}
// SoA vec3 with the component planes as uniform bases: one 32-bit offset
// serves all three components.
__device__ inline V3 ld3o(const float* p, int N, int i) {
    return {ldo(p, (uint32_t)i), ldo(p + N, (uint32_t)i), ldo(p + 2 * (size_t)N, (uint32_t)i)};
}
__device__ inline void st3o(float* p, int N, int i, V3 v) {
    sto(p, (uint32_t)i, v.x);
    sto(p + N, (uint32_t)i, v.y);
    sto(p + 2 * (size_t)N, (uint32_t)i, v.z);
}
__device__ inline V3 add3(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ inline V3 sub3(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ inline V3 scl3(float s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
__device__ inline float dot3(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// symmetric (xx,yy,zz,xy,xz,yz) times vector, mat3*vec3 order
__device__ inline V3 symmv(const float* c, V3 w) {
    return {c[0] * w.x + c[3] * w.y + c[4] * w.z, c[3] * w.x + c[1] * w.y + c[5] * w.z,
            c[4] * w.x + c[5] * w.y + c[2] * w.z};
}

// Force on p1 (OpenCloth_ImplicitEuler/.../main.cpp:663-675), stiffness
// Jacobian K (:669-671) and P = dp dp^T / l^2, in the reference's expression
// order.
__device__ inline void spring_terms(V3 x1, V3 x2, V3 v1, V3 v2, float L0, float Ks, float Kd,
                                    V3& f, float* K, float* P) {
    V3 dp = {x1.x - x2.x, x1.y - x2.y, x1.z - x2.z};
    V3 dv = {v1.x - v2.x, v1.y - v2.y, v1.z - v2.z};
    float sq = dp.x * dp.x + dp.y * dp.y + dp.z * dp.z;
    float dist = sqrtf(sq);
    float dist2 = dist * dist;
    float lo_l = L0 / dist;
    P[0] = (dp.x * dp.x) / dist2;
    P[1] = (dp.y * dp.y) / dist2;
    P[2] = (dp.z * dp.z) / dist2;
    P[3] = (dp.y * dp.x) / dist2;
    P[4] = (dp.z * dp.x) / dist2;
    P[5] = (dp.z * dp.y) / dist2;
#pragma unroll
    for (int c = 0; c < 3; c++) K[c] = Ks * (-1.0f + lo_l * (1.0f - P[c]));
#pragma unroll
    for (int c = 3; c < 6; c++) K[c] = Ks * (lo_l * (0.0f - P[c]));
    float leftTerm = -Ks * (dist - L0);
    float rightTerm = Kd * ((dv.x * dp.x + dv.y * dp.y + dv.z * dp.z) / dist);
    float inv = 1.0f / sqrtf(sq);
    float s = leftTerm + rightTerm;
    f = {s * (dp.x * inv), s * (dp.y * inv), s * (dp.z * inv)};
}

// One spring from the x1 end: force on x1, K (v1 - v2) and the pair-block
// term h (kd P) + h^2 K, in the per-end assembly's expressions.
struct SpringOut {
    V3 f, kw;
    float c[6];
};
__device__ inline SpringOut eval_spring(V3 x1, V3 x2, V3 v1, V3 v2, float L0, float ks, float kd,
                                        float h, float hh) {
    SpringOut o;
    float K[6], P[6];
    spring_terms(x1, x2, v1, v2, L0, ks, kd, o.f, K, P);
#pragma unroll
    for (int q = 0; q < 6; q++) o.c[q] = h * (kd * P[q]) + hh * K[q];
    o.kw = symmv(K, V3{v1.x - v2.x, v1.y - v2.y, v1.z - v2.z});
    return o;
}

// Force only (no Jacobian), same arithmetic as spring_terms.
__device__ inline V3 spring_force(V3 x1, V3 x2, V3 v1, V3 v2, float L0, float Ks, float Kd) {
    V3 dp = {x1.x - x2.x, x1.y - x2.y, x1.z - x2.z};
    V3 dv = {v1.x - v2.x, v1.y - v2.y, v1.z - v2.z};
    float sq = dp.x * dp.x + dp.y * dp.y + dp.z * dp.z;
    float dist = sqrtf(sq);
    float leftTerm = -Ks * (dist - L0);
    float rightTerm = Kd * ((dv.x * dp.x + dv.y * dp.y + dv.z * dp.z) / dist);
    float inv = 1.0f / sqrtf(sq);
    float s = leftTerm + rightTerm;
    return {s * (dp.x * inv), s * (dp.y * inv), s * (dp.z * inv)};
}

// Wave tree: lane l += lane l+off for off = 32..1; lane 0 ends with the sum.
__device__ inline float wave_sum(float t) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) t += __shfl_down(t, off, 64);
    return t;
}

// Tile partial: wave sums added in wave order.  Result valid in thread 0.
template <int NT>
__device__ inline float tile_sum(float t, float* ws) {
    t = wave_sum(t);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) ws[w] = t;
    __syncthreads();
    float p = 0.0f;
    if (threadIdx.x == 0) {
        p = ws[0];
#pragma unroll
        for (int k = 1; k < NT / 64; k++) p = p + ws[k];
    }
    return p;
}

// Cross-tile sum in two halves so the loads can be issued early.  Lane l of
// wave 0 sums tiles l, l+64, l+128, ... in order, then the wave tree.
__device__ inline void partial_load(const float* part, int nt, float* v) {
#pragma unroll
    for (int j = 0; j < kPartialRegs; j++) {
        const int k = threadIdx.x + 64 * j;
        v[j] = k < nt ? part[k] : 0.0f;
    }
}
__device__ inline float partial_finish(const float* part, int nt, const float* v) {
    float s = 0.0f;
#pragma unroll
    for (int j = 0; j < kPartialRegs; j++) s += v[j];
    for (int k = threadIdx.x + 64 * kPartialRegs; k < nt; k += 64) s += part[k];
    return wave_sum(s);
}

// Sum of nt tile partials, returned to all threads (one-shot form).
__device__ inline float partial_sum(const float* part, int nt, float* bc) {
    if (threadIdx.x < 64) {
        float v[kPartialRegs];
        partial_load(part, nt, v);
        const float s = partial_finish(part, nt, v);
        if (threadIdx.x == 0) *bc = s;
    }
    __syncthreads();
    return *bc;
}

// Per-spring adjoint terms (oracle.c spring_adjoint, same expression order):
// derivative of h mu.f + h^2 mu.K v' + h mu.D dv with respect to e = x1-x2
// (ge), to v1 (gvv) and to ks (gk).
__device__ inline void spring_adjoint(V3 xa, V3 xb, V3 va, V3 vb, V3 ua, V3 ub, V3 wa, V3 wb,
                                      V3 pa, V3 pb, float L0, float ks, float Kd, float h, float hh,
                                      float* ge, float* gvv, float& gk) {
    const float e[3] = {xa.x - xb.x, xa.y - xb.y, xa.z - xb.z};
    const float dvel[3] = {va.x - vb.x, va.y - vb.y, va.z - vb.z};
    const float u[3] = {ua.x - ub.x, ua.y - ub.y, ua.z - ub.z};
    const float w[3] = {wa.x - wb.x, wa.y - wb.y, wa.z - wb.z};
    const float wp[3] = {pa.x - pb.x, pa.y - pb.y, pa.z - pb.z};
    const float l = sqrtf(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
    const float il = 1.0f / l;
    const float il2 = il * il, il3 = il2 * il, il4 = il2 * il2, il5 = il4 * il;
    const float ue = u[0] * e[0] + u[1] * e[1] + u[2] * e[2];
    const float we = w[0] * e[0] + w[1] * e[1] + w[2] * e[2];
    const float wpe = wp[0] * e[0] + wp[1] * e[1] + wp[2] * e[2];
    const float uw = u[0] * w[0] + u[1] * w[1] + u[2] * w[2];
    const float dve = dvel[0] * e[0] + dvel[1] * e[1] + dvel[2] * e[2];
    const float phi = -ks * (l - L0) + Kd * (dve * il);
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const float dphi = -ks * e[c] * il + Kd * (dvel[c] * il - dve * e[c] * il3);
        const float A = h * ((ue * il) * dphi + phi * (u[c] * il - ue * e[c] * il3));
        const float B = hh * ks * L0 *
                        (-uw * e[c] * il3 - (u[c] * we + w[c] * ue) * il3 + 3.0f * ue * we * e[c] * il5);
        const float Cc = h * Kd * ((u[c] * wpe + wp[c] * ue) * il2 - 2.0f * ue * wpe * e[c] * il4);
        ge[c] = A + B + Cc;
    }
    V3 f;
    float K[6], P[6];
    spring_terms(xa, xb, va, vb, L0, ks, Kd, f, K, P);
    const V3 ku = symmv(K, V3{u[0], u[1], u[2]});
    gvv[0] = h * Kd * ue * e[0] * il2 + hh * ku.x;
    gvv[1] = h * Kd * ue * e[1] * il2 + hh * ku.y;
    gvv[2] = h * Kd * ue * e[2] * il2 + hh * ku.z;
    gk = h * (ue * il) * (-(l - L0)) + hh * ((L0 * il - 1.0f) * uw - L0 * ue * we * il3);
}

}  // namespace dev
}  // namespace oc
