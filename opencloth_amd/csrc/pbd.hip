// HIP kernels of the position-based-dynamics cloth step (gfx950, wave64),
// after OpenCloth_PositionBasedDynamics/.../main.cpp (PBD) with the triangle
// bending constraint (PBD:57).
//
//   pass_a    : gravity, global damping, V += (F dt) W; block partials of
//               m X, m V, m                         (PBD:601-631)
//   reduce    : one wave per scalar: block partials lane-strided, 64-lane tree
//   pass_b    : R = X - Xcm; partials of cross(R, m V) and m [R]x[R]x^T
//   pass_c    : w = I^-1 L, centre-of-mass damping, prediction tmp_X
//                                                   (PBD:633-668)
//   distance  : one colour of distance constraints  (PBD:680-701)
//   bending   : one colour of bending constraints   (PBD:703-731)
//   ground    : tmp_X.y < 0 -> 0                    (PBD:839-845)
//   integrate : V = (tmp_X - X)/dt, X = tmp_X       (PBD:670-678)
//
// The constraints of one colour share no particle, so a thread per
// constraint reproduces the sequential projection of that colour exactly;
// colours run in order.  The sums of the damping step use one fixed tree
// (64-lane shuffle trees, the 4 waves of a block in order, then the block
// partials lane-strided and one more tree), so results are deterministic
// and equal the CPU restatement with the same order bit for bit.
// Arithmetic follows the reference's GLM expression order (built with
// -ffp-contract=off).
#include "oc_device.hpp"

namespace oc {

namespace {

using namespace dev;

constexpr int NB = kPbdBlock;
constexpr int NWB = NB / 64;

__device__ inline V3 cross3(V3 x, V3 y) {
    return {x.y * y.z - y.y * x.z, x.z * y.x - y.z * x.x, x.x * y.y - y.x * x.y};
}

// Velocity update of particle i (gravity, global damping, V += (F dt) W;
// PBD:601-631); returns the updated velocity.
__device__ inline V3 external_velocity(const DevPbd& c, const PbdScalars& s, int i) {
    const float wi = c.w[i];
    V3 v = ld3(c.v, c.N, i);
    const V3 F = wi > 0.0f ? V3{0.0f + s.gx, 0.0f + s.gy, 0.0f + s.gz} : V3{0.0f, 0.0f, 0.0f};
    v = {v.x * s.gd, v.y * s.gd, v.z * s.gd};
    v = {v.x + (F.x * s.dt) * wi, v.y + (F.y * s.dt) * wi, v.z + (F.z * s.dt) * wi};
    st3(c.v, c.N, i, v);
    return v;
}

// Block partials of Q values per thread: wave trees, waves in order.
template <int Q>
__device__ inline void block_partials(const float (&t)[Q], float* part, int nblk, int first) {
    __shared__ float ws[Q][NWB];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < Q; q++) {
        const float s = wave_sum(t[q]);
        if (lane == 0) ws[q][wv] = s;
    }
    __syncthreads();
    if (threadIdx.x < Q) {
        const int q = threadIdx.x;
        float p = ws[q][0];
#pragma unroll
        for (int k = 1; k < NWB; k++) p = p + ws[q][k];
        part[(size_t)(first + q) * nblk + blockIdx.x] = p;
    }
}

__global__ __launch_bounds__(NB) void k_pbd_pass_a(DevPbd c, PbdScalars s) {
    const int i = blockIdx.x * NB + threadIdx.x;
    float t[7] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    if (i < c.N) {
        const V3 x = ld3(c.x, c.N, i);
        const V3 v = external_velocity(c, s, i);
        t[0] = x.x * s.mass;
        t[1] = x.y * s.mass;
        t[2] = x.z * s.mass;
        t[3] = v.x * s.mass;
        t[4] = v.y * s.mass;
        t[5] = v.z * s.mass;
        t[6] = s.mass;
    }
    block_partials<7>(t, c.part, c.nblk, 0);
}

// sums[first + q] = lane-strided sum of the block partials, then a 64-lane
// tree; one wave per value.
__global__ void k_pbd_reduce(DevPbd c, int first) {
    const int q = first + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    const float* p = c.part + (size_t)q * c.nblk;
    float s = 0.0f;
    for (int b = lane; b < c.nblk; b += 64) s += p[b];
    s = wave_sum(s);
    if (lane == 0) c.sums[q] = s;
}

__device__ inline V3 centre(const float* sums, int first) {
    return {sums[first] / sums[6], sums[first + 1] / sums[6], sums[first + 2] / sums[6]};
}

__global__ __launch_bounds__(NB) void k_pbd_pass_b(DevPbd c, PbdScalars s) {
    const int i = blockIdx.x * NB + threadIdx.x;
    float t[9] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    if (i < c.N) {
        const V3 xcm = centre(c.sums, 0);
        const V3 x = ld3(c.x, c.N, i), v = ld3(c.v, c.N, i);
        const V3 R = {x.x - xcm.x, x.y - xcm.y, x.z - xcm.z};
        const V3 l = cross3(R, V3{s.mass * v.x, s.mass * v.y, s.mass * v.z});
        t[0] = l.x;
        t[1] = l.y;
        t[2] = l.z;
        // m [R]x [R]x^T reduces exactly to these entries (xx yy zz xy xz yz)
        t[3] = (R.y * R.y + R.z * R.z) * s.mass;
        t[4] = (R.x * R.x + R.z * R.z) * s.mass;
        t[5] = (R.x * R.x + R.y * R.y) * s.mass;
        t[6] = -(R.x * R.y) * s.mass;
        t[7] = -(R.x * R.z) * s.mass;
        t[8] = -(R.y * R.z) * s.mass;
    }
    block_partials<9>(t, c.part, c.nblk, 7);
}

// Centre-of-mass damping and prediction of particle i (PBD:633-668) from the
// sums: I6 = I (xx yy zz xy xz yz), L = sum cross(R, m V), the centres xcm,
// vcm.  I is inverted as glm::inverse does (every thread alike).
__device__ inline void damp_predict(const DevPbd& c, const PbdScalars& s, int i, const float* I6,
                                    const float* Lv, V3 xcm, V3 vcm) {
    const float xx = I6[0], yy = I6[1], zz = I6[2], xy = I6[3], xz = I6[4], yz = I6[5];
    const float M[3][3] = {{xx, xy, xz}, {xy, yy, yz}, {xz, yz, zz}};
    const float det = +M[0][0] * (M[1][1] * M[2][2] - M[2][1] * M[1][2]) -
                      M[1][0] * (M[0][1] * M[2][2] - M[2][1] * M[0][2]) +
                      M[2][0] * (M[0][1] * M[1][2] - M[1][1] * M[0][2]);
    float Inv[3][3];
    Inv[0][0] = +(M[1][1] * M[2][2] - M[2][1] * M[1][2]);
    Inv[1][0] = -(M[1][0] * M[2][2] - M[2][0] * M[1][2]);
    Inv[2][0] = +(M[1][0] * M[2][1] - M[2][0] * M[1][1]);
    Inv[0][1] = -(M[0][1] * M[2][2] - M[2][1] * M[0][2]);
    Inv[1][1] = +(M[0][0] * M[2][2] - M[2][0] * M[0][2]);
    Inv[2][1] = -(M[0][0] * M[2][1] - M[2][0] * M[0][1]);
    Inv[0][2] = +(M[0][1] * M[1][2] - M[1][1] * M[0][2]);
    Inv[1][2] = -(M[0][0] * M[1][2] - M[1][0] * M[0][2]);
    Inv[2][2] = +(M[0][0] * M[1][1] - M[1][0] * M[0][1]);
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = 0; b < 3; b++) Inv[a][b] = Inv[a][b] / det;
    const float L[3] = {Lv[0], Lv[1], Lv[2]};
    V3 w;
    w.x = Inv[0][0] * L[0] + Inv[1][0] * L[1] + Inv[2][0] * L[2];
    w.y = Inv[0][1] * L[0] + Inv[1][1] * L[1] + Inv[2][1] * L[2];
    w.z = Inv[0][2] * L[0] + Inv[1][2] * L[1] + Inv[2][2] * L[2];

    const V3 x = ld3(c.x, c.N, i);
    V3 v = ld3(c.v, c.N, i);
    const V3 R = {x.x - xcm.x, x.y - xcm.y, x.z - xcm.z};
    const V3 wr = cross3(w, R);
    const V3 del = {(vcm.x + wr.x) - v.x, (vcm.y + wr.y) - v.y, (vcm.z + wr.z) - v.z};
    v = {v.x + s.kdamp * del.x, v.y + s.kdamp * del.y, v.z + s.kdamp * del.z};
    st3(c.v, c.N, i, v);
    const V3 t = c.w[i] <= 0.0f ? x : V3{x.x + v.x * s.dt, x.y + v.y * s.dt, x.z + v.z * s.dt};
    st3(c.t, c.N, i, t);
}

__global__ __launch_bounds__(NB) void k_pbd_pass_c(DevPbd c, PbdScalars s) {
    const int i = blockIdx.x * NB + threadIdx.x;
    if (i >= c.N) return;
    const float* S = c.sums;
    // I = identity + sum (symmetric, column-major m[c][r])
    const float I6[6] = {1.0f + S[10], 1.0f + S[11], 1.0f + S[12], 0.0f + S[13], 0.0f + S[14], 0.0f + S[15]};
    damp_predict(c, s, i, I6, S + 7, centre(S, 0), centre(S, 3));
}

// The prediction with the reference's sequential sums (list order): one
// workgroup; thread 0 adds the centre-of-mass sums in particle order and
// accumulates I onto the identity, as PBD:612-650 does; every particle's
// update runs in parallel.
constexpr int kSeqThreads = 1024;
__global__ __launch_bounds__(kSeqThreads) void k_pbd_predict_seq(DevPbd c, PbdScalars s) {
    __shared__ float sx[3], sv[3], sI[6], sL[3];
    for (int i = threadIdx.x; i < c.N; i += kSeqThreads) external_velocity(c, s, i);
    __syncthreads();
    if (threadIdx.x == 0) {
        float xs[3] = {0.0f, 0.0f, 0.0f}, vs[3] = {0.0f, 0.0f, 0.0f}, ms = 0.0f;
        for (int i = 0; i < c.N; i++) {
            const V3 x = ld3(c.x, c.N, i), v = ld3(c.v, c.N, i);
            xs[0] += x.x * s.mass;
            xs[1] += x.y * s.mass;
            xs[2] += x.z * s.mass;
            vs[0] += v.x * s.mass;
            vs[1] += v.y * s.mass;
            vs[2] += v.z * s.mass;
            ms += s.mass;
        }
        const V3 xcm = {xs[0] / ms, xs[1] / ms, xs[2] / ms};
        float L[3] = {0.0f, 0.0f, 0.0f}, I6[6] = {1.0f, 1.0f, 1.0f, 0.0f, 0.0f, 0.0f};
        for (int i = 0; i < c.N; i++) {
            const V3 x = ld3(c.x, c.N, i), v = ld3(c.v, c.N, i);
            const V3 R = {x.x - xcm.x, x.y - xcm.y, x.z - xcm.z};
            const V3 l = cross3(R, V3{s.mass * v.x, s.mass * v.y, s.mass * v.z});
            L[0] += l.x;
            L[1] += l.y;
            L[2] += l.z;
            I6[0] += (R.y * R.y + R.z * R.z) * s.mass;
            I6[1] += (R.x * R.x + R.z * R.z) * s.mass;
            I6[2] += (R.x * R.x + R.y * R.y) * s.mass;
            I6[3] += -(R.x * R.y) * s.mass;
            I6[4] += -(R.x * R.z) * s.mass;
            I6[5] += -(R.y * R.z) * s.mass;
        }
        for (int k = 0; k < 3; k++) {
            sx[k] = k == 0 ? xcm.x : k == 1 ? xcm.y : xcm.z;
            sv[k] = vs[k] / ms;
            sL[k] = L[k];
        }
        for (int k = 0; k < 6; k++) sI[k] = I6[k];
    }
    __syncthreads();
    for (int i = threadIdx.x; i < c.N; i += kSeqThreads)
        damp_predict(c, s, i, sI, sL, V3{sx[0], sx[1], sx[2]}, V3{sv[0], sv[1], sv[2]});
}

// UpdateDistanceConstraint (PBD:680-701) of stored constraint k.
__device__ inline void project_distance(const DevPbd& c, int k) {
    const int a = c.d_p1[k], b = c.d_p2[k];
    const float rest = c.d_rest[k], kp = c.d_kp[k];
    const float w1 = c.w[a], w2 = c.w[b];
    const V3 ta = ld3(c.t, c.N, a), tb = ld3(c.t, c.N, b);
    const V3 dir = {ta.x - tb.x, ta.y - tb.y, ta.z - tb.z};
    const float len = sqrtf(dir.x * dir.x + dir.y * dir.y + dir.z * dir.z);
    if (len <= 0.0000001f) return;
    const float inv_mass = w1 + w2;
    if (inv_mass <= 0.0000001f) return;
    const float sc = (1.0f / inv_mass) * (len - rest);
    const V3 dP = {(sc * (dir.x / len)) * kp, (sc * (dir.y / len)) * kp, (sc * (dir.z / len)) * kp};
    if (w1 > 0.0f) st3(c.t, c.N, a, V3{ta.x - dP.x * w1, ta.y - dP.y * w1, ta.z - dP.z * w1});
    if (w2 > 0.0f) st3(c.t, c.N, b, V3{tb.x + dP.x * w2, tb.y + dP.y * w2, tb.z + dP.z * w2});
}

__global__ __launch_bounds__(256) void k_pbd_distance(DevPbd c, int first, int count) {
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j < count) project_distance(c, first + j);
}

// UpdateBendingConstraint (PBD:703-731) of stored constraint k.
__device__ inline void project_bending(const DevPbd& c, const PbdScalars& s, int k) {
    const int a = c.b_p1[k], b = c.b_p2[k], d = c.b_p3[k];
    const float rest = c.b_rest[k], wsum = c.b_w[k], kp = c.b_kp[k];
    const float wa = c.w[a], wb = c.w[b], wd = c.w[d];
    const V3 ta = ld3(c.t, c.N, a), tb = ld3(c.t, c.N, b), td = ld3(c.t, c.N, d);
    const V3 ctr = {0.3333f * ((ta.x + tb.x) + td.x), 0.3333f * ((ta.y + tb.y) + td.y),
                    0.3333f * ((ta.z + tb.z) + td.z)};
    const V3 dc = {td.x - ctr.x, td.y - ctr.y, td.z - ctr.z};
    const float dist = sqrtf(dc.x * dc.x + dc.y * dc.y + dc.z * dc.z);
    const float diff = 1.0f - ((s.global_k + rest) / dist);
    const V3 f = {dc.x * diff, dc.y * diff, dc.z * diff};
    const float sa = kp * ((2.0f * wa) / wsum);
    const float sb = kp * ((2.0f * wb) / wsum);
    const float sd = -kp * ((4.0f * wd) / wsum);
    if (wa > 0.0f) st3(c.t, c.N, a, V3{ta.x + sa * f.x, ta.y + sa * f.y, ta.z + sa * f.z});
    if (wb > 0.0f) st3(c.t, c.N, b, V3{tb.x + sb * f.x, tb.y + sb * f.y, tb.z + sb * f.z});
    if (wd > 0.0f) st3(c.t, c.N, d, V3{td.x + sd * f.x, td.y + sd * f.y, td.z + sd * f.z});
}

__global__ __launch_bounds__(256) void k_pbd_bending(DevPbd c, PbdScalars s, int first, int count) {
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j < count) project_bending(c, s, first + j);
}

// The list-order schedule in one workgroup.  The constraints of a level
// touch disjoint particles and every earlier constraint sharing a particle
// with one of them sits in an earlier level, so projecting the levels in
// order, a level's constraints in parallel, gives the sequential list order's
// results (UpdateInternalConstraints, PBD:876-889).  The workgroup barrier
// after each level orders the global memory accesses of the one workgroup.
constexpr int kLevelThreads = 1024;
__global__ __launch_bounds__(kLevelThreads) void k_pbd_levels(DevPbd c, PbdScalars s, int iterations,
                                                             int ground) {
    for (int it = 0; it < iterations; it++) {
        for (int L = 0; L < c.nd_lv; L++) {
            const int first = c.d_lv[L], last = c.d_lv[L + 1];
            for (int k = first + (int)threadIdx.x; k < last; k += kLevelThreads) project_distance(c, k);
            __syncthreads();
        }
        for (int L = 0; L < c.nb_lv; L++) {
            const int first = c.b_lv[L], last = c.b_lv[L + 1];
            for (int k = first + (int)threadIdx.x; k < last; k += kLevelThreads) project_bending(c, s, k);
            __syncthreads();
        }
        if (ground) {  // GroundCollision (PBD:839-845)
            for (int i = threadIdx.x; i < c.N; i += kLevelThreads) {
                float* ty = c.t + c.N + i;
                if (*ty < 0.0f) *ty = 0.0f;
            }
            __syncthreads();
        }
    }
}

__global__ __launch_bounds__(256) void k_pbd_ground(DevPbd c) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= c.N) return;
    float* ty = c.t + c.N + i;
    if (*ty < 0.0f) *ty = 0.0f;
}

__global__ __launch_bounds__(256) void k_pbd_integrate(DevPbd c, PbdScalars s) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= c.N) return;
    const V3 t = ld3(c.t, c.N, i), x = ld3(c.x, c.N, i);
    st3(c.v, c.N, i, V3{(t.x - x.x) * s.inv_dt, (t.y - x.y) * s.inv_dt, (t.z - x.z) * s.inv_dt});
    st3(c.x, c.N, i, t);
}

}  // namespace

hipError_t launch_pbd_predict(const DevPbd& c, const PbdScalars& s, hipStream_t st) {
    hipLaunchKernelGGL(k_pbd_pass_a, dim3(c.nblk), dim3(NB), 0, st, c, s);
    hipLaunchKernelGGL(k_pbd_reduce, dim3(1), dim3(64 * 7), 0, st, c, 0);
    hipLaunchKernelGGL(k_pbd_pass_b, dim3(c.nblk), dim3(NB), 0, st, c, s);
    hipLaunchKernelGGL(k_pbd_reduce, dim3(1), dim3(64 * 9), 0, st, c, 7);
    hipLaunchKernelGGL(k_pbd_pass_c, dim3(c.nblk), dim3(NB), 0, st, c, s);
    return hipGetLastError();
}

hipError_t launch_pbd_distance(const DevPbd& c, int first, int count, hipStream_t st) {
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_pbd_distance, dim3((count + 255) / 256), dim3(256), 0, st, c, first, count);
    return hipGetLastError();
}

hipError_t launch_pbd_bending(const DevPbd& c, const PbdScalars& s, int first, int count,
                              hipStream_t st) {
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_pbd_bending, dim3((count + 255) / 256), dim3(256), 0, st, c, s, first,
                       count);
    return hipGetLastError();
}

hipError_t launch_pbd_predict_seq(const DevPbd& c, const PbdScalars& s, hipStream_t st) {
    hipLaunchKernelGGL(k_pbd_predict_seq, dim3(1), dim3(kSeqThreads), 0, st, c, s);
    return hipGetLastError();
}

hipError_t launch_pbd_levels(const DevPbd& c, const PbdScalars& s, int iterations, int ground,
                             hipStream_t st) {
    hipLaunchKernelGGL(k_pbd_levels, dim3(1), dim3(kLevelThreads), 0, st, c, s, iterations, ground);
    return hipGetLastError();
}

hipError_t launch_pbd_ground(const DevPbd& c, hipStream_t st) {
    hipLaunchKernelGGL(k_pbd_ground, dim3((c.N + 255) / 256), dim3(256), 0, st, c);
    return hipGetLastError();
}

hipError_t launch_pbd_integrate(const DevPbd& c, const PbdScalars& s, hipStream_t st) {
    hipLaunchKernelGGL(k_pbd_integrate, dim3((c.N + 255) / 256), dim3(256), 0, st, c, s);
    return hipGetLastError();
}

}  // namespace oc
