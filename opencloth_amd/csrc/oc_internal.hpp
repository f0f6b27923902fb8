// Internal declarations shared by the host C ABI and the HIP kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "opencloth.h"

namespace oc {

// A spring may join vertices at most kHalo rows and kHalo columns apart
// (the reference's bend springs span 2, IE:519-530).  The SpMV stages a
// kHalo-wide halo around each tile in LDS.
constexpr int kHalo = 2;
constexpr int kMaxPlanes = 16;
constexpr int kMaxAssemblyWidth = 32;
constexpr int kMaxPairWidth = 24;
constexpr uint16_t kEmptySlot = 0xFFFFu;

// ---- per-vertex slot encodings (16 bits) ----------------------------------
// assembly slot (one incident spring):
//   bits 0-3 di+8, 4-7 dj+8 of the other end relative to this vertex,
//   bit 8 this vertex is the spring's p2, bits 9-12 plane of the pair,
//   bit 13 last spring of its pair in this row, bit 14 this vertex owns the
//   pair block (it is the pair's lower index).
// pair slot (one incident pair, SpMV):
//   bits 0-3 di+8, 4-7 dj+8, bits 8-11 plane, bit 12 this vertex owns it.
__host__ __device__ inline int slot_di(uint16_t m) { return int(m & 15u) - 8; }
__host__ __device__ inline int slot_dj(uint16_t m) { return int((m >> 4) & 15u) - 8; }

// The reference's 12-neighbour stencil in the order its spring list numbers
// a vertex's pairs (IE:496-530: horizontal, vertical, the two shears of each
// quad, bend rows, bend columns): (di, dj) of the neighbour for position k.
// A cloth whose every vertex lists its pairs as a subsequence of this order
// (the reference's list and any sub-list of it) is described by a 12-bit
// mask per vertex; the SpMV then takes offsets, ownership and planes as
// constants.
constexpr int kStencil = 12;
__host__ __device__ constexpr int stencil_di(int k) {
    return k == 0 ? -1 : k == 1 ? 1 : k == 4 ? -1 : k == 5 ? 1 : k == 6 ? -1 : k == 7 ? 1
         : k == 8 ? -2 : k == 9 ? 2 : 0;
}
__host__ __device__ constexpr int stencil_dj(int k) {
    return k == 2 ? -1 : k == 3 ? 1 : k == 4 ? -1 : k == 5 ? -1 : k == 6 ? 1 : k == 7 ? 1
         : k == 10 ? -2 : k == 11 ? 2 : 0;
}
// The vertex owns (stores) the pair's block when the neighbour has the
// larger index: dj > 0, or dj == 0 and di > 0 (|di| < nx for dj != 0).
__host__ __device__ constexpr bool stencil_owner(int k) {
    return stencil_dj(k) > 0 || (stencil_dj(k) == 0 && stencil_di(k) > 0);
}

// The six stencil positions whose pair the vertex owns (its neighbour has
// the larger index), in the order the spring-once assembly visits them:
// +1 (k 1), +nx (k 3), +nx+1 (k 7), +nx-1 (k 6), +2 (k 9), +2nx (k 11).  The
// partner's position of each is kOwnedPartner (the opposite offset).
constexpr int kOwned = 6;
__host__ __device__ constexpr int owned_k(int o) {
    return o == 0 ? 1 : o == 1 ? 3 : o == 2 ? 7 : o == 3 ? 6 : o == 4 ? 9 : 11;
}
__host__ __device__ constexpr int owned_partner_k(int o) {
    return o == 0 ? 0 : o == 1 ? 2 : o == 2 ? 4 : o == 3 ? 5 : o == 4 ? 8 : 10;
}
// Owned direction of stencil position k: the pair's owned index, whether the
// vertex owns it (k = owned_k(o)) or is its partner (k = owned_partner_k(o)).
__host__ __device__ constexpr int owned_index(int k) {
    return k <= 1 ? 0 : k <= 3 ? 1 : k == 4 || k == 7 ? 2 : k <= 6 ? 3 : k <= 9 ? 4 : 5;
}
__host__ __device__ constexpr int own_dx(int o) { return stencil_di(owned_k(o)); }
__host__ __device__ constexpr int own_dy(int o) { return stencil_dj(owned_k(o)); }

// The strip of owned direction o around a TX x TY tile: owners outside the
// tile whose partner (owner + (dx, dy)) is inside, in tile coordinates.
// dx > 0, dy == 0: the dx columns left of the tile; dy > 0, dx == 0: the dy
// rows above it; diagonals: the row above (TX owners) and the column on the
// side the diagonal comes from (TY - 1 owners).
template <int TX, int TY>
__host__ __device__ constexpr int strip_count(int o) {
    return own_dy(o) == 0 ? own_dx(o) * TY : own_dx(o) == 0 ? own_dy(o) * TX : TX + TY - 1;
}
// The strips of all owned directions one after another: direction o's strip
// starts at column strip_base(o); strip_total() columns in all.
template <int TX, int TY>
__host__ __device__ constexpr int strip_base(int o) {
    return (o > 0 ? strip_count<TX, TY>(0) : 0) + (o > 1 ? strip_count<TX, TY>(1) : 0) +
           (o > 2 ? strip_count<TX, TY>(2) : 0) + (o > 3 ? strip_count<TX, TY>(3) : 0) +
           (o > 4 ? strip_count<TX, TY>(4) : 0) + (o > 5 ? strip_count<TX, TY>(5) : 0);
}
template <int TX, int TY>
__host__ __device__ constexpr int strip_total() {
    static_assert(kOwned == 6, "strip_base sums six directions");
    return strip_base<TX, TY>(kOwned);
}
// Direction o and strip index s of column t (0 <= t < strip_total()).
template <int TX, int TY>
__host__ __device__ inline void strip_of_column(int t, int& o, int& s) {
    o = 0;
    while (o + 1 < kOwned && t >= strip_base<TX, TY>(o + 1)) o++;
    s = t - strip_base<TX, TY>(o);
}
template <int TX, int TY>
__host__ __device__ constexpr int strip_max() {
    return (2 * TX > 2 * TY ? 2 * TX : 2 * TY) > TX + TY - 1 ? (2 * TX > 2 * TY ? 2 * TX : 2 * TY)
                                                             : TX + TY - 1;
}
// Strip index of the owner at tile coordinates (wx, wy) (outside the tile,
// its partner inside).
template <int TX, int TY>
__host__ __device__ inline int strip_index(int o, int wx, int wy) {
    const int dx = own_dx(o), dy = own_dy(o);
    if (dy == 0) return (wx + dx) * TY + wy;
    if (dx == 0) return (wy + dy) * TX + wx;
    if (wy == -1) return dx > 0 ? wx + 1 : wx - 1;
    return TX + wy;
}
// Owner of strip index s.
template <int TX, int TY>
__host__ __device__ inline void strip_owner(int o, int s, int& wx, int& wy) {
    const int dx = own_dx(o), dy = own_dy(o);
    if (dy == 0) {
        wx = s / TY - dx;
        wy = s % TY;
    } else if (dx == 0) {
        wx = s % TX;
        wy = s / TX - dy;
    } else if (s < TX) {
        wx = dx > 0 ? s - 1 : s + 1;
        wy = -1;
    } else {
        wx = dx > 0 ? -1 : TX;
        wy = s - TX;
    }
}

// Constants of the owner-plane copy: (multiplicity m, owned direction o,
// constant q = L0, ks, kd) of the pair owned by vertex i at
// cst[((m * kOwned + o) * 3 + q) * N + i].
constexpr int kCstPlanes = 2 * kOwned * 3;

// Grid topology in the device layout, built on the host from a spring list.
struct GridTopology {
    int nx = 0, ny = 0, N = 0, S = 0, P = 0;
    int n_planes = 0;
    int di[kMaxPlanes] = {}, dj[kMaxPlanes] = {};  // offset of the plane's upper vertex
    int wa = 0;                                    // assembly slots per vertex
    int wp = 0;                                    // pair slots per vertex
    std::vector<int32_t> ainc_spring;              // wa*N, -1 = empty (slot-major)
    std::vector<uint16_t> ainc_meta;               // wa*N
    std::vector<uint16_t> pair_meta;               // wp*N, kEmptySlot = empty
    bool stencil = false;                          // pairs follow the kStencil order
    std::vector<uint16_t> pair_mask;               // N, bit k: pair at stencil position k
    int stencil_plane[kStencil] = {};              // plane of the pair at position k
    // stencil assembly form: the stencil form with at most two springs per
    // pair.  asm_mask: bit k pair at stencil position k, bit 16 + k two
    // springs there.  spring_cst: per spring, the index of its L0 in the
    // owner-plane constants (kOwnedPlanes planes of N per multiplicity, see
    // owned_slot), ks and kd follow at +N and +2N.
    bool stencil_asm = false;
    std::vector<uint32_t> asm_mask;
    std::vector<int32_t> spring_cst;
    std::vector<int32_t> pair_a, pair_b, pair_plane;
    std::vector<int32_t> spring_pair;
};

// Builds the tables; returns OC_OK or an error code with a message in *err.
int build_grid_topology(int nx, int ny, const oc_spring* springs, int n_springs,
                        GridTopology* out, std::string* err);

// ---- device views ---------------------------------------------------------
// SoA vectors: component c of vertex i at ptr[c*N + i].
struct DevCloth {
    int nx, ny, N, ntx, nty, nt;
    int wa, wp;
    // what the CG launches read, first and together (a launch's scalar
    // loads of its arguments then touch few cache lines)
    float* dv;
    float* r;
    float* d0;
    float* d1;
    float* q;
    float* s;       // z = P^-1 r of the Jacobi-preconditioned solve
    float* pinv;    // Jacobi preconditioner 1/diag(A)
    const uint8_t* pin;
    float* blocks;  // n_planes*6*N
    const uint16_t* pair_m;
    const uint16_t* pair_mask;  // stencil form: bit k = pair at stencil position k
    float* del_part;  // (max_it+1)*nt
    float* sig_part;  // max_it*nt
    float* scal;      // max_it+1 reduced r.r values
    int* ctl;         // [0] done, [1] iterations, [2] status bits, [3] sticky status,
                      // [4] grid-barrier time-out, [6] grid-barrier counter
    int stencil;      // nonzero: pairs follow the kStencil order
    int stencil_plane[kStencil];
    int persistent;   // > 0: run all CG iterations in one cooperative launch of
                      // this many blocks (each serving several tiles)
    // the assembly's and the other kernels' fields
    float* x;
    float* v;
    float* b;
    float* f;
    const int32_t* ainc_s;
    const uint16_t* ainc_m;
    const float* L0;
    const float* ks;
    const float* kd;
    // fused update (assembly only): x + h(v + dv) and v + dv are written here
    float* xo;
    float* vo;
    // stencil assembly form (each spring evaluated once per tile)
    int stencil_asm;
    const uint32_t* asm_mask;  // per vertex: bit k pair at position k, bit 16 + k two springs
    const float* cst;          // owner-plane constants, kCstPlanes * N
    // single-reduction CG: second buffers of r and w (q), the two s buffers,
    // and the offset of the alpha values in scal
    float* r2;
    float* q2;
    float* s1;
    float* s2;
    int cg1_alpha;
};

// Implicit steps of a cloth handle, the body of oc_cloth_step (capi.cpp).
// partials: NULL for the handle's own CG partials and scalars, else a device
// buffer of cloth_partial_floats(c) floats used in their place.
size_t cloth_partial_floats(const ::oc_cloth_s* c);
int cloth_device(const ::oc_cloth_s* c);  // the device the handle was created on
int cloth_step(::oc_cloth_s* c, int nsteps, float* partials);

// Blocks of the persistent CG launch (all resident at once), 0 if unsupported.
int persistent_cg_grid(const DevCloth& c);

struct StepScalars {
    float mass, gx, gy, gz, h, hh, tol2;
    int max_it, tol_mode, warm, flags;
    float damp;  // velocity damping c (force c v on free vertices)
    float md;    // diagonal mass of A: mass - h c (mass when c = 0)
};

// Per-launch timing of the implicit step's kernels (oc_cloth_profile_steps).
// While a LaunchProfile is installed on the calling thread, each launcher of
// implicit.hip starts its kernel with hipExtLaunchKernelGGL and the next
// start/stop event pair of the profile, so every kernel's own duration is
// recorded inside the real launch sequence, with no event packets between
// the launches.  kind: 0 assembly, 1 CG iteration (part 1), 2 CG part 2,
// 3 other solver launches, 4 update.
struct LaunchProfile {
    hipEvent_t* start;
    hipEvent_t* stop;
    int* kind;
    int n, cap;
};
void set_launch_profile(LaunchProfile* p);  // NULL: off

// Kernel launchers (implicit.hip).  All return hipError_t of the launch.
hipError_t launch_forces(const DevCloth& c, const StepScalars& s, hipStream_t st);
// fused_update: first apply the previous step's update (v += dv, x += h v,
// ground clamp) to x and v, writing the result to xo and vo, then assemble
// at the updated state (the same bits as an update launch followed by the
// assembly).
hipError_t launch_assemble(const DevCloth& c, const StepScalars& s, hipStream_t st,
                           bool fused_update = false);
// Fill the owner-plane constants (stencil assembly form) from the per-spring
// L0, ks, kd: constants q in [q0, q0 + nq) (0 L0, 1 ks, 2 kd).
hipError_t launch_scatter_cst(const DevCloth& c, const int32_t* spring_cst, int S, int q0, int nq,
                              hipStream_t st);
// finish = false leaves the iteration count and status to launch_update(...,
// finish = true), the step's last launch.
hipError_t launch_cg(const DevCloth& c, const StepScalars& s, hipStream_t st, bool finish = true);
hipError_t launch_cg_spmv(const DevCloth& c, const StepScalars& s, int k, hipStream_t st);
hipError_t launch_cg_axpy(const DevCloth& c, const StepScalars& s, int k, hipStream_t st);
hipError_t launch_cg_finish(const DevCloth& c, const StepScalars& s, hipStream_t st);
hipError_t launch_update(const DevCloth& c, const StepScalars& s, hipStream_t st,
                         bool finish = false);
// Buffers of one adjoint step (SoA vectors of the cloth's size).
struct AdjointBufs {
    const float* gx_out;  // dL/dx after the step
    const float* gv_out;  // dL/dv after the step
    float* gvt;           // gv' + h gx'
    const float* mu;      // adjoint solve
    float* gx_in;         // dL/dx before the step
    float* gv_in;         // dL/dv before the step
    float* gks;           // dL/dks, accumulated (per spring)
};
hipError_t launch_adj_init(const DevCloth& c, const StepScalars& s, const AdjointBufs& a,
                           hipStream_t st);
hipError_t launch_adj_gather(const DevCloth& c, const StepScalars& s, const AdjointBufs& a,
                             hipStream_t st);

// Post-step passes on an implicit cloth (post.hip).
hipError_t launch_ellipsoid(float* x, float* w, int N, const oc_ellipsoid& e, int mode,
                            hipStream_t st);
hipError_t launch_provot(const DevCloth& c, float slack, hipStream_t st);

// Position Verlet (post.hip): positions rotate through three buffers.
struct DevVerlet {
    int nx, ny, N, wa;
    const float* x;   // current
    const float* xl;  // previous
    float* xn;        // next
    float* f;
    const uint8_t* pin;
    const int32_t* ainc_s;
    const uint16_t* ainc_m;
    const float* L0;
    const float* ks;
    const float* kd;
    // stencil assembly form: ownership mask per vertex and owner-plane
    // constants (the layout of DevCloth::cst); stencil_asm = 0 without it
    const uint32_t* asm_mask;
    const float* cst;
    int stencil_asm;
};
struct VerletScalars {
    float mass, gx, gy, gz, damping, dt;
    int flags;
};
hipError_t launch_verlet_forces(const DevVerlet& c, const VerletScalars& s, hipStream_t st);
hipError_t launch_verlet_step(const DevVerlet& c, const VerletScalars& s, hipStream_t st);

// Grid Verlet (grid_verlet.hip): the reference's GPU Verlet kernel on a
// width x height grid with a fixed 12-neighbour stencil.  Two position sets
// (x, x_last) are ping-ponged: a step reads one set and writes the other.
struct DevVgrid {
    int nx, ny, N;
    const float* x;   // current positions, SoA planes
    const float* xl;  // previous positions
    float* xo;        // next positions
    float* xlo;       // next previous positions (the current ones, clamped)
    const uint8_t* pin;
    const float* w;   // w of the position buffer, carried into the vertex buffer
    float* vbo;       // NULL or N float4 (16-byte aligned)
};
struct VgridScalars {
    float mass, gx, gy, gz, damping, dt;
    float rest[12];  // rest length of stencil position k
    float ks[3], kd[3];
};
hipError_t launch_vgrid(const DevVgrid& c, const VgridScalars& s, hipStream_t st);

// Position-based dynamics (pbd.hip).  Particle vectors are SoA planes;
// constraints are stored colour by colour (list order inside a colour).
constexpr int kPbdBlock = 256;     // particles per reduction block
constexpr int kPbdSums = 16;       // reduced scalars: [0..6] pass A, [7..15] pass B
struct DevPbd {
    int N, nblk;
    float* x;
    float* v;
    float* t;          // predicted positions tmp_X
    const float* w;    // inverse masses
    const int32_t* d_p1;
    const int32_t* d_p2;
    const float* d_rest;
    const float* d_kp;
    const int32_t* b_p1;
    const int32_t* b_p2;
    const int32_t* b_p3;
    const float* b_rest;
    const float* b_w;
    const float* b_kp;
    float* part;       // kPbdSums * nblk block partials
    float* sums;       // kPbdSums reduced values
    // list order (OC_FLAG_LIST_ORDER): the constraints are stored level by
    // level; level offsets (levels + 1 each) on the device
    const int32_t* d_lv;
    const int32_t* b_lv;
    int nd_lv, nb_lv;
};
struct PbdScalars {
    float mass, gx, gy, gz, dt, gd, kdamp, global_k, inv_dt;
    int flags;
};
hipError_t launch_pbd_predict(const DevPbd& c, const PbdScalars& s, hipStream_t st);
hipError_t launch_pbd_distance(const DevPbd& c, int first, int count, hipStream_t st);
hipError_t launch_pbd_bending(const DevPbd& c, const PbdScalars& s, int first, int count,
                              hipStream_t st);
hipError_t launch_pbd_ground(const DevPbd& c, hipStream_t st);
// The prediction with sequential centre-of-mass sums (list order): one
// workgroup, the reference's summation order.
hipError_t launch_pbd_predict_seq(const DevPbd& c, const PbdScalars& s, hipStream_t st);
// All solver iterations of the list-order schedule in one workgroup: per
// iteration the distance levels, the bending levels and (ground != 0) the
// ground clamp, with a workgroup barrier after each.
hipError_t launch_pbd_levels(const DevPbd& c, const PbdScalars& s, int iterations, int ground,
                             hipStream_t st);
hipError_t launch_pbd_integrate(const DevPbd& c, const PbdScalars& s, hipStream_t st);

// Tile shape of the implicit kernels (one workgroup per tile; a multiple of
// 64 threads).  The reduction order depends on it; oc_cloth_tile_shape
// reports it.  64x8 measured fastest at 512x512 among 32x8, 32x16, 32x32,
// 64x4, 64x8, 64x16, 128x4 and 128x8.
#ifndef OC_TILE_X
#define OC_TILE_X 64
#endif
#ifndef OC_TILE_Y
#define OC_TILE_Y 8
#endif
constexpr int kTileX = OC_TILE_X;
constexpr int kTileY = OC_TILE_Y;

// ---- batched cloths --------------------------------------------------------
// All cloths' vertices are concatenated (SoA planes of stride Ntot); slot
// tables are slot-major with stride Ntot and hold global spring ids.  A work
// item is a pack of `count` consecutive cloths (in the size-sorted order)
// served by T threads each.
struct DevBatch {
    int n_items, Ntot, wa, wp;
    float* x;
    float* v;
    float* dv;
    float* r;
    float* d;
    float* q;
    float* b;
    float* pinv;    // Jacobi preconditioner 1/diag(A)
    const uint8_t* pin;
    float* blocks;  // n_planes*6*Ntot
    const int32_t* ainc_s;
    const uint16_t* ainc_m;
    const uint16_t* pair_m;
    const uint16_t* pair_mask;  // stencil form (every cloth, one plane table)
    int stencil;
    int stencil_plane[kStencil];
    const float* L0;
    const float* ks;
    const float* kd;
    // the constants again per assembly slot, coalesced: plane q (0 L0, 1 ks,
    // 2 kd) of slot k of vertex i at slot_c[(q * wa + k) * Ntot + i]
    float* slot_c;
    const int* inst_voff;
    const int* inst_nx;
    const int* inst_n;
    const int* item_inst;   // per item: first entry in `order`
    const int* item_count;  // cloths in the item
    const int* item_T;      // threads per cloth
    const int* order;       // cloth ids sorted by size (descending)
    int* queue;
    int* inst_iters;
    int* inst_status;
    int* inst_sticky;  // OR of every step's status bits since the last query
    // stencil assembly form (every cloth has it): per vertex, bit k a pair at
    // stencil position k, bit 16 + k that pair's second spring
    const uint32_t* asm_mask;
    int stencil_asm;
    // spring-once assembly (stencil assembly form and the schedule's LDS
    // fits): each pair evaluated once by its lower vertex, the partner reads
    // the result from LDS; lds_floats of dynamic LDS per workgroup
    int asm_once;
    int lds_floats;
    // owner-plane constants of the stencil form: constant q (L0, ks, kd) of
    // the first spring of the pair vertex i owns in direction o at
    // cst[(o * 3 + q) * Ntot + i]; a partner reads its owner's entry
    float* cst;
};

// Spring-once batched assembly: owner results of direction o for the
// vertices before the current band live in a ring of ring_len(o, nx)
// entries (a power of two >= the direction's index offset); a cloth's rings
// take ring_floats(nx) floats (12 per entry).
__host__ __device__ constexpr int own_offset(int o, int nx) { return own_dy(o) * nx + own_dx(o); }
__host__ __device__ inline int pow2_at_least(int x) {
    int p = 1;
    while (p < x) p <<= 1;
    return p;
}
__host__ __device__ inline int ring_len(int o, int nx) { return pow2_at_least(own_offset(o, nx)); }
__host__ __device__ inline int ring_base(int o, int nx) {  // in entries
    int b = 0;
    for (int q = 0; q < o; q++) b += ring_len(q, nx);
    return b;
}
__host__ __device__ inline int ring_floats(int nx) { return 12 * ring_base(kOwned, nx); }
// Dynamic LDS of the spring-once batched assembly: 12 floats per thread for
// the current direction's results, then ring_floats of the widest cloth per
// cloth of the work item; schedules needing more than the device allows per
// workgroup (at most 160 KiB) keep the per-vertex form.
constexpr int kBatchLdsFloats = 160 * 1024 / 4;
// Batched CG: vertices per thread whose d, r, q and solution stay in LDS for
// a whole solve (12 floats each: 36 KiB per 256 threads, 144 KiB for 1024).
constexpr int kResidentPerThread = 3;
constexpr int batch_resident_floats(int block_threads) { return 12 * kResidentPerThread * block_threads; }

// Buffers of one batched adjoint step (SoA, stride Ntot).
struct BatchAdjBufs {
    const float* gx_out;
    const float* gv_out;
    float* gvt;
    const float* dvsol;  // the forward step's solution
    float* gx_in;
    float* gv_in;
    float* gks;
};

hipError_t launch_batch_adjoint(const DevBatch& c, const StepScalars& s, const BatchAdjBufs& a,
                                int block_threads, int grid, hipStream_t st);
hipError_t launch_batch_reset(const DevBatch& c, hipStream_t st);
// Fill slot_c planes q0 .. q0+nq-1 from L0/ks/kd through the slot tables.
hipError_t launch_batch_slot_constants(const DevBatch& c, int q0, int nq, hipStream_t st);
hipError_t launch_batch_step(const DevBatch& c, const StepScalars& s, int block_threads,
                             int grid, hipStream_t st);
// Fill the owner-plane constants (c.cst) from the slot-major copies.
hipError_t launch_batch_owner_constants(const DevBatch& c, hipStream_t st);

}  // namespace oc
