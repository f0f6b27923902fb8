// Grid Verlet: the reference's GPU Verlet kernel
// (OpenCloth_Verlet_CUDA_GLSL_OPENCL_CPU/OpenCloth_Verlet_CUDA/verlet_kernel.cu:70-197,
// "CUK"; its OpenCL twin verlet_kernel.cl:67-153 computes the same) as one
// fused gfx950 kernel per step.
//
// Per vertex (CUK:119-195):
//   v = (x - x_last)/dt;  F = gravity*mass + v*damping
//   for k = 0..11 in the stencil order of getNextNeighbor (CUK:14-64), when
//   the neighbour lies on the grid: F += spring force with rest length
//   |(dx*sx, dy*sy)| (CUK:164) and the class's ks, kd
//   a = F/mass, or 0 on pinned vertices (the kernel zeroes their mass,
//   CUK:126-127); y < 0 of the current position clamped to 0 (CUK:183-184)
//   x' = (2x - x_last) + (a dt) dt;  x_last' = x (clamped)
//
// Layout: SoA planes; tiles of 32x8 vertices with a 2-wide halo of x and v
// staged in LDS, so every neighbour's velocity is computed once per tile
// instead of once per spring end (the same division, so the same bits).
// The spring force uses spring_force's expression order (normalize as
// v * (1/sqrt(v.v))), built with -ffp-contract=off.
#include "oc_device.hpp"

namespace oc {

namespace {

using namespace dev;

// Column and row offset of stencil position k (CUK:25-63): structural
// (1,0) (0,-1) (-1,0) (0,1), shear (1,-1) (-1,-1) (-1,1) (1,1), bend (2,0)
// (0,-2) (-2,0) (0,2).
__device__ constexpr int vg_dx(int k) {
    return k == 0 ? 1 : k == 2 ? -1 : k == 4 ? 1 : k == 5 ? -1 : k == 6 ? -1 : k == 7 ? 1
         : k == 8 ? 2 : k == 10 ? -2 : 0;
}
__device__ constexpr int vg_dy(int k) {
    return k == 1 ? -1 : k == 3 ? 1 : k == 4 ? -1 : k == 5 ? -1 : k == 6 ? 1 : k == 7 ? 1
         : k == 9 ? -2 : k == 11 ? 2 : 0;
}

template <int TX, int TY>
struct VgridShared {
    static constexpr int WX = TX + 4, WY = TY + 4;
    float wx[3][WY][WX];
    float wv[3][WY][WX];
};

template <int TX, int TY, bool VBO>
__global__ __launch_bounds__(TX* TY) void k_vgrid(DevVgrid c, VgridScalars s) {
    constexpr int NT = TX * TY, H = 2;
    constexpr int WX = TX + 2 * H, WY = TY + 2 * H;
    constexpr int NWIN = (WX * WY + NT - 1) / NT;
    __shared__ VgridShared<TX, TY> sh;
    const int ntx = (c.nx + TX - 1) / TX, nt = ntx * ((c.ny + TY - 1) / TY);
    const int tile = tile_of_block(blockIdx.x, nt);
    const int bx = tile % ntx, by = tile / ntx;
    const int lx = threadIdx.x % TX, ly = threadIdx.x / TX;
    const int gi = bx * TX + lx, gj = by * TY + ly;
    const int N = c.N;
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
    if (gi >= c.nx || gj >= c.ny) return;
    const int i = gj * c.nx + gi;
    const int cy = ly + H, cx = lx + H;
    const V3 xi = {sh.wx[0][cy][cx], sh.wx[1][cy][cx], sh.wx[2][cy][cx]};
    const V3 vi = {sh.wv[0][cy][cx], sh.wv[1][cy][cx], sh.wv[2][cy][cx]};
    V3 F = {s.gx * s.mass + vi.x * s.damping, s.gy * s.mass + vi.y * s.damping,
            s.gz * s.mass + vi.z * s.damping};
#pragma unroll
    for (int k = 0; k < 12; k++) {
        const int dx = vg_dx(k), dy = vg_dy(k), cls = k >> 2;
        if (gi + dx < 0 || gi + dx >= c.nx || gj + dy < 0 || gj + dy >= c.ny) continue;
        const int wy = cy + dy, wx = cx + dx;
        const V3 xj = {sh.wx[0][wy][wx], sh.wx[1][wy][wx], sh.wx[2][wy][wx]};
        const V3 vj = {sh.wv[0][wy][wx], sh.wv[1][wy][wx], sh.wv[2][wy][wx]};
        F = add3(F, spring_force(xi, xj, vi, vj, s.rest[k], s.ks[cls], s.kd[cls]));
    }
    const bool pinned = c.pin[i] != 0;
    const V3 a = pinned ? V3{0.0f, 0.0f, 0.0f} : V3{F.x / s.mass, F.y / s.mass, F.z / s.mass};
    V3 xc = xi;
    if (xc.y < 0.0f) xc.y = 0.0f;
    const V3 xl = ld3(c.xl, N, i);
    const V3 xn = {(xc.x * 2.0f - xl.x) + (a.x * s.dt) * s.dt, (xc.y * 2.0f - xl.y) + (a.y * s.dt) * s.dt,
                   (xc.z * 2.0f - xl.z) + (a.z * s.dt) * s.dt};
    st3(c.xo, N, i, xn);
    st3(c.xlo, N, i, xc);
    if (VBO) reinterpret_cast<float4*>(c.vbo)[i] = make_float4(xn.x, xn.y, xn.z, c.w[i]);
}

// Same tile shape as the spring-list Verlet kernel, which measured 32x8
// fastest at 512x512.
constexpr int kVgridTileX = 32, kVgridTileY = 8;

}  // namespace

hipError_t launch_vgrid(const DevVgrid& c, const VgridScalars& s, hipStream_t st) {
    constexpr int TX = kVgridTileX, TY = kVgridTileY;
    const int nt = ((c.nx + TX - 1) / TX) * ((c.ny + TY - 1) / TY);
    if (c.vbo)
        hipLaunchKernelGGL((k_vgrid<TX, TY, true>), dim3(nt), dim3(TX * TY), 0, st, c, s);
    else
        hipLaunchKernelGGL((k_vgrid<TX, TY, false>), dim3(nt), dim3(TX * TY), 0, st, c, s);
    return hipGetLastError();
}

}  // namespace oc
