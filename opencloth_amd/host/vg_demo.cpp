// The GPU Verlet demo of OpenCloth (CUM, its CUDA path) as a C++ program
// without the window: InitGL's float4 grid, then one VerletCUDA per frame,
// through the C++ host layer (opencloth.hpp) over the C ABI.
//
//   vg_demo --steps N --out FILE     N steps; FILE gets X and X_last
//   vg_demo --scene-only --out FILE  the scene only (no GPU call): X and the
//                                    parameters
//
// FILE layout (little-endian): int32 n_vertices, int32 texsize_x,
// int32 texsize_y, then X (4 floats per vertex), and either X_last or, with
// --scene-only, the parameters (oc_vgrid_params, 14 floats).
//
// CUM = OpenCloth_Verlet_CUDA_GLSL_OPENCL_CPU/OpenCloth_Verlet_CUDA/main.cpp,
// CUK = .../verlet_kernel.cu (paths relative to the reference repository).
#include "demo_common.hpp"

namespace {

struct vec4 {  // the layout of glm::vec4
    float x, y, z, w;
};

// Grid and tunables (CUM:56-102) and the kernel's constants (CUK:8-10, 123).
const int numX = 63, numY = 63;
const int sizeX = 4, sizeY = 4;
const float DEFAULT_DAMPING = -0.0125f;
const float mass = 1.0f;
const float timeStep = 1.0f / 60.0f;

}  // namespace

int main(int argc, char** argv) {
    demo::Args a;
    if (!demo::parse(argc, argv, &a)) return demo::usage(argv[0]);
    const int u = numX + 1, v = numY + 1;
    // CUM:438-444: vec4(((i/(u-1))*2-1)*halfsize, sizeX+1, (j/(v-1))*sizeY, 1)
    const std::vector<demo::vec3> g = demo::grid(numX, numY, (float)sizeX);
    std::vector<vec4> X(g.size());
    for (size_t k = 0; k < g.size(); k++) X[k] = vec4{g[k].x, g[k].y, g[k].z, 1.0f};
    std::vector<vec4> X_last = X;

    oc_vgrid_params p = {};
    p.mass = mass;
    p.gravity[1] = -0.00981f;  // CUK:123
    p.damping = DEFAULT_DAMPING;
    p.dt = timeStep;
    p.spacing[0] = float(sizeX) / numX;  // inv_cloth_size (CUM:601)
    p.spacing[1] = float(sizeY) / numY;
    p.ks[0] = p.ks[1] = 50.75f;  // KsStruct, KsShear
    p.ks[2] = 50.95f;            // KsBend
    p.kd[0] = p.kd[1] = p.kd[2] = -0.25f;
    const int32_t head[3] = {(int32_t)X.size(), u, v};
    if (a.scene_only)
        return demo::write_file(a.out, {{head, sizeof head}, {X.data(), X.size() * sizeof(vec4)}, {&p, sizeof p}})
                   ? 0 : 1;
    try {
        opencloth::GridVerlet<vec4> grid(u, v, X, X_last, p);  // InitCUDA + UploadCUDA (CUM:513, 602)
        for (int s = 0; s < a.steps; s++) grid.Verlet();      // VerletCUDA (CUM:605)
        grid.Download();
    } catch (const opencloth::Error& e) {
        fprintf(stderr, "vg_demo: %s (status %d)\n", e.what(), e.code());
        return 1;
    }
    if (!demo::write_file(a.out, {{head, sizeof head}, {X.data(), X.size() * sizeof(vec4)},
                                  {X_last.data(), X_last.size() * sizeof(vec4)}}))
        return 1;
    printf("vg_demo: %d steps on a %dx%d grid\n", a.steps, u, v);
    return 0;
}
