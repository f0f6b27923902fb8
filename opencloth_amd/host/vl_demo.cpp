// The position-Verlet demo of OpenCloth (VL) as a C++ program without the
// window: InitGL's scene, then StepPhysics in a loop, through the C++ host
// layer (opencloth.hpp) over the C ABI.
//
//   vl_demo --steps N --out FILE     N StepPhysics calls; FILE gets X and X_last
//   vl_demo --scene-only --out FILE  the scene only (no GPU call): X,
//                                    springs, pins and the collider
//
// FILE layout (little-endian): int32 n_vertices, int32 n_springs, then
// X (3 floats per vertex), and either X_last or, with --scene-only, the
// springs (24-byte records), the pins (1 byte per vertex) and the collider
// (oc_ellipsoid, 36 floats).
//
// VL = OpenCloth_Verlet/OpenCloth_Verlet/main.cpp (paths relative to the
// reference repository).
#include "demo_common.hpp"

using demo::vec3;

namespace {

// Grid and tunables (VL:59-106).
const int numX = 20, numY = 20;
const float fullsize = 4.0f;
const float Ks[3] = {50.75f, 50.75f, 50.95f};
const float Kd[3] = {-0.25f, -0.25f, -0.25f};
const float DEFAULT_DAMPING = -0.0125f;
const float mass = 1.0f;
const vec3 gravity = {0.0f, -0.00981f, 0.0f};  // an acceleration (the force is gravity*mass)
const float timeStep = 1 / 60.0f;

}  // namespace

int main(int argc, char** argv) {
    demo::Args a;
    if (!demo::parse(argc, argv, &a)) return demo::usage(argv[0]);
    std::vector<vec3> X = demo::grid(numX, numY, fullsize);  // VL:240-260
    std::vector<vec3> X_last = X;                             // at rest
    const std::vector<oc_spring> springs = demo::springs(numX, numY, X, Ks, Kd);
    std::vector<uint8_t> pinned(X.size(), 0);
    pinned[0] = pinned[numX] = 1;
    const oc_ellipsoid ellipsoid = demo::ellipsoid();
    const int32_t head[2] = {(int32_t)X.size(), (int32_t)springs.size()};
    if (a.scene_only)
        return demo::write_file(a.out, {{head, sizeof head}, {X.data(), X.size() * sizeof(vec3)},
                                        {springs.data(), springs.size() * sizeof(oc_spring)},
                                        {pinned.data(), pinned.size()}, {&ellipsoid, sizeof ellipsoid}})
                   ? 0 : 1;

    oc_verlet_params p = {};
    p.mass = mass;
    p.gravity[0] = gravity.x;
    p.gravity[1] = gravity.y;
    p.gravity[2] = gravity.z;
    p.damping = DEFAULT_DAMPING;
    p.dt = timeStep;
    p.flags = OC_FLAG_GROUND_CLAMP;  // VL:440-442
    try {
        opencloth::Verlet<vec3> cloth(numX, numY, X, X_last, springs, pinned, p);
        for (int s = 0; s < a.steps; s++) cloth.StepPhysics(timeStep, &ellipsoid);  // VL:557-562
    } catch (const opencloth::Error& e) {
        fprintf(stderr, "vl_demo: %s (status %d)\n", e.what(), e.code());
        return 1;
    }
    if (!demo::write_file(a.out, {{head, sizeof head}, {X.data(), X.size() * sizeof(vec3)},
                                  {X_last.data(), X_last.size() * sizeof(vec3)}}))
        return 1;
    printf("vl_demo: %d StepPhysics calls on %zu vertices, %zu springs\n", a.steps, X.size(), springs.size());
    return 0;
}
