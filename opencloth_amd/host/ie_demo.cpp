// The implicit-Euler demo of OpenCloth (IE) as a C++ program without the
// window: InitGL's scene, then StepPhysics in a loop, through the C++ host
// layer (opencloth.hpp) over the C ABI.  This is the drop-in the reference's
// own main.cpp would make; the scene set-up follows InitGL, the step loop
// follows OnIdle's fixed time stepping.
//
//   ie_demo --steps N --out FILE     N StepPhysics calls; FILE gets X and V
//   ie_demo --scene-only --out FILE  the scene only (no GPU call): X,
//                                    springs, pins and the collider
//
// FILE layout (little-endian): int32 n_vertices, int32 n_springs, then
// X (3 floats per vertex), and either V (3 floats per vertex) or, with
// --scene-only, the springs (24-byte records), the pins (1 byte per vertex)
// and the collider (oc_ellipsoid, 36 floats).
//
// IE = OpenCloth_ImplicitEuler/OpenCloth_ImplicitEuler/main.cpp (paths
// relative to the reference repository).
#include "demo_common.hpp"

using demo::vec3;

namespace {

// Grid and tunables (IE:59-62, IE:75-77, IE:310-316).
const int numX = 20, numY = 20;
const float fullsize = 4.0f;
const float Ks[3] = {0.75f, 0.75f, 0.95f};      // KsStruct, KsShear, KsBend
const float Kd[3] = {-0.25f, -0.25f, -0.25f};   // KdStruct, KdShear, KdBend
const vec3 gravity = {0.0f, -0.00981f, 0.0f};   // a force per vertex (IE:653)
const float mass = 0.5f;
const float timeStep = 1 / 60.0f;
const float EPS = 0.001f;
const int i_max = 10;

}  // namespace

int main(int argc, char** argv) {
    demo::Args a;
    if (!demo::parse(argc, argv, &a)) return demo::usage(argv[0]);
    std::vector<vec3> X = demo::grid(numX, numY, fullsize);  // IE:454-466
    std::vector<vec3> V(X.size(), vec3{0.0f, 0.0f, 0.0f});
    const std::vector<oc_spring> springs = demo::springs(numX, numY, X, Ks, Kd);  // IE:496-530
    std::vector<uint8_t> pinned(X.size(), 0);
    pinned[0] = pinned[numX] = 1;  // IE:652
    const oc_ellipsoid ellipsoid = demo::ellipsoid();
    const int32_t head[2] = {(int32_t)X.size(), (int32_t)springs.size()};
    if (a.scene_only)
        return demo::write_file(a.out, {{head, sizeof head}, {X.data(), X.size() * sizeof(vec3)},
                                        {springs.data(), springs.size() * sizeof(oc_spring)},
                                        {pinned.data(), pinned.size()}, {&ellipsoid, sizeof ellipsoid}})
                   ? 0 : 1;

    oc_params p = {};
    p.mass = mass;
    p.gravity[0] = gravity.x / mass;
    p.gravity[1] = gravity.y / mass;
    p.gravity[2] = gravity.z / mass;
    p.dt = timeStep;
    p.cg_max_iters = i_max;
    p.cg_tol = EPS;
    p.cg_tol_mode = OC_TOL_ABSOLUTE;  // IE:238
    p.warm_start = 1;                 // IE:691
    // IE:695-697; the CG in one launch that stops at convergence: at the
    // demo's size, where the solve needs 1-2 of its 10 iterations, that is
    // faster than one launch pair per iteration (same results)
    p.flags = OC_FLAG_GROUND_CLAMP | OC_FLAG_CG_PERSISTENT;
    int iters = 0, status = 0;
    try {
        opencloth::ImplicitEuler<vec3> cloth(numX, numY, X, V, springs, pinned, p);
        for (int s = 0; s < a.steps; s++) cloth.StepPhysics(timeStep, &ellipsoid, 1.0f);
        cloth.LastSolve(&iters, &status);
    } catch (const opencloth::Error& e) {
        fprintf(stderr, "ie_demo: %s (status %d)\n", e.what(), e.code());
        return 1;
    }
    if (!demo::write_file(a.out, {{head, sizeof head}, {X.data(), X.size() * sizeof(vec3)},
                                  {V.data(), V.size() * sizeof(vec3)}}))
        return 1;
    printf("ie_demo: %d StepPhysics calls on %zu vertices, %zu springs; last CG %d iterations, status %d\n",
           a.steps, X.size(), springs.size(), iters, status);
    return 0;
}
