// The Baraff-Witkin implicit demo of OpenCloth (BW) as a C++ program without
// the window: InitGL's scene, then StepPhysics in a loop, through the C++
// host layer (opencloth.hpp) over the C ABI.  The step is BW's: the global
// velocity damping DEFAULT_DAMPING on every vertex (BW:706-708), the
// Jacobi-preconditioned CG of SolveConjugateGradientPreconditioned
// (BW:308-331) with its relative stopping test (BW:319), the ground clamp
// (BW:794-796), Provot with the 1% slack and the collider (BW:800-848).
// BW's own dt applied twice to dV (BW:791) and the damping left out of its
// Jacobian are not copied: the damping's Jacobian c*I is in A.
//
//   bw_demo --steps N --out FILE     N StepPhysics calls; FILE gets X and V
//   bw_demo --scene-only --out FILE  the scene only (no GPU call): X,
//                                    springs, pins and the collider
//
// FILE layout as ie_demo's.  BW = OpenCloth_Implicit/OpenCloth_Implicit/main.cpp
// (paths relative to the reference repository).
#include "demo_common.hpp"

using demo::vec3;

namespace {

// Grid and tunables (BW:60-65, BW:104-109, BW:126-128).
const int numX = 20, numY = 20;
const float fullsize = 4.0f;
const float Ks[3] = {0.75f, 0.75f, 0.95f};     // KsStruct, KsShear, KsBend
const float Kd[3] = {-0.25f, -0.25f, -0.25f};  // KdStruct, KdShear, KdBend
const vec3 gravity = {0.0f, -0.00981f, 0.0f};  // a force on each free vertex (BW:704-706)
const float mass = 1.0f;
const float timeStep = 5 / 60.0f;
const float DEFAULT_DAMPING = -0.125f;
const float EPS = 0.001f;
const int i_max = 10;

}  // namespace

int main(int argc, char** argv) {
    demo::Args a;
    if (!demo::parse(argc, argv, &a)) return demo::usage(argv[0]);
    std::vector<vec3> X = demo::grid(numX, numY, fullsize);  // BW:497-502
    std::vector<vec3> V(X.size(), vec3{0.0f, 0.0f, 0.0f});
    const std::vector<oc_spring> springs = demo::springs(numX, numY, X, Ks, Kd);  // BW:531-565
    std::vector<uint8_t> pinned(X.size(), 0);
    pinned[0] = pinned[numX] = 1;  // no gravity there (BW:705), Provot's pins (BW:812-817)
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
    p.cg_tol_mode = OC_TOL_RELATIVE;                  // BW:319
    p.warm_start = 1;
    p.damping = DEFAULT_DAMPING;                      // BW:707-708
    p.flags = OC_FLAG_GROUND_CLAMP | OC_FLAG_JACOBI;  // BW:794-796, BW:784-788
    int iters = 0, status = 0;
    try {
        opencloth::ImplicitEuler<vec3> cloth(numX, numY, X, V, springs, pinned, p);
        for (int s = 0; s < a.steps; s++) cloth.StepPhysics(timeStep, &ellipsoid, 1.01f);  // BW:808
        cloth.LastSolve(&iters, &status);
    } catch (const opencloth::Error& e) {
        fprintf(stderr, "bw_demo: %s (status %d)\n", e.what(), e.code());
        return 1;
    }
    if (!demo::write_file(a.out, {{head, sizeof head}, {X.data(), X.size() * sizeof(vec3)},
                                  {V.data(), V.size() * sizeof(vec3)}}))
        return 1;
    printf("bw_demo: %d StepPhysics calls on %zu vertices, %zu springs; last CG %d iterations, status %d\n",
           a.steps, X.size(), springs.size(), iters, status);
    return 0;
}
