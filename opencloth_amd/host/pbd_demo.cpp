// The position-based-dynamics demo of OpenCloth (PBD) as a C++ program
// without the window: InitGL's particles and constraints, then StepPhysics
// in a loop, through the C++ host layer (opencloth.hpp) over the C ABI.
//
//   pbd_demo --steps N --out FILE     N StepPhysics calls; FILE gets X and V
//   pbd_demo --scene-only --out FILE  the scene only (no GPU call): X, the
//                                     constraint lists, pins and the collider
//
// FILE layout (little-endian): int32 n_vertices, int32 n_distance,
// int32 n_bending, then X (3 floats per vertex), and either V or, with
// --scene-only, the distance constraints (p1, p2, k), the bending
// constraints (p1, p2, p3, k), the pins (1 byte per vertex) and the
// collider (oc_ellipsoid, 36 floats).
//
// PBD = OpenCloth_PositionBasedDynamics/OpenCloth_PositionBasedDynamics/main.cpp
// (paths relative to the reference repository), built with
// USE_TRIANGLE_BENDING_CONSTRAINT (PBD:57).
#include "demo_common.hpp"

using demo::vec3;

namespace {

// Grid and tunables (PBD:66-134).
const int numX = 20, numY = 20;
const float fullsize = 4.0f;
const float kStretch = 0.25f, kBend = 0.5f, kDamp = 0.00125f;
const float global_dampening = 0.98f;
const int solver_iterations = 2;
const vec3 gravity = {0.0f, -0.00981f, 0.0f};
const float timeStep = 1.0f / 60.0f;

int getIndex(int i, int j) { return j * (numX + 1) + i; }  // PBD:306-308

}  // namespace

int main(int argc, char** argv) {
    demo::Args a;
    if (!demo::parse(argc, argv, &a)) return demo::usage(argv[0]);
    const int u = numX + 1, v = numY + 1;
    std::vector<vec3> X = demo::grid(numX, numY, fullsize);  // PBD:330-343
    std::vector<vec3> V(X.size(), vec3{0.0f, 0.0f, 0.0f});
    std::vector<uint8_t> pinned(X.size(), 0);
    pinned[0] = pinned[numX] = 1;  // W = 0 (PBD:350-351)
    // AddDistanceConstraint calls (PBD:399-416): horizontal, vertical, shear
    std::vector<oc_pbd_distance> d;
    for (int l1 = 0; l1 < v; l1++)
        for (int l2 = 0; l2 < u - 1; l2++) d.push_back({getIndex(l2, l1), getIndex(l2 + 1, l1), kStretch});
    for (int l1 = 0; l1 < u; l1++)
        for (int l2 = 0; l2 < v - 1; l2++) d.push_back({getIndex(l1, l2), getIndex(l1, l2 + 1), kStretch});
    for (int l1 = 0; l1 < v - 1; l1++)
        for (int l2 = 0; l2 < u - 1; l2++) {
            d.push_back({getIndex(l2, l1), getIndex(l2 + 1, l1 + 1), kStretch});
            d.push_back({getIndex(l2, l1 + 1), getIndex(l2 + 1, l1), kStretch});
        }
    // AddBendingConstraint calls, triangle form (PBD:418-433): along columns, then rows
    std::vector<oc_pbd_bending> b;
    for (int i = 0; i < u; i++)
        for (int j = 0; j < v - 2; j++) b.push_back({getIndex(i, j), getIndex(i, j + 1), getIndex(i, j + 2), kBend});
    for (int i = 0; i < u - 2; i++)
        for (int j = 0; j < v; j++) b.push_back({getIndex(i, j), getIndex(i + 1, j), getIndex(i + 2, j), kBend});
    const oc_ellipsoid ellipsoid = demo::ellipsoid();
    const int32_t head[3] = {(int32_t)X.size(), (int32_t)d.size(), (int32_t)b.size()};
    if (a.scene_only)
        return demo::write_file(a.out, {{head, sizeof head}, {X.data(), X.size() * sizeof(vec3)},
                                        {d.data(), d.size() * sizeof(oc_pbd_distance)},
                                        {b.data(), b.size() * sizeof(oc_pbd_bending)},
                                        {pinned.data(), pinned.size()}, {&ellipsoid, sizeof ellipsoid}})
                   ? 0 : 1;

    oc_pbd_params p = {};
    p.mass = 1.0f / (float)X.size();  // PBD:114
    p.gravity[0] = gravity.x;
    p.gravity[1] = gravity.y;
    p.gravity[2] = gravity.z;
    p.dt = timeStep;
    p.global_damping = global_dampening;
    p.k_damp = kDamp;
    p.solver_iterations = solver_iterations;
    p.flags = OC_FLAG_GROUND_CLAMP;  // GroundCollision (PBD:887)
    try {
        opencloth::PositionBasedDynamics<vec3> cloth(X, V, pinned, d, b, p);
        for (int s = 0; s < a.steps; s++) cloth.StepPhysics(timeStep, &ellipsoid);  // PBD:915-925
    } catch (const opencloth::Error& e) {
        fprintf(stderr, "pbd_demo: %s (status %d)\n", e.what(), e.code());
        return 1;
    }
    if (!demo::write_file(a.out, {{head, sizeof head}, {X.data(), X.size() * sizeof(vec3)},
                                  {V.data(), V.size() * sizeof(vec3)}}))
        return 1;
    printf("pbd_demo: %d StepPhysics calls on %zu particles, %zu distance and %zu bending constraints\n",
           a.steps, X.size(), d.size(), b.size());
    return 0;
}
