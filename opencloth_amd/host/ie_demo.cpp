// The implicit-Euler demo of OpenCloth (IE) as a C++ program without the
// window: InitGL's scene, then StepPhysics in a loop, through the C++ host
// layer (opencloth.hpp) over the C ABI.  This is the drop-in the reference's
// own main.cpp would make; the scene set-up below follows InitGL, the step
// loop follows OnIdle's fixed time stepping.
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
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "opencloth.hpp"

namespace {

struct vec3 {  // the layout of glm::vec3
    float x, y, z;
};

// Grid and tunables (IE:59-62, IE:75-77, IE:310-316).
const int numX = 20, numY = 20;
const float fullsize = 4.0f, halfsize = fullsize / 2.0f;
const float KsStruct = 0.75f, KdStruct = -0.25f;
const float KsShear = 0.75f, KdShear = -0.25f;
const float KsBend = 0.95f, KdBend = -0.25f;
const vec3 gravity = {0.0f, -0.00981f, 0.0f};  // a force per vertex (IE:653)
const float mass = 0.5f;
const float timeStep = 1 / 60.0f;
const float EPS = 0.001f;
const int i_max = 10;
enum { STRUCTURAL_SPRING = 0, SHEAR_SPRING = 1, BEND_SPRING = 2 };

std::vector<vec3> X, V;
std::vector<oc_spring> springs;

// AddSpring (IE:336-346): rest length = |X[a] - X[b]| in float.
void AddSpring(int a, int b, float ks, float kd, int type) {
    const vec3 d = {X[a].x - X[b].x, X[a].y - X[b].y, X[a].z - X[b].z};
    oc_spring s;
    s.p1 = a;
    s.p2 = b;
    s.rest_length = std::sqrt(d.x * d.x + d.y * d.y + d.z * d.z);
    s.Ks = ks;
    s.Kd = kd;
    s.type = type;
    springs.push_back(s);
}

// The state and spring parts of InitGL (IE:454-530).
void InitScene() {
    const int u = numX + 1, v = numY + 1;
    X.resize((size_t)u * v);
    V.assign((size_t)u * v, vec3{0.0f, 0.0f, 0.0f});
    size_t k = 0;
    for (int j = 0; j < v; j++)
        for (int i = 0; i < u; i++)
            X[k++] = vec3{((float(i) / (u - 1)) * 2 - 1) * halfsize, fullsize + 1,
                          (float(j) / (v - 1)) * fullsize};
    for (int l1 = 0; l1 < v; l1++)  // horizontal structural
        for (int l2 = 0; l2 < u - 1; l2++)
            AddSpring(l1 * u + l2, l1 * u + l2 + 1, KsStruct, KdStruct, STRUCTURAL_SPRING);
    for (int l1 = 0; l1 < u; l1++)  // vertical structural
        for (int l2 = 0; l2 < v - 1; l2++)
            AddSpring(l2 * u + l1, (l2 + 1) * u + l1, KsStruct, KdStruct, STRUCTURAL_SPRING);
    for (int l1 = 0; l1 < v - 1; l1++)  // two shears per quad
        for (int l2 = 0; l2 < u - 1; l2++) {
            AddSpring(l1 * u + l2, (l1 + 1) * u + l2 + 1, KsShear, KdShear, SHEAR_SPRING);
            AddSpring((l1 + 1) * u + l2, l1 * u + l2 + 1, KsShear, KdShear, SHEAR_SPRING);
        }
    for (int l1 = 0; l1 < v; l1++) {  // bend rows, the last one twice
        for (int l2 = 0; l2 < u - 2; l2++)
            AddSpring(l1 * u + l2, l1 * u + l2 + 2, KsBend, KdBend, BEND_SPRING);
        AddSpring(l1 * u + (u - 3), l1 * u + (u - 1), KsBend, KdBend, BEND_SPRING);
    }
    for (int l1 = 0; l1 < u; l1++) {  // bend columns, the last one twice
        for (int l2 = 0; l2 < v - 2; l2++)
            AddSpring(l2 * u + l1, (l2 + 2) * u + l1, KsBend, KdBend, BEND_SPRING);
        AddSpring((v - 3) * u + l1, (v - 1) * u + l1, KsBend, KdBend, BEND_SPRING);
    }
}

// The collider of InitGL (IE:541-544): translate(0,2,0) * rotate(45 degrees
// about x) * scale(1, 1, 1/2), and its inverse scale^-1 * rotate^T *
// translate^-1, both evaluated in double and rounded, column-major.
oc_ellipsoid MakeEllipsoid() {
    const double a = 45.0 * (M_PI / 180.0), c = std::cos(a), s = std::sin(a);
    const double r = 1.0, t[3] = {0.0, 2.0, 0.0};
    const double sc[3] = {r, r, r / 2};
    // rows of the rotation about x
    const double R[3][3] = {{1, 0, 0}, {0, c, -s}, {0, s, c}};
    double M[4][4] = {}, Mi[4][4] = {};
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++) M[i][j] = R[i][j] * sc[j];
        M[i][3] = t[i];
    }
    M[3][3] = 1;
    for (int i = 0; i < 3; i++) {
        double rt = 0;  // (R^T t)_i
        for (int j = 0; j < 3; j++) {
            Mi[i][j] = R[j][i] / sc[i];
            rt += R[j][i] * t[j];
        }
        Mi[i][3] = 0.0 - rt / sc[i];
    }
    Mi[3][3] = 1;
    oc_ellipsoid e;
    for (int col = 0; col < 4; col++)
        for (int row = 0; row < 4; row++) {
            e.transform[col * 4 + row] = (float)M[row][col];
            e.inverse[col * 4 + row] = (float)Mi[row][col];
        }
    e.center[0] = e.center[1] = e.center[2] = 0.0f;  // IE:85
    e.radius = 1.0f;                                  // IE:86
    return e;
}

bool write_all(FILE* f, const void* p, size_t n) { return n == 0 || fwrite(p, 1, n, f) == n; }

}  // namespace

int main(int argc, char** argv) {
    int steps = 0;
    bool scene_only = false;
    const char* out = nullptr;
    for (int k = 1; k < argc; k++) {
        if (!strcmp(argv[k], "--steps") && k + 1 < argc)
            steps = atoi(argv[++k]);
        else if (!strcmp(argv[k], "--out") && k + 1 < argc)
            out = argv[++k];
        else if (!strcmp(argv[k], "--scene-only"))
            scene_only = true;
        else {
            fprintf(stderr, "usage: %s [--steps N | --scene-only] --out FILE\n", argv[0]);
            return 2;
        }
    }
    if (!out || steps < 0) {
        fprintf(stderr, "usage: %s [--steps N | --scene-only] --out FILE\n", argv[0]);
        return 2;
    }
    InitScene();
    std::vector<uint8_t> pinned(X.size(), 0);
    pinned[0] = pinned[numX] = 1;  // IE:652
    const oc_ellipsoid ellipsoid = MakeEllipsoid();

    int iters = 0, status = 0;
    if (!scene_only) {
        oc_params p = {};
        p.mass = mass;
        p.gravity[0] = gravity.x / mass;
        p.gravity[1] = gravity.y / mass;
        p.gravity[2] = gravity.z / mass;
        p.dt = timeStep;
        p.cg_max_iters = i_max;
        p.cg_tol = EPS;
        p.cg_tol_mode = OC_TOL_ABSOLUTE;       // IE:238
        p.warm_start = 1;                      // IE:691
        p.flags = OC_FLAG_GROUND_CLAMP;        // IE:695-697
        try {
            opencloth::ImplicitEuler<vec3> cloth(numX, numY, X, V, springs, pinned, p);
            for (int s = 0; s < steps; s++) cloth.StepPhysics(timeStep, &ellipsoid, 1.0f);
            cloth.LastSolve(&iters, &status);
        } catch (const opencloth::Error& e) {
            fprintf(stderr, "ie_demo: %s (status %d)\n", e.what(), e.code());
            return 1;
        }
    }

    FILE* f = fopen(out, "wb");
    if (!f) {
        perror(out);
        return 1;
    }
    const int32_t head[2] = {(int32_t)X.size(), (int32_t)springs.size()};
    bool ok = write_all(f, head, sizeof head) && write_all(f, X.data(), X.size() * sizeof(vec3));
    if (scene_only)
        ok = ok && write_all(f, springs.data(), springs.size() * sizeof(oc_spring)) &&
             write_all(f, pinned.data(), pinned.size()) && write_all(f, &ellipsoid, sizeof ellipsoid);
    else
        ok = ok && write_all(f, V.data(), V.size() * sizeof(vec3));
    ok = (fclose(f) == 0) && ok;
    if (!ok) {
        fprintf(stderr, "ie_demo: writing %s failed\n", out);
        return 1;
    }
    if (!scene_only)
        printf("ie_demo: %d StepPhysics calls on %zu vertices, %zu springs; last CG %d iterations, status %d\n",
               steps, X.size(), springs.size(), iters, status);
    return 0;
}
