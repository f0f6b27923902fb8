// Pieces shared by the demo programs: the grid of InitGL, the collider and
// the output file.  Each demo follows its own reference main.cpp for the
// rest (paths relative to the reference repository):
//   IE  = OpenCloth_ImplicitEuler/OpenCloth_ImplicitEuler/main.cpp
//   VL  = OpenCloth_Verlet/OpenCloth_Verlet/main.cpp
//   PBD = OpenCloth_PositionBasedDynamics/OpenCloth_PositionBasedDynamics/main.cpp
#pragma once

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "opencloth.hpp"

namespace demo {

struct vec3 {  // the layout of glm::vec3
    float x, y, z;
};

// Vertex k = j*(numX+1) + i at (((i/(u-1))*2-1)*halfsize, fullsize+1,
// (j/(v-1))*fullsize) (IE:462-466, VL and PBD alike).
inline std::vector<vec3> grid(int numX, int numY, float fullsize) {
    const int u = numX + 1, v = numY + 1;
    const float halfsize = fullsize / 2.0f;
    std::vector<vec3> X((size_t)u * v);
    size_t k = 0;
    for (int j = 0; j < v; j++)
        for (int i = 0; i < u; i++)
            X[k++] = vec3{((float(i) / (u - 1)) * 2 - 1) * halfsize, fullsize + 1,
                          (float(j) / (v - 1)) * fullsize};
    return X;
}

// The reference's spring list (IE:496-530, VL alike): horizontal, vertical,
// two shears per quad, bend rows and bend columns, each ending with a
// duplicate of its last bend; rest length |X[a] - X[b]| (AddSpring,
// IE:336-346).
inline std::vector<oc_spring> springs(int numX, int numY, const std::vector<vec3>& X, const float ks[3],
                                      const float kd[3]) {
    const int u = numX + 1, v = numY + 1;
    std::vector<oc_spring> out;
    auto add = [&](int a, int b, int type) {
        const vec3 d = {X[a].x - X[b].x, X[a].y - X[b].y, X[a].z - X[b].z};
        oc_spring s;
        s.p1 = a;
        s.p2 = b;
        s.rest_length = std::sqrt(d.x * d.x + d.y * d.y + d.z * d.z);
        s.Ks = ks[type];
        s.Kd = kd[type];
        s.type = type;
        out.push_back(s);
    };
    for (int l1 = 0; l1 < v; l1++)
        for (int l2 = 0; l2 < u - 1; l2++) add(l1 * u + l2, l1 * u + l2 + 1, 0);
    for (int l1 = 0; l1 < u; l1++)
        for (int l2 = 0; l2 < v - 1; l2++) add(l2 * u + l1, (l2 + 1) * u + l1, 0);
    for (int l1 = 0; l1 < v - 1; l1++)
        for (int l2 = 0; l2 < u - 1; l2++) {
            add(l1 * u + l2, (l1 + 1) * u + l2 + 1, 1);
            add((l1 + 1) * u + l2, l1 * u + l2 + 1, 1);
        }
    for (int l1 = 0; l1 < v; l1++) {
        for (int l2 = 0; l2 < u - 2; l2++) add(l1 * u + l2, l1 * u + l2 + 2, 2);
        add(l1 * u + (u - 3), l1 * u + (u - 1), 2);
    }
    for (int l1 = 0; l1 < u; l1++) {
        for (int l2 = 0; l2 < v - 2; l2++) add(l2 * u + l1, (l2 + 2) * u + l1, 2);
        add((v - 3) * u + l1, (v - 1) * u + l1, 2);
    }
    return out;
}

// The collider of InitGL (IE:541-544, VL and PBD alike): translate(0,2,0) *
// rotate(45 degrees about x) * scale(1, 1, 1/2), and its inverse scale^-1 *
// rotate^T * translate^-1, both evaluated in double and rounded,
// column-major; object-space centre 0 and radius 1 (IE:85-86).
inline oc_ellipsoid ellipsoid() {
    const double a = 45.0 * (M_PI / 180.0), c = std::cos(a), s = std::sin(a);
    const double r = 1.0, t[3] = {0.0, 2.0, 0.0};
    const double sc[3] = {r, r, r / 2};
    const double R[3][3] = {{1, 0, 0}, {0, c, -s}, {0, s, c}};  // rows
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
    e.center[0] = e.center[1] = e.center[2] = 0.0f;
    e.radius = 1.0f;
    return e;
}

struct Args {
    int steps = 0;
    bool scene_only = false;
    const char* out = nullptr;
};

inline bool parse(int argc, char** argv, Args* a) {
    for (int k = 1; k < argc; k++) {
        if (!strcmp(argv[k], "--steps") && k + 1 < argc)
            a->steps = atoi(argv[++k]);
        else if (!strcmp(argv[k], "--out") && k + 1 < argc)
            a->out = argv[++k];
        else if (!strcmp(argv[k], "--scene-only"))
            a->scene_only = true;
        else
            return false;
    }
    return a->out != nullptr && a->steps >= 0;
}

inline int usage(const char* prog) {
    fprintf(stderr, "usage: %s [--steps N | --scene-only] --out FILE\n", prog);
    return 2;
}

// Writes each (pointer, bytes) piece in order; returns false on failure.
struct Piece {
    const void* p;
    size_t n;
};
inline bool write_file(const char* path, std::initializer_list<Piece> pieces) {
    FILE* f = fopen(path, "wb");
    if (!f) {
        perror(path);
        return false;
    }
    bool ok = true;
    for (const Piece& q : pieces) ok = ok && (q.n == 0 || fwrite(q.p, 1, q.n, f) == q.n);
    ok = (fclose(f) == 0) && ok;
    if (!ok) fprintf(stderr, "writing %s failed\n", path);
    return ok;
}

}  // namespace demo
