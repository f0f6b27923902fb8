// glm_ops.cpp -- TEST INFRASTRUCTURE ONLY (the oracle's own checker).
//
// Evaluates the vec3/mat3 operations of the implicit step with the
// reference's vendored GLM 0.9.0.0 headers (dep/glm/glm/core), compiled where
// they lie in the reference tree by oracle/Makefile's `ref` target into
// oracle/_ref/libglmref.so.  tests/test_oracle_glm.py compares oracle.c's
// restated arithmetic with these results bit for bit.
//
// Only GLM's core headers are included (glm.hpp also pulls in
// func_integer.inl, which current compilers reject); <cmath> and <limits> are
// included first because GLM 0.9.0 uses them without including them.
#include <cmath>
#include <limits>

#include <glm/core/_detail.hpp>
#include <glm/core/type.hpp>
#include <glm/core/func_exponential.hpp>
#include <glm/core/func_common.hpp>
#include <glm/core/func_geometric.hpp>
#include <glm/core/func_matrix.hpp>

extern "C" {

// Per row i of n: dot(a,b), length(a), normalize(a), outerProduct(a,b)
// (column-major 9), m*b with m = the 9 floats of M[i] (column-major).
__attribute__((visibility("default"))) void glmref_ops(int n, const float* a, const float* b,
                                                       const float* M, float* out_dot,
                                                       float* out_len, float* out_norm,
                                                       float* out_outer, float* out_mv) {
    for (int i = 0; i < n; i++) {
        glm::vec3 va(a[3 * i], a[3 * i + 1], a[3 * i + 2]);
        glm::vec3 vb(b[3 * i], b[3 * i + 1], b[3 * i + 2]);
        out_dot[i] = glm::dot(va, vb);
        out_len[i] = glm::length(va);
        glm::vec3 nn = glm::normalize(va);
        out_norm[3 * i] = nn.x;
        out_norm[3 * i + 1] = nn.y;
        out_norm[3 * i + 2] = nn.z;
        glm::mat3 o = glm::outerProduct(va, vb);
        glm::mat3 m(M[9 * i + 0], M[9 * i + 1], M[9 * i + 2], M[9 * i + 3], M[9 * i + 4],
                    M[9 * i + 5], M[9 * i + 6], M[9 * i + 7], M[9 * i + 8]);
        glm::vec3 mv = m * vb;
        for (int c = 0; c < 3; c++)
            for (int r = 0; r < 3; r++) out_outer[9 * i + 3 * c + r] = o[c][r];
        out_mv[3 * i] = mv.x;
        out_mv[3 * i + 1] = mv.y;
        out_mv[3 * i + 2] = mv.z;
    }
}

// The per-spring force and stiffness Jacobian of ComputeForces (IE:659-675),
// written with GLM types in the reference's expression order, for n springs
// given as endpoint positions/velocities.  K is column-major 9.
__attribute__((visibility("default"))) void glmref_spring(int n, const float* x1, const float* x2,
                                                          const float* v1, const float* v2,
                                                          const float* L0, const float* Ks,
                                                          const float* Kd, float* f, float* K) {
    const glm::mat3 I = glm::mat3(1);
    for (int i = 0; i < n; i++) {
        glm::vec3 p1(x1[3 * i], x1[3 * i + 1], x1[3 * i + 2]);
        glm::vec3 p2(x2[3 * i], x2[3 * i + 1], x2[3 * i + 2]);
        glm::vec3 w1(v1[3 * i], v1[3 * i + 1], v1[3 * i + 2]);
        glm::vec3 w2(v2[3 * i], v2[3 * i + 1], v2[3 * i + 2]);
        glm::vec3 deltaP = p1 - p2;
        glm::vec3 deltaV = w1 - w2;
        float dist = glm::length(deltaP);
        float dist2 = dist * dist;
        float lo_l = L0[i] / dist;
        glm::mat3 Ki = Ks[i] * (-I + (lo_l * (I - glm::outerProduct(deltaP, deltaP) / dist2)));
        float leftTerm = -Ks[i] * (dist - L0[i]);
        float rightTerm = Kd[i] * (glm::dot(deltaV, deltaP) / dist);
        glm::vec3 springForce = (leftTerm + rightTerm) * glm::normalize(deltaP);
        f[3 * i] = springForce.x;
        f[3 * i + 1] = springForce.y;
        f[3 * i + 2] = springForce.z;
        for (int c = 0; c < 3; c++)
            for (int r = 0; r < 3; r++) K[9 * i + 3 * c + r] = Ki[c][r];
    }
}
}
