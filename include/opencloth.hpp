/*
 * opencloth.hpp -- C++ host layer over the C ABI (opencloth.h) with the
 * reference's entry-point names, for C++ programs built like the
 * reference's demos.
 *
 * The reference keeps each demo's cloth in file-scope globals
 * (std::vector<glm::vec3> X, V, vector<Spring> springs) and calls free
 * `void` functions on them from the GLUT idle loop.  The classes below keep
 * references to the caller's X and V instead of globals.  Each member named
 * after a reference function does that function's work on the GPU and
 * leaves X and V updated in host memory on return, as the reference's
 * globals are.  Step() runs several steps without that copy-back.
 *
 * Errors: the reference's functions cannot fail (they return void and let
 * NaNs propagate).  Here a failing call throws opencloth::Error, which
 * carries the oc_status code and the library's message.  The solver's own
 * outcome (iterations, OC_SOLVE_* bits) is a result, not an error; it is
 * available from LastSolve().
 *
 * Vec3 is the caller's 3-float vector type: glm::vec3 in the reference, or
 * any 12-byte struct of three floats (checked at compile time).
 *
 * Reference files (paths relative to the reference repository):
 *   IE  = OpenCloth_ImplicitEuler/OpenCloth_ImplicitEuler/main.cpp
 *   VL  = OpenCloth_Verlet/OpenCloth_Verlet/main.cpp
 *   PBD = OpenCloth_PositionBasedDynamics/OpenCloth_PositionBasedDynamics/main.cpp
 *   CUH = OpenCloth_Verlet_CUDA_GLSL_OPENCL_CPU/OpenCloth_Verlet_CUDA/verlet.cu
 */
#ifndef OPENCLOTH_HPP
#define OPENCLOTH_HPP

#include <cstdint>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "opencloth.h"

namespace opencloth {

class Error : public std::runtime_error {
   public:
    Error(int code, const std::string& what) : std::runtime_error(what), code_(code) {}
    int code() const { return code_; }

   private:
    int code_;
};

inline void check(int rc, const char* call) {
    if (rc != OC_OK) throw Error(rc, std::string(call) + ": " + oc_last_error());
}

template <class Vec4>
inline float* floats4(std::vector<Vec4>& v) {
    static_assert(sizeof(Vec4) == 4 * sizeof(float) && std::is_standard_layout<Vec4>::value,
                  "Vec4 must be four packed floats, like glm::vec4");
    return reinterpret_cast<float*>(v.data());
}
template <class Vec3>
inline float* floats(std::vector<Vec3>& v) {
    static_assert(sizeof(Vec3) == 3 * sizeof(float) && std::is_standard_layout<Vec3>::value,
                  "Vec3 must be three packed floats, like glm::vec3");
    return reinterpret_cast<float*>(v.data());
}
template <class Vec3>
inline const float* floats(const std::vector<Vec3>& v) {
    return floats(const_cast<std::vector<Vec3>&>(v));
}

inline void need(bool ok, const char* what) {
    if (!ok) throw Error(OC_ERR_INVALID_ARGUMENT, what);
}

/* The implicit-Euler demo's cloth (IE).  The scene comes from the caller as
 * InitGL builds it (IE:434-546): X, V on the (numX+1) x (numY+1) grid, the
 * spring list, the pinned vertices (IE:652 pins {0, numX}) and the tunables
 * in oc_params.  The reference adds `gravity` as a force (IE:653); oc_params
 * holds an acceleration, so a caller passes gravity / mass (exact for the
 * reference's mass 0.5). */
template <class Vec3>
class ImplicitEuler {
   public:
    ImplicitEuler(int numX, int numY, std::vector<Vec3>& X, std::vector<Vec3>& V,
                  const std::vector<oc_spring>& springs, const std::vector<uint8_t>& pinned,
                  const oc_params& params)
        : X_(X), V_(V), dt_(params.dt) {
        const size_t n = (size_t)(numX + 1) * (size_t)(numY + 1);
        need(X.size() == n && V.size() == n && pinned.size() == n,
             "ImplicitEuler: X, V and pinned need (numX+1)*(numY+1) entries");
        check(oc_cloth_create(numX + 1, numY + 1, floats(X), floats(V), springs.data(),
                              (int32_t)springs.size(), pinned.data(), &params, &h_),
              "oc_cloth_create");
    }
    ~ImplicitEuler() { oc_cloth_destroy(h_); }
    ImplicitEuler(const ImplicitEuler&) = delete;
    ImplicitEuler& operator=(const ImplicitEuler&) = delete;

    /* ComputeForces (IE:645-682): gravity on free vertices plus the spring
     * forces, into F (resized to the vertex count). */
    void ComputeForces(std::vector<Vec3>& F) {
        F.resize(X_.size());
        check(oc_cloth_compute_forces(h_, floats(F)), "oc_cloth_compute_forces");
    }

    /* IntegrateImplicit (IE:684-700): assemble the dv system, solve it by
     * filtered CG, v += dv, x += dt v (ground clamp when the flag is set). */
    void IntegrateImplicit(float deltaTime) {
        check(oc_cloth_integrate_implicit(h_, deltaTime), "oc_cloth_integrate_implicit");
        download();
    }

    /* IntegrateImplicit's parts, for callers that keep the reference's split:
     * the system at the current state, SolveConjugateGradient (IE:227-249) on
     * it, and the update. */
    void AssembleSystem(float deltaTime) { check(oc_cloth_assemble(h_, deltaTime), "oc_cloth_assemble"); }
    int SolveConjugateGradient() {
        int32_t it = 0, st = 0;
        check(oc_cloth_solve(h_, &it, &st), "oc_cloth_solve");
        return it;
    }
    void UpdateState(float deltaTime) {
        check(oc_cloth_update(h_, deltaTime), "oc_cloth_update");
        download();
    }

    /* EllipsoidCollision (IE:727-751). */
    void EllipsoidCollision(const oc_ellipsoid& e) {
        check(oc_cloth_ellipsoid_collision(h_, &e), "oc_cloth_ellipsoid_collision");
        download();
    }

    /* ApplyProvotDynamicInverse (IE:702-726); slack 1 as in IE. */
    void ApplyProvotDynamicInverse(float slack = 1.0f) {
        check(oc_cloth_provot(h_, slack), "oc_cloth_provot");
        download();
    }

    /* StepPhysics (IE:774-779): ComputeForces, IntegrateImplicit(timeStep),
     * EllipsoidCollision (when e is given), ApplyProvotDynamicInverse (when
     * slack > 0).  Like the reference, the step integrates with the
     * configured time step; dt is accepted for the same signature. */
    void StepPhysics(float dt, const oc_ellipsoid* e = nullptr, float slack = 1.0f) {
        (void)dt;
        check(oc_cloth_step_physics(h_, 1, e, slack), "oc_cloth_step_physics");
        download();
    }

    /* n StepPhysics calls without copying X and V back in between. */
    void Step(int n, const oc_ellipsoid* e = nullptr, float slack = 1.0f) {
        check(oc_cloth_step_physics(h_, n, e, slack), "oc_cloth_step_physics");
        download();
    }

    /* After the caller writes X or V directly, as the reference's mouse drag
     * does between steps (IE:405-410). */
    void Upload() { check(oc_cloth_set_state(h_, floats(X_), floats(V_)), "oc_cloth_set_state"); }

    /* Iterations and OC_SOLVE_* bits of the last CG solve. */
    void LastSolve(int* iters, int* status) {
        int32_t it = 0, st = 0;
        check(oc_cloth_last_status(h_, &it, &st), "oc_cloth_last_status");
        if (iters) *iters = it;
        if (status) *status = st;
    }

    oc_cloth handle() const { return h_; }
    float timeStep() const { return dt_; }

   private:
    void download() { check(oc_cloth_get_state(h_, floats(X_), floats(V_)), "oc_cloth_get_state"); }

    std::vector<Vec3>& X_;
    std::vector<Vec3>& V_;
    float dt_;
    oc_cloth h_ = nullptr;
};

/* The position-Verlet demo's cloth (VL): X and X_last (VL:240-328). */
template <class Vec3>
class Verlet {
   public:
    Verlet(int numX, int numY, std::vector<Vec3>& X, std::vector<Vec3>& X_last,
           const std::vector<oc_spring>& springs, const std::vector<uint8_t>& pinned,
           const oc_verlet_params& params)
        : X_(X), XL_(X_last) {
        const size_t n = (size_t)(numX + 1) * (size_t)(numY + 1);
        need(X.size() == n && X_last.size() == n && pinned.size() == n,
             "Verlet: X, X_last and pinned need (numX+1)*(numY+1) entries");
        check(oc_verlet_create(numX + 1, numY + 1, floats(X), floats(X_last), springs.data(),
                               (int32_t)springs.size(), pinned.data(), &params, &h_),
              "oc_verlet_create");
    }
    ~Verlet() { oc_verlet_destroy(h_); }
    Verlet(const Verlet&) = delete;
    Verlet& operator=(const Verlet&) = delete;

    /* ComputeForces(dt) (VL:448-484). */
    void ComputeForces(std::vector<Vec3>& F) {
        F.resize(X_.size());
        check(oc_verlet_compute_forces(h_, floats(F)), "oc_verlet_compute_forces");
    }
    /* StepPhysics (VL:557-562): ComputeForces, IntegrateVerlet, and the
     * ellipsoid pass when e is given. */
    void StepPhysics(float dt, const oc_ellipsoid* e = nullptr) {
        (void)dt;
        check(oc_verlet_step(h_, 1, e), "oc_verlet_step");
        download();
    }
    void Upload() { check(oc_verlet_set_state(h_, floats(X_), floats(XL_)), "oc_verlet_set_state"); }
    oc_verlet handle() const { return h_; }

   private:
    void download() { check(oc_verlet_get_state(h_, floats(X_), floats(XL_)), "oc_verlet_get_state"); }

    std::vector<Vec3>& X_;
    std::vector<Vec3>& XL_;
    oc_verlet h_ = nullptr;
};

/* The position-based-dynamics demo's cloth (PBD): particles X, V and the
 * constraint lists that AddDistanceConstraint / AddBendingConstraint build
 * (PBD:148-178). */
template <class Vec3>
class PositionBasedDynamics {
   public:
    PositionBasedDynamics(std::vector<Vec3>& X, std::vector<Vec3>& V, const std::vector<uint8_t>& pinned,
                          const std::vector<oc_pbd_distance>& d, const std::vector<oc_pbd_bending>& b,
                          const oc_pbd_params& params)
        : X_(X), V_(V) {
        need(V.size() == X.size() && pinned.size() == X.size(),
             "PositionBasedDynamics: X, V and pinned need the same length");
        check(oc_pbd_create((int32_t)X.size(), floats(X), pinned.data(), d.data(), (int32_t)d.size(),
                            b.data(), (int32_t)b.size(), &params, &h_),
              "oc_pbd_create");
        check(oc_pbd_set_state(h_, floats(X), floats(V)), "oc_pbd_set_state");
    }
    ~PositionBasedDynamics() { oc_pbd_destroy(h_); }
    PositionBasedDynamics(const PositionBasedDynamics&) = delete;
    PositionBasedDynamics& operator=(const PositionBasedDynamics&) = delete;

    /* StepPhysics (PBD:915-925). */
    void StepPhysics(float dt, const oc_ellipsoid* e = nullptr) {
        (void)dt;
        check(oc_pbd_step(h_, 1, e), "oc_pbd_step");
        check(oc_pbd_get_state(h_, floats(X_), floats(V_)), "oc_pbd_get_state");
    }
    void Upload() { check(oc_pbd_set_state(h_, floats(X_), floats(V_)), "oc_pbd_set_state"); }
    oc_pbd handle() const { return h_; }

   private:
    std::vector<Vec3>& X_;
    std::vector<Vec3>& V_;
    oc_pbd h_ = nullptr;
};

/* The GPU Verlet demo's device state (CUH): the constructor does InitCUDA
 * and the first UploadCUDA (CUH:16-25, CUH:53-79), Verlet() is VerletCUDA
 * (CUH:81-100) and the destructor ShutdownCUDA (CUH:27-45).  Vec4 is the
 * caller's four-float vector (glm::vec4 in the demo).  As in the demo, the
 * positions stay on the device between steps and reach the host only
 * through the vertex buffer; Download() copies X and X_last back. */
template <class Vec4>
class GridVerlet {
   public:
    GridVerlet(int texsize_x, int texsize_y, std::vector<Vec4>& X, std::vector<Vec4>& X_last,
               const oc_vgrid_params& params, const uint8_t* pinned = nullptr)
        : X_(X), XL_(X_last) {
        const size_t n = (size_t)texsize_x * (size_t)texsize_y;
        need(X.size() == n && X_last.size() == n, "GridVerlet: X and X_last need texsize_x*texsize_y entries");
        check(oc_vgrid_create(texsize_x, texsize_y, floats4(X), floats4(X_last), pinned, &params, &h_),
              "oc_vgrid_create");
    }
    ~GridVerlet() { oc_vgrid_destroy(h_); }
    GridVerlet(const GridVerlet&) = delete;
    GridVerlet& operator=(const GridVerlet&) = delete;

    /* VerletCUDA: one step; pos_vbo is NULL or a device array of float4 that
     * receives the new positions.  Returns when the step is done. */
    void Verlet(float* pos_vbo = nullptr) {
        check(oc_vgrid_step(h_, 1, pos_vbo), "oc_vgrid_step");
        check(oc_vgrid_synchronize(h_), "oc_vgrid_synchronize");
    }
    /* n steps, stream-ordered (no wait). */
    void Step(int n, float* pos_vbo = nullptr) { check(oc_vgrid_step(h_, n, pos_vbo), "oc_vgrid_step"); }
    void SetParams(const oc_vgrid_params& p) { check(oc_vgrid_set_params(h_, &p), "oc_vgrid_set_params"); }
    void Upload() { check(oc_vgrid_set_state(h_, floats4(X_), floats4(XL_)), "oc_vgrid_set_state"); }
    void Download() { check(oc_vgrid_get_state(h_, floats4(X_), floats4(XL_)), "oc_vgrid_get_state"); }
    oc_vgrid handle() const { return h_; }

   private:
    std::vector<Vec4>& X_;
    std::vector<Vec4>& XL_;
    oc_vgrid h_ = nullptr;
};

}  // namespace opencloth

#endif /* OPENCLOTH_HPP */
