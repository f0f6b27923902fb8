/*
 * opencloth.h -- C ABI of the MI355X implicit mass-spring cloth library.
 *
 * The reference (mmmovania/opencloth) has no library interface: its demos
 * keep the cloth in file-scope globals and call free `void` functions from
 * the GLUT idle loop.  Each entry point below replaces one of those
 * functions for a C or C++ caller; the comment names the function it
 * replaces (paths relative to the reference repository):
 *
 *   IE  = OpenCloth_ImplicitEuler/OpenCloth_ImplicitEuler/main.cpp
 *   BW  = OpenCloth_Implicit/OpenCloth_Implicit/main.cpp
 *   VL  = OpenCloth_Verlet/OpenCloth_Verlet/main.cpp
 *   PBD = OpenCloth_PositionBasedDynamics/OpenCloth_PositionBasedDynamics/main.cpp
 *
 * Conventions
 *   - Host arrays are AoS float3 with a 12-byte stride, the layout of the
 *     reference's vector<glm::vec3> (IE:279-281); the caller owns them and
 *     the library copies in and out.  Device storage is internal SoA.
 *   - Springs use the reference's record (IE:68-73).  Kd keeps the
 *     reference's sign: negative values damp (IE:310-312).
 *   - Every function returns an oc_status; nothing throws across the ABI.
 *   - A handle is used by one thread at a time.  Work is ordered on the
 *     handle's HIP stream (oc_cloth_set_stream); functions that return host
 *     data synchronise that stream.
 *   - Results are deterministic: forces and Jacobian blocks are gathered per
 *     vertex with no atomics and every reduction has a fixed order.
 */
#ifndef OPENCLOTH_H
#define OPENCLOTH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
typedef enum {
    OC_OK = 0,
    OC_ERR_INVALID_ARGUMENT = -1,
    OC_ERR_UNSUPPORTED_TOPOLOGY = -2, /* a spring spans more than the tile halo */
    OC_ERR_DEVICE = -3,               /* a HIP call failed */
    OC_ERR_OUT_OF_MEMORY = -4,
    OC_ERR_NO_DEVICE = -5
} oc_status;

/* solver status bits reported per step (per instance in batches) */
#define OC_SOLVE_NOT_CONVERGED 1 /* tolerance not met within the iteration cap */
#define OC_SOLVE_BREAKDOWN 2     /* d.q <= 0 or non-finite in CG */
#define OC_SOLVE_NONFINITE 4     /* NaN or Inf in the state */

/* CG stopping tests */
#define OC_TOL_ABSOLUTE 0 /* continue while r.r > tol^2 (IE:238) */
#define OC_TOL_RELATIVE 1 /* continue while r.r > tol^2 * r0.r0 (BW:292, BW:319) */
#define OC_TOL_NONE 2     /* run exactly cg_max_iters iterations */

/* oc_params.flags */
#define OC_FLAG_GROUND_CLAMP 1  /* clamp y < 0 to 0 after the update (IE:695-697) */
#define OC_FLAG_CG_PERSISTENT 2 /* run all CG iterations of a solve in one cooperative
                                   launch, resident blocks serving several tiles each
                                   (same arithmetic and results, grid barriers between
                                   phases); ignored where cooperative launch is absent */
#define OC_FLAG_JACOBI 4        /* Jacobi-preconditioned CG: Baraff and Witkin's modified PCG
                                   with P = diag(A), the intent of BW:308-330 and BW:784-785;
                                   the stopping tests then measure r.(P^-1 r) */
#define OC_FLAG_CG_SINGLE_REDUCTION 16 /* the CG recurrence of Chronopoulos and Gear (1989) with one
                                   reduction point per iteration: both dots of an iteration come
                                   from one launch, so an iteration is one launch instead of two.
                                   The same solution in exact arithmetic, other rounding; the
                                   default keeps the reference's recurrence (IE:227-249).
                                   Single cloths only (batches ignore it); takes precedence over
                                   OC_FLAG_CG_PERSISTENT */
#define OC_FLAG_LIST_ORDER 8    /* oc_pbd_params.flags: project the constraints in list order,
                                   one after another as UpdateInternalConstraints does
                                   (PBD:876-889), instead of colour by colour: one workgroup
                                   runs them level by level, a constraint's level being one
                                   more than the highest level among the earlier constraints
                                   that share a particle with it (the reference's results;
                                   for demo-sized cloths) */

/* Spring record, identical to the reference's `struct Spring` (IE:68-73). */
typedef struct {
    int32_t p1, p2;
    float rest_length;
    float Ks, Kd;
    int32_t type;
} oc_spring;

/* Scalars of the implicit step.  The reference keeps them in globals:
 * mass (IE:314), gravity (IE:313, an acceleration here), timeStep (IE:316),
 * EPS and i_max (IE:75-77). */
typedef struct {
    float mass;
    float gravity[3];
    float dt;
    int32_t cg_max_iters;
    float cg_tol;
    int32_t cg_tol_mode; /* OC_TOL_* */
    int32_t warm_start;  /* nonzero: CG starts from the previous step's dv (IE:691, BW:788) */
    int32_t flags;       /* OC_FLAG_* */
    float damping;       /* velocity damping c: every free vertex gets the force c*v, as
                            BW:706-708 adds DEFAULT_DAMPING*V (negative damps; 0 in IE,
                            which comments it out at IE:655); its Jacobian c*I enters
                            A = M - h df/dv - h^2 df/dx, which BW:737-772 omits */
} oc_params;

/* ---- library ----------------------------------------------------------- */
/* Version string of the library. */
const char* oc_version(void);
/* Text of the last error on this thread (empty if none). */
const char* oc_last_error(void);
/* Number of visible HIP devices (0 when none). */
int oc_device_count(void);

/* Tile shape (tile_x by tile_y vertices, one workgroup each) of the
 * single-cloth kernels.  The CG dot products are reduced per tile and then
 * across tiles in a fixed order, so this shape is what a CPU restatement
 * needs to reproduce their bits.  Host only. */
int oc_cloth_tile_shape(int32_t* tile_x, int32_t* tile_y);

/* Render triangles of an nx*ny grid: 2*(nx-1)*(ny-1) triangles, 3 vertex
 * indices each, written to out (host memory, 6*(nx-1)*(ny-1) int32), in the
 * reference's order and winding (IE:473-488) but 32-bit (the reference's
 * GLushort indices overflow past 65,535 vertices, IE:276).  Host only. */
int oc_grid_triangles(int32_t nx, int32_t ny, int32_t* out);

/* ---- implicit cloth (the hot path) ------------------------------------- */
typedef struct oc_cloth_s* oc_cloth;

/* Create a cloth on the current HIP device.  Vertices form an nx*ny grid,
 * vertex k = j*nx + i (IE:462-466).  x0/v0: host AoS 3*nx*ny floats (v0 may
 * be NULL for rest).  springs: the reference spring list (IE:496-530 or any
 * list whose springs join vertices at most 2 rows and 2 columns apart).
 * pinned: nx*ny bytes, nonzero = fixed vertex (IE:652 pins {0, numX}); NULL
 * pins nothing.  Replaces InitGL's state set-up (IE:434-546). */
int oc_cloth_create(int32_t nx, int32_t ny, const float* x0, const float* v0,
                    const oc_spring* springs, int32_t n_springs, const uint8_t* pinned,
                    const oc_params* params, oc_cloth* out);
/* Free a cloth.  Replaces OnShutdown (IE:635-643). */
int oc_cloth_destroy(oc_cloth cloth);

/* Order all later work on `hip_stream` (a hipStream_t; NULL = default stream). */
int oc_cloth_set_stream(oc_cloth cloth, void* hip_stream);
int oc_cloth_set_params(oc_cloth cloth, const oc_params* params);
int oc_cloth_get_params(oc_cloth cloth, oc_params* params);
/* Copy positions/velocities in (either may be NULL).  The reference's mouse
 * drag writes X and V between steps (IE:405-410). */
int oc_cloth_set_state(oc_cloth cloth, const float* x, const float* v);
/* Copy positions/velocities out (either may be NULL); synchronises. */
int oc_cloth_get_state(oc_cloth cloth, float* x, float* v);
/* Copy the CG solution dv out (host AoS); synchronises. */
int oc_cloth_get_dv(oc_cloth cloth, float* dv);
/* Set the CG start value dv (host AoS, may be NULL for zero). */
int oc_cloth_set_dv(oc_cloth cloth, const float* dv);

/* Spring forces plus gravity on every free vertex, zero on pinned ones.
 * Replaces ComputeForces (IE:645-682).  f_out: host AoS 3N (synchronises),
 * or NULL to leave the result on the device. */
int oc_cloth_compute_forces(oc_cloth cloth, float* f_out);

/* Assemble (M - h df/dv - h^2 df/dx) dv = h (f + h df/dx v) at the current
 * state with h = dt, and set up CG from the start value.  The first half of
 * IntegrateImplicit (IE:684-689, in the dv form of BW:775-786). */
int oc_cloth_assemble(oc_cloth cloth, float dt);
/* Solve the assembled system by filtered CG (pinned rows and columns
 * removed).  Replaces SolveConjugateGradient (IE:227-249).  iters/status
 * (either may be NULL) receive the iteration count and OC_SOLVE_* bits;
 * synchronises when either is given. */
int oc_cloth_solve(oc_cloth cloth, int32_t* iters, int32_t* status);
/* v += dv, x += dt*v, optional ground clamp (IE:693-699). */
int oc_cloth_update(oc_cloth cloth, float dt);
/* assemble + solve + update.  Replaces IntegrateImplicit (IE:684-700). */
int oc_cloth_integrate_implicit(oc_cloth cloth, float dt);
/* nsteps implicit steps with the handle's parameters, fully on the device.
 * Replaces StepPhysics (IE:774-779) for the implicit integrator; the
 * reference's post-step collision and Provot passes are separate calls. */
int oc_cloth_step(oc_cloth cloth, int32_t nsteps);
/* oc_cloth_step that allocates nothing.  No counterpart in the reference:
 * for callers that cannot allocate device memory around the step (HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc).  The same
 * launches and results as oc_cloth_step, but the CG partials and scalars
 * come from a fixed device array of 2^20 floats (4 MiB) that each GPU holds
 * until the process ends, instead of the handle's own buffer.  Cloths whose
 * partials need more (2 (tiles + 1)(cg_max_iters + 1) + 2 floats,
 * tiles = ceil(nx/64) * ceil(ny/8)) are rejected with
 * OC_ERR_INVALID_ARGUMENT.  The cloth's device must be current.  Every call
 * on one GPU shares that GPU's array: calls that overlap on one GPU (other
 * streams or threads) race. */
int oc_cloth_step_noalloc(oc_cloth cloth, int32_t nsteps);
/* Wait for the handle's stream. */
int oc_cloth_synchronize(oc_cloth cloth);
/* Status of the last solve: CG iterations and OC_SOLVE_* bits; synchronises. */
int oc_cloth_last_status(oc_cloth cloth, int32_t* iters, int32_t* status);
/* OR of the OC_SOLVE_* bits of every solve and state update since the last
 * call of this function (or since creation), then cleared: a breakdown,
 * non-convergence or non-finite state in any step of a multi-step call,
 * step_physics or rollout is kept here even when a later step is clean.
 * The reference's CG (IE:227-249) reports nothing; synchronises. */
int oc_cloth_sticky_status(oc_cloth cloth, int32_t* status);

/* Measurement (no counterpart in the reference).  Average duration of the
 * CG SpMV launch: assemble at the current state, run one CG iteration, then
 * time `reps` back-to-back launches of the second iteration's SpMV (same
 * inputs each time) between two HIP events on the handle's stream.  x and v
 * are unchanged; dv and the CG vectors are left as that iteration leaves
 * them.  Needs cg_max_iters >= 2; synchronises. */
int oc_cloth_time_cg_spmv(oc_cloth cloth, int32_t reps, float* avg_us);

/* Kernel classes of oc_cloth_time_kernel. */
#define OC_KERNEL_ASSEMBLE 0       /* assembly + CG start */
#define OC_KERNEL_ASSEMBLE_FUSED 1 /* the same with the previous step's update applied
                                      (steps after the first of oc_cloth_step) */
#define OC_KERNEL_CG_SPMV 2        /* CG direction update and q = S A d */
#define OC_KERNEL_CG_AXPY 3        /* CG dv, r update and r.r */
#define OC_KERNEL_UPDATE 4         /* v += dv, x += dt v */
/* Measurement (no counterpart in the reference).  Average duration of one
 * kernel class: assemble at the current state and run one CG iteration,
 * then time `reps` back-to-back launches of the class (the CG kernels of
 * the second iteration, with the stopping test off) between two HIP events
 * on the handle's stream.  x, v, dv and the status words are unchanged (the
 * launches that write them write copies); the other CG vectors are left as
 * the launches leave them.  Needs cg_max_iters >= 2, and a warm start for
 * OC_KERNEL_ASSEMBLE_FUSED; synchronises. */
int oc_cloth_time_kernel(oc_cloth cloth, int32_t kernel, int32_t reps, float* avg_us);

/* ---- post-step passes of StepPhysics (IE:774-779) ----------------------- */
/* Ellipsoid collider: transform and inverse as column-major 4x4 matrices
 * (glm::mat4, IE:541-544), object-space center and radius (IE:85-86). */
typedef struct {
    float transform[16];
    float inverse[16];
    float center[3];
    float radius;
} oc_ellipsoid;

/* Push vertices inside the ellipsoid to its surface and zero their velocity.
 * Replaces EllipsoidCollision (IE:727-751). */
int oc_cloth_ellipsoid_collision(oc_cloth cloth, const oc_ellipsoid* ellipsoid);
/* Move half the excess length of every spring longer than slack*rest into
 * the velocities.  Replaces ApplyProvotDynamicInverse (IE:702-726, slack 1;
 * BW:800-823, slack 1.01). */
int oc_cloth_provot(oc_cloth cloth, float slack);
/* nsteps full steps: implicit step, then the ellipsoid pass (when given) and
 * the Provot pass (when slack > 0).  Replaces StepPhysics (IE:774-779). */
int oc_cloth_step_physics(oc_cloth cloth, int32_t nsteps, const oc_ellipsoid* ellipsoid,
                          float provot_slack);

/* ---- position Verlet (VL) ---------------------------------------------- */
typedef struct oc_verlet_s* oc_verlet;

/* Scalars of the Verlet demo: mass (VL:102), gravity as an acceleration
 * (VL:101, the force is gravity*mass), damping (DEFAULT_DAMPING, VL:97),
 * dt (VL:104) and OC_FLAG_GROUND_CLAMP (VL:440-442). */
typedef struct {
    float mass;
    float gravity[3];
    float damping;
    float dt;
    int32_t flags;
} oc_verlet_params;

/* x_last0 may be NULL (start at rest: x_last = x).  Springs as for
 * oc_cloth_create.  Replaces the Verlet demo's InitGL state (VL:240-328). */
int oc_verlet_create(int32_t nx, int32_t ny, const float* x0, const float* x_last0,
                     const oc_spring* springs, int32_t n_springs, const uint8_t* pinned,
                     const oc_verlet_params* params, oc_verlet* out);
int oc_verlet_destroy(oc_verlet verlet);
int oc_verlet_set_stream(oc_verlet verlet, void* hip_stream);
int oc_verlet_set_state(oc_verlet verlet, const float* x, const float* x_last);
int oc_verlet_get_state(oc_verlet verlet, float* x, float* x_last);
/* Forces with Verlet velocities (x - x_last)/dt.  Replaces ComputeForces(dt)
 * (VL:448-484); f_out host AoS; synchronises. */
int oc_verlet_compute_forces(oc_verlet verlet, float* f_out);
/* nsteps of ComputeForces + IntegrateVerlet (VL:428-444) fused per vertex,
 * then the ellipsoid pass (x_last = x on contact, VL:509-533) when given.
 * Replaces StepPhysics of the Verlet demo (VL:557-562). */
int oc_verlet_step(oc_verlet verlet, int32_t nsteps, const oc_ellipsoid* ellipsoid);

/* ---- grid Verlet: the GPU demos' kernel ---------------------------------- */
/* CUK = OpenCloth_Verlet_CUDA_GLSL_OPENCL_CPU/OpenCloth_Verlet_CUDA/verlet_kernel.cu,
 * CUH = .../verlet.cu (its host glue), CUM = .../main.cpp (the caller).
 * One vertex per texel of a width x height grid, a fixed 12-neighbour
 * stencil (structural +-1, shear diagonals, bend +-2; CUK:14-64) with rest
 * lengths from the grid spacing (CUK:164) and one ks/kd per class, no spring
 * list.  Positions are host AoS float4 (x, y, z, w), as the demo's
 * glm::vec4 arrays (CUM:441); w is carried unchanged. */
typedef struct oc_vgrid_s* oc_vgrid;

typedef struct {
    float mass;        /* CUM:100 */
    float gravity[3];  /* acceleration; the kernel's is (0, -0.00981, 0) (CUK:123) */
    float damping;     /* DEFAULT_DAMPING (CUM:95) */
    float dt;          /* timeStep (CUM:102) */
    float spacing[2];  /* inv_cloth_size = (sizeX/numX, sizeY/numY) (CUM:601) */
    float ks[3];       /* structural, shear, bend (CUK:8-10) */
    float kd[3];       /* the same classes; the demo's are negative (CUK:8-10) */
} oc_vgrid_params;

/* Allocates the device state and uploads x0 and x_last0 (width*height
 * float4 each; x_last0 may be NULL: start at rest).  pinned: one byte per
 * vertex, non-zero = zero acceleration; NULL pins vertices 0 and width-1 as
 * the kernel does (CUK:126-127).  Replaces InitCUDA (CUH:16-25) plus the
 * first UploadCUDA (CUH:53-79). */
int oc_vgrid_create(int32_t width, int32_t height, const float* x0, const float* x_last0,
                    const uint8_t* pinned, const oc_vgrid_params* params, oc_vgrid* out);
/* Replaces ShutdownCUDA (CUH:27-45). */
int oc_vgrid_destroy(oc_vgrid grid);
int oc_vgrid_set_stream(oc_vgrid grid, void* hip_stream);
int oc_vgrid_set_params(oc_vgrid grid, const oc_vgrid_params* params);
/* Host AoS float4 in and out (either may be NULL); get synchronises. */
int oc_vgrid_set_state(oc_vgrid grid, const float* x, const float* x_last);
int oc_vgrid_get_state(oc_vgrid grid, float* x, float* x_last);
/* nsteps of the kernel's step (CUK:70-197): force = gravity*mass +
 * damping*v plus the stencil's springs, ground clamp of the current
 * position, x' = 2x - x_last + a dt^2, x_last' = x.  vbo: NULL or a device
 * array of width*height float4 that receives the last step's positions, as
 * the mapped vertex buffer of VerletCUDA (CUH:81-100, CUM:597-607).  Stream
 * ordered. */
int oc_vgrid_step(oc_vgrid grid, int32_t nsteps, float* vbo);
/* Waits for the grid's stream (VerletCUDA blocks until its kernel is done,
 * CUH:96). */
int oc_vgrid_synchronize(oc_vgrid grid);

/* ---- position-based dynamics (PBD) --------------------------------------- */
/* PBD = OpenCloth_PositionBasedDynamics/OpenCloth_PositionBasedDynamics/main.cpp,
 * built with USE_TRIANGLE_BENDING_CONSTRAINT (PBD:57). */
typedef struct oc_pbd_s* oc_pbd;

/* AddDistanceConstraint(p1, p2, k) (PBD:148-162). */
typedef struct {
    int32_t p1, p2;
    float k;
} oc_pbd_distance;
/* AddBendingConstraint(p1, p2, p3, k), triangle form (PBD:164-178). */
typedef struct {
    int32_t p1, p2, p3;
    float k;
} oc_pbd_bending;

typedef struct {
    float mass;               /* per particle (PBD:114); inverse mass 1/mass, 0 if pinned (PBD:345-351) */
    float gravity[3];         /* acceleration (PBD:112) */
    float dt;                 /* PBD:74 */
    float global_damping;     /* velocity factor per step (PBD:78), capped at 1 (PBD:395-396) */
    float k_damp;             /* centre-of-mass damping (PBD:111), clamped to [0,1] (PBD:391-394) */
    int32_t solver_iterations; /* PBD:107 */
    int32_t flags;            /* OC_FLAG_GROUND_CLAMP: GroundCollision in the solver loop (PBD:887) */
} oc_pbd_params;

/* n_vertices particles at x0 (host AoS float3), pinned may be NULL.  Rest
 * lengths come from x0, k is clamped to [0,1] and k' = 1-(1-k)^(1/ns) as in
 * the reference's InitGL (PBD:148-178, PBD:383-396).  Each constraint family
 * is split into colours (greedy, in list order) whose constraints touch
 * disjoint particles; colours are projected one after another.  Replaces the
 * PBD demo's InitGL state (PBD:310-464). */
int oc_pbd_create(int32_t n_vertices, const float* x0, const uint8_t* pinned,
                  const oc_pbd_distance* distance, int32_t n_distance,
                  const oc_pbd_bending* bending, int32_t n_bending, const oc_pbd_params* params,
                  oc_pbd* out);
int oc_pbd_destroy(oc_pbd pbd);
int oc_pbd_set_stream(oc_pbd pbd, void* hip_stream);
int oc_pbd_set_state(oc_pbd pbd, const float* x, const float* v);
int oc_pbd_get_state(oc_pbd pbd, float* x, float* v);
/* Colours of the distance and bending constraints (with OC_FLAG_LIST_ORDER,
 * their levels). */
int oc_pbd_num_colors(oc_pbd pbd, int32_t* n_distance_colors, int32_t* n_bending_colors);
/* nsteps of StepPhysics (PBD:915-925): gravity, global and centre-of-mass
 * damping, prediction (PBD:601-668), solver_iterations passes of distance
 * and bending projections and the ground clamp (PBD:876-889), the ellipsoid
 * pass on the predicted positions when given (PBD:846-874), then
 * V = (tmp_X - X)/dt, X = tmp_X (PBD:670-678). */
int oc_pbd_step(oc_pbd pbd, int32_t nsteps, const oc_ellipsoid* ellipsoid);
int oc_pbd_synchronize(oc_pbd pbd);

/* ---- differentiable rollouts (no counterpart in the reference) ---------- */
/* Per-spring stiffness Ks (n_springs floats, the spring list's order); the
 * reference fixes Ks per spring type (IE:310-312). */
int oc_cloth_set_spring_stiffness(oc_cloth cloth, const float* ks);
int oc_cloth_get_spring_stiffness(oc_cloth cloth, float* ks);
/* Run nsteps implicit steps from the current state with dv reset to zero,
 * keeping a device checkpoint (x, v, dv) every checkpoint_every steps for
 * oc_cloth_rollout_backward.  The state ends at step nsteps. */
int oc_cloth_rollout_forward(oc_cloth cloth, int32_t nsteps, int32_t checkpoint_every);
/* Hand-written adjoint of the last rollout_forward: given dL/dx and dL/dv at
 * its final state (host AoS; grad_v may be NULL for zero), return dL/dks
 * (n_springs floats), dL/dx0 and dL/dv0 (host AoS; any output may be NULL).
 * Each segment between checkpoints is recomputed forward, then every step's
 * adjoint solves the same symmetric system (filtered CG) and gathers the
 * derivatives of the spring forces and Jacobians per vertex.  The state is
 * restored to the rollout's final state; synchronises. */
int oc_cloth_rollout_backward(oc_cloth cloth, const float* grad_x, const float* grad_v,
                              float* grad_ks, float* grad_x0, float* grad_v0);

/* Measurement hook (no counterpart in the reference, whose only timing is a
 * frame counter, IE:562-584): runs nsteps exactly as oc_cloth_step does
 * (same launches, same results), each kernel launched with a start and a
 * stop event attached to its own dispatch, and returns per kernel class
 * (0 assembly, 1 first CG launch of an iteration (cg_spmv), 2 second
 * (cg_axpy), 3 other solver launches, 4 update) the summed kernel
 * milliseconds and the launch counts.  ms[5], launches[5]; synchronises. */
int oc_cloth_profile_steps(oc_cloth cloth, int32_t nsteps, float* ms, int32_t* launches);

/* Introspection of the assembled system (for tests and tools).
 * Pairs are the unique unordered vertex pairs joined by springs, numbered by
 * their first spring; a[p] < b[p]. */
int oc_cloth_num_pairs(oc_cloth cloth, int32_t* n_pairs);
int oc_cloth_get_pairs(oc_cloth cloth, int32_t* a, int32_t* b);
/* Off-diagonal blocks C_ab = A_ab of every pair as 6 floats
 * (xx,yy,zz,xy,xz,yz), and the right-hand side h(f + h df/dx v) with pinned
 * rows zero (host AoS); either may be NULL; synchronises. */
int oc_cloth_get_system(oc_cloth cloth, float* blocks, float* rhs);

/* ---- batched cloths ---------------------------------------------------- */
/* Many independent cloths of uneven sizes stepped together: one persistent
 * kernel launch per step runs assembly, filtered CG (per-cloth stopping
 * test, converged cloths masked out) and the update for every cloth.  No
 * counterpart in the reference, whose demos each step one global cloth; each
 * cloth follows the same step as oc_cloth_step. */
typedef struct oc_batch_s* oc_batch;

/* Schedule of the persistent kernel: workgroup size (256, 512 or 1024) and
 * the target vertices per thread, which sets the threads serving one cloth
 * (a power of two from 64 to block_threads); cloths served by fewer threads
 * than a workgroup are packed several to a workgroup. */
typedef struct {
    int32_t block_threads;
    int32_t vertices_per_thread;
} oc_batch_schedule;

/* n_cloths grids of nx[c]*ny[c] vertices.  x0/v0/pinned: all cloths'
 * vertices concatenated in cloth order (AoS float3 / bytes; v0 and pinned may
 * be NULL).  springs: all cloths' spring lists concatenated, n_springs[c]
 * records for cloth c, whose p1/p2 index that cloth's own vertices. */
int oc_batch_create(int32_t n_cloths, const int32_t* nx, const int32_t* ny, const float* x0,
                    const float* v0, const oc_spring* springs, const int32_t* n_springs,
                    const uint8_t* pinned, const oc_params* params, oc_batch* out);
int oc_batch_destroy(oc_batch batch);
int oc_batch_set_stream(oc_batch batch, void* hip_stream);
int oc_batch_set_params(oc_batch batch, const oc_params* params);
int oc_batch_set_schedule(oc_batch batch, const oc_batch_schedule* schedule);
int oc_batch_get_schedule(oc_batch batch, oc_batch_schedule* schedule);
/* Time every candidate schedule (the fastest of three steps from the same
 * saved state), keep the fastest and return it in *best (may be NULL); the
 * state is restored; synchronises. */
int oc_batch_autotune(oc_batch batch, oc_batch_schedule* best);
/* Threads that serve each cloth under the current schedule (n_cloths ints). */
int oc_batch_threads_per_cloth(oc_batch batch, int32_t* threads);
/* Copy positions/velocities of all cloths in (host AoS, vertex offsets in
 * cloth order; either may be NULL, which leaves it unchanged).  New
 * positions also clear the CG warm start, so the next solve starts at zero. */
int oc_batch_set_state(oc_batch batch, const float* x, const float* v);
int oc_batch_get_state(oc_batch batch, float* x, float* v);
/* nsteps steps, each a replay of the captured HIP graph of one step. */
int oc_batch_step(oc_batch batch, int32_t nsteps);
/* nsteps steps launched directly (no graph). */
int oc_batch_step_eager(oc_batch batch, int32_t nsteps);
/* CG iterations and OC_SOLVE_* bits of every cloth's last step (or, after
 * oc_batch_rollout_backward, of its last adjoint solve); synchronises. */
int oc_batch_get_status(oc_batch batch, int32_t* iters, int32_t* status);
/* Per cloth (n_cloths ints), the OR of the OC_SOLVE_* bits of every step
 * and adjoint solve since the last call of this function (or since
 * creation), then cleared; synchronises. */
int oc_batch_sticky_status(oc_batch batch, int32_t* status);
/* Per-spring stiffness of all cloths (concatenated spring lists). */
int oc_batch_set_spring_stiffness(oc_batch batch, const float* ks);
int oc_batch_get_spring_stiffness(oc_batch batch, float* ks);
/* Batched differentiable rollout: as oc_cloth_rollout_forward/backward for
 * every cloth at once; each step's adjoint is one persistent launch that
 * re-assembles, solves the adjoint system per cloth (masked CG) and gathers
 * the derivatives.  Arrays are all cloths' vertices/springs concatenated. */
int oc_batch_rollout_forward(oc_batch batch, int32_t nsteps, int32_t checkpoint_every);
int oc_batch_rollout_backward(oc_batch batch, const float* grad_x, const float* grad_v,
                              float* grad_ks, float* grad_x0, float* grad_v0);
int oc_batch_synchronize(oc_batch batch);

#ifdef __cplusplus
}
#endif

#endif /* OPENCLOTH_H */
