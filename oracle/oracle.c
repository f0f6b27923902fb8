/*
 * oracle.c -- CPU restatement of OpenCloth's implicit mass-spring step.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP library in
 * opencloth_amd/.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it; the product never links or calls it.
 *
 * The same source is compiled twice (see oracle/Makefile):
 *   liboracle32.so  OC_REAL=float   -ffp-contract=off  (FP32 restatement)
 *   liboracle64.so  OC_REAL=double                     (FP64 accuracy oracle)
 *
 * Reference files followed (paths relative to the reference repository):
 *   IE  = OpenCloth_ImplicitEuler/OpenCloth_ImplicitEuler/main.cpp
 *   BW  = OpenCloth_Implicit/OpenCloth_Implicit/main.cpp
 *   VL  = OpenCloth_Verlet/OpenCloth_Verlet/main.cpp
 *   PBD = OpenCloth_PositionBasedDynamics/OpenCloth_PositionBasedDynamics/main.cpp
 *   GLM = the .inl files of dep/glm/glm/core (GLM 0.9.0.0, the arithmetic of
 *         every vec3 and mat3 operation on the path)
 *
 * Arithmetic conventions (GLM 0.9.0; pinned bit-for-bit by
 * tests/test_oracle_glm.py against the vendored GLM headers):
 *   dot(a,b)      = a.x*b.x + a.y*b.y + a.z*b.z              GLM func_geometric.inl:148
 *   length(a)     = sqrt(a.x*a.x + a.y*a.y + a.z*a.z)        GLM func_geometric.inl:49-50
 *   normalize(a)  = a * (1 / sqrt(a.x*a.x+a.y*a.y+a.z*a.z))  GLM func_geometric.inl:228-229,
 *                                                            func_exponential.inl:316
 *   mat3*vec3     = row r: m[0][r]*x + m[1][r]*y + m[2][r]*z  GLM type_mat3x3.inl:508-518
 *   outerProduct(c,r)[i][j] = c[j]*r[i]                      GLM func_matrix.inl:51-63
 * Symmetric 3x3 blocks are stored as 6 values (xx, yy, zz, xy, xz, yz).
 *
 * Each exported function's comment says what pins it ("pinned by"): the
 * independent checks in tests/ that fail on a dropped term, a wrong sign or
 * index, or a transposed operand.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifndef OC_REAL
#define OC_REAL float
#define OC_IS_DOUBLE 0
#endif
typedef OC_REAL real;

#if OC_IS_DOUBLE
#define SQRT_(x) sqrt(x)
#define ISFINITE_(x) isfinite(x)
#else
#define SQRT_(x) sqrtf(x)
#define ISFINITE_(x) isfinite(x)
#endif

#define EXPORT __attribute__((visibility("default")))

/* Worker threads of the per-vertex, per-spring and per-tile loops (OpenMP
 * when built with -fopenmp; 1 by default).  Every parallel loop writes
 * disjoint outputs and every sum keeps its serial order (per-pair sums in
 * list order, dot-product tiles combined in tile order, mode 0 dots and the
 * batched order serial), so results are bitwise independent of the count.
 * Pinned by: tests/test_oracle_threads.py (1 vs 4 threads, bit for bit). */
static int g_threads = 1;
#ifdef _OPENMP
#define OMP_FOR _Pragma("omp parallel for schedule(static) num_threads(g_threads) if (g_threads > 1)")
#else
#define OMP_FOR
#endif

EXPORT int orc_set_threads(int n) {
    int old = g_threads;
    g_threads = n > 0 ? n : 1;
    return old;
}

/* Spring record in the reference's own layout (IE:68-73, 24 bytes).
 * Kd follows the reference's sign convention: negative for damping
 * (IE:310-312), so a damping coefficient kd > 0 is stored as Kd = -kd. */
typedef struct {
    int32_t p1, p2;
    float rest_length, Ks, Kd;
    int32_t type;
} orc_spring;

/* Scalar parameters of one implicit step.  Doubles so that one ctypes layout
 * serves both builds; each is rounded to `real` once on entry. */
typedef struct {
    double mass, gx, gy, gz, h, tol;
    int32_t max_it;     /* CG iteration cap (IE:77 i_max) */
    int32_t tol_mode;   /* 0: stop when r.r <= tol^2 (IE:238); 1: when r.r <= tol^2 * r0.r0
                           (BW:292, BW:319); 2: no test, run max_it iterations */
    int32_t warm_start; /* 1: CG starts from the previous step's dv (IE:691, BW:788) */
    int32_t dot_mode;   /* 0: sequential dots (IE:219-225); 1: tile-tree order of the
                           single-cloth HIP kernels; 2: order of the batched kernels */
    int32_t nx, ny, tx, ty; /* grid and tile shape (dot_mode 1) */
    int32_t flags;      /* bit0: clamp y<0 to 0 after the update (IE:695-697) */
    int32_t pad;
    double damping;     /* velocity damping c: f_i += c*v_i on free vertices (BW:706-708
                           DEFAULT_DAMPING, negative damps), with its Jacobian c*I in A */
} orc_params;

/* Diagonal mass of A: m - h*c, the velocity damping's term of
 * A = M - h df/dv - h^2 df/dx (m when c = 0, exactly). */
static real diag_mass(const orc_params* prm) {
    return (real)prm->mass - (real)prm->h * (real)prm->damping;
}

#define ORC_NOT_CONVERGED 1
#define ORC_BREAKDOWN 2
#define ORC_NONFINITE 4

/* orc_params.flags */
#define ORC_FLAG_GROUND_CLAMP 1
#define ORC_FLAG_JACOBI 4 /* Jacobi-preconditioned CG (Baraff-Witkin modified PCG) */
#define ORC_FLAG_CG1 16   /* single-reduction CG (Chronopoulos and Gear 1989) */

/* ------------------------------------------------------------------------ */
/* vec3 / symmetric-block helpers in GLM evaluation order                    */
/* ------------------------------------------------------------------------ */
static inline real dot3(const real* a, const real* b) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
static inline real len3(const real* a) {
    real s = a[0] * a[0] + a[1] * a[1] + a[2] * a[2];
    return SQRT_(s);
}
/* symmetric block (xx,yy,zz,xy,xz,yz) times vector in mat3*vec3 order */
static inline void symmv(const real* c, const real* w, real* out) {
    out[0] = c[0] * w[0] + c[3] * w[1] + c[4] * w[2];
    out[1] = c[3] * w[0] + c[1] * w[1] + c[5] * w[2];
    out[2] = c[4] * w[0] + c[5] * w[1] + c[2] * w[2];
}

/* ------------------------------------------------------------------------ */
/* Scene: grid positions and the reference spring list                       */
/* ------------------------------------------------------------------------ */

/* IE:462-466.  Vertex k = j*nx + i sits at
 * (((i/(nx-1))*2-1)*halfsize, y0, (j/(ny-1))*fullsize); the reference uses
 * y0 = fullsize+1.  Pinned by: tests/test_oracle_scene.py (closed form). */
EXPORT void orc_grid_positions(int nx, int ny, float fullsize, float y0, float* X) {
    float halfsize = fullsize / 2.0f;
    int count = 0;
    for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++) {
            X[3 * count + 0] = (((float)i / (nx - 1)) * 2 - 1) * halfsize;
            X[3 * count + 1] = y0;
            X[3 * count + 2] = (((float)j / (ny - 1)) * fullsize);
            count++;
        }
}

/* AddSpring (IE:336-346): rest length sqrt(dot(dp,dp)) of the given positions. */
static void add_spring(const float* X, int a, int b, float ks, float kd, int type,
                       orc_spring* out, int* n) {
    if (out) {
        float dx = X[3 * a + 0] - X[3 * b + 0];
        float dy = X[3 * a + 1] - X[3 * b + 1];
        float dz = X[3 * a + 2] - X[3 * b + 2];
        orc_spring s;
        s.p1 = a;
        s.p2 = b;
        s.Ks = ks;
        s.Kd = kd;
        s.type = type;
        s.rest_length = sqrtf(dx * dx + dy * dy + dz * dz);
        out[*n] = s;
    }
    (*n)++;
}

/* IE:496-530.  Spring list of an nx*ny vertex grid in the reference's order:
 * horizontal structural, vertical structural, two shears per quad, bend rows
 * (each row's last bend added twice, IE:521+523), bend columns (likewise,
 * IE:527+529).  Returns the count; writes nothing when out is NULL.
 * Pinned by: tests/test_oracle_scene.py (count formula 4s(s-1)+2(s-1)^2,
 * duplicate positions, pair offsets, rest lengths against the spacing). */
EXPORT int orc_grid_springs(int nx, int ny, const float* X, float ks_struct, float kd_struct,
                            float ks_shear, float kd_shear, float ks_bend, float kd_bend,
                            orc_spring* out) {
    int n = 0;
    int v = ny, u = nx;
    for (int l1 = 0; l1 < v; l1++)
        for (int l2 = 0; l2 < u - 1; l2++)
            add_spring(X, l1 * u + l2, l1 * u + l2 + 1, ks_struct, kd_struct, 0, out, &n);
    for (int l1 = 0; l1 < u; l1++)
        for (int l2 = 0; l2 < v - 1; l2++)
            add_spring(X, l2 * u + l1, (l2 + 1) * u + l1, ks_struct, kd_struct, 0, out, &n);
    for (int l1 = 0; l1 < v - 1; l1++)
        for (int l2 = 0; l2 < u - 1; l2++) {
            add_spring(X, l1 * u + l2, (l1 + 1) * u + l2 + 1, ks_shear, kd_shear, 1, out, &n);
            add_spring(X, (l1 + 1) * u + l2, l1 * u + l2 + 1, ks_shear, kd_shear, 1, out, &n);
        }
    for (int l1 = 0; l1 < v; l1++) {
        for (int l2 = 0; l2 < u - 2; l2++)
            add_spring(X, l1 * u + l2, l1 * u + l2 + 2, ks_bend, kd_bend, 2, out, &n);
        add_spring(X, l1 * u + (u - 3), l1 * u + (u - 1), ks_bend, kd_bend, 2, out, &n);
    }
    for (int l1 = 0; l1 < u; l1++) {
        for (int l2 = 0; l2 < v - 2; l2++)
            add_spring(X, l2 * u + l1, (l2 + 2) * u + l1, ks_bend, kd_bend, 2, out, &n);
        add_spring(X, (v - 3) * u + l1, (v - 1) * u + l1, ks_bend, kd_bend, 2, out, &n);
    }
    return n;
}

/* ------------------------------------------------------------------------ */
/* System topology                                                          */
/* ------------------------------------------------------------------------ */
/* Unique unordered vertex pairs are numbered by their first spring in list
 * order.  Each vertex sees its incident springs grouped by pair (pairs by
 * number, springs by id inside a pair): for the reference grids, whose only
 * duplicate springs are adjacent in the list, this is exactly the spring
 * order of the reference's scatter loop (IE:658-681). */
typedef struct {
    int N, S, P;
    orc_spring* springs;
    real* ks; /* per-spring stiffness in the oracle's precision (orc_set_ks) */
    uint8_t* pinned;
    int* pair_a; /* lower vertex of each pair */
    int* pair_b;
    int* spring_pair;
    int* inc_ptr; /* N+1 */
    int* inc_s;   /* incident springs, pair-grouped */
    int* vp_ptr;  /* N+1 */
    int* vp_pair; /* incident pairs by pair number */
} orc_sys;

typedef struct {
    int64_t key;
    int id;
} keyid;

static int cmp_keyid(const void* a, const void* b) {
    const keyid* x = (const keyid*)a;
    const keyid* y = (const keyid*)b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    return x->id - y->id;
}

EXPORT orc_sys* orc_create(int N, int S, const orc_spring* springs, const uint8_t* pinned) {
    orc_sys* s = (orc_sys*)calloc(1, sizeof(orc_sys));
    s->N = N;
    s->S = S;
    s->springs = (orc_spring*)malloc(sizeof(orc_spring) * (S > 0 ? S : 1));
    memcpy(s->springs, springs, sizeof(orc_spring) * S);
    s->ks = (real*)malloc(sizeof(real) * (S > 0 ? S : 1));
    for (int i = 0; i < S; i++) s->ks[i] = (real)springs[i].Ks;
    s->pinned = (uint8_t*)malloc(N > 0 ? N : 1);
    if (pinned)
        memcpy(s->pinned, pinned, N);
    else
        memset(s->pinned, 0, N);

    /* group springs by unordered pair, number pairs by first spring id */
    keyid* k = (keyid*)malloc(sizeof(keyid) * (S > 0 ? S : 1));
    for (int i = 0; i < S; i++) {
        int a = springs[i].p1, b = springs[i].p2;
        int lo = a < b ? a : b, hi = a < b ? b : a;
        k[i].key = (int64_t)lo * N + hi;
        k[i].id = i;
    }
    qsort(k, S, sizeof(keyid), cmp_keyid);
    /* first id of each group; group representative */
    int* grp_first = (int*)malloc(sizeof(int) * (S > 0 ? S : 1));
    int ng = 0;
    int* spring_grp = (int*)malloc(sizeof(int) * (S > 0 ? S : 1));
    for (int i = 0; i < S; i++) {
        if (i == 0 || k[i].key != k[i - 1].key) grp_first[ng++] = k[i].id;
        spring_grp[k[i].id] = ng - 1;
    }
    /* order groups by first id */
    keyid* g = (keyid*)malloc(sizeof(keyid) * (ng > 0 ? ng : 1));
    for (int i = 0; i < ng; i++) {
        g[i].key = grp_first[i];
        g[i].id = i;
    }
    qsort(g, ng, sizeof(keyid), cmp_keyid);
    int* grp_to_pair = (int*)malloc(sizeof(int) * (ng > 0 ? ng : 1));
    for (int i = 0; i < ng; i++) grp_to_pair[g[i].id] = i;
    s->P = ng;
    s->pair_a = (int*)malloc(sizeof(int) * (ng > 0 ? ng : 1));
    s->pair_b = (int*)malloc(sizeof(int) * (ng > 0 ? ng : 1));
    s->spring_pair = (int*)malloc(sizeof(int) * (S > 0 ? S : 1));
    for (int i = 0; i < S; i++) {
        int p = grp_to_pair[spring_grp[i]];
        s->spring_pair[i] = p;
        int a = springs[i].p1, b = springs[i].p2;
        s->pair_a[p] = a < b ? a : b;
        s->pair_b[p] = a < b ? b : a;
    }
    free(k);
    free(grp_first);
    free(spring_grp);
    free(g);
    free(grp_to_pair);

    /* per-vertex incidence, pair-grouped */
    s->inc_ptr = (int*)calloc(N + 1, sizeof(int));
    for (int i = 0; i < S; i++) {
        s->inc_ptr[springs[i].p1 + 1]++;
        s->inc_ptr[springs[i].p2 + 1]++;
    }
    for (int i = 0; i < N; i++) s->inc_ptr[i + 1] += s->inc_ptr[i];
    s->inc_s = (int*)malloc(sizeof(int) * (2 * S > 0 ? 2 * S : 1));
    int* fill = (int*)calloc(N > 0 ? N : 1, sizeof(int));
    for (int i = 0; i < S; i++) {
        int a = springs[i].p1, b = springs[i].p2;
        s->inc_s[s->inc_ptr[a] + fill[a]++] = i;
        s->inc_s[s->inc_ptr[b] + fill[b]++] = i;
    }
    free(fill);
    /* insertion sort each row by (pair, spring id) */
    for (int v = 0; v < N; v++) {
        int b0 = s->inc_ptr[v], b1 = s->inc_ptr[v + 1];
        for (int i = b0 + 1; i < b1; i++) {
            int x = s->inc_s[i];
            int j = i - 1;
            while (j >= b0) {
                int y = s->inc_s[j];
                int px = s->spring_pair[x], py = s->spring_pair[y];
                if (py > px || (py == px && y > x)) {
                    s->inc_s[j + 1] = y;
                    j--;
                } else
                    break;
            }
            s->inc_s[j + 1] = x;
        }
    }
    /* per-vertex pair lists */
    s->vp_ptr = (int*)calloc(N + 1, sizeof(int));
    s->vp_pair = (int*)malloc(sizeof(int) * (2 * S > 0 ? 2 * S : 1));
    int nvp = 0;
    for (int v = 0; v < N; v++) {
        s->vp_ptr[v] = nvp;
        int last = -1;
        for (int e = s->inc_ptr[v]; e < s->inc_ptr[v + 1]; e++) {
            int p = s->spring_pair[s->inc_s[e]];
            if (p != last) {
                s->vp_pair[nvp++] = p;
                last = p;
            }
        }
    }
    s->vp_ptr[N] = nvp;
    return s;
}

EXPORT void orc_destroy(orc_sys* s) {
    if (!s) return;
    free(s->springs);
    free(s->ks);
    free(s->pinned);
    free(s->pair_a);
    free(s->pair_b);
    free(s->spring_pair);
    free(s->inc_ptr);
    free(s->inc_s);
    free(s->vp_ptr);
    free(s->vp_pair);
    free(s);
}

EXPORT int orc_num_pairs(const orc_sys* s) { return s->P; }

/* Per-spring stiffness in the oracle's own precision (FP64 finite
 * differences in ks need steps finer than float resolution). */
EXPORT void orc_set_ks(orc_sys* s, const double* ks) {
    for (int i = 0; i < s->S; i++) s->ks[i] = (real)ks[i];
}
EXPORT void orc_pairs(const orc_sys* s, int* a, int* b, int* spring_pair) {
    memcpy(a, s->pair_a, sizeof(int) * s->P);
    memcpy(b, s->pair_b, sizeof(int) * s->P);
    memcpy(spring_pair, s->spring_pair, sizeof(int) * s->S);
}

/* ------------------------------------------------------------------------ */
/* Per-spring terms                                                          */
/* ------------------------------------------------------------------------ */

/* Force on p1 of one spring (IE:659-675):
 *   dp = x1-x2, dv = v1-v2, l = length(dp)
 *   f  = (-Ks*(l-L0) + Kd*(dot(dv,dp)/l)) * normalize(dp)
 * and, when K is given, the stiffness Jacobian df1/dx1 (IE:669-671)
 *   K  = Ks*(-I + (L0/l)*(I - outerProduct(dp,dp)/(l*l)))
 * and, when P is given, outerProduct(dp,dp)/(l*l) (used for df1/dv1 = Kd*P).
 * Pinned by: tests/test_oracle_physics.py -- GLM-built operations (bitwise),
 * central finite differences of f in x and v (FP64), zero force at rest
 * length with zero relative velocity, and the eigen-structure of K
 * (eigenvalue -Ks along dp, -Ks*(1-L0/l) across). */
static void spring_terms(const real* x1, const real* x2, const real* v1, const real* v2,
                         real L0, real Ks, real Kd, real* f, real* K, real* P) {
    real dp[3] = {x1[0] - x2[0], x1[1] - x2[1], x1[2] - x2[2]};
    real dv[3] = {v1[0] - v2[0], v1[1] - v2[1], v1[2] - v2[2]};
    real dist = len3(dp);
    if (K || P) {
        real dist2 = dist * dist;
        real lo_l = L0 / dist;
        real p[6];
        p[0] = (dp[0] * dp[0]) / dist2;
        p[1] = (dp[1] * dp[1]) / dist2;
        p[2] = (dp[2] * dp[2]) / dist2;
        p[3] = (dp[1] * dp[0]) / dist2;
        p[4] = (dp[2] * dp[0]) / dist2;
        p[5] = (dp[2] * dp[1]) / dist2;
        if (K) {
            for (int c = 0; c < 3; c++) K[c] = Ks * ((real)-1 + lo_l * ((real)1 - p[c]));
            for (int c = 3; c < 6; c++) K[c] = Ks * (lo_l * ((real)0 - p[c]));
        }
        if (P)
            for (int c = 0; c < 6; c++) P[c] = p[c];
    }
    real leftTerm = -Ks * (dist - L0);
    real rightTerm = Kd * (dot3(dv, dp) / dist);
    real inv = (real)1 / SQRT_(dp[0] * dp[0] + dp[1] * dp[1] + dp[2] * dp[2]);
    real s = leftTerm + rightTerm;
    f[0] = s * (dp[0] * inv);
    f[1] = s * (dp[1] * inv);
    f[2] = s * (dp[2] * inv);
}

/* Per-spring force, stiffness Jacobian (sym-6) and P (sym-6) for all springs,
 * for fixtures and the per-spring tests.  Any output may be NULL. */
EXPORT void orc_spring_terms(const orc_sys* s, const real* X, const real* V, real* Fs, real* Ks6,
                             real* P6) {
    for (int i = 0; i < s->S; i++) {
        const orc_spring* sp = &s->springs[i];
        real f[3], K[6], P[6];
        spring_terms(X + 3 * sp->p1, X + 3 * sp->p2, V + 3 * sp->p1, V + 3 * sp->p2,
                     (real)sp->rest_length, s->ks[sp - s->springs], (real)sp->Kd, f, K, P);
        if (Fs) memcpy(Fs + 3 * i, f, sizeof(f));
        if (Ks6) memcpy(Ks6 + 6 * i, K, sizeof(K));
        if (P6) memcpy(P6 + 6 * i, P, sizeof(P));
    }
}

/* ------------------------------------------------------------------------ */
/* Forces                                                                    */
/* ------------------------------------------------------------------------ */

/* ComputeForces (IE:645-682) in the reference's scatter order, with the
 * hard-coded pins {0, numX} (IE:652, IE:677-680) generalised to a pin mask
 * and gravity given as an acceleration (force mass*g; the reference adds
 * `gravity` as a force, the same thing at mass 1).
 * Pinned by: equality with orc_forces_gather (an independent loop order) and
 * the physics tests of spring_terms. */
EXPORT void orc_forces_scatter(const orc_sys* s, const real* X, const real* V, double mass,
                               const double* g, double damping, real* F) {
    real m = (real)mass, dmp = (real)damping;
    real gv[3] = {(real)g[0], (real)g[1], (real)g[2]};
    for (int i = 0; i < s->N; i++) {
        F[3 * i + 0] = F[3 * i + 1] = F[3 * i + 2] = 0;
        if (!s->pinned[i])
            for (int c = 0; c < 3; c++) F[3 * i + c] += m * gv[c];
        /* BW:708 adds DEFAULT_DAMPING*V[i] after gravity */
        if (!s->pinned[i] && damping != 0)
            for (int c = 0; c < 3; c++) F[3 * i + c] += dmp * V[3 * i + c];
    }
    for (int i = 0; i < s->S; i++) {
        const orc_spring* sp = &s->springs[i];
        real f[3];
        spring_terms(X + 3 * sp->p1, X + 3 * sp->p2, V + 3 * sp->p1, V + 3 * sp->p2,
                     (real)sp->rest_length, s->ks[sp - s->springs], (real)sp->Kd, f, NULL, NULL);
        if (!s->pinned[sp->p1])
            for (int c = 0; c < 3; c++) F[3 * sp->p1 + c] += f[c];
        if (!s->pinned[sp->p2])
            for (int c = 0; c < 3; c++) F[3 * sp->p2 + c] -= f[c];
    }
}

/* Gather form of the same forces: each vertex sums its incident springs in
 * pair-grouped order (the order the HIP assembly kernel uses).  Also returns
 * KV = (df/dx) v, the stiffness Jacobian applied to the velocities. */
EXPORT void orc_forces_gather(const orc_sys* s, const real* X, const real* V, double mass,
                              const double* g, double damping, real* F, real* KV) {
    real m = (real)mass, dmp = (real)damping;
    real gv[3] = {(real)g[0], (real)g[1], (real)g[2]};
    for (int i = 0; i < s->N; i++) {
        real acc[3] = {0, 0, 0}, kv[3] = {0, 0, 0};
        if (!s->pinned[i])
            for (int c = 0; c < 3; c++) acc[c] += m * gv[c];
        if (!s->pinned[i] && damping != 0)
            for (int c = 0; c < 3; c++) acc[c] += dmp * V[3 * i + c];
        for (int e = s->inc_ptr[i]; e < s->inc_ptr[i + 1]; e++) {
            const orc_spring* sp = &s->springs[s->inc_s[e]];
            real f[3], K[6];
            spring_terms(X + 3 * sp->p1, X + 3 * sp->p2, V + 3 * sp->p1, V + 3 * sp->p2,
                         (real)sp->rest_length, s->ks[sp - s->springs], (real)sp->Kd, f, K, NULL);
            real w[3] = {V[3 * sp->p1 + 0] - V[3 * sp->p2 + 0], V[3 * sp->p1 + 1] - V[3 * sp->p2 + 1],
                         V[3 * sp->p1 + 2] - V[3 * sp->p2 + 2]};
            real kw[3];
            symmv(K, w, kw);
            if (sp->p1 == i) {
                if (!s->pinned[i])
                    for (int c = 0; c < 3; c++) acc[c] += f[c];
                for (int c = 0; c < 3; c++) kv[c] += kw[c];
            } else {
                if (!s->pinned[i])
                    for (int c = 0; c < 3; c++) acc[c] -= f[c];
                for (int c = 0; c < 3; c++) kv[c] -= kw[c];
            }
        }
        for (int c = 0; c < 3; c++) {
            F[3 * i + c] = acc[c];
            if (KV) KV[3 * i + c] = kv[c];
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Assembly of (M - h df/dv - h^2 df/dx) dv = h (f + h df/dx v)              */
/* ------------------------------------------------------------------------ */
/* Off-diagonal block of pair (i,j):  C_ij = sum_s (h*Kd_s*P_s + h^2*K_s), the
 * sum over the pair's springs in list order.  Diagonal block:
 * A_ii = (m - h*c)*I - sum_j C_ij, never stored (c the velocity damping,
 * usually 0; diag_mass): A d is evaluated as
 *   (A d)_i = (m - h*c)*d_i + sum_j C_ij (d_j - d_i)      (pairs in vertex order).
 * df/dv of one spring is Kd*P (the derivative of the force the reference
 * applies, IE:674); BW:768-770's -Kd*(dp1+dp2+dp3) has the opposite sign
 * with BW's own negative Kd and is not followed.  df/dx is the reference's
 * stiffness Jacobian K (IE:671). */
static void spring_block(const orc_sys* s, const orc_spring* sp, const real* X, const real* V, real h, real hh,
                         real* f, real* K, real* C) {
    real P[6];
    spring_terms(X + 3 * sp->p1, X + 3 * sp->p2, V + 3 * sp->p1, V + 3 * sp->p2,
                 (real)sp->rest_length, s->ks[sp - s->springs], (real)sp->Kd, f, K, P);
    real kd = (real)sp->Kd;
    for (int c = 0; c < 6; c++) C[c] = h * (kd * P[c]) + hh * K[c];
}

/* Assemble the right-hand side B (filtered: zero on pins), the pair blocks
 * C[P][6], and the Jacobi diagonal diag(A) (may be NULL).
 * Pinned by: tests/test_oracle_physics.py -- dense A built from these blocks
 * equals M - h*J_v - h^2*J_x with J_v, J_x from central finite differences of
 * the force function (FP64); A symmetric; row sums of the stiffness part zero;
 * B against an explicit dense evaluation. */
EXPORT void orc_assemble(const orc_sys* s, const real* X, const real* V, const orc_params* prm,
                         real* B, real* Cp, real* diag) {
    real h = (real)prm->h, hh = h * h, m = (real)prm->mass, md = diag_mass(prm);
    real dmp = (real)prm->damping;
    real gv[3] = {(real)prm->gx, (real)prm->gy, (real)prm->gz};
    /* pair blocks, in list order within each pair (each spring's block is
     * evaluated first, then the blocks are added serially) */
    real* Cs = (real*)malloc(sizeof(real) * 6 * (s->S > 0 ? s->S : 1));
    OMP_FOR
    for (int i = 0; i < s->S; i++) {
        real f[3], K[6];
        spring_block(s, &s->springs[i], X, V, h, hh, f, K, Cs + 6 * i);
    }
    for (int p = 0; p < s->P; p++)
        for (int c = 0; c < 6; c++) Cp[6 * p + c] = 0;
    for (int i = 0; i < s->S; i++) {
        int p = s->spring_pair[i];
        for (int c = 0; c < 6; c++) Cp[6 * p + c] += Cs[6 * i + c];
    }
    free(Cs);
    OMP_FOR
    for (int i = 0; i < s->N; i++) {
        real acc[3] = {0, 0, 0}, kv[3] = {0, 0, 0};
        if (!s->pinned[i])
            for (int c = 0; c < 3; c++) acc[c] += m * gv[c];
        if (!s->pinned[i] && prm->damping != 0)
            for (int c = 0; c < 3; c++) acc[c] += dmp * V[3 * i + c];
        for (int e = s->inc_ptr[i]; e < s->inc_ptr[i + 1]; e++) {
            const orc_spring* sp = &s->springs[s->inc_s[e]];
            real f[3], K[6], C[6];
            spring_block(s, sp, X, V, h, hh, f, K, C);
            real w[3] = {V[3 * sp->p1 + 0] - V[3 * sp->p2 + 0], V[3 * sp->p1 + 1] - V[3 * sp->p2 + 1],
                         V[3 * sp->p1 + 2] - V[3 * sp->p2 + 2]};
            real kw[3];
            symmv(K, w, kw);
            if (sp->p1 == i) {
                for (int c = 0; c < 3; c++) acc[c] += f[c];
                for (int c = 0; c < 3; c++) kv[c] += kw[c];
            } else {
                for (int c = 0; c < 3; c++) acc[c] -= f[c];
                for (int c = 0; c < 3; c++) kv[c] -= kw[c];
            }
        }
        for (int c = 0; c < 3; c++) {
            real t = h * kv[c];
            t = acc[c] + t;
            B[3 * i + c] = s->pinned[i] ? (real)0 : h * t;
        }
        if (diag) {
            real dg[3] = {md, md, md};
            for (int k = s->vp_ptr[i]; k < s->vp_ptr[i + 1]; k++) {
                const real* C = Cp + 6 * s->vp_pair[k];
                for (int c = 0; c < 3; c++) dg[c] -= C[c];
            }
            for (int c = 0; c < 3; c++) diag[3 * i + c] = dg[c];
        }
    }
}

/* q = S A d with the on-the-fly diagonal (see above; `mass` is the diagonal
 * mass m - h*c); S zeroes pinned rows.
 * Pinned by: the dense-matrix tests (q equals the dense A times d). */
EXPORT void orc_matvec(const orc_sys* s, const real* Cp, double mass, const real* d, real* q) {
    real m = (real)mass;
    OMP_FOR
    for (int i = 0; i < s->N; i++) {
        real acc[3] = {m * d[3 * i + 0], m * d[3 * i + 1], m * d[3 * i + 2]};
        for (int k = s->vp_ptr[i]; k < s->vp_ptr[i + 1]; k++) {
            int p = s->vp_pair[k];
            int j = s->pair_a[p] == i ? s->pair_b[p] : s->pair_a[p];
            real w[3] = {d[3 * j + 0] - d[3 * i + 0], d[3 * j + 1] - d[3 * i + 1],
                         d[3 * j + 2] - d[3 * i + 2]};
            real t[3];
            symmv(Cp + 6 * p, w, t);
            for (int c = 0; c < 3; c++) acc[c] += t[c];
        }
        for (int c = 0; c < 3; c++) q[3 * i + c] = s->pinned[i] ? (real)0 : acc[c];
    }
}

/* ------------------------------------------------------------------------ */
/* Dot products                                                              */
/* ------------------------------------------------------------------------ */
/* mode 0: sequential, sum += dot(a_i,b_i) (IE:219-225).
 * mode 2: the batched kernels' order (tx = threads per cloth), below.
 * mode 1: the fixed reduction tree of the HIP kernels.  The grid is cut into
 * tx*ty tiles (tiles row-major, threads row-major in a tile, tx*ty a multiple
 * of 64).  Per 64-lane wave: for off = 32,16,..,1 lane l += lane l+off; the
 * wave sums of a tile are added in wave order; the tile partials are then
 * summed lane-strided (lane l takes tiles l, l+64, ... in order) and the 64
 * lane sums reduced by the same 32..1 tree. */
static real wave_tree(real* v) {
    for (int off = 32; off >= 1; off >>= 1)
        for (int l = 0; l < off; l++) v[l] = v[l] + v[l + off];
    return v[0];
}

static real dot_vec(const orc_sys* s, const orc_params* prm, const real* a, const real* b) {
    int N = s->N;
    if (prm->dot_mode == 0) {
        real sum = 0;
        for (int i = 0; i < N; i++) sum += dot3(a + 3 * i, b + 3 * i);
        return sum;
    }
    if (prm->dot_mode == 2) {
        /* batched kernels: T = prm->tx threads per cloth; thread t sums
         * vertices t, t+T, ... in order; 64-lane wave trees; waves in order */
        int T = prm->tx, nw = T / 64;
        real lanes[64];
        real total = 0;
        for (int w = 0; w < nw; w++) {
            for (int l = 0; l < 64; l++) {
                real s = 0;
                for (int v = w * 64 + l; v < N; v += T) s += dot3(a + 3 * v, b + 3 * v);
                lanes[l] = s;
            }
            real ws = wave_tree(lanes);
            total = (w == 0) ? ws : total + ws;
        }
        return total;
    }
    int nx = prm->nx, ny = prm->ny, tx = prm->tx, ty = prm->ty;
    int ntx = (nx + tx - 1) / tx, nty = (ny + ty - 1) / ty;
    int nt = ntx * nty;
    int nthr = tx * ty, nw = nthr / 64;
    real cross[64];
    for (int l = 0; l < 64; l++) cross[l] = 0;
    real* part = (real*)malloc(sizeof(real) * (nt > 0 ? nt : 1));
    OMP_FOR
    for (int t = 0; t < nt; t++) {
        int bx = t % ntx, by = t / ntx;
        real lanes[64];
        real pt = 0;
        for (int w = 0; w < nw; w++) {
            for (int l = 0; l < 64; l++) {
                int tid = w * 64 + l;
                int i = bx * tx + tid % tx, j = by * ty + tid / tx;
                lanes[l] = (i < nx && j < ny) ? dot3(a + 3 * (j * nx + i), b + 3 * (j * nx + i)) : (real)0;
            }
            real ws = wave_tree(lanes);
            pt = (w == 0) ? ws : pt + ws;
        }
        part[t] = pt;
    }
    for (int t = 0; t < nt; t++) cross[t % 64] += part[t];
    free(part);
    return wave_tree(cross);
}

EXPORT double orc_dot(const orc_sys* s, const orc_params* prm, const real* a, const real* b) {
    return (double)dot_vec(s, prm, a, b);
}

/* ------------------------------------------------------------------------ */
/* Filtered conjugate gradient                                               */
/* ------------------------------------------------------------------------ */
/* Baraff-Witkin filtered CG without preconditioner: the recurrence of
 * SolveConjugateGradient (IE:227-249) on S A S dv = S b, with the stopping
 * test of IE:238 (absolute) or BW:292 (relative), or none.  dv holds the
 * start value (warm start) on entry and the solution on exit; pinned entries
 * of dv stay at their start value (zero).  Guards the reference does not
 * have: d.q < 0 or non-finite stops with ORC_BREAKDOWN; d.q == 0 gives
 * alpha = 0 and r.r_old == 0 gives beta = 0 (both only after r underflows,
 * so a fixed iteration count never divides 0 by 0).
 * The recurrence is written once (cg_core) over a matrix-vector callback so
 * that the same code also runs on block-diagonal systems.
 * With a Jacobi inverse pinv (flags ORC_FLAG_JACOBI) it is Baraff and
 * Witkin's modified PCG (1998, sec. 5.3), the intent of BW:308-330 and
 * BW:784-785: z = P^-1 r, delta = r.z (also for the stopping tests),
 * d = z + beta d; the filter keeps pinned entries of r, z and d at zero.
 * (The reference's version takes delta0 = r.(P r) and drops P after the
 * first iteration; that is not restated.)
 * Returns iterations run; *status gets ORC_* bits.
 * Pinned by: a dense FP64 solve of S A S on tiny grids; bitwise equality of
 * cg_core on block-diagonal systems with a transcription of IE:227-249
 * (orc_cg_reference_blockdiag); the residual recurrence r_k = S(b - A x_k);
 * for the preconditioned form, the dense FP64 solve and a dense
 * preconditioned-CG transcription in the tests. */
typedef void (*matvec_fn)(const void* ctx, const real* in, real* out);
typedef real (*dot_fn)(const void* ctx, const real* a, const real* b);

static int cg_core(int N, const uint8_t* pinned, matvec_fn mv, dot_fn dot, const void* ctx,
                   int max_it, real tol, int tol_mode, const real* pinv, const real* b, real* dv,
                   int* status, double* out_delta) {
    real* r = (real*)malloc(sizeof(real) * 3 * (N > 0 ? N : 1));
    real* d = (real*)malloc(sizeof(real) * 3 * (N > 0 ? N : 1));
    real* q = (real*)malloc(sizeof(real) * 3 * (N > 0 ? N : 1));
    real* z = pinv ? (real*)malloc(sizeof(real) * 3 * (N > 0 ? N : 1)) : r; /* P^-1 r */
    mv(ctx, dv, q);
    OMP_FOR
    for (int i = 0; i < 3 * N; i++) {
        r[i] = (pinned && pinned[i / 3]) ? (real)0 : b[i] - q[i];
        if (pinv) z[i] = pinv[i] * r[i];
        d[i] = z[i];
    }
    real tol2 = tol * tol;
    real delta = dot(ctx, r, z);
    real delta0 = delta;
    real thresh = tol_mode == 1 ? tol2 * delta0 : tol2;
    int it = 0, st = 0;
    while (it < max_it) {
        if (tol_mode != 2 && !(delta > thresh)) break;
        mv(ctx, d, q);
        real sigma = dot(ctx, d, q);
        if (sigma < 0 || !ISFINITE_(sigma)) {
            st |= ORC_BREAKDOWN;
            break;
        }
        /* sigma == 0 only once r (hence d) has underflowed: keep the iteration
         * count fixed without producing 0/0 */
        real alpha = sigma == 0 ? (real)0 : delta / sigma;
        OMP_FOR
        for (int i = 0; i < 3 * N; i++) {
            dv[i] = dv[i] + alpha * d[i];
            r[i] = r[i] - alpha * q[i];
            if (pinv) z[i] = pinv[i] * r[i];
        }
        real delta_old = delta;
        delta = dot(ctx, r, z);
        real beta = delta_old == 0 ? (real)0 : delta / delta_old;
        OMP_FOR
        for (int i = 0; i < 3 * N; i++) d[i] = z[i] + beta * d[i];
        it++;
    }
    if (tol_mode != 2 && delta > thresh && !(st & ORC_BREAKDOWN)) st |= ORC_NOT_CONVERGED;
    if (!ISFINITE_(delta)) st |= ORC_NONFINITE;
    if (status) *status = st;
    if (out_delta) *out_delta = (double)delta;
    if (pinv) free(z);
    free(r);
    free(d);
    free(q);
    return it;
}

/* The same solve with one reduction point per iteration: the CG variant of
 * A. T. Chronopoulos and C. W. Gear, "s-step iterative methods for
 * symmetric linear systems" (J. Comput. Appl. Math. 25, 1989), s = 1.  With
 * u = P^-1 r and w = S A u, both dots gamma = r.u and delta = w.u come from
 * the same vectors, and
 *   beta_k  = gamma_k / gamma_{k-1}                      (0 at k = 0)
 *   alpha_k = gamma_k / (delta_k - beta_k gamma_k / alpha_{k-1})
 *   p_k = u_k + beta_k p_{k-1},  s_k = w_k + beta_k s_{k-1}   (s = S A p)
 *   dv += alpha_k p_k,  r_{k+1} = r_k - alpha_k s_k,  u_{k+1} = P^-1 r_{k+1},
 *   w_{k+1} = S A u_{k+1}.
 * In exact arithmetic it is the recurrence of cg_core (the denominator is
 * p_k.A p_k, by the A-conjugacy of p and the orthogonality of the
 * residuals); in floating point the iterates differ by rounding.  The stop
 * tests use gamma (= delta of cg_core).  Guards as in cg_core: a denominator
 * < 0 or non-finite stops with ORC_BREAKDOWN, a zero one gives alpha = 0,
 * gamma_{k-1} == 0 gives beta = 0, and alpha_{k-1} == 0 drops the
 * correction term (both only after the residual underflowed).
 * Pinned by: the FP64 dense solve; agreement with cg_core to 1e-4 (FP32) and
 * 1e-10 (FP64) on converged solves; a Python transcription of the
 * published recurrence on the dense filtered matrix, every iterate. */
static int cg_core_cg1(int N, const uint8_t* pinned, matvec_fn mv, dot_fn dot, const void* ctx,
                       int max_it, real tol, int tol_mode, const real* pinv, const real* b, real* dv,
                       int* status, double* out_delta) {
    const size_t n3 = 3 * (size_t)(N > 0 ? N : 1);
    real* r = (real*)malloc(sizeof(real) * n3);
    real* w = (real*)malloc(sizeof(real) * n3);
    real* p = (real*)malloc(sizeof(real) * n3);
    real* sv = (real*)malloc(sizeof(real) * n3);
    real* u = pinv ? (real*)malloc(sizeof(real) * n3) : r; /* P^-1 r */
    mv(ctx, dv, w);
    OMP_FOR
    for (int i = 0; i < 3 * N; i++) {
        r[i] = (pinned && pinned[i / 3]) ? (real)0 : b[i] - w[i];
        if (pinv) u[i] = pinv[i] * r[i];
    }
    mv(ctx, u, w);
    real tol2 = tol * tol;
    real gamma = dot(ctx, r, u), delta = dot(ctx, w, u);
    real gamma0 = gamma;
    real thresh = tol_mode == 1 ? tol2 * gamma0 : tol2;
    real gamma_old = 0, alpha_old = 0;
    int it = 0, st = 0;
    while (it < max_it) {
        if (tol_mode != 2 && !(gamma > thresh)) break;
        real beta = (it > 0 && gamma_old != 0) ? gamma / gamma_old : (real)0;
        real den = delta;
        if (beta != 0 && alpha_old != 0) {
            real t = beta * gamma;
            t = t / alpha_old;
            den = delta - t;
        }
        if (den < 0 || !ISFINITE_(den)) {
            st |= ORC_BREAKDOWN;
            break;
        }
        real alpha = den == 0 ? (real)0 : gamma / den;
        const int first = it == 0;
        OMP_FOR
        for (int i = 0; i < 3 * N; i++) {
            real ui = pinv ? pinv[i] * r[i] : r[i];
            p[i] = first ? ui : ui + beta * p[i];
            sv[i] = first ? w[i] : w[i] + beta * sv[i];
            dv[i] = dv[i] + alpha * p[i];
            r[i] = r[i] - alpha * sv[i];
            if (pinv) u[i] = pinv[i] * r[i];
        }
        mv(ctx, u, w);
        gamma_old = gamma;
        alpha_old = alpha;
        gamma = dot(ctx, r, u);
        delta = dot(ctx, w, u);
        it++;
    }
    if (tol_mode != 2 && gamma > thresh && !(st & ORC_BREAKDOWN)) st |= ORC_NOT_CONVERGED;
    if (!ISFINITE_(gamma)) st |= ORC_NONFINITE;
    if (status) *status = st;
    if (out_delta) *out_delta = (double)gamma;
    if (pinv) free(u);
    free(r);
    free(w);
    free(p);
    free(sv);
    return it;
}

/* cg_core or, with ORC_FLAG_CG1, cg_core_cg1. */
static int cg_solve(const orc_params* prm, int N, const uint8_t* pinned, matvec_fn mv, dot_fn dot,
                    const void* ctx, const real* pinv, const real* b, real* dv, int* status,
                    double* out_delta) {
    if (prm->flags & ORC_FLAG_CG1)
        return cg_core_cg1(N, pinned, mv, dot, ctx, prm->max_it, (real)prm->tol, prm->tol_mode, pinv, b,
                           dv, status, out_delta);
    return cg_core(N, pinned, mv, dot, ctx, prm->max_it, (real)prm->tol, prm->tol_mode, pinv, b, dv,
                   status, out_delta);
}

typedef struct {
    const orc_sys* s;
    const real* Cp;
    const orc_params* prm;
} pair_ctx;

static void pair_mv(const void* ctx, const real* in, real* out) {
    const pair_ctx* c = (const pair_ctx*)ctx;
    orc_matvec(c->s, c->Cp, (double)diag_mass(c->prm), in, out);
}
static real pair_dot(const void* ctx, const real* a, const real* b) {
    const pair_ctx* c = (const pair_ctx*)ctx;
    return dot_vec(c->s, c->prm, a, b);
}

/* Jacobi preconditioner (flags bit ORC_FLAG_JACOBI): P^-1 = 1/diag(A) per
 * component, diag(A)_i = m - sum of the pair blocks' diagonals in the
 * vertex's pair order; a non-positive or non-finite diagonal entry (possible
 * under strong compression, where A can be indefinite) falls back to 1.
 * Returns NULL when the flag is off (plain CG). */
static real* jacobi_inverse(const orc_sys* s, const real* Cp, const orc_params* prm) {
    if (!(prm->flags & ORC_FLAG_JACOBI)) return NULL;
    real* pinv = (real*)malloc(sizeof(real) * 3 * (s->N > 0 ? s->N : 1));
    real m = diag_mass(prm);
    OMP_FOR
    for (int i = 0; i < s->N; i++) {
        real dg[3] = {m, m, m};
        for (int k = s->vp_ptr[i]; k < s->vp_ptr[i + 1]; k++)
            for (int c = 0; c < 3; c++) dg[c] -= Cp[6 * s->vp_pair[k] + c];
        for (int c = 0; c < 3; c++)
            pinv[3 * i + c] = (dg[c] > 0 && ISFINITE_(dg[c])) ? (real)1 / dg[c] : (real)1;
    }
    return pinv;
}

EXPORT int orc_cg(const orc_sys* s, const real* Cp, const orc_params* prm, const real* b, real* dv,
                  int* status, double* out_delta) {
    pair_ctx c = {s, Cp, prm};
    real* pinv = jacobi_inverse(s, Cp, prm);
    int it = cg_solve(prm, s->N, s->pinned, pair_mv, pair_dot, &c, pinv, b, dv, status, out_delta);
    free(pinv);
    return it;
}

typedef struct {
    int N;
    const real* A;
} bd_ctx;

static void bd_mv(const void* ctx, const real* in, real* out) {
    const bd_ctx* c = (const bd_ctx*)ctx;
    for (int i = 0; i < c->N; i++) {
        const real* Ai = c->A + 9 * i;
        const real* v = in + 3 * i;
        for (int k = 0; k < 3; k++) out[3 * i + k] = Ai[k] * v[0] + Ai[3 + k] * v[1] + Ai[6 + k] * v[2];
    }
}
static real bd_dot(const void* ctx, const real* a, const real* b) {
    const bd_ctx* c = (const bd_ctx*)ctx;
    real sum = 0;
    for (int i = 0; i < c->N; i++) sum += dot3(a + 3 * i, b + 3 * i);
    return sum;
}

/* cg_core on a block-diagonal system: A is N full 3x3 blocks, column-major
 * like glm::mat3 (A[9*i + 3*col + row]); no pins, sequential dots, absolute
 * stopping test. */
EXPORT int orc_cg_blockdiag(int N, const real* A, const real* b, real* x, int i_max, double eps) {
    bd_ctx c = {N, A};
    return cg_core(N, NULL, bd_mv, bd_dot, &c, i_max, (real)eps, 0, NULL, b, x, NULL, NULL);
}

/* Line-by-line transcription of SolveConjugateGradient (IE:227-249) with the
 * LargeVector<mat3>*LargeVector<vec3> product (IE:165-171) and dot
 * (IE:219-225); same A layout as orc_cg_blockdiag. */
EXPORT int orc_cg_reference_blockdiag(int N, const real* A, const real* b, real* x, int i_max,
                                      double eps) {
    real* r = (real*)malloc(sizeof(real) * 3 * N);
    real* d = (real*)malloc(sizeof(real) * 3 * N);
    real* q = (real*)malloc(sizeof(real) * 3 * N);
#define MV(Ai, v, o)                                                     \
    do {                                                                 \
        (o)[0] = (Ai)[0] * (v)[0] + (Ai)[3] * (v)[1] + (Ai)[6] * (v)[2]; \
        (o)[1] = (Ai)[1] * (v)[0] + (Ai)[4] * (v)[1] + (Ai)[7] * (v)[2]; \
        (o)[2] = (Ai)[2] * (v)[0] + (Ai)[5] * (v)[1] + (Ai)[8] * (v)[2]; \
    } while (0)
    for (int i = 0; i < N; i++) {
        real ax[3];
        MV(A + 9 * i, x + 3 * i, ax);
        for (int c = 0; c < 3; c++) r[3 * i + c] = b[3 * i + c] - ax[c];
    }
    memcpy(d, r, sizeof(real) * 3 * N);
    real e = (real)eps;
    real eps2 = e * e;
    real delta_new = 0;
    for (int i = 0; i < N; i++) delta_new += dot3(r + 3 * i, r + 3 * i);
    int it = 0;
    while (it < i_max && delta_new > eps2) {
        for (int i = 0; i < N; i++) MV(A + 9 * i, d + 3 * i, q + 3 * i);
        real dq = 0;
        for (int i = 0; i < N; i++) dq += dot3(d + 3 * i, q + 3 * i);
        real alpha = delta_new / dq;
        for (int i = 0; i < 3 * N; i++) x[i] = x[i] + alpha * d[i];
        for (int i = 0; i < 3 * N; i++) r[i] = r[i] - alpha * q[i];
        real delta_old = delta_new;
        delta_new = 0;
        for (int i = 0; i < N; i++) delta_new += dot3(r + 3 * i, r + 3 * i);
        real beta = delta_new / delta_old;
        for (int i = 0; i < 3 * N; i++) d[i] = r[i] + beta * d[i];
        it++;
    }
#undef MV
    free(r);
    free(d);
    free(q);
    return it;
}

/* ------------------------------------------------------------------------ */
/* The reference's own step, bug for bug                                     */
/* ------------------------------------------------------------------------ */
/* ComputeForces + IntegrateImplicit of IE as written (IE:645-700), for the
 * record and the CPU timing only; nothing in the library follows it:
 *   F   = gravity on free vertices plus the spring forces, scattered in list
 *         order with pinned ends skipped (IE:645-682);
 *   A_i = M_i - h^2 K_i for i < N, K_i the Jacobian of SPRING i (IE:671):
 *         LargeVector's operator- runs over M's N entries (IE:187-193,
 *         IE:688), so spring i's block lands on vertex i;
 *   b   = M V + h F (IE:689), with GLM's mat3*vec3 and vec3*float;
 *   SolveConjugateGradient(A, V_new, b) on that block-diagonal system, warm
 *         from the previous V_new (IE:227-249), then x += h V_new, clamp
 *         y < 0 to 0, V = V_new (IE:693-699).
 * The reference adds `gravity` as a force (IE:653); here g is an
 * acceleration and the force is mass*g, which is the same float when mass
 * is a power of two (the demo's 0.5).  Pins come from the pinned mask (the
 * reference tests i != 0 && i != numX, IE:652, IE:677-680).  Needs S >= N.
 * Returns the CG iterations, or -1 when S < N.
 * Pinned by: tests/test_oracle_literal.py -- free fall with inert springs
 * (closed form), and a three-spring system solved densely in float64 with
 * each vertex's block built from its spring's Jacobian in the spectral form
 * of IE:671 (-Ks along the spring, -Ks(1-L0/l) across). */
EXPORT int orc_ie_literal_step(const orc_sys* s, real* X, real* V, real* Vnew, double mass,
                               const double* g, double h, int i_max, double eps) {
    int N = s->N;
    if (s->S < N) return -1;
    real m = (real)mass, dt = (real)h, dt2 = dt * dt;
    real* F = (real*)malloc(sizeof(real) * 3 * (N > 0 ? N : 1));
    real* A = (real*)malloc(sizeof(real) * 9 * (N > 0 ? N : 1));
    real* b = (real*)malloc(sizeof(real) * 3 * (N > 0 ? N : 1));
    orc_forces_scatter(s, X, V, mass, g, 0.0, F);
    for (int i = 0; i < N; i++) {
        const orc_spring* sp = &s->springs[i];
        real f[3], K[6];
        spring_terms(X + 3 * sp->p1, X + 3 * sp->p2, V + 3 * sp->p1, V + 3 * sp->p2,
                     (real)sp->rest_length, s->ks[i], (real)sp->Kd, f, K, NULL);
        /* column-major 3x3 (glm::mat3), entry (row r, column c) at c*3+r */
        const real Kf[9] = {K[0], K[3], K[4], K[3], K[1], K[5], K[4], K[5], K[2]};
        real Mi[9];
        for (int e = 0; e < 9; e++) {
            Mi[e] = (e == 0 || e == 4 || e == 8) ? m : (real)0;
            A[9 * i + e] = Mi[e] - Kf[e] * dt2;
        }
        for (int r = 0; r < 3; r++) {
            real mv = Mi[r] * V[3 * i + 0] + Mi[3 + r] * V[3 * i + 1] + Mi[6 + r] * V[3 * i + 2];
            b[3 * i + r] = mv + F[3 * i + r] * dt;
        }
    }
    int it = orc_cg_reference_blockdiag(N, A, b, Vnew, i_max, eps);
    for (int i = 0; i < N; i++) {
        for (int c = 0; c < 3; c++) X[3 * i + c] = X[3 * i + c] + Vnew[3 * i + c] * dt;
        if (X[3 * i + 1] < 0) X[3 * i + 1] = 0;
        for (int c = 0; c < 3; c++) V[3 * i + c] = Vnew[3 * i + c];
    }
    free(F);
    free(A);
    free(b);
    return it;
}

/* ------------------------------------------------------------------------ */
/* One implicit step and rollouts                                            */
/* ------------------------------------------------------------------------ */
/* The implicit step: assemble, filtered CG for dv (warm start from DV when
 * prm->warm_start, else from zero), then v += dv, x += h*v (IE:693-699 with
 * the block-CSR dv form of BW:781-792, velocity updated once by dv).
 * Optional ground clamp (flags bit0, IE:695-697).  DV receives the solution.
 * Returns CG iterations; *status gets ORC_* bits. */
EXPORT int orc_step(const orc_sys* s, real* X, real* V, real* DV, const orc_params* prm,
                    int* status) {
    int N = s->N;
    real* B = (real*)malloc(sizeof(real) * 3 * N);
    real* Cp = (real*)malloc(sizeof(real) * 6 * (s->P > 0 ? s->P : 1));
    orc_assemble(s, X, V, prm, B, Cp, NULL);
    if (!prm->warm_start)
        for (int i = 0; i < 3 * N; i++) DV[i] = 0;
    int st = 0;
    int it = orc_cg(s, Cp, prm, B, DV, &st, NULL);
    real h = (real)prm->h;
    int bad = 0;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(g_threads) if (g_threads > 1) reduction(| : bad)
#endif
    for (int i = 0; i < N; i++) {
        for (int c = 0; c < 3; c++) {
            V[3 * i + c] = V[3 * i + c] + DV[3 * i + c];
            X[3 * i + c] = X[3 * i + c] + h * V[3 * i + c];
        }
        if ((prm->flags & 1) && X[3 * i + 1] < 0) X[3 * i + 1] = 0;
        for (int c = 0; c < 3; c++)
            if (!ISFINITE_(X[3 * i + c]) || !ISFINITE_(V[3 * i + c])) bad = 1;
    }
    if (bad) st |= ORC_NONFINITE;
    free(B);
    free(Cp);
    if (status) *status = st;
    return it;
}

/* nsteps of orc_step; Xhist (may be NULL) receives X after every step
 * (nsteps*3N values); iters (may be NULL) the CG iterations per step.
 * Returns the OR of all step status bits. */
EXPORT int orc_rollout(const orc_sys* s, real* X, real* V, real* DV, const orc_params* prm,
                       int nsteps, real* Xhist, int* iters) {
    int all = 0;
    for (int k = 0; k < nsteps; k++) {
        int st = 0;
        int it = orc_step(s, X, V, DV, prm, &st);
        all |= st;
        if (iters) iters[k] = it;
        if (Xhist) memcpy(Xhist + (size_t)k * 3 * s->N, X, sizeof(real) * 3 * s->N);
    }
    return all;
}

/* ------------------------------------------------------------------------ */
/* Adjoint of one implicit step                                              */
/* ------------------------------------------------------------------------ */
/* Forward step (orc_step): A(x) dv = b(x,v) on the free vertices, with
 *   A = M - h D(x) - h^2 K(x),  b = h f(x,v) + h^2 K(x) v,
 *   v' = v + dv,  x' = x + h v'.
 * Given gx' = dL/dx', gv' = dL/dv':
 *   gvt = gv' + h gx'
 *   A mu = S gvt           (filtered CG on the same symmetric matrix)
 *   J(x,v,ks) = h mu.f + h^2 mu.K v' + h mu.D dv   (mu, v', dv held fixed)
 *   gx = gx' + dJ/dx,  gv = gvt + dJ/dv,  gks += dJ/dks.
 * Per spring (e = x1-x2, u = mu1-mu2, w = v'1-v'2, w' = dv1-dv2, l = |e|):
 *   mu.f     -> phi (u.e)/l,  phi = -ks (l-L0) + Kd (dvel.e)/l
 *   mu.K v'  -> ks [(L0/l - 1)(u.w) - L0 (u.e)(w.e)/l^3]
 *   mu.D dv  -> Kd (u.e)(w'.e)/l^2
 * plus the explicit v in K v (dJ/dv1 += h^2 K u).  Each vertex gathers its
 * springs in the forward's order; a spring's ks gradient is added by its p1.
 * The implicit-function form assumes the forward solve converged.
 * Pinned by: tests/test_oracle_adjoint.py -- central finite differences of
 * FP64 rollouts in ks, v0 and x0 directions and the dot-product test. */
static void spring_adjoint(const real* xa, const real* xb, const real* va, const real* vb,
                           const real* ua, const real* ub, const real* wa, const real* wb,
                           const real* pa, const real* pb, real L0, real ks, real Kd, real h,
                           real hh, real* ge, real* gvv, real* gk) {
    real e[3], dvel[3], u[3], w[3], wp[3];
    for (int c = 0; c < 3; c++) {
        e[c] = xa[c] - xb[c];
        dvel[c] = va[c] - vb[c];
        u[c] = ua[c] - ub[c];
        w[c] = wa[c] - wb[c];
        wp[c] = pa[c] - pb[c];
    }
    real l = len3(e);
    real il = (real)1 / l;
    real il2 = il * il, il3 = il2 * il, il4 = il2 * il2, il5 = il4 * il;
    real ue = dot3(u, e), we = dot3(w, e), wpe = dot3(wp, e), uw = dot3(u, w), dve = dot3(dvel, e);
    real phi = -ks * (l - L0) + Kd * (dve * il);
    for (int c = 0; c < 3; c++) {
        real dphi = -ks * e[c] * il + Kd * (dvel[c] * il - dve * e[c] * il3);
        real A = h * ((ue * il) * dphi + phi * (u[c] * il - ue * e[c] * il3));
        real B = hh * ks * L0 * (-uw * e[c] * il3 - (u[c] * we + w[c] * ue) * il3 +
                                 (real)3 * ue * we * e[c] * il5);
        real Cc = h * Kd * ((u[c] * wpe + wp[c] * ue) * il2 - (real)2 * ue * wpe * e[c] * il4);
        ge[c] = A + B + Cc;
    }
    real f[3], K[6], P[6];
    spring_terms(xa, xb, va, vb, L0, ks, Kd, f, K, P);
    real ku[3];
    symmv(K, u, ku);
    for (int c = 0; c < 3; c++) gvv[c] = h * Kd * ue * e[c] * il2 + hh * ku[c];
    *gk = h * (ue * il) * (-(l - L0)) + hh * ((L0 * il - (real)1) * uw - L0 * ue * we * il3);
}

/* One step's adjoint at the pre-step state (X, V) whose solve gave DV.
 * gks (S values) is accumulated; gx_in/gv_in are overwritten; mu_out (may be
 * NULL) receives the adjoint solve.  Returns the adjoint CG iterations. */
EXPORT int orc_step_adjoint(const orc_sys* s, const real* X, const real* V, const real* DV,
                            const orc_params* prm, const real* gx_out, const real* gv_out,
                            real* gx_in, real* gv_in, real* gks, real* mu_out, int* status) {
    int N = s->N;
    real h = (real)prm->h, hh = h * h;
    real* gvt = (real*)malloc(sizeof(real) * 3 * N);
    real* mu = (real*)calloc(3 * (size_t)N, sizeof(real));
    real* B = (real*)malloc(sizeof(real) * 3 * N);
    real* Cp = (real*)malloc(sizeof(real) * 6 * (s->P > 0 ? s->P : 1));
    real* Vp = (real*)malloc(sizeof(real) * 3 * N);
    for (int i = 0; i < 3 * N; i++) {
        gvt[i] = gv_out[i] + h * gx_out[i];
        Vp[i] = V[i] + DV[i];
    }
    orc_assemble(s, X, V, prm, B, Cp, NULL);
    pair_ctx pc = {s, Cp, prm};
    real* pinv = jacobi_inverse(s, Cp, prm);
    int it = cg_solve(prm, N, s->pinned, pair_mv, pair_dot, &pc, pinv, gvt, mu, status, NULL);
    free(pinv);
    real dmp = (real)prm->damping;
    for (int i = 0; i < N; i++) {
        real gx[3] = {gx_out[3 * i], gx_out[3 * i + 1], gx_out[3 * i + 2]};
        real gv[3] = {gvt[3 * i], gvt[3 * i + 1], gvt[3 * i + 2]};
        /* h mu.f with f_i += c v_i on free vertices: dJ/dv_i += h c mu_i */
        if (!s->pinned[i] && prm->damping != 0)
            for (int c = 0; c < 3; c++) gv[c] = gv[c] + h * (dmp * mu[3 * i + c]);
        for (int e = s->inc_ptr[i]; e < s->inc_ptr[i + 1]; e++) {
            int sid = s->inc_s[e];
            const orc_spring* sp = &s->springs[sid];
            int a = sp->p1, b = sp->p2;
            real ge[3], gvv[3], gk;
            spring_adjoint(X + 3 * a, X + 3 * b, V + 3 * a, V + 3 * b, mu + 3 * a, mu + 3 * b,
                           Vp + 3 * a, Vp + 3 * b, DV + 3 * a, DV + 3 * b, (real)sp->rest_length,
                           s->ks[sid], (real)sp->Kd, h, hh, ge, gvv, &gk);
            if (a == i) {
                for (int c = 0; c < 3; c++) {
                    gx[c] += ge[c];
                    gv[c] += gvv[c];
                }
                gks[sid] += gk;
            } else {
                for (int c = 0; c < 3; c++) {
                    gx[c] -= ge[c];
                    gv[c] -= gvv[c];
                }
            }
        }
        for (int c = 0; c < 3; c++) {
            gx_in[3 * i + c] = gx[c];
            gv_in[3 * i + c] = gv[c];
        }
    }
    if (mu_out) memcpy(mu_out, mu, sizeof(real) * 3 * N);
    free(gvt);
    free(mu);
    free(B);
    free(Cp);
    free(Vp);
    return it;
}

/* ------------------------------------------------------------------------ */
/* Post-step passes of StepPhysics                                           */
/* ------------------------------------------------------------------------ */

/* EllipsoidCollision (IE:727-751; VL:509-533; PBD:846-870).  M, Minv:
 * column-major 4x4 (glm::mat4, M[4*col + row]).  A vertex inside the unit
 * sphere of the ellipsoid's object space is pushed to its surface scaled by
 * radius; then mode 0 zeroes its velocity W (IE, PBD) and mode 1 sets its
 * previous position W = X (VL:530).
 * Pinned by: tests/test_oracle_post.py (outside points unchanged, points on
 * the axis of a sphere pushed to the closed-form surface point, GLM order). */
EXPORT void orc_ellipsoid_points(int N, real* X, real* W, const float* M, const float* Minv,
                                 const float* center, float radius, int mode) {
    for (int i = 0; i < N; i++) {
        real* x = X + 3 * i;
        real d0[3];
        for (int r = 0; r < 3; r++) {
            real t = (real)Minv[0 * 4 + r] * x[0] + (real)Minv[1 * 4 + r] * x[1] +
                     (real)Minv[2 * 4 + r] * x[2] + (real)Minv[3 * 4 + r] * (real)1;
            d0[r] = t - (real)center[r];
        }
        real dist = len3(d0);
        if (dist < (real)1) {
            for (int c = 0; c < 3; c++) d0[c] = (((real)radius - dist) * d0[c]) / dist;
            real delta[3];
            for (int a = 0; a < 3; a++) {
                real t[3] = {(real)M[0 * 4 + a], (real)M[1 * 4 + a], (real)M[2 * 4 + a]};
                real tt = dot3(t, t);
                for (int c = 0; c < 3; c++) t[c] = t[c] / tt;
                delta[a] = dot3(d0, t);
            }
            for (int c = 0; c < 3; c++) x[c] += delta[c];
            for (int c = 0; c < 3; c++) W[3 * i + c] = mode == 1 ? x[c] : (real)0;
        }
    }
}

EXPORT void orc_ellipsoid(const orc_sys* s, real* X, real* W, const float* M, const float* Minv,
                          const float* center, float radius, int mode) {
    orc_ellipsoid_points(s->N, X, W, M, Minv, center, radius, mode);
}

/* ApplyProvotDynamicInverse (IE:702-726; BW:800-823 with slack 1.01): for
 * every spring longer than slack*rest, in list order, half the excess along
 * the spring is moved into the velocities V (never into X, so all springs
 * see the same positions); a pinned p1 sends it all to p2, else a pinned p2
 * skips p2's share.  Pins generalise the reference's {0, numX}.
 * Pinned by: tests/test_oracle_post.py (excess-length closed form on a
 * stretched pair, no change below the threshold, momentum of an unpinned
 * system). */
EXPORT void orc_provot(const orc_sys* s, const real* X, real* V, double slack) {
    for (int i = 0; i < s->S; i++) {
        const orc_spring* sp = &s->springs[i];
        real dp[3] = {X[3 * sp->p1] - X[3 * sp->p2], X[3 * sp->p1 + 1] - X[3 * sp->p2 + 1],
                      X[3 * sp->p1 + 2] - X[3 * sp->p2 + 2]};
        real dist = len3(dp);
        real thr = (real)sp->rest_length * (real)slack;
        if (dist > thr) {
            dist -= thr;
            dist /= (real)2;
            real inv = (real)1 / SQRT_(dp[0] * dp[0] + dp[1] * dp[1] + dp[2] * dp[2]);
            for (int c = 0; c < 3; c++) dp[c] = dp[c] * inv;
            for (int c = 0; c < 3; c++) dp[c] *= dist;
            if (s->pinned[sp->p1]) {
                for (int c = 0; c < 3; c++) V[3 * sp->p2 + c] += dp[c];
            } else if (s->pinned[sp->p2]) {
                for (int c = 0; c < 3; c++) V[3 * sp->p1 + c] -= dp[c];
            } else {
                for (int c = 0; c < 3; c++) V[3 * sp->p1 + c] -= dp[c];
                for (int c = 0; c < 3; c++) V[3 * sp->p2 + c] += dp[c];
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Position Verlet (VL)                                                      */
/* ------------------------------------------------------------------------ */

/* ComputeForces(dt) (VL:448-484): velocities (x - x_last)/dt
 * (GetVerletVelocity, VL:445-447); gravity*mass on free vertices, damping*v on
 * all (VL:455-458); spring forces scattered in list order, pinned endpoints
 * skipped (VL:462-483).
 * Pinned by: equality with orc_forces_scatter at v = (x - x_last)/dt plus
 * the damping term (tests/test_oracle_post.py). */
EXPORT void orc_verlet_forces(const orc_sys* s, const real* X, const real* XL, double dt,
                              double mass, const double* g, double damping, real* F) {
    real h = (real)dt, m = (real)mass, kd = (real)damping;
    real gv[3] = {(real)g[0], (real)g[1], (real)g[2]};
    real* V = (real*)malloc(sizeof(real) * 3 * (s->N > 0 ? s->N : 1));
    for (int i = 0; i < 3 * s->N; i++) V[i] = (X[i] - XL[i]) / h;
    for (int i = 0; i < s->N; i++) {
        for (int c = 0; c < 3; c++) F[3 * i + c] = 0;
        if (!s->pinned[i])
            for (int c = 0; c < 3; c++) F[3 * i + c] += gv[c] * m;
        for (int c = 0; c < 3; c++) F[3 * i + c] += kd * V[3 * i + c];
    }
    for (int i = 0; i < s->S; i++) {
        const orc_spring* sp = &s->springs[i];
        real f[3];
        spring_terms(X + 3 * sp->p1, X + 3 * sp->p2, V + 3 * sp->p1, V + 3 * sp->p2,
                     (real)sp->rest_length, s->ks[i], (real)sp->Kd, f, NULL, NULL);
        if (!s->pinned[sp->p1])
            for (int c = 0; c < 3; c++) F[3 * sp->p1 + c] += f[c];
        if (!s->pinned[sp->p2])
            for (int c = 0; c < 3; c++) F[3 * sp->p2 + c] -= f[c];
    }
    free(V);
}

/* IntegrateVerlet(dt) (VL:428-444): x' = x + (x - x_last) + (dt^2/mass) F,
 * x_last' = x, then (flags bit0) y < 0 clamped to 0.
 * Pinned by: free fall (no springs) against the closed form of the
 * recurrence x_{n+1} = 2x_n - x_{n-1} + dt^2 g. */
EXPORT void orc_verlet_integrate(const orc_sys* s, real* X, real* XL, const real* F, double dt,
                                 double mass, int flags) {
    real h = (real)dt;
    real c2 = (h * h) / (real)mass;
    for (int i = 0; i < s->N; i++) {
        for (int c = 0; c < 3; c++) {
            real buffer = X[3 * i + c];
            X[3 * i + c] = X[3 * i + c] + (X[3 * i + c] - XL[3 * i + c]) + c2 * F[3 * i + c];
            XL[3 * i + c] = buffer;
        }
        if ((flags & 1) && X[3 * i + 1] < 0) X[3 * i + 1] = 0;
    }
}

/* StepPhysics of the Verlet demo (VL:557-562): forces, integrate, and the
 * ellipsoid pass when M is given (mode 1). */
EXPORT void orc_verlet_step(const orc_sys* s, real* X, real* XL, double dt, double mass,
                            const double* g, double damping, int flags, const float* M,
                            const float* Minv, const float* center, float radius) {
    real* F = (real*)malloc(sizeof(real) * 3 * (s->N > 0 ? s->N : 1));
    orc_verlet_forces(s, X, XL, dt, mass, g, damping, F);
    orc_verlet_integrate(s, X, XL, F, dt, mass, flags);
    if (M) orc_ellipsoid(s, X, XL, M, Minv, center, radius, 1);
    free(F);
}

/* ------------------------------------------------------------------------ */
/* Grid Verlet: the GPU demos' kernel                                        */
/* ------------------------------------------------------------------------ */
/* CUK = OpenCloth_Verlet_CUDA_GLSL_OPENCL_CPU/OpenCloth_Verlet_CUDA/verlet_kernel.cu.
 * The kernel includes cutil_math.h, a helper header of the CUDA SDK of its
 * time (GPU Computing SDK 3.x) that is not part of the reference.  It
 * defines the float2/float3 operators the kernel uses: componentwise + - *,
 * dot as a left-to-right sum of products, length = sqrtf(dot(v, v)) and
 * normalize = v * rsqrtf(dot(v, v)).  rsqrtf is an approximate reciprocal
 * square root and nvcc contracts a*b+c into FMAs by default, so the CUDA
 * kernel's bits are not reproducible; this restatement (and the HIP kernel)
 * uses correctly rounded operations in the kernel's expression order, with
 * normalize = v * (1/sqrt(dot(v, v))) as in spring_terms and scalar
 * division componentwise.
 *
 * Stencil position k = 0..11 of getNextNeighbor (CUK:14-64), (column, row)
 * offsets; k < 4 structural, k < 8 shear, else bend. */
static const int vg_dx[12] = {1, 0, -1, 0, 1, -1, -1, 1, 2, 0, -2, 0};
static const int vg_dy[12] = {0, -1, 0, 1, -1, -1, 1, 1, 0, -2, 0, 2};

/* nsteps of the kernel's step (CUK:119-195) on a width x height grid of
 * float4 positions X4 and previous positions XL4 (AoS, w carried):
 *   v = (x - x_last)/dt,  F = gravity*mass + v*damping            CUK:121-124
 *   F += spring force of each on-grid stencil neighbour, rest
 *        length |(dx*sx, dy*sy)|, the class's ks and kd          CUK:130-176
 *   a = F/mass, 0 on pinned vertices (mass zeroed)               CUK:126-127, 178-180
 *   y < 0 of the current position clamped to 0                   CUK:183-184
 *   x' = (2x - x_last) + (a dt) dt,  x_last' = x (clamped)       CUK:187-189
 * prm = {mass, gx, gy, gz, damping, dt, sx, sy, ks[3], kd[3]}.
 * Pinned by: tests/test_oracle_vgrid.py -- free fall of an unpinned grid at
 * rest spacing against the closed form x_n = x_0 + g dt^2 n(n+1)/2 (springs
 * at rest length exert nothing, so a wrong rest length, class or offset
 * fails), an independent scatter over the enumerated springs (FP64), zero
 * total force with no gravity, damping or pins, and the clamp. */
EXPORT void orc_vgrid_step(int width, int height, real* X4, real* XL4, const uint8_t* pin,
                           const double* prm, int nsteps) {
    const int N = width * height;
    const real m = (real)prm[0], h = (real)prm[5], damp = (real)prm[4];
    const real g[3] = {(real)prm[1], (real)prm[2], (real)prm[3]};
    const real sx = (real)prm[6], sy = (real)prm[7];
    real rest[12], ks[3], kd[3];
    for (int k = 0; k < 12; k++) {
        real a = (real)vg_dx[k] * sx, b = (real)vg_dy[k] * sy;
        rest[k] = SQRT_(a * a + b * b);
    }
    for (int c = 0; c < 3; c++) {
        ks[c] = (real)prm[8 + c];
        kd[c] = (real)prm[11 + c];
    }
    real* V = (real*)malloc(sizeof(real) * 3 * (N > 0 ? N : 1));
    real* Xn = (real*)malloc(sizeof(real) * 4 * (N > 0 ? N : 1));
    real* XLn = (real*)malloc(sizeof(real) * 4 * (N > 0 ? N : 1));
    for (int step = 0; step < nsteps; step++) {
        OMP_FOR
        for (int i = 0; i < N; i++)
            for (int c = 0; c < 3; c++) V[3 * i + c] = (X4[4 * i + c] - XL4[4 * i + c]) / h;
        OMP_FOR
        for (int i = 0; i < N; i++) {
            const int ix = i % width, iy = i / width;
            real x[3] = {X4[4 * i], X4[4 * i + 1], X4[4 * i + 2]};
            real F[3];
            for (int c = 0; c < 3; c++) F[c] = g[c] * m + V[3 * i + c] * damp;
            for (int k = 0; k < 12; k++) {
                const int jx = ix + vg_dx[k], jy = iy + vg_dy[k];
                if (jy < 0 || jy > height - 1 || jx < 0 || jx > width - 1) continue;
                const int j = jy * width + jx;
                real xj[3] = {X4[4 * j], X4[4 * j + 1], X4[4 * j + 2]};
                real f[3];
                spring_terms(x, xj, V + 3 * i, V + 3 * j, rest[k], ks[k >> 2], kd[k >> 2], f, NULL,
                             NULL);
                for (int c = 0; c < 3; c++) F[c] += f[c];
            }
            real a[3] = {0, 0, 0};
            if (!pin[i])
                for (int c = 0; c < 3; c++) a[c] = F[c] / m;
            if (x[1] < 0) x[1] = 0;
            for (int c = 0; c < 3; c++) {
                Xn[4 * i + c] = (x[c] * (real)2 - XL4[4 * i + c]) + (a[c] * h) * h;
                XLn[4 * i + c] = x[c];
            }
            Xn[4 * i + 3] = X4[4 * i + 3];
            XLn[4 * i + 3] = XL4[4 * i + 3];
        }
        memcpy(X4, Xn, sizeof(real) * 4 * N);
        memcpy(XL4, XLn, sizeof(real) * 4 * N);
    }
    free(V);
    free(Xn);
    free(XLn);
}
