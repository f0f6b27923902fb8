// C ABI of the position-based-dynamics cloth (oc_pbd_*): the PBD demo's step
// (OpenCloth_PositionBasedDynamics/.../main.cpp) over caller-given distance
// and bending constraint lists.  Constraints are coloured on the host
// (greedy, in list order: a constraint takes the lowest colour not yet used
// at any of its particles) and stored colour by colour; one launch projects
// one colour.  With OC_FLAG_LIST_ORDER they are stored level by level
// instead (a constraint's level is one more than the highest level among the
// earlier constraints sharing a particle with it) and one single-workgroup
// launch projects all levels of a step in order: the reference's sequential
// results.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <new>
#include <string>
#include <vector>

#include "capi_common.hpp"
#include "oc_internal.hpp"

namespace {

using oc::align_up;
using oc::set_error;

float clamp01(float k) { return k > 1.0f ? 1.0f : (k < 0.0f ? 0.0f : k); }

// k' = 1 - (1 - k)^(1/ns), capped at 1 (PBD:153-156, PBD:174-176).
float k_prime(float k, int ns) {
    const float kp = 1.0f - powf(1.0f - k, 1.0f / (float)ns);
    return kp > 1.0f ? 1.0f : kp;
}

// Constraints grouped by group id col[c] (ncol groups): order gets the ids
// group by group, ptr the group offsets.  Inside a group the constraints are
// ordered by first particle: they touch disjoint particles, so the order
// does not change the result, and neighbouring lanes then read neighbouring
// particles (the reference lists column constraints column by column, which
// in list order would put a whole grid row between neighbouring lanes).
void bucket(int n, int arity, const std::vector<int32_t>& p, const std::vector<int>& col, int ncol,
            std::vector<int32_t>& order, std::vector<int32_t>& ptr) {
    ptr.assign((size_t)ncol + 1, 0);
    for (int c = 0; c < n; c++) ptr[col[c] + 1]++;
    for (int k = 0; k < ncol; k++) ptr[k + 1] += ptr[k];
    std::vector<int32_t> fill(ptr.begin(), ptr.end());
    order.assign((size_t)n, 0);
    for (int c = 0; c < n; c++) order[fill[col[c]]++] = c;
    for (int k = 0; k < ncol; k++)
        std::stable_sort(order.begin() + ptr[k], order.begin() + ptr[k + 1], [&](int32_t x, int32_t y) {
            return p[(size_t)arity * x] < p[(size_t)arity * y];
        });
}

// Greedy colouring in list order (see bucket for the order inside a colour).
// Returns the colour count, -1 past 64.
int colour(int N, int n, int arity, const std::vector<int32_t>& p, std::vector<int32_t>& order,
           std::vector<int32_t>& ptr) {
    std::vector<uint64_t> used((size_t)N, 0);
    std::vector<int> col((size_t)n);
    int ncol = 0;
    for (int c = 0; c < n; c++) {
        uint64_t m = 0;
        for (int a = 0; a < arity; a++) m |= used[p[(size_t)arity * c + a]];
        if (m == ~uint64_t(0)) return -1;
        const int k = __builtin_ctzll(~m);
        col[c] = k;
        for (int a = 0; a < arity; a++) used[p[(size_t)arity * c + a]] |= uint64_t(1) << k;
        ncol = std::max(ncol, k + 1);
    }
    bucket(n, arity, p, col, ncol, order, ptr);
    return ncol;
}

// List-order levels: constraint c's level is one more than the highest level
// among the earlier constraints sharing a particle with it, so every earlier
// constraint it depends on sits in a lower level and the constraints of one
// level share no particle.  Returns the level count.
int levels(int N, int n, int arity, const std::vector<int32_t>& p, std::vector<int32_t>& order,
           std::vector<int32_t>& ptr) {
    std::vector<int> last((size_t)N, -1);
    std::vector<int> lev((size_t)n);
    int nlev = 0;
    for (int c = 0; c < n; c++) {
        int l = 0;
        for (int a = 0; a < arity; a++) l = std::max(l, last[p[(size_t)arity * c + a]] + 1);
        lev[c] = l;
        for (int a = 0; a < arity; a++) last[p[(size_t)arity * c + a]] = l;
        nlev = std::max(nlev, l + 1);
    }
    bucket(n, arity, p, lev, nlev, order, ptr);
    return nlev;
}

}  // namespace

struct oc_pbd_s {
    int N = 0, nd = 0, nb = 0;
    oc_pbd_params prm{};
    hipStream_t stream = nullptr;
    void* arena = nullptr;
    oc::DevPbd dev{};
    std::vector<int32_t> d_ptr, b_ptr;  // colour offsets
};

namespace {

oc::PbdScalars pscalars(const oc_pbd_s* p) {
    oc::PbdScalars s;
    s.mass = p->prm.mass;
    s.gx = p->prm.gravity[0];
    s.gy = p->prm.gravity[1];
    s.gz = p->prm.gravity[2];
    s.dt = p->prm.dt;
    // InitGL caps global_dampening at 1 and clamps kDamp to [0,1] (PBD:391-396)
    s.gd = p->prm.global_damping > 1.0f ? 1.0f : p->prm.global_damping;
    s.kdamp = clamp01(p->prm.k_damp);
    s.global_k = s.gd * 0.01f;  // PBD:712
    s.inv_dt = 1.0f / p->prm.dt;
    s.flags = p->prm.flags;
    return s;
}

int upload_aos(oc_pbd_s* p, float* dst, const float* aos) {
    const size_t N = p->N;
    std::vector<float> t(3 * N);
    for (size_t i = 0; i < N; i++)
        for (int k = 0; k < 3; k++) t[k * N + i] = aos ? aos[3 * i + k] : 0.0f;
    OC_HIP(hipMemcpyAsync(dst, t.data(), t.size() * 4, hipMemcpyHostToDevice, p->stream));
    OC_HIP(hipStreamSynchronize(p->stream));
    return OC_OK;
}

int download_aos(oc_pbd_s* p, const float* src, float* aos) {
    const size_t N = p->N;
    std::vector<float> t(3 * N);
    OC_HIP(hipMemcpyAsync(t.data(), src, t.size() * 4, hipMemcpyDeviceToHost, p->stream));
    OC_HIP(hipStreamSynchronize(p->stream));
    for (size_t i = 0; i < N; i++)
        for (int k = 0; k < 3; k++) aos[3 * i + k] = t[k * N + i];
    return OC_OK;
}

int check_pbd_params(const oc_pbd_params* p) {
    if (!p) return set_error(OC_ERR_INVALID_ARGUMENT, "params is NULL");
    if (!(p->mass > 0.0f) || !std::isfinite(p->mass))
        return set_error(OC_ERR_INVALID_ARGUMENT, "mass must be positive");
    if (!(p->dt > 0.0f) || !std::isfinite(p->dt))
        return set_error(OC_ERR_INVALID_ARGUMENT, "dt must be positive");
    if (p->solver_iterations < 1 || p->solver_iterations > 1000)
        return set_error(OC_ERR_INVALID_ARGUMENT, "solver_iterations out of range");
    return OC_OK;
}

}  // namespace

extern "C" {

int oc_pbd_create(int32_t n_vertices, const float* x0, const uint8_t* pinned,
                  const oc_pbd_distance* distance, int32_t n_distance,
                  const oc_pbd_bending* bending, int32_t n_bending, const oc_pbd_params* params,
                  oc_pbd* out) {
    if (!out) return set_error(OC_ERR_INVALID_ARGUMENT, "out is NULL");
    *out = nullptr;
    if (!x0) return set_error(OC_ERR_INVALID_ARGUMENT, "x0 is NULL");
    if (n_vertices < 1 || n_distance < 0 || n_bending < 0)
        return set_error(OC_ERR_INVALID_ARGUMENT, "bad counts");
    if ((n_distance > 0 && !distance) || (n_bending > 0 && !bending))
        return set_error(OC_ERR_INVALID_ARGUMENT, "constraint list is NULL");
    int rc = check_pbd_params(params);
    if (rc != OC_OK) return rc;
    const int N = n_vertices;
    for (int c = 0; c < n_distance; c++) {
        const oc_pbd_distance& d = distance[c];
        if (d.p1 < 0 || d.p1 >= N || d.p2 < 0 || d.p2 >= N || d.p1 == d.p2)
            return set_error(OC_ERR_INVALID_ARGUMENT,
                             "distance constraint " + std::to_string(c) + ": bad particle index");
    }
    for (int c = 0; c < n_bending; c++) {
        const oc_pbd_bending& b = bending[c];
        if (b.p1 < 0 || b.p1 >= N || b.p2 < 0 || b.p2 >= N || b.p3 < 0 || b.p3 >= N ||
            b.p1 == b.p2 || b.p1 == b.p3 || b.p2 == b.p3)
            return set_error(OC_ERR_INVALID_ARGUMENT,
                             "bending constraint " + std::to_string(c) + ": bad particle index");
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return set_error(OC_ERR_NO_DEVICE, "no HIP device");

    // set-up of InitGL (PBD:148-178, PBD:345-351)
    const int ns = params->solver_iterations;
    std::vector<float> W((size_t)N);
    for (int i = 0; i < N; i++) W[i] = (pinned && pinned[i]) ? 0.0f : 1.0f / params->mass;
    std::vector<int32_t> dp(2 * (size_t)n_distance), bp(3 * (size_t)n_bending);
    for (int c = 0; c < n_distance; c++) {
        dp[2 * c] = distance[c].p1;
        dp[2 * c + 1] = distance[c].p2;
    }
    for (int c = 0; c < n_bending; c++) {
        bp[3 * c] = bending[c].p1;
        bp[3 * c + 1] = bending[c].p2;
        bp[3 * c + 2] = bending[c].p3;
    }
    oc_pbd_s* p = new (std::nothrow) oc_pbd_s();
    if (!p) return set_error(OC_ERR_OUT_OF_MEMORY, "host allocation failed");
    std::vector<int32_t> d_order, b_order;
    const bool list_order = (params->flags & OC_FLAG_LIST_ORDER) != 0;
    if (list_order) {
        levels(N, n_distance, 2, dp, d_order, p->d_ptr);
        levels(N, n_bending, 3, bp, b_order, p->b_ptr);
    } else if (colour(N, n_distance, 2, dp, d_order, p->d_ptr) < 0 ||
               colour(N, n_bending, 3, bp, b_order, p->b_ptr) < 0) {
        delete p;
        return set_error(OC_ERR_UNSUPPORTED_TOPOLOGY, "constraints need more than 64 colours");
    }
    const size_t nd = n_distance, nb = n_bending;
    std::vector<int32_t> d1(nd), d2(nd), b1(nb), b2(nb), b3(nb);
    std::vector<float> drest(nd), dkp(nd), brest(nb), bw(nb), bkp(nb);
    for (size_t j = 0; j < nd; j++) {
        const int c = d_order[j];
        const int a = dp[2 * c], b = dp[2 * c + 1];
        d1[j] = a;
        d2[j] = b;
        const float dx = x0[3 * a] - x0[3 * b], dy = x0[3 * a + 1] - x0[3 * b + 1],
                    dz = x0[3 * a + 2] - x0[3 * b + 2];
        drest[j] = sqrtf(dx * dx + dy * dy + dz * dz);
        dkp[j] = k_prime(clamp01(distance[c].k), ns);
    }
    for (size_t j = 0; j < nb; j++) {
        const int c = b_order[j];
        const int a = bp[3 * c], b = bp[3 * c + 1], d = bp[3 * c + 2];
        b1[j] = a;
        b2[j] = b;
        b3[j] = d;
        bw[j] = W[a] + W[b] + 2.0f * W[d];
        float dc[3];
        for (int k = 0; k < 3; k++) {
            const float ctr = 0.3333f * ((x0[3 * a + k] + x0[3 * b + k]) + x0[3 * d + k]);
            dc[k] = x0[3 * d + k] - ctr;
        }
        brest[j] = sqrtf(dc[0] * dc[0] + dc[1] * dc[1] + dc[2] * dc[2]);
        bkp[j] = k_prime(clamp01(bending[c].k), ns);
    }

    p->N = N;
    p->nd = n_distance;
    p->nb = n_bending;
    p->prm = *params;
    const int nblk = (N + oc::kPbdBlock - 1) / oc::kPbdBlock;
    const size_t vec = align_up(sizeof(float) * 3 * N);
    const size_t sN = align_up(sizeof(float) * N);
    const size_t sd = align_up(4 * std::max<size_t>(nd, 1)), sb = align_up(4 * std::max<size_t>(nb, 1));
    const size_t spart = align_up(sizeof(float) * oc::kPbdSums * nblk);
    const size_t ssum = align_up(sizeof(float) * oc::kPbdSums);
    const size_t sdl = align_up(4 * p->d_ptr.size()), sbl = align_up(4 * p->b_ptr.size());
    const size_t total = 3 * vec + sN + 4 * sd + 6 * sb + spart + ssum + sdl + sbl;
    if (hipMalloc(&p->arena, total) != hipSuccess) {
        delete p;
        return set_error(OC_ERR_OUT_OF_MEMORY, "hipMalloc failed");
    }
    char* q = (char*)p->arena;
    auto take = [&](size_t bytes) {
        char* r = q;
        q += bytes;
        return r;
    };
    oc::DevPbd& d = p->dev;
    d.N = N;
    d.nblk = nblk;
    d.x = (float*)take(vec);
    d.v = (float*)take(vec);
    d.t = (float*)take(vec);
    float* w = (float*)take(sN);
    int32_t* dd1 = (int32_t*)take(sd);
    int32_t* dd2 = (int32_t*)take(sd);
    float* ddr = (float*)take(sd);
    float* ddk = (float*)take(sd);
    int32_t* bb1 = (int32_t*)take(sb);
    int32_t* bb2 = (int32_t*)take(sb);
    int32_t* bb3 = (int32_t*)take(sb);
    float* bbr = (float*)take(sb);
    float* bbw = (float*)take(sb);
    float* bbk = (float*)take(sb);
    d.part = (float*)take(spart);
    d.sums = (float*)take(ssum);
    int32_t* dlv = (int32_t*)take(sdl);
    int32_t* blv = (int32_t*)take(sbl);
    d.d_lv = dlv;
    d.b_lv = blv;
    d.nd_lv = (int)p->d_ptr.size() - 1;
    d.nb_lv = (int)p->b_ptr.size() - 1;
    d.w = w;
    d.d_p1 = dd1;
    d.d_p2 = dd2;
    d.d_rest = ddr;
    d.d_kp = ddk;
    d.b_p1 = bb1;
    d.b_p2 = bb2;
    d.b_p3 = bb3;
    d.b_rest = bbr;
    d.b_w = bbw;
    d.b_kp = bbk;
    bool ok = hipMemset(p->arena, 0, total) == hipSuccess &&
              hipMemcpy(w, W.data(), 4 * (size_t)N, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(dlv, p->d_ptr.data(), 4 * p->d_ptr.size(), hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(blv, p->b_ptr.data(), 4 * p->b_ptr.size(), hipMemcpyHostToDevice) == hipSuccess;
    if (ok && nd)
        ok = hipMemcpy(dd1, d1.data(), 4 * nd, hipMemcpyHostToDevice) == hipSuccess &&
             hipMemcpy(dd2, d2.data(), 4 * nd, hipMemcpyHostToDevice) == hipSuccess &&
             hipMemcpy(ddr, drest.data(), 4 * nd, hipMemcpyHostToDevice) == hipSuccess &&
             hipMemcpy(ddk, dkp.data(), 4 * nd, hipMemcpyHostToDevice) == hipSuccess;
    if (ok && nb)
        ok = hipMemcpy(bb1, b1.data(), 4 * nb, hipMemcpyHostToDevice) == hipSuccess &&
             hipMemcpy(bb2, b2.data(), 4 * nb, hipMemcpyHostToDevice) == hipSuccess &&
             hipMemcpy(bb3, b3.data(), 4 * nb, hipMemcpyHostToDevice) == hipSuccess &&
             hipMemcpy(bbr, brest.data(), 4 * nb, hipMemcpyHostToDevice) == hipSuccess &&
             hipMemcpy(bbw, bw.data(), 4 * nb, hipMemcpyHostToDevice) == hipSuccess &&
             hipMemcpy(bbk, bkp.data(), 4 * nb, hipMemcpyHostToDevice) == hipSuccess;
    if (!ok) {
        (void)hipFree(p->arena);
        delete p;
        return set_error(OC_ERR_DEVICE, "upload failed");
    }
    *out = p;
    rc = oc_pbd_set_state(p, x0, nullptr);
    if (rc != OC_OK) {
        std::string e = oc_last_error();
        oc_pbd_destroy(p);
        *out = nullptr;
        return set_error(rc, e);
    }
    return OC_OK;
}

int oc_pbd_destroy(oc_pbd p) {
    if (!p) return OC_OK;
    (void)hipStreamSynchronize(p->stream);
    if (p->arena) (void)hipFree(p->arena);
    delete p;
    return OC_OK;
}

int oc_pbd_set_stream(oc_pbd p, void* stream) {
    if (!p) return set_error(OC_ERR_INVALID_ARGUMENT, "pbd is NULL");
    p->stream = (hipStream_t)stream;
    return OC_OK;
}

// x and v may be NULL to keep that vector, except that v defaults to zero on
// the first call from oc_pbd_create.
int oc_pbd_set_state(oc_pbd p, const float* x, const float* v) {
    if (!p) return set_error(OC_ERR_INVALID_ARGUMENT, "pbd is NULL");
    int rc;
    if (x && (rc = upload_aos(p, p->dev.x, x)) != OC_OK) return rc;
    if (v && (rc = upload_aos(p, p->dev.v, v)) != OC_OK) return rc;
    return OC_OK;
}

int oc_pbd_get_state(oc_pbd p, float* x, float* v) {
    if (!p) return set_error(OC_ERR_INVALID_ARGUMENT, "pbd is NULL");
    int rc;
    if (x && (rc = download_aos(p, p->dev.x, x)) != OC_OK) return rc;
    if (v && (rc = download_aos(p, p->dev.v, v)) != OC_OK) return rc;
    return OC_OK;
}

int oc_pbd_num_colors(oc_pbd p, int32_t* n_distance_colors, int32_t* n_bending_colors) {
    if (!p) return set_error(OC_ERR_INVALID_ARGUMENT, "pbd is NULL");
    if (n_distance_colors) *n_distance_colors = (int32_t)p->d_ptr.size() - 1;
    if (n_bending_colors) *n_bending_colors = (int32_t)p->b_ptr.size() - 1;
    return OC_OK;
}

int oc_pbd_step(oc_pbd p, int32_t nsteps, const oc_ellipsoid* e) {
    if (!p) return set_error(OC_ERR_INVALID_ARGUMENT, "pbd is NULL");
    if (nsteps < 0) return set_error(OC_ERR_INVALID_ARGUMENT, "nsteps < 0");
    const oc::PbdScalars s = pscalars(p);
    const oc::DevPbd& d = p->dev;
    const int nd_col = (int)p->d_ptr.size() - 1, nb_col = (int)p->b_ptr.size() - 1;
    for (int n = 0; n < nsteps; n++) {
        if (p->prm.flags & OC_FLAG_LIST_ORDER) {
            OC_HIP(oc::launch_pbd_predict_seq(d, s, p->stream));
            OC_HIP(oc::launch_pbd_levels(d, s, p->prm.solver_iterations,
                                         (p->prm.flags & OC_FLAG_GROUND_CLAMP) ? 1 : 0, p->stream));
        } else {
            OC_HIP(oc::launch_pbd_predict(d, s, p->stream));
            for (int it = 0; it < p->prm.solver_iterations; it++) {
                for (int c = 0; c < nd_col; c++)
                    OC_HIP(oc::launch_pbd_distance(d, p->d_ptr[c], p->d_ptr[c + 1] - p->d_ptr[c], p->stream));
                for (int c = 0; c < nb_col; c++)
                    OC_HIP(oc::launch_pbd_bending(d, s, p->b_ptr[c], p->b_ptr[c + 1] - p->b_ptr[c], p->stream));
                if (p->prm.flags & OC_FLAG_GROUND_CLAMP) OC_HIP(oc::launch_pbd_ground(d, p->stream));
            }
        }
        if (e) OC_HIP(oc::launch_ellipsoid(d.t, d.v, d.N, *e, 0, p->stream));
        OC_HIP(oc::launch_pbd_integrate(d, s, p->stream));
    }
    return OC_OK;
}

int oc_pbd_synchronize(oc_pbd p) {
    if (!p) return set_error(OC_ERR_INVALID_ARGUMENT, "pbd is NULL");
    OC_HIP(hipStreamSynchronize(p->stream));
    return OC_OK;
}

}  // extern "C"
