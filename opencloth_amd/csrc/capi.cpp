// C ABI of the implicit cloth (include/opencloth.h): handle lifetime, host
// <-> device copies (AoS float3 on the host, SoA planes on the device) and
// the launch sequence of one step.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <utility>
#include <vector>

#include "capi_common.hpp"
#include "oc_internal.hpp"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

using oc::align_up;

}  // namespace

namespace oc {
int set_error(int code, const std::string& msg) { return fail(code, msg); }
}  // namespace oc

struct oc_cloth_s {
    oc::GridTopology topo;
    oc_params prm{};
    hipStream_t stream = nullptr;
    int device = 0;
    int max_it_alloc = 0;
    void* arena = nullptr;        // vertex vectors, topology, blocks
    void* scratch = nullptr;      // CG partials and scalars
    int* ctl_host = nullptr;      // pinned copy of ctl
    oc::DevCloth dc{};
    float* alt_x = nullptr;  // the other (x, v) set: a fused update writes here,
    float* alt_v = nullptr;  // then the two sets swap
    int32_t* spring_cst = nullptr;  // stencil assembly form: owner-plane index per spring
    bool assembled = false;
    // differentiable rollouts: checkpoints, one segment of states, adjoint vectors
    void* adj_arena = nullptr;
    size_t adj_bytes = 0;
    int roll_steps = 0, roll_every = 0;
    bool roll_valid = false;
};

namespace {

oc::StepScalars scalars(const oc_cloth_s* c, float dt) {
    oc::StepScalars s;
    s.mass = c->prm.mass;
    s.gx = c->prm.gravity[0];
    s.gy = c->prm.gravity[1];
    s.gz = c->prm.gravity[2];
    s.h = dt;
    s.hh = dt * dt;
    s.tol2 = c->prm.cg_tol * c->prm.cg_tol;
    s.max_it = c->prm.cg_max_iters;
    s.tol_mode = c->prm.cg_tol_mode;
    s.warm = c->prm.warm_start ? 1 : 0;
    s.flags = c->prm.flags;
    s.damp = c->prm.damping;
    s.md = s.mass - s.h * s.damp;
    return s;
}

int check_params(const oc_params* p) {
    if (!p) return fail(OC_ERR_INVALID_ARGUMENT, "params is NULL");
    if (!(p->mass > 0.0f) || !std::isfinite(p->mass))
        return fail(OC_ERR_INVALID_ARGUMENT, "mass must be positive");
    if (!(p->dt > 0.0f) || !std::isfinite(p->dt))
        return fail(OC_ERR_INVALID_ARGUMENT, "dt must be positive");
    if (p->cg_max_iters < 0 || p->cg_max_iters > 100000)
        return fail(OC_ERR_INVALID_ARGUMENT, "cg_max_iters out of range");
    if (p->cg_tol_mode < 0 || p->cg_tol_mode > 2)
        return fail(OC_ERR_INVALID_ARGUMENT, "unknown cg_tol_mode");
    if (!(p->cg_tol >= 0.0f)) return fail(OC_ERR_INVALID_ARGUMENT, "cg_tol must be >= 0");
    if (!std::isfinite(p->damping)) return fail(OC_ERR_INVALID_ARGUMENT, "damping must be finite");
    return OC_OK;
}

// (Re)allocate the CG partials for an iteration cap of max_iters.  The new
// buffer is allocated before the old one is released, so on failure the
// handle keeps working with its previous cap.
int ensure_scratch(oc_cloth_s* c, int max_iters) {
    const int need = std::max(max_iters, 1);
    if (c->scratch && need <= c->max_it_alloc) return OC_OK;
    const size_t nt = c->dc.nt;
    // scal holds the reduced r.r values and (single-reduction CG) the alphas
    const size_t del = align_up(sizeof(float) * (need + 1) * nt), sig = align_up(sizeof(float) * (need + 1) * nt),
                 scal = align_up(sizeof(float) * 2 * (need + 2));
    const size_t bytes = del + sig + scal + align_up(sizeof(int) * 8);
    void* fresh = nullptr;
    if (hipMalloc(&fresh, bytes) != hipSuccess)
        return fail(OC_ERR_OUT_OF_MEMORY, "hipMalloc of CG scratch failed");
    if (c->scratch) {
        if (hipStreamSynchronize(c->stream) != hipSuccess) {
            (void)hipFree(fresh);
            return fail(OC_ERR_DEVICE, "hipStreamSynchronize failed");
        }
        (void)hipFree(c->scratch);
    }
    c->scratch = fresh;
    char* p = (char*)fresh;
    c->dc.del_part = (float*)p;
    c->dc.sig_part = (float*)(p + del);
    c->dc.scal = (float*)(p + del + sig);
    c->dc.ctl = (int*)(p + del + sig + scal);
    c->dc.cg1_alpha = need + 2;
    c->max_it_alloc = need;
    OC_HIP(hipMemsetAsync(fresh, 0, bytes, c->stream));
    return OC_OK;
}

void aos_to_soa(const float* a, int N, std::vector<float>& out) {
    out.resize(3 * (size_t)N);
    for (int i = 0; i < N; i++)
        for (int k = 0; k < 3; k++) out[(size_t)k * N + i] = a[3 * (size_t)i + k];
}
void soa_to_aos(const std::vector<float>& s, int N, float* a) {
    for (int i = 0; i < N; i++)
        for (int k = 0; k < 3; k++) a[3 * (size_t)i + k] = s[(size_t)k * N + i];
}

int upload_vec(oc_cloth_s* c, float* dst, const float* host_aos) {
    std::vector<float> tmp;
    if (host_aos)
        aos_to_soa(host_aos, c->dc.N, tmp);
    else
        tmp.assign(3 * (size_t)c->dc.N, 0.0f);
    OC_HIP(hipMemcpyAsync(dst, tmp.data(), tmp.size() * sizeof(float), hipMemcpyHostToDevice,
                          c->stream));
    OC_HIP(hipStreamSynchronize(c->stream));
    return OC_OK;
}

int download_vec(oc_cloth_s* c, const float* src, float* host_aos) {
    std::vector<float> tmp(3 * (size_t)c->dc.N);
    OC_HIP(hipMemcpyAsync(tmp.data(), src, tmp.size() * sizeof(float), hipMemcpyDeviceToHost,
                          c->stream));
    OC_HIP(hipStreamSynchronize(c->stream));
    soa_to_aos(tmp, c->dc.N, host_aos);
    return OC_OK;
}

}  // namespace

namespace oc {

int cloth_device(const oc_cloth_s* c) { return c->device; }

size_t cloth_partial_floats(const oc_cloth_s* c) {
    const size_t it = (size_t)std::max(c->prm.cg_max_iters, 1), nt = (size_t)c->dc.nt;
    return 2 * (it + 1) * nt + 2 * (it + 2);
}

int cloth_step(oc_cloth_s* c, int nsteps, float* partials) {
    const StepScalars s = scalars(c, c->prm.dt);
    DevCloth d = c->dc;
    if (partials) {
        const size_t it = (size_t)std::max(s.max_it, 1), nt = (size_t)d.nt;
        d.del_part = partials;
        d.sig_part = partials + (it + 1) * nt;
        d.scal = partials + 2 * (it + 1) * nt;
        d.cg1_alpha = (int)it + 2;
    }
    // With a warm CG start (and at least one CG iteration), steps after the
    // first apply the previous step's update inside their assembly (one
    // launch fewer per step, the same bits); the last step's update, which
    // also records the solve's status, is a launch of its own.  A cold start
    // is not fused: its assembly zeroes dv, which neighbouring blocks of a
    // fused assembly would still be reading for the update.
    const bool fuse = s.warm && s.max_it >= 1;
    bool pending = false;  // a solve whose update has not been launched yet
    hipError_t e = hipSuccess;
    for (int k = 0; k < nsteps; k++) {
        if (pending) {
            DevCloth u = d;
            u.xo = c->alt_x;
            u.vo = c->alt_v;
            if ((e = launch_assemble(u, s, c->stream, true)) != hipSuccess) break;
            std::swap(d.x, c->alt_x);
            std::swap(d.v, c->alt_v);
            pending = false;
        } else if ((e = launch_assemble(d, s, c->stream)) != hipSuccess) {
            break;
        }
        if ((e = launch_cg(d, s, c->stream, false)) != hipSuccess) break;
        if (fuse) {
            pending = true;
            continue;
        }
        if ((e = launch_update(d, s, c->stream, true)) != hipSuccess) break;
    }
    // The last solve's update; after a failed launch, the update of the last
    // complete solve, so that x and v are never left one update behind it.
    const hipError_t eu = pending ? launch_update(d, s, c->stream, true) : hipSuccess;
    c->dc.x = d.x;
    c->dc.v = d.v;
    c->assembled = false;
    if (e == hipSuccess) e = eu;
    if (e != hipSuccess) return fail(OC_ERR_DEVICE, std::string("step launch failed: ") + hipGetErrorString(e));
    return OC_OK;
}

}  // namespace oc

extern "C" {

const char* oc_version(void) { return "opencloth_amd 0.1 (gfx950)"; }

const char* oc_last_error(void) { return g_err.c_str(); }

int oc_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int oc_cloth_tile_shape(int32_t* tile_x, int32_t* tile_y) {
    if (!tile_x || !tile_y) return fail(OC_ERR_INVALID_ARGUMENT, "NULL argument");
    *tile_x = oc::kTileX;
    *tile_y = oc::kTileY;
    return OC_OK;
}

int oc_grid_triangles(int32_t nx, int32_t ny, int32_t* out) {
    if (!out) return fail(OC_ERR_INVALID_ARGUMENT, "out is NULL");
    if (nx < 2 || ny < 2) return fail(OC_ERR_INVALID_ARGUMENT, "grid must be at least 2x2");
    int32_t* o = out;
    for (int i = 0; i < ny - 1; i++)
        for (int j = 0; j < nx - 1; j++) {
            const int32_t i0 = i * nx + j, i1 = i0 + 1, i2 = i0 + nx, i3 = i2 + 1;
            const int32_t odd[6] = {i0, i2, i1, i1, i2, i3}, even[6] = {i0, i2, i3, i0, i3, i1};
            const int32_t* t = ((i + j) & 1) ? odd : even;
            for (int k = 0; k < 6; k++) *o++ = t[k];
        }
    return OC_OK;
}

int oc_cloth_create(int32_t nx, int32_t ny, const float* x0, const float* v0,
                    const oc_spring* springs, int32_t n_springs, const uint8_t* pinned,
                    const oc_params* params, oc_cloth* out) {
    g_err.clear();
    if (!out) return fail(OC_ERR_INVALID_ARGUMENT, "out is NULL");
    *out = nullptr;
    if (!x0) return fail(OC_ERR_INVALID_ARGUMENT, "x0 is NULL");
    int rc = check_params(params);
    if (rc != OC_OK) return rc;
    if (oc_device_count() < 1) return fail(OC_ERR_NO_DEVICE, "no HIP device");

    oc_cloth_s* c = new (std::nothrow) oc_cloth_s();
    if (!c) return fail(OC_ERR_OUT_OF_MEMORY, "host allocation failed");
    std::string err;
    rc = oc::build_grid_topology(nx, ny, springs, n_springs, &c->topo, &err);
    if (rc != OC_OK) {
        delete c;
        return fail(rc, err);
    }
    c->prm = *params;
    if (hipGetDevice(&c->device) != hipSuccess) {
        delete c;
        return fail(OC_ERR_DEVICE, "hipGetDevice failed");
    }
    const oc::GridTopology& t = c->topo;
    const size_t N = t.N, S = std::max(t.S, 1);
    oc::DevCloth& d = c->dc;
    d.nx = t.nx;
    d.ny = t.ny;
    d.N = t.N;
    d.ntx = (t.nx + oc::kTileX - 1) / oc::kTileX;
    d.nty = (t.ny + oc::kTileY - 1) / oc::kTileY;
    d.nt = d.ntx * d.nty;
    d.wa = t.wa;
    d.wp = t.wp;
    d.persistent = (params->flags & OC_FLAG_CG_PERSISTENT) ? oc::persistent_cg_grid(d) : 0;

    // the CG and assembly kernels address every array with 32-bit byte offsets
    const uint64_t largest = std::max<uint64_t>(
        {4ull * 6 * N * (uint64_t)std::max(t.n_planes, 1), 4ull * N * (uint64_t)std::max(t.wa, 1),
         2ull * N * (uint64_t)std::max(t.wp, 1), 4ull * S, 12ull * N, 4ull * oc::kCstPlanes * N});
    if (largest >= (1ull << 32)) {
        delete c;
        return fail(OC_ERR_UNSUPPORTED_TOPOLOGY,
                    "cloth too large: a device array would exceed 4 GiB (" + std::to_string(largest) +
                        " bytes)");
    }
    const size_t vec = align_up(sizeof(float) * 3 * N);
    const size_t blocks = align_up(sizeof(float) * 6 * N * std::max(t.n_planes, 1));
    const size_t ainc_s = align_up(sizeof(int32_t) * std::max<size_t>(t.ainc_spring.size(), 1));
    const size_t ainc_m = align_up(sizeof(uint16_t) * std::max<size_t>(t.ainc_meta.size(), 1));
    const size_t pair_m = align_up(sizeof(uint16_t) * std::max<size_t>(t.pair_meta.size(), 1));
    const size_t pair_k = align_up(sizeof(uint16_t) * std::max<size_t>(t.pair_mask.size(), 1));
    const size_t sp = align_up(sizeof(float) * S);
    const size_t pin = align_up(N);
    // stencil assembly form (each spring evaluated once per tile), unless
    // OPENCLOTH_ASSEMBLY=per-end asks for the per-end kernel (measurement)
    const char* asm_env = std::getenv("OPENCLOTH_ASSEMBLY");
    const bool st_asm = t.stencil_asm && !(asm_env && std::strcmp(asm_env, "per-end") == 0);
    const size_t amask = st_asm ? align_up(sizeof(uint32_t) * N) : 0;
    const size_t cst = st_asm ? align_up(sizeof(float) * oc::kCstPlanes * N) : 0;
    const size_t scst = st_asm ? align_up(sizeof(int32_t) * S) : 0;
    const size_t total = 17 * vec + blocks + ainc_s + ainc_m + pair_m + pair_k + 3 * sp + pin + amask + cst + scst;
    if (hipMalloc(&c->arena, total) != hipSuccess) {
        delete c;
        return fail(OC_ERR_OUT_OF_MEMORY, "hipMalloc of " + std::to_string(total) + " bytes failed");
    }
    char* p = (char*)c->arena;
    auto take = [&](size_t n) {
        char* q = p;
        p += n;
        return q;
    };
    d.x = (float*)take(vec);
    d.v = (float*)take(vec);
    d.dv = (float*)take(vec);
    d.r = (float*)take(vec);
    d.d0 = (float*)take(vec);
    d.d1 = (float*)take(vec);
    d.q = (float*)take(vec);
    d.b = (float*)take(vec);
    d.f = (float*)take(vec);
    d.s = (float*)take(vec);
    d.pinv = (float*)take(vec);
    d.blocks = (float*)take(blocks);
    d.ainc_s = (const int32_t*)take(ainc_s);
    d.ainc_m = (const uint16_t*)take(ainc_m);
    d.pair_m = (const uint16_t*)take(pair_m);
    d.pair_mask = (const uint16_t*)take(pair_k);
    d.stencil = t.stencil ? 1 : 0;
    for (int k = 0; k < oc::kStencil; k++) d.stencil_plane[k] = t.stencil_plane[k];
    d.L0 = (const float*)take(sp);
    d.ks = (const float*)take(sp);
    d.kd = (const float*)take(sp);
    d.pin = (const uint8_t*)take(pin);
    c->alt_x = (float*)take(vec);  // last, so the other arrays keep their offsets
    c->alt_v = (float*)take(vec);
    d.stencil_asm = st_asm ? 1 : 0;
    d.asm_mask = st_asm ? (const uint32_t*)take(amask) : nullptr;
    d.cst = st_asm ? (const float*)take(cst) : nullptr;
    c->spring_cst = st_asm ? (int32_t*)take(scst) : nullptr;
    d.r2 = (float*)take(vec);
    d.q2 = (float*)take(vec);
    d.s1 = (float*)take(vec);
    d.s2 = (float*)take(vec);

    auto cleanup = [&](int code) {
        (void)hipFree(c->arena);
        if (c->scratch) (void)hipFree(c->scratch);
        std::string keep = g_err;
        delete c;
        g_err = keep;
        return code;
    };
    if (hipMemset(c->arena, 0, total) != hipSuccess)
        return cleanup(fail(OC_ERR_DEVICE, "hipMemset failed"));
    std::vector<float> L0(S, 0.0f), ks(S, 0.0f), kd(S, 0.0f);
    for (int s = 0; s < t.S; s++) {
        L0[s] = springs[s].rest_length;
        ks[s] = springs[s].Ks;
        kd[s] = springs[s].Kd;
    }
    std::vector<uint8_t> pins(N, 0);
    if (pinned)
        for (size_t i = 0; i < N; i++) pins[i] = pinned[i] ? 1 : 0;
    bool ok = hipMemcpy((void*)d.ainc_s, t.ainc_spring.data(), t.ainc_spring.size() * 4,
                        hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy((void*)d.ainc_m, t.ainc_meta.data(), t.ainc_meta.size() * 2,
                        hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy((void*)d.pair_m, t.pair_meta.data(), t.pair_meta.size() * 2,
                        hipMemcpyHostToDevice) == hipSuccess &&
              (t.pair_mask.empty() ||
               hipMemcpy((void*)d.pair_mask, t.pair_mask.data(), t.pair_mask.size() * 2,
                         hipMemcpyHostToDevice) == hipSuccess) &&
              hipMemcpy((void*)d.L0, L0.data(), S * 4, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy((void*)d.ks, ks.data(), S * 4, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy((void*)d.kd, kd.data(), S * 4, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy((void*)d.pin, pins.data(), N, hipMemcpyHostToDevice) == hipSuccess;
    if (ok && st_asm)
        ok = hipMemcpy((void*)d.asm_mask, t.asm_mask.data(), N * 4, hipMemcpyHostToDevice) == hipSuccess &&
             (t.S == 0 || hipMemcpy(c->spring_cst, t.spring_cst.data(), (size_t)t.S * 4,
                                    hipMemcpyHostToDevice) == hipSuccess) &&
             oc::launch_scatter_cst(d, c->spring_cst, t.S, 0, 3, nullptr) == hipSuccess &&
             hipDeviceSynchronize() == hipSuccess;
    if (!ok) return cleanup(fail(OC_ERR_DEVICE, "upload of topology failed"));
    if (hipHostMalloc((void**)&c->ctl_host, sizeof(int) * 8) != hipSuccess)
        return cleanup(fail(OC_ERR_OUT_OF_MEMORY, "hipHostMalloc failed"));
    if ((rc = ensure_scratch(c, params->cg_max_iters)) != OC_OK) return cleanup(rc);
    if ((rc = upload_vec(c, d.x, x0)) != OC_OK) return cleanup(rc);
    if ((rc = upload_vec(c, d.v, v0)) != OC_OK) return cleanup(rc);
    *out = c;
    return OC_OK;
}

int oc_cloth_destroy(oc_cloth c) {
    if (!c) return OC_OK;
    (void)hipStreamSynchronize(c->stream);
    if (c->arena) (void)hipFree(c->arena);
    if (c->scratch) (void)hipFree(c->scratch);
    if (c->ctl_host) (void)hipHostFree(c->ctl_host);
    if (c->adj_arena) (void)hipFree(c->adj_arena);
    delete c;
    return OC_OK;
}

int oc_cloth_set_stream(oc_cloth c, void* stream) {
    if (!c) return fail(OC_ERR_INVALID_ARGUMENT, "cloth is NULL");
    c->stream = (hipStream_t)stream;
    return OC_OK;
}

int oc_cloth_set_params(oc_cloth c, const oc_params* p) {
    if (!c) return fail(OC_ERR_INVALID_ARGUMENT, "cloth is NULL");
    int rc = check_params(p);
    if (rc != OC_OK) return rc;
    if ((rc = ensure_scratch(c, p->cg_max_iters)) != OC_OK) return rc;  // old params stay on failure
    c->prm = *p;
    c->dc.persistent = (p->flags & OC_FLAG_CG_PERSISTENT) ? oc::persistent_cg_grid(c->dc) : 0;
    return OC_OK;
}

int oc_cloth_get_params(oc_cloth c, oc_params* p) {
    if (!c || !p) return fail(OC_ERR_INVALID_ARGUMENT, "NULL argument");
    *p = c->prm;
    return OC_OK;
}

int oc_cloth_set_state(oc_cloth c, const float* x, const float* v) {
    if (!c) return fail(OC_ERR_INVALID_ARGUMENT, "cloth is NULL");
    int rc;
    if (x && (rc = upload_vec(c, c->dc.x, x)) != OC_OK) return rc;
    if (v && (rc = upload_vec(c, c->dc.v, v)) != OC_OK) return rc;
    c->assembled = false;
    return OC_OK;
}

int oc_cloth_get_state(oc_cloth c, float* x, float* v) {
    if (!c) return fail(OC_ERR_INVALID_ARGUMENT, "cloth is NULL");
    int rc;
    if (x && (rc = download_vec(c, c->dc.x, x)) != OC_OK) return rc;
    if (v && (rc = download_vec(c, c->dc.v, v)) != OC_OK) return rc;
    return OC_OK;
}

int oc_cloth_get_dv(oc_cloth c, float* dv) {
    if (!c || !dv) return fail(OC_ERR_INVALID_ARGUMENT, "NULL argument");
    return download_vec(c, c->dc.dv, dv);
}

int oc_cloth_set_dv(oc_cloth c, const float* dv) {
    if (!c) return fail(OC_ERR_INVALID_ARGUMENT, "cloth is NULL");
    return upload_vec(c, c->dc.dv, dv);
}

int oc_cloth_compute_forces(oc_cloth c, float* f_out) {
    if (!c) return fail(OC_ERR_INVALID_ARGUMENT, "cloth is NULL");
    OC_HIP(oc::launch_forces(c->dc, scalars(c, c->prm.dt), c->stream));
    if (f_out) return download_vec(c, c->dc.f, f_out);
    return OC_OK;
}

int oc_cloth_assemble(oc_cloth c, float dt) {
    if (!c) return fail(OC_ERR_INVALID_ARGUMENT, "cloth is NULL");
    if (!(dt > 0.0f)) return fail(OC_ERR_INVALID_ARGUMENT, "dt must be positive");
    OC_HIP(oc::launch_assemble(c->dc, scalars(c, dt), c->stream));
    c->assembled = true;
    return OC_OK;
}

int oc_cloth_solve(oc_cloth c, int32_t* iters, int32_t* status) {
    if (!c) return fail(OC_ERR_INVALID_ARGUMENT, "cloth is NULL");
    if (!c->assembled) return fail(OC_ERR_INVALID_ARGUMENT, "solve before assemble");
    OC_HIP(oc::launch_cg(c->dc, scalars(c, c->prm.dt), c->stream));
    if (iters || status) {
        OC_HIP(hipMemcpyAsync(c->ctl_host, c->dc.ctl, sizeof(int) * 4, hipMemcpyDeviceToHost,
                              c->stream));
        OC_HIP(hipStreamSynchronize(c->stream));
        if (iters) *iters = c->ctl_host[1];
        if (status) *status = c->ctl_host[2];
    }
    return OC_OK;
}

int oc_cloth_update(oc_cloth c, float dt) {
    if (!c) return fail(OC_ERR_INVALID_ARGUMENT, "cloth is NULL");
    if (!(dt > 0.0f)) return fail(OC_ERR_INVALID_ARGUMENT, "dt must be positive");
    OC_HIP(oc::launch_update(c->dc, scalars(c, dt), c->stream));
    c->assembled = false;
    return OC_OK;
}

int oc_cloth_integrate_implicit(oc_cloth c, float dt) {
    if (!c) return fail(OC_ERR_INVALID_ARGUMENT, "cloth is NULL");
    if (!(dt > 0.0f)) return fail(OC_ERR_INVALID_ARGUMENT, "dt must be positive");
    const oc::StepScalars s = scalars(c, dt);
    OC_HIP(oc::launch_assemble(c->dc, s, c->stream));
    OC_HIP(oc::launch_cg(c->dc, s, c->stream, false));
    OC_HIP(oc::launch_update(c->dc, s, c->stream, true));
    c->assembled = false;
    return OC_OK;
}

int oc_cloth_step(oc_cloth c, int32_t nsteps) {
    if (!c) return fail(OC_ERR_INVALID_ARGUMENT, "cloth is NULL");
    if (nsteps < 0) return fail(OC_ERR_INVALID_ARGUMENT, "nsteps < 0");
    return oc::cloth_step(c, nsteps, nullptr);
}

int oc_cloth_sticky_status(oc_cloth c, int32_t* status) {
    if (!c || !status) return fail(OC_ERR_INVALID_ARGUMENT, "NULL argument");
    OC_HIP(hipMemcpyAsync(c->ctl_host, c->dc.ctl, sizeof(int) * 8, hipMemcpyDeviceToHost,
                          c->stream));
    OC_HIP(hipMemsetAsync(c->dc.ctl + 3, 0, sizeof(int), c->stream));
    OC_HIP(hipStreamSynchronize(c->stream));
    *status = c->ctl_host[3];
    if (c->ctl_host[4]) return fail(OC_ERR_DEVICE, "persistent CG grid barrier timed out");
    return OC_OK;
}

int oc_cloth_profile_steps(oc_cloth c, int32_t nsteps, float* ms, int32_t* launches) {
    if (!c || !ms || !launches) return fail(OC_ERR_INVALID_ARGUMENT, "NULL argument");
    if (nsteps < 0) return fail(OC_ERR_INVALID_ARGUMENT, "nsteps < 0");
    const oc::StepScalars s = scalars(c, c->prm.dt);
    // the launches oc_cloth_step makes: per step an assembly and the solver's
    // launches (at most 2 per iteration, a barrier reset and a finish), one
    // update per step or per call
    const int cap = nsteps * (2 * s.max_it + 4) + 2;
    std::vector<hipEvent_t> e0(cap), e1(cap);
    std::vector<int> kind(cap);
    int made = 0;
    int rc = OC_OK;
    for (; made < cap; made++) {
        if (hipEventCreate(&e0[made]) != hipSuccess) break;
        if (hipEventCreate(&e1[made]) != hipSuccess) {
            (void)hipEventDestroy(e0[made]);
            break;
        }
    }
    if (made < cap) rc = fail(OC_ERR_DEVICE, "hipEventCreate failed");
    oc::LaunchProfile lp{e0.data(), e1.data(), kind.data(), 0, cap};
    if (rc == OC_OK) {
        oc::set_launch_profile(&lp);
        rc = oc::cloth_step(c, nsteps, nullptr);
        oc::set_launch_profile(nullptr);
    }
    for (int k = 0; k < 5; k++) {
        ms[k] = 0.0f;
        launches[k] = 0;
    }
    if (rc == OC_OK && lp.n > 0 && hipEventSynchronize(e1[lp.n - 1]) != hipSuccess)
        rc = fail(OC_ERR_DEVICE, "hipEventSynchronize failed");
    for (int k = 0; rc == OC_OK && k < lp.n; k++) {
        float t = 0.0f;
        if (hipEventElapsedTime(&t, e0[k], e1[k]) != hipSuccess) {
            rc = fail(OC_ERR_DEVICE, "hipEventElapsedTime failed");
            break;
        }
        ms[kind[k]] += t;
        launches[kind[k]]++;
    }
    for (int k = 0; k < made; k++) {
        (void)hipEventDestroy(e0[k]);
        (void)hipEventDestroy(e1[k]);
    }
    return rc;
}

int oc_cloth_time_cg_spmv(oc_cloth c, int32_t reps, float* avg_us) {
    if (!c || !avg_us) return fail(OC_ERR_INVALID_ARGUMENT, "NULL argument");
    if (reps < 1) return fail(OC_ERR_INVALID_ARGUMENT, "reps < 1");
    const oc::StepScalars s = scalars(c, c->prm.dt);
    if (s.max_it < 2) return fail(OC_ERR_INVALID_ARGUMENT, "cg_max_iters must be >= 2");
    OC_HIP(oc::launch_assemble(c->dc, s, c->stream));
    OC_HIP(oc::launch_cg_spmv(c->dc, s, 0, c->stream));
    OC_HIP(oc::launch_cg_axpy(c->dc, s, 0, c->stream));
    OC_HIP(oc::launch_cg_spmv(c->dc, s, 1, c->stream));  // warm-up
    hipEvent_t e0, e1;
    OC_HIP(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return fail(OC_ERR_DEVICE, "hipEventCreate failed");
    }
    bool ok = hipEventRecord(e0, c->stream) == hipSuccess;
    for (int r = 0; ok && r < reps; r++) ok = oc::launch_cg_spmv(c->dc, s, 1, c->stream) == hipSuccess;
    float ms = 0.0f;
    ok = ok && hipEventRecord(e1, c->stream) == hipSuccess && hipEventSynchronize(e1) == hipSuccess &&
         hipEventElapsedTime(&ms, e0, e1) == hipSuccess;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    c->assembled = false;
    if (!ok) return fail(OC_ERR_DEVICE, "timed SpMV launches failed");
    *avg_us = 1000.0f * ms / (float)reps;
    return OC_OK;
}

int oc_cloth_time_kernel(oc_cloth c, int32_t kernel, int32_t reps, float* avg_us) {
    if (!c || !avg_us) return fail(OC_ERR_INVALID_ARGUMENT, "NULL argument");
    if (reps < 1) return fail(OC_ERR_INVALID_ARGUMENT, "reps < 1");
    if (kernel < OC_KERNEL_ASSEMBLE || kernel > OC_KERNEL_UPDATE)
        return fail(OC_ERR_INVALID_ARGUMENT, "unknown kernel class");
    oc::StepScalars s = scalars(c, c->prm.dt);
    if (s.max_it < 2) return fail(OC_ERR_INVALID_ARGUMENT, "cg_max_iters must be >= 2");
    if (kernel == OC_KERNEL_ASSEMBLE_FUSED && !s.warm)
        return fail(OC_ERR_INVALID_ARGUMENT, "the fused assembly needs a warm CG start");
    s.tol_mode = OC_TOL_NONE;  // every timed launch does its full work
    const size_t v3 = sizeof(float) * 3 * (size_t)c->dc.N;
    // the status words, and x, v, dv: the timed launches write copies
    OC_HIP(hipMemcpyAsync(c->ctl_host, c->dc.ctl, sizeof(int) * 8, hipMemcpyDeviceToHost, c->stream));
    OC_HIP(hipStreamSynchronize(c->stream));
    int saved[8];
    std::memcpy(saved, c->ctl_host, sizeof(saved));
    oc::DevCloth d = c->dc;
    OC_HIP(oc::launch_assemble(d, s, c->stream));
    OC_HIP(oc::launch_cg_spmv(d, s, 0, c->stream));
    OC_HIP(oc::launch_cg_axpy(d, s, 0, c->stream));
    OC_HIP(oc::launch_cg_spmv(d, s, 1, c->stream));
    if (kernel == OC_KERNEL_CG_AXPY) {  // dv += alpha d, r -= alpha q on copies
        OC_HIP(hipMemcpyAsync(c->alt_x, d.dv, v3, hipMemcpyDeviceToDevice, c->stream));
        OC_HIP(hipMemcpyAsync(c->alt_v, d.r, v3, hipMemcpyDeviceToDevice, c->stream));
        d.dv = c->alt_x;
        d.r = c->alt_v;
    } else if (kernel == OC_KERNEL_UPDATE) {  // x, v updated in place on copies
        OC_HIP(hipMemcpyAsync(c->alt_x, d.x, v3, hipMemcpyDeviceToDevice, c->stream));
        OC_HIP(hipMemcpyAsync(c->alt_v, d.v, v3, hipMemcpyDeviceToDevice, c->stream));
        d.x = c->alt_x;
        d.v = c->alt_v;
    } else if (kernel == OC_KERNEL_ASSEMBLE_FUSED) {  // updated x, v go to the other set
        d.xo = c->alt_x;
        d.vo = c->alt_v;
    }
    auto launch = [&]() -> hipError_t {
        switch (kernel) {
            case OC_KERNEL_ASSEMBLE: return oc::launch_assemble(d, s, c->stream);
            case OC_KERNEL_ASSEMBLE_FUSED: return oc::launch_assemble(d, s, c->stream, true);
            case OC_KERNEL_CG_SPMV: return oc::launch_cg_spmv(d, s, 1, c->stream);
            case OC_KERNEL_CG_AXPY: return oc::launch_cg_axpy(d, s, 1, c->stream);
            default: return oc::launch_update(d, s, c->stream);
        }
    };
    OC_HIP(launch());  // warm-up
    hipEvent_t e0, e1;
    OC_HIP(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return fail(OC_ERR_DEVICE, "hipEventCreate failed");
    }
    bool ok = hipEventRecord(e0, c->stream) == hipSuccess;
    for (int r = 0; ok && r < reps; r++) ok = launch() == hipSuccess;
    float ms = 0.0f;
    ok = ok && hipEventRecord(e1, c->stream) == hipSuccess && hipEventSynchronize(e1) == hipSuccess &&
         hipEventElapsedTime(&ms, e0, e1) == hipSuccess;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    std::memcpy(c->ctl_host, saved, sizeof(saved));
    ok = ok && hipMemcpyAsync(c->dc.ctl, c->ctl_host, sizeof(int) * 8, hipMemcpyHostToDevice, c->stream) ==
                   hipSuccess &&
         hipStreamSynchronize(c->stream) == hipSuccess;
    c->assembled = false;
    if (!ok) return fail(OC_ERR_DEVICE, "timed launches failed");
    *avg_us = 1000.0f * ms / (float)reps;
    return OC_OK;
}

int oc_cloth_ellipsoid_collision(oc_cloth c, const oc_ellipsoid* e) {
    if (!c || !e) return fail(OC_ERR_INVALID_ARGUMENT, "NULL argument");
    OC_HIP(oc::launch_ellipsoid(c->dc.x, c->dc.v, c->dc.N, *e, 0, c->stream));
    return OC_OK;
}

int oc_cloth_provot(oc_cloth c, float slack) {
    if (!c) return fail(OC_ERR_INVALID_ARGUMENT, "cloth is NULL");
    if (!(slack > 0.0f)) return fail(OC_ERR_INVALID_ARGUMENT, "slack must be positive");
    OC_HIP(oc::launch_provot(c->dc, slack, c->stream));
    return OC_OK;
}

int oc_cloth_step_physics(oc_cloth c, int32_t nsteps, const oc_ellipsoid* e, float slack) {
    if (!c) return fail(OC_ERR_INVALID_ARGUMENT, "cloth is NULL");
    if (nsteps < 0) return fail(OC_ERR_INVALID_ARGUMENT, "nsteps < 0");
    const oc::StepScalars s = scalars(c, c->prm.dt);
    for (int k = 0; k < nsteps; k++) {
        OC_HIP(oc::launch_assemble(c->dc, s, c->stream));
        OC_HIP(oc::launch_cg(c->dc, s, c->stream, false));
        OC_HIP(oc::launch_update(c->dc, s, c->stream, true));
        if (e) OC_HIP(oc::launch_ellipsoid(c->dc.x, c->dc.v, c->dc.N, *e, 0, c->stream));
        if (slack > 0.0f) OC_HIP(oc::launch_provot(c->dc, slack, c->stream));
    }
    c->assembled = false;
    return OC_OK;
}

int oc_cloth_synchronize(oc_cloth c) {
    if (!c) return fail(OC_ERR_INVALID_ARGUMENT, "cloth is NULL");
    OC_HIP(hipStreamSynchronize(c->stream));
    return OC_OK;
}

int oc_cloth_last_status(oc_cloth c, int32_t* iters, int32_t* status) {
    if (!c) return fail(OC_ERR_INVALID_ARGUMENT, "cloth is NULL");
    OC_HIP(hipMemcpyAsync(c->ctl_host, c->dc.ctl, sizeof(int) * 8, hipMemcpyDeviceToHost,
                          c->stream));
    OC_HIP(hipStreamSynchronize(c->stream));
    if (iters) *iters = c->ctl_host[1];
    if (status) *status = c->ctl_host[2];
    if (c->ctl_host[4]) return fail(OC_ERR_DEVICE, "persistent CG grid barrier timed out");
    return OC_OK;
}

int oc_cloth_num_pairs(oc_cloth c, int32_t* n) {
    if (!c || !n) return fail(OC_ERR_INVALID_ARGUMENT, "NULL argument");
    *n = c->topo.P;
    return OC_OK;
}

int oc_cloth_get_pairs(oc_cloth c, int32_t* a, int32_t* b) {
    if (!c) return fail(OC_ERR_INVALID_ARGUMENT, "cloth is NULL");
    if (a) std::memcpy(a, c->topo.pair_a.data(), sizeof(int32_t) * c->topo.P);
    if (b) std::memcpy(b, c->topo.pair_b.data(), sizeof(int32_t) * c->topo.P);
    return OC_OK;
}

int oc_cloth_get_system(oc_cloth c, float* blocks, float* rhs) {
    if (!c) return fail(OC_ERR_INVALID_ARGUMENT, "cloth is NULL");
    const oc::GridTopology& t = c->topo;
    const size_t N = t.N;
    if (blocks) {
        std::vector<float> planes(6 * N * std::max(t.n_planes, 1));
        OC_HIP(hipMemcpyAsync(planes.data(), c->dc.blocks, planes.size() * sizeof(float),
                              hipMemcpyDeviceToHost, c->stream));
        OC_HIP(hipStreamSynchronize(c->stream));
        for (int p = 0; p < t.P; p++)
            for (int q = 0; q < 6; q++)
                blocks[6 * (size_t)p + q] =
                    planes[((size_t)t.pair_plane[p] * 6 + q) * N + (size_t)t.pair_a[p]];
    }
    if (rhs) return download_vec(c, c->dc.b, rhs);
    return OC_OK;
}

int oc_cloth_set_spring_stiffness(oc_cloth c, const float* ks) {
    if (!c || !ks) return fail(OC_ERR_INVALID_ARGUMENT, "NULL argument");
    OC_HIP(hipMemcpyAsync((void*)c->dc.ks, ks, sizeof(float) * c->topo.S, hipMemcpyHostToDevice,
                          c->stream));
    if (c->dc.stencil_asm) OC_HIP(oc::launch_scatter_cst(c->dc, c->spring_cst, c->topo.S, 1, 1, c->stream));
    OC_HIP(hipStreamSynchronize(c->stream));
    c->roll_valid = false;
    return OC_OK;
}

int oc_cloth_get_spring_stiffness(oc_cloth c, float* ks) {
    if (!c || !ks) return fail(OC_ERR_INVALID_ARGUMENT, "NULL argument");
    OC_HIP(hipMemcpyAsync(ks, c->dc.ks, sizeof(float) * c->topo.S, hipMemcpyDeviceToHost,
                          c->stream));
    OC_HIP(hipStreamSynchronize(c->stream));
    return OC_OK;
}

}  // extern "C"

namespace {

// Views into the rollout arena: checkpoints [(nseg+1) x (x,v,dv)], one
// segment of pre-step states and solutions [every x (x,v,dv)], 11 adjoint
// vectors and the per-spring gradient.
struct RolloutViews {
    float* ckpt;
    float* seg;
    float* vec[11];
    float* gks;
};

RolloutViews rollout_views(oc_cloth_s* c, int nseg, int every) {
    const size_t v3 = 3 * (size_t)c->dc.N;
    RolloutViews r;
    float* p = (float*)c->adj_arena;
    r.ckpt = p;
    p += (size_t)(nseg + 1) * 3 * v3;
    r.seg = p;
    p += (size_t)every * 3 * v3;
    for (int k = 0; k < 11; k++) {
        r.vec[k] = p;
        p += v3;
    }
    r.gks = p;
    return r;
}

int save_state(oc_cloth_s* c, float* dst) {
    const size_t b = sizeof(float) * 3 * (size_t)c->dc.N;
    OC_HIP(hipMemcpyAsync(dst, c->dc.x, b, hipMemcpyDeviceToDevice, c->stream));
    OC_HIP(hipMemcpyAsync((char*)dst + b, c->dc.v, b, hipMemcpyDeviceToDevice, c->stream));
    OC_HIP(hipMemcpyAsync((char*)dst + 2 * b, c->dc.dv, b, hipMemcpyDeviceToDevice, c->stream));
    return OC_OK;
}

int load_state(oc_cloth_s* c, const float* src) {
    const size_t b = sizeof(float) * 3 * (size_t)c->dc.N;
    OC_HIP(hipMemcpyAsync(c->dc.x, src, b, hipMemcpyDeviceToDevice, c->stream));
    OC_HIP(hipMemcpyAsync(c->dc.v, (const char*)src + b, b, hipMemcpyDeviceToDevice, c->stream));
    OC_HIP(hipMemcpyAsync(c->dc.dv, (const char*)src + 2 * b, b, hipMemcpyDeviceToDevice, c->stream));
    return OC_OK;
}

int one_step(oc_cloth_s* c, const oc::StepScalars& s) {
    OC_HIP(oc::launch_assemble(c->dc, s, c->stream));
    OC_HIP(oc::launch_cg(c->dc, s, c->stream, false));
    OC_HIP(oc::launch_update(c->dc, s, c->stream, true));
    return OC_OK;
}

}  // namespace

extern "C" {

int oc_cloth_rollout_forward(oc_cloth c, int32_t nsteps, int32_t every) {
    if (!c) return fail(OC_ERR_INVALID_ARGUMENT, "cloth is NULL");
    if (nsteps < 1 || every < 1) return fail(OC_ERR_INVALID_ARGUMENT, "nsteps and checkpoint_every must be >= 1");
    c->roll_valid = false;  // until this rollout has completed
    every = std::min(every, nsteps);
    const int nseg = (nsteps + every - 1) / every;
    const size_t v3 = 3 * (size_t)c->dc.N;
    const size_t need = sizeof(float) * ((size_t)(nseg + 1) * 3 * v3 + (size_t)every * 3 * v3 +
                                         11 * v3 + (size_t)std::max(c->topo.S, 1));
    if (need > c->adj_bytes) {
        if (c->adj_arena) {
            OC_HIP(hipStreamSynchronize(c->stream));
            OC_HIP(hipFree(c->adj_arena));
            c->adj_arena = nullptr;
            c->adj_bytes = 0;
        }
        if (hipMalloc(&c->adj_arena, need) != hipSuccess)
            return fail(OC_ERR_OUT_OF_MEMORY, "hipMalloc of rollout checkpoints failed");
        c->adj_bytes = need;
    }
    RolloutViews rv = rollout_views(c, nseg, every);
    const oc::StepScalars s = scalars(c, c->prm.dt);
    OC_HIP(hipMemsetAsync(c->dc.dv, 0, sizeof(float) * v3, c->stream));
    int rc;
    for (int n = 0; n < nsteps; n++) {
        if (n % every == 0 && (rc = save_state(c, rv.ckpt + (size_t)(n / every) * 3 * v3)) != OC_OK)
            return rc;
        if ((rc = one_step(c, s)) != OC_OK) return rc;
    }
    if ((rc = save_state(c, rv.ckpt + (size_t)nseg * 3 * v3)) != OC_OK) return rc;
    c->roll_steps = nsteps;
    c->roll_every = every;
    c->roll_valid = true;
    c->assembled = false;
    return OC_OK;
}

int oc_cloth_rollout_backward(oc_cloth c, const float* grad_x, const float* grad_v, float* grad_ks,
                              float* grad_x0, float* grad_v0) {
    if (!c) return fail(OC_ERR_INVALID_ARGUMENT, "cloth is NULL");
    if (!grad_x) return fail(OC_ERR_INVALID_ARGUMENT, "grad_x is NULL");
    if (!c->roll_valid) return fail(OC_ERR_INVALID_ARGUMENT, "no rollout_forward to differentiate");
    const int nsteps = c->roll_steps, every = c->roll_every;
    const int nseg = (nsteps + every - 1) / every;
    const size_t v3 = 3 * (size_t)c->dc.N;
    RolloutViews rv = rollout_views(c, nseg, every);
    float* mu = rv.vec[0];
    float* gx[2] = {rv.vec[7], rv.vec[8]};
    float* gv[2] = {rv.vec[9], rv.vec[10]};
    int rc;
    if ((rc = upload_vec(c, gx[0], grad_x)) != OC_OK) return rc;
    if ((rc = upload_vec(c, gv[0], grad_v)) != OC_OK) return rc;
    OC_HIP(hipMemsetAsync(rv.gks, 0, sizeof(float) * std::max(c->topo.S, 1), c->stream));
    const oc::StepScalars s = scalars(c, c->prm.dt);
    oc::StepScalars s_cold = s;
    s_cold.warm = 0;
    int cur = 0;
    for (int seg = nseg - 1; seg >= 0; seg--) {
        const int n0 = seg * every, n1 = std::min(n0 + every, nsteps);
        if ((rc = load_state(c, rv.ckpt + (size_t)seg * 3 * v3)) != OC_OK) return rc;
        for (int n = n0; n < n1; n++) {
            float* slot = rv.seg + (size_t)(n - n0) * 3 * v3;
            const size_t b = sizeof(float) * v3;
            OC_HIP(hipMemcpyAsync(slot, c->dc.x, b, hipMemcpyDeviceToDevice, c->stream));
            OC_HIP(hipMemcpyAsync(slot + v3, c->dc.v, b, hipMemcpyDeviceToDevice, c->stream));
            if ((rc = one_step(c, s)) != OC_OK) return rc;
            OC_HIP(hipMemcpyAsync(slot + 2 * v3, c->dc.dv, b, hipMemcpyDeviceToDevice, c->stream));
        }
        for (int n = n1 - 1; n >= n0; n--) {
            float* slot = rv.seg + (size_t)(n - n0) * 3 * v3;
            oc::DevCloth ca = c->dc;  // adjoint solve view: own CG vectors
            ca.x = slot;
            ca.v = slot + v3;
            ca.dv = mu;
            ca.r = rv.vec[1];
            ca.d0 = rv.vec[2];
            ca.d1 = rv.vec[3];
            ca.q = rv.vec[4];
            ca.b = rv.vec[5];
            oc::AdjointBufs ab;
            ab.gx_out = gx[cur];
            ab.gv_out = gv[cur];
            ab.gvt = rv.vec[6];
            ab.mu = mu;
            ab.gx_in = gx[1 - cur];
            ab.gv_in = gv[1 - cur];
            ab.gks = rv.gks;
            OC_HIP(oc::launch_assemble(ca, s_cold, c->stream));
            OC_HIP(oc::launch_adj_init(ca, s, ab, c->stream));
            OC_HIP(oc::launch_cg(ca, s, c->stream));
            oc::DevCloth cg = c->dc;  // pre-step state and the step's solution
            cg.x = slot;
            cg.v = slot + v3;
            cg.dv = slot + 2 * v3;
            OC_HIP(oc::launch_adj_gather(cg, s, ab, c->stream));
            cur = 1 - cur;
        }
    }
    if ((rc = load_state(c, rv.ckpt + (size_t)nseg * 3 * v3)) != OC_OK) return rc;
    if (grad_ks) {
        OC_HIP(hipMemcpyAsync(grad_ks, rv.gks, sizeof(float) * c->topo.S, hipMemcpyDeviceToHost,
                              c->stream));
        OC_HIP(hipStreamSynchronize(c->stream));
    }
    if (grad_x0 && (rc = download_vec(c, gx[cur], grad_x0)) != OC_OK) return rc;
    if (grad_v0 && (rc = download_vec(c, gv[cur], grad_v0)) != OC_OK) return rc;
    OC_HIP(hipStreamSynchronize(c->stream));
    c->assembled = false;
    return OC_OK;
}

}  // extern "C"
