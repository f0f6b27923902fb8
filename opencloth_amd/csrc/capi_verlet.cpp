// C ABI of the position-Verlet cloth (oc_verlet_*): the Verlet demo's step
// (OpenCloth_Verlet/OpenCloth_Verlet/main.cpp) on the same grid topology
// tables as the implicit step.  Positions rotate through three device
// buffers: current, previous and next.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "capi_common.hpp"
#include "oc_internal.hpp"

namespace {

using oc::align_up;
using oc::set_error;

}  // namespace

struct oc_verlet_s {
    oc::GridTopology topo;
    oc_verlet_params prm{};
    hipStream_t stream = nullptr;
    void* arena = nullptr;
    float* pos[3] = {nullptr, nullptr, nullptr};  // rotated: cur, last, next
    int cur = 0, last = 1;
    float* f = nullptr;
    uint8_t* pin = nullptr;
    int32_t* ainc_s = nullptr;
    uint16_t* ainc_m = nullptr;
    float* L0 = nullptr;
    float* ks = nullptr;
    float* kd = nullptr;
    uint32_t* asm_mask = nullptr;  // stencil assembly form (else NULL)
    float* cst = nullptr;
};

namespace {

oc::DevVerlet view(const oc_verlet_s* v) {
    oc::DevVerlet d;
    d.nx = v->topo.nx;
    d.ny = v->topo.ny;
    d.N = v->topo.N;
    d.wa = v->topo.wa;
    d.x = v->pos[v->cur];
    d.xl = v->pos[v->last];
    d.xn = v->pos[3 - v->cur - v->last];
    d.f = v->f;
    d.pin = v->pin;
    d.ainc_s = v->ainc_s;
    d.ainc_m = v->ainc_m;
    d.L0 = v->L0;
    d.ks = v->ks;
    d.kd = v->kd;
    d.asm_mask = v->asm_mask;
    d.cst = v->cst;
    d.stencil_asm = v->asm_mask ? 1 : 0;
    return d;
}

oc::VerletScalars vscalars(const oc_verlet_s* v) {
    oc::VerletScalars s;
    s.mass = v->prm.mass;
    s.gx = v->prm.gravity[0];
    s.gy = v->prm.gravity[1];
    s.gz = v->prm.gravity[2];
    s.damping = v->prm.damping;
    s.dt = v->prm.dt;
    s.flags = v->prm.flags;
    return s;
}

int upload_aos(oc_verlet_s* v, float* dst, const float* aos) {
    const size_t N = v->topo.N;
    std::vector<float> t(3 * N);
    for (size_t i = 0; i < N; i++)
        for (int k = 0; k < 3; k++) t[k * N + i] = aos[3 * i + k];
    OC_HIP(hipMemcpyAsync(dst, t.data(), t.size() * 4, hipMemcpyHostToDevice, v->stream));
    OC_HIP(hipStreamSynchronize(v->stream));
    return OC_OK;
}

int download_aos(oc_verlet_s* v, const float* src, float* aos) {
    const size_t N = v->topo.N;
    std::vector<float> t(3 * N);
    OC_HIP(hipMemcpyAsync(t.data(), src, t.size() * 4, hipMemcpyDeviceToHost, v->stream));
    OC_HIP(hipStreamSynchronize(v->stream));
    for (size_t i = 0; i < N; i++)
        for (int k = 0; k < 3; k++) aos[3 * i + k] = t[k * N + i];
    return OC_OK;
}

}  // namespace

extern "C" {

int oc_verlet_create(int32_t nx, int32_t ny, const float* x0, const float* x_last0,
                     const oc_spring* springs, int32_t n_springs, const uint8_t* pinned,
                     const oc_verlet_params* params, oc_verlet* out) {
    if (!out) return set_error(OC_ERR_INVALID_ARGUMENT, "out is NULL");
    *out = nullptr;
    if (!x0 || !params) return set_error(OC_ERR_INVALID_ARGUMENT, "NULL argument");
    if (!(params->mass > 0.0f) || !(params->dt > 0.0f))
        return set_error(OC_ERR_INVALID_ARGUMENT, "mass and dt must be positive");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return set_error(OC_ERR_NO_DEVICE, "no HIP device");
    oc_verlet_s* v = new (std::nothrow) oc_verlet_s();
    if (!v) return set_error(OC_ERR_OUT_OF_MEMORY, "host allocation failed");
    std::string err;
    int rc = oc::build_grid_topology(nx, ny, springs, n_springs, &v->topo, &err);
    if (rc != OC_OK) {
        delete v;
        return set_error(rc, err);
    }
    v->prm = *params;
    const oc::GridTopology& t = v->topo;
    const size_t N = t.N, S = std::max(t.S, 1);
    const size_t vec = align_up(sizeof(float) * 3 * N);
    const size_t as = align_up(sizeof(int32_t) * std::max<size_t>(t.ainc_spring.size(), 1));
    const size_t am = align_up(sizeof(uint16_t) * std::max<size_t>(t.ainc_meta.size(), 1));
    const size_t sp = align_up(sizeof(float) * S);
    // stencil assembly form: springs from a per-vertex mask, constants from
    // owner planes (OPENCLOTH_ASSEMBLY=per-end keeps the slot tables)
    const char* asm_env = std::getenv("OPENCLOTH_ASSEMBLY");
    const bool st_asm = t.stencil_asm && !(asm_env && std::strcmp(asm_env, "per-end") == 0);
    const size_t amask = st_asm ? align_up(sizeof(uint32_t) * N) : 0;
    const size_t cst = st_asm ? align_up(sizeof(float) * oc::kCstPlanes * N) : 0;
    const size_t total = 4 * vec + align_up(N) + as + am + 3 * sp + amask + cst;
    if (hipMalloc(&v->arena, total) != hipSuccess) {
        delete v;
        return set_error(OC_ERR_OUT_OF_MEMORY, "hipMalloc failed");
    }
    char* p = (char*)v->arena;
    for (int k = 0; k < 3; k++) {
        v->pos[k] = (float*)p;
        p += vec;
    }
    v->f = (float*)p;
    p += vec;
    v->pin = (uint8_t*)p;
    p += align_up(N);
    v->ainc_s = (int32_t*)p;
    p += as;
    v->ainc_m = (uint16_t*)p;
    p += am;
    v->L0 = (float*)p;
    p += sp;
    v->ks = (float*)p;
    p += sp;
    v->kd = (float*)p;
    p += sp;
    if (st_asm) {
        v->asm_mask = (uint32_t*)p;
        p += amask;
        v->cst = (float*)p;
    }
    std::vector<float> L0(S, 0.0f), ks(S, 0.0f), kd(S, 0.0f);
    for (int s = 0; s < t.S; s++) {
        L0[s] = springs[s].rest_length;
        ks[s] = springs[s].Ks;
        kd[s] = springs[s].Kd;
    }
    std::vector<uint8_t> pins(N, 0);
    if (pinned)
        for (size_t i = 0; i < N; i++) pins[i] = pinned[i] ? 1 : 0;
    bool ok = hipMemset(v->arena, 0, total) == hipSuccess &&
              hipMemcpy(v->ainc_s, t.ainc_spring.data(), t.ainc_spring.size() * 4, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(v->ainc_m, t.ainc_meta.data(), t.ainc_meta.size() * 2, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(v->L0, L0.data(), S * 4, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(v->ks, ks.data(), S * 4, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(v->kd, kd.data(), S * 4, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(v->pin, pins.data(), N, hipMemcpyHostToDevice) == hipSuccess;
    if (ok && st_asm) {
        std::vector<float> hc((size_t)oc::kCstPlanes * N, 0.0f);
        for (int s = 0; s < t.S; s++) {
            const size_t base = (size_t)t.spring_cst[s];
            hc[base] = L0[s];
            hc[base + N] = ks[s];
            hc[base + 2 * N] = kd[s];
        }
        ok = hipMemcpy(v->asm_mask, t.asm_mask.data(), N * 4, hipMemcpyHostToDevice) == hipSuccess &&
             hipMemcpy(v->cst, hc.data(), hc.size() * 4, hipMemcpyHostToDevice) == hipSuccess;
    }
    if (!ok) {
        (void)hipFree(v->arena);
        delete v;
        return set_error(OC_ERR_DEVICE, "upload failed");
    }
    *out = v;
    rc = oc_verlet_set_state(v, x0, x_last0 ? x_last0 : x0);
    if (rc != OC_OK) {
        std::string e = oc_last_error();
        oc_verlet_destroy(v);
        *out = nullptr;
        return set_error(rc, e);
    }
    return OC_OK;
}

int oc_verlet_destroy(oc_verlet v) {
    if (!v) return OC_OK;
    (void)hipStreamSynchronize(v->stream);
    if (v->arena) (void)hipFree(v->arena);
    delete v;
    return OC_OK;
}

int oc_verlet_set_stream(oc_verlet v, void* stream) {
    if (!v) return set_error(OC_ERR_INVALID_ARGUMENT, "verlet is NULL");
    v->stream = (hipStream_t)stream;
    return OC_OK;
}

int oc_verlet_set_state(oc_verlet v, const float* x, const float* x_last) {
    if (!v) return set_error(OC_ERR_INVALID_ARGUMENT, "verlet is NULL");
    int rc;
    if (x && (rc = upload_aos(v, v->pos[v->cur], x)) != OC_OK) return rc;
    if (x_last && (rc = upload_aos(v, v->pos[v->last], x_last)) != OC_OK) return rc;
    return OC_OK;
}

int oc_verlet_get_state(oc_verlet v, float* x, float* x_last) {
    if (!v) return set_error(OC_ERR_INVALID_ARGUMENT, "verlet is NULL");
    int rc;
    if (x && (rc = download_aos(v, v->pos[v->cur], x)) != OC_OK) return rc;
    if (x_last && (rc = download_aos(v, v->pos[v->last], x_last)) != OC_OK) return rc;
    return OC_OK;
}

int oc_verlet_compute_forces(oc_verlet v, float* f_out) {
    if (!v) return set_error(OC_ERR_INVALID_ARGUMENT, "verlet is NULL");
    OC_HIP(oc::launch_verlet_forces(view(v), vscalars(v), v->stream));
    if (f_out) return download_aos(v, v->f, f_out);
    return OC_OK;
}

int oc_verlet_step(oc_verlet v, int32_t nsteps, const oc_ellipsoid* e) {
    if (!v) return set_error(OC_ERR_INVALID_ARGUMENT, "verlet is NULL");
    if (nsteps < 0) return set_error(OC_ERR_INVALID_ARGUMENT, "nsteps < 0");
    const oc::VerletScalars s = vscalars(v);
    for (int k = 0; k < nsteps; k++) {
        oc::DevVerlet d = view(v);
        OC_HIP(oc::launch_verlet_step(d, s, v->stream));
        const int next = 3 - v->cur - v->last;
        v->last = v->cur;  // x_last' = x (VL:438)
        v->cur = next;
        if (e) OC_HIP(oc::launch_ellipsoid(v->pos[v->cur], v->pos[v->last], v->topo.N, *e, 1, v->stream));
    }
    return OC_OK;
}

}  // extern "C"
