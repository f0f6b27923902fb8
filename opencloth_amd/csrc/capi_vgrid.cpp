// C ABI of the grid Verlet (oc_vgrid_*): the reference's GPU Verlet path,
// whose host glue (OpenCloth_Verlet_CUDA_GLSL_OPENCL_CPU/OpenCloth_Verlet_CUDA/
// verlet.cu, "CUH") allocates two ping-pong pairs of float4 position
// buffers, uploads the host arrays once and swaps the pairs every step.
// Here the device state is SoA planes (x, y, z) per buffer plus one w plane
// per buffer kind, which never changes (the kernel carries w through,
// CUK:193-195).
#include <hip/hip_runtime.h>

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

}  // namespace

struct oc_vgrid_s {
    int nx = 0, ny = 0, N = 0;
    oc_vgrid_params prm{};
    hipStream_t stream = nullptr;
    void* arena = nullptr;
    float* x[2] = {nullptr, nullptr};   // position sets, ping-ponged
    float* xl[2] = {nullptr, nullptr};  // previous positions, ping-ponged
    float* wx = nullptr;                // w of the position buffer
    float* wxl = nullptr;               // w of the previous-position buffer
    uint8_t* pin = nullptr;
    int cur = 0;                        // the set the next step reads
};

namespace {

int check_params(const oc_vgrid_params* p) {
    if (!p) return set_error(OC_ERR_INVALID_ARGUMENT, "params is NULL");
    if (!(p->mass > 0.0f) || !std::isfinite(p->mass) || !(p->dt > 0.0f) || !std::isfinite(p->dt))
        return set_error(OC_ERR_INVALID_ARGUMENT, "mass and dt must be positive");
    return OC_OK;
}

// Rest length of stencil position k, |(dx*sx, dy*sy)| in FP32 as the kernel
// evaluates length(make_float2(coord.x*inv_cloth_size.x, ...)) (CUK:164).
oc::VgridScalars scalars(const oc_vgrid_s* g) {
    static const int dx[12] = {1, 0, -1, 0, 1, -1, -1, 1, 2, 0, -2, 0};
    static const int dy[12] = {0, -1, 0, 1, -1, -1, 1, 1, 0, -2, 0, 2};
    const oc_vgrid_params& p = g->prm;
    oc::VgridScalars s;
    s.mass = p.mass;
    s.gx = p.gravity[0];
    s.gy = p.gravity[1];
    s.gz = p.gravity[2];
    s.damping = p.damping;
    s.dt = p.dt;
    for (int k = 0; k < 12; k++) {
        const float a = (float)dx[k] * p.spacing[0], b = (float)dy[k] * p.spacing[1];
        s.rest[k] = std::sqrt(a * a + b * b);
    }
    for (int c = 0; c < 3; c++) {
        s.ks[c] = p.ks[c];
        s.kd[c] = p.kd[c];
    }
    return s;
}

// Host AoS float4 -> xyz planes at dst, w into w (when given).
int upload4(oc_vgrid_s* g, float* dst, float* w, const float* aos) {
    const size_t N = g->N;
    std::vector<float> t(4 * N);
    for (size_t i = 0; i < N; i++)
        for (int k = 0; k < 4; k++) t[k * N + i] = aos[4 * i + k];
    OC_HIP(hipMemcpyAsync(dst, t.data(), 3 * N * sizeof(float), hipMemcpyHostToDevice, g->stream));
    OC_HIP(hipMemcpyAsync(w, t.data() + 3 * N, N * sizeof(float), hipMemcpyHostToDevice, g->stream));
    OC_HIP(hipStreamSynchronize(g->stream));
    return OC_OK;
}

int download4(oc_vgrid_s* g, const float* src, const float* w, float* aos) {
    const size_t N = g->N;
    std::vector<float> t(4 * N);
    OC_HIP(hipMemcpyAsync(t.data(), src, 3 * N * sizeof(float), hipMemcpyDeviceToHost, g->stream));
    OC_HIP(hipMemcpyAsync(t.data() + 3 * N, w, N * sizeof(float), hipMemcpyDeviceToHost, g->stream));
    OC_HIP(hipStreamSynchronize(g->stream));
    for (size_t i = 0; i < N; i++)
        for (int k = 0; k < 4; k++) aos[4 * i + k] = t[k * N + i];
    return OC_OK;
}

}  // namespace

extern "C" {

int oc_vgrid_create(int32_t width, int32_t height, const float* x0, const float* x_last0,
                    const uint8_t* pinned, const oc_vgrid_params* params, oc_vgrid* out) {
    if (!out) return set_error(OC_ERR_INVALID_ARGUMENT, "out is NULL");
    *out = nullptr;
    if (!x0) return set_error(OC_ERR_INVALID_ARGUMENT, "x0 is NULL");
    if (width < 1 || height < 1 || (int64_t)width * height > (1 << 28))
        return set_error(OC_ERR_INVALID_ARGUMENT, "grid size out of range");
    int rc = check_params(params);
    if (rc != OC_OK) return rc;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return set_error(OC_ERR_NO_DEVICE, "no HIP device");
    oc_vgrid_s* g = new (std::nothrow) oc_vgrid_s();
    if (!g) return set_error(OC_ERR_OUT_OF_MEMORY, "host allocation failed");
    g->nx = width;
    g->ny = height;
    g->N = width * height;
    g->prm = *params;
    const size_t N = g->N;
    const size_t vec = align_up(sizeof(float) * 3 * N), pl = align_up(sizeof(float) * N);
    const size_t total = 4 * vec + 2 * pl + align_up(N);
    if (hipMalloc(&g->arena, total) != hipSuccess) {
        delete g;
        return set_error(OC_ERR_OUT_OF_MEMORY, "hipMalloc of " + std::to_string(total) + " bytes failed");
    }
    char* p = (char*)g->arena;
    for (int k = 0; k < 2; k++) {
        g->x[k] = (float*)p;
        p += vec;
        g->xl[k] = (float*)p;
        p += vec;
    }
    g->wx = (float*)p;
    p += pl;
    g->wxl = (float*)p;
    p += pl;
    g->pin = (uint8_t*)p;
    std::vector<uint8_t> pins(N, 0);
    if (pinned) {
        for (size_t i = 0; i < N; i++) pins[i] = pinned[i] ? 1 : 0;
    } else {  // index 0 and texsize.x - 1 (CUK:126)
        pins[0] = 1;
        pins[width - 1] = 1;
    }
    *out = g;
    if (hipMemset(g->arena, 0, total) != hipSuccess ||
        hipMemcpy(g->pin, pins.data(), N, hipMemcpyHostToDevice) != hipSuccess) {
        oc_vgrid_destroy(g);
        *out = nullptr;
        return set_error(OC_ERR_DEVICE, "upload failed");
    }
    rc = oc_vgrid_set_state(g, x0, x_last0 ? x_last0 : x0);
    if (rc != OC_OK) {
        std::string e = oc_last_error();
        oc_vgrid_destroy(g);
        *out = nullptr;
        return set_error(rc, e);
    }
    return OC_OK;
}

int oc_vgrid_destroy(oc_vgrid g) {
    if (!g) return OC_OK;
    (void)hipStreamSynchronize(g->stream);
    if (g->arena) (void)hipFree(g->arena);
    delete g;
    return OC_OK;
}

int oc_vgrid_set_stream(oc_vgrid g, void* stream) {
    if (!g) return set_error(OC_ERR_INVALID_ARGUMENT, "grid is NULL");
    g->stream = (hipStream_t)stream;
    return OC_OK;
}

int oc_vgrid_set_params(oc_vgrid g, const oc_vgrid_params* p) {
    if (!g) return set_error(OC_ERR_INVALID_ARGUMENT, "grid is NULL");
    int rc = check_params(p);
    if (rc != OC_OK) return rc;
    g->prm = *p;
    return OC_OK;
}

int oc_vgrid_set_state(oc_vgrid g, const float* x, const float* x_last) {
    if (!g) return set_error(OC_ERR_INVALID_ARGUMENT, "grid is NULL");
    int rc;
    if (x && (rc = upload4(g, g->x[g->cur], g->wx, x)) != OC_OK) return rc;
    if (x_last && (rc = upload4(g, g->xl[g->cur], g->wxl, x_last)) != OC_OK) return rc;
    return OC_OK;
}

int oc_vgrid_get_state(oc_vgrid g, float* x, float* x_last) {
    if (!g) return set_error(OC_ERR_INVALID_ARGUMENT, "grid is NULL");
    int rc;
    if (x && (rc = download4(g, g->x[g->cur], g->wx, x)) != OC_OK) return rc;
    if (x_last && (rc = download4(g, g->xl[g->cur], g->wxl, x_last)) != OC_OK) return rc;
    return OC_OK;
}

int oc_vgrid_step(oc_vgrid g, int32_t nsteps, float* vbo) {
    if (!g) return set_error(OC_ERR_INVALID_ARGUMENT, "grid is NULL");
    if (nsteps < 0) return set_error(OC_ERR_INVALID_ARGUMENT, "nsteps < 0");
    if (vbo && (reinterpret_cast<uintptr_t>(vbo) & 15u))
        return set_error(OC_ERR_INVALID_ARGUMENT, "vbo must be 16-byte aligned (float4)");
    const oc::VgridScalars s = scalars(g);
    for (int k = 0; k < nsteps; k++) {
        oc::DevVgrid d;
        d.nx = g->nx;
        d.ny = g->ny;
        d.N = g->N;
        d.x = g->x[g->cur];
        d.xl = g->xl[g->cur];
        d.xo = g->x[1 - g->cur];
        d.xlo = g->xl[1 - g->cur];
        d.pin = g->pin;
        d.w = g->wx;
        d.vbo = (k + 1 == nsteps) ? vbo : nullptr;  // the last step's positions
        OC_HIP(oc::launch_vgrid(d, s, g->stream));
        g->cur = 1 - g->cur;
    }
    return OC_OK;
}

int oc_vgrid_synchronize(oc_vgrid g) {
    if (!g) return set_error(OC_ERR_INVALID_ARGUMENT, "grid is NULL");
    OC_HIP(hipStreamSynchronize(g->stream));
    return OC_OK;
}

}  // extern "C"
