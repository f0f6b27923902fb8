// C ABI of the batched cloths (oc_batch_*): per-cloth topology, packing of
// cloths into work items, the captured one-step HIP graph and the schedule
// autotuner.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <new>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "capi_common.hpp"
#include "oc_internal.hpp"

namespace {

using oc::align_up;
using oc::set_error;

}  // namespace

struct oc_batch_s {
    int n = 0;                 // cloths
    int Ntot = 0, Stot = 0;
    int n_planes = 0;
    std::vector<int> nx, ny, voff, nverts;
    oc_params prm{};
    oc_batch_schedule sched{512, 8};
    hipStream_t stream = nullptr;
    hipStream_t capture_stream = nullptr;
    hipGraphExec_t graph = nullptr;
    bool graph_dirty = true;
    void* arena = nullptr;
    void* sched_buf = nullptr;  // items, order (rebuilt per schedule)
    int* host_T = nullptr;
    std::vector<int> T_of;      // threads per cloth under the schedule
    int grid = 0;
    oc::DevBatch dc{};
    // differentiable rollouts
    void* adj_arena = nullptr;
    size_t adj_bytes = 0;
    int roll_steps = 0, roll_every = 0;
    bool roll_valid = false;
    // spring-once batched assembly (OPENCLOTH_ASSEMBLY=once) where the
    // schedule's LDS fits; off by default: on the 4096-cloth workload it ran
    // 9.30 ms per step against 8.45 ms for the per-end stencil form
    bool allow_once = false;
};

namespace {

oc::StepScalars batch_scalars(const oc_batch_s* b) {
    oc::StepScalars s;
    s.mass = b->prm.mass;
    s.gx = b->prm.gravity[0];
    s.gy = b->prm.gravity[1];
    s.gz = b->prm.gravity[2];
    s.h = b->prm.dt;
    s.hh = b->prm.dt * b->prm.dt;
    s.tol2 = b->prm.cg_tol * b->prm.cg_tol;
    s.max_it = b->prm.cg_max_iters;
    s.tol_mode = b->prm.cg_tol_mode;
    s.warm = b->prm.warm_start ? 1 : 0;
    s.flags = b->prm.flags;
    s.damp = b->prm.damping;
    s.md = s.mass - s.h * s.damp;
    return s;
}

int check_params(const oc_params* p) {
    if (!p) return set_error(OC_ERR_INVALID_ARGUMENT, "params is NULL");
    if (!(p->mass > 0.0f) || !(p->dt > 0.0f) || p->cg_max_iters < 0 || p->cg_tol_mode < 0 ||
        p->cg_tol_mode > 2 || !(p->cg_tol >= 0.0f) || !std::isfinite(p->damping))
        return set_error(OC_ERR_INVALID_ARGUMENT, "invalid step parameters");
    return OC_OK;
}

int pow2ceil(int x) {
    int p = 1;
    while (p < x) p <<= 1;
    return p;
}

// Pack cloths into work items for the schedule and upload the tables.
int apply_schedule(oc_batch_s* b) {
    const int NT = b->sched.block_threads;
    if (NT != 256 && NT != 512 && NT != 1024)
        return set_error(OC_ERR_INVALID_ARGUMENT, "block_threads must be 256, 512 or 1024");
    if (b->sched.vertices_per_thread < 1)
        return set_error(OC_ERR_INVALID_ARGUMENT, "vertices_per_thread must be >= 1");
    std::vector<int> order(b->n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int x, int y) { return b->nverts[x] > b->nverts[y]; });
    std::vector<int> T_of(b->n, 64);
    for (int c = 0; c < b->n; c++) {
        const int want = (b->nverts[c] + b->sched.vertices_per_thread - 1) / b->sched.vertices_per_thread;
        T_of[c] = std::min(NT, std::max(64, pow2ceil(want)));
    }
    std::vector<int> it_inst, it_count, it_T;
    for (int k = 0; k < b->n;) {
        const int T = T_of[order[k]];
        const int G = NT / T;
        int cnt = 0;
        while (k + cnt < b->n && cnt < G && T_of[order[k + cnt]] == T) cnt++;
        it_inst.push_back(k);
        it_count.push_back(cnt);
        it_T.push_back(T);
        k += cnt;
    }
    const int n_items = (int)it_inst.size();
    // spring-once assembly: 12 floats per thread plus each cloth's rings
    // (sized for the item's widest cloth), when that fits the LDS budget
    int lds_floats = 12 * NT;
    for (int k = 0; k < n_items; k++) {
        int nxm = 1;
        for (int q = 0; q < it_count[k]; q++) nxm = std::max(nxm, b->nx[order[it_inst[k] + q]]);
        lds_floats = std::max(lds_floats, 12 * NT + it_count[k] * oc::ring_floats(nxm));
    }
    int dev_lds = 0, cur = 0;
    OC_HIP(hipGetDevice(&cur));
    OC_HIP(hipDeviceGetAttribute(&dev_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, cur));
    // the kernel's static LDS (a few hundred bytes) comes on top
    const int lds_cap = std::min(oc::kBatchLdsFloats, (dev_lds - 1024) / 4);
    const bool once = b->dc.stencil_asm && b->allow_once && lds_floats <= lds_cap;
    if (!once && oc::batch_resident_floats(NT) > lds_cap)
        return set_error(OC_ERR_DEVICE, "the device's LDS per workgroup cannot hold the CG's resident state (" +
                                            std::to_string(4 * oc::batch_resident_floats(NT)) + " bytes)");
    // persistent grid: as many workgroups as fit on the device at once
    int dev = 0, cus = 0;
    OC_HIP(hipGetDevice(&dev));
    OC_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const int grid = std::max(1, std::min(n_items, cus * std::max(1, 2048 / NT)));
    // the new tables go to a new buffer; the old one (which a captured step
    // may still name) is released only once the new one is in place
    const size_t bi = align_up(sizeof(int) * std::max(n_items, 1));
    const size_t bo = align_up(sizeof(int) * std::max(b->n, 1));
    void* fresh = nullptr;
    if (hipMalloc(&fresh, 3 * bi + bo) != hipSuccess)
        return set_error(OC_ERR_OUT_OF_MEMORY, "hipMalloc of the schedule failed");
    char* p = (char*)fresh;
    int* d_inst = (int*)p;
    int* d_count = (int*)(p + bi);
    int* d_T = (int*)(p + 2 * bi);
    int* d_order = (int*)(p + 3 * bi);
    const bool ok =
        hipMemcpy(d_inst, it_inst.data(), sizeof(int) * n_items, hipMemcpyHostToDevice) == hipSuccess &&
        hipMemcpy(d_count, it_count.data(), sizeof(int) * n_items, hipMemcpyHostToDevice) == hipSuccess &&
        hipMemcpy(d_T, it_T.data(), sizeof(int) * n_items, hipMemcpyHostToDevice) == hipSuccess &&
        hipMemcpy(d_order, order.data(), sizeof(int) * b->n, hipMemcpyHostToDevice) == hipSuccess &&
        (!b->sched_buf || hipStreamSynchronize(b->stream) == hipSuccess);
    if (!ok) {
        (void)hipFree(fresh);
        return set_error(OC_ERR_DEVICE, "upload of the schedule failed");
    }
    if (b->sched_buf) (void)hipFree(b->sched_buf);
    b->sched_buf = fresh;
    b->T_of.swap(T_of);
    b->dc.item_inst = d_inst;
    b->dc.item_count = d_count;
    b->dc.item_T = d_T;
    b->dc.order = d_order;
    b->dc.n_items = n_items;
    b->dc.asm_once = once ? 1 : 0;
    b->dc.lds_floats = once ? lds_floats : 0;
    b->grid = grid;
    b->graph_dirty = true;
    return OC_OK;
}

int launch_step(oc_batch_s* b, hipStream_t st) {
    OC_HIP(oc::launch_batch_reset(b->dc, st));
    OC_HIP(oc::launch_batch_step(b->dc, batch_scalars(b), b->sched.block_threads, b->grid, st));
    return OC_OK;
}

int ensure_graph(oc_batch_s* b) {
    if (!b->graph_dirty && b->graph) return OC_OK;
    if (b->graph) {
        OC_HIP(hipGraphExecDestroy(b->graph));
        b->graph = nullptr;
    }
    if (!b->capture_stream) OC_HIP(hipStreamCreateWithFlags(&b->capture_stream, hipStreamNonBlocking));
    hipGraph_t g = nullptr;
    OC_HIP(hipStreamBeginCapture(b->capture_stream, hipStreamCaptureModeThreadLocal));
    int rc = launch_step(b, b->capture_stream);
    hipError_t e = hipStreamEndCapture(b->capture_stream, &g);
    if (rc != OC_OK) return rc;
    if (e != hipSuccess) return set_error(OC_ERR_DEVICE, "hipStreamEndCapture failed");
    e = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return set_error(OC_ERR_DEVICE, "hipGraphInstantiate failed");
    b->graph_dirty = false;
    return OC_OK;
}

struct ClothTables {
    oc::GridTopology topo;
    int rc = OC_OK;
    std::string err;
};

}  // namespace

extern "C" {

int oc_batch_create(int32_t n_cloths, const int32_t* nx, const int32_t* ny, const float* x0,
                    const float* v0, const oc_spring* springs, const int32_t* n_springs,
                    const uint8_t* pinned, const oc_params* params, oc_batch* out) {
    if (!out) return set_error(OC_ERR_INVALID_ARGUMENT, "out is NULL");
    *out = nullptr;
    if (n_cloths < 1 || !nx || !ny || !x0 || !n_springs || !springs)
        return set_error(OC_ERR_INVALID_ARGUMENT, "invalid batch description");
    int rc = check_params(params);
    if (rc != OC_OK) return rc;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return set_error(OC_ERR_NO_DEVICE, "no HIP device");

    oc_batch_s* b = new (std::nothrow) oc_batch_s();
    if (!b) return set_error(OC_ERR_OUT_OF_MEMORY, "host allocation failed");
    b->n = n_cloths;
    b->prm = *params;
    b->nx.assign(nx, nx + n_cloths);
    b->ny.assign(ny, ny + n_cloths);
    b->voff.resize(n_cloths);
    b->nverts.resize(n_cloths);
    std::vector<int64_t> soff(n_cloths + 1, 0);
    int64_t Nt = 0;
    for (int c = 0; c < n_cloths; c++) {
        if (nx[c] < 1 || ny[c] < 1 || n_springs[c] < 0 || (int64_t)nx[c] * ny[c] > (1 << 28)) {
            delete b;
            return set_error(OC_ERR_INVALID_ARGUMENT, "invalid cloth size");
        }
        b->voff[c] = (int)Nt;
        b->nverts[c] = nx[c] * ny[c];
        Nt += (int64_t)nx[c] * ny[c];
        soff[c + 1] = soff[c] + n_springs[c];
    }
    if (Nt > (1 << 28) || soff[n_cloths] > (int64_t)INT32_MAX) {
        delete b;
        return set_error(OC_ERR_INVALID_ARGUMENT, "batch too large");
    }
    b->Ntot = (int)Nt;
    b->Stot = (int)soff[n_cloths];

    // per-cloth topology, built in parallel
    std::vector<ClothTables> tabs(n_cloths);
    {
        std::atomic<int> next(0);
        const int nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        std::vector<std::thread> th;
        for (int t = 0; t < nth; t++)
            th.emplace_back([&]() {
                for (int c; (c = next.fetch_add(1)) < n_cloths;) {
                    tabs[c].rc = oc::build_grid_topology(nx[c], ny[c], springs + soff[c],
                                                         n_springs[c], &tabs[c].topo, &tabs[c].err);
                }
            });
        for (auto& t : th) t.join();
    }
    int wa = 0, wp = 0;
    for (int c = 0; c < n_cloths; c++) {
        if (tabs[c].rc != OC_OK) {
            std::string e = "cloth " + std::to_string(c) + ": " + tabs[c].err;
            int code = tabs[c].rc;
            delete b;
            return set_error(code, e);
        }
        wa = std::max(wa, tabs[c].topo.wa);
        wp = std::max(wp, tabs[c].topo.wp);
        b->n_planes = std::max(b->n_planes, tabs[c].topo.n_planes);
    }
    // stencil-form SpMV when every cloth has it with the same plane table
    bool stencil = n_cloths > 0;
    for (int c = 0; c < n_cloths && stencil; c++) {
        const oc::GridTopology& t = tabs[c].topo;
        stencil = t.stencil;
        for (int k = 0; k < oc::kStencil && stencil; k++)
            stencil = t.stencil_plane[k] == tabs[0].topo.stencil_plane[k];
    }
    // stencil assembly form (neighbours and slots from a per-vertex mask)
    // when every cloth has it, unless OPENCLOTH_ASSEMBLY=per-end asks for the
    // slot-table assembly (measurement)
    bool st_asm = stencil;
    for (int c = 0; c < n_cloths && st_asm; c++) st_asm = tabs[c].topo.stencil_asm;
    {  // OPENCLOTH_ASSEMBLY=once asks for the spring-once form
        const char* asm_env = std::getenv("OPENCLOTH_ASSEMBLY");
        if (asm_env && std::strcmp(asm_env, "per-end") == 0) st_asm = false;
        if (asm_env && std::strcmp(asm_env, "once") == 0) b->allow_once = true;
    }
    const size_t N = b->Ntot, S = std::max(b->Stot, 1);
    const size_t vec = align_up(sizeof(float) * 3 * N);
    const size_t blocks = align_up(sizeof(float) * 6 * N * std::max(b->n_planes, 1));
    const size_t as = align_up(sizeof(int32_t) * wa * N + 4);
    const size_t am = align_up(sizeof(uint16_t) * wa * N + 4);
    const size_t pm = align_up(sizeof(uint16_t) * wp * N + 4);
    const size_t pk = align_up(sizeof(uint16_t) * (stencil ? N : 1));
    const size_t sp = align_up(sizeof(float) * S);
    const size_t slotc = align_up(sizeof(float) * 3 * (size_t)std::max(wa, 1) * N);
    const size_t pin = align_up(N);
    const size_t per = align_up(sizeof(int) * n_cloths);
    const size_t amask = align_up(sizeof(uint32_t) * (st_asm ? N : 1));
    const size_t cst = align_up(sizeof(float) * 3 * oc::kOwned * (st_asm ? N : 1));
    const size_t total = 8 * vec + blocks + as + am + pm + pk + 3 * sp + slotc + pin + 6 * per + amask + cst + 256;
    if (hipMalloc(&b->arena, total) != hipSuccess) {
        delete b;
        return set_error(OC_ERR_OUT_OF_MEMORY, "hipMalloc of " + std::to_string(total) + " bytes failed");
    }
    char* p = (char*)b->arena;
    auto take = [&](size_t n) {
        char* q = p;
        p += n;
        return q;
    };
    oc::DevBatch& d = b->dc;
    d.Ntot = b->Ntot;
    d.wa = wa;
    d.wp = wp;
    d.x = (float*)take(vec);
    d.v = (float*)take(vec);
    d.dv = (float*)take(vec);
    d.r = (float*)take(vec);
    d.d = (float*)take(vec);
    d.q = (float*)take(vec);
    d.b = (float*)take(vec);
    d.pinv = (float*)take(vec);
    d.blocks = (float*)take(blocks);
    d.ainc_s = (const int32_t*)take(as);
    d.ainc_m = (const uint16_t*)take(am);
    d.pair_m = (const uint16_t*)take(pm);
    d.pair_mask = (const uint16_t*)take(pk);
    d.stencil = stencil ? 1 : 0;
    for (int k = 0; k < oc::kStencil; k++) d.stencil_plane[k] = stencil ? tabs[0].topo.stencil_plane[k] : 0;
    d.L0 = (const float*)take(sp);
    d.ks = (const float*)take(sp);
    d.kd = (const float*)take(sp);
    d.slot_c = (float*)take(slotc);
    d.pin = (const uint8_t*)take(pin);
    d.inst_voff = (const int*)take(per);
    d.inst_nx = (const int*)take(per);
    d.inst_n = (const int*)take(per);
    d.inst_iters = (int*)take(per);
    d.inst_status = (int*)take(per);
    d.inst_sticky = (int*)take(per);
    d.queue = (int*)take(256);
    d.asm_mask = (const uint32_t*)take(amask);
    d.stencil_asm = st_asm ? 1 : 0;
    d.cst = (float*)take(cst);

    auto fail_free = [&](int code, const std::string& msg) {
        (void)hipFree(b->arena);
        delete b;
        return set_error(code, msg);
    };
    if (hipMemset(b->arena, 0, total) != hipSuccess) return fail_free(OC_ERR_DEVICE, "hipMemset failed");

    // concatenated tables (slot-major, stride Ntot)
    {
        std::vector<int32_t> h_as((size_t)wa * N, -1);
        std::vector<uint16_t> h_am((size_t)wa * N, oc::kEmptySlot);
        std::vector<uint16_t> h_pm((size_t)wp * N, oc::kEmptySlot);
        std::vector<uint16_t> h_mask(stencil ? N : 0, 0);
        std::vector<uint32_t> h_amask(st_asm ? N : 0, 0);
        for (int c = 0; c < n_cloths; c++) {
            const oc::GridTopology& t = tabs[c].topo;
            const size_t off = b->voff[c];
            for (int k = 0; k < t.wa; k++)
                for (int v = 0; v < t.N; v++) {
                    const int32_t s = t.ainc_spring[(size_t)k * t.N + v];
                    h_as[(size_t)k * N + off + v] = s < 0 ? -1 : (int32_t)(s + soff[c]);
                    h_am[(size_t)k * N + off + v] = t.ainc_meta[(size_t)k * t.N + v];
                }
            for (int k = 0; k < t.wp; k++)
                for (int v = 0; v < t.N; v++)
                    h_pm[(size_t)k * N + off + v] = t.pair_meta[(size_t)k * t.N + v];
            if (stencil)
                for (int v = 0; v < t.N; v++) h_mask[off + v] = t.pair_mask[v];
            if (st_asm)
                for (int v = 0; v < t.N; v++) h_amask[off + v] = t.asm_mask[v];
        }
        if (hipMemcpy((void*)d.ainc_s, h_as.data(), h_as.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy((void*)d.ainc_m, h_am.data(), h_am.size() * 2, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy((void*)d.pair_m, h_pm.data(), h_pm.size() * 2, hipMemcpyHostToDevice) != hipSuccess ||
            (stencil && hipMemcpy((void*)d.pair_mask, h_mask.data(), h_mask.size() * 2,
                                  hipMemcpyHostToDevice) != hipSuccess) ||
            (st_asm && hipMemcpy((void*)d.asm_mask, h_amask.data(), h_amask.size() * 4,
                                 hipMemcpyHostToDevice) != hipSuccess))
            return fail_free(OC_ERR_DEVICE, "upload of slot tables failed");
    }
    tabs.clear();
    tabs.shrink_to_fit();
    {
        std::vector<float> L0(S, 0.0f), ks(S, 0.0f), kd(S, 0.0f);
        for (int s = 0; s < b->Stot; s++) {
            L0[s] = springs[s].rest_length;
            ks[s] = springs[s].Ks;
            kd[s] = springs[s].Kd;
        }
        std::vector<uint8_t> pins(N, 0);
        if (pinned)
            for (size_t i = 0; i < N; i++) pins[i] = pinned[i] ? 1 : 0;
        std::vector<int> nxv(b->nx), nv(b->nverts);
        if (hipMemcpy((void*)d.L0, L0.data(), S * 4, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy((void*)d.ks, ks.data(), S * 4, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy((void*)d.kd, kd.data(), S * 4, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy((void*)d.pin, pins.data(), N, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy((void*)d.inst_voff, b->voff.data(), sizeof(int) * n_cloths, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy((void*)d.inst_nx, nxv.data(), sizeof(int) * n_cloths, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy((void*)d.inst_n, nv.data(), sizeof(int) * n_cloths, hipMemcpyHostToDevice) != hipSuccess)
            return fail_free(OC_ERR_DEVICE, "upload failed");
        if (oc::launch_batch_slot_constants(d, 0, 3, nullptr) != hipSuccess ||
            oc::launch_batch_owner_constants(d, nullptr) != hipSuccess ||
            hipDeviceSynchronize() != hipSuccess)
            return fail_free(OC_ERR_DEVICE, "slot constants failed");
    }
    if ((rc = apply_schedule(b)) != OC_OK) {
        std::string e = oc_last_error();
        return fail_free(rc, e);
    }
    *out = b;
    rc = oc_batch_set_state(b, x0, v0);
    if (rc != OC_OK) {
        std::string e = oc_last_error();
        oc_batch_destroy(b);
        *out = nullptr;
        return set_error(rc, e);
    }
    return OC_OK;
}

int oc_batch_destroy(oc_batch b) {
    if (!b) return OC_OK;
    (void)hipStreamSynchronize(b->stream);
    if (b->graph) (void)hipGraphExecDestroy(b->graph);
    if (b->capture_stream) (void)hipStreamDestroy(b->capture_stream);
    if (b->sched_buf) (void)hipFree(b->sched_buf);
    if (b->arena) (void)hipFree(b->arena);
    if (b->adj_arena) (void)hipFree(b->adj_arena);
    delete b;
    return OC_OK;
}

int oc_batch_set_stream(oc_batch b, void* stream) {
    if (!b) return set_error(OC_ERR_INVALID_ARGUMENT, "batch is NULL");
    b->stream = (hipStream_t)stream;
    return OC_OK;
}

int oc_batch_set_params(oc_batch b, const oc_params* p) {
    if (!b) return set_error(OC_ERR_INVALID_ARGUMENT, "batch is NULL");
    int rc = check_params(p);
    if (rc != OC_OK) return rc;
    b->prm = *p;
    b->graph_dirty = true;
    return OC_OK;
}

int oc_batch_set_schedule(oc_batch b, const oc_batch_schedule* s) {
    if (!b || !s) return set_error(OC_ERR_INVALID_ARGUMENT, "NULL argument");
    oc_batch_schedule old = b->sched;
    b->sched = *s;
    int rc = apply_schedule(b);
    if (rc != OC_OK) b->sched = old;  // the old tables are still in place
    return rc;
}

int oc_batch_get_schedule(oc_batch b, oc_batch_schedule* s) {
    if (!b || !s) return set_error(OC_ERR_INVALID_ARGUMENT, "NULL argument");
    *s = b->sched;
    return OC_OK;
}

int oc_batch_threads_per_cloth(oc_batch b, int32_t* threads) {
    if (!b || !threads) return set_error(OC_ERR_INVALID_ARGUMENT, "NULL argument");
    for (int c = 0; c < b->n; c++) threads[c] = b->T_of[c];
    return OC_OK;
}

int oc_batch_set_state(oc_batch b, const float* x, const float* v) {
    if (!b) return set_error(OC_ERR_INVALID_ARGUMENT, "batch is NULL");
    const size_t N = b->Ntot;
    std::vector<float> t(3 * N);
    // ordered on the batch's stream; t is reused, so each copy completes first
    auto up = [&](float* dst, const float* src) -> int {
        for (size_t i = 0; i < N; i++)
            for (int k = 0; k < 3; k++) t[k * N + i] = src[3 * i + k];
        OC_HIP(hipMemcpyAsync(dst, t.data(), t.size() * 4, hipMemcpyHostToDevice, b->stream));
        OC_HIP(hipStreamSynchronize(b->stream));
        return OC_OK;
    };
    int rc;
    if (x && (rc = up(b->dc.x, x)) != OC_OK) return rc;
    if (v && (rc = up(b->dc.v, v)) != OC_OK) return rc;
    if (x) {
        OC_HIP(hipMemsetAsync(b->dc.dv, 0, sizeof(float) * 3 * N, b->stream));
        OC_HIP(hipStreamSynchronize(b->stream));
    }
    return OC_OK;
}

int oc_batch_get_state(oc_batch b, float* x, float* v) {
    if (!b) return set_error(OC_ERR_INVALID_ARGUMENT, "batch is NULL");
    const size_t N = b->Ntot;
    std::vector<float> t(3 * N);
    OC_HIP(hipStreamSynchronize(b->stream));
    auto down = [&](float* dst, const float* src) -> int {
        OC_HIP(hipMemcpy(t.data(), src, t.size() * 4, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < N; i++)
            for (int k = 0; k < 3; k++) dst[3 * i + k] = t[k * N + i];
        return OC_OK;
    };
    int rc;
    if (x && (rc = down(x, b->dc.x)) != OC_OK) return rc;
    if (v && (rc = down(v, b->dc.v)) != OC_OK) return rc;
    return OC_OK;
}

int oc_batch_step(oc_batch b, int32_t nsteps) {
    if (!b) return set_error(OC_ERR_INVALID_ARGUMENT, "batch is NULL");
    if (nsteps < 0) return set_error(OC_ERR_INVALID_ARGUMENT, "nsteps < 0");
    int rc = ensure_graph(b);
    if (rc != OC_OK) return rc;
    for (int k = 0; k < nsteps; k++) OC_HIP(hipGraphLaunch(b->graph, b->stream));
    return OC_OK;
}

int oc_batch_step_eager(oc_batch b, int32_t nsteps) {
    if (!b) return set_error(OC_ERR_INVALID_ARGUMENT, "batch is NULL");
    if (nsteps < 0) return set_error(OC_ERR_INVALID_ARGUMENT, "nsteps < 0");
    for (int k = 0; k < nsteps; k++) {
        int rc = launch_step(b, b->stream);
        if (rc != OC_OK) return rc;
    }
    return OC_OK;
}

int oc_batch_get_status(oc_batch b, int32_t* iters, int32_t* status) {
    if (!b) return set_error(OC_ERR_INVALID_ARGUMENT, "batch is NULL");
    OC_HIP(hipStreamSynchronize(b->stream));
    if (iters)
        OC_HIP(hipMemcpy(iters, b->dc.inst_iters, sizeof(int) * b->n, hipMemcpyDeviceToHost));
    if (status)
        OC_HIP(hipMemcpy(status, b->dc.inst_status, sizeof(int) * b->n, hipMemcpyDeviceToHost));
    return OC_OK;
}

int oc_batch_sticky_status(oc_batch b, int32_t* status) {
    if (!b || !status) return set_error(OC_ERR_INVALID_ARGUMENT, "NULL argument");
    OC_HIP(hipStreamSynchronize(b->stream));
    OC_HIP(hipMemcpy(status, b->dc.inst_sticky, sizeof(int) * b->n, hipMemcpyDeviceToHost));
    OC_HIP(hipMemsetAsync(b->dc.inst_sticky, 0, sizeof(int) * b->n, b->stream));
    OC_HIP(hipStreamSynchronize(b->stream));
    return OC_OK;
}

int oc_batch_synchronize(oc_batch b) {
    if (!b) return set_error(OC_ERR_INVALID_ARGUMENT, "batch is NULL");
    OC_HIP(hipStreamSynchronize(b->stream));
    return OC_OK;
}

int oc_batch_autotune(oc_batch b, oc_batch_schedule* best) {
    if (!b) return set_error(OC_ERR_INVALID_ARGUMENT, "batch is NULL");
    const size_t bytes = sizeof(float) * 3 * (size_t)b->Ntot;
    void* save = nullptr;
    if (hipMalloc(&save, 3 * bytes) != hipSuccess)
        return set_error(OC_ERR_OUT_OF_MEMORY, "hipMalloc of the autotune snapshot failed");
    char* sv = (char*)save;
    hipEvent_t e0, e1;
    OC_HIP(hipEventCreate(&e0));
    OC_HIP(hipEventCreate(&e1));
    auto snapshot = [&](bool restore) -> int {
        float* bufs[3] = {b->dc.x, b->dc.v, b->dc.dv};
        for (int k = 0; k < 3; k++) {
            if (restore)
                OC_HIP(hipMemcpyAsync(bufs[k], sv + k * bytes, bytes, hipMemcpyDeviceToDevice, b->stream));
            else
                OC_HIP(hipMemcpyAsync(sv + k * bytes, bufs[k], bytes, hipMemcpyDeviceToDevice, b->stream));
        }
        return OC_OK;
    };
    const oc_batch_schedule start = b->sched;
    oc_batch_schedule win = start;
    float win_ms = 1e30f;
    int rc = snapshot(false);
    const int nts[3] = {256, 512, 1024};
    const int vpts[6] = {1, 2, 4, 8, 16, 32};
    constexpr int kReps = 3;  // each schedule: the fastest of 3 timed steps from the same state
    for (int a = 0; a < 3 && rc == OC_OK; a++)
        for (int v = 0; v < 6 && rc == OC_OK; v++) {
            b->sched = {nts[a], vpts[v]};
            if ((rc = apply_schedule(b)) != OC_OK) break;
            if ((rc = launch_step(b, b->stream)) != OC_OK) break;  // warm-up
            for (int r = 0; r < kReps && rc == OC_OK; r++) {
                if ((rc = snapshot(true)) != OC_OK) break;
                if (hipEventRecord(e0, b->stream) != hipSuccess) { rc = set_error(OC_ERR_DEVICE, "event"); break; }
                if ((rc = launch_step(b, b->stream)) != OC_OK) break;
                if (hipEventRecord(e1, b->stream) != hipSuccess) { rc = set_error(OC_ERR_DEVICE, "event"); break; }
                if (hipEventSynchronize(e1) != hipSuccess) { rc = set_error(OC_ERR_DEVICE, "event"); break; }
                float ms = 0.0f;
                (void)hipEventElapsedTime(&ms, e0, e1);
                if (ms < win_ms) {
                    win_ms = ms;
                    win = b->sched;
                }
            }
            if (rc == OC_OK) rc = snapshot(true);
        }
    (void)hipStreamSynchronize(b->stream);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(save);
    b->sched = rc == OC_OK ? win : start;
    int rc2 = apply_schedule(b);
    if (rc != OC_OK) return rc;
    if (rc2 != OC_OK) return rc2;
    if (best) *best = win;
    return OC_OK;
}

int oc_batch_set_spring_stiffness(oc_batch b, const float* ks) {
    if (!b || !ks) return set_error(OC_ERR_INVALID_ARGUMENT, "NULL argument");
    OC_HIP(hipStreamSynchronize(b->stream));
    OC_HIP(hipMemcpy((void*)b->dc.ks, ks, sizeof(float) * b->Stot, hipMemcpyHostToDevice));
    OC_HIP(oc::launch_batch_slot_constants(b->dc, 1, 1, b->stream));
    OC_HIP(oc::launch_batch_owner_constants(b->dc, b->stream));
    OC_HIP(hipStreamSynchronize(b->stream));
    b->roll_valid = false;
    return OC_OK;
}

int oc_batch_get_spring_stiffness(oc_batch b, float* ks) {
    if (!b || !ks) return set_error(OC_ERR_INVALID_ARGUMENT, "NULL argument");
    OC_HIP(hipStreamSynchronize(b->stream));
    OC_HIP(hipMemcpy(ks, b->dc.ks, sizeof(float) * b->Stot, hipMemcpyDeviceToHost));
    return OC_OK;
}

}  // extern "C"

namespace {

// Rollout arena: checkpoints [(nseg+1) x (x,v,dv)], one segment of pre-step
// states and solutions [every x (x,v,dv)], mu, gvt, gx[2], gv[2], gks.
struct BatchRollout {
    float* ckpt;
    float* seg;
    float* mu;
    float* gvt;
    float* gx[2];
    float* gv[2];
    float* gks;
};

BatchRollout batch_rollout_views(oc_batch_s* b, int nseg, int every) {
    const size_t v3 = 3 * (size_t)b->Ntot;
    BatchRollout r;
    float* p = (float*)b->adj_arena;
    r.ckpt = p;
    p += (size_t)(nseg + 1) * 3 * v3;
    r.seg = p;
    p += (size_t)every * 3 * v3;
    r.mu = p;
    p += v3;
    r.gvt = p;
    p += v3;
    for (int k = 0; k < 2; k++) {
        r.gx[k] = p;
        p += v3;
        r.gv[k] = p;
        p += v3;
    }
    r.gks = p;
    return r;
}

int batch_copy3(oc_batch_s* b, float* dst, const float* x, const float* v, const float* dv) {
    const size_t n = sizeof(float) * 3 * (size_t)b->Ntot;
    OC_HIP(hipMemcpyAsync(dst, x, n, hipMemcpyDeviceToDevice, b->stream));
    OC_HIP(hipMemcpyAsync((char*)dst + n, v, n, hipMemcpyDeviceToDevice, b->stream));
    OC_HIP(hipMemcpyAsync((char*)dst + 2 * n, dv, n, hipMemcpyDeviceToDevice, b->stream));
    return OC_OK;
}

int batch_restore(oc_batch_s* b, const float* src) {
    const size_t n = sizeof(float) * 3 * (size_t)b->Ntot;
    OC_HIP(hipMemcpyAsync(b->dc.x, src, n, hipMemcpyDeviceToDevice, b->stream));
    OC_HIP(hipMemcpyAsync(b->dc.v, (const char*)src + n, n, hipMemcpyDeviceToDevice, b->stream));
    OC_HIP(hipMemcpyAsync(b->dc.dv, (const char*)src + 2 * n, n, hipMemcpyDeviceToDevice, b->stream));
    return OC_OK;
}

int batch_upload_aos(oc_batch_s* b, float* dst, const float* aos) {
    const size_t N = b->Ntot;
    std::vector<float> t(3 * N, 0.0f);
    if (aos)
        for (size_t i = 0; i < N; i++)
            for (int k = 0; k < 3; k++) t[k * N + i] = aos[3 * i + k];
    OC_HIP(hipMemcpyAsync(dst, t.data(), t.size() * 4, hipMemcpyHostToDevice, b->stream));
    OC_HIP(hipStreamSynchronize(b->stream));
    return OC_OK;
}

int batch_download_aos(oc_batch_s* b, const float* src, float* aos) {
    const size_t N = b->Ntot;
    std::vector<float> t(3 * N);
    OC_HIP(hipMemcpyAsync(t.data(), src, t.size() * 4, hipMemcpyDeviceToHost, b->stream));
    OC_HIP(hipStreamSynchronize(b->stream));
    for (size_t i = 0; i < N; i++)
        for (int k = 0; k < 3; k++) aos[3 * i + k] = t[k * N + i];
    return OC_OK;
}

}  // namespace

extern "C" {

int oc_batch_rollout_forward(oc_batch b, int32_t nsteps, int32_t every) {
    if (!b) return set_error(OC_ERR_INVALID_ARGUMENT, "batch is NULL");
    if (nsteps < 1 || every < 1)
        return set_error(OC_ERR_INVALID_ARGUMENT, "nsteps and checkpoint_every must be >= 1");
    b->roll_valid = false;  // until this rollout has completed
    every = std::min(every, nsteps);
    const int nseg = (nsteps + every - 1) / every;
    const size_t v3 = 3 * (size_t)b->Ntot;
    const size_t need = sizeof(float) * ((size_t)(nseg + 1) * 3 * v3 + (size_t)every * 3 * v3 +
                                         6 * v3 + (size_t)std::max(b->Stot, 1));
    if (need > b->adj_bytes) {
        if (b->adj_arena) {
            OC_HIP(hipStreamSynchronize(b->stream));
            OC_HIP(hipFree(b->adj_arena));
            b->adj_arena = nullptr;
            b->adj_bytes = 0;
        }
        if (hipMalloc(&b->adj_arena, need) != hipSuccess)
            return set_error(OC_ERR_OUT_OF_MEMORY, "hipMalloc of rollout checkpoints failed");
        b->adj_bytes = need;
    }
    BatchRollout rv = batch_rollout_views(b, nseg, every);
    int rc = ensure_graph(b);
    if (rc != OC_OK) return rc;
    OC_HIP(hipMemsetAsync(b->dc.dv, 0, sizeof(float) * v3, b->stream));
    for (int n = 0; n < nsteps; n++) {
        if (n % every == 0 &&
            (rc = batch_copy3(b, rv.ckpt + (size_t)(n / every) * 3 * v3, b->dc.x, b->dc.v, b->dc.dv)) != OC_OK)
            return rc;
        OC_HIP(hipGraphLaunch(b->graph, b->stream));
    }
    if ((rc = batch_copy3(b, rv.ckpt + (size_t)nseg * 3 * v3, b->dc.x, b->dc.v, b->dc.dv)) != OC_OK)
        return rc;
    b->roll_steps = nsteps;
    b->roll_every = every;
    b->roll_valid = true;
    return OC_OK;
}

int oc_batch_rollout_backward(oc_batch b, const float* grad_x, const float* grad_v, float* grad_ks,
                              float* grad_x0, float* grad_v0) {
    if (!b) return set_error(OC_ERR_INVALID_ARGUMENT, "batch is NULL");
    if (!grad_x) return set_error(OC_ERR_INVALID_ARGUMENT, "grad_x is NULL");
    if (!b->roll_valid) return set_error(OC_ERR_INVALID_ARGUMENT, "no rollout_forward to differentiate");
    const int nsteps = b->roll_steps, every = b->roll_every;
    const int nseg = (nsteps + every - 1) / every;
    const size_t v3 = 3 * (size_t)b->Ntot;
    BatchRollout rv = batch_rollout_views(b, nseg, every);
    int rc;
    if ((rc = ensure_graph(b)) != OC_OK) return rc;
    if ((rc = batch_upload_aos(b, rv.gx[0], grad_x)) != OC_OK) return rc;
    if ((rc = batch_upload_aos(b, rv.gv[0], grad_v)) != OC_OK) return rc;
    OC_HIP(hipMemsetAsync(rv.gks, 0, sizeof(float) * std::max(b->Stot, 1), b->stream));
    oc::StepScalars s = batch_scalars(b);
    s.warm = 0;  // the adjoint assembly must not read or keep a warm start
    int cur = 0;
    for (int seg = nseg - 1; seg >= 0; seg--) {
        const int n0 = seg * every, n1 = std::min(n0 + every, nsteps);
        if ((rc = batch_restore(b, rv.ckpt + (size_t)seg * 3 * v3)) != OC_OK) return rc;
        for (int n = n0; n < n1; n++) {
            float* slot = rv.seg + (size_t)(n - n0) * 3 * v3;
            const size_t bytes = sizeof(float) * v3;
            OC_HIP(hipMemcpyAsync(slot, b->dc.x, bytes, hipMemcpyDeviceToDevice, b->stream));
            OC_HIP(hipMemcpyAsync(slot + v3, b->dc.v, bytes, hipMemcpyDeviceToDevice, b->stream));
            OC_HIP(hipGraphLaunch(b->graph, b->stream));
            OC_HIP(hipMemcpyAsync(slot + 2 * v3, b->dc.dv, bytes, hipMemcpyDeviceToDevice, b->stream));
        }
        for (int n = n1 - 1; n >= n0; n--) {
            float* slot = rv.seg + (size_t)(n - n0) * 3 * v3;
            oc::DevBatch ca = b->dc;
            ca.x = slot;
            ca.v = slot + v3;
            ca.dv = rv.mu;
            oc::BatchAdjBufs ab;
            ab.gx_out = rv.gx[cur];
            ab.gv_out = rv.gv[cur];
            ab.gvt = rv.gvt;
            ab.dvsol = slot + 2 * v3;
            ab.gx_in = rv.gx[1 - cur];
            ab.gv_in = rv.gv[1 - cur];
            ab.gks = rv.gks;
            OC_HIP(oc::launch_batch_reset(b->dc, b->stream));
            OC_HIP(oc::launch_batch_adjoint(ca, s, ab, b->sched.block_threads, b->grid, b->stream));
            cur = 1 - cur;
        }
    }
    if ((rc = batch_restore(b, rv.ckpt + (size_t)nseg * 3 * v3)) != OC_OK) return rc;
    if (grad_ks) {
        OC_HIP(hipMemcpyAsync(grad_ks, rv.gks, sizeof(float) * b->Stot, hipMemcpyDeviceToHost, b->stream));
        OC_HIP(hipStreamSynchronize(b->stream));
    }
    if (grad_x0 && (rc = batch_download_aos(b, rv.gx[cur], grad_x0)) != OC_OK) return rc;
    if (grad_v0 && (rc = batch_download_aos(b, rv.gv[cur], grad_v0)) != OC_OK) return rc;
    OC_HIP(hipStreamSynchronize(b->stream));
    return OC_OK;
}

}  // extern "C"
