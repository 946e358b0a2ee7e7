// lfpm_ctx.hip — frame-level context: owns device buffers across frames of the same layout
// (the reference reallocates and resets the device every frame, DepthSolver.cpp:327-377) and
// drives slice -> solve -> filter for one host frame.
#include <hip/hip_runtime.h>

#include <cstring>
#include <vector>

#include "../../include/lfpm.h"
#include "lfpm_internal.h"

using lfpm::set_error;

struct lfpm_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    // cached layout
    lfpm_layout_params lp{};
    bool have_layout = false;
    int V = 0, crop = 0, K = 0;
    size_t raw_bytes = 0;
    uint8_t* raw = nullptr;
    float* centers = nullptr;
    int* origins = nullptr;
    int* nbr = nullptr;
    int* cnt = nullptr;
    float* tiles = nullptr;
    float* gray = nullptr;
    float* plane = nullptr;
    float* cost = nullptr;
    float* disp = nullptr;
    uint32_t* sel = nullptr;
    float* fplane = nullptr;
    float* fcost = nullptr;
    float* fdisp = nullptr;
    uint32_t* fsel = nullptr;
    uint8_t* reason = nullptr;
    void* ws = nullptr;
    size_t ws_bytes = 0;
};

#define CTX_TRY(expr)                                                                   \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess)                                                           \
            return set_error(LFPM_E_HIP, "%s: %s", #expr, hipGetErrorString(e_));       \
    } while (0)

static void free_layout(lfpm_ctx* c) {
    void* ps[] = {c->raw, c->centers, c->origins, c->nbr, c->cnt, c->tiles, c->gray,
                  c->plane, c->cost, c->disp, c->sel, c->fplane, c->fcost, c->fdisp,
                  c->fsel, c->reason, c->ws};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    c->raw = nullptr; c->centers = nullptr; c->origins = nullptr; c->nbr = nullptr;
    c->cnt = nullptr; c->tiles = nullptr; c->gray = nullptr; c->plane = nullptr;
    c->cost = nullptr; c->disp = nullptr; c->sel = nullptr; c->fplane = nullptr;
    c->fcost = nullptr; c->fdisp = nullptr; c->fsel = nullptr; c->reason = nullptr;
    c->ws = nullptr; c->ws_bytes = 0;
    c->have_layout = false;
}

static bool same_layout(const lfpm_layout_params& a, const lfpm_layout_params& b) {
    return std::memcmp(&a, &b, sizeof(a)) == 0;
}

static int ensure_layout(lfpm_ctx* c, const lfpm_layout_params* lp, const lfpm_pm_params* p) {
    if (c->have_layout && same_layout(c->lp, *lp) && c->K == p->max_neighbors) {
        const size_t need = lfpm_patchmatch_workspace_size(c->V, p);
        if (need <= c->ws_bytes) return LFPM_OK;
    }
    free_layout(c);
    lfpm_layout_info info;
    int rc = lfpm_layout_query(lp, &info);
    if (rc) return rc;
    if (info.crop != p->crop)
        return set_error(LFPM_E_INVALID, "ctx: layout crop %d != params crop %d", info.crop,
                         p->crop);
    if (lp->max_neighbors != p->max_neighbors)
        return set_error(LFPM_E_INVALID, "ctx: layout K %d != params K %d", lp->max_neighbors,
                         p->max_neighbors);
    const int V = info.num_mis, K = lp->max_neighbors, W = info.crop;
    std::vector<float> centers(2 * (size_t)V);
    std::vector<int> origins(2 * (size_t)V), nbr((size_t)V * K), cnt(V);
    rc = lfpm_layout_build(lp, centers.data(), origins.data(), nullptr, nullptr, nbr.data(),
                           cnt.data(), nullptr);
    if (rc) return rc;
    const size_t npx = (size_t)V * W * W;
    c->raw_bytes = (size_t)lp->raw_w * lp->raw_h;
    CTX_TRY(hipMalloc((void**)&c->raw, c->raw_bytes));
    CTX_TRY(hipMalloc((void**)&c->centers, centers.size() * 4));
    CTX_TRY(hipMalloc((void**)&c->origins, origins.size() * 4));
    CTX_TRY(hipMalloc((void**)&c->nbr, nbr.size() * 4));
    CTX_TRY(hipMalloc((void**)&c->cnt, cnt.size() * 4));
    CTX_TRY(hipMalloc((void**)&c->tiles, npx * 4));
    CTX_TRY(hipMalloc((void**)&c->gray, npx * 4));
    CTX_TRY(hipMalloc((void**)&c->plane, npx * 16));
    CTX_TRY(hipMalloc((void**)&c->cost, npx * 4));
    CTX_TRY(hipMalloc((void**)&c->disp, npx * 16));
    CTX_TRY(hipMalloc((void**)&c->sel, npx * 4));
    CTX_TRY(hipMalloc((void**)&c->fplane, npx * 16));
    CTX_TRY(hipMalloc((void**)&c->fcost, npx * 4));
    CTX_TRY(hipMalloc((void**)&c->fdisp, npx * 16));
    CTX_TRY(hipMalloc((void**)&c->fsel, npx * 4));
    CTX_TRY(hipMalloc((void**)&c->reason, npx));
    c->ws_bytes = lfpm_patchmatch_workspace_size(V, p);
    if (c->ws_bytes == 0) return set_error(LFPM_E_INVALID, "ctx: bad solve parameters");
    CTX_TRY(hipMalloc(&c->ws, c->ws_bytes));
    CTX_TRY(hipMemcpy(c->centers, centers.data(), centers.size() * 4, hipMemcpyHostToDevice));
    CTX_TRY(hipMemcpy(c->origins, origins.data(), origins.size() * 4, hipMemcpyHostToDevice));
    CTX_TRY(hipMemcpy(c->nbr, nbr.data(), nbr.size() * 4, hipMemcpyHostToDevice));
    CTX_TRY(hipMemcpy(c->cnt, cnt.data(), cnt.size() * 4, hipMemcpyHostToDevice));
    c->lp = *lp;
    c->V = V; c->crop = W; c->K = K;
    c->have_layout = true;
    return LFPM_OK;
}

extern "C" {

int lfpm_create(int device, lfpm_ctx** out) {
    if (!out) return set_error(LFPM_E_INVALID, "create: null out");
    *out = nullptr;
    int n = 0;
    CTX_TRY(hipGetDeviceCount(&n));
    if (device < 0 || device >= n)
        return set_error(LFPM_E_INVALID, "create: device %d of %d", device, n);
    CTX_TRY(hipSetDevice(device));
    lfpm_ctx* c = new lfpm_ctx();
    c->device = device;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return set_error(LFPM_E_HIP, "create: %s", hipGetErrorString(e));
    }
    *out = c;
    return LFPM_OK;
}

void lfpm_destroy(lfpm_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    free_layout(c);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int lfpm_ctx_solve_frame(lfpm_ctx* c, const lfpm_layout_params* lp, const uint8_t* raw_host,
                         const lfpm_pm_params* p, const lfpm_filter_params* f, float* plane,
                         float* cost, float* disp, uint32_t* sel, uint8_t* reason) {
    if (!c || !lp || !raw_host || !p) return set_error(LFPM_E_INVALID, "solve_frame: null arg");
    CTX_TRY(hipSetDevice(c->device));
    int rc = ensure_layout(c, lp, p);
    if (rc) return rc;
    CTX_TRY(hipMemcpyAsync(c->raw, raw_host, c->raw_bytes, hipMemcpyHostToDevice, c->stream));
    if ((rc = lfpm_slice_tiles(c->raw, lp->raw_w, lp->raw_h, (size_t)lp->raw_w, c->origins, c->V,
                               c->crop, c->tiles, c->stream)))
        return rc;
    if ((rc = lfpm_patchmatch_frame(c->tiles, c->centers, c->nbr, c->cnt, c->V, p, c->plane,
                                    c->cost, c->disp, c->sel, c->ws, c->ws_bytes, 0, c->stream)))
        return rc;
    float* po = c->plane;
    float* co = c->cost;
    float* dout = c->disp;
    uint32_t* so = c->sel;
    if (reason) {
        if (!f) return set_error(LFPM_E_INVALID, "solve_frame: reason requested without filter");
        if ((rc = lfpm_slice_gray(c->raw, lp->raw_w, lp->raw_h, (size_t)lp->raw_w, c->origins,
                                  c->V, c->crop, c->gray, c->stream)))
            return rc;
        if ((rc = lfpm_filter_frame(c->gray, c->centers, c->nbr, c->cnt, c->V, c->crop, c->K,
                                    p->base, f, c->plane, c->cost, c->disp, c->sel, c->fplane,
                                    c->fcost, c->fdisp, c->fsel, c->reason, c->stream)))
            return rc;
        po = c->fplane; co = c->fcost; dout = c->fdisp; so = c->fsel;
    }
    const size_t npx = (size_t)c->V * c->crop * c->crop;
    if (plane) CTX_TRY(hipMemcpyAsync(plane, po, npx * 16, hipMemcpyDeviceToHost, c->stream));
    if (cost) CTX_TRY(hipMemcpyAsync(cost, co, npx * 4, hipMemcpyDeviceToHost, c->stream));
    if (disp) CTX_TRY(hipMemcpyAsync(disp, dout, npx * 16, hipMemcpyDeviceToHost, c->stream));
    if (sel) CTX_TRY(hipMemcpyAsync(sel, so, npx * 4, hipMemcpyDeviceToHost, c->stream));
    if (reason)
        CTX_TRY(hipMemcpyAsync(reason, c->reason, npx, hipMemcpyDeviceToHost, c->stream));
    CTX_TRY(hipStreamSynchronize(c->stream));
    return LFPM_OK;
}

int lfpm_ctx_solve_tiles(lfpm_ctx* c, const uint8_t* tiles_u8, const float* centers,
                         const int* neighbor_ids, const int* neighbor_counts, int num_mis,
                         const lfpm_pm_params* p, const lfpm_filter_params* f, float* plane,
                         float* cost, float* disp, uint32_t* sel, uint8_t* reason) {
    if (!c || !tiles_u8 || !centers || !neighbor_ids || !neighbor_counts || !p || num_mis <= 0)
        return set_error(LFPM_E_INVALID, "solve_tiles: null argument");
    CTX_TRY(hipSetDevice(c->device));
    const int V = num_mis, W = p->crop, K = p->max_neighbors;
    if (W < 3 || W > LFPM_MAX_CROP || K < 1 || K > LFPM_MAX_NEIGHBORS)
        return set_error(LFPM_E_INVALID, "solve_tiles: bad crop or K");
    // the crops stacked vertically form a crop-wide "raw" image; the tile gather then reads
    // crop v at origin (0, v*crop)
    lfpm_layout_params lp{};
    lp.pitch = p->base; lp.crop = W; lp.raw_w = W; lp.raw_h = V * W; lp.max_neighbors = K;
    lp.lens_types = -V;  // cache key distinct from any frame layout
    free_layout(c);
    const size_t npx = (size_t)V * W * W;
    c->raw_bytes = npx;
    std::vector<int> origins(2 * (size_t)V);
    for (int v = 0; v < V; ++v) { origins[2 * v] = 0; origins[2 * v + 1] = v * W; }
    CTX_TRY(hipMalloc((void**)&c->raw, npx));
    CTX_TRY(hipMalloc((void**)&c->centers, 8 * (size_t)V));
    CTX_TRY(hipMalloc((void**)&c->origins, 8 * (size_t)V));
    CTX_TRY(hipMalloc((void**)&c->nbr, 4 * (size_t)V * K));
    CTX_TRY(hipMalloc((void**)&c->cnt, 4 * (size_t)V));
    CTX_TRY(hipMalloc((void**)&c->tiles, npx * 4));
    CTX_TRY(hipMalloc((void**)&c->gray, npx * 4));
    CTX_TRY(hipMalloc((void**)&c->plane, npx * 16));
    CTX_TRY(hipMalloc((void**)&c->cost, npx * 4));
    CTX_TRY(hipMalloc((void**)&c->disp, npx * 16));
    CTX_TRY(hipMalloc((void**)&c->sel, npx * 4));
    CTX_TRY(hipMalloc((void**)&c->fplane, npx * 16));
    CTX_TRY(hipMalloc((void**)&c->fcost, npx * 4));
    CTX_TRY(hipMalloc((void**)&c->fdisp, npx * 16));
    CTX_TRY(hipMalloc((void**)&c->fsel, npx * 4));
    CTX_TRY(hipMalloc((void**)&c->reason, npx));
    c->ws_bytes = lfpm_patchmatch_workspace_size(V, p);
    if (c->ws_bytes == 0) return set_error(LFPM_E_INVALID, "solve_tiles: bad solve parameters");
    CTX_TRY(hipMalloc(&c->ws, c->ws_bytes));
    c->V = V; c->crop = W; c->K = K;
    CTX_TRY(hipMemcpyAsync(c->raw, tiles_u8, npx, hipMemcpyHostToDevice, c->stream));
    CTX_TRY(hipMemcpyAsync(c->centers, centers, 8 * (size_t)V, hipMemcpyHostToDevice, c->stream));
    CTX_TRY(hipMemcpyAsync(c->origins, origins.data(), 8 * (size_t)V, hipMemcpyHostToDevice,
                           c->stream));
    CTX_TRY(hipMemcpyAsync(c->nbr, neighbor_ids, 4 * (size_t)V * K, hipMemcpyHostToDevice,
                           c->stream));
    CTX_TRY(hipMemcpyAsync(c->cnt, neighbor_counts, 4 * (size_t)V, hipMemcpyHostToDevice,
                           c->stream));
    int rc;
    if ((rc = lfpm_slice_tiles(c->raw, W, V * W, (size_t)W, c->origins, V, W, c->tiles,
                               c->stream)))
        return rc;
    if ((rc = lfpm_patchmatch_frame(c->tiles, c->centers, c->nbr, c->cnt, V, p, c->plane,
                                    c->cost, c->disp, c->sel, c->ws, c->ws_bytes, 0, c->stream)))
        return rc;
    float* po = c->plane;
    float* co = c->cost;
    float* dout = c->disp;
    uint32_t* so = c->sel;
    if (reason) {
        if (!f) return set_error(LFPM_E_INVALID, "solve_tiles: reason requested without filter");
        if ((rc = lfpm_slice_gray(c->raw, W, V * W, (size_t)W, c->origins, V, W, c->gray,
                                  c->stream)))
            return rc;
        if ((rc = lfpm_filter_frame(c->gray, c->centers, c->nbr, c->cnt, V, W, K, p->base, f,
                                    c->plane, c->cost, c->disp, c->sel, c->fplane, c->fcost,
                                    c->fdisp, c->fsel, c->reason, c->stream)))
            return rc;
        po = c->fplane; co = c->fcost; dout = c->fdisp; so = c->fsel;
    }
    if (plane) CTX_TRY(hipMemcpyAsync(plane, po, npx * 16, hipMemcpyDeviceToHost, c->stream));
    if (cost) CTX_TRY(hipMemcpyAsync(cost, co, npx * 4, hipMemcpyDeviceToHost, c->stream));
    if (disp) CTX_TRY(hipMemcpyAsync(disp, dout, npx * 16, hipMemcpyDeviceToHost, c->stream));
    if (sel) CTX_TRY(hipMemcpyAsync(sel, so, npx * 4, hipMemcpyDeviceToHost, c->stream));
    if (reason)
        CTX_TRY(hipMemcpyAsync(reason, c->reason, npx, hipMemcpyDeviceToHost, c->stream));
    CTX_TRY(hipStreamSynchronize(c->stream));
    c->lp = lp;
    c->have_layout = false;  // buffers belong to this tile set; the next frame call re-plans
    return LFPM_OK;
}

}  // extern "C"
