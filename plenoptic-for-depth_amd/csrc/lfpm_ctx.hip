// lfpm_ctx.hip — frame-level context: owns device buffers across frames of the same layout
// (the reference reallocates and resets the device every frame, DepthSolver.cpp:327-377),
// drives slice -> solve -> filter for host frames (one at a time or a batch of frames in one
// lfpm_batch_run), and keeps each frame's results resident for the per-frame products.
#include <hip/hip_runtime.h>

#include <cstring>
#include <vector>

#include "../../include/lfpm.h"
#include "lfpm_internal.h"

using lfpm::set_error;

namespace {

// Device buffers of one frame.
struct FrameBufs {
    uint8_t* raw = nullptr;
    float* tiles = nullptr;
    float* gray = nullptr;
    float* plane = nullptr;
    float* cost = nullptr;
    float* disp = nullptr;
    uint32_t* sel = nullptr;
    float* plane_b = nullptr;   // red/black partner state (batched solves)
    float* cost_b = nullptr;
    uint32_t* sel_b = nullptr;
    float* fplane = nullptr;
    float* fcost = nullptr;
    float* fdisp = nullptr;
    uint32_t* fsel = nullptr;
    uint8_t* reason = nullptr;
    // results the products read: the filtered copies when the filter ran
    const float* out_plane = nullptr;
    const float* out_cost = nullptr;
    bool gray_ok = false;
};

void free_frame(FrameBufs& f) {
    void* ps[] = {f.raw, f.tiles, f.gray, f.plane, f.cost, f.disp, f.sel, f.plane_b, f.cost_b,
                  f.sel_b, f.fplane, f.fcost, f.fdisp, f.fsel, f.reason};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    f = FrameBufs();
}

}  // namespace

struct lfpm_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    // tables of the current MI set (a frame layout, or caller-supplied tiles)
    lfpm_layout_params lp{};
    bool have_layout = false;   // tables belong to `lp` (false: caller-supplied tile set)
    int V = 0, crop = 0, K = 0;
    int raw_w = 0, raw_h = 0;   // geometry of the frames' raw buffers
    float* centers = nullptr;
    int* origins = nullptr;
    int* nbr = nullptr;
    int* cnt = nullptr;
    std::vector<FrameBufs> frames;
    bool frames_batched = false;  // frames carry B state
    bool frames_blur = false;     // tiles hold the blur tiles after the intensities
    void* ws = nullptr;
    size_t ws_bytes = 0;
    int solved = 0;               // frames of the last solve call
    int variant = 0;              // LFPM_VARIANT_* of single-frame solves
    int filter_order = LFPM_FILTER_SNAPSHOT;
    // level schedule of the ordered filter over the current MI set (lfpm_filter_order_levels)
    int* lvl_order = nullptr;     // device, V ints
    std::vector<int> lvl_start;   // host, levels + 1 ints
    // per-frame products (sized for the last raw frame)
    void* pws = nullptr;
    size_t pws_bytes = 0;
    float* pout = nullptr;        // mosaic, vd (raw px each), 4 range floats
    uint8_t* pcol = nullptr;      // colour image
    size_t pout_px = 0;
};

#define CTX_TRY(expr)                                                                   \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess)                                                           \
            return set_error(LFPM_E_HIP, "%s: %s", #expr, hipGetErrorString(e_));       \
    } while (0)

static void free_frames(lfpm_ctx* c) {
    for (FrameBufs& f : c->frames) free_frame(f);
    c->frames.clear();
    if (c->ws) (void)hipFree(c->ws);
    c->ws = nullptr;
    c->ws_bytes = 0;
    c->solved = 0;
}

static void free_tables(lfpm_ctx* c) {
    free_frames(c);
    void* ps[] = {c->centers, c->origins, c->nbr, c->cnt, c->lvl_order};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    c->centers = nullptr; c->origins = nullptr; c->nbr = nullptr; c->cnt = nullptr;
    c->lvl_order = nullptr;
    c->lvl_start.clear();
    c->have_layout = false;
    c->V = c->crop = c->K = 0;
}

static void free_products(lfpm_ctx* c) {
    if (c->pws) (void)hipFree(c->pws);
    if (c->pout) (void)hipFree(c->pout);
    if (c->pcol) (void)hipFree(c->pcol);
    c->pws = nullptr; c->pws_bytes = 0; c->pout = nullptr; c->pcol = nullptr; c->pout_px = 0;
}

static int alloc_tables(lfpm_ctx* c, int V, int W, int K) {
    CTX_TRY(hipMalloc((void**)&c->centers, 8 * (size_t)V));
    CTX_TRY(hipMalloc((void**)&c->origins, 8 * (size_t)V));
    CTX_TRY(hipMalloc((void**)&c->nbr, 4 * (size_t)V * K));
    CTX_TRY(hipMalloc((void**)&c->cnt, 4 * (size_t)V));
    CTX_TRY(hipMalloc((void**)&c->lvl_order, 4 * (size_t)V));
    c->V = V; c->crop = W; c->K = K;
    return LFPM_OK;
}

// Level schedule of the ordered filter from the host neighbour tables of the current MI set;
// the device copy is ordered on the context's stream before any filter launch.
static int build_levels(lfpm_ctx* c, const int* nbr, const int* cnt) {
    std::vector<int> order((size_t)c->V);
    c->lvl_start.assign((size_t)c->V + 1, 0);
    int levels = 0;
    const int rc = lfpm_filter_order_levels(nbr, cnt, c->V, c->K, order.data(),
                                            c->lvl_start.data(), &levels);
    if (rc) return rc;
    c->lvl_start.resize((size_t)levels + 1);
    CTX_TRY(hipMemcpy(c->lvl_order, order.data(), 4 * (size_t)c->V, hipMemcpyHostToDevice));
    return LFPM_OK;
}

static int run_filter(lfpm_ctx* c, FrameBufs& f, const lfpm_pm_params* p,
                      const lfpm_filter_params* flt) {
    if (c->filter_order == LFPM_FILTER_REFERENCE_ORDER)
        return lfpm_filter_frame_ordered(f.gray, c->centers, c->nbr, c->cnt, c->V, c->crop, c->K,
                                         p->base, flt, f.plane, f.cost, f.disp, f.sel, f.fplane,
                                         f.fcost, f.fdisp, f.fsel, f.reason, c->lvl_order,
                                         c->lvl_start.data(), (int)c->lvl_start.size() - 1,
                                         c->stream);
    return lfpm_filter_frame(f.gray, c->centers, c->nbr, c->cnt, c->V, c->crop, c->K, p->base,
                             flt, f.plane, f.cost, f.disp, f.sel, f.fplane, f.fcost, f.fdisp,
                             f.fsel, f.reason, c->stream);
}

// Device buffers for n frames (kept when the count and batching match) and the workspace.
// With the blur term (p->blur_lambda != 0) the tiles carry the frame's Laplacian blur tiles after
// the intensities; that needs the raw frame (not caller-supplied tiles).
static int ensure_frames(lfpm_ctx* c, int n, bool batched, const lfpm_pm_params* p,
                         bool raw_frames) {
    const bool blur = p && p->blur_lambda != 0.0f;
    if (blur && !raw_frames)
        return set_error(LFPM_E_UNSUPPORTED, "ctx: the blur term needs the raw frame's blur "
                         "tiles");
    const size_t npx = (size_t)c->V * c->crop * c->crop;
    const size_t rawb = (size_t)c->raw_w * c->raw_h;
    if ((int)c->frames.size() != n || c->frames_batched != batched || c->frames_blur != blur) {
        free_frames(c);
        c->frames.resize(n);
        c->frames_blur = blur;
        for (FrameBufs& f : c->frames) {
            CTX_TRY(hipMalloc((void**)&f.raw, rawb));
            CTX_TRY(hipMalloc((void**)&f.tiles, npx * 4 * (blur ? 2 : 1)));
            CTX_TRY(hipMalloc((void**)&f.gray, npx * 4));
            CTX_TRY(hipMalloc((void**)&f.plane, npx * 16));
            CTX_TRY(hipMalloc((void**)&f.cost, npx * 4));
            CTX_TRY(hipMalloc((void**)&f.disp, npx * 16));
            CTX_TRY(hipMalloc((void**)&f.sel, npx * 4));
            if (batched) {
                CTX_TRY(hipMalloc((void**)&f.plane_b, npx * 16));
                CTX_TRY(hipMalloc((void**)&f.cost_b, npx * 4));
                CTX_TRY(hipMalloc((void**)&f.sel_b, npx * 4));
            }
            CTX_TRY(hipMalloc((void**)&f.fplane, npx * 16));
            CTX_TRY(hipMalloc((void**)&f.fcost, npx * 4));
            CTX_TRY(hipMalloc((void**)&f.fdisp, npx * 16));
            CTX_TRY(hipMalloc((void**)&f.fsel, npx * 4));
            CTX_TRY(hipMalloc((void**)&f.reason, npx));
        }
        c->frames_batched = batched;
    }
    if (!batched) {
        const size_t need = c->variant == LFPM_VARIANT_ACMM
                                ? lfpm_patchmatch_acmm_workspace_size(c->V, p)
                                : lfpm_patchmatch_workspace_size(c->V, p);
        if (need == 0) return set_error(LFPM_E_INVALID, "ctx: bad solve parameters");
        if (need > c->ws_bytes) {
            if (c->ws) (void)hipFree(c->ws);
            c->ws = nullptr;
            c->ws_bytes = 0;
            CTX_TRY(hipMalloc(&c->ws, need));
            c->ws_bytes = need;
        }
    }
    for (FrameBufs& f : c->frames) {
        f.out_plane = nullptr;
        f.out_cost = nullptr;
        f.gray_ok = false;
    }
    c->solved = 0;
    return LFPM_OK;
}

static bool same_layout(const lfpm_layout_params& a, const lfpm_layout_params& b) {
    return std::memcmp(&a, &b, sizeof(a)) == 0;
}

// Tables of a frame layout (kept while the layout is unchanged).
static int ensure_layout(lfpm_ctx* c, const lfpm_layout_params* lp, const lfpm_pm_params* p) {
    if (c->have_layout && same_layout(c->lp, *lp) && c->K == p->max_neighbors &&
        c->crop == p->crop)
        return LFPM_OK;
    free_tables(c);
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
    if ((rc = alloc_tables(c, V, W, K))) return rc;
    CTX_TRY(hipMemcpy(c->centers, centers.data(), centers.size() * 4, hipMemcpyHostToDevice));
    CTX_TRY(hipMemcpy(c->origins, origins.data(), origins.size() * 4, hipMemcpyHostToDevice));
    CTX_TRY(hipMemcpy(c->nbr, nbr.data(), nbr.size() * 4, hipMemcpyHostToDevice));
    CTX_TRY(hipMemcpy(c->cnt, cnt.data(), cnt.size() * 4, hipMemcpyHostToDevice));
    if ((rc = build_levels(c, nbr.data(), cnt.data()))) return rc;
    c->lp = *lp;
    c->raw_w = lp->raw_w;
    c->raw_h = lp->raw_h;
    c->have_layout = true;
    return LFPM_OK;
}

// slice -> solve -> (filter) of a frame whose raw is already on the device (single-frame form)
static int solve_one(lfpm_ctx* c, FrameBufs& f, const lfpm_pm_params* p,
                     const lfpm_filter_params* flt) {
    int rc;
    if ((rc = lfpm_slice_tiles(f.raw, c->raw_w, c->raw_h, (size_t)c->raw_w, c->origins, c->V,
                               c->crop, f.tiles, c->stream)))
        return rc;
    if (p->blur_lambda != 0.0f &&
        (rc = lfpm_slice_blur_tiles(f.raw, c->raw_w, c->raw_h, (size_t)c->raw_w, c->origins,
                                    c->V, c->crop, f.tiles + (size_t)c->V * c->crop * c->crop,
                                    c->stream)))
        return rc;
    if (c->variant == LFPM_VARIANT_ACMM)
        rc = lfpm_patchmatch_frame_acmm(f.tiles, c->centers, c->nbr, c->cnt, c->V, p, f.plane,
                                        f.cost, f.disp, f.sel, c->ws, c->ws_bytes, c->stream);
    else
        rc = lfpm_patchmatch_frame(f.tiles, c->centers, c->nbr, c->cnt, c->V, p, f.plane, f.cost,
                                   f.disp, f.sel, c->ws, c->ws_bytes, 0, c->stream);
    if (rc) return rc;
    f.out_plane = f.plane;
    f.out_cost = f.cost;
    if (flt) {
        if ((rc = lfpm_slice_gray(f.raw, c->raw_w, c->raw_h, (size_t)c->raw_w, c->origins, c->V,
                                  c->crop, f.gray, c->stream)))
            return rc;
        f.gray_ok = true;
        if ((rc = run_filter(c, f, p, flt))) return rc;
        f.out_plane = f.fplane;
        f.out_cost = f.fcost;
    }
    return LFPM_OK;
}

static int download(lfpm_ctx* c, const FrameBufs& f, bool filtered, float* plane, float* cost,
                    float* disp, uint32_t* sel, uint8_t* reason) {
    const size_t npx = (size_t)c->V * c->crop * c->crop;
    const float* po = filtered ? f.fplane : f.plane;
    const float* co = filtered ? f.fcost : f.cost;
    const float* dout = filtered ? f.fdisp : f.disp;
    const uint32_t* so = filtered ? f.fsel : f.sel;
    if (plane) CTX_TRY(hipMemcpyAsync(plane, po, npx * 16, hipMemcpyDeviceToHost, c->stream));
    if (cost) CTX_TRY(hipMemcpyAsync(cost, co, npx * 4, hipMemcpyDeviceToHost, c->stream));
    if (disp) CTX_TRY(hipMemcpyAsync(disp, dout, npx * 16, hipMemcpyDeviceToHost, c->stream));
    if (sel) CTX_TRY(hipMemcpyAsync(sel, so, npx * 4, hipMemcpyDeviceToHost, c->stream));
    if (reason && filtered)
        CTX_TRY(hipMemcpyAsync(reason, f.reason, npx, hipMemcpyDeviceToHost, c->stream));
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
    free_tables(c);
    free_products(c);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int lfpm_ctx_set_variant(lfpm_ctx* c, int variant) {
    if (!c) return set_error(LFPM_E_INVALID, "set_variant: null ctx");
    if (variant != LFPM_VARIANT_PLAIN && variant != LFPM_VARIANT_ACMM)
        return set_error(LFPM_E_INVALID, "set_variant: unknown variant %d", variant);
    c->variant = variant;  // the next solve re-sizes the workspace for it
    return LFPM_OK;
}

int lfpm_ctx_set_filter_order(lfpm_ctx* c, int order) {
    if (!c) return set_error(LFPM_E_INVALID, "set_filter_order: null ctx");
    if (order != LFPM_FILTER_SNAPSHOT && order != LFPM_FILTER_REFERENCE_ORDER)
        return set_error(LFPM_E_INVALID, "set_filter_order: unknown order %d", order);
    c->filter_order = order;
    return LFPM_OK;
}

int lfpm_ctx_solve_frame(lfpm_ctx* c, const lfpm_layout_params* lp, const uint8_t* raw_host,
                         const lfpm_pm_params* p, const lfpm_filter_params* f, float* plane,
                         float* cost, float* disp, uint32_t* sel, uint8_t* reason) {
    if (!c || !lp || !raw_host || !p) return set_error(LFPM_E_INVALID, "solve_frame: null arg");
    if (reason && !f)
        return set_error(LFPM_E_INVALID, "solve_frame: reason requested without filter");
    CTX_TRY(hipSetDevice(c->device));
    int rc = ensure_layout(c, lp, p);
    if (rc) return rc;
    if ((rc = ensure_frames(c, 1, false, p, true))) return rc;
    FrameBufs& fb = c->frames[0];
    CTX_TRY(hipMemcpyAsync(fb.raw, raw_host, (size_t)c->raw_w * c->raw_h, hipMemcpyHostToDevice,
                           c->stream));
    const lfpm_filter_params* flt = reason ? f : nullptr;
    if ((rc = solve_one(c, fb, p, flt))) return rc;
    if ((rc = download(c, fb, flt != nullptr, plane, cost, disp, sel, reason))) return rc;
    CTX_TRY(hipStreamSynchronize(c->stream));
    c->solved = 1;
    return LFPM_OK;
}

int lfpm_ctx_solve_frames(lfpm_ctx* c, const lfpm_layout_params* lp, int n,
                          const uint8_t* const* raws, const lfpm_pm_params* p,
                          const lfpm_filter_params* f, float* const* plane, float* const* cost,
                          float* const* disp, uint32_t* const* sel, uint8_t* const* reason) {
    if (!c || !lp || !raws || !p || n <= 0)
        return set_error(LFPM_E_INVALID, "solve_frames: null argument");
    if (n > LFPM_MAX_BATCH)
        return set_error(LFPM_E_INVALID, "solve_frames: %d frames > %d", n, LFPM_MAX_BATCH);
    if (reason && !f) return set_error(LFPM_E_INVALID, "solve_frames: reason without filter");
    if (c->variant != LFPM_VARIANT_PLAIN)
        return set_error(LFPM_E_UNSUPPORTED, "solve_frames: the batched form runs the plain "
                         "variant");
    for (int i = 0; i < n; ++i)
        if (!raws[i]) return set_error(LFPM_E_INVALID, "solve_frames: raw %d is null", i);
    CTX_TRY(hipSetDevice(c->device));
    int rc = ensure_layout(c, lp, p);
    if (rc) return rc;
    if ((rc = ensure_frames(c, n, true, p, true))) return rc;
    const lfpm_filter_params* flt = reason ? f : nullptr;
    std::vector<lfpm_pm_params> pms(n, *p);
    std::vector<lfpm_batch_frame> bf(n);
    for (int i = 0; i < n; ++i) {
        FrameBufs& fb = c->frames[i];
        pms[i].frame_uid = p->frame_uid + (uint32_t)i;
        CTX_TRY(hipMemcpyAsync(fb.raw, raws[i], (size_t)c->raw_w * c->raw_h,
                               hipMemcpyHostToDevice, c->stream));
        if ((rc = lfpm_slice_tiles(fb.raw, c->raw_w, c->raw_h, (size_t)c->raw_w, c->origins, c->V,
                                   c->crop, fb.tiles, c->stream)))
            return rc;
        if (p->blur_lambda != 0.0f &&
            (rc = lfpm_slice_blur_tiles(fb.raw, c->raw_w, c->raw_h, (size_t)c->raw_w, c->origins,
                                        c->V, c->crop,
                                        fb.tiles + (size_t)c->V * c->crop * c->crop, c->stream)))
            return rc;
        if ((rc = lfpm_slice_gray(fb.raw, c->raw_w, c->raw_h, (size_t)c->raw_w, c->origins, c->V,
                                  c->crop, fb.gray, c->stream)))
            return rc;
        fb.gray_ok = true;
        lfpm_batch_frame& b = bf[i];
        std::memset(&b, 0, sizeof(b));
        b.tiles_dev = fb.tiles; b.gray_dev = fb.gray; b.centers_dev = c->centers;
        b.neighbor_ids_dev = c->nbr; b.neighbor_counts_dev = c->cnt;
        b.num_mis = c->V; b.raw_w = c->raw_w; b.raw_h = c->raw_h;
        b.params = &pms[i];
        b.filter = flt;
        b.plane_dev = fb.plane; b.cost_dev = fb.cost; b.sel_dev = fb.sel; b.disp_dev = fb.disp;
        b.plane_b_dev = fb.plane_b; b.cost_b_dev = fb.cost_b; b.sel_b_dev = fb.sel_b;
        b.f_plane_dev = fb.fplane; b.f_cost_dev = fb.fcost; b.f_disp_dev = fb.fdisp;
        b.f_sel_dev = fb.fsel; b.reason_dev = fb.reason;
        if (flt && c->filter_order == LFPM_FILTER_REFERENCE_ORDER) {
            b.filter_mi_order_dev = c->lvl_order;
            b.filter_level_start = c->lvl_start.data();
            b.filter_num_levels = (int)c->lvl_start.size() - 1;
        }
        fb.out_plane = flt ? fb.fplane : fb.plane;
        fb.out_cost = flt ? fb.fcost : fb.cost;
    }
    const size_t need = lfpm_batch_workspace_size(bf.data(), n);
    if (need == 0) return set_error(LFPM_E_INVALID, "solve_frames: bad batch");
    if (need > c->ws_bytes) {
        if (c->ws) (void)hipFree(c->ws);
        c->ws = nullptr;
        c->ws_bytes = 0;
        CTX_TRY(hipMalloc(&c->ws, need));
        c->ws_bytes = need;
    }
    if ((rc = lfpm_batch_run(bf.data(), n, c->ws, c->ws_bytes, c->stream))) return rc;
    for (int i = 0; i < n; ++i)
        if ((rc = download(c, c->frames[i], flt != nullptr, plane ? plane[i] : nullptr,
                           cost ? cost[i] : nullptr, disp ? disp[i] : nullptr,
                           sel ? sel[i] : nullptr, reason ? reason[i] : nullptr)))
            return rc;
    CTX_TRY(hipStreamSynchronize(c->stream));
    c->solved = n;
    return LFPM_OK;
}

int lfpm_ctx_solve_tiles(lfpm_ctx* c, const uint8_t* tiles_u8, const float* centers,
                         const int* neighbor_ids, const int* neighbor_counts, int num_mis,
                         const lfpm_pm_params* p, const lfpm_filter_params* f, float* plane,
                         float* cost, float* disp, uint32_t* sel, uint8_t* reason) {
    if (!c || !tiles_u8 || !centers || !neighbor_ids || !neighbor_counts || !p || num_mis <= 0)
        return set_error(LFPM_E_INVALID, "solve_tiles: null argument");
    if (reason && !f)
        return set_error(LFPM_E_INVALID, "solve_tiles: reason requested without filter");
    CTX_TRY(hipSetDevice(c->device));
    const int V = num_mis, W = p->crop, K = p->max_neighbors;
    if (W < 3 || W > LFPM_MAX_CROP || K < 1 || K > LFPM_MAX_NEIGHBORS)
        return set_error(LFPM_E_INVALID, "solve_tiles: bad crop or K");
    for (int v = 0; v < V; ++v) {
        if (neighbor_counts[v] < 0 || neighbor_counts[v] > K)
            return set_error(LFPM_E_INVALID, "solve_tiles: count %d of MI %d", neighbor_counts[v],
                             v);
        for (int k = 0; k < neighbor_counts[v]; ++k) {
            const int nb = neighbor_ids[(size_t)v * K + k];
            if (nb < 0 || nb >= V || nb == v)
                return set_error(LFPM_E_INVALID, "solve_tiles: neighbour %d of MI %d is %d", k,
                                 v, nb);
        }
    }
    // the crops stacked vertically form a crop-wide "raw" image; the tile gather then reads
    // crop v at origin (0, v*crop)
    free_tables(c);
    int rc;
    if ((rc = alloc_tables(c, V, W, K))) return rc;
    c->raw_w = W;
    c->raw_h = V * W;
    std::vector<int> origins(2 * (size_t)V);
    for (int v = 0; v < V; ++v) { origins[2 * v] = 0; origins[2 * v + 1] = v * W; }
    CTX_TRY(hipMemcpyAsync(c->centers, centers, 8 * (size_t)V, hipMemcpyHostToDevice, c->stream));
    CTX_TRY(hipMemcpyAsync(c->origins, origins.data(), 8 * (size_t)V, hipMemcpyHostToDevice,
                           c->stream));
    CTX_TRY(hipMemcpyAsync(c->nbr, neighbor_ids, 4 * (size_t)V * K, hipMemcpyHostToDevice,
                           c->stream));
    CTX_TRY(hipMemcpyAsync(c->cnt, neighbor_counts, 4 * (size_t)V, hipMemcpyHostToDevice,
                           c->stream));
    if ((rc = build_levels(c, neighbor_ids, neighbor_counts))) return rc;
    if ((rc = ensure_frames(c, 1, false, p, false))) return rc;
    FrameBufs& fb = c->frames[0];
    CTX_TRY(hipMemcpyAsync(fb.raw, tiles_u8, (size_t)V * W * W, hipMemcpyHostToDevice,
                           c->stream));
    const lfpm_filter_params* flt = reason ? f : nullptr;
    if ((rc = solve_one(c, fb, p, flt))) return rc;
    if ((rc = download(c, fb, flt != nullptr, plane, cost, disp, sel, reason))) return rc;
    CTX_TRY(hipStreamSynchronize(c->stream));
    c->solved = 1;
    return LFPM_OK;
}

int lfpm_ctx_frame_products(lfpm_ctx* c, int frame_index, int raw_w, int raw_h, float max_cost,
                            float mla_base, float* mosaic, float* mosaic_range, float* vd,
                            uint8_t* color, float* vd_range) {
    if (!c) return set_error(LFPM_E_INVALID, "frame_products: null ctx");
    if (frame_index < 0 || frame_index >= c->solved)
        return set_error(LFPM_E_INVALID, "frame_products: frame %d of %d solved", frame_index,
                         c->solved);
    FrameBufs& f = c->frames[frame_index];
    if (!f.out_plane) return set_error(LFPM_E_INVALID, "frame_products: frame not solved");
    CTX_TRY(hipSetDevice(c->device));
    const int V = c->V, W = c->crop;
    int rc;
    if (!f.gray_ok) {
        if ((rc = lfpm_slice_gray(f.raw, c->raw_w, c->raw_h, (size_t)c->raw_w, c->origins, V, W,
                                  f.gray, c->stream)))
            return rc;
        f.gray_ok = true;
    }
    const size_t need = lfpm_products_workspace_size(V, W, raw_w, raw_h);
    if (need == 0) return set_error(LFPM_E_INVALID, "frame_products: bad sizes");
    const size_t px = (size_t)raw_w * raw_h;
    if (need > c->pws_bytes || px != c->pout_px) {
        free_products(c);
        CTX_TRY(hipMalloc(&c->pws, need));
        c->pws_bytes = need;
        CTX_TRY(hipMalloc((void**)&c->pout, (2 * px + 4) * 4));
        CTX_TRY(hipMalloc((void**)&c->pcol, px));
        c->pout_px = px;
    }
    float* dm = c->pout;
    float* dv = c->pout + px;
    float* dr = c->pout + 2 * px;
    if (mosaic || mosaic_range) {
        if ((rc = lfpm_disparity_mosaic(f.out_plane, f.out_cost, c->centers, V, W, max_cost,
                                        raw_w, raw_h, dm, dr, c->pws, c->pws_bytes, c->stream)))
            return rc;
        if (mosaic) CTX_TRY(hipMemcpyAsync(mosaic, dm, px * 4, hipMemcpyDeviceToHost, c->stream));
        if (mosaic_range)
            CTX_TRY(hipMemcpyAsync(mosaic_range, dr, 8, hipMemcpyDeviceToHost, c->stream));
    }
    if (vd || color || vd_range) {
        if ((rc = lfpm_virtual_depth_map(f.out_plane, c->centers, f.gray, V, W, mla_base, raw_w,
                                         raw_h, dv, color ? c->pcol : nullptr, dr + 2, c->pws,
                                         c->pws_bytes, c->stream)))
            return rc;
        if (vd) CTX_TRY(hipMemcpyAsync(vd, dv, px * 4, hipMemcpyDeviceToHost, c->stream));
        if (color) CTX_TRY(hipMemcpyAsync(color, c->pcol, px, hipMemcpyDeviceToHost, c->stream));
        if (vd_range)
            CTX_TRY(hipMemcpyAsync(vd_range, dr + 2, 8, hipMemcpyDeviceToHost, c->stream));
    }
    CTX_TRY(hipStreamSynchronize(c->stream));
    return LFPM_OK;
}

}  // extern "C"
