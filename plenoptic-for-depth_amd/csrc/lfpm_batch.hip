// lfpm_batch.hip — batched and served forms over the frame-table kernel (k_stage_w):
// class grouping, the stream-ordered batch schedule, its hipGraph capture, the continuous-
// batching step and the occupancy autotuner.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "../../include/lfpm.h"
#include "lfpm_internal.h"

using lfpm::set_error;
using lfpm::StageFrame;

namespace {

#define B_TRY(expr)                                                                     \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess)                                                           \
            return set_error(LFPM_E_HIP, "%s: %s", #expr, hipGetErrorString(e_));       \
    } while (0)

// Parameter-class equality: everything but the per-frame random key.
bool same_class(const lfpm_pm_params& a, const lfpm_pm_params& b) {
    lfpm_pm_params x = a, y = b;
    x.seed = y.seed = 0;
    x.frame_uid = y.frame_uid = 0;
    return std::memcmp(&x, &y, sizeof(x)) == 0;
}

struct ClassPlan {
    const lfpm_pm_params* p;
    std::vector<int> frames;
    int total = 0;
    size_t scratch_off = 0, scratch_bytes = 0;
};

size_t align_up(size_t x) { return (x + 255) & ~(size_t)255; }

int plan(const lfpm_batch_frame* fr, int n, std::vector<ClassPlan>& classes, size_t& total) {
    if (!fr || n <= 0 || n > LFPM_MAX_BATCH)
        return set_error(LFPM_E_INVALID, "batch: frame count %d outside [1, %d]", n, LFPM_MAX_BATCH);
    classes.clear();
    for (int i = 0; i < n; ++i) {
        const lfpm_batch_frame& f = fr[i];
        if (!f.params || !f.tiles_dev || !f.centers_dev || !f.neighbor_ids_dev ||
            !f.neighbor_counts_dev || f.num_mis <= 0 || !f.plane_dev || !f.cost_dev ||
            !f.sel_dev || !f.plane_b_dev || !f.cost_b_dev || !f.sel_b_dev)
            return set_error(LFPM_E_INVALID, "batch: frame %d is missing inputs or state", i);
        if (f.params->max_neighbors < 1 || f.params->max_neighbors > LFPM_MAX_NEIGHBORS)
            return set_error(LFPM_E_INVALID, "batch: frame %d bad max_neighbors", i);
        int c = -1;
        for (size_t k = 0; k < classes.size(); ++k)
            if (same_class(*classes[k].p, *f.params)) { c = (int)k; break; }
        if (c < 0) {
            classes.push_back(ClassPlan{f.params, {}, 0, 0, 0});
            c = (int)classes.size() - 1;
        }
        classes[c].frames.push_back(i);
        classes[c].total += f.num_mis;
    }
    total = 0;
    for (auto& c : classes) {
        c.scratch_bytes = lfpm::stage_scratch_size(c.p, c.total);
        if (c.scratch_bytes == 0) return set_error(LFPM_E_INVALID, "batch: bad class parameters");
        c.scratch_off = total;
        total += align_up(c.scratch_bytes);
    }
    return LFPM_OK;
}

StageFrame stage_frame(const lfpm_batch_frame& f, int stage, int iter, int phase) {
    StageFrame s;
    std::memset(&s, 0, sizeof(s));
    s.tiles = f.tiles_dev; s.centers = f.centers_dev; s.nbr = f.neighbor_ids_dev;
    s.cnt = f.neighbor_counts_dev; s.V = f.num_mis;
    s.stage = stage; s.iter = iter; s.phase = phase;
    s.seed = f.params->seed; s.uid = f.params->frame_uid;
    if (stage == lfpm::kStageInit) {
        s.plane_out = f.plane_dev; s.cost_out = f.cost_dev; s.sel_out = f.sel_dev;
    } else if (stage == lfpm::kStagePhase) {
        const bool ab = phase == 0;  // phase 0: A -> B, phase 1: B -> A
        s.plane_in = ab ? f.plane_dev : f.plane_b_dev;
        s.cost_in = ab ? f.cost_dev : f.cost_b_dev;
        s.sel_in = ab ? f.sel_dev : f.sel_b_dev;
        s.plane_out = ab ? f.plane_b_dev : f.plane_dev;
        s.cost_out = ab ? f.cost_b_dev : f.cost_dev;
        s.sel_out = ab ? f.sel_b_dev : f.sel_dev;
    } else if (stage == lfpm::kStageFinal) {
        s.plane_in = f.plane_dev; s.disp = f.disp_dev;
    }
    return s;
}

// Filter and refocus of one finished frame.
int frame_products(const lfpm_batch_frame& f, void* stream) {
    int rc;
    if (f.filter) {
        if (!f.gray_dev || !f.f_plane_dev || !f.f_cost_dev || !f.f_disp_dev || !f.f_sel_dev ||
            !f.reason_dev)
            return set_error(LFPM_E_INVALID, "batch: filter requested without buffers");
        if (f.filter_num_levels > 0)
            rc = lfpm_filter_frame_ordered(
                f.gray_dev, f.centers_dev, f.neighbor_ids_dev, f.neighbor_counts_dev, f.num_mis,
                f.params->crop, f.params->max_neighbors, f.params->base, f.filter, f.plane_dev,
                f.cost_dev, f.disp_dev, f.sel_dev, f.f_plane_dev, f.f_cost_dev, f.f_disp_dev,
                f.f_sel_dev, f.reason_dev, f.filter_mi_order_dev, f.filter_level_start,
                f.filter_num_levels, stream);
        else
            rc = lfpm_filter_frame(f.gray_dev, f.centers_dev, f.neighbor_ids_dev,
                                   f.neighbor_counts_dev, f.num_mis, f.params->crop,
                                   f.params->max_neighbors, f.params->base, f.filter,
                                   f.plane_dev, f.cost_dev, f.disp_dev, f.sel_dev, f.f_plane_dev,
                                   f.f_cost_dev, f.f_disp_dev, f.f_sel_dev, f.reason_dev, stream);
        if (rc) return rc;
    }
    if (f.refocus_slices > 0) {
        if (!f.refocus_dev || !f.lattice_to_mi_dev)
            return set_error(LFPM_E_INVALID, "batch: refocus requested without buffers");
        if ((rc = lfpm_refocus(f.tiles_dev, f.centers_dev, f.lattice_to_mi_dev, f.lattice_cols,
                               f.lattice_rows, f.params->base, f.params->crop, f.raw_w, f.raw_h,
                               f.refocus_slices, f.refocus_vmin, f.refocus_vmax, f.refocus_dev,
                               stream)))
            return rc;
    }
    return LFPM_OK;
}

// The whole schedule of one parameter class on `stream`.
int run_class(const lfpm_batch_frame* fr, const ClassPlan& c, unsigned char* ws, void* stream) {
    std::vector<StageFrame> sf(c.frames.size());
    void* scratch = ws + c.scratch_off;
    // each frame's cost records sit at its MI offset in the class scratch; this schedule runs
    // init, the phases and the final stage of every frame in order on that scratch
    std::vector<float*> rec(c.frames.size());
    int mi_off = 0;
    for (size_t i = 0; i < c.frames.size(); ++i) {
        rec[i] = lfpm::stage_records(c.p, scratch, mi_off);
        mi_off += fr[c.frames[i]].num_mis;
    }
    int rc;
    auto step = [&](int stage, int iter, int phase) -> int {
        for (size_t i = 0; i < c.frames.size(); ++i) {
            sf[i] = stage_frame(fr[c.frames[i]], stage, iter, phase);
            sf[i].rec = rec[i];
        }
        return lfpm::stage_launch(c.p, sf.data(), (int)sf.size(), scratch, c.scratch_bytes, stream);
    };
    if ((rc = step(lfpm::kStageInit, 0, 0))) return rc;
    for (int it = 0; it < c.p->iterations; ++it)
        for (int ph = 0; ph < 2; ++ph)
            if ((rc = step(lfpm::kStagePhase, it, ph))) return rc;
    bool need_final = false;
    for (int i : c.frames) need_final |= fr[i].disp_dev != nullptr;
    if (need_final) {
        for (size_t i = 0; i < c.frames.size(); ++i) {
            sf[i] = stage_frame(fr[c.frames[i]], fr[c.frames[i]].disp_dev ? lfpm::kStageFinal
                                                                          : lfpm::kStageIdle, 0, 0);
            sf[i].rec = rec[i];
        }
        if ((rc = lfpm::stage_launch(c.p, sf.data(), (int)sf.size(), scratch, c.scratch_bytes,
                                     stream)))
            return rc;
    }
    for (int i : c.frames)
        if ((rc = frame_products(fr[i], stream))) return rc;
    return LFPM_OK;
}

// Streams on which lfpm_batch_step runs the parameter classes of a step side by side, one pool
// per host thread and device (default priority, non-blocking), with the fork and join events.
constexpr int kStepStreams = 4;
struct StepStreams {
    int device = -1;
    hipStream_t s[kStepStreams] = {};
    hipEvent_t done[kStepStreams] = {};
    hipEvent_t fork = nullptr;
    ~StepStreams() {
        for (int i = 0; i < kStepStreams; ++i) {
            if (done[i]) (void)hipEventDestroy(done[i]);
            if (s[i]) (void)hipStreamDestroy(s[i]);
        }
        if (fork) (void)hipEventDestroy(fork);
    }
};
thread_local std::vector<StepStreams> t_step_streams;

StepStreams* step_streams() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    for (auto& p : t_step_streams)
        if (p.device == dev) return &p;
    t_step_streams.emplace_back();
    StepStreams& p = t_step_streams.back();
    p.device = dev;
    if (hipEventCreateWithFlags(&p.fork, hipEventDisableTiming) != hipSuccess) return nullptr;
    for (int i = 0; i < kStepStreams; ++i)
        if (hipStreamCreateWithFlags(&p.s[i], hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&p.done[i], hipEventDisableTiming) != hipSuccess)
            return nullptr;
    return &p;
}

}  // namespace

struct lfpm_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

extern "C" {

size_t lfpm_batch_workspace_size(const lfpm_batch_frame* frames, int n) {
    std::vector<ClassPlan> classes;
    size_t total = 0;
    if (plan(frames, n, classes, total)) return 0;
    return total;
}

int lfpm_batch_run(const lfpm_batch_frame* frames, int n, void* ws, size_t ws_bytes,
                   void* stream) {
    std::vector<ClassPlan> classes;
    size_t total = 0;
    int rc = plan(frames, n, classes, total);
    if (rc) return rc;
    if (!ws || ws_bytes < total)
        return set_error(LFPM_E_INVALID, "batch: workspace too small (%zu < %zu)", ws_bytes, total);
    for (auto& c : classes)
        if ((rc = run_class(frames, c, (unsigned char*)ws, stream))) return rc;
    return LFPM_OK;
}

int lfpm_batch_graph_create(const lfpm_batch_frame* frames, int n, void* ws, size_t ws_bytes,
                            lfpm_graph** out) {
    if (!out) return set_error(LFPM_E_INVALID, "graph: null out");
    *out = nullptr;
    std::vector<ClassPlan> classes;
    size_t total = 0;
    int rc = plan(frames, n, classes, total);
    if (rc) return rc;
    if (!ws || ws_bytes < total)
        return set_error(LFPM_E_INVALID, "graph: workspace too small (%zu < %zu)", ws_bytes, total);
    hipStream_t s0;
    B_TRY(hipStreamCreateWithFlags(&s0, hipStreamNonBlocking));
    std::vector<hipStream_t> side(classes.size(), nullptr);
    std::vector<hipEvent_t> ev(classes.size() + 1, nullptr);
    auto cleanup = [&]() {
        for (auto s : side)
            if (s) (void)hipStreamDestroy(s);
        for (auto e : ev)
            if (e) (void)hipEventDestroy(e);
        (void)hipStreamDestroy(s0);
    };
    for (size_t i = 0; i < classes.size(); ++i) {
        if (hipStreamCreateWithFlags(&side[i], hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&ev[i], hipEventDisableTiming) != hipSuccess) {
            cleanup();
            return set_error(LFPM_E_HIP, "graph: stream/event creation failed");
        }
    }
    if (hipEventCreateWithFlags(&ev[classes.size()], hipEventDisableTiming) != hipSuccess) {
        cleanup();
        return set_error(LFPM_E_HIP, "graph: event creation failed");
    }
    lfpm_graph* g = new lfpm_graph();
    hipError_t e = hipStreamBeginCapture(s0, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        delete g;
        cleanup();
        return set_error(LFPM_E_HIP, "graph: begin capture: %s", hipGetErrorString(e));
    }
    // fork: every class on its own branch
    hipError_t fe = hipEventRecord(ev[classes.size()], s0);
    for (size_t i = 0; i < classes.size() && fe == hipSuccess; ++i)
        fe = hipStreamWaitEvent(side[i], ev[classes.size()], 0);
    for (size_t i = 0; i < classes.size() && fe == hipSuccess && rc == LFPM_OK; ++i) {
        rc = run_class(frames, classes[i], (unsigned char*)ws, side[i]);
        if (rc == LFPM_OK) fe = hipEventRecord(ev[i], side[i]);
        if (rc == LFPM_OK && fe == hipSuccess) fe = hipStreamWaitEvent(s0, ev[i], 0);
    }
    e = hipStreamEndCapture(s0, &g->graph);
    if (rc == LFPM_OK && (fe != hipSuccess || e != hipSuccess))
        rc = set_error(LFPM_E_HIP, "graph: capture failed: %s",
                       hipGetErrorString(fe != hipSuccess ? fe : e));
    if (rc == LFPM_OK) {
        e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
        if (e != hipSuccess)
            rc = set_error(LFPM_E_HIP, "graph: instantiate: %s", hipGetErrorString(e));
    }
    cleanup();
    if (rc) {
        lfpm_graph_destroy(g);
        return rc;
    }
    *out = g;
    return LFPM_OK;
}

int lfpm_graph_launch(lfpm_graph* g, void* stream) {
    if (!g || !g->exec) return set_error(LFPM_E_INVALID, "graph: not instantiated");
    B_TRY(hipGraphLaunch(g->exec, (hipStream_t)stream));
    return LFPM_OK;
}

void lfpm_graph_destroy(lfpm_graph* g) {
    if (!g) return;
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    delete g;
}

int lfpm_batch_step(const lfpm_batch_frame* frames, int n, const int* stages, const int* iters,
                    const int* phases, void* ws, size_t ws_bytes, void* stream) {
    if (!stages || !iters || !phases) return set_error(LFPM_E_INVALID, "step: null stage arrays");
    std::vector<ClassPlan> classes;
    size_t total = 0;
    int rc = plan(frames, n, classes, total);
    if (rc) return rc;
    if (!ws || ws_bytes < total)
        return set_error(LFPM_E_INVALID, "step: workspace too small (%zu < %zu)", ws_bytes, total);
    for (int i = 0; i < n; ++i) {
        if (stages[i] < LFPM_STAGE_IDLE || stages[i] > LFPM_STAGE_FINAL)
            return set_error(LFPM_E_INVALID, "step: frame %d bad stage %d", i, stages[i]);
        if (stages[i] == LFPM_STAGE_PHASE && (phases[i] < 0 || phases[i] > 1 || iters[i] < 0))
            return set_error(LFPM_E_INVALID, "step: frame %d bad iter/phase", i);
    }
    // several parameter classes: each on a stream of its own after a fork from `stream`, joined
    // back into `stream` at the end, so their launches overlap on the device
    StepStreams* ps = classes.size() > 1 ? step_streams() : nullptr;
    const int nstreams = ps ? std::min((int)classes.size(), kStepStreams) : 0;
    if (ps) {
        if (hipEventRecord(ps->fork, (hipStream_t)stream) != hipSuccess)
            return set_error(LFPM_E_HIP, "step: fork event");
        for (int i = 0; i < nstreams; ++i)
            if (hipStreamWaitEvent(ps->s[i], ps->fork, 0) != hipSuccess)
                return set_error(LFPM_E_HIP, "step: fork wait");
    }
    for (size_t ci = 0; ci < classes.size(); ++ci) {
        auto& c = classes[ci];
        void* cs = ps ? (void*)ps->s[ci % kStepStreams] : stream;
        std::vector<StageFrame> sf;
        for (int i : c.frames) {
            const int st = stages[i] == LFPM_STAGE_FINAL && !frames[i].disp_dev ? LFPM_STAGE_IDLE
                                                                               : stages[i];
            sf.push_back(stage_frame(frames[i], st, iters[i], phases[i]));
        }
        if ((rc = lfpm::stage_launch(c.p, sf.data(), (int)sf.size(),
                                     (unsigned char*)ws + c.scratch_off, c.scratch_bytes, cs)))
            return rc;
        for (int i : c.frames)
            if (stages[i] == LFPM_STAGE_FINAL && (rc = frame_products(frames[i], cs)))
                return rc;
    }
    if (ps) {
        for (int i = 0; i < nstreams; ++i)
            if (hipEventRecord(ps->done[i], ps->s[i]) != hipSuccess ||
                hipStreamWaitEvent((hipStream_t)stream, ps->done[i], 0) != hipSuccess)
                return set_error(LFPM_E_HIP, "step: join");
    }
    return LFPM_OK;
}

int lfpm_autotune_class(const lfpm_batch_frame* frame, void* ws, size_t ws_bytes,
                        int* best_per_cu, float* best_ms) {
    if (!frame) return set_error(LFPM_E_INVALID, "tune: null frame");
    const lfpm_pm_params* p = frame->params;
    // launch shapes: wavefronts per workgroup (threads sharing one MI's staged tiles, and with
    // them the workgroup's LDS footprint) x persistent workgroups per CU
    const int waves[5] = {1, 2, 4, 8, 12};
    const int per_cu[4] = {2, 4, 6, 8};
    hipStream_t s;
    B_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    hipEvent_t e0, e1;
    B_TRY(hipEventCreate(&e0));
    B_TRY(hipEventCreate(&e1));
    int best = -1, bestw = -1, rc = LFPM_OK;
    float bms = 0.0f;
    const int prev = lfpm::stage_slots_per_cu(p), prevw = lfpm::stage_waves(p);
    for (int wi = 0; wi < 5 && rc == LFPM_OK; ++wi)
        for (int ci = 0; ci < 4 && rc == LFPM_OK; ++ci) {
            lfpm::set_stage_config(p, per_cu[ci], waves[wi]);
            const size_t need = lfpm::stage_scratch_size(p, frame->num_mis);
            if (!ws || ws_bytes < need) {
                rc = set_error(LFPM_E_INVALID, "tune: workspace too small (%zu < %zu)", ws_bytes,
                               need);
                break;
            }
            StageFrame sf = stage_frame(*frame, lfpm::kStagePhase, 1, 0);
            // warm-up launch, then one timed launch
            for (int rep = 0; rep < 2 && rc == LFPM_OK; ++rep) {
                if (rep == 1 && hipEventRecord(e0, s) != hipSuccess) rc = LFPM_E_HIP;
                if (rc == LFPM_OK) rc = lfpm::stage_launch(p, &sf, 1, ws, ws_bytes, s);
                if (rep == 1 && rc == LFPM_OK && hipEventRecord(e1, s) != hipSuccess)
                    rc = LFPM_E_HIP;
            }
            if (rc == LFPM_OK && hipEventSynchronize(e1) != hipSuccess) rc = LFPM_E_HIP;
            float ms = 0.0f;
            if (rc == LFPM_OK && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = LFPM_E_HIP;
            if (rc == LFPM_OK && (best < 0 || ms < bms)) {
                best = per_cu[ci];
                bestw = waves[wi];
                bms = ms;
            }
        }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(s);
    if (rc) {
        lfpm::set_stage_config(p, prev, prevw);
        return rc == LFPM_E_HIP ? set_error(LFPM_E_HIP, "tune: HIP event failure") : rc;
    }
    lfpm::set_stage_config(p, best, bestw);
    if (best_per_cu) *best_per_cu = best;
    if (best_ms) *best_ms = bms;
    return LFPM_OK;
}

int lfpm_autotune_get(const lfpm_pm_params* params, int* per_cu, int* waves_per_mi) {
    if (!params) return set_error(LFPM_E_INVALID, "tune: null params");
    if (per_cu) *per_cu = lfpm::stage_slots_per_cu(params);
    if (waves_per_mi) *waves_per_mi = lfpm::stage_waves_used(params);
    return LFPM_OK;
}

}  // extern "C"
