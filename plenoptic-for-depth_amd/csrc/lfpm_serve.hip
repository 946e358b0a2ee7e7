// lfpm_serve.hip — the served form behind the C ABI: a continuous-batching frame server.
//
// lfpm_frame_submit hands a host raw frame (its layout, solve / filter / refocus settings) to a
// server thread that owns the device work; lfpm_frame_wait returns that frame's results.  The
// server admits queued frames into device slots while their bytes fit the HBM budget, keeps
// every admitted frame's tiles, plane / cost / selection state and cost records resident, and
// advances all of them one stage per step (random init, one red/black phase, or the final
// record with the filter and the refocus stack).  Frames of one parameter class (equal solve
// parameters but the random key) share one persistent stage launch per step; each class has
// its own stream, so the classes' launches overlap on the device, and no step waits on the host
// for the one before it (at most two steps per class are in flight).  A frame's results depend
// only on its inputs and (seed, frame_uid), never on what it was batched with, so they equal
// the same frame solved alone.
//
// Replaces the per-frame loop of DepthSolver::ProcessProblemsImpSequence (DepthSolver.cpp:
// 2897-2901) driven by DepthSolverManager::DepthSolverPipelineBySequence / ByBatch
// (DepthSolverManager.cpp:41-135), which solve one frame at a time and reset the device after
// each (DepthSolver.cpp:368-374).
#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/lfpm.h"
#include "lfpm_internal.h"

using lfpm::set_error;

namespace {

double now_ms() {
    using clk = std::chrono::steady_clock;
    return std::chrono::duration<double, std::milli>(clk::now().time_since_epoch()).count();
}

size_t a256(size_t x) { return (x + 255) & ~(size_t)255; }

bool same_class(const lfpm_pm_params& a, const lfpm_pm_params& b) {
    lfpm_pm_params x = a, y = b;
    x.seed = y.seed = 0;
    x.frame_uid = y.frame_uid = 0;
    return std::memcmp(&x, &y, sizeof(x)) == 0;
}

struct Stage {
    int stage, iter, phase;
};

// One submitted frame.
struct Job {
    uint64_t ticket = 0;
    lfpm_layout_params lp{};
    lfpm_pm_params pm{};
    lfpm_filter_params fp{};
    bool filt = false;
    bool ordered = false;                // filter in the reference's MI order
    int slices = 0;
    float vmin = 0.0f, vmax = 0.0f;
    int raw_w = 0, raw_h = 0, V = 0, crop = 0, K = 0, cols = 0, rows = 0;
    // host tables of the layout
    std::vector<float> centers;
    std::vector<int> origins, nbr, cnt, l2m;
    std::vector<int> lvl_order, lvl_start;  // level schedule of the ordered filter
    uint8_t* raw_pinned = nullptr;       // the raw frame, copied at submit
    // device block and its buffers
    void* dev = nullptr;
    size_t dev_bytes = 0;
    uint8_t* raw_d = nullptr;
    float *tiles = nullptr, *gray = nullptr, *centers_d = nullptr;
    int *origins_d = nullptr, *nbr_d = nullptr, *cnt_d = nullptr, *l2m_d = nullptr;
    int* lvl_order_d = nullptr;
    float *plane = nullptr, *cost = nullptr, *disp = nullptr, *rec = nullptr;
    uint32_t* sel = nullptr;
    float *plane_b = nullptr, *cost_b = nullptr;
    uint32_t* sel_b = nullptr;
    float *fplane = nullptr, *fcost = nullptr, *fdisp = nullptr, *refocus = nullptr;
    uint32_t* fsel = nullptr;
    uint8_t* reason = nullptr;
    // pinned result staging: plane, cost, disp, sel, reason, refocus
    uint8_t* res_pinned = nullptr;
    size_t res_bytes = 0;
    // progress
    int cls = -1;
    std::vector<Stage> sched;
    size_t pos = 0;
    hipEvent_t done_ev = nullptr;
    bool issued = false;  // every stage and the result copies are on the stream
    bool done = false;
    int rc = LFPM_OK;
    std::string err;
    double t_submit = 0.0, t_admit = 0.0, t_done = 0.0;
};

struct Class {
    lfpm_pm_params pm{};
    int stream = 0;
    void* tickets = nullptr;  // 256-byte stage scratch header (per-XCD MI tickets)
    hipEvent_t ev[2] = {nullptr, nullptr};
    int ev_next = 0;
    bool ev_used[2] = {false, false};
};

// Device bytes of one frame's slot (inputs, state, records, filtered copies, refocus stack).
size_t slot_bytes(const Job& j) {
    const size_t n = (size_t)j.V * j.crop * j.crop;
    size_t b = 0;
    b += a256((size_t)j.raw_w * j.raw_h);
    b += 2 * a256(n * 4);                                // tiles, gray
    b += a256((size_t)j.V * 8) + a256((size_t)j.V * 8);  // centers, origins
    b += a256((size_t)j.V * j.K * 4) + a256((size_t)j.V * 4);
    b += a256((size_t)j.cols * j.rows * 4);
    b += 2 * (a256(n * 16) + a256(n * 4) + a256(n * 4));  // A and B state
    b += a256(n * 16);                                     // disp
    b += a256(n * (5 + (size_t)j.K) * 4);                  // cost records
    if (j.filt) b += a256(n * 16) * 2 + a256(n * 4) * 2 + a256(n);
    if (j.ordered) b += a256((size_t)j.V * 4);
    b += a256((size_t)j.slices * j.raw_w * j.raw_h * 4);
    return b;
}

size_t result_bytes(const Job& j) {
    const size_t n = (size_t)j.V * j.crop * j.crop;
    return n * (16 + 4 + 16 + 4) + (j.filt ? n : 0) + (size_t)j.slices * j.raw_w * j.raw_h * 4;
}

}  // namespace

struct lfpm_server {
    int device = 0;
    size_t budget = 0;
    int max_slots = 16;
    int filter_order = LFPM_FILTER_SNAPSHOT;
    std::vector<hipStream_t> streams;
    std::thread worker;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    bool stop = false;
    uint64_t next_ticket = 1;
    std::deque<Job*> queue;
    std::vector<Job*> active;   // admitted, not yet finished
    std::unordered_map<uint64_t, Job*> jobs;
    size_t in_use = 0;
    std::vector<Class> classes;
    lfpm_server_stats stats{};
};

namespace {

#define S_TRY(expr)                                                                       \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess) {                                                           \
            j->err = std::string(#expr) + ": " + hipGetErrorString(e_);                   \
            return LFPM_E_HIP;                                                            \
        }                                                                                 \
    } while (0)

void release_device(Job* j) {
    if (j->dev) (void)hipFree(j->dev);
    j->dev = nullptr;
    if (j->done_ev) (void)hipEventDestroy(j->done_ev);
    j->done_ev = nullptr;
}

void release_host(Job* j) {
    if (j->raw_pinned) (void)hipHostFree(j->raw_pinned);
    j->raw_pinned = nullptr;
    if (j->res_pinned) (void)hipHostFree(j->res_pinned);
    j->res_pinned = nullptr;
}

// Allocate the slot, upload the frame and its tables, slice tiles and gray crops.
int admit(Job* j, hipStream_t st) {
    j->dev_bytes = slot_bytes(*j);
    S_TRY(hipMalloc(&j->dev, j->dev_bytes));
    S_TRY(hipEventCreateWithFlags(&j->done_ev, hipEventDisableTiming));
    unsigned char* q = (unsigned char*)j->dev;
    const size_t n = (size_t)j->V * j->crop * j->crop;
    auto take = [&](size_t bytes) {
        unsigned char* p = q;
        q += a256(bytes);
        return (void*)p;
    };
    j->raw_d = (uint8_t*)take((size_t)j->raw_w * j->raw_h);
    j->tiles = (float*)take(n * 4);
    j->gray = (float*)take(n * 4);
    j->centers_d = (float*)take((size_t)j->V * 8);
    j->origins_d = (int*)take((size_t)j->V * 8);
    j->nbr_d = (int*)take((size_t)j->V * j->K * 4);
    j->cnt_d = (int*)take((size_t)j->V * 4);
    j->l2m_d = (int*)take((size_t)j->cols * j->rows * 4);
    j->plane = (float*)take(n * 16);
    j->cost = (float*)take(n * 4);
    j->sel = (uint32_t*)take(n * 4);
    j->plane_b = (float*)take(n * 16);
    j->cost_b = (float*)take(n * 4);
    j->sel_b = (uint32_t*)take(n * 4);
    j->disp = (float*)take(n * 16);
    j->rec = (float*)take(n * (5 + (size_t)j->K) * 4);
    if (j->filt) {
        j->fplane = (float*)take(n * 16);
        j->fdisp = (float*)take(n * 16);
        j->fcost = (float*)take(n * 4);
        j->fsel = (uint32_t*)take(n * 4);
        j->reason = (uint8_t*)take(n);
    }
    if (j->ordered) j->lvl_order_d = (int*)take((size_t)j->V * 4);
    if (j->slices > 0) j->refocus = (float*)take((size_t)j->slices * j->raw_w * j->raw_h * 4);
    S_TRY(hipMemcpyAsync(j->raw_d, j->raw_pinned, (size_t)j->raw_w * j->raw_h,
                         hipMemcpyHostToDevice, st));
    S_TRY(hipMemcpyAsync(j->centers_d, j->centers.data(), j->centers.size() * 4,
                         hipMemcpyHostToDevice, st));
    S_TRY(hipMemcpyAsync(j->origins_d, j->origins.data(), j->origins.size() * 4,
                         hipMemcpyHostToDevice, st));
    S_TRY(hipMemcpyAsync(j->nbr_d, j->nbr.data(), j->nbr.size() * 4, hipMemcpyHostToDevice, st));
    S_TRY(hipMemcpyAsync(j->cnt_d, j->cnt.data(), j->cnt.size() * 4, hipMemcpyHostToDevice, st));
    S_TRY(hipMemcpyAsync(j->l2m_d, j->l2m.data(), j->l2m.size() * 4, hipMemcpyHostToDevice, st));
    if (j->ordered)
        S_TRY(hipMemcpyAsync(j->lvl_order_d, j->lvl_order.data(), j->lvl_order.size() * 4,
                             hipMemcpyHostToDevice, st));
    int rc;
    if ((rc = lfpm_slice_tiles(j->raw_d, j->raw_w, j->raw_h, (size_t)j->raw_w, j->origins_d, j->V,
                               j->crop, j->tiles, st)) ||
        (rc = lfpm_slice_gray(j->raw_d, j->raw_w, j->raw_h, (size_t)j->raw_w, j->origins_d, j->V,
                              j->crop, j->gray, st))) {
        j->err = lfpm_last_error();
        return rc;
    }
    j->sched.clear();
    j->sched.push_back({lfpm::kStageInit, 0, 0});
    for (int it = 0; it < j->pm.iterations; ++it) {
        j->sched.push_back({lfpm::kStagePhase, it, 0});
        j->sched.push_back({lfpm::kStagePhase, it, 1});
    }
    j->sched.push_back({lfpm::kStageFinal, 0, 0});
    j->pos = 0;
    return LFPM_OK;
}

lfpm::StageFrame stage_of(const Job* j) {
    lfpm::StageFrame f;
    std::memset(&f, 0, sizeof(f));
    const Stage& g = j->sched[j->pos];
    f.tiles = j->tiles; f.centers = j->centers_d; f.nbr = j->nbr_d; f.cnt = j->cnt_d;
    f.V = j->V;
    f.stage = g.stage; f.iter = g.iter; f.phase = g.phase;
    f.seed = j->pm.seed; f.uid = j->pm.frame_uid;
    f.rec = j->rec;
    if (g.stage == lfpm::kStageInit) {
        f.plane_out = j->plane; f.cost_out = j->cost; f.sel_out = j->sel;
    } else if (g.stage == lfpm::kStagePhase) {
        const bool ab = g.phase == 0;  // phase 0: A -> B, phase 1: B -> A
        f.plane_in = ab ? j->plane : j->plane_b;
        f.cost_in = ab ? j->cost : j->cost_b;
        f.sel_in = ab ? j->sel : j->sel_b;
        f.plane_out = ab ? j->plane_b : j->plane;
        f.cost_out = ab ? j->cost_b : j->cost;
        f.sel_out = ab ? j->sel_b : j->sel;
    } else {
        f.plane_in = j->plane;
        f.disp = j->disp;
    }
    return f;
}

// Filter, refocus and the result copies of a frame whose final stage is on `st`, then its
// completion event.
int finish(Job* j, hipStream_t st) {
    int rc;
    const size_t n = (size_t)j->V * j->crop * j->crop;
    if (j->filt) {
        if (j->ordered)
            rc = lfpm_filter_frame_ordered(j->gray, j->centers_d, j->nbr_d, j->cnt_d, j->V,
                                           j->crop, j->K, j->pm.base, &j->fp, j->plane, j->cost,
                                           j->disp, j->sel, j->fplane, j->fcost, j->fdisp,
                                           j->fsel, j->reason, j->lvl_order_d,
                                           j->lvl_start.data(), (int)j->lvl_start.size() - 1, st);
        else
            rc = lfpm_filter_frame(j->gray, j->centers_d, j->nbr_d, j->cnt_d, j->V, j->crop,
                                   j->K, j->pm.base, &j->fp, j->plane, j->cost, j->disp, j->sel,
                                   j->fplane, j->fcost, j->fdisp, j->fsel, j->reason, st);
        if (rc) {
            j->err = lfpm_last_error();
            return rc;
        }
    }
    if (j->slices > 0 &&
        (rc = lfpm_refocus(j->tiles, j->centers_d, j->l2m_d, j->cols, j->rows, j->lp.pitch,
                           j->crop, j->raw_w, j->raw_h, j->slices, j->vmin, j->vmax, j->refocus,
                           st))) {
        j->err = lfpm_last_error();
        return rc;
    }
    uint8_t* h = j->res_pinned;
    const void* src[5] = {j->filt ? j->fplane : j->plane, j->filt ? j->fcost : j->cost,
                          j->filt ? j->fdisp : j->disp, j->filt ? (void*)j->fsel : (void*)j->sel,
                          j->reason};
    const size_t sz[5] = {n * 16, n * 4, n * 16, n * 4, j->filt ? n : 0};
    for (int k = 0; k < 5; ++k) {
        if (sz[k]) S_TRY(hipMemcpyAsync(h, src[k], sz[k], hipMemcpyDeviceToHost, st));
        h += sz[k];
    }
    if (j->slices > 0)
        S_TRY(hipMemcpyAsync(h, j->refocus, (size_t)j->slices * j->raw_w * j->raw_h * 4,
                             hipMemcpyDeviceToHost, st));
    S_TRY(hipEventRecord(j->done_ev, st));
    return LFPM_OK;
}

void complete(lfpm_server* s, Job* j, int rc) {
    // with s->mu held
    j->rc = rc;
    j->done = true;
    j->t_done = now_ms();
    release_device(j);
    s->in_use -= std::min(s->in_use, j->dev_bytes);
    s->stats.completed += 1;
    s->cv_done.notify_all();
}

int class_of(lfpm_server* s, const lfpm_pm_params& p) {
    for (size_t c = 0; c < s->classes.size(); ++c)
        if (same_class(s->classes[c].pm, p)) return (int)c;
    Class c;
    c.pm = p;
    c.stream = (int)(s->classes.size() % s->streams.size());
    if (hipMalloc(&c.tickets, 256) != hipSuccess) return -1;
    for (int k = 0; k < 2; ++k)
        if (hipEventCreateWithFlags(&c.ev[k], hipEventDisableTiming) != hipSuccess) return -1;
    s->classes.push_back(c);
    return (int)s->classes.size() - 1;
}

void worker(lfpm_server* s) {
    (void)hipSetDevice(s->device);
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        s->cv_work.wait(lk, [&] { return s->stop || !s->queue.empty() || !s->active.empty(); });
        if (s->queue.empty() && s->active.empty()) {
            if (s->stop) break;
            continue;
        }
        // admission by HBM budget and slot count
        std::vector<Job*> admitted;
        while (!s->queue.empty() && (int)s->active.size() + (int)admitted.size() < s->max_slots) {
            Job* j = s->queue.front();
            const size_t need = slot_bytes(*j);
            if (s->in_use + need > s->budget) {
                if (s->active.empty() && admitted.empty()) {
                    s->queue.pop_front();
                    j->err = "frame needs " + std::to_string(need) + " device bytes, budget " +
                             std::to_string(s->budget);
                    j->dev_bytes = 0;
                    complete(s, j, LFPM_E_NOMEM);
                    continue;
                }
                break;
            }
            s->queue.pop_front();
            s->in_use += need;
            admitted.push_back(j);
        }
        lk.unlock();
        for (Job* j : admitted) {
            j->t_admit = now_ms();
            j->cls = class_of(s, j->pm);
            int rc = j->cls < 0 ? LFPM_E_HIP : LFPM_OK;
            if (rc) j->err = "class stream setup failed";
            if (rc == LFPM_OK)
                rc = admit(j, s->streams[s->classes[j->cls].stream]);
            if (rc) {
                lk.lock();
                j->dev_bytes = slot_bytes(*j);
                complete(s, j, rc);
                lk.unlock();
                continue;
            }
            lk.lock();
            s->active.push_back(j);
            s->stats.max_active = std::max(s->stats.max_active, (int)s->active.size());
            lk.unlock();
        }
        lk.lock();
        std::vector<Job*> act = s->active;
        lk.unlock();
        // one step: every frame with stages left advances one stage, one launch per class on
        // the class stream; at most two steps per class in flight
        bool stepped = false;
        for (size_t c = 0; c < s->classes.size(); ++c) {
            Class& cl = s->classes[c];
            std::vector<Job*> fr;
            for (Job* j : act)
                if (j->cls == (int)c && !j->issued) fr.push_back(j);
            if (fr.empty()) continue;
            hipStream_t st = s->streams[cl.stream];
            // throttle: wait for the step before the previous one of this class
            hipEvent_t ev = cl.ev[cl.ev_next];
            if (cl.ev_used[cl.ev_next]) (void)hipEventSynchronize(ev);
            std::vector<lfpm::StageFrame> sf;
            for (Job* j : fr) sf.push_back(stage_of(j));
            int rc = lfpm::stage_launch(&cl.pm, sf.data(), (int)sf.size(), cl.tickets, 256, st);
            std::string err = rc ? lfpm_last_error() : "";
            s->stats.launches += 1;
            for (Job* j : fr) {
                if (rc == LFPM_OK && j->sched[j->pos].stage == lfpm::kStageFinal) {
                    const int frc = finish(j, st);
                    if (frc) {
                        lk.lock();
                        complete(s, j, frc);
                        lk.unlock();
                        continue;
                    }
                    j->issued = true;
                }
                if (rc) {
                    j->err = err;
                    lk.lock();
                    complete(s, j, rc);
                    lk.unlock();
                    continue;
                }
                ++j->pos;
            }
            (void)hipEventRecord(ev, st);
            cl.ev_used[cl.ev_next] = true;
            cl.ev_next ^= 1;
            stepped = true;
        }
        if (stepped) s->stats.steps += 1;
        // completions
        lk.lock();
        std::vector<Job*> still;
        bool any_pending = false;
        for (Job* j : s->active) {
            if (j->done) continue;
            if (j->issued) {
                const hipError_t q = hipEventQuery(j->done_ev);
                if (q == hipSuccess) {
                    complete(s, j, LFPM_OK);
                    continue;
                }
                if (q != hipErrorNotReady) {
                    j->err = std::string("completion: ") + hipGetErrorString(q);
                    complete(s, j, LFPM_E_HIP);
                    continue;
                }
                any_pending = true;
            }
            still.push_back(j);
        }
        s->active.swap(still);
        if (!stepped && any_pending) {
            // nothing to launch until a frame completes or a request arrives: poll the
            // completions every 100 us, woken early by a submit
            s->cv_work.wait_for(lk, std::chrono::microseconds(100));
        }
    }
}

}  // namespace

extern "C" {

int lfpm_server_create(int device, const lfpm_serve_opts* opts, lfpm_server** out) {
    if (!out) return set_error(LFPM_E_INVALID, "server: null out");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev)
        return set_error(LFPM_E_HIP, "server: no HIP device %d", device);
    if (hipSetDevice(device) != hipSuccess) return set_error(LFPM_E_HIP, "server: set device");
    lfpm_server* s = new lfpm_server();
    s->device = device;
    size_t freeb = 0, totalb = 0;
    (void)hipMemGetInfo(&freeb, &totalb);
    s->budget = opts && opts->hbm_budget_bytes ? opts->hbm_budget_bytes : freeb / 2;
    s->max_slots = opts && opts->max_slots > 0 ? opts->max_slots : 16;
    const int ns = opts && opts->streams > 0 ? opts->streams : 4;
    if (opts && opts->filter_order != LFPM_FILTER_SNAPSHOT &&
        opts->filter_order != LFPM_FILTER_REFERENCE_ORDER) {
        delete s;
        return set_error(LFPM_E_INVALID, "server: unknown filter order %d", opts->filter_order);
    }
    s->filter_order = opts ? opts->filter_order : LFPM_FILTER_SNAPSHOT;
    for (int i = 0; i < ns; ++i) {
        hipStream_t st;
        if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
            for (auto x : s->streams) (void)hipStreamDestroy(x);
            delete s;
            return set_error(LFPM_E_HIP, "server: stream creation failed");
        }
        s->streams.push_back(st);
    }
    s->worker = std::thread(worker, s);
    *out = s;
    return LFPM_OK;
}

void lfpm_server_destroy(lfpm_server* s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    if (s->worker.joinable()) s->worker.join();
    (void)hipSetDevice(s->device);
    for (auto& kv : s->jobs) {
        release_device(kv.second);
        release_host(kv.second);
        delete kv.second;
    }
    for (auto& c : s->classes) {
        if (c.tickets) (void)hipFree(c.tickets);
        for (auto e : c.ev)
            if (e) (void)hipEventDestroy(e);
    }
    for (auto st : s->streams) (void)hipStreamDestroy(st);
    delete s;
}

int lfpm_frame_submit(lfpm_server* s, const lfpm_frame_request* r, uint64_t* ticket) {
    if (!s || !r || !ticket || !r->layout || !r->raw_host || !r->params)
        return set_error(LFPM_E_INVALID, "submit: null argument");
    lfpm_layout_info info;
    int rc = lfpm_layout_query(r->layout, &info);
    if (rc) return rc;
    if (r->params->crop != info.crop || r->params->max_neighbors != r->layout->max_neighbors)
        return set_error(LFPM_E_INVALID, "submit: params crop/max_neighbors (%d, %d) differ from "
                         "the layout's (%d, %d)", r->params->crop, r->params->max_neighbors,
                         info.crop, r->layout->max_neighbors);
    if (r->params->storage != 0 && r->params->storage != 1)
        return set_error(LFPM_E_UNSUPPORTED, "submit: storage %d", r->params->storage);
    if (r->params->blur_lambda != 0.0f)
        return set_error(LFPM_E_UNSUPPORTED, "submit: the served form has no blur term");
    if (r->refocus_slices < 0 || (r->refocus_slices > 0 && !(r->refocus_vmin > 0.0f &&
                                                              r->refocus_vmax >= r->refocus_vmin)))
        return set_error(LFPM_E_INVALID, "submit: bad refocus range");
    const size_t pitch = r->raw_pitch ? r->raw_pitch : (size_t)r->layout->raw_w;
    if (pitch < (size_t)r->layout->raw_w) return set_error(LFPM_E_INVALID, "submit: raw pitch");
    Job* j = new Job();
    j->lp = *r->layout;
    j->pm = *r->params;
    j->filt = r->filter != nullptr;
    if (j->filt) j->fp = *r->filter;
    j->slices = r->refocus_slices;
    j->vmin = r->refocus_vmin;
    j->vmax = r->refocus_vmax;
    j->raw_w = r->layout->raw_w;
    j->raw_h = r->layout->raw_h;
    j->V = info.num_mis;
    j->crop = info.crop;
    j->K = r->layout->max_neighbors;
    j->cols = info.lattice_cols;
    j->rows = info.lattice_rows;
    j->centers.resize((size_t)2 * j->V);
    j->origins.resize((size_t)2 * j->V);
    j->nbr.resize((size_t)j->V * j->K);
    j->cnt.resize(j->V);
    j->l2m.resize((size_t)j->cols * j->rows);
    if ((rc = lfpm_layout_build(r->layout, j->centers.data(), j->origins.data(), nullptr, nullptr,
                                j->nbr.data(), j->cnt.data(), j->l2m.data()))) {
        delete j;
        return rc;
    }
    j->ordered = j->filt && s->filter_order == LFPM_FILTER_REFERENCE_ORDER;
    if (j->ordered) {
        j->lvl_order.resize(j->V);
        j->lvl_start.resize((size_t)j->V + 1);
        int levels = 0;
        if ((rc = lfpm_filter_order_levels(j->nbr.data(), j->cnt.data(), j->V, j->K,
                                           j->lvl_order.data(), j->lvl_start.data(),
                                           &levels))) {
            delete j;
            return rc;
        }
        j->lvl_start.resize((size_t)levels + 1);
    }
    if (hipSetDevice(s->device) != hipSuccess ||
        hipHostMalloc((void**)&j->raw_pinned, (size_t)j->raw_w * j->raw_h) != hipSuccess) {
        delete j;
        return set_error(LFPM_E_NOMEM, "submit: pinned staging");
    }
    for (int y = 0; y < j->raw_h; ++y)
        std::memcpy(j->raw_pinned + (size_t)y * j->raw_w, r->raw_host + (size_t)y * pitch,
                    (size_t)j->raw_w);
    j->res_bytes = result_bytes(*j);
    if (hipHostMalloc((void**)&j->res_pinned, j->res_bytes) != hipSuccess) {
        release_host(j);
        delete j;
        return set_error(LFPM_E_NOMEM, "submit: pinned result staging");
    }
    {
        std::lock_guard<std::mutex> lk(s->mu);
        j->ticket = s->next_ticket++;
        j->t_submit = now_ms();
        s->jobs[j->ticket] = j;
        s->queue.push_back(j);
        s->stats.submitted += 1;
    }
    s->cv_work.notify_all();
    *ticket = j->ticket;
    return LFPM_OK;
}

int lfpm_frame_wait(lfpm_server* s, uint64_t ticket, lfpm_frame_result* out) {
    if (!s) return set_error(LFPM_E_INVALID, "wait: null server");
    Job* j = nullptr;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        auto it = s->jobs.find(ticket);
        if (it == s->jobs.end()) return set_error(LFPM_E_INVALID, "wait: unknown ticket %llu",
                                                  (unsigned long long)ticket);
        j = it->second;
        s->cv_done.wait(lk, [&] { return j->done; });
        s->jobs.erase(it);
    }
    const int rc = j->rc;
    if (rc == LFPM_OK && out) {
        const size_t n = (size_t)j->V * j->crop * j->crop;
        const uint8_t* h = j->res_pinned;
        void* dst[5] = {out->plane, out->cost, out->disp, out->sel, out->reason};
        const size_t sz[5] = {n * 16, n * 4, n * 16, n * 4, j->filt ? n : 0};
        for (int k = 0; k < 5; ++k) {
            if (sz[k] && dst[k]) std::memcpy(dst[k], h, sz[k]);
            h += sz[k];
        }
        if (j->slices > 0 && out->refocus)
            std::memcpy(out->refocus, h, (size_t)j->slices * j->raw_w * j->raw_h * 4);
        out->num_mis = j->V;
        out->crop = j->crop;
        out->latency_ms = j->t_done - j->t_submit;
        out->queue_ms = j->t_admit > 0.0 ? j->t_admit - j->t_submit : 0.0;
    }
    const std::string err = j->err;
    release_host(j);
    delete j;
    if (rc) return set_error(rc, "served frame: %s", err.c_str());
    return LFPM_OK;
}

int lfpm_server_get_stats(lfpm_server* s, lfpm_server_stats* out) {
    if (!s || !out) return set_error(LFPM_E_INVALID, "stats: null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    *out = s->stats;
    out->classes = (int)s->classes.size();
    out->in_use_bytes = s->in_use;
    out->budget_bytes = s->budget;
    return LFPM_OK;
}

}  // extern "C"
