// lfpm_internal.h — declarations shared by the host and device translation units.
#pragma once
#include "../../include/lfpm.h"

namespace lfpm {
int set_error(int code, const char* fmt, ...);
int fixed_neighbors_core(int V, const int* mi_i, const int* mi_j, const float* cx,
                         const float* cy, const int* cell_to_mi, int cols, int rows, int K,
                         int* neighbor_ids, int* neighbor_counts, int* garbage);
const char* last_error();
int resolve_crop(const lfpm_layout_params* p);

// One frame of a k_stage launch (host description; device pointers).
struct StageFrame {
    const float* tiles; const float* centers; const int* nbr; const int* cnt;
    int V;
    int stage, iter, phase;  // STAGE_* below
    uint32_t seed, uid;
    const float* plane_in; const float* cost_in; const uint32_t* sel_in;
    float* plane_out; float* cost_out; uint32_t* sel_out;
    float* disp;
    float* rec;  // cost records of the frame's pixels (wave engine; null: none)
};
enum : int { kStageIdle = 0, kStageInit = 1, kStagePhase = 2, kStageFinal = 3 };

// Scratch bytes one k_stage launch over `total_mis` MIs of parameter class `p` needs.
size_t stage_scratch_size(const lfpm_pm_params* p, int total_mis);
// Cost records of the frame whose MIs start at MI `mi_offset` of a class scratch of
// stage_scratch_size bytes (null when the engine keeps none or scratch is null).  A frame's
// records are written by its init stage and read by its later phase and final stages, so those
// stages must run on the same scratch region.
float* stage_records(const lfpm_pm_params* p, void* scratch, int mi_offset);
// Launch k_stage for `n` frames of one parameter class (chunks of 16 frames per launch).
int stage_launch(const lfpm_pm_params* p, const StageFrame* frames, int n, void* scratch,
                 size_t scratch_bytes, void* stream);
// Workgroups per CU used by k_stage for a parameter class (tuned value or the default), and
// the setter the autotuner uses.
int stage_slots_per_cu(const lfpm_pm_params* p);
void set_stage_slots_per_cu(const lfpm_pm_params* p, int per_cu);
// Wavefronts per k_stage workgroup for a parameter class as configured (0 = chosen at launch:
// the shape that keeps the most wavefronts per CU), the number a launch uses, and the setter of
// both values.
int stage_waves(const lfpm_pm_params* p);
int stage_waves_used(const lfpm_pm_params* p);
void set_stage_config(const lfpm_pm_params* p, int per_cu, int waves);

// Device helpers of lfpm_products.hip.  exclusive_scan: out[i] = sum of in[0..i) for i < n, on
// `stream`, with `sums` of scan_sums_bytes(n) bytes.  median5_passes: `passes` (>= 1) 5x5
// median passes with replicated borders over the raw_w x raw_h plane `in`, ping-ponging through
// t0 / t1 (raw_w*raw_h bytes each; `in` is not written), the last one into out at
// out[(y*raw_w + x) * ostride].
size_t scan_sums_bytes(long long n);
void exclusive_scan(const int* in, long long n, int* out, int* sums, void* stream);
void median5_passes(const uint8_t* in, uint8_t* t0, uint8_t* t1, int raw_w, int raw_h, int passes,
                    uint8_t* out, int ostride, void* stream);
}  // namespace lfpm
