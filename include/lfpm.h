/*
 * lfpm.h — C ABI of the MI355X-native micro-image PatchMatch depth solver.
 *
 * Conventions
 *  - Every entry point returns an int status: LFPM_OK (0) or a negative LFPM_E* code;
 *    lfpm_last_error() gives the message of the last failure on the calling thread.
 *  - "_dev" pointers are device (HBM) pointers on the current HIP device; all other pointers
 *    are host pointers.  `stream` is a hipStream_t passed as void* (NULL = default stream).
 *    Device entry points are asynchronous on `stream` unless stated otherwise.
 *  - A micro-image (MI) is a crop of crop x crop pixels of the raw frame.  Per-pixel arrays
 *    are MI-major: element (v, y, x) sits at v*crop*crop + y*crop + x.
 *  - Planes are float4 (nx, ny, nz, w) where w is the disparity at the pixel that owns the
 *    plane; disparity d at local pixel q is w + (-nx/nz)*(q.x-p.x) + (-ny/nz)*(q.y-p.y).
 *  - Correspondence between a reference MI r and neighbour n: local pixel p maps to
 *    p + (c_r - c_n) * d / base, with c the MI centres in raw pixels.
 *
 * Each declaration cites the reference interface it replaces (paths relative to the
 * reference repository root).
 */
#ifndef LFPM_H
#define LFPM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LFPM_OK 0
#define LFPM_E_INVALID (-1)   /* bad argument / shape */
#define LFPM_E_HIP (-2)       /* HIP runtime error */
#define LFPM_E_NOMEM (-3)     /* allocation / budget failure */
#define LFPM_E_UNSUPPORTED (-4)

#define LFPM_MAX_NEIGHBORS 32
#define LFPM_MAX_CROP 32
#define LFPM_MAX_RADIUS 10

/* Filter reject reasons (same codes as MIDisparityFilterCPU::PixelRejectReason,
 * CheckModule/MIDisparityFilter.h:102-109). */
#define LFPM_PR_KEEP 0
#define LFPM_PR_COST 80
#define LFPM_PR_NOSUPPORT 140
#define LFPM_PR_CYCLE 200
#define LFPM_PR_SPIKE 255

/* ---------------------------------------------------------------------------------------
 * Lens layout (hex micro-lens grid) and neighbour tables.
 * Replaces DepthSolver::CreateProblemsAndDisNormals_Frame crop/abandon rule
 * (DepthSolver.cpp:1880-1915), LightFieldParams::ComputeMIA_Math_Info (Common/Common.cpp:34-38),
 * SelectNeighbors::CollectMIANeighImagesForMatch + SortNeighScoreForMatch
 * (MINeighbor/SelectNeighbors.cpp:113-193; MVStereo/LFDepthInfo.cpp:214-241) and
 * AdaptMIPMFrame::BuildViewIndexMap (MIStereo/AdaptMIPMFrame.cpp:145-222).
 * ------------------------------------------------------------------------------------- */
typedef struct lfpm_layout_params {
    float pitch;               /* lens pitch = MI diameter in raw pixels (reference Base/baseline) */
    int crop;                  /* MI crop side; <= 0 selects floor(pitch * 0.707106781186) */
    int raw_w, raw_h;          /* raw frame size in pixels */
    int max_neighbors;         /* K: row stride of the neighbour table (<= LFPM_MAX_NEIGHBORS) */
    int lens_types;            /* 1 = single lens type, 3 = multi-focus 3-type hex pattern */
    int neighbors_per_type[3]; /* neighbour cap per lens type (each <= max_neighbors) */
} lfpm_layout_params;

typedef struct lfpm_layout_info {
    int num_mis;               /* V: kept micro-images */
    int crop;                  /* resolved crop side */
    int lattice_cols, lattice_rows;
    float row_step;            /* pitch * sqrt(3)/2 */
} lfpm_layout_info;

/* Sizes of a layout (host only, no GPU). */
int lfpm_layout_query(const lfpm_layout_params* p, lfpm_layout_info* out);

/* Fill the layout arrays (host only).  Sizes: centers 2V floats, origins 2V ints,
 * tilekeys 2V ints (lattice col,row), lens_type V ints, neighbor_ids V*K ints (-1 padded),
 * neighbor_counts V ints, lattice_to_mi cols*rows ints (-1 where no kept MI).
 * Any output pointer may be NULL.  MI order = (row, col), the QuadTreeTileKey order
 * (Common/QuadTree.h:137-148).  Neighbours are ordered by distance, then angle. */
int lfpm_layout_build(const lfpm_layout_params* p, float* centers, int* origins, int* tilekeys,
                      int* lens_type, int* neighbor_ids, int* neighbor_counts,
                      int* lattice_to_mi);

/* FixedPosition neighbour table (SelectNeighbors::FixedPositionAlgorithm with
 * SelectFixedPositionMIByMode1, MINeighbor/SelectNeighbors.cpp:1035-1286): for each kept MI of
 * the layout (lfpm_layout_build order), the mode-1 key offsets of its place in the lens
 * lattice (interior: 12; edges and corners: 6) that name kept MIs, sorted by centre distance
 * (MLA_Problem::CreateResTileKeysFromNeiKeyVec, MVStereo/LFDepthInfo.cpp:431-453; equal
 * distances keep the collection order), capped at max_neighbors.  garbage[v] = 1 where
 * DepthSolver::ConfirmProblemForEstimation (DepthSolver.cpp:2933-2946) excludes the MI from
 * estimation (key column or row 0, or no neighbour).  Host only; any output may be NULL.
 * Sizes: neighbor_ids V*K, neighbor_counts V, garbage V. */
int lfpm_layout_fixed_neighbors(const lfpm_layout_params* p, int* neighbor_ids,
                                int* neighbor_counts, int* garbage);

/* ---------------------------------------------------------------------------------------
 * Calibration ingest (host only).
 *
 * lfpm_read_intrinsics_xml replaces DepthSolver::ReadIntrinsicsFromXML (DepthSolver.cpp:
 * 872-1110): root element "_"; D -> bl0 (g_bl0), d -> B (g_B), sensor/pose/rotation (three "_"
 * columns of three values, stored here row-major), sensor/pose/translation/_ (three values),
 * sensor/scale (mm per pixel), main_lens/f, mia/mesh/{width, height, pitch/_/...} (the last
 * value under pitch wins, as the reference's loop assigns each in turn), distortions/{depth,
 * radial, tangential}/_ values.  Numbers are read with the reference's sscanf formats; missing
 * tags leave 0.
 *
 * lfpm_read_calib_xml replaces DepthSolver::ReadCalibrationParamsFromXML (:1112-1251): root
 * "Calib"; BlurInfo/MLAType id="t" -> blur_radius[t] (t in 0..2, 0 when absent); MLAInfo/MLA
 * {key "row,col" (%i: a leading 0 reads octal, 0x hex), center "x,y", MLAType, aera} per lens,
 * returned in (row, col) order (the MLA map's QuadTreeTileKey order; a repeated key keeps its
 * last entry, as the map assignment of :1242 does).  tilekeys: 2*count ints (col, row) as
 * lfpm_layout_build writes them.  With every array NULL only count (and blur_radius) is
 * returned; otherwise capacity must hold count.  blur_radius holds the BlurInfo values wherever
 * the group sits in the document; the reference gives each lens the radius its type has when
 * that MLA element is parsed, so a BlurInfo group placed after MLAInfo leaves its lenses at 0
 * (a lens's radius is blur_radius[type] only when BlurInfo comes first, as in the files the
 * reference's calibration writes).
 *
 * lfpm_table_fixed_neighbors is lfpm_layout_fixed_neighbors over such a table: the lens
 * lattice is lattice_cols x lattice_rows keys (the intrinsics' mla_u_size x mla_v_size); keys
 * must lie inside it and be distinct.
 * ------------------------------------------------------------------------------------- */
typedef struct lfpm_intrinsics {
    float D;                         /* main lens to MLA distance (g_bl0), mm */
    float d;                         /* MLA to sensor distance (g_B), mm */
    float sensor_pixel_size;         /* sensor/scale, mm per pixel */
    float main_lens_f;               /* main_lens/f, mm */
    int mla_u_size, mla_v_size;      /* mia/mesh width, height: lens lattice columns, rows */
    float baseline;                  /* mia/mesh/pitch: lens pitch, pixels */
    double sensor_rotation[9];       /* row-major 3x3 */
    double sensor_translation[3];
    double distortion_depth[3], distortion_radial[3], distortion_tangential[2];
} lfpm_intrinsics;

int lfpm_read_intrinsics_xml(const char* path, lfpm_intrinsics* out);
int lfpm_read_calib_xml(const char* path, int capacity, int* tilekeys, float* centers,
                        int* lens_type, float* areas, float* blur_radius, int* count);
int lfpm_table_fixed_neighbors(const int* tilekeys, const float* centers, int num_mis,
                               int lattice_cols, int lattice_rows, int max_neighbors,
                               int* neighbor_ids, int* neighbor_counts, int* garbage);

/* Matching neighbour table of a read lens table (SelectNeighbors::CollectMIANeighImagesForMatch
 * -> ComputeNeighborScoresForMatch -> MLA_Problem::SortNeighScoreForMatch,
 * MINeighbor/SelectNeighbors.cpp:113-193,770-830, MVStereo/LFDepthInfo.cpp:214-241): the keys
 * on hex rings 1..rings around each lens (cube distance in the odd-r offset lattice, odd key rows
 * half a pitch to the right), ordered by the reference's score 1 - normalised baseline (weights
 * 1, 0, 0), i.e. by the centre distance (float, as the reference computes it); equal distances by
 * angle in [0, 2 pi), then collection order; capped at max_neighbors (-1 padded).  rings in
 * 1..5 (2 = ring-1 + ring-2, 18 neighbours in the interior).  The reference also drops
 * candidates by a RANSAC test on image-similarity scores (Hu moments,
 * MLA_Problem::OutlierImpByRANSAC); this table keeps every lens of the rings. */
int lfpm_table_ring_neighbors(const int* tilekeys, const float* centers, int num_mis,
                              int lattice_cols, int lattice_rows, int rings, int max_neighbors,
                              int* neighbor_ids, int* neighbor_counts);

/* Crop origins of a read lens table (DepthSolver::CreateProblemsAndDisNormals_Frame,
 * DepthSolver.cpp:1894-1905, and the slicing rectangle of :1977): corner = centre - (crop-1)/2
 * in FP32; kept[v] = 0 (origin -1, -1) when the crop leaves mla_u_size*baseline x
 * mla_v_size*baseline, or the raw frame; else origins = the corner truncated to int. */
int lfpm_table_origins(const float* centers, int num_mis, int crop, int mla_u_size,
                       int mla_v_size, float baseline, int raw_w, int raw_h, int* origins,
                       int* kept);

/* ---------------------------------------------------------------------------------------
 * Synthetic focused-plenoptic frame (host only; test and benchmark workload generator — the
 * reference ships no data).  Seeded slanted planes in virtual space on a Voronoi partition,
 * band-limited value-noise texture, Gaussian sensor noise.
 * ------------------------------------------------------------------------------------- */
typedef struct lfpm_synth_params {
    int raw_w, raw_h;
    float pitch;
    int num_planes;
    float vmin, vmax;          /* virtual depth range of the scene */
    uint32_t scene_seed;
    uint32_t noise_seed;
    float noise_sigma;         /* in 0..255 units */
} lfpm_synth_params;

/* raw: raw_h*raw_w uint8.  gt_virtual_depth (optional, may be NULL): raw_h*raw_w float,
 * 0 outside micro-lens discs. */
int lfpm_synth_frame(const lfpm_synth_params* p, uint8_t* raw, float* gt_virtual_depth);

/* Philox4x32-10 block (Salmon et al., SC'11) exposed for known-answer tests. */
void lfpm_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* ---------------------------------------------------------------------------------------
 * PatchMatch solve parameters.  Mirrors PatchMatchParamsLF (MVStereo/LFDepthInfo.h:19-46)
 * as filled by AdaptMIPMFrame::InitGPUParamsFromCPUParams (MIStereo/AdaptMIPMFrame.cpp:126-143).
 * ------------------------------------------------------------------------------------- */
typedef struct lfpm_pm_params {
    int crop;                  /* MLA_Mask_Width_Cuda == MLA_Mask_Height_Cuda */
    float base;                /* Base: disparity unit (lens pitch in px) */
    int radius;                /* patch_size / 2 */
    int radius_increment;      /* window stride (reference radius_increment) */
    float sigma_spatial;
    float sigma_color;         /* on [0,1] intensities */
    float min_var;             /* variance gate (reference kMinVar=1e-5 on 0..255 -> 1e-5/255^2) */
    int top_k;
    float depth_min, depth_max;/* disparity search range */
    int iterations;            /* red-black iterations */
    int max_neighbors;         /* K: neighbour table row stride */
    uint32_t seed;             /* Philox key word 0 */
    uint32_t frame_uid;        /* Philox key word 1: results depend on (seed, frame_uid) only */
    int storage;               /* 0 = FP32 tiles in LDS; 1 = FP16 tiles in LDS holding the 8-bit
                                  intensity codes 0..255 exactly, each cell with its code and the
                                  code step to the right (tiles must be k/255, as
                                  lfpm_slice_tiles writes them; other values round to the nearest
                                  1/255), FP32 arithmetic on the 0..255 scale */
    float cos_max_slant;       /* planes with |nz| < cos_max_slant score 2 (0 = no gate, the
                                  reference behaviour; default 0.5 = 60 degrees) */
    float blur_lambda;         /* blur term of the cost vector: cost = min(2, ncc + blur_lambda *
                                  (1 - exp(-0.007368 * (sum b_ref - sum b_src)^2))) over the
                                  window samples that enter the NCC (ComputeMultiViewCostVector_
                                  Frame, MIStereo/AdaptMIPMFrame.cu:211-218: 2.0; the ACMM twin
                                  AdaptMIPMFrameACMM.cu:164: 0.8).  0 (default) = off.  When > 0
                                  the tiles buffer holds 2*V*crop^2 floats: the intensities, then
                                  the MIs' blur maps (blureness / 255) in the same layout.
                                  Provided with FP32 tiles (storage 0) and the default solver by
                                  lfpm_patchmatch_frame/_init/_phase/_finalize, lfpm_eval_costs,
                                  the batched forms, the ACMM variant (its coarse level takes
                                  the blur maps area-resized like the intensities) and the host
                                  context on raw frames (Laplacian blur tiles). */
} lfpm_pm_params;

/* Fill defaults for a given crop/base (reference defaults + config conventions). */
void lfpm_pm_params_default(lfpm_pm_params* p, int crop, float base);

/* Colour raw frame -> gray raw frame, the input of every slice entry point below.  Replaces the
 * per-crop cv::cvtColor(problem.m_Image_rgb, problem.m_Image_gray, cv::COLOR_BGR2GRAY) of
 * DepthSolver::Slice_RawMLAImage (DepthSolver.cpp:1985-1990) on the frame read with
 * cv::imread(path, 1) (:2503); the conversion is per pixel, so converting the frame and then
 * slicing gives the same crops.  OpenCV's 8-bit rule: Y = (1868 B + 9617 G + 4899 R + 8192) >> 14.
 * bgr_dev: height rows of bgr_pitch bytes, pixels B, G, R.  gray_dev: height rows of gray_pitch
 * bytes.  16-byte aligned rows take the vector path; any pitch >= 3 * width (>= width) works. */
int lfpm_bgr_to_gray(const uint8_t* bgr_dev, int width, int height, size_t bgr_pitch,
                     uint8_t* gray_dev, size_t gray_pitch, void* stream);

/* Gather MI crops from a raw uint8 frame into FP32 tiles in [0,1].
 * Replaces DepthSolver::Slice_RawMLAImage (DepthSolver.cpp:1951-1997) and the uint8->float
 * convertTo of AdaptMIPMFrame::BuildViewIndexMap (MIStereo/AdaptMIPMFrame.cpp:173-179).
 * raw_dev: raw_h rows of raw_pitch bytes.  origins_dev: 2V ints.  tiles_dev: V*crop*crop. */
int lfpm_slice_tiles(const uint8_t* raw_dev, int raw_w, int raw_h, size_t raw_pitch,
                     const int* origins_dev, int num_mis, int crop, float* tiles_dev,
                     void* stream);

/* Same gather, intensities kept on the reference's 0..255 scale (the filter's gray input,
 * MIDisparityFilterCPU::BuildTileInputsFromMaps, CheckModule/MIDisparityFilter.cpp:362-421). */
int lfpm_slice_gray(const uint8_t* raw_dev, int raw_w, int raw_h, size_t raw_pitch,
                    const int* origins_dev, int num_mis, int crop, float* gray_dev, void* stream);

/* Colour crops: bytes B, G, R of each MI pixel of a colour frame (bgr_dev: raw_h rows of
 * bgr_pitch bytes, 3 per pixel) -> bgr_tiles_dev (V*crop*crop*3 bytes), the problem.m_Image_rgb
 * crops of DepthSolver::Slice_RawMLAImage (DepthSolver.cpp:1951-1997) that colour the VD map. */
int lfpm_slice_bgr_tiles(const uint8_t* bgr_dev, int raw_w, int raw_h, size_t bgr_pitch,
                         const int* origins_dev, int num_mis, int crop, uint8_t* bgr_tiles_dev,
                         void* stream);

/* Blur tiles for the blur term (lfpm_pm_params.blur_lambda): the Laplacian blur score of the
 * gray frame (ImageQualityEstimate::QuantizeBlurByLaplacian, Util/ImageQualityEstimate.cpp:
 * 211-230 — cv::Laplacian with ksize 1, i.e. the 4-neighbour stencil, FP32, border reflected
 * without repeating the edge pixel), cropped per MI (DepthSolver::SliceAndIndicatlizeForMI,
 * DepthSolver.cpp:2709-2745) and divided by 255 as ComputeBilateralNCC_Frame reads it.
 * Write it to tiles_dev + V*crop*crop to pass both halves to the solver. */
int lfpm_slice_blur_tiles(const uint8_t* raw_dev, int raw_w, int raw_h, size_t raw_pitch,
                          const int* origins_dev, int num_mis, int crop, float* blur_tiles_dev,
                          void* stream);

/* Device workspace (bytes) needed by lfpm_patchmatch_frame / _init / _phase. */
size_t lfpm_patchmatch_workspace_size(int num_mis, const lfpm_pm_params* p);

/* Full single-frame solve: random init, `iterations` x {red phase, black phase}, then the
 * per-pixel disparity/baseline record for the final plane.
 * Replaces RunPatchMatchCUDAForFrame_Impl (MIStereo/AdaptMIPMFrame.cu:949-995, declared at
 * MIStereo/AdaptMIPMFrame.cpp:337-341) and its kernels RandomInitializationForFrame
 * (:474-518) and CheckerboardUpdateFrame (:520-947).
 * tiles_dev V*crop^2; centers_dev 2V; neighbor_ids_dev V*K; neighbor_counts_dev V.
 * Outputs (device, caller-owned): plane 4*V*crop^2 (float4), cost V*crop^2,
 * disp 4*V*crop^2 (float4: d, d*|b|/base, |b|, |b|/(d_real+1e-7) of the lowest-cost view),
 * sel V*crop^2 (uint32 bitmask of selected views).
 * sync_each_phase != 0 synchronises the stream after every launch (the reference's launch
 * pattern); 0 leaves all launches queued. */
int lfpm_patchmatch_frame(const float* tiles_dev, const float* centers_dev,
                          const int* neighbor_ids_dev, const int* neighbor_counts_dev,
                          int num_mis, const lfpm_pm_params* p,
                          float* plane_dev, float* cost_dev, float* disp_dev, uint32_t* sel_dev,
                          void* workspace_dev, size_t workspace_bytes, int sync_each_phase,
                          void* stream);

/* Building blocks of lfpm_patchmatch_frame for callers that drive the loop themselves.
 * _init writes the random-initialised state; _phase reads (plane,cost,sel)_in and writes
 * (plane,cost,sel)_out for one red/black phase; _finalize writes disp for the final planes.
 * A workspace of lfpm_patchmatch_workspace_size bytes also carries per-pixel cost records
 * (the costs of each pixel's plane over all its views) from _init through the phases of one
 * solve, so a phase does not score a pixel's unchanged plane again: the phases of a solve
 * take the workspace its _init was given, with the same inputs and parameters, and two solves
 * in flight take two workspaces.  Results are the same with or without the records; a smaller
 * workspace (the solve's state excluded) runs without them. */
int lfpm_patchmatch_init(const float* tiles_dev, const float* centers_dev,
                         const int* neighbor_ids_dev, const int* neighbor_counts_dev,
                         int num_mis, const lfpm_pm_params* p,
                         float* plane_dev, float* cost_dev, uint32_t* sel_dev,
                         void* workspace_dev, size_t workspace_bytes, void* stream);
int lfpm_patchmatch_phase(const float* tiles_dev, const float* centers_dev,
                          const int* neighbor_ids_dev, const int* neighbor_counts_dev,
                          int num_mis, const lfpm_pm_params* p, int iter, int phase,
                          const float* plane_in_dev, const float* cost_in_dev,
                          const uint32_t* sel_in_dev,
                          float* plane_out_dev, float* cost_out_dev, uint32_t* sel_out_dev,
                          void* workspace_dev, size_t workspace_bytes, void* stream);
int lfpm_patchmatch_finalize(const float* tiles_dev, const float* centers_dev,
                             const int* neighbor_ids_dev, const int* neighbor_counts_dev,
                             int num_mis, const lfpm_pm_params* p, const float* plane_dev,
                             float* disp_dev, void* stream);
/* _finalize on the solve's workspace (the one its _init and phases were given): pixels whose
 * cost record still holds their final plane are not scored again.  Same outputs as
 * _finalize. */
int lfpm_patchmatch_finalize_ws(const float* tiles_dev, const float* centers_dev,
                                const int* neighbor_ids_dev, const int* neighbor_counts_dev,
                                int num_mis, const lfpm_pm_params* p, const float* plane_dev,
                                float* disp_dev, void* workspace_dev, size_t workspace_bytes,
                                void* stream);

/* ACMM multi-scale variant.  Replaces AdaptMIPMFrameACMM::RunPatchMatchCUDAForFrameACMM
 * (MIStereo/AdaptMIPMFrameACMM.cpp:572-665) and RunPatchMatchCUDAForFrameACMM_Impl
 * (MIStereo/AdaptMIPMFrameACMM.cu:931-991).  Level schedule of the reference
 * (AdaptMIPMFrameACMM.cpp:155-200, 300-345): crops of at least 24 px run a half-scale level
 * (INTER_AREA area-average tiles, centres and Base scaled, patch 7, 3 iterations, top-k 6)
 * whose reliable planes (cost < 1, >= 3 selected views, 3x3 disparity spread <=
 * max(0.03, 0.12 Base0)) warm-start the full level (patch 5, 6 iterations, top-k 5); smaller
 * crops run one level (patch 5, 5 iterations, top-k 6).  Every update ends with the warm-seed
 * detail restorer and the final plane's own top-k cost and selection; a post pass writes the
 * disp record.  p supplies crop, base, sigmas, depth range (scaled per level), neighbours, seed
 * and frame_uid; its radius/iterations/top_k are replaced by the schedule.  storage must be 0.
 * Outputs as lfpm_patchmatch_frame (cost = top-k mean, sel = top-k mask). */
size_t lfpm_patchmatch_acmm_workspace_size(int num_mis, const lfpm_pm_params* p);
int lfpm_patchmatch_frame_acmm(const float* tiles_dev, const float* centers_dev,
                               const int* neighbor_ids_dev, const int* neighbor_counts_dev,
                               int num_mis, const lfpm_pm_params* p, float* plane_dev,
                               float* cost_dev, float* disp_dev, uint32_t* sel_dev,
                               void* workspace_dev, size_t workspace_bytes, void* stream);

/* Cost of fixed plane hypotheses.  hyp_dev: n records of {int mi, int pixel, float4 plane}
 * packed as 6 x 32-bit words (mi, pixel as int32; then nx, ny, nz, w as float).
 * cost_out_dev: n*K floats (per-view bilateral NCC cost, 2 for views beyond the MI's count).
 * Replaces ComputeMultiViewCostVector_Frame / ComputeBilateralNCC_Frame
 * (MIStereo/AdaptMIPMFrame.cu:73-225). */
int lfpm_eval_costs(const float* tiles_dev, const float* centers_dev,
                    const int* neighbor_ids_dev, const int* neighbor_counts_dev,
                    int num_mis, const lfpm_pm_params* p, const void* hyp_dev, int n,
                    float* cost_out_dev, void* stream);

/* lfpm_eval_costs without any allocation: no counterpart in the reference.  For callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  Same arguments, kernels and results as lfpm_eval_costs; the per-MI
 * segment table lives in a library-owned device array of LFPM_NOALLOC_MAX_MIS + 1 ints (1 MiB
 * per device, held until the process ends).  num_mis above LFPM_NOALLOC_MAX_MIS is rejected with
 * LFPM_E_INVALID.  Calls on one device share the array: calls that overlap (other streams or
 * threads) race and must be serialised by the caller. */
#define LFPM_NOALLOC_MAX_MIS 262143
int lfpm_eval_costs_noalloc(const float* tiles_dev, const float* centers_dev,
                            const int* neighbor_ids_dev, const int* neighbor_counts_dev,
                            int num_mis, const lfpm_pm_params* p, const void* hyp_dev, int n,
                            float* cost_out_dev, void* stream);

/* ---------------------------------------------------------------------------------------
 * Fused CheckModule pass.  Replaces MIDisparityFilterCPU::FilterGlobal/FilterTile
 * (CheckModule/MIDisparityFilter.cpp:423-693, 962-1009) with configuration
 * MIStereoMatch::FilterAndStatic (MIStereo/MIStereoMatch.cpp:759-778).  lfpm_filter_frame
 * reads a snapshot of the solve (all reads see pre-filter values, independent of MI order) and
 * writes filtered copies + reason mask; lfpm_filter_frame_ordered below keeps the reference's
 * MI order.
 * gray_dev: V*crop^2 floats in 0..255 (the uint8 crops as float).
 * ------------------------------------------------------------------------------------- */
typedef struct lfpm_filter_params {
    int max_triplet;
    float min_valid_disp;
    int use_selected_views_only;
    int clear_selected_views_when_invalid;
    int enable_cost_filter;
    float max_cost;
    int enable_cycle_geo_filter;
    float max_geo_err_px;
    int enable_cycle_photo_filter;
    float max_photo_err_u;
    int min_good_neighbors;
    int enable_spike_filter;
    float spike_abs_diff;
    int spike_min_neighbors;
} lfpm_filter_params;

void lfpm_filter_params_default(lfpm_filter_params* f);

int lfpm_filter_frame(const float* gray_dev, const float* centers_dev,
                      const int* neighbor_ids_dev, const int* neighbor_counts_dev,
                      int num_mis, int crop, int max_neighbors, float base,
                      const lfpm_filter_params* f,
                      const float* plane_in_dev, const float* cost_in_dev,
                      const float* disp_in_dev, const uint32_t* sel_in_dev,
                      float* plane_out_dev, float* cost_out_dev, float* disp_out_dev,
                      uint32_t* sel_out_dev, uint8_t* reason_dev, void* stream);

/* The same pass in FilterGlobal's own order (CheckModule/MIDisparityFilter.cpp:973-985, with
 * FilterTile's invalidation at :671-686): MIs in index order (the reference's tile-key map
 * order, tile row then column), each reading the filtered disparities and costs its neighbours
 * with a smaller index left and the pre-filter values of the others.  The result is the
 * reference's order-dependent one (rejections cascade to later MIs).
 *
 * lfpm_filter_order_levels (host memory, once per layout) groups the MIs into levels: an MI's
 * level is one more than the largest level of its smaller-index neighbours (0 without any), so
 * the MIs of one level read no filtered state of each other.  mi_order (num_mis ints) lists
 * the MIs level by level, level l at mi_order[level_start[l] .. level_start[l+1]);
 * level_start holds num_levels + 1 ints (at most num_mis + 1).
 * lfpm_filter_frame_ordered runs one launch per level: mi_order_dev is mi_order copied to the
 * device, level_start stays in host memory.  Outputs must not alias the inputs. */
int lfpm_filter_order_levels(const int* neighbor_ids, const int* neighbor_counts, int num_mis,
                             int max_neighbors, int* mi_order, int* level_start,
                             int* num_levels);
int lfpm_filter_frame_ordered(const float* gray_dev, const float* centers_dev,
                              const int* neighbor_ids_dev, const int* neighbor_counts_dev,
                              int num_mis, int crop, int max_neighbors, float base,
                              const lfpm_filter_params* f,
                              const float* plane_in_dev, const float* cost_in_dev,
                              const float* disp_in_dev, const uint32_t* sel_in_dev,
                              float* plane_out_dev, float* cost_out_dev, float* disp_out_dev,
                              uint32_t* sel_out_dev, uint8_t* reason_dev,
                              const int* mi_order_dev, const int* level_start, int num_levels,
                              void* stream);

/* ---------------------------------------------------------------------------------------
 * Focal-stack refocus: S slices at virtual depths v_k = vmin + (vmax-vmin)*k/(S-1).
 * Output pixel X (raw grid) of slice k = mean over MIs i with local coordinate
 * u = (X - c_i)/v_k + (crop-1)/2 inside [0,crop-1]^2 of bilinear(tile_i, u); 0 if none.
 * Built from LFRefocus::ComputeVirtual3D / BackProject2MI (FocalStack/LFRefocus.cpp:2100-2125)
 * and the AIF mapping (FocalStack/LFRefocus.cpp:244-281).
 * lattice_to_mi_dev: cols*rows ints from lfpm_layout_build.  out_dev: S*out_h*out_w floats.
 * ------------------------------------------------------------------------------------- */
int lfpm_refocus(const float* tiles_dev, const float* centers_dev, const int* lattice_to_mi_dev,
                 int lattice_cols, int lattice_rows, float pitch, int crop,
                 int out_w, int out_h, int num_slices, float vmin, float vmax,
                 float* out_dev, void* stream);

/* ---------------------------------------------------------------------------------------
 * Per-frame products.  Device workspace for either call: lfpm_products_workspace_size().
 *
 * lfpm_disparity_mosaic replaces DepthSolver::WriteDisMap_TileKey_new_AccuSequence
 * (DepthSolver.cpp:6716-6897): pixel (x, y) of MI v lands at raw position
 * trunc(c_v - (round(crop/2) - 1)) + (x, y) when its cost <= max_cost and its disparity
 * d = plane.w > 0; mosaic_dev (raw_h*raw_w floats) holds that d (0 elsewhere; where crops
 * overlap the highest MI index wins, the reference's last-writer order).  range_dev (2 floats,
 * may be NULL) = {min, max} of every d > 0 (no cost gate), the reference's colour-scale range
 * (+inf/-inf when empty).  The reference's uint8/JET renderings and its tolower() on the
 * integer corner are not reproduced (the corner is plainly truncated).
 *
 * lfpm_virtual_depth_map replaces DepthSolver::Virtual_depth_map_TileKeySequence
 * (DepthSolver.cpp:7832-8000): pixel (x, y) of MI v with d = plane.w (d <= 0 -> 1) splats
 * f_v = mla_base / d at round(c_v + trunc((x, y) - ((crop-1)/2 + 1)) * f_v) when that position
 * is strictly inside the frame.  vd_dev (raw_h*raw_w floats) = the median of each pixel's
 * samples by the reference's index rule (s[n/2+1] + s[(n+1)/2]) / 2 on the sorted samples,
 * indices clamped to n-1 (the rule reads past the end for n <= 2), 0 where none.
 * color_dev (optional, raw_h*raw_w bytes): median gray (gray_dev, 0..255 floats, the colour
 * image of the gray frames) with the reference's fill of empty pixels (probe order p+s, p-s,
 * p+s*(raw_w-1), p-s*raw_w) and ten 5x5 median passes with replicated borders.
 * range_dev (2 floats, may be NULL) = {min, max} of the medians.  mla_base = baseline *
 * bl0 / (bl0 + B) of the reference; the synthetic workloads use the lens pitch.
 * ------------------------------------------------------------------------------------- */
size_t lfpm_products_workspace_size(int num_mis, int crop, int raw_w, int raw_h);

int lfpm_disparity_mosaic(const float* plane_dev, const float* cost_dev, const float* centers_dev,
                          int num_mis, int crop, float max_cost, int raw_w, int raw_h,
                          float* mosaic_dev, float* range_dev, void* workspace_dev,
                          size_t workspace_bytes, void* stream);

int lfpm_virtual_depth_map(const float* plane_dev, const float* centers_dev, const float* gray_dev,
                           int num_mis, int crop, float mla_base, int raw_w, int raw_h,
                           float* vd_dev, uint8_t* color_dev, float* range_dev,
                           void* workspace_dev, size_t workspace_bytes, void* stream);

/* lfpm_virtual_depth_map coloured from the colour crops (lfpm_slice_bgr_tiles): the reference's
 * real-colour image (DepthSolver.cpp:7895-7993): per raw pixel each of B, G, R sorted on its own
 * with the colour index rule (color0/1/2, :7919-7927), the fill copying all three channels of
 * the pixel it finds, ten 5x5 median passes on each channel.  color_dev: raw_h*raw_w*3 bytes
 * B, G, R (required). */
int lfpm_virtual_depth_map_bgr(const float* plane_dev, const float* centers_dev,
                               const uint8_t* bgr_tiles_dev, int num_mis, int crop,
                               float mla_base, int raw_w, int raw_h, float* vd_dev,
                               uint8_t* color_dev, float* range_dev, void* workspace_dev,
                               size_t workspace_bytes, void* stream);

/* VD point cloud of DepthSolver::Virtual_depth_map_TileKeySequence (DepthSolver.cpp:8019-8053),
 * built on the device as the binary payload of StoreColorPlyFileBinaryPointCloud
 * (MVStereo/ACMP.cpp:451-500).  One 27-byte record {x, y, z, nx, ny, nz: float32 LE; red, green,
 * blue: uint8} per raw pixel with vd_dev > 0 (a pixel that received a VD sample), in row-major
 * pixel order (the order the reference collects them in; its OpenMP writer interleaves them in an
 * unspecified order).  With vdm = vd_dev[y][x]:
 *   tx = (x - origin_x) * pixel_size, ty = (y - origin_y) * pixel_size     (FP64)
 *   tz = vdm * B + bl0                       (FP32, two roundings: the float line of the reference)
 *   z = f / (tz - f) * tz, x = -f / (tz - f) * tx, y = -f / (tz - f) * ty  (FP64, then FP32)
 * The reference passes origin = (crop / 2, crop / 2) (mi_width/height_for_match * 0.5), its
 * sensor_pixel_size, mainlen_flocal_length and the calibration's B = d, bl0 = D (g_B, g_bl0).
 * Points with a non-finite x, y or z are written as (0, 0, 0), the writer's test; the normal is
 * (0, 0, 0) (the reference leaves it unset); the colour is color_dev's gray (NULL: 0) on all
 * three channels.  records_dev: records_bytes >= 27 * raw_w * raw_h.  count_dev: one uint32, the
 * number of records.  workspace: lfpm_products_workspace_size() bytes. */
int lfpm_vd_point_cloud(const float* vd_dev, const uint8_t* color_dev, int raw_w, int raw_h,
                        float origin_x, float origin_y, float pixel_size, float main_lens_f,
                        float B, float bl0, uint8_t* records_dev, size_t records_bytes,
                        unsigned* count_dev, void* workspace_dev, size_t workspace_bytes,
                        void* stream);

/* lfpm_vd_point_cloud with the B, G, R image of lfpm_virtual_depth_map_bgr (raw_h*raw_w*3
 * bytes): red, green, blue = channels 2, 1, 0, the writer's r/g/b = color.z/.y/.x
 * (MVStereo/ACMP.cpp:479-481 with point3D.color of DepthSolver.cpp:8044-8046). */
int lfpm_vd_point_cloud_bgr(const float* vd_dev, const uint8_t* color_bgr_dev, int raw_w,
                            int raw_h, float origin_x, float origin_y, float pixel_size,
                            float main_lens_f, float B, float bl0, uint8_t* records_dev,
                            size_t records_bytes, unsigned* count_dev, void* workspace_dev,
                            size_t workspace_bytes, void* stream);

/* ---------------------------------------------------------------------------------------
 * Refocus-side composites of a solved frame (device pointers, asynchronous on `stream`).
 *
 * lfpm_aif_composite replaces LFRefocus::AIFImageCompositeForMIA (FocalStack/LFRefocus.cpp:
 * 183-425) with FilterMatchingPoints (:2283-2301), GenerateVirtualImage (:2303-2328) and
 * HoleFilling (:2362-2498).  For every MI v not flagged in garbage_dev (V bytes, may be NULL)
 * and every crop pixel (x, y) with d = plane.w > 0: vd = baseline / d (FP32; the reference passes
 * the MI-image baseline here), (X, Y) = round(((x, y) - (crop - 1) / 2) * vd + centre_v) with
 * round half away from zero; positions inside the raw frame are candidates of that virtual
 * pixel.  Each virtual pixel keeps its first candidate in collection order (MI, row, column)
 * among those with the highest score rich / 255 * 0.9 + blur / 255 * 0.1 (integer quotients:
 * 1 only for a saturated 255), the score maps read at row = source column, column = source row
 * as the reference indexes them.  richness_tiles_dev: the binarised texture-richness crops
 * (V*crop^2 bytes, 255 = rich; NULL = all 0); blur_tiles_dev: the blur-score crops (NULL = all
 * 0).  The reference's std::sort leaves the order of equal scores unspecified; below 17
 * candidates its implementation keeps collection order, as this one always does.  Outputs per
 * raw pixel (0 where no candidate): vd_dev (vd), inv_vd_dev (1 / vd), real_mm_dev (f * t / (t -
 * f) with t = vd * B + bl0, f = main_lens_f; metres = this / 1000), aif_dev (raw_h*raw_w*3 bytes
 * B, G, R of the kept source pixel), filled_dev (the composite after HoleFilling: raster order,
 * a hole takes the source pixel next to its right, left, lower or upper neighbour's source, in
 * that order of preference, the right and lower neighbours as they were when the pass started;
 * a filled pixel's source is attributed to the MI of the last candidate collected, as the
 * reference's member m_vIPoint does; passes until >= 98 % of pixels are counted, a pass fills
 * nothing or more than 30000 fills were made), denoised_dev (five 5x5 median passes with
 * replicated borders of aif_dev, each channel on its own).  stats_dev (8 ints): collected
 * candidates, kept pixels, hole-filling passes, fills, most candidates at one pixel, the
 * attribution MI (-1: none), pixels with more than 16 candidates, pixels with every channel >= 1
 * before filling.  Limits: crop <= 127, raw_w <= 8192, V*crop^2 < 2^30.  The hole filling runs
 * as row iterations whose number depends on the data: the call synchronises `stream` to read
 * their progress and returns when all outputs are written.
 * Display renderings (MAD clamp, normalisation, gamma, colour maps) are not produced.
 * ------------------------------------------------------------------------------------- */
size_t lfpm_aif_workspace_size(int num_mis, int crop, int raw_w, int raw_h);

int lfpm_aif_composite(const float* plane_dev, const float* centers_dev,
                       const uint8_t* bgr_tiles_dev, const uint8_t* richness_tiles_dev,
                       const uint8_t* blur_tiles_dev, const uint8_t* garbage_dev, int num_mis,
                       int crop, float baseline, float B, float bl0, float main_lens_f,
                       int raw_w, int raw_h, float* vd_dev, float* inv_vd_dev,
                       float* real_mm_dev, uint8_t* aif_dev, uint8_t* filled_dev,
                       uint8_t* denoised_dev, int* stats_dev, void* workspace_dev,
                       size_t workspace_bytes, void* stream);

/* lfpm_fuse_backproject replaces LFRefocus::FuseVirtualDepth_BackProject (FocalStack/
 * LFRefocus.cpp:1553-1793) with ComputeVirtualDepth, ComputeVirtual3D and BackProject2MI
 * (:2086-2122).  The reference walks the MIs in order and each MI's pixels in raster order; a
 * pixel p of MI a not masked, with d >= 1, has the virtual point P = ((c, r) - ((crop - 1) / 2 +
 * 1)) * vd + centre_a, vd = mla_base / d (FP64 quotient, FP32 result), mla_base = baseline * bl0
 * / (bl0 + B).  For each matching neighbour b (nbr_dev / cnt_dev, the solve's table, -1 and
 * garbage entries skipped) P is projected into b, rounded by +0.5 and truncation; an unmasked
 * pixel there with d > 0 whose own virtual point projects back into a within 1 px, with virtual
 * depths within 5 % of vd, is consistent.  With >= 2 consistent neighbours the point, normal and
 * colour are averaged over the pixel and them (a product with the float reciprocal of count + 1,
 * as cv::Vec3f /= float is); the point splats at cvRound(x, y) (ties to even) into vd_dev (z),
 * aif_dev (B, G, R, rounded and saturated) and conf_dev (min(255, (count + 1) * 60)), later
 * points overwriting earlier ones, and the neighbour pixels used are masked.  Outputs are zero
 * where nothing splats.  points_dev (optional, points_bytes >= V*crop^2*36): the fused points
 * inside the frame in walk order, 9 floats {x, y, z, nx, ny, nz, b, g, r} each.  counts_dev (2
 * uint64): points inside the frame, fused points.  The walk runs as a schedule from
 * lfpm_fuse_schedule (order_dev on the device, level_start on the host): MIs of one level touch
 * no common mask, so they run together, and every MI whose walk touches a mask another one
 * touches runs in a later level than it when it comes later in MI order — the outcome equals
 * the sequential walk.  Limits: max_neighbors <= 31, crop <= 127.  A garbage neighbour is
 * skipped (the reference indexes an empty mask for it).  Display renderings are not produced. */
size_t lfpm_fuse_workspace_size(int num_mis, int crop, int raw_w, int raw_h);

/* Host: the level schedule of a neighbour table (nbr: V*max_neighbors, cnt: V, garbage: V bytes
 * or NULL).  order: V ints (the non-garbage MIs grouped by level, ascending within a level);
 * level_start: V + 1 ints, level l = order[level_start[l] .. level_start[l + 1]); num_levels. */
int lfpm_fuse_schedule(const int32_t* nbr, const int32_t* cnt, const uint8_t* garbage,
                       int num_mis, int max_neighbors, int32_t* order, int32_t* level_start,
                       int* num_levels);

int lfpm_fuse_backproject(const float* plane_dev, const float* centers_dev,
                          const uint8_t* bgr_tiles_dev, const int32_t* nbr_dev,
                          const int32_t* cnt_dev, int max_neighbors, const uint8_t* garbage_dev,
                          const int32_t* order_dev, const int32_t* level_start, int num_levels,
                          int num_mis, int crop, float baseline, float B, float bl0, int raw_w,
                          int raw_h, float* vd_dev, uint8_t* aif_dev, uint8_t* conf_dev,
                          float* points_dev, size_t points_bytes,
                          unsigned long long* counts_dev, void* workspace_dev,
                          size_t workspace_bytes, void* stream);

/* Host file writers.  lfpm_write_ply_points writes the header of StoreColorPlyFileBinaryPointCloud
 * (MVStereo/ACMP.cpp:458-470) and count records of lfpm_vd_point_cloud (copied to the host).
 * lfpm_write_dmb / lfpm_read_dmb are writeDepthDmb / readDepthDmb (MVStereo/ACMP.cpp:338-390):
 * int32 {1, height, width, 1} then height*width float32, native (little-endian) order; reading
 * rejects a type other than 1.  lfpm_read_dmb with data == NULL returns the size only. */
int lfpm_write_ply_points(const char* path, const uint8_t* records, uint64_t count);
int lfpm_write_dmb(const char* path, const float* data, int height, int width);
int lfpm_read_dmb(const char* path, float* data, size_t capacity, int* height, int* width);

/* ---------------------------------------------------------------------------------------
 * Virtual-to-real depth behavioural model, FP64 on the GPU.
 * lfpm_fit_v2r replaces VirtualToRealDepthFunc::BehavioralModel
 * (FocalStack/VirtualToRealDepthFunc.cpp:3390-3418): least squares of a_L on [a_L*v, v, 1].
 * a_l_dev, v_dev: n floats (device).  c_out: 3 doubles (host), written after a stream sync.
 * lfpm_v2r_convert replaces convertVirtualToRealDepth (:3421-3451):
 * z = (c1*v + c2)/(1 - c0*v), non-finite or negative -> 0.
 * ------------------------------------------------------------------------------------- */
int lfpm_fit_v2r(const float* a_l_dev, const float* v_dev, int n, double c_out[3], void* stream);
int lfpm_v2r_convert(const float* v_dev, int n, const double c[3], double* z_dev, void* stream);

/* ---------------------------------------------------------------------------------------
 * Batched and served forms (no counterpart in the reference, which solves one frame per
 * device reset).  Frames are grouped into parameter classes (equal lfpm_pm_params apart from
 * seed/frame_uid); each class runs as persistent k_stage launches over the concatenated MIs of
 * its frames.  Every frame's results equal its single-frame solve bit for bit (the random
 * stream is keyed by (seed, frame_uid), never by batch position).
 * ------------------------------------------------------------------------------------- */
#define LFPM_MAX_BATCH 64

#define LFPM_STAGE_IDLE 0
#define LFPM_STAGE_INIT 1
#define LFPM_STAGE_PHASE 2
#define LFPM_STAGE_FINAL 3    /* disparity record, then filter and refocus when requested */

typedef struct lfpm_batch_frame {
    /* inputs (device) */
    const float* tiles_dev;           /* V*crop^2 in [0,1]; with params->blur_lambda != 0 the
                                         V*crop^2 blur tiles follow (lfpm_slice_blur_tiles) */
    const float* gray_dev;            /* V*crop^2 in 0..255 (filter input), may be NULL */
    const float* centers_dev;         /* 2V */
    const int* neighbor_ids_dev;      /* V*K */
    const int* neighbor_counts_dev;   /* V */
    const int* lattice_to_mi_dev;     /* cols*rows (refocus), may be NULL */
    int num_mis, lattice_cols, lattice_rows, raw_w, raw_h;
    const lfpm_pm_params* params;     /* solve parameters incl. seed and frame_uid */
    const lfpm_filter_params* filter; /* NULL: no filter pass */
    /* state: A ends with the result, B is the red/black ping-pong partner (device) */
    float* plane_dev; float* cost_dev; uint32_t* sel_dev; float* disp_dev;
    float* plane_b_dev; float* cost_b_dev; uint32_t* sel_b_dev;
    /* filter outputs (device), used when filter != NULL */
    float* f_plane_dev; float* f_cost_dev; float* f_disp_dev; uint32_t* f_sel_dev;
    uint8_t* reason_dev;
    /* focal stack (device), used when refocus_slices > 0 */
    int refocus_slices;
    float refocus_vmin, refocus_vmax;
    float* refocus_dev;               /* slices*raw_h*raw_w */
    /* filter order: filter_num_levels == 0 runs lfpm_filter_frame (snapshot); > 0 runs
     * lfpm_filter_frame_ordered on this schedule (lfpm_filter_order_levels; mi_order on the
     * device, level_start on the host) */
    const int* filter_mi_order_dev;
    const int* filter_level_start;
    int filter_num_levels;
} lfpm_batch_frame;

/* Workspace for lfpm_batch_run / lfpm_batch_graph_create / lfpm_batch_step on these frames. */
size_t lfpm_batch_workspace_size(const lfpm_batch_frame* frames, int n);

/* Whole batch, stream-ordered: init, iterations x 2 phases, final record (+ filter, refocus). */
int lfpm_batch_run(const lfpm_batch_frame* frames, int n, void* workspace_dev,
                   size_t workspace_bytes, void* stream);

/* The same schedule captured once as a hipGraph: each parameter class on its own branch, so
 * classes run concurrently; replay with lfpm_graph_launch.  Frame pointers are baked in. */
typedef struct lfpm_graph lfpm_graph;
int lfpm_batch_graph_create(const lfpm_batch_frame* frames, int n, void* workspace_dev,
                            size_t workspace_bytes, lfpm_graph** out);
int lfpm_graph_launch(lfpm_graph* g, void* stream);
void lfpm_graph_destroy(lfpm_graph* g);

/* One continuous-batching step: every frame advances by its own stage (stages[i], iters[i],
 * phases[i]); LFPM_STAGE_PHASE reads A and writes B for phase 0, B to A for phase 1.  With
 * several parameter classes, each class runs on a stream of its own (forked from `stream` and
 * joined back into it before the call returns), so their launches overlap. */
int lfpm_batch_step(const lfpm_batch_frame* frames, int n, const int* stages, const int* iters,
                    const int* phases, void* workspace_dev, size_t workspace_bytes, void* stream);

/* Autotuner (the reference fixes one 16x16 block shape, MIStereo/AdaptMIPMFrame.cu:957-966):
 * time phase launches of this frame's parameter class (crop, neighbours, radius, tile storage,
 * blur term) at every launch shape of {1, 2, 4, 8, 12} wavefronts per workgroup (the threads that
 * share one MI's staged tiles, which sets the workgroup's LDS footprint) x {2, 4, 6, 8}
 * persistent workgroups per CU, keep the fastest for the class (process-wide), and return its
 * workgroups per CU.  Results do not depend on the choice.  Uses and overwrites B state of
 * `frame`.  lfpm_autotune_get returns the launch shape a class uses: until tuned, 8 workgroups
 * per CU (as many as fit) of whichever of 4, 8 or 12 wavefronts keeps the most wavefronts per
 * CU (4 for crops up to 16 with 18 views; 12 for the crop-21 class); it has no reference
 * counterpart. */
int lfpm_autotune_class(const lfpm_batch_frame* frame, void* workspace_dev,
                        size_t workspace_bytes, int* best_per_cu, float* best_ms);
int lfpm_autotune_get(const lfpm_pm_params* params, int* per_cu, int* waves_per_mi);

/* Work accounting for roofline reports (no reference counterpart).  While counters_dev (four
 * device uint64: evaluated NCC window samples, evaluated (pixel, view, hypothesis) triples, and
 * the samples and triples the two-round candidate scoring leaves out, i.e. the views of rows
 * whose cost over the first half of the selection already reaches the best) is set, solver
 * launches from the calling thread use counting kernel variants that accumulate into it; they
 * score the left-out views too (into the last two counters only) and their results are identical
 * to the non-counting kernels.  NULL switches counting off. */
int lfpm_set_work_counters(unsigned long long* counters_dev);

/* ---------------------------------------------------------------------------------------
 * Context API: frame-level drop-in used by the C++ facades (DepthSolver, AdaptMIPMFrame…).
 * Replaces AdaptMIPMFrame::{Initialize, InitializeGPU, RunPatchMatchCUDAForFrame,
 * WriteBackResults, ReleaseMemory} (MIStereo/AdaptMIPMFrame.cpp:48-124, 278-384).
 * ------------------------------------------------------------------------------------- */
/* ---------------------------------------------------------------------------------------
 * Served form: a continuous-batching frame server.  Replaces the per-frame loop of
 * DepthSolver::ProcessProblemsImpSequence (DepthSolver.cpp:2897-2901) under
 * DepthSolverManager::DepthSolverPipelineBySequence / ByBatch (DepthSolverManager.cpp:41-135),
 * which solves one frame at a time and resets the device after each (DepthSolver.cpp:368-374).
 *
 * lfpm_frame_submit copies the host raw frame (gray, raw_h rows of raw_pitch bytes; 0 = raw_w)
 * and returns at once with a ticket; a server thread admits queued frames into device slots
 * while their bytes fit hbm_budget_bytes (0 = half the free device memory) and max_slots
 * (0 = 16), keeps each admitted frame's tiles, state and cost records resident, and advances all
 * of them one stage per step (init, one red/black phase, final record + filter + refocus).
 * Frames of one parameter class (equal lfpm_pm_params but seed / frame_uid) share one launch per
 * step; classes run on their own streams (`streams`, 0 = 4) and overlap.  lfpm_frame_wait blocks
 * until that frame is done and copies its results into the caller's buffers (NULL = not wanted):
 * plane V*crop^2*4 floats, cost, disp *4, sel, reason (with a filter), refocus slices*raw_h*raw_w
 * floats; with a filter plane/cost/disp/sel are the filtered copies (as lfpm_ctx_solve_frame
 * returns them).  Results equal the same frame solved alone, bit for bit.  latency_ms = submit to
 * completion (queueing included), queue_ms = submit to admission.  Plain variant, no blur term.
 * A server is used from any number of host threads; lfpm_server_destroy finishes every
 * submitted frame first.
 * ------------------------------------------------------------------------------------- */
typedef struct lfpm_server lfpm_server;

typedef struct lfpm_serve_opts {
    size_t hbm_budget_bytes;
    int max_slots;
    int streams;
    int filter_order;  /* LFPM_FILTER_SNAPSHOT (0) or LFPM_FILTER_REFERENCE_ORDER: order of every
                          request's CheckModule pass (lfpm_ctx_set_filter_order) */
} lfpm_serve_opts;

typedef struct lfpm_frame_request {
    const lfpm_layout_params* layout;
    const uint8_t* raw_host;
    size_t raw_pitch;
    const lfpm_pm_params* params;      /* crop and max_neighbors must match the layout */
    const lfpm_filter_params* filter;  /* NULL: no CheckModule pass */
    int refocus_slices;                /* 0: no refocus stack */
    float refocus_vmin, refocus_vmax;
} lfpm_frame_request;

typedef struct lfpm_frame_result {
    float* plane;
    float* cost;
    float* disp;
    uint32_t* sel;
    uint8_t* reason;
    float* refocus;
    int num_mis, crop;                 /* out */
    double latency_ms, queue_ms;       /* out */
} lfpm_frame_result;

typedef struct lfpm_server_stats {
    long long submitted, completed, steps, launches;
    int max_active, classes;
    size_t in_use_bytes, budget_bytes;
} lfpm_server_stats;

int lfpm_server_create(int device, const lfpm_serve_opts* opts, lfpm_server** out);
void lfpm_server_destroy(lfpm_server* server);
int lfpm_frame_submit(lfpm_server* server, const lfpm_frame_request* request, uint64_t* ticket);
int lfpm_frame_wait(lfpm_server* server, uint64_t ticket, lfpm_frame_result* out);
int lfpm_server_get_stats(lfpm_server* server, lfpm_server_stats* out);

typedef struct lfpm_ctx lfpm_ctx;

int lfpm_create(int device, lfpm_ctx** out);
void lfpm_destroy(lfpm_ctx* ctx);
const char* lfpm_last_error(void);

/* Solver variant of the context's single-frame solves (lfpm_ctx_solve_frame, _solve_tiles):
 * LFPM_VARIANT_PLAIN = AdaptMIPMFrame (lfpm_patchmatch_frame, the default),
 * LFPM_VARIANT_ACMM = AdaptMIPMFrameACMM (lfpm_patchmatch_frame_acmm). */
#define LFPM_VARIANT_PLAIN 0
#define LFPM_VARIANT_ACMM 1
int lfpm_ctx_set_variant(lfpm_ctx* ctx, int variant);

/* Order of the context's CheckModule pass (every solve call that filters):
 * LFPM_FILTER_SNAPSHOT = lfpm_filter_frame (the default), LFPM_FILTER_REFERENCE_ORDER =
 * lfpm_filter_frame_ordered on the schedule the context builds with each MI set, the result of
 * the reference's MIDisparityFilterCPU::FilterGlobal (CheckModule/MIDisparityFilter.cpp:
 * 973-985). */
#define LFPM_FILTER_SNAPSHOT 0
#define LFPM_FILTER_REFERENCE_ORDER 1
int lfpm_ctx_set_filter_order(lfpm_ctx* ctx, int order);

/* Upload a layout and a raw host frame, solve it, filter it, and copy results to host.
 * Host outputs (any may be NULL): plane 4*V*crop^2, cost, disp 4*V*crop^2, sel, reason
 * (filter reason mask; NULL skips the filter pass). */
int lfpm_ctx_solve_frame(lfpm_ctx* ctx, const lfpm_layout_params* lp, const uint8_t* raw_host,
                         const lfpm_pm_params* p, const lfpm_filter_params* f,
                         float* plane, float* cost, float* disp, uint32_t* sel, uint8_t* reason);

/* Solve caller-supplied MI crops (the reference's MLA_Problem gray images, in view order) with a
 * caller-supplied neighbour table — the data model of AdaptMIPMFrame::BuildViewIndexMap
 * (MIStereo/AdaptMIPMFrame.cpp:145-222).  tiles_u8: V*crop*crop host bytes; centers 2V floats;
 * neighbor_ids V*K; neighbor_counts V (all host).  Outputs as lfpm_ctx_solve_frame. */
int lfpm_ctx_solve_tiles(lfpm_ctx* ctx, const uint8_t* tiles_u8, const float* centers,
                         const int* neighbor_ids, const int* neighbor_counts, int num_mis,
                         const lfpm_pm_params* p, const lfpm_filter_params* f, float* plane,
                         float* cost, float* disp, uint32_t* sel, uint8_t* reason);

/* Batched form of lfpm_ctx_solve_frame: n host frames of one layout solved together through
 * lfpm_batch_run (persistent launches over all frames' MIs).  Frame i keys its random stream
 * with frame_uid = p->frame_uid + i, so its results equal lfpm_ctx_solve_frame of that frame
 * with that frame_uid, bit for bit.  Output arguments are arrays of n host pointers (an array,
 * or any entry, may be NULL).  No reference counterpart: the reference's batch pipeline solves
 * its frames one after another (DepthSolverManager::DepthSolverPipelineByBatch,
 * DepthSolverManager.cpp:93-135). */
int lfpm_ctx_solve_frames(lfpm_ctx* ctx, const lfpm_layout_params* lp, int n,
                          const uint8_t* const* raws, const lfpm_pm_params* p,
                          const lfpm_filter_params* f, float* const* plane, float* const* cost,
                          float* const* disp, uint32_t* const* sel, uint8_t* const* reason);

/* Per-frame products of frame `frame_index` of the context's last solve call (its filtered
 * results when that call ran the filter): lfpm_disparity_mosaic and lfpm_virtual_depth_map on
 * the resident device state, copied to host.  The solve's centres must be raw-frame
 * coordinates of a raw_w x raw_h frame.  Host outputs (any may be NULL): mosaic and vd
 * raw_h*raw_w floats, color raw_h*raw_w bytes, mosaic_range and vd_range 2 floats each.
 * Replaces the writers that follow the solve in the reference's pipelines
 * (DepthSolver.cpp:6716-6897, 7832-8000). */
int lfpm_ctx_frame_products(lfpm_ctx* ctx, int frame_index, int raw_w, int raw_h,
                            float max_cost, float mla_base, float* mosaic, float* mosaic_range,
                            float* vd, uint8_t* color, float* vd_range);

#ifdef __cplusplus
}
#endif
#endif /* LFPM_H */
