// lfmvs.hpp — C++ facades of the micro-image depth path with the reference's class and method
// names (namespace LFMVS), implemented over the lfpm C ABI.  A caller of the reference's
// DepthSolverManager / DepthSolver / AdaptMIPMFrame / AdaptMIPM keeps its call sequence; OpenCV
// types are replaced by the small image structs below (OpenCV is not a dependency here).
//
// Reference interfaces (paths relative to the reference repository root):
//   DepthSolverManager          DepthSolverManager.h:14-29
//   DepthSolver                 DepthSolver.h:62-200 (the members on the matching path)
//   AdaptMIPMFrame              MIStereo/AdaptMIPMFrame.h:16-29
//   AdaptMIPM                   MIStereo/AdaptMIPM.h:47-107
//   AdaptMIPMPFPGR              MIStereo/AdaptMIPMPFPGR.h:14-46
//   DisparityAndNormal          MVStereo/LFDepthInfo.h:273-320
//   MLA_Problem                 MVStereo/LFDepthInfo.h:225-268 (fields used by the matcher)
//   SelectNeighbors             MINeighbor/SelectNeighbors.h (FixedPosition selection)
//   QuadTreeTileKey             Common/QuadTree.h:65-148
//   LightFieldParams            Common/Common.h:154-187
//   eStereoType / eStereoStage  Common/Common.h:133-151
//
// Error behaviour: the reference's Initialize() returns false for an empty frame; its device
// errors abort the process (CUDA_SAFE_CALL, Common/CommonCUDA.cpp:77-86).  Here Initialize()
// returns false with LastError() set, and device errors throw std::runtime_error carrying the
// lfpm_last_error() text.
#pragma once

#include <cstdint>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "lfpm.h"

namespace LFMVS {

struct Float4 { float x, y, z, w; };
struct Point2f { float x, y; };
struct Int3 { int x, y, z; };

// 8-bit single-channel image (stands in for cv::Mat CV_8UC1).
struct GrayImage {
    int rows = 0, cols = 0;
    std::vector<uint8_t> data;
    GrayImage() = default;
    GrayImage(int r, int c) : rows(r), cols(c), data((size_t)r * c, 0) {}
    uint8_t& at(int y, int x) { return data[(size_t)y * cols + x]; }
    uint8_t at(int y, int x) const { return data[(size_t)y * cols + x]; }
    bool empty() const { return data.empty(); }
};

// 32-bit float single-channel image (stands in for cv::Mat CV_32FC1).
struct FloatImage {
    int rows = 0, cols = 0;
    std::vector<float> data;
    float at(int y, int x) const { return data[(size_t)y * cols + x]; }
};

enum eStereoType { eST_ACMH = 0, eST_PlannerPrior, eST_Horizontal_Propagation, eST_BlurFeature };
enum eStereoStage {
    eSS_ACMH_Begin = 0, eSS_ACMH_Finished, eSS_PlannerPrior_Begin, eSS_PlannerPrior_Finished,
    eSS_LRGeometric_Check, eSS_BlurFeature_Coarse, eSS_BlurFeature_Fine, eSS_BlurFeature_Finest
};

class QuadTreeTileKey {
public:
    QuadTreeTileKey(int32_t lod, int32_t tile_x, int32_t tile_y);
    int32_t GetLOD() const { return m_LOD; }
    int32_t GetTileX() const { return m_TileX; }
    int32_t GetTileY() const { return m_TileY; }
    const std::string& StrRemoveLOD() const { return m_strKey_RemoveLOD; }
    const std::string& Str() const { return m_strKey; }
    bool operator==(const QuadTreeTileKey& r) const {
        return m_LOD == r.m_LOD && m_TileX == r.m_TileX && m_TileY == r.m_TileY;
    }
    bool operator!=(const QuadTreeTileKey& r) const { return !(*this == r); }
    // (LOD, row, column): the order every per-frame table follows
    bool operator<(const QuadTreeTileKey& r) const {
        if (m_LOD != r.m_LOD) return m_LOD < r.m_LOD;
        if (m_TileY != r.m_TileY) return m_TileY < r.m_TileY;
        return m_TileX < r.m_TileX;
    }

private:
    int32_t m_LOD, m_TileX, m_TileY;
    std::string m_strKey, m_strKey_RemoveLOD;
};
typedef std::shared_ptr<QuadTreeTileKey> QuadTreeTileKeyPtr;
typedef std::vector<QuadTreeTileKeyPtr> QuadTreeTileKeyPtrVec;
struct QuadTreeTileKeyMapCmpLess {
    bool operator()(const QuadTreeTileKeyPtr& a, const QuadTreeTileKeyPtr& b) const {
        return *a < *b;
    }
};

class MLA_Info {
public:
    MLA_Info(QuadTreeTileKeyPtr key, Point2f center, int lens_type)
        : m_ptrKey(std::move(key)), m_Center(center), m_LensType(lens_type) {}
    const Point2f& GetCenter() const { return m_Center; }
    int GetLensType() const { return m_LensType; }
    QuadTreeTileKeyPtr GetTileKey() const { return m_ptrKey; }

private:
    QuadTreeTileKeyPtr m_ptrKey;
    Point2f m_Center;
    int m_LensType;
};
typedef std::shared_ptr<MLA_Info> MLA_InfoPtr;
typedef std::map<QuadTreeTileKeyPtr, MLA_InfoPtr, QuadTreeTileKeyMapCmpLess> QuadTreeTileInfoMap;

// Neighbour with its centre distance (Res_image_Key, MVStereo/LFDepthInfo.h:160-224).
struct Res_image_Key {
    QuadTreeTileKeyPtr m_ptrKey;
    float Base_line = 0.0f;
};

struct MLA_Problem {
    QuadTreeTileKeyPtr m_ptrKey;
    GrayImage m_Image_gray;                       // crop x crop micro-image
    QuadTreeTileKeyPtrVec m_NeighsSortVecForMatch;  // matching neighbours, best first
    std::vector<Res_image_Key> m_Res_Image_KeyVec;  // FixedPosition neighbours, nearest first
    bool m_bGarbage = false;
    bool m_bNeedMatch = true;
};

enum eSelectNeighborsType { eSNT_FixedPosition = 0, eSNT_Features, eSNT_Gradients };
typedef std::map<QuadTreeTileKeyPtr, MLA_Problem, QuadTreeTileKeyMapCmpLess> QuadTreeProblemMap;
typedef std::map<std::string, QuadTreeProblemMap> QuadTreeProblemMapMap;

struct LightFieldParams {
    LightFieldParams();
    // crop side from the baseline when unset (Common/Common.cpp:34-38)
    void ComputeMIA_Math_Info();

    int mla_u_size = 0, mla_v_size = 0;
    int mi_width_for_match = 0, mi_height_for_match = 0;
    float baseline = 0.0f;                // lens pitch in raw pixels
    float mainlen_flocal_length = 0.0f;
    float sensor_pixel_size = 0.0f;
    // virtual-depth scale: f_v = mla_base / d with mla_base = baseline * bl0 / (bl0 + B)
    // (DepthSolver.cpp:7890-7891); 0 selects baseline (the synthetic lens model)
    float mla_base = 0.0f;
    // matcher settings.  When solve_params_set is false the frame matcher uses
    // lfpm_pm_params_default() with the reference frame path's neighbour cap (16) and
    // iteration count (3) (MIStereo/AdaptMIPMFrame.cpp:17, :128).
    lfpm_pm_params solve_params{};
    bool solve_params_set = false;
    // frame matcher of ProcessProblemsImpSequence: false = AdaptMIPMFrame
    // (MIStereoMatch::StereoMatchingForMIA_FrameCrossViews, MIStereo/MIStereoMatch.cpp:626-696),
    // true = AdaptMIPMFrameACMM (StereoMatchingForMIA_FrameCrossViews_ACMM, :698-740)
    bool frame_matcher_acmm = false;
    // CheckModule filter after the solve (MIStereoMatch.cpp:759-778); on by default
    lfpm_filter_params filter_params{};
    bool run_filter = true;
    // true: MIs filtered in the reference's map order, each reading the filtered state its
    // smaller-index neighbours left (MIDisparityFilterCPU::FilterGlobal,
    // CheckModule/MIDisparityFilter.cpp:973-985), the reference's result; false: every MI reads
    // the pre-filter snapshot (order-independent, one launch)
    bool filter_reference_order = true;
    // disparity-mosaic cost gate (DepthSolver.cpp:6802-6805)
    float mosaic_max_cost = 0.2f;
};

class DisparityAndNormal {
public:
    explicit DisparityAndNormal(LightFieldParams& params);
    void Release();

    LightFieldParams m_Params;
    QuadTreeTileKeyPtr m_ptrKey;
    std::vector<Float4> ph_cuda;        // plane (nx, ny, nz, w = disparity)
    std::vector<Float4> disp_v_cuda;    // (d, d_real, |baseline|, virtual depth) of best view
    std::vector<float> c_cuda;          // matching cost
    std::vector<float> d_cuda;          // disparity (= ph_cuda[i].w)
    std::vector<unsigned int> selected_views;
    std::vector<uint8_t> reject_reason; // CheckModule reason per pixel (LFPM_PR_*)
    eStereoStage m_StereoStage = eSS_ACMH_Begin;
    int row = 0, col = 0, num = 0;
};
typedef std::shared_ptr<DisparityAndNormal> DisparityAndNormalPtr;
typedef std::map<QuadTreeTileKeyPtr, DisparityAndNormalPtr, QuadTreeTileKeyMapCmpLess>
    QuadTreeDisNormalMap;
typedef std::map<std::string, QuadTreeDisNormalMap> QuadTreeDisNormalMapMap;

// Whole-frame matcher (AdaptMIPMFrame, MIStereo/AdaptMIPMFrame.h:16-29).
class AdaptMIPMFrame {
public:
    explicit AdaptMIPMFrame(LightFieldParams& params);
    ~AdaptMIPMFrame();

    bool Initialize(QuadTreeTileInfoMap& MLA_info_map, QuadTreeProblemMap& problem_map);
    void RunPatchMatchCUDAForFrame();
    void WriteBackResults(QuadTreeDisNormalMap& disNormals_map);
    void ReleaseMemory();
    int GetValidViewCount() const { return m_num_views; }

    const std::string& LastError() const { return m_error; }
    const lfpm_pm_params& GetParamsGPU() const { return params; }
    lfpm_ctx* Context();  // the device context (created on first use)
    // frame position in a sequence: the random stream is keyed by frame_uid + index, so a
    // frame solves identically alone, in sequence and in a batch
    void SetFrameIndex(int index) { m_frame_index = index; }

protected:
    void InitGPUParamsFromCPUParams();
    bool BuildViewIndexMap(QuadTreeTileInfoMap& MLA_info_map, QuadTreeProblemMap& problem_map);
    void Solve(int variant);

    LightFieldParams m_Params;
    lfpm_pm_params params{};
    int m_num_views = 0;
    int m_max_neighbors = 16;
    int m_crop = 0;
    int m_frame_index = 0;
    bool m_bReleased = true;
    std::string m_error;
    std::vector<QuadTreeTileKeyPtr> m_view_keys;
    std::map<QuadTreeTileKeyPtr, int, QuadTreeTileKeyMapCmpLess> m_key_to_viewid;
    std::vector<uint8_t> tiles_u8;
    std::vector<float> centers;
    std::vector<int> neighbor_ids_host, neighbor_counts_host;
    std::vector<float> plane_final_host, cost_final_host, disp_final_host;
    std::vector<uint32_t> selected_final_host;
    std::vector<uint8_t> reason_final_host;
    lfpm_ctx* m_ctx = nullptr;
};

// Multi-scale frame matcher (AdaptMIPMFrameACMM, MIStereo/AdaptMIPMFrameACMM.h:25-42): the ACMM
// variant (level schedule, warm start, detail restorer, top-k metadata) on the same views.
class AdaptMIPMFrameACMM : public AdaptMIPMFrame {
public:
    explicit AdaptMIPMFrameACMM(LightFieldParams& params) : AdaptMIPMFrame(params) {}
    void RunPatchMatchCUDAForFrameACMM();
    int GetNumLevels() const;
};

// Per-micro-image matcher (AdaptMIPM, MIStereo/AdaptMIPM.h:47-107): the same solver on a work
// list of one reference MI plus its matching neighbours (views), results for the reference MI.
// The planar-prior inputs (planeVec, costVec) of Initialize are accepted and not used.
class AdaptMIPM {
public:
    explicit AdaptMIPM(LightFieldParams& params);
    virtual ~AdaptMIPM();

    void RunPatchMatchCUDAForMI();
    // The reference's planar-prior and soft-proxy patch-repair variants (AdaptMIPM.h:54-57) run
    // other per-MI solvers; here both run the solve of RunPatchMatchCUDAForMI.
    void RunPatchMatchCUDAForMI_plane() { RunPatchMatchCUDAForMI(); }
    void RunPatchMatchCUDAForMI_SoftProxy_PatchRepair() { RunPatchMatchCUDAForMI(); }
    void SetTileKey(int tile_x, int tile_y);
    void Initialize(QuadTreeTileInfoMap& MLA_info_map, MLA_Problem& problem,
                    QuadTreeProblemMap& problem_map, std::vector<Float4>& planeVec,
                    std::vector<float>& costVec);

    Float4 GetPlaneHypothesis(const int index);
    Float4* GetPlaneHypothesisVector();
    float GetCost(const int index);
    float* GetCostVector();
    Float4 GetDisparityBaseline(const int index);
    Float4* GetDisparityBaselineVector();
    unsigned int GetSelected_viewIndexs(const int pt_col, const int pt_row);
    unsigned int* GetSelected_view_vector();
    float GetMinDepth() const { return params.depth_min; }
    float GetMaxDepth() const { return params.depth_max; }
    const lfpm_pm_params& GetParamsGPU() const { return params; }
    void ReleaseMemory();

    // the one-MI work list (view 0 = reference MI), for inspection
    int NumViews() const { return (int)centers.size() / 2; }
    const std::vector<uint8_t>& Tiles() const { return tiles_u8; }
    const std::vector<float>& Centers() const { return centers; }
    const std::vector<int>& NeighborIds() const { return neighbor_ids; }
    const std::vector<int>& NeighborCounts() const { return neighbor_counts; }

private:
    LightFieldParams m_Params;
    lfpm_pm_params params{};
    int m_tile_x = 0, m_tile_y = 0;
    int m_crop = 0;
    std::vector<uint8_t> tiles_u8;
    std::vector<float> centers;
    std::vector<int> neighbor_ids, neighbor_counts;
    std::vector<Float4> plane, disp;
    std::vector<float> cost;
    std::vector<unsigned int> sel;
    lfpm_ctx* m_ctx = nullptr;
};

// Proxy planes of one pixel for the PFPGR repair stage (Proxy_DisPlane, Common/CommonUtil.h:15-25):
// 8 slots, unset slots hold -9999.
struct Proxy_DisPlane {
    Float4 plane[8];
    Proxy_DisPlane() {
        for (Float4& p : plane) p = Float4{-9999.0f, -9999.0f, -9999.0f, -9999.0f};
    }
};

// First-stage results of one MI handed to the second stage (StereoResultInfo,
// MIStereo/AdaptMIPM.h:18-45; host arrays only).
struct StereoResultInfo {
    StereoResultInfo(const int mi_width, const int mi_height)
        : mi_image_width(mi_width), mi_image_height(mi_height),
          plane_hypotheses_host((size_t)mi_width * mi_height),
          costs_host((size_t)mi_width * mi_height),
          neighbor_patchFill_host((size_t)mi_width * mi_height, Int3{-1, -1, -1}),
          disp_baseline_host((size_t)mi_width * mi_height),
          selected_views_host((size_t)mi_width * mi_height) {}
    int mi_image_width, mi_image_height;
    std::vector<Float4> plane_hypotheses_host;
    std::vector<float> costs_host;
    std::vector<Int3> neighbor_patchFill_host;
    std::vector<Float4> disp_baseline_host;
    std::vector<unsigned int> selected_views_host;
};

// Two-stage per-MI matcher (AdaptMIPMPFPGR, MIStereo/AdaptMIPMPFPGR.h:14-46) with the reference's
// method names.  The edge-aware propagation and the boundary/broken-pixel proxy search of the
// reference's collect and repair kernels are not part of this library: the collect stage runs
// the per-MI solve and reports no proxy for any pixel (GetNeighborPGR = (-1, -1, -1)), and the
// repair stage, having no proxies to apply, runs the same solve again (same seed, same result).
class AdaptMIPMPFPGR : public AdaptMIPM {
public:
    explicit AdaptMIPMPFPGR(LightFieldParams& params) : AdaptMIPM(params) {}

    void RunPatchMatchCUDAForMI_PFPGR_Collect();
    void RunPatchMatchCUDAForMI_PFPGR_Repair();
    void Initialize_SecondStage(QuadTreeTileInfoMap& MLA_info_map, MLA_Problem& problem,
                                QuadTreeProblemMap& problem_map, std::vector<Float4>& planeVec,
                                std::vector<float>& costVec, Proxy_DisPlane* proxy_Plane,
                                StereoResultInfo* pStereoResult);
    Int3 GetNeighborPGR(const int index) const { return neighbor_PGR.at(index); }

private:
    std::vector<Int3> neighbor_PGR;
    Proxy_DisPlane* proxy_plane = nullptr;
    StereoResultInfo* stereo_result = nullptr;
};

// Frame-level driver (DepthSolver, DepthSolver.h:62-200): frames are supplied in memory (the
// reference's XML calibration and image ingest are outside this library).
class DepthSolver {
public:
    DepthSolver(std::string& strRootPath, std::string& strSavePath);
    ~DepthSolver();

    LightFieldParams& GetLightFieldParams() { return m_Params; }
    QuadTreeTileInfoMap& GetMLAInfoMap() { return m_MLA_info_map; }
    std::map<std::string, GrayImage>& GetRawImageMap() { return m_RawImageMap; }
    QuadTreeProblemMapMap& GetMIAProblemsMapMap() { return m_MIA_problem_map_map; }
    QuadTreeProblemMap& GetMIAProblemsMap(std::string& strName) {
        return m_MIA_problem_map_map[strName];
    }
    QuadTreeDisNormalMap& GetDisNormalMap(const std::string& strName) {
        return m_MIA_dispNormal_map_map[strName];
    }
    std::string& GetSavePath() { return m_strSavePath; }
    std::string& GetRootPath() { return m_strRootPath; }
    void SetStereoType(eStereoType type) { m_eStereoType = type; }
    eStereoType GetStereoType() { return m_eStereoType; }

    // Micro-lens layout of every frame (replaces the CalibInfo.xml MLA table and the neighbour
    // selection of MINeighbor): hex lattice centres, lens types and sorted matching neighbours.
    // Unless the caller set solve parameters, the solver uses lfpm_pm_params_default() with the
    // layout's neighbour count and 3 iterations.
    bool SetMLALayout(const lfpm_layout_params& lp);
    void AddRawFrame(const std::string& strName, const GrayImage& raw);

    // Crops of one raw frame into per-MI problems (DepthSolver::GetMLAImagesSequence).
    bool GetMLAImagesSequence(std::string strNameLessExt, QuadTreeProblemMap& problem_map);
    // Matching neighbours of every problem (SelectNeighborsForProblemsBySequence).
    void SelectNeighborsForProblemsBySequence(std::string& strName,
                                              QuadTreeProblemMap& problem_map);
    // Solve (+ CheckModule filter) one frame into its DisparityAndNormal map with the matcher
    // LightFieldParams::frame_matcher_acmm names.
    void ProcessProblemsImpSequence(QuadTreeProblemMapMap::iterator itr, bool geom_consistency);
    // Per-frame products of the last solved frame.
    void WriteDisMap_TileKey_new_AccuSequence(std::string& strName,
                                              QuadTreeProblemMap& problem_map);
    void Virtual_depth_map_TileKeySequence(std::string& strName, QuadTreeProblemMap& problem_map);

    const FloatImage& GetDisparityMap(const std::string& n) { return m_DisparityMaps[n]; }
    const FloatImage& GetVirtualDepthMap(const std::string& n) { return m_VirtualDepthMaps[n]; }
    const GrayImage& GetVirtualDepthColor(const std::string& n) { return m_VDColorImages[n]; }
    // {min, max} disparity over d > 0 and {min, max} of the virtual-depth medians
    const float* GetDisparityRange(const std::string& n) { return m_DisRange[n].data(); }
    const float* GetVirtualDepthRange(const std::string& n) { return m_VdRange[n].data(); }

    // Batched form: every added frame solved in one lfpm_batch_run (no reference counterpart;
    // DepthSolverManager::DepthSolverPipelineByBatch uses it).
    int ProcessAllFramesBatched();

    void ReleaseResources(std::string& strFrame);

private:
    friend class DepthSolverManager;
    friend class SelectNeighbors;
    std::string m_strRootPath, m_strSavePath;
    LightFieldParams m_Params;
    eStereoType m_eStereoType = eST_ACMH;
    lfpm_layout_params m_lp{};
    bool m_have_layout = false;
    std::vector<QuadTreeTileKeyPtr> m_keys;       // MI order of the layout
    std::vector<int> m_origins;                   // 2 per MI
    std::vector<int> m_nbr, m_cnt;                // layout neighbour table
    QuadTreeTileInfoMap m_MLA_info_map;
    std::map<std::string, GrayImage> m_RawImageMap;
    QuadTreeProblemMapMap m_MIA_problem_map_map;
    QuadTreeDisNormalMapMap m_MIA_dispNormal_map_map;
    std::map<std::string, FloatImage> m_DisparityMaps, m_VirtualDepthMaps;
    std::map<std::string, GrayImage> m_VDColorImages;
    std::map<std::string, std::vector<float>> m_DisRange, m_VdRange;
    std::unique_ptr<AdaptMIPMFrame> m_frame;      // matcher of the last solved frame
    std::string m_lastSolved;
};

// Neighbour selection (SelectNeighbors, MINeighbor/SelectNeighbors.cpp:981-1030).  The
// FixedPosition type fills every problem's m_Res_Image_KeyVec with the mode-1 key offsets of
// its place in the lens lattice, nearest first (capped at the layout's max_neighbors), and sets
// m_bGarbage as DepthSolver::ConfirmProblemForEstimation does; the feature and gradient types
// are outside this library and throw.
class SelectNeighbors {
public:
    explicit SelectNeighbors(DepthSolver* solver) : m_ptrDepthSolver(solver) {}
    void SelectNeighborsForMIA(QuadTreeProblemMapMap::iterator& itrPM,
                               eSelectNeighborsType eSelectType);

private:
    DepthSolver* m_ptrDepthSolver;
};

// Pipeline entry points (DepthSolverManager, DepthSolverManager.h:14-29).
class DepthSolverManager {
public:
    DepthSolverManager(std::string& strRootPath, std::string& strSavePath);
    ~DepthSolverManager();

    // every added frame in name order: crops, neighbours, solve + filter, mosaic, VD map
    int DepthSolverPipelineBySequence(eStereoType eStereoType);
    // every added frame in one batched launch sequence, then the per-frame products
    int DepthSolverPipelineByBatch();

    std::shared_ptr<DepthSolver> GetDepthSolver() { return m_ptrDepthSolver; }

private:
    std::shared_ptr<DepthSolver> m_ptrDepthSolver;
};

}  // namespace LFMVS
