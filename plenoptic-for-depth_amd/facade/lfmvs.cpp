// lfmvs.cpp — C++ facades (namespace LFMVS) over the lfpm C ABI.  Host-side data marshalling
// only: every solve, filter and product runs through liblfpm's HIP kernels.
#include "../../include/lfmvs.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <iterator>

namespace LFMVS {

namespace {

[[noreturn]] void raise(const char* what) {
    throw std::runtime_error(std::string(what) + ": " + lfpm_last_error());
}

void check(int rc, const char* what) {
    if (rc != LFPM_OK) raise(what);
}

lfpm_ctx* make_ctx() {
    lfpm_ctx* c = nullptr;
    check(lfpm_create(0, &c), "lfpm_create");
    return c;
}

// Solve parameters of a matcher from the light-field parameters (the values the reference's
// InitGPUParamsFromCPUParams would fill, MIStereo/AdaptMIPMFrame.cpp:126-143, when the caller
// set none).
lfpm_pm_params solve_params(const LightFieldParams& lf, int crop, int max_neighbors,
                            int iterations) {
    lfpm_pm_params p;
    if (lf.solve_params_set) {
        p = lf.solve_params;
    } else {
        lfpm_pm_params_default(&p, crop, lf.baseline);
        p.max_neighbors = max_neighbors;
        p.iterations = iterations;
    }
    p.crop = crop;
    p.base = lf.baseline;
    return p;
}

int crop_of(const LightFieldParams& lf) {
    return lf.mi_width_for_match > 0 ? lf.mi_width_for_match
                                     : (int)std::floor(lf.baseline * 0.707106781186);
}

}  // namespace

// ------------------------------------------------------------------------------------------
QuadTreeTileKey::QuadTreeTileKey(int32_t lod, int32_t tile_x, int32_t tile_y)
    : m_LOD(lod), m_TileX(tile_x), m_TileY(tile_y) {
    m_strKey = std::to_string(lod) + "_" + std::to_string(tile_x) + "_" + std::to_string(tile_y);
    m_strKey_RemoveLOD = std::to_string(tile_x) + "_" + std::to_string(tile_y);
}

LightFieldParams::LightFieldParams() { lfpm_filter_params_default(&filter_params); }

void LightFieldParams::ComputeMIA_Math_Info() {
    if (mi_width_for_match <= 0) mi_width_for_match = (int)std::floor(baseline * 0.707106781186);
    if (mi_height_for_match <= 0) mi_height_for_match = mi_width_for_match;
}

DisparityAndNormal::DisparityAndNormal(LightFieldParams& params) : m_Params(params) {
    const int W = crop_of(params);
    const size_t n = (size_t)W * W;
    ph_cuda.assign(n, Float4{0, 0, 0, 0});
    disp_v_cuda.assign(n, Float4{0, 0, 0, 0});
    c_cuda.assign(n, 0.0f);
    d_cuda.assign(n, 0.0f);
    selected_views.assign(n, 0u);
    reject_reason.assign(n, 0);
}

void DisparityAndNormal::Release() {
    std::vector<Float4>().swap(ph_cuda);
    std::vector<Float4>().swap(disp_v_cuda);
    std::vector<float>().swap(c_cuda);
    std::vector<float>().swap(d_cuda);
    std::vector<unsigned int>().swap(selected_views);
    std::vector<uint8_t>().swap(reject_reason);
}

// ------------------------------------------------------------------------------------------
// AdaptMIPMFrame
// ------------------------------------------------------------------------------------------
AdaptMIPMFrame::AdaptMIPMFrame(LightFieldParams& params) : m_Params(params) {
    m_crop = crop_of(params);
    if (params.solve_params_set) m_max_neighbors = params.solve_params.max_neighbors;
}

AdaptMIPMFrame::~AdaptMIPMFrame() {
    ReleaseMemory();
    if (m_ctx) lfpm_destroy(m_ctx);
}

lfpm_ctx* AdaptMIPMFrame::Context() {
    if (!m_ctx) m_ctx = make_ctx();
    return m_ctx;
}

void AdaptMIPMFrame::InitGPUParamsFromCPUParams() {
    params = solve_params(m_Params, m_crop, m_max_neighbors, 3);
    params.frame_uid += (uint32_t)m_frame_index;
}

bool AdaptMIPMFrame::BuildViewIndexMap(QuadTreeTileInfoMap& MLA_info_map,
                                       QuadTreeProblemMap& problem_map) {
    m_view_keys.clear();
    m_key_to_viewid.clear();
    tiles_u8.clear();
    centers.clear();
    const int W = m_crop;
    for (auto& kv : problem_map) {
        MLA_Problem& problem = kv.second;
        if (problem.m_bGarbage || !problem.m_bNeedMatch) continue;
        auto it = MLA_info_map.find(kv.first);
        if (it == MLA_info_map.end()) continue;
        if (problem.m_Image_gray.rows != W || problem.m_Image_gray.cols != W) {
            m_error = "micro-image " + kv.first->StrRemoveLOD() + " is not " +
                      std::to_string(W) + "x" + std::to_string(W);
            return false;
        }
        const int vid = (int)m_view_keys.size();
        m_view_keys.push_back(kv.first);
        m_key_to_viewid[kv.first] = vid;
        tiles_u8.insert(tiles_u8.end(), problem.m_Image_gray.data.begin(),
                        problem.m_Image_gray.data.end());
        centers.push_back(it->second->GetCenter().x);
        centers.push_back(it->second->GetCenter().y);
    }
    m_num_views = (int)m_view_keys.size();
    if (m_num_views <= 0) {
        m_error = "no valid micro-images for the current frame";
        return false;
    }
    const int K = m_max_neighbors;
    neighbor_ids_host.assign((size_t)m_num_views * K, -1);
    neighbor_counts_host.assign(m_num_views, 0);
    for (int vid = 0; vid < m_num_views; ++vid) {
        MLA_Problem& problem = problem_map[m_view_keys[vid]];
        int count = 0;
        for (size_t i = 0; i < problem.m_NeighsSortVecForMatch.size() && count < K; ++i) {
            auto it = m_key_to_viewid.find(problem.m_NeighsSortVecForMatch[i]);
            if (it == m_key_to_viewid.end()) continue;
            neighbor_ids_host[(size_t)vid * K + count] = it->second;
            ++count;
        }
        neighbor_counts_host[vid] = count;
    }
    return true;
}

bool AdaptMIPMFrame::Initialize(QuadTreeTileInfoMap& MLA_info_map,
                                QuadTreeProblemMap& problem_map) {
    m_error.clear();
    if (m_crop < 3 || m_crop > LFPM_MAX_CROP) {
        m_error = "micro-image crop out of range";
        return false;
    }
    if (m_max_neighbors < 1 || m_max_neighbors > LFPM_MAX_NEIGHBORS) {
        m_error = "neighbour cap out of range";
        return false;
    }
    if (!BuildViewIndexMap(MLA_info_map, problem_map)) return false;
    InitGPUParamsFromCPUParams();
    m_bReleased = false;
    return true;
}

void AdaptMIPMFrame::RunPatchMatchCUDAForFrame() { Solve(LFPM_VARIANT_PLAIN); }

void AdaptMIPMFrameACMM::RunPatchMatchCUDAForFrameACMM() { Solve(LFPM_VARIANT_ACMM); }

int AdaptMIPMFrameACMM::GetNumLevels() const {
    // level rule of AdaptMIPMFrameACMM::BuildBlurAlignedPyramids (AdaptMIPMFrameACMM.cpp:
    // 300-345): a half-scale level exists when the crop is at least 24 px
    return m_crop >= 24 ? 2 : 1;
}

void AdaptMIPMFrame::Solve(int variant) {
    if (m_bReleased || m_num_views <= 0)
        throw std::runtime_error("AdaptMIPMFrame: RunPatchMatchCUDAForFrame before Initialize");
    check(lfpm_ctx_set_variant(Context(), variant), "AdaptMIPMFrame: solver variant");
    check(lfpm_ctx_set_filter_order(Context(), m_Params.filter_reference_order
                                                   ? LFPM_FILTER_REFERENCE_ORDER
                                                   : LFPM_FILTER_SNAPSHOT),
          "AdaptMIPMFrame: filter order");
    const size_t n = (size_t)m_num_views * m_crop * m_crop;
    plane_final_host.assign(4 * n, 0.0f);
    cost_final_host.assign(n, 0.0f);
    disp_final_host.assign(4 * n, 0.0f);
    selected_final_host.assign(n, 0u);
    const bool filt = m_Params.run_filter;
    reason_final_host.assign(filt ? n : 0, 0);
    check(lfpm_ctx_solve_tiles(Context(), tiles_u8.data(), centers.data(),
                               neighbor_ids_host.data(), neighbor_counts_host.data(), m_num_views,
                               &params, filt ? &m_Params.filter_params : nullptr,
                               plane_final_host.data(), cost_final_host.data(),
                               disp_final_host.data(), selected_final_host.data(),
                               filt ? reason_final_host.data() : nullptr),
          "AdaptMIPMFrame::RunPatchMatchCUDAForFrame");
}

void AdaptMIPMFrame::WriteBackResults(QuadTreeDisNormalMap& disNormals_map) {
    const int pixels = m_crop * m_crop;
    if (plane_final_host.size() < (size_t)m_num_views * pixels * 4) return;
    for (int vid = 0; vid < m_num_views; ++vid) {
        auto it = disNormals_map.find(m_view_keys[vid]);
        if (it == disNormals_map.end()) continue;
        DisparityAndNormal& dn = *it->second;
        dn.ph_cuda.resize(pixels);
        dn.disp_v_cuda.resize(pixels);
        dn.c_cuda.resize(pixels);
        dn.d_cuda.resize(pixels);
        dn.selected_views.resize(pixels);
        dn.reject_reason.assign(pixels, 0);
        const size_t off = (size_t)vid * pixels;
        for (int i = 0; i < pixels; ++i) {
            const float* p = &plane_final_host[4 * (off + i)];
            const float* d = &disp_final_host[4 * (off + i)];
            dn.ph_cuda[i] = Float4{p[0], p[1], p[2], p[3]};
            dn.d_cuda[i] = p[3];
            dn.c_cuda[i] = cost_final_host[off + i];
            dn.disp_v_cuda[i] = Float4{d[0], d[1], d[2], d[3]};
            dn.selected_views[i] = selected_final_host[off + i];
            if (!reason_final_host.empty()) dn.reject_reason[i] = reason_final_host[off + i];
        }
        dn.m_StereoStage = eSS_ACMH_Finished;
    }
}

void AdaptMIPMFrame::ReleaseMemory() {
    if (m_bReleased) return;
    std::vector<float>().swap(plane_final_host);
    std::vector<float>().swap(cost_final_host);
    std::vector<float>().swap(disp_final_host);
    std::vector<uint32_t>().swap(selected_final_host);
    std::vector<uint8_t>().swap(reason_final_host);
    m_bReleased = true;
}

// ------------------------------------------------------------------------------------------
// AdaptMIPM (one reference MI)
// ------------------------------------------------------------------------------------------
AdaptMIPM::AdaptMIPM(LightFieldParams& params) : m_Params(params) { m_crop = crop_of(params); }

AdaptMIPM::~AdaptMIPM() {
    if (m_ctx) lfpm_destroy(m_ctx);
}

void AdaptMIPM::SetTileKey(int tile_x, int tile_y) {
    m_tile_x = tile_x;
    m_tile_y = tile_y;
}

void AdaptMIPM::Initialize(QuadTreeTileInfoMap& MLA_info_map, MLA_Problem& problem,
                           QuadTreeProblemMap& problem_map, std::vector<Float4>& planeVec,
                           std::vector<float>& costVec) {
    (void)planeVec;
    (void)costVec;
    const int W = m_crop;
    // per-MI path neighbour cap (MIStereo/AdaptMIPM.cpp:157-160)
    const int K = m_Params.solve_params_set ? m_Params.solve_params.max_neighbors : 8;
    params = solve_params(m_Params, W, K, 3);
    tiles_u8.clear();
    centers.clear();
    auto add = [&](const QuadTreeTileKeyPtr& key, const MLA_Problem& pr) -> bool {
        auto it = MLA_info_map.find(key);
        if (it == MLA_info_map.end()) return false;
        if (pr.m_Image_gray.rows != W || pr.m_Image_gray.cols != W) return false;
        tiles_u8.insert(tiles_u8.end(), pr.m_Image_gray.data.begin(), pr.m_Image_gray.data.end());
        centers.push_back(it->second->GetCenter().x);
        centers.push_back(it->second->GetCenter().y);
        return true;
    };
    if (!add(problem.m_ptrKey, problem))
        throw std::runtime_error("AdaptMIPM::Initialize: reference micro-image missing or "
                                 "not crop-sized");
    int count = 0;
    for (size_t i = 0; i < problem.m_NeighsSortVecForMatch.size() && count < K; ++i) {
        const QuadTreeTileKeyPtr& nk = problem.m_NeighsSortVecForMatch[i];
        auto pit = problem_map.find(nk);
        if (pit == problem_map.end() || pit->second.m_bGarbage) continue;
        if (add(nk, pit->second)) ++count;
    }
    const int V = 1 + count;
    neighbor_ids.assign((size_t)V * K, -1);
    neighbor_counts.assign(V, 0);
    for (int k = 0; k < count; ++k) neighbor_ids[k] = 1 + k;
    neighbor_counts[0] = count;
    SetTileKey(problem.m_ptrKey->GetTileX(), problem.m_ptrKey->GetTileY());
}

void AdaptMIPM::RunPatchMatchCUDAForMI() {
    const int V = NumViews();
    if (V <= 0) throw std::runtime_error("AdaptMIPM: RunPatchMatchCUDAForMI before Initialize");
    if (!m_ctx) m_ctx = make_ctx();
    const size_t n = (size_t)V * m_crop * m_crop;
    std::vector<float> pl(4 * n), ds(4 * n), cs(n);
    std::vector<uint32_t> sl(n);
    check(lfpm_ctx_solve_tiles(m_ctx, tiles_u8.data(), centers.data(), neighbor_ids.data(),
                               neighbor_counts.data(), V, &params, nullptr, pl.data(), cs.data(),
                               ds.data(), sl.data(), nullptr),
          "AdaptMIPM::RunPatchMatchCUDAForMI");
    const int px = m_crop * m_crop;  // results of view 0 (the reference MI)
    plane.resize(px);
    disp.resize(px);
    cost.assign(cs.begin(), cs.begin() + px);
    sel.assign(sl.begin(), sl.begin() + px);
    for (int i = 0; i < px; ++i) {
        plane[i] = Float4{pl[4 * i], pl[4 * i + 1], pl[4 * i + 2], pl[4 * i + 3]};
        disp[i] = Float4{ds[4 * i], ds[4 * i + 1], ds[4 * i + 2], ds[4 * i + 3]};
    }
}

void AdaptMIPMPFPGR::RunPatchMatchCUDAForMI_PFPGR_Collect() {
    RunPatchMatchCUDAForMI();
    const int px = (int)(GetParamsGPU().crop * GetParamsGPU().crop);
    neighbor_PGR.assign(px, Int3{-1, -1, -1});
}

void AdaptMIPMPFPGR::Initialize_SecondStage(QuadTreeTileInfoMap& MLA_info_map,
                                            MLA_Problem& problem, QuadTreeProblemMap& problem_map,
                                            std::vector<Float4>& planeVec,
                                            std::vector<float>& costVec,
                                            Proxy_DisPlane* proxy_Plane,
                                            StereoResultInfo* pStereoResult) {
    Initialize(MLA_info_map, problem, problem_map, planeVec, costVec);
    proxy_plane = proxy_Plane;
    stereo_result = pStereoResult;
}

void AdaptMIPMPFPGR::RunPatchMatchCUDAForMI_PFPGR_Repair() {
    // no proxies were collected, so the repair is the solve itself; the second-stage inputs
    // (proxy planes, first-stage results) are kept for the caller but not read
    RunPatchMatchCUDAForMI();
}

Float4 AdaptMIPM::GetPlaneHypothesis(const int index) { return plane.at(index); }
Float4* AdaptMIPM::GetPlaneHypothesisVector() { return plane.data(); }
float AdaptMIPM::GetCost(const int index) { return cost.at(index); }
float* AdaptMIPM::GetCostVector() { return cost.data(); }
Float4 AdaptMIPM::GetDisparityBaseline(const int index) { return disp.at(index); }
Float4* AdaptMIPM::GetDisparityBaselineVector() { return disp.data(); }
unsigned int AdaptMIPM::GetSelected_viewIndexs(const int pt_col, const int pt_row) {
    return sel.at((size_t)pt_row * m_crop + pt_col);
}
unsigned int* AdaptMIPM::GetSelected_view_vector() { return sel.data(); }

void AdaptMIPM::ReleaseMemory() {
    std::vector<Float4>().swap(plane);
    std::vector<Float4>().swap(disp);
    std::vector<float>().swap(cost);
    std::vector<unsigned int>().swap(sel);
}

// ------------------------------------------------------------------------------------------
// DepthSolver
// ------------------------------------------------------------------------------------------
DepthSolver::DepthSolver(std::string& strRootPath, std::string& strSavePath)
    : m_strRootPath(strRootPath), m_strSavePath(strSavePath) {}

DepthSolver::~DepthSolver() = default;

bool DepthSolver::SetMLALayout(const lfpm_layout_params& lp) {
    lfpm_layout_info info;
    if (lfpm_layout_query(&lp, &info) != LFPM_OK) return false;
    const int V = info.num_mis, K = lp.max_neighbors;
    std::vector<float> centers(2 * (size_t)V);
    std::vector<int> keys(2 * (size_t)V), types(V);
    m_origins.assign(2 * (size_t)V, 0);
    m_nbr.assign((size_t)V * K, -1);
    m_cnt.assign(V, 0);
    if (lfpm_layout_build(&lp, centers.data(), m_origins.data(), keys.data(), types.data(),
                          m_nbr.data(), m_cnt.data(), nullptr) != LFPM_OK)
        return false;
    m_lp = lp;
    m_have_layout = true;
    m_Params.baseline = lp.pitch;
    m_Params.mi_width_for_match = m_Params.mi_height_for_match = info.crop;
    if (!m_Params.solve_params_set) {
        lfpm_pm_params_default(&m_Params.solve_params, info.crop, lp.pitch);
        m_Params.solve_params.max_neighbors = K;
        m_Params.solve_params.iterations = 3;
        m_Params.solve_params_set = true;
    }
    m_Params.mla_u_size = info.lattice_cols;
    m_Params.mla_v_size = info.lattice_rows;
    m_keys.clear();
    m_MLA_info_map.clear();
    for (int v = 0; v < V; ++v) {
        auto key = std::make_shared<QuadTreeTileKey>(0, keys[2 * v], keys[2 * v + 1]);
        m_keys.push_back(key);
        m_MLA_info_map[key] = std::make_shared<MLA_Info>(
            key, Point2f{centers[2 * v], centers[2 * v + 1]}, types[v]);
    }
    return true;
}

void DepthSolver::AddRawFrame(const std::string& strName, const GrayImage& raw) {
    m_RawImageMap[strName] = raw;
}

bool DepthSolver::GetMLAImagesSequence(std::string strNameLessExt,
                                       QuadTreeProblemMap& problem_map) {
    auto it = m_RawImageMap.find(strNameLessExt);
    if (it == m_RawImageMap.end() || !m_have_layout) return false;
    const GrayImage& raw = it->second;
    if (raw.cols != m_lp.raw_w || raw.rows != m_lp.raw_h) return false;
    const int W = m_Params.mi_width_for_match;
    problem_map.clear();
    for (size_t v = 0; v < m_keys.size(); ++v) {
        MLA_Problem pr;
        pr.m_ptrKey = m_keys[v];
        pr.m_Image_gray = GrayImage(W, W);
        const int ox = m_origins[2 * v], oy = m_origins[2 * v + 1];
        for (int y = 0; y < W; ++y)
            std::memcpy(&pr.m_Image_gray.data[(size_t)y * W],
                        &raw.data[(size_t)(oy + y) * raw.cols + ox], W);
        problem_map[m_keys[v]] = std::move(pr);
    }
    return true;
}

void DepthSolver::SelectNeighborsForProblemsBySequence(std::string& strName,
                                                       QuadTreeProblemMap& problem_map) {
    (void)strName;
    const int K = m_lp.max_neighbors;
    for (size_t v = 0; v < m_keys.size(); ++v) {
        auto it = problem_map.find(m_keys[v]);
        if (it == problem_map.end()) continue;
        QuadTreeTileKeyPtrVec& nv = it->second.m_NeighsSortVecForMatch;
        nv.clear();
        for (int k = 0; k < m_cnt[v]; ++k) nv.push_back(m_keys[m_nbr[v * K + k]]);
    }
}

void DepthSolver::ProcessProblemsImpSequence(QuadTreeProblemMapMap::iterator itr,
                                             bool geom_consistency) {
    (void)geom_consistency;
    const std::string& name = itr->first;
    QuadTreeProblemMap& problem_map = itr->second;
    const bool acmm = m_Params.frame_matcher_acmm;
    m_frame.reset(acmm ? new AdaptMIPMFrameACMM(m_Params) : new AdaptMIPMFrame(m_Params));
    // frame index = position in the raw-frame map (the order the batched form uses)
    auto pos = m_RawImageMap.find(name);
    m_frame->SetFrameIndex(pos == m_RawImageMap.end()
                               ? 0
                               : (int)std::distance(m_RawImageMap.begin(), pos));
    if (!m_frame->Initialize(m_MLA_info_map, problem_map))
        throw std::runtime_error("DepthSolver: frame " + name + ": " + m_frame->LastError());
    if (acmm) static_cast<AdaptMIPMFrameACMM*>(m_frame.get())->RunPatchMatchCUDAForFrameACMM();
    else m_frame->RunPatchMatchCUDAForFrame();
    QuadTreeDisNormalMap& dn = m_MIA_dispNormal_map_map[name];
    for (auto& kv : problem_map) {
        if (dn.find(kv.first) == dn.end()) {
            auto p = std::make_shared<DisparityAndNormal>(m_Params);
            p->m_ptrKey = kv.first;
            p->col = kv.first->GetTileX();
            p->row = kv.first->GetTileY();
            dn[kv.first] = p;
        }
    }
    m_frame->WriteBackResults(dn);
    m_lastSolved = name;
}

void DepthSolver::WriteDisMap_TileKey_new_AccuSequence(std::string& strName,
                                                       QuadTreeProblemMap& problem_map) {
    (void)problem_map;
    if (!m_frame || m_lastSolved != strName)
        throw std::runtime_error("DepthSolver: frame " + strName + " is not the last solved");
    const int w = m_lp.raw_w, h = m_lp.raw_h;
    FloatImage& m = m_DisparityMaps[strName];
    m.rows = h;
    m.cols = w;
    m.data.assign((size_t)w * h, 0.0f);
    std::vector<float>& r = m_DisRange[strName];
    r.assign(2, 0.0f);
    check(lfpm_ctx_frame_products(m_frame->Context(), 0, w, h, m_Params.mosaic_max_cost,
                                  m_Params.mla_base > 0 ? m_Params.mla_base : m_Params.baseline,
                                  m.data.data(), r.data(), nullptr, nullptr, nullptr),
          "DepthSolver::WriteDisMap_TileKey_new_AccuSequence");
}

void DepthSolver::Virtual_depth_map_TileKeySequence(std::string& strName,
                                                    QuadTreeProblemMap& problem_map) {
    (void)problem_map;
    if (!m_frame || m_lastSolved != strName)
        throw std::runtime_error("DepthSolver: frame " + strName + " is not the last solved");
    const int w = m_lp.raw_w, h = m_lp.raw_h;
    FloatImage& vd = m_VirtualDepthMaps[strName];
    vd.rows = h;
    vd.cols = w;
    vd.data.assign((size_t)w * h, 0.0f);
    GrayImage& col = m_VDColorImages[strName];
    col = GrayImage(h, w);
    std::vector<float>& r = m_VdRange[strName];
    r.assign(2, 0.0f);
    check(lfpm_ctx_frame_products(m_frame->Context(), 0, w, h, m_Params.mosaic_max_cost,
                                  m_Params.mla_base > 0 ? m_Params.mla_base : m_Params.baseline,
                                  nullptr, nullptr, vd.data.data(), col.data.data(), r.data()),
          "DepthSolver::Virtual_depth_map_TileKeySequence");
}

int DepthSolver::ProcessAllFramesBatched() {
    if (!m_have_layout || m_RawImageMap.empty()) return -1;
    const int n = (int)m_RawImageMap.size();
    if (n > LFPM_MAX_BATCH) return -1;
    const int W = m_Params.mi_width_for_match, K = m_lp.max_neighbors;
    const int V = (int)m_keys.size();
    const size_t px = (size_t)V * W * W;
    lfpm_pm_params p = solve_params(m_Params, W, K, 3);
    if (p.max_neighbors != K) return -1;
    std::vector<std::string> names;
    std::vector<const uint8_t*> raws;
    for (auto& kv : m_RawImageMap) {
        if (kv.second.cols != m_lp.raw_w || kv.second.rows != m_lp.raw_h) return -1;
        names.push_back(kv.first);
        raws.push_back(kv.second.data.data());
    }
    std::vector<std::vector<float>> pl(n, std::vector<float>(4 * px)), cs(n, std::vector<float>(px)),
        ds(n, std::vector<float>(4 * px));
    std::vector<std::vector<uint32_t>> sl(n, std::vector<uint32_t>(px));
    std::vector<std::vector<uint8_t>> rs(n, std::vector<uint8_t>(px));
    std::vector<float*> plp(n), csp(n), dsp(n);
    std::vector<uint32_t*> slp(n);
    std::vector<uint8_t*> rsp(n);
    for (int i = 0; i < n; ++i) {
        plp[i] = pl[i].data(); csp[i] = cs[i].data(); dsp[i] = ds[i].data();
        slp[i] = sl[i].data(); rsp[i] = rs[i].data();
    }
    if (!m_frame) m_frame.reset(new AdaptMIPMFrame(m_Params));
    check(lfpm_ctx_set_variant(m_frame->Context(), LFPM_VARIANT_PLAIN),
          "DepthSolver::ProcessAllFramesBatched");
    check(lfpm_ctx_set_filter_order(m_frame->Context(), m_Params.filter_reference_order
                                                            ? LFPM_FILTER_REFERENCE_ORDER
                                                            : LFPM_FILTER_SNAPSHOT),
          "DepthSolver::ProcessAllFramesBatched");
    const bool filt = m_Params.run_filter;
    check(lfpm_ctx_solve_frames(m_frame->Context(), &m_lp, n, raws.data(), &p,
                                filt ? &m_Params.filter_params : nullptr, plp.data(), csp.data(),
                                dsp.data(), slp.data(), filt ? rsp.data() : nullptr),
          "DepthSolver::ProcessAllFramesBatched");
    const int pixels = W * W;
    const float vd_base = m_Params.mla_base > 0 ? m_Params.mla_base : m_Params.baseline;
    for (int i = 0; i < n; ++i) {
        QuadTreeDisNormalMap& dn = m_MIA_dispNormal_map_map[names[i]];
        for (int v = 0; v < V; ++v) {
            auto p2 = std::make_shared<DisparityAndNormal>(m_Params);
            p2->m_ptrKey = m_keys[v];
            p2->col = m_keys[v]->GetTileX();
            p2->row = m_keys[v]->GetTileY();
            const size_t off = (size_t)v * pixels;
            for (int k = 0; k < pixels; ++k) {
                const float* q = &pl[i][4 * (off + k)];
                const float* d = &ds[i][4 * (off + k)];
                p2->ph_cuda[k] = Float4{q[0], q[1], q[2], q[3]};
                p2->d_cuda[k] = q[3];
                p2->c_cuda[k] = cs[i][off + k];
                p2->disp_v_cuda[k] = Float4{d[0], d[1], d[2], d[3]};
                p2->selected_views[k] = sl[i][off + k];
                p2->reject_reason[k] = filt ? rs[i][off + k] : 0;
            }
            p2->m_StereoStage = eSS_ACMH_Finished;
            dn[m_keys[v]] = p2;
        }
        const int w = m_lp.raw_w, h = m_lp.raw_h;
        FloatImage& m = m_DisparityMaps[names[i]];
        m.rows = h; m.cols = w; m.data.assign((size_t)w * h, 0.0f);
        FloatImage& vd = m_VirtualDepthMaps[names[i]];
        vd.rows = h; vd.cols = w; vd.data.assign((size_t)w * h, 0.0f);
        GrayImage& col = m_VDColorImages[names[i]];
        col = GrayImage(h, w);
        m_DisRange[names[i]].assign(2, 0.0f);
        m_VdRange[names[i]].assign(2, 0.0f);
        check(lfpm_ctx_frame_products(m_frame->Context(), i, w, h, m_Params.mosaic_max_cost,
                                      vd_base, m.data.data(), m_DisRange[names[i]].data(),
                                      vd.data.data(), col.data.data(),
                                      m_VdRange[names[i]].data()),
              "DepthSolver::ProcessAllFramesBatched products");
    }
    m_lastSolved.clear();
    return 0;
}

void DepthSolver::ReleaseResources(std::string& strFrame) {
    m_MIA_problem_map_map.erase(strFrame);
    m_MIA_dispNormal_map_map.erase(strFrame);
    m_DisparityMaps.erase(strFrame);
    m_VirtualDepthMaps.erase(strFrame);
    m_VDColorImages.erase(strFrame);
    m_DisRange.erase(strFrame);
    m_VdRange.erase(strFrame);
    if (m_lastSolved == strFrame) m_lastSolved.clear();
}

// ------------------------------------------------------------------------------------------
// DepthSolverManager
// ------------------------------------------------------------------------------------------
DepthSolverManager::DepthSolverManager(std::string& strRootPath, std::string& strSavePath)
    : m_ptrDepthSolver(std::make_shared<DepthSolver>(strRootPath, strSavePath)) {}

DepthSolverManager::~DepthSolverManager() = default;

void SelectNeighbors::SelectNeighborsForMIA(QuadTreeProblemMapMap::iterator& itrPM,
                                            eSelectNeighborsType eSelectType) {
    if (!m_ptrDepthSolver) return;
    if (eSelectType != eSNT_FixedPosition)
        throw std::runtime_error("SelectNeighbors: only eSNT_FixedPosition is provided");
    DepthSolver& s = *m_ptrDepthSolver;
    if (!s.m_have_layout) throw std::runtime_error("SelectNeighbors: no MLA layout set");
    const int V = (int)s.m_keys.size(), K = s.m_lp.max_neighbors;
    std::vector<int> nbr((size_t)V * K), cnt(V), garbage(V);
    check(lfpm_layout_fixed_neighbors(&s.m_lp, nbr.data(), cnt.data(), garbage.data()),
          "SelectNeighbors::SelectNeighborsForMIA");
    std::vector<float> cen(2 * (size_t)V);
    check(lfpm_layout_build(&s.m_lp, cen.data(), nullptr, nullptr, nullptr, nullptr, nullptr,
                            nullptr),
          "SelectNeighbors::SelectNeighborsForMIA");
    QuadTreeProblemMap& problem_map = itrPM->second;
    for (int v = 0; v < V; ++v) {
        auto it = problem_map.find(s.m_keys[v]);
        if (it == problem_map.end()) continue;
        MLA_Problem& problem = it->second;
        problem.m_Res_Image_KeyVec.clear();
        for (int k = 0; k < cnt[v]; ++k) {
            const int u = nbr[(size_t)v * K + k];
            const float dx = cen[2 * u] - cen[2 * v], dy = cen[2 * u + 1] - cen[2 * v + 1];
            problem.m_Res_Image_KeyVec.push_back(
                Res_image_Key{s.m_keys[u], (float)std::sqrt((double)dy * dy + (double)dx * dx)});
        }
        problem.m_bGarbage = garbage[v] != 0;
    }
}

int DepthSolverManager::DepthSolverPipelineBySequence(eStereoType eStereoType) {
    if (!m_ptrDepthSolver) return -1;
    DepthSolver& s = *m_ptrDepthSolver;
    s.SetStereoType(eStereoType);
    std::vector<std::string> names;
    for (auto& kv : s.GetRawImageMap()) names.push_back(kv.first);
    for (std::string& name : names) {
        QuadTreeProblemMap& pm = s.GetMIAProblemsMap(name);
        if (!s.GetMLAImagesSequence(name, pm)) return -1;
        s.SelectNeighborsForProblemsBySequence(name, pm);
        auto itr = s.GetMIAProblemsMapMap().find(name);
        s.ProcessProblemsImpSequence(itr, false);
        s.WriteDisMap_TileKey_new_AccuSequence(name, pm);
        s.Virtual_depth_map_TileKeySequence(name, pm);
    }
    return 0;
}

int DepthSolverManager::DepthSolverPipelineByBatch() {
    if (!m_ptrDepthSolver) return -1;
    return m_ptrDepthSolver->ProcessAllFramesBatched();
}

}  // namespace LFMVS
