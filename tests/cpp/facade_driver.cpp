// facade_driver.cpp — drives the LFMVS C++ facades the way the reference's main() drives
// DepthSolverManager, on synthetic frames, and writes every result as raw binary for the GPU
// test (tests/test_gpu_facade.py) to compare with the C-ABI path.
//
// usage: facade_driver <cfg.bin> <out_dir>
// cfg.bin: lfpm_layout_params, lfpm_synth_params, lfpm_pm_params, lfpm_filter_params (raw
// struct bytes), then int32 nframes, int32 mode (0 = DepthSolverPipelineBySequence,
// 1 = DepthSolverPipelineByBatch, 2 = AdaptMIPM on the middle micro-image of frame 0,
// 3 = DepthSolverPipelineBySequence with the AdaptMIPMFrameACMM matcher).
// Frame i uses scene/noise seeds + 17*i.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "lfmvs.hpp"

using namespace LFMVS;

static void put(const std::string& path, const void* p, size_t n) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f || std::fwrite(p, 1, n, f) != n) {
        std::fprintf(stderr, "write %s failed\n", path.c_str());
        std::exit(3);
    }
    std::fclose(f);
}

int main(int argc, char** argv) {
    if (argc != 3) {
        std::fprintf(stderr, "usage: %s cfg.bin out_dir\n", argv[0]);
        return 2;
    }
    lfpm_layout_params lp;
    lfpm_synth_params sp;
    lfpm_pm_params pm;
    lfpm_filter_params fp;
    int nframes = 0, mode = 0;
    FILE* f = std::fopen(argv[1], "rb");
    if (!f) return 2;
    bool ok = std::fread(&lp, sizeof lp, 1, f) == 1 && std::fread(&sp, sizeof sp, 1, f) == 1 &&
              std::fread(&pm, sizeof pm, 1, f) == 1 && std::fread(&fp, sizeof fp, 1, f) == 1 &&
              std::fread(&nframes, 4, 1, f) == 1 && std::fread(&mode, 4, 1, f) == 1;
    std::fclose(f);
    if (!ok || nframes < 1) return 2;
    const std::string out = argv[2];
    std::string root = ".", save = out;
    try {
        DepthSolverManager mgr(root, save);
        DepthSolver& s = *mgr.GetDepthSolver();
        LightFieldParams& lf = s.GetLightFieldParams();
        lf.solve_params = pm;
        lf.solve_params_set = true;
        lf.filter_params = fp;
        lf.run_filter = true;
        lf.frame_matcher_acmm = mode == 3;
        if (!s.SetMLALayout(lp)) {
            std::fprintf(stderr, "layout: %s\n", lfpm_last_error());
            return 1;
        }
        std::vector<std::string> names;
        for (int i = 0; i < nframes; ++i) {
            lfpm_synth_params spi = sp;
            spi.scene_seed += 17u * (uint32_t)i;
            spi.noise_seed += 17u * (uint32_t)i;
            GrayImage raw(sp.raw_h, sp.raw_w);
            if (lfpm_synth_frame(&spi, raw.data.data(), nullptr) != LFPM_OK) return 1;
            char name[32];
            std::snprintf(name, sizeof name, "frame%03d", i);
            names.push_back(name);
            s.AddRawFrame(name, raw);
        }
        if (mode == 2) {
            QuadTreeProblemMap& prob = s.GetMIAProblemsMap(names[0]);
            if (!s.GetMLAImagesSequence(names[0], prob)) return 1;
            s.SelectNeighborsForProblemsBySequence(names[0], prob);
            auto it = prob.begin();
            std::advance(it, prob.size() / 2);
            AdaptMIPM mi(lf);
            std::vector<Float4> planeVec;
            std::vector<float> costVec;
            mi.Initialize(s.GetMLAInfoMap(), it->second, prob, planeVec, costVec);
            mi.RunPatchMatchCUDAForMI();
            const int px = lf.mi_width_for_match * lf.mi_width_for_match;
            const int V = mi.NumViews();
            put(out + "/mi_plane.bin", mi.GetPlaneHypothesisVector(), px * 16);
            put(out + "/mi_cost.bin", mi.GetCostVector(), px * 4);
            put(out + "/mi_tiles.bin", mi.Tiles().data(), mi.Tiles().size());
            put(out + "/mi_centers.bin", mi.Centers().data(), mi.Centers().size() * 4);
            put(out + "/mi_nbr.bin", mi.NeighborIds().data(), mi.NeighborIds().size() * 4);
            put(out + "/mi_cnt.bin", mi.NeighborCounts().data(), mi.NeighborCounts().size() * 4);
            const int meta[3] = {V, mi.GetParamsGPU().max_neighbors, px};
            put(out + "/mi_meta.bin", meta, sizeof meta);
            // AdaptMIPMPFPGR on the same MI: collect, then the second stage (repair)
            AdaptMIPMPFPGR pf(lf);
            pf.Initialize(s.GetMLAInfoMap(), it->second, prob, planeVec, costVec);
            pf.RunPatchMatchCUDAForMI_PFPGR_Collect();
            put(out + "/pf_collect_plane.bin", pf.GetPlaneHypothesisVector(), px * 16);
            std::vector<int> pgr;
            for (int i = 0; i < px; ++i) {
                const Int3 g = pf.GetNeighborPGR(i);
                pgr.insert(pgr.end(), {g.x, g.y, g.z});
            }
            put(out + "/pf_pgr.bin", pgr.data(), pgr.size() * 4);
            std::vector<Proxy_DisPlane> proxies(px);
            StereoResultInfo first(lf.mi_width_for_match, lf.mi_width_for_match);
            pf.Initialize_SecondStage(s.GetMLAInfoMap(), it->second, prob, planeVec, costVec,
                                      proxies.data(), &first);
            pf.RunPatchMatchCUDAForMI_PFPGR_Repair();
            put(out + "/pf_repair_plane.bin", pf.GetPlaneHypothesisVector(), px * 16);
            std::printf("facade_driver: AdaptMIPM views=%d\n", V);
            return 0;
        }
        const int rc = mode != 1 ? mgr.DepthSolverPipelineBySequence(eST_ACMH)
                                 : mgr.DepthSolverPipelineByBatch();
        if (rc != 0) {
            std::fprintf(stderr, "pipeline rc=%d: %s\n", rc, lfpm_last_error());
            return 1;
        }
        for (const std::string& name : names) {
            QuadTreeDisNormalMap& dn = s.GetDisNormalMap(name);
            std::vector<float> plane, cost;
            std::vector<uint32_t> sel;
            std::vector<uint8_t> reason;
            for (auto& kv : dn) {  // key order = the layout's MI order
                const DisparityAndNormal& d = *kv.second;
                for (size_t i = 0; i < d.ph_cuda.size(); ++i) {
                    plane.insert(plane.end(), {d.ph_cuda[i].x, d.ph_cuda[i].y, d.ph_cuda[i].z,
                                               d.ph_cuda[i].w});
                    cost.push_back(d.c_cuda[i]);
                    sel.push_back(d.selected_views[i]);
                    reason.push_back(d.reject_reason[i]);
                }
            }
            const std::string b = out + "/" + name;
            put(b + "_plane.bin", plane.data(), plane.size() * 4);
            put(b + "_cost.bin", cost.data(), cost.size() * 4);
            put(b + "_sel.bin", sel.data(), sel.size() * 4);
            put(b + "_reason.bin", reason.data(), reason.size());
            const FloatImage& m = s.GetDisparityMap(name);
            const FloatImage& vd = s.GetVirtualDepthMap(name);
            const GrayImage& col = s.GetVirtualDepthColor(name);
            put(b + "_mosaic.bin", m.data.data(), m.data.size() * 4);
            put(b + "_vd.bin", vd.data.data(), vd.data.size() * 4);
            put(b + "_color.bin", col.data.data(), col.data.size());
            std::printf("facade_driver: %s MIs=%zu mosaic=%dx%d\n", name.c_str(), dn.size(),
                        m.cols, m.rows);
        }
    } catch (const std::exception& e) {
        std::fprintf(stderr, "facade_driver: %s\n", e.what());
        return 1;
    }
    return 0;
}
