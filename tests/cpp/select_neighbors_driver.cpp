// select_neighbors_driver.cpp — drives LFMVS::SelectNeighbors::SelectNeighborsForMIA with the
// FixedPosition type the way DepthSolver's pipeline calls it (host only, no GPU), and writes
// each problem's neighbour keys and garbage flag for tests/test_fixed_neighbors.py.
//
// usage: select_neighbors_driver <pitch> <raw_w> <raw_h> <max_neighbors> <out.bin>
// out.bin: int32 V, K, then per MI in layout order: X, Y, garbage, n, n x (X, Y) neighbour keys
// padded to K pairs with (-1, -1); finally int32 1 if the feature type threw, else 0.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "lfmvs.hpp"

using namespace LFMVS;

int main(int argc, char** argv) {
    if (argc != 6) {
        std::fprintf(stderr, "usage: %s pitch raw_w raw_h max_neighbors out.bin\n", argv[0]);
        return 2;
    }
    lfpm_layout_params lp{};
    lp.pitch = (float)std::atof(argv[1]);
    lp.raw_w = std::atoi(argv[2]);
    lp.raw_h = std::atoi(argv[3]);
    lp.max_neighbors = std::atoi(argv[4]);
    lp.lens_types = 1;
    lp.neighbors_per_type[0] = lp.neighbors_per_type[1] = lp.neighbors_per_type[2] =
        lp.max_neighbors;
    std::string root = ".", save = ".";
    try {
        DepthSolverManager mgr(root, save);
        DepthSolver& s = *mgr.GetDepthSolver();
        if (!s.SetMLALayout(lp)) {
            std::fprintf(stderr, "layout: %s\n", lfpm_last_error());
            return 1;
        }
        std::string name = "frame000";
        s.AddRawFrame(name, GrayImage(lp.raw_h, lp.raw_w));
        QuadTreeProblemMap& prob = s.GetMIAProblemsMap(name);
        if (!s.GetMLAImagesSequence(name, prob)) return 1;
        auto itr = s.GetMIAProblemsMapMap().find(name);
        SelectNeighbors sn(&s);
        sn.SelectNeighborsForMIA(itr, eSNT_FixedPosition);
        std::vector<int> out = {(int)prob.size(), lp.max_neighbors};
        for (auto& kv : prob) {  // key order = the layout's MI order
            const MLA_Problem& p = kv.second;
            out.push_back(kv.first->GetTileX());
            out.push_back(kv.first->GetTileY());
            out.push_back(p.m_bGarbage ? 1 : 0);
            out.push_back((int)p.m_Res_Image_KeyVec.size());
            for (int k = 0; k < lp.max_neighbors; ++k) {
                const bool has = k < (int)p.m_Res_Image_KeyVec.size();
                out.push_back(has ? p.m_Res_Image_KeyVec[k].m_ptrKey->GetTileX() : -1);
                out.push_back(has ? p.m_Res_Image_KeyVec[k].m_ptrKey->GetTileY() : -1);
            }
        }
        int threw = 0;
        try {
            sn.SelectNeighborsForMIA(itr, eSNT_Features);
        } catch (const std::runtime_error&) {
            threw = 1;
        }
        out.push_back(threw);
        FILE* f = std::fopen(argv[5], "wb");
        if (!f || std::fwrite(out.data(), 4, out.size(), f) != out.size()) return 3;
        std::fclose(f);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
