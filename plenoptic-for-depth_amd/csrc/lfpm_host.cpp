// lfpm_host.cpp — host-side parts of the C ABI: lens layout and neighbour tables, parameter
// defaults, the synthetic frame generator and the thread-local error channel.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/lfpm.h"
#include "lfpm_common.h"
#include "lfpm_internal.h"

namespace lfpm {

static thread_local std::string g_last_error;

int set_error(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

const char* last_error() { return g_last_error.c_str(); }

// ------------------------------------------------------------------------------------------
// Layout
// ------------------------------------------------------------------------------------------
int resolve_crop(const lfpm_layout_params* p) {
    // LightFieldParams::ComputeMIA_Math_Info (Common/Common.cpp:34-38): inscribed square
    return p->crop > 0 ? p->crop : (int)(p->pitch * 0.707106781186f);
}

// Ring-1 and ring-2 neighbour offsets of a cell in an even (par=0) or odd (par=1) row,
// enumerated in cube coordinates and ordered by (4*dist^2 in pitch units, angle in [0, 2pi)).
struct NbrOffset { int di, dj, d4; };

static std::vector<NbrOffset> make_offsets(int pr) {
    std::vector<NbrOffset> tab;
    {
        {
            struct E { NbrOffset o; double a; };
            std::vector<E> es;
            // centre cell (i=0, j=pr) has axial (q, r) = (0, pr) in the odd-r offset layout
            // q = i - (j - (j&1))/2; a target at axial (dq, pr+dr) sits in column
            // i = dq + (r - (r&1))/2 and row offset dj = dr.
            for (int dq = -2; dq <= 2; ++dq)
                for (int dr = -2; dr <= 2; ++dr) {
                    const int ring = std::max(std::abs(dq), std::max(std::abs(dr), std::abs(dq + dr)));
                    if (ring == 0 || ring > 2) continue;
                    const int rt = pr + dr;
                    const int di = dq + (rt - (rt & 1)) / 2;
                    const int dx2 = 2 * di + ((rt & 1) - pr);  // 2*dx / pitch
                    const int d4 = dx2 * dx2 + 3 * dr * dr;     // 4*dist^2 / pitch^2
                    double a = std::atan2(std::sqrt(3.0) * dr, (double)dx2);
                    if (a < 0) a += 2.0 * M_PI;
                    es.push_back(E{NbrOffset{di, dr, d4}, a});
                }
            std::sort(es.begin(), es.end(), [](const E& x, const E& y) {
                return x.o.d4 != y.o.d4 ? x.o.d4 < y.o.d4 : x.a < y.a;
            });
            for (auto& e : es) tab.push_back(e.o);
        }
    }
    return tab;
}

static const std::vector<NbrOffset>& neighbour_offsets(int par) {
    static const std::vector<NbrOffset> even = make_offsets(0), odd = make_offsets(1);
    return par ? odd : even;
}

struct Lattice {
    int crop, cols, rows;
    double row_step;
    std::vector<int> cell_to_mi;  // cols*rows
    std::vector<int> mi_i, mi_j;
    std::vector<float> cx, cy;
    std::vector<int> ox, oy;
};

static int build_lattice(const lfpm_layout_params* p, Lattice& L) {
    if (!p || !(p->pitch >= 4.0f) || p->raw_w <= 0 || p->raw_h <= 0)
        return set_error(LFPM_E_INVALID, "layout: bad pitch or raw size");
    L.crop = resolve_crop(p);
    if (L.crop < 3 || L.crop > LFPM_MAX_CROP)
        return set_error(LFPM_E_INVALID, "layout: crop %d outside [3, %d]", L.crop, LFPM_MAX_CROP);
    L.row_step = (double)p->pitch * 0.8660254037844386;
    L.cols = (int)std::ceil(p->raw_w / (double)p->pitch) + 1;
    L.rows = (int)std::ceil(p->raw_h / L.row_step) + 1;
    L.cell_to_mi.assign((size_t)L.cols * L.rows, -1);
    const double half = (L.crop - 1) * 0.5;
    for (int j = 0; j < L.rows; ++j)
        for (int i = 0; i < L.cols; ++i) {
            float cx, cy;
            lattice_center(p->pitch, L.crop, i, j, cx, cy);
            // abandon rule: crop must lie inside the frame (DepthSolver.cpp:1894-1903)
            const double lx = (double)cx - half, ly = (double)cy - half;
            if (lx < 0 || ly < 0) continue;
            const int ox = (int)std::floor(lx), oy = (int)std::floor(ly);
            if (ox + L.crop > p->raw_w || oy + L.crop > p->raw_h) continue;
            L.cell_to_mi[(size_t)j * L.cols + i] = (int)L.mi_i.size();
            L.mi_i.push_back(i); L.mi_j.push_back(j);
            L.cx.push_back(cx); L.cy.push_back(cy);
            L.ox.push_back(ox); L.oy.push_back(oy);
        }
    return LFPM_OK;
}

// FixedPosition neighbours, mode 1 (SelectNeighbors::SelectFixedPositionMIByMode1,
// MINeighbor/SelectNeighbors.cpp:1055-1286): key offsets (dX, dY) in the order the reference
// collects them, one list per place of the key in the u x v lens lattice.
enum FixedPlace {
    FP_INTERIOR, FP_TOP_LEFT, FP_TOP, FP_BOTTOM_LEFT, FP_BOTTOM_1, FP_BOTTOM_RIGHT,
    FP_BOTTOM_RIGHT_2, FP_TOP_RIGHT, FP_TOP_RIGHT_2, FP_LEFT, FP_LEFT_1, FP_BOTTOM, FP_RIGHT,
    FP_RIGHT_2, FP_NONE
};

struct FixedList { int n; int8_t d[12][2]; };

static const FixedList kFixedLists[FP_NONE] = {
    {12, {{1, 0}, {-1, 0}, {2, 0}, {0, -1}, {-2, -1}, {0, 1}, {-2, 0}, {-1, -1}, {1, -1},
          {-1, 1}, {1, 1}, {2, 1}}},                                              // :1065-1087
    {6, {{1, 0}, {2, 0}, {0, 1}, {1, 1}, {2, 2}, {0, 2}}},                        // :1094-1104
    {6, {{1, 0}, {-1, 0}, {2, 0}, {0, 1}, {-1, 1}, {1, 1}}},                      // :1110-1120
    {6, {{1, 0}, {2, 0}, {0, -1}, {1, -1}, {2, -2}, {0, -2}}},                    // :1125-1135
    {6, {{1, 0}, {-1, 0}, {2, 0}, {0, -1}, {-1, -1}, {1, -1}}},                   // :1140-1150
    {6, {{-1, 0}, {0, -1}, {-2, -1}, {-2, 0}, {0, -2}, {-1, -1}}},                // :1155-1165
    {6, {{1, 0}, {-1, 0}, {0, -1}, {-2, -1}, {-1, -1}, {1, -1}}},                 // :1170-1180
    {6, {{-1, 0}, {0, 1}, {-2, 2}, {-2, 0}, {0, 2}, {-1, 1}}},                    // :1185-1195
    {6, {{1, 0}, {-1, 0}, {0, 1}, {-2, 0}, {-1, 1}, {1, 1}}},                     // :1200-1210
    {6, {{1, 0}, {2, 0}, {0, -1}, {0, 1}, {1, -1}, {1, 1}}},                      // :1215-1225
    {6, {{1, 0}, {-1, 0}, {2, 0}, {0, -1}, {0, 1}, {-1, -1}}},                    // :1230-1240
    {6, {{1, 0}, {-1, 0}, {2, 0}, {0, -1}, {-2, -1}, {1, -1}}},                   // :1245-1255
    {6, {{-1, 0}, {0, -1}, {-2, -1}, {0, 1}, {-1, -1}, {-1, 1}}},                 // :1260-1270
    {6, {{1, 0}, {-1, 0}, {0, -1}, {-2, -1}, {0, 1}, {1, 1}}},                    // :1275-1285
};

// Place of key (X, Y), tested in the reference's order.  Its top-row test
// `(X == 1 && Y == 0) || ((X != u-2 || X != u-1) && Y == 0)` holds for every X on row 0, so
// FP_TOP_RIGHT and FP_TOP_RIGHT_2 are never chosen, as in the reference.
static FixedPlace fixed_place(int X, int Y, int u, int v) {
    if (X > 1 && X < u - 2 && Y > 0 && Y < v - 1) return FP_INTERIOR;
    if (X == 0 && Y == 0) return FP_TOP_LEFT;
    if (Y == 0) return FP_TOP;
    if (X == 0 && Y == v - 1) return FP_BOTTOM_LEFT;
    if (X == 1 && Y == v - 1) return FP_BOTTOM_1;
    if (X == u - 1 && Y == v - 1) return FP_BOTTOM_RIGHT;
    if (X == u - 2 && Y == v - 1) return FP_BOTTOM_RIGHT_2;
    if (X == 0) return FP_LEFT;
    if (X == 1) return FP_LEFT_1;
    if (Y == v - 1) return FP_BOTTOM;
    if (X == u - 1) return FP_RIGHT;
    if (X == u - 2) return FP_RIGHT_2;
    return FP_NONE;
}

} // namespace lfpm

using namespace lfpm;

namespace lfpm {
// FixedPosition neighbour table over a lens lattice of cols x rows keys: MI v sits at key
// (mi_i[v], mi_j[v]) with centre (cx[v], cy[v]); cell_to_mi maps a key to its MI (-1: none).
int fixed_neighbors_core(int V, const int* mi_i, const int* mi_j, const float* cx,
                         const float* cy, const int* cell_to_mi, int cols, int rows, int K,
                         int* neighbor_ids, int* neighbor_counts, int* garbage) {
    if (K <= 0 || K > LFPM_MAX_NEIGHBORS)
        return set_error(LFPM_E_INVALID, "fixed neighbours: max_neighbors %d outside [1, %d]", K,
                         LFPM_MAX_NEIGHBORS);
    struct Nb { int id; float len; };
    std::vector<Nb> nb;
    for (int v = 0; v < V; ++v) {
        const int X = mi_i[v], Y = mi_j[v];
        nb.clear();
        const FixedPlace place = fixed_place(X, Y, cols, rows);
        if (place != FP_NONE) {
            const FixedList& fl = kFixedLists[place];
            for (int e = 0; e < fl.n; ++e) {
                // CollectValidNeighborKey (:1330-1352): the key must name a kept lens
                const int nx = X + fl.d[e][0], ny = Y + fl.d[e][1];
                if (nx < 0 || ny < 0 || nx >= cols || ny >= rows) continue;
                const int u = cell_to_mi[(size_t)ny * cols + nx];
                if (u < 0) continue;
                // CreateResTileKeysFromNeiKeyVec (MVStereo/LFDepthInfo.cpp:431-453): centre
                // distance from float differences, squared and summed in double
                const float dx = cx[u] - cx[v], dy = cy[u] - cy[v];
                nb.push_back(Nb{u, (float)std::sqrt((double)dy * dy + (double)dx * dx)});
            }
        }
        // ascending distance; ties keep the collection order (std::sort leaves them unspecified)
        std::stable_sort(nb.begin(), nb.end(), [](const Nb& a, const Nb& b) { return a.len < b.len; });
        const int c = std::min((int)nb.size(), K);
        if (neighbor_ids)
            for (int k = 0; k < K; ++k) neighbor_ids[(size_t)v * K + k] = k < c ? nb[k].id : -1;
        if (neighbor_counts) neighbor_counts[v] = c;
        // DepthSolver::ConfirmProblemForEstimation (DepthSolver.cpp:2933-2946) with
        // m_iGarbageRows = 1 (:117): first key row/column, or no neighbour at all
        if (garbage) garbage[v] = (nb.empty() || X < 1 || Y < 1) ? 1 : 0;
    }
    return LFPM_OK;
}
}  // namespace lfpm

extern "C" {

const char* lfpm_last_error(void) { return lfpm::last_error(); }

void lfpm_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    philox4x32_10(ctr[0], ctr[1], ctr[2], ctr[3], key[0], key[1], out);
}

int lfpm_layout_query(const lfpm_layout_params* p, lfpm_layout_info* out) {
    Lattice L;
    const int rc = build_lattice(p, L);
    if (rc) return rc;
    if (out) {
        out->num_mis = (int)L.mi_i.size();
        out->crop = L.crop;
        out->lattice_cols = L.cols;
        out->lattice_rows = L.rows;
        out->row_step = (float)L.row_step;
    }
    return LFPM_OK;
}

int lfpm_layout_build(const lfpm_layout_params* p, float* centers, int* origins, int* tilekeys,
                      int* lens_type, int* neighbor_ids, int* neighbor_counts,
                      int* lattice_to_mi) {
    Lattice L;
    const int rc = build_lattice(p, L);
    if (rc) return rc;
    const int K = p->max_neighbors;
    if (K <= 0 || K > LFPM_MAX_NEIGHBORS)
        return set_error(LFPM_E_INVALID, "layout: max_neighbors %d outside [1, %d]", K,
                         LFPM_MAX_NEIGHBORS);
    if (p->lens_types != 1 && p->lens_types != 3)
        return set_error(LFPM_E_INVALID, "layout: lens_types must be 1 or 3");
    const int V = (int)L.mi_i.size();
    for (int v = 0; v < V; ++v) {
        const int i = L.mi_i[v], j = L.mi_j[v];
        if (centers) { centers[2 * v] = L.cx[v]; centers[2 * v + 1] = L.cy[v]; }
        if (origins) { origins[2 * v] = L.ox[v]; origins[2 * v + 1] = L.oy[v]; }
        if (tilekeys) { tilekeys[2 * v] = i; tilekeys[2 * v + 1] = j; }
        // multi-focus lens type: 3-colouring of the hex lattice, (q - r) mod 3 in axial coords
        const int q = i - (j - (j & 1)) / 2;
        const int t = p->lens_types == 3 ? (((q - j) % 3) + 3) % 3 : 0;
        if (lens_type) lens_type[v] = t;
        if (!neighbor_ids && !neighbor_counts) continue;
        int cap = p->lens_types == 3 ? p->neighbors_per_type[t] : K;
        cap = std::max(0, std::min(cap, K));
        const int d4max = cap <= 6 ? 4 : (cap <= 12 ? 12 : 16);
        int c = 0;
        for (const NbrOffset& o : neighbour_offsets(j & 1)) {
            if (c >= cap || o.d4 > d4max) break;
            const int ni = i + o.di, nj = j + o.dj;
            if (ni < 0 || nj < 0 || ni >= L.cols || nj >= L.rows) continue;
            const int u = L.cell_to_mi[(size_t)nj * L.cols + ni];
            if (u < 0) continue;
            if (neighbor_ids) neighbor_ids[(size_t)v * K + c] = u;
            ++c;
        }
        if (neighbor_ids)
            for (int k = c; k < K; ++k) neighbor_ids[(size_t)v * K + k] = -1;
        if (neighbor_counts) neighbor_counts[v] = c;
    }
    if (lattice_to_mi) std::memcpy(lattice_to_mi, L.cell_to_mi.data(), L.cell_to_mi.size() * 4);
    return LFPM_OK;
}

int lfpm_layout_fixed_neighbors(const lfpm_layout_params* p, int* neighbor_ids,
                                int* neighbor_counts, int* garbage) {
    Lattice L;
    const int rc = build_lattice(p, L);
    if (rc) return rc;
    return lfpm::fixed_neighbors_core((int)L.mi_i.size(), L.mi_i.data(), L.mi_j.data(),
                                      L.cx.data(), L.cy.data(), L.cell_to_mi.data(), L.cols,
                                      L.rows, p->max_neighbors, neighbor_ids, neighbor_counts,
                                      garbage);
}

void lfpm_pm_params_default(lfpm_pm_params* p, int crop, float base) {
    std::memset(p, 0, sizeof(*p));
    p->crop = crop;
    p->base = base;
    p->radius = 2;                  // patch_size 5 (AdaptMIPMFrame.cpp:129)
    p->radius_increment = 1;
    p->sigma_spatial = 5.0f;
    p->sigma_color = 3.0f / 255.0f; // reference default 3.0 on 0..255
    p->min_var = 1e-5f / (255.0f * 255.0f);
    p->top_k = 5;
    p->depth_min = 0.0f;            // AdaptMIPMFrame.cpp:133-134
    p->depth_max = base * 0.5f;
    p->iterations = 3;
    p->max_neighbors = 16;
    p->seed = 1234u;
    p->frame_uid = 0u;
    p->storage = 0;
    p->cos_max_slant = 0.5f;
    p->blur_lambda = 0.0f;          // blur term off: it needs per-MI blur maps
}

void lfpm_filter_params_default(lfpm_filter_params* f) {
    // MIStereoMatch::FilterAndStatic configuration (MIStereo/MIStereoMatch.cpp:759-778)
    f->max_triplet = 4;
    f->min_valid_disp = 0.0f;
    f->use_selected_views_only = 1;
    f->clear_selected_views_when_invalid = 1;
    f->enable_cost_filter = 1;
    f->max_cost = 1.8f;
    f->enable_cycle_geo_filter = 1;
    f->max_geo_err_px = 0.5f;
    f->enable_cycle_photo_filter = 1;
    f->max_photo_err_u = 0.5f;
    f->min_good_neighbors = 2;
    f->enable_spike_filter = 1;
    f->spike_abs_diff = 1.0f;
    f->spike_min_neighbors = 3;
}

// ------------------------------------------------------------------------------------------
// Synthetic focused-plenoptic frame
// ------------------------------------------------------------------------------------------
namespace {

struct PlaneV { double sx, sy, a, b, e; };

static uint32_t ph_word(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t k0, uint32_t k1, int w) {
    uint32_t o[4];
    philox4x32_10(c0, c1, c2, 0u, k0, k1, o);
    return o[w];
}

static double u01d(uint32_t x) { return ((double)(x >> 8) + 0.5) * 5.9604644775390625e-08; }

struct ValueNoise {
    int period;
    int gx0, gy0, gw, gh;  // lattice window
    std::vector<float> val;
    double amp;
    double sample(double X, double Y) const {
        const double fx = X / period - gx0, fy = Y / period - gy0;
        int ix = (int)std::floor(fx), iy = (int)std::floor(fy);
        double tx = fx - ix, ty = fy - iy;
        ix = std::min(std::max(ix, 0), gw - 2);
        iy = std::min(std::max(iy, 0), gh - 2);
        tx = tx * tx * (3.0 - 2.0 * tx);
        ty = ty * ty * (3.0 - 2.0 * ty);
        const double a = val[(size_t)iy * gw + ix], b = val[(size_t)iy * gw + ix + 1];
        const double c = val[(size_t)(iy + 1) * gw + ix], d = val[(size_t)(iy + 1) * gw + ix + 1];
        const double top = a + tx * (b - a), bot = c + tx * (d - c);
        return amp * (top + ty * (bot - top));
    }
};

} // namespace

int lfpm_synth_frame(const lfpm_synth_params* sp, uint8_t* raw, float* gt_v) {
    if (!sp || !raw || sp->raw_w <= 0 || sp->raw_h <= 0 || !(sp->pitch >= 4.0f) ||
        sp->num_planes <= 0 || !(sp->vmax > sp->vmin) || !(sp->vmin > 1.0f))
        return set_error(LFPM_E_INVALID, "synth: bad parameters");
    const int Wr = sp->raw_w, Hr = sp->raw_h;
    const int crop = (int)(sp->pitch * 0.707106781186f);
    const double pitch = sp->pitch, rs = pitch * 0.8660254037844386;
    const double diag = std::sqrt((double)Wr * Wr + (double)Hr * Hr);
    // planes v(X) = v_k + g*(cos t*(X-s).x + sin t*(X-s).y) on a Voronoi partition of seeds s_k;
    // |grad v| * diag <= (vmax-vmin)/4 keeps v inside [vmin, vmax] over the virtual image.
    std::vector<PlaneV> planes(sp->num_planes);
    const double gmax = 0.25 * (sp->vmax - sp->vmin) / diag;
    for (int k = 0; k < sp->num_planes; ++k) {
        uint32_t o[4];
        philox4x32_10((uint32_t)k, 0u, 0u, 0u, sp->scene_seed, 0x5CE0E0u, o);
        const double sx = u01d(o[0]) * Wr, sy = u01d(o[1]) * Hr;
        const double th = 2.0 * M_PI * u01d(o[2]);
        const double g = gmax * u01d(o[3]);
        philox4x32_10((uint32_t)k, 1u, 0u, 0u, sp->scene_seed, 0x5CE0E0u, o);
        const double margin = g * diag + 0.02 * (sp->vmax - sp->vmin);
        const double vk = sp->vmin + margin + u01d(o[0]) * (sp->vmax - sp->vmin - 2.0 * margin);
        const double a = g * std::cos(th), b = g * std::sin(th);
        planes[k] = PlaneV{sx, sy, a, b, vk - a * sx - b * sy};
    }
    auto owner = [&](double X, double Y) {
        int best = 0;
        double bd = 1e300;
        for (int k = 0; k < (int)planes.size(); ++k) {
            const double dx = X - planes[k].sx, dy = Y - planes[k].sy;
            const double d = dx * dx + dy * dy;
            if (d < bd) { bd = d; best = k; }
        }
        return best;
    };
    // value-noise texture: 3 octaves, lattice periods 32/16/8 virtual px (2-16 raw px at virtual
    // depth 2-8, below the ~1 px aliasing limit of finer octaves), amplitudes 0.5/0.3/0.2
    const int periods[3] = {32, 16, 8};
    const double amps[3] = {0.5, 0.3, 0.2};
    const double ext = 0.5 * pitch * sp->vmax + 2.0 * pitch;
    std::vector<ValueNoise> oct(3);
    for (int o = 0; o < 3; ++o) {
        ValueNoise& vn = oct[o];
        vn.period = periods[o];
        vn.amp = amps[o];
        vn.gx0 = (int)std::floor(-ext / vn.period) - 1;
        vn.gy0 = (int)std::floor(-ext / vn.period) - 1;
        vn.gw = (int)std::ceil((Wr + ext) / vn.period) - vn.gx0 + 2;
        vn.gh = (int)std::ceil((Hr + ext) / vn.period) - vn.gy0 + 2;
        vn.val.resize((size_t)vn.gw * vn.gh);
        for (int y = 0; y < vn.gh; ++y)
            for (int x = 0; x < vn.gw; x += 4) {
                uint32_t w4[4];
                philox4x32_10((uint32_t)x, (uint32_t)y, (uint32_t)o, 0u, sp->scene_seed, 0x7E47u,
                              w4);
                for (int q = 0; q < 4 && x + q < vn.gw; ++q)
                    vn.val[(size_t)y * vn.gw + x + q] = (float)u01d(w4[q]);
            }
    }
    const double rdisc = 0.5 * pitch;
    for (int y = 0; y < Hr; ++y) {
        for (int x = 0; x < Wr; ++x) {
            // nearest lattice centre
            const int j0 = (int)std::floor((y - 0.5 * pitch) / rs + 0.5);
            double bd = 1e300, cx = 0, cy = 0;
            for (int j = j0 - 1; j <= j0 + 1; ++j) {
                const double off = 0.5 * pitch * (1 + (j & 1));
                const int i0 = (int)std::floor((x - off) / pitch + 0.5);
                for (int i = i0 - 1; i <= i0 + 1; ++i) {
                    float fcx, fcy;
                    lattice_center(sp->pitch, crop, i, j, fcx, fcy);
                    const double dx = x - (double)fcx, dy = y - (double)fcy;
                    const double d = dx * dx + dy * dy;
                    if (d < bd) { bd = d; cx = fcx; cy = fcy; }
                }
            }
            const size_t idx = (size_t)y * Wr + x;
            if (std::sqrt(bd) > rdisc) {
                raw[idx] = 0;
                if (gt_v) gt_v[idx] = 0.0f;
                continue;
            }
            // solve X = c + (u - c) * v(X) for the plane owning X (fixed point over <= 2 cells)
            const double ddx = x - cx, ddy = y - cy;
            int k = owner(cx, cy);
            double v = 0, X = cx, Y = cy;
            for (int it = 0; it < 3; ++it) {
                const PlaneV& P = planes[k];
                v = (P.a * cx + P.b * cy + P.e) / (1.0 - P.a * ddx - P.b * ddy);
                X = cx + ddx * v;
                Y = cy + ddy * v;
                const int k2 = owner(X, Y);
                if (k2 == k) break;
                k = k2;
            }
            double T = 0;
            for (int o = 0; o < 3; ++o) T += oct[o].sample(X, Y);
            uint32_t nz[4];
            philox4x32_10((uint32_t)x, (uint32_t)y, 0u, 0u, sp->noise_seed, 0xB0B0u, nz);
            const double u1 = u01d(nz[0]), u2 = u01d(nz[1]);
            const double gauss = std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
            double val = 255.0 * T + sp->noise_sigma * gauss;
            val = std::floor(val + 0.5);
            val = std::min(255.0, std::max(0.0, val));
            raw[idx] = (uint8_t)val;
            if (gt_v) gt_v[idx] = (float)v;
        }
    }
    (void)ph_word;
    return LFPM_OK;
}

// ------------------------------------------------------------------------------------------
// File writers of the per-frame products (host side; the payloads are built on the device)
// ------------------------------------------------------------------------------------------
int lfpm_write_ply_points(const char* path, const uint8_t* records, uint64_t count) {
    if (!path || (count && !records)) return lfpm::set_error(LFPM_E_INVALID, "write_ply: null");
    FILE* f = std::fopen(path, "wb");
    if (!f) return lfpm::set_error(LFPM_E_INVALID, "write_ply: cannot open %s", path);
    // the header of StoreColorPlyFileBinaryPointCloud (MVStereo/ACMP.cpp:458-470)
    std::fprintf(f, "ply\nformat binary_little_endian 1.0\nelement vertex %llu\n",
                 (unsigned long long)count);
    static const char* const props[] = {"float x", "float y", "float z", "float nx",
                                        "float ny", "float nz", "uchar red", "uchar green",
                                        "uchar blue"};
    for (const char* q : props) std::fprintf(f, "property %s\n", q);
    std::fprintf(f, "end_header\n");
    const size_t n = (size_t)count * 27;
    const bool ok = n == 0 || std::fwrite(records, 1, n, f) == n;
    if (std::fclose(f) != 0 || !ok)
        return lfpm::set_error(LFPM_E_INVALID, "write_ply: short write to %s", path);
    return LFPM_OK;
}

int lfpm_write_dmb(const char* path, const float* data, int height, int width) {
    if (!path || !data || height < 0 || width < 0)
        return lfpm::set_error(LFPM_E_INVALID, "write_dmb: bad arguments");
    FILE* f = std::fopen(path, "wb");
    if (!f) return lfpm::set_error(LFPM_E_INVALID, "write_dmb: cannot open %s", path);
    // writeDepthDmb (MVStereo/ACMP.cpp:367-390): int32 type = 1, rows, cols, bands = 1, floats
    const int32_t hdr[4] = {1, height, width, 1};
    const size_t n = (size_t)height * width;
    bool ok = std::fwrite(hdr, sizeof(int32_t), 4, f) == 4;
    ok = ok && (n == 0 || std::fwrite(data, sizeof(float), n, f) == n);
    if (std::fclose(f) != 0 || !ok)
        return lfpm::set_error(LFPM_E_INVALID, "write_dmb: short write to %s", path);
    return LFPM_OK;
}

int lfpm_read_dmb(const char* path, float* data, size_t capacity, int* height, int* width) {
    if (!path || !height || !width) return lfpm::set_error(LFPM_E_INVALID, "read_dmb: null");
    FILE* f = std::fopen(path, "rb");
    if (!f) return lfpm::set_error(LFPM_E_INVALID, "read_dmb: cannot open %s", path);
    // readDepthDmb (MVStereo/ACMP.cpp:338-365): type must be 1; h * w * bands floats follow
    int32_t hdr[4] = {-1, 0, 0, 0};
    if (std::fread(hdr, sizeof(int32_t), 4, f) != 4 || hdr[0] != 1 || hdr[1] < 0 || hdr[2] < 0 ||
        hdr[3] < 1) {
        std::fclose(f);
        return lfpm::set_error(LFPM_E_INVALID, "read_dmb: %s is not a type-1 dmb file", path);
    }
    *height = hdr[1];
    *width = hdr[2];
    const size_t n = (size_t)hdr[1] * hdr[2] * hdr[3];
    if (!data) {  // size query
        std::fclose(f);
        return LFPM_OK;
    }
    if (capacity < n) {
        std::fclose(f);
        return lfpm::set_error(LFPM_E_INVALID, "read_dmb: capacity %zu < %zu floats", capacity, n);
    }
    const bool ok = n == 0 || std::fread(data, sizeof(float), n, f) == n;
    std::fclose(f);
    if (!ok) return lfpm::set_error(LFPM_E_INVALID, "read_dmb: %s is truncated", path);
    return LFPM_OK;
}

// Level schedule of the ordered CheckModule pass: MI v reads the filtered state of every
// neighbour n < v, so its level is one more than the largest level among those neighbours
// (0 without any).  MIs of one level read no filtered state of each other; levels run in order.
int lfpm_filter_order_levels(const int* neighbor_ids, const int* neighbor_counts, int num_mis,
                             int max_neighbors, int* mi_order, int* level_start,
                             int* num_levels) {
    if (!neighbor_ids || !neighbor_counts || !mi_order || !level_start || !num_levels ||
        num_mis <= 0 || max_neighbors < 1 || max_neighbors > LFPM_MAX_NEIGHBORS)
        return lfpm::set_error(LFPM_E_INVALID, "filter_order_levels: bad arguments");
    std::vector<int> level((size_t)num_mis, 0);
    int top = 0;
    for (int v = 0; v < num_mis; ++v) {
        const int c = std::min(neighbor_counts[v], max_neighbors);
        int l = 0;
        for (int b = 0; b < c; ++b) {
            const int n = neighbor_ids[(size_t)v * max_neighbors + b];
            if (n >= 0 && n < v) l = std::max(l, level[(size_t)n] + 1);
        }
        level[(size_t)v] = l;
        top = std::max(top, l);
    }
    const int L = top + 1;
    std::vector<int> count((size_t)L + 1, 0);
    for (int v = 0; v < num_mis; ++v) ++count[(size_t)level[(size_t)v] + 1];
    for (int l = 0; l < L; ++l) count[(size_t)l + 1] += count[(size_t)l];
    for (int l = 0; l <= L; ++l) level_start[l] = count[(size_t)l];
    for (int v = 0; v < num_mis; ++v) mi_order[count[(size_t)level[(size_t)v]]++] = v;
    *num_levels = L;
    return LFPM_OK;
}

} // extern "C"
